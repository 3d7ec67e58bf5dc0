"""Snapshot output (les_output.cpp).

NetCDF: the six files of module_LES_write_netcdf.f95 are read back with scipy's reader of the
netCDF classic format, an implementation independent of the writer: dimensions, variables,
units, coordinate values, the written time level and the fill value elsewhere.
anime: data22/data23 are compared byte for byte with the files anime.f95 itself writes, built
from the reference sources into the reference harness (stage 13), on the same arrays.
"""
import os

import numpy as np
import pytest

import common
from les_halo_amd import fields as fl
from les_halo_amd import inputs
from les_halo_amd.output import NetCDFOutput, anime_arrays

FILL = np.float32(9.9692099683868690e+36)  # NC_FILL_FLOAT
NAMES = ("p", "u", "v", "w", "usum", "vsum", "wsum")
FILES = {"": [("pressure", "p", "hPa")], "_p": [("pressure", "p", "hPa")],
         "_u": [("velocity_x", "u", "m/s")], "_v": [("velocity_y", "v", "m/s")],
         "_w": [("velocity_z", "w", "m/s")],
         "_uvwsum": [("velocity_sum_x", "usum", "m/s"), ("velocity_sum_y", "vsum", "m/s"),
                     ("velocity_sum_z", "wsum", "m/s")]}


def _random(name, dims, rng):
    return rng.standard_normal(fl.size(name, *dims)).astype(np.float32)


def _check_netcdf(directory, dims, nt, t, arrs):
    netcdf_file = pytest.importorskip("scipy.io").netcdf_file
    for suffix, variables in FILES.items():
        path = os.path.join(str(directory), "LES_output%s.nc" % suffix)
        with open(path, "rb") as f:
            assert f.read(4) == b"CDF\x01"
        ni, nj, nk = fl.shape(variables[0][1], *dims)
        nc = netcdf_file(path, "r", mmap=False)
        try:
            assert dict(nc.dimensions) == {"level": nk, "latitude": nj, "longitude": ni,
                                           "time": nt}
            assert list(nc.variables)[:3] == ["latitude", "longitude", "time"]
            lat, lon, tim = nc.variables["latitude"], nc.variables["longitude"], nc.variables["time"]
            assert lat.units == b"degrees_north" and lon.units == b"degrees_east"
            assert tim.units == b"hours since 2014-08-05 16:00:00"
            assert lat.data.dtype == np.dtype(">f4") and tim.data.dtype == np.dtype(">f8")
            np.testing.assert_array_equal(lat.data, np.arange(1, nj + 1, dtype=np.float32))
            np.testing.assert_array_equal(lon.data, np.arange(1, ni + 1, dtype=np.float32))
            np.testing.assert_array_equal(tim.data, np.arange(1, nt + 1, dtype=np.float64))
            for vname, field, units in variables:
                var = nc.variables[vname]
                assert var.dimensions == ("time", "level", "latitude", "longitude")
                assert var.units == units.encode()
                data = np.asarray(var.data)
                assert data.shape == (nt, nk, nj, ni)
                # the Fortran array (i fastest) is the C-ordered (level, lat, lon) slab
                want = arrs[field].reshape(nk, nj, ni)
                got = data[t - 1].astype(np.float32)  # big-endian in the file
                assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), vname
                for other in range(nt):
                    if other != t - 1:
                        assert np.all(data[other] == FILL), (vname, other)
        finally:
            nc.close()


def test_netcdf_files_read_back(tmp_path):
    dims, nt, t = (6, 5, 4), 3, 2
    rng = np.random.default_rng(5)
    arrs = {n: _random(n, dims, rng) for n in NAMES}
    with NetCDFOutput(tmp_path, *dims, ntimes=nt) as nc:
        nc.write_arrays(t, *[arrs[n] for n in NAMES])
    _check_netcdf(tmp_path, dims, nt, t, arrs)


def test_netcdf_rejects_time_outside_the_dimension(tmp_path):
    dims = (4, 3, 2)
    rng = np.random.default_rng(1)
    arrs = [_random(n, dims, rng) for n in NAMES]
    with NetCDFOutput(tmp_path, *dims, ntimes=2) as nc:
        for t in (0, 3):
            with pytest.raises(RuntimeError):
                nc.write_arrays(t, *arrs)


def _anime_inputs(dims, seed):
    case = inputs.make_case(*dims, seed=seed, hmax=12.0, coverage=0.3)
    rng = np.random.default_rng(seed)
    state = {n: np.zeros(fl.size(n, *dims), np.float32) for n in fl.STATE_3D + fl.COEFFS}
    for n in ("u", "v", "w"):
        state[n] = (3.0 * rng.standard_normal(fl.size(n, *dims))).astype(np.float32)
    return case, state


def test_anime_schedule_writes_only_at_n0_nmax_and_every_1000(tmp_path):
    dims = (5, 4, 3)
    case, state = _anime_inputs(dims, 2)
    args = [case["dx1"], case["dy1"], case["z2"], state["u"], state["v"], state["w"],
            case["amask1"]]
    d22, d23 = tmp_path / "d22", tmp_path / "d23"
    anime_arrays(7, 1, 10, *dims, *args, d22, d23)  # anime.f95:26: not written
    assert not d22.exists() and not d23.exists()
    for n in (1, 10, 2000):
        anime_arrays(n, 1, 10, *dims, *args, d22, d23)
        assert d22.exists() and d23.exists()
        os.remove(d22)
        os.remove(d23)


def test_anime_files_equal_the_reference(tmp_path):
    from oracle.ref import run_ref
    if not run_ref.available():
        pytest.skip("reference sources absent")
    dims = (10, 8, 6)
    case, state = _anime_inputs(dims, 9)
    ref22, ref23 = run_ref.anime_files(case, common.REF_PHYS, state, 5)
    d22, d23 = tmp_path / "d22", tmp_path / "d23"
    anime_arrays(5, 5, 5, *dims, case["dx1"], case["dy1"], case["z2"], state["u"], state["v"],
                 state["w"], case["amask1"], d22, d23)
    assert d22.read_bytes() == ref22
    assert d23.read_bytes() == ref23


@pytest.mark.gpu
def test_netcdf_and_anime_from_a_handle(tmp_path):
    from les_halo_amd import LES
    from les_halo_amd.output import anime
    case = common.tiny_case(12, 10, 8, seed=4)
    ip, jp, kp = case["dims"]
    L = LES(ip, jp, kp, **common.REF_PHYS, omega=1.5, pjuge=0.0, nmaxp=20)
    L.load_case(case)
    L.ifdata()
    L.run(1, 3)
    arrs = {n: L.download(n) for n in NAMES}
    with NetCDFOutput(tmp_path, ip, jp, kp, ntimes=4) as nc:
        nc.write(L, 3)
    _check_netcdf(tmp_path, (ip, jp, kp), 4, 3, arrs)
    anime(L, 3, 1, 3, tmp_path / "h22", tmp_path / "h23")
    anime_arrays(3, 1, 3, ip, jp, kp, L.download("dx1"), L.download("dy1"), L.download("z2"),
                 arrs["u"], arrs["v"], arrs["w"], L.download("amask1"), tmp_path / "a22",
                 tmp_path / "a23")
    assert (tmp_path / "h22").read_bytes() == (tmp_path / "a22").read_bytes()
    assert (tmp_path / "h23").read_bytes() == (tmp_path / "a23").read_bytes()
    L.close()
