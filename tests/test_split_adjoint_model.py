"""The transpose of the colour-split SOR solve (k_split_gather_t, k_split_sweep_t,
k_split_scatter_t in les_halo_amd/csrc/les_sor.hip, driven by sor_adjoint_external), checked
on the CPU by a dot-product test of float64 models of both directions.

Forward (press.f95:87-115 as the split solve runs it): the interior of p is the iterate; each
iteration sweeps colour 1 then colour 0, a cell reading the other colour's neighbours and, across
i and k faces, its own value (the mirror image); afterwards every box cell of
[0,ip+1]x[0,jp+1]x[0,kp+1] takes its mirror interior cell's value (the gather).  The transposed
model applies, in the kernels' order, the gather's transpose, the red-black adjoint sweeps (the
first gathers only, the last scales only) and the scatter's transpose.  <ybar, F(x+dx) - F(x)>
must equal <F^T ybar, dx> (F is linear in p and rhs).  Grids with odd and even ip and the
j seam (even jp) are covered.
"""
import numpy as np
import pytest


def _coeffs(kp, seed):
    rng = np.random.default_rng(seed)
    return dict(om=1.3, c2l=0.7, c2s=0.7, c3l=0.4, c3s=0.4,
                c4l=rng.uniform(0.2, 0.6, kp), c4s=rng.uniform(0.2, 0.6, kp),
                cn1=rng.uniform(0.3, 0.5, kp + 2))  # cn1[k], k = 1..kp


def _forward(p, r, dims, C, nit):
    ip, jp, kp = dims
    x = p[1:ip + 1, 1:jp + 1, 1:kp + 1].copy()

    def val(i, j, k, own):
        if i < 1 or i > ip or k < 1 or k > kp:
            return own  # mirror image across an i or k face
        return x[i - 1, (j - 1) % jp, k - 1]  # periodic j

    for _ in range(nit):
        for c in (1, 0):
            new = x.copy()
            for i in range(1, ip + 1):
                for j in range(1, jp + 1):
                    for k in range(1, kp + 1):
                        if (i + j + k) & 1 != c:
                            continue
                        pc = x[i - 1, j - 1, k - 1]
                        s = (C["c2l"] * val(i + 1, j, k, pc) + C["c2s"] * val(i - 1, j, k, pc) +
                             C["c3l"] * val(i, j + 1, k, pc) + C["c3s"] * val(i, j - 1, k, pc) +
                             C["c4l"][k - 1] * val(i, j, k + 1, pc) +
                             C["c4s"][k - 1] * val(i, j, k - 1, pc) - r[i, j, k])
                        new[i - 1, j - 1, k - 1] = pc + C["om"] * (C["cn1"][k] * s - pc)
            x = new
    out = np.zeros_like(p)
    for a in range(ip + 2):
        for b in range(jp + 2):
            for c in range(kp + 2):
                i = 1 if a == 0 else (ip if a == ip + 1 else a)
                j = jp if b == 0 else (1 if b == jp + 1 else b)
                k = 1 if c == 0 else (kp if c == kp + 1 else c)
                out[a, b, c] = x[i - 1, j - 1, k - 1]
    return out


def _transpose(pb, dims, C, nit):
    ip, jp, kp = dims
    pb = pb.copy()
    rb = np.zeros_like(pb)
    X = np.zeros((ip, jp, kp))
    R = np.zeros((ip, jp, kp))
    for i in range(1, ip + 1):  # k_split_gather_t
        for j in range(1, jp + 1):
            for k in range(1, kp + 1):
                iss = [i] + ([0] if i == 1 else ([ip + 1] if i == ip else []))
                jss = [j] + ([jp + 1] if j == 1 else ([0] if j == jp else []))
                kss = [k] + ([0] if k == 1 else ([kp + 1] if k == kp else []))
                v = 0.0
                for c in kss:
                    for b in jss:
                        for a in iss:
                            v += pb[a, b, c]
                            pb[a, b, c] = 0.0
                X[i - 1, j - 1, k - 1] = v
    om, cn1 = C["om"], C["cn1"]

    def sweep(c, mode):  # k_split_sweep_t
        for i in range(1, ip + 1):
            for j in range(1, jp + 1):
                for k in range(1, kp + 1):
                    if (i + j + k) & 1 != c:
                        continue
                    own = X[i - 1, j - 1, k - 1]
                    out = own
                    if mode != 0:
                        img = ((C["c2s"] if i == 1 else 0.0) + (C["c2l"] if i == ip else 0.0) +
                               (C["c4s"][k - 1] if k == 1 else 0.0) +
                               (C["c4l"][k - 1] if k == kp else 0.0))
                        R[i - 1, j - 1, k - 1] -= om * cn1[k] * own
                        out = ((1.0 - om) + om * cn1[k] * img) * own
                    if mode != 2:
                        jn, js = (1 if j == jp else j + 1), (jp if j == 1 else j - 1)
                        g = 0.0
                        if i > 1:
                            g += C["c2l"] * X[i - 2, j - 1, k - 1]
                        if i < ip:
                            g += C["c2s"] * X[i, j - 1, k - 1]
                        g += C["c3l"] * X[i - 1, js - 1, k - 1] + C["c3s"] * X[i - 1, jn - 1, k - 1]
                        v = om * cn1[k] * g
                        if k > 1:
                            v += om * cn1[k - 1] * C["c4l"][k - 2] * X[i - 1, j - 1, k - 2]
                        if k < kp:
                            v += om * cn1[k + 1] * C["c4s"][k] * X[i - 1, j - 1, k]
                        out += v
                    X[i - 1, j - 1, k - 1] = out

    sweep(1, 0)  # sor_adjoint_external's order
    for it in range(nit, 0, -1):
        sweep(0, 1)
        sweep(1, 1 if it > 1 else 2)
    pb[1:ip + 1, 1:jp + 1, 1:kp + 1] += X  # k_split_scatter_t
    rb[1:ip + 1, 1:jp + 1, 1:kp + 1] += R
    return pb, rb


@pytest.mark.parametrize("dims,nit", [((7, 4, 3), 3), ((6, 6, 4), 2), ((5, 2, 2), 4)])
def test_split_sor_transpose_dot_product(dims, nit):
    C = _coeffs(dims[2], 5)
    rng = np.random.default_rng(1)
    shape = tuple(n + 2 for n in dims)
    p0, r0 = rng.standard_normal(shape), rng.standard_normal(shape)
    dp, dr, yb = rng.standard_normal(shape), rng.standard_normal(shape), rng.standard_normal(shape)
    lhs = np.sum(yb * (_forward(p0 + dp, r0 + dr, dims, C, nit) - _forward(p0, r0, dims, C, nit)))
    pb, rb = _transpose(yb, dims, C, nit)
    rhs = np.sum(pb * dp) + np.sum(rb * dr)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (lhs, rhs)


@pytest.mark.parametrize("mistake", ["c2s", "c4l"])
def test_split_sor_transpose_check_is_sharp(mistake):
    """A transpose with one coefficient dropped (planted in the model, the forward unchanged)
    fails the same check."""
    dims, nit = (5, 4, 3), 2
    C = _coeffs(dims[2], 5)
    rng = np.random.default_rng(2)
    shape = tuple(n + 2 for n in dims)
    p0, r0, dp, yb = (rng.standard_normal(shape) for _ in range(4))
    lhs = np.sum(yb * (_forward(p0 + dp, r0, dims, C, nit) - _forward(p0, r0, dims, C, nit)))
    bad = dict(C)
    bad[mistake] = 0.0 * C[mistake]
    pb, _ = _transpose(yb, dims, bad, nit)
    assert abs(lhs - np.sum(pb * dp)) > 1e-6 * abs(lhs)
