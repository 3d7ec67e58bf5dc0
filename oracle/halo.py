"""CPU oracle of the halo-exchange pack/unpack (test infrastructure only: imported by tests/,
never by the product library).

Restates src/OpenCL/Kernels/halo_exchange.cl (write :3-51, read :156-185) and the inner-ring
read of halo_exchange_correct.cl (:156-185) for an (im, jm, km) array per vector component,
with the components stored as separate arrays (the reference's float2/float4/float16 element
components become v = 0..v_dim-1).  Buffer block v starts at v * v_sz and holds
north | south | west | east, each ordered along the row/column first, then w, then k.

Pinned by tests/test_halo.py against the output of the reference's own C test programs
(src/OpenCL/test_halo_{read,write}{,_correct}.c) and by the read(write(b)) == b identity.
"""
import numpy as np


def halo_size(v_dim, im, jm, km, h_w, correct=0):
    """Buffer length (floats): halo_exchange.cl:15 / halo_exchange_correct.cl:168 times v_dim."""
    if correct:
        return v_dim * 2 * km * (im + jm - 4 * h_w) * h_w
    return v_dim * (2 * km * im * h_w + 2 * km * (jm - 2 * h_w) * h_w)


def _plan(im, jm, km, h_w, correct):
    """(array index, buffer index) pairs of one component, in the reference's formulas."""
    a_idx, b_idx = [], []
    if not correct:  # halo_exchange.cl:156-185 (read) == :3-51 (write), same index pairs
        for k in range(km):
            for w in range(h_w):
                for i in range(im):  # north, south rows (corners included)
                    a_idx.append(i + im * w + k * im * jm)
                    b_idx.append(i + im * w + k * im * h_w)
                    a_idx.append(i + im * w + k * im * jm + im * (jm - h_w))
                    b_idx.append(km * im * h_w + i + im * w + k * im * h_w)
                for i in range(h_w, jm - h_w):  # west, east columns
                    a_idx.append(i * im + w + k * im * jm)
                    b_idx.append(km * im * h_w * 2 + (i - h_w) * h_w + w + k * (jm - 2 * h_w) * h_w)
                    a_idx.append(i * im + w + k * im * jm + im - h_w)
                    b_idx.append(km * im * h_w * 2 + km * (jm - 2 * h_w) * h_w + (i - h_w) * h_w + w
                                 + k * (jm - 2 * h_w) * h_w)
    else:  # halo_exchange_correct.cl:156-185: inner ring, corners excluded
        nsw = (im - 2 * h_w)
        for k in range(km):
            for w in range(h_w):
                for i in range(h_w, im - h_w):
                    a_idx.append(im * h_w + i + im * w + k * im * jm)
                    b_idx.append(i - h_w + nsw * w + k * nsw * h_w)
                    a_idx.append(i + k * im * jm + im * (jm - 2 * h_w) + im * w)
                    b_idx.append(km * nsw * h_w + i - h_w + nsw * w + k * nsw * h_w)
                for i in range(h_w, jm - h_w):
                    a_idx.append(i * im + w + h_w + k * im * jm)
                    b_idx.append(km * nsw * h_w * 2 + (i - h_w) * h_w + w + k * (jm - 2 * h_w) * h_w)
                    a_idx.append(i * im + w + k * im * jm + im - 2 * h_w)
                    b_idx.append(km * nsw * h_w * 2 + km * (jm - 2 * h_w) * h_w + (i - h_w) * h_w + w
                                 + k * (jm - 2 * h_w) * h_w)
    return np.array(a_idx, np.int64), np.array(b_idx, np.int64)


def halo_read(arrays, im, jm, km, h_w, correct=0, buffer=None):
    """Pack: buffer[v*v_sz + b] = arrays[v][a] for every index pair, in the reference's order
    (a later pair overwrites an earlier one at the same buffer position)."""
    v_dim = len(arrays)
    v_sz = halo_size(1, im, jm, km, h_w, correct)
    buf = np.zeros(v_dim * v_sz, np.float32) if buffer is None else np.array(buffer, np.float32)
    a, b = _plan(im, jm, km, h_w, correct)
    for v, arr in enumerate(arrays):
        flat = np.asarray(arr, np.float32).reshape(-1)
        buf[v * v_sz + b] = flat[a]
    return buf


def halo_write(arrays, buffer, im, jm, km, h_w):
    """Unpack (halo_exchange.cl:3-51): arrays[v][a] = buffer[v*v_sz + b]; returns new arrays."""
    v_sz = halo_size(1, im, jm, km, h_w, 0)
    a, b = _plan(im, jm, km, h_w, 0)
    out = []
    buf = np.asarray(buffer, np.float32)
    for v, arr in enumerate(arrays):
        flat = np.array(arr, np.float32).reshape(-1)
        flat[a] = buf[v * v_sz + b]
        out.append(flat)
    return out
