"""The adjoint form's statement list (les_halo_amd/adjgen.py), checked on the CPU before any of
its kernels runs: every element a statement reads or writes lies inside its array's allocation
(the flat Fortran offset, as the generated kernels compute it), and no point of a statement's
box reads an element of the target array that another point of the box writes - the property
that lets a statement's points run in parallel and its adjoint scatter with atomics."""
import itertools

import numpy as np
import pytest

from les_halo_amd import adjgen


def _ev(expr, **env):
    return int(eval(expr, {}, env))


def _box(s, im, jm, km):
    env = dict(im=im, jm=jm, km=km)
    (klo, khi), (jlo, jhi), (ilo, ihi) = s.rk, s.rj, s.ri
    r = [range(_ev(lo, **env), _ev(hi, **env) + 1) for lo, hi in ((ilo, ihi), (jlo, jhi), (klo, khi))]
    for i, j, k in itertools.product(*r):
        if s.colour is None or ((i + j + k + s.colour) & 1) == 1:
            yield i, j, k


def _refs(e, acc):
    if isinstance(e, adjgen.Ref):
        acc.append(e)
    elif isinstance(e, adjgen.Op):
        _refs(e.a, acc)
        _refs(e.b, acc)
    elif isinstance(e, (adjgen.Neg, adjgen.Sqrt)):
        _refs(e.a, acc)
    return acc


def _size(arr, im, jm, km):
    return _ev(adjgen.size_expr(arr).replace("(long)", ""), im=im, jm=jm, km=km)


def _flat(arr, idx, i, j, k, im, jm, km):
    return _ev(adjgen.flat(arr, idx).replace("(long)", ""), i=i, j=j, k=k, im=im, jm=jm, km=km)


STMTS = [s for s in adjgen.build_step() if isinstance(s, adjgen.Stmt)] + adjgen.build_sor()


@pytest.mark.parametrize("dims", [(6, 4, 5), (9, 8, 6)])
def test_statements_stay_in_bounds_and_conflict_free(dims):
    im, jm, km = dims
    for s in STMTS:
        refs = _refs(s.expr, [])
        written = set()
        reads = []
        for i, j, k in _box(s, im, jm, km):
            t = _flat(s.target.arr, s.target.idx, i, j, k, im, jm, km)
            assert 0 <= t < _size(s.target.arr, im, jm, km), (s.name, i, j, k)
            written.add(t)
            for r in refs:
                q = _flat(r.arr, r.idx, i, j, k, im, jm, km)
                assert 0 <= q < _size(r.arr, im, jm, km), (s.name, r.arr, r.idx, i, j, k)
                if r.arr == s.target.arr and q != t:
                    reads.append(q)
                if r.arr == s.target.arr and q == t:
                    # the adjoint recognises the element's own value by its index text
                    assert r.idx == s.target.idx, (s.name, r.idx)
        assert not written.intersection(reads), s.name


def test_generated_source_is_current():
    import os
    with open(adjgen.OUT) as fh:
        assert fh.read() == adjgen.generate(), "run python -m les_halo_amd.adjgen"
