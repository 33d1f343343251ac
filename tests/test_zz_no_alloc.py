"""The encounter check that allocates nothing
(swg_encounter_check_all_plpl_noalloc), through the C ABI, against the
allocating entry points it stands for (swg_encounter_check_all_plpl followed
by swg_encounter_list_copy): identical lists on two calls in a row, the second
with fewer bodies than the first, and the stated limits rejected with an
error."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import swiftest_amd as S
from swiftest_amd import core

from helpers import compact_disk

pytestmark = pytest.mark.gpu

MAX_NPL = 2048
SWG_ERR_ARG = -1


def _lib():
    L = core.library()
    vp = C.c_void_p
    L.swg_encounter_check_all_plpl_noalloc.argtypes = [vp, C.c_int32, vp, vp, vp, C.c_double, C.c_int64,
                                                        C.POINTER(C.c_int64), vp, vp, vp]
    return L


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _state(npl, seed):
    d = compact_disk(seed=seed, npl=npl)
    renc = np.ascontiguousarray(6.5 * 0.01 * np.ones(npl) * (1 + 0.3 * np.sin(np.arange(npl))))
    return np.ascontiguousarray(d.pl_rh), np.ascontiguousarray(d.pl_vh), renc


def _noalloc(r, v, renc, dt, cap):
    L = _lib()
    n = C.c_int64(-1)
    a, b, c = (np.full(max(cap, 1), -7, np.int32) for _ in range(3))
    rc = L.swg_encounter_check_all_plpl_noalloc(core.context(), r.shape[0], _p(r), _p(v), _p(renc), dt, cap,
                                                C.byref(n), _p(a), _p(b), _p(c))
    return rc, n.value, a[:max(n.value, 0)], b[:max(n.value, 0)], c[:max(n.value, 0)]


@pytest.mark.parametrize("sizes", [(1900, 700), (MAX_NPL, 64)])
def test_noalloc_encounter_list_equals_allocating_path(sizes):
    for k, npl in enumerate(sizes):
        r, v, renc = _state(npl, seed=40 + k)
        dt = 0.05
        i1, i2, lv = S.encounter_check_all_plpl(npl, r, v, renc, dt)
        assert len(i1) > 0
        rc, n, a, b, c = _noalloc(r, v, renc, dt, cap=len(i1) + 5)
        assert rc == 0, core.library().swg_last_error().decode()
        assert n == len(i1)
        np.testing.assert_array_equal(a, i1)
        np.testing.assert_array_equal(b, i2)
        np.testing.assert_array_equal(c, lv)


def test_noalloc_limits_are_rejected():
    r, v, renc = _state(MAX_NPL + 1, seed=45)
    rc, n, *_ = _noalloc(r, v, renc, 0.05, cap=10)
    assert rc == SWG_ERR_ARG
    assert "exceeds the no-allocation limit" in core.library().swg_last_error().decode()
    # a list longer than the caller's arrays: its length is reported, nothing copied
    r, v, renc = _state(500, seed=46)
    i1, _, _ = S.encounter_check_all_plpl(500, r, v, renc, 0.05)
    assert len(i1) > 3
    rc, n, *_ = _noalloc(r, v, renc, 0.05, cap=3)
    assert rc == SWG_ERR_ARG and n == len(i1)
    # and the next call in range still works
    rc, n, a, b, c = _noalloc(r, v, renc, 0.05, cap=len(i1))
    assert rc == 0 and n == len(i1)
    np.testing.assert_array_equal(a, i1)
