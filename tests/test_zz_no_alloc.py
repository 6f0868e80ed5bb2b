"""The allocation-free solve (blco_solve_noalloc) against blco_solve.

Both run the same kernels on the same inputs, so their results must agree
bit for bit.  Two calls in a row on one GPU share the library's scratch
array; the second is smaller (fewer rows, lower rank) and must not be
affected by what the first left there.  Every call goes through the C ABI
with ctypes."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

vp = C.c_void_p


def _noalloc(B):
    L = B.lib()
    f = L.blco_solve_noalloc
    f.restype = C.c_int
    f.argtypes = [C.POINTER(vp), C.c_int, C.c_int, C.c_int, vp, vp, C.c_int, C.c_uint64,
                  C.POINTER(C.c_int), vp]
    L.blco_solve_noalloc_scratch_bytes.restype = C.c_uint64
    L.blco_solve_noalloc_scratch_bytes.argtypes = []
    return f


def _case(B, rng, dims, R, mode, rows, dtype, singular):
    U = [rng.random((d, R)) for d in dims]
    if singular:
        for u in U:
            u[:, R - 1] = u[:, 0]
    grams = [u.T @ u for u in U]
    M = rng.random((rows, R))
    Gd = []
    for g in grams:
        Gd.append(B.DeviceArray.from_numpy(g, B.F64))
    mdt = B.F64 if dtype == B.F64 else B.F32
    Md = B.DeviceArray.from_numpy(M, mdt)
    return Gd, Md


def _run(B, fn, Gd, Md, R, mode, rows, dtype):
    Ud = B.DeviceArray((rows, R), dtype)
    fb = C.c_int(-1)
    ptrs = (vp * len(Gd))(*[vp(g.data_ptr()) for g in Gd])
    rc = fn(ptrs, len(Gd), mode, R, vp(Md.data_ptr()), vp(Ud.data_ptr()), dtype, rows, C.byref(fb), None)
    if rc != 0:
        raise AssertionError(B.last_error())
    B.synchronize()
    out = Ud.numpy()
    Ud.free()
    return out, fb.value


@pytest.mark.parametrize("dtype_name", ["F64", "F32"])
def test_noalloc_solve_matches_solve_twice(B, gpu, dtype_name):
    dtype = getattr(B, dtype_name)
    noalloc = _noalloc(B)
    assert B.lib().blco_solve_noalloc_scratch_bytes() >= 8 * (3 * 64 * 64 + 64 + 2048 * 64)
    rng = np.random.default_rng(2027)
    # first call: 70000 rows at rank 64 (the row solve's full 2048 workgroups),
    # then a smaller one: 333 rows at rank 12 with a singular V (pseudo-inverse)
    cases = [((400, 300, 200), 64, 1, 70000, False), ((90, 80, 70), 12, 0, 333, True)]
    for dims, R, mode, rows, singular in cases:
        Gd, Md = _case(B, rng, dims, R, mode, rows, dtype, singular)
        want, fb_want = _run(B, B.lib().blco_solve, Gd, Md, R, mode, rows, dtype)
        got, fb_got = _run(B, noalloc, Gd, Md, R, mode, rows, dtype)
        assert fb_got == fb_want == (1 if singular else 0)
        assert np.array_equal(got, want), (dims, R, rows, np.abs(got - want).max())
        for a in Gd + [Md]:
            a.free()


def test_noalloc_solve_rejects_rank_beyond_scratch(B, gpu):
    noalloc = _noalloc(B)
    ptrs = (vp * 3)(None, None, None)
    rc = noalloc(ptrs, 3, 0, 65, vp(1), vp(1), B.F64, 10, None, None)
    assert rc != 0 and "rank 65" in B.last_error()
