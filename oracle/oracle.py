"""ctypes view of the CPU oracle (oracle/liboracle.so) and of the compiled
reference harness (oracle/_ref/libaltoref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, never by the product package.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def lib():
    """Load liboracle.so (built by `make -C oracle`)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(path)
        L.orc_pdep64.restype = C.c_uint64
        L.orc_pdep64.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_pext64.restype = C.c_uint64
        L.orc_pext64.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_pdep128.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64,
                                  C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.orc_pext128.restype = C.c_uint64
        L.orc_pext128.argtypes = [C.c_uint64] * 4
        L.orc_alto_setup.argtypes = [C.c_int, u64p, C.c_int, u64p, u64p, i32p, i32p]
        L.orc_alto_encode.argtypes = [C.c_int, u64p, C.c_uint64, u64p, f64p, C.c_int,
                                      u64p, u64p, f64p, u64p]
        L.orc_blco_encode.argtypes = [C.c_int, u64p, C.c_uint64, u64p, f64p, C.c_uint64,
                                      C.c_int, u64p, f64p, u64p, i32p,
                                      C.POINTER(C.c_uint64), u64p, u64p, u64p, C.c_uint64]
        L.orc_mttkrp_coo.argtypes = [C.c_int, C.c_uint64, u64p, f64p, C.c_int,
                                     C.POINTER(C.c_void_p), C.c_int, f64p]
        L.orc_alto_create.restype = C.c_void_p
        L.orc_alto_create.argtypes = [C.c_int, u64p, C.c_uint64, u64p, f64p, C.c_int]
        L.orc_alto_free.argtypes = [C.c_void_p]
        L.orc_alto_mttkrp.argtypes = [C.c_void_p, C.c_int, C.c_int,
                                      C.POINTER(C.c_void_p), f64p]
        L.orc_cpd_alto.restype = C.c_int
        L.orc_cpd_alto.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_void_p), f64p,
                                   C.c_int, C.c_double, f64p, f64p, f64p, C.c_int]
        L.orc_cpd_alto_ex.restype = C.c_int
        L.orc_cpd_alto_ex.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_void_p), f64p,
                                      C.c_int, C.c_double, f64p, f64p, f64p, C.c_int, C.c_double]
        L.orc_cholesky.restype = C.c_int
        L.orc_cholesky.argtypes = [C.c_int, f64p, f64p]
        L.orc_cholesky_solve_rows.argtypes = [C.c_int, f64p, C.c_uint64, f64p]
        L.orc_pinv_sym.argtypes = [C.c_int, f64p, f64p]
        L.orc_kruskal_norm.argtypes = [C.c_uint64, C.c_int, f64p, f64p, C.c_int]
        L.orc_cpd_fit.restype = C.c_double
        L.orc_cpd_fit.argtypes = [C.c_int, C.c_uint64, C.c_int, f64p, f64p, f64p,
                                  C.POINTER(C.c_void_p), C.c_double]
        _LIB = L
    return _LIB


def ref_lib():
    """Load the compiled reference harness, or return None when absent."""
    global _REF
    if _REF is None:
        path = os.path.join(HERE, "_ref", "libaltoref.so")
        if not os.path.exists(path):
            return None
        R = C.CDLL(path)
        R.ref_alto_setup.argtypes = [C.c_int, u64p, C.c_int, u64p, u64p, i32p]
        R.ref_gen_alto.argtypes = [C.c_int, u64p, C.c_uint64, u64p, f64p, C.c_int,
                                   u64p, u64p, f64p]
        R.ref_mttkrp_alto_par.argtypes = [C.c_int, u64p, C.c_uint64, u64p, f64p, C.c_int,
                                          C.c_int, C.POINTER(C.c_void_p), C.c_int, f64p]
        R.ref_kruskal_norm.argtypes = [C.c_uint64, C.c_int, f64p, f64p, C.c_int]
        R.ref_mttkrp_alto_par_time.argtypes = [C.c_int, u64p, C.c_uint64, u64p, f64p, C.c_int, C.c_int,
                                               C.POINTER(C.c_void_p), C.c_int, f64p, C.c_int]
        R.ref_mttkrp_alto_par_time.restype = C.c_double
        _REF = R
    return _REF


def _ptrs(arrs):
    return (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])


def _cidx(cidx):
    return np.ascontiguousarray(cidx, dtype=np.uint64)


def alto_setup(dims, lit_bits=64):
    dims = np.asarray(dims, dtype=np.uint64)
    n = len(dims)
    lo = np.zeros(n, np.uint64); hi = np.zeros(n, np.uint64)
    pos = np.zeros(n, np.int32); bits = np.zeros(n, np.int32)
    rc = lib().orc_alto_setup(n, dims, lit_bits, lo, hi, pos, bits)
    if rc != 0:
        raise ValueError("key needs more than %d bits" % lit_bits)
    return lo, hi, pos, bits


def alto_encode(dims, cidx, vals, lit_bits=64):
    dims = np.asarray(dims, dtype=np.uint64)
    cidx = _cidx(cidx)
    nnz = cidx.shape[1]
    klo = np.zeros(nnz, np.uint64); khi = np.zeros(nnz, np.uint64)
    vo = np.zeros(nnz, np.float64); perm = np.zeros(nnz, np.uint64)
    rc = lib().orc_alto_encode(len(dims), dims, nnz, cidx.ravel(),
                               np.ascontiguousarray(vals, np.float64), lit_bits,
                               klo, khi, vo, perm)
    if rc != 0:
        raise ValueError("alto_encode failed: %d" % rc)
    return klo, khi, vo, perm


def blco_encode(dims, cidx, vals, max_block_size=0, lit_bits=128):
    """Returns dict(coords, vals, masks, pos, block_start, block_nnz, block_coords)."""
    dims = np.asarray(dims, dtype=np.uint64)
    cidx = _cidx(cidx)
    n, nnz = cidx.shape
    coords = np.zeros(nnz, np.uint64); vo = np.zeros(nnz, np.float64)
    masks = np.zeros(n, np.uint64); pos = np.zeros(n, np.int32)
    cap = max(16, (nnz // max_block_size + 8) if max_block_size else 16) * 4
    while True:
        bs = np.zeros(cap, np.uint64); bn = np.zeros(cap, np.uint64)
        bc = np.zeros(cap * n, np.uint64); nb = C.c_uint64(0)
        rc = lib().orc_blco_encode(n, dims, nnz, cidx.ravel(),
                                   np.ascontiguousarray(vals, np.float64),
                                   max_block_size, lit_bits, coords, vo, masks, pos,
                                   C.byref(nb), bs, bn, bc, cap)
        if rc == -2:
            cap = nb.value + 1
            continue
        if rc != 0:
            raise ValueError("blco_encode failed: %d" % rc)
        break
    k = nb.value
    return dict(coords=coords, vals=vo, masks=masks, pos=pos,
                block_start=bs[:k].copy(), block_nnz=bn[:k].copy(),
                block_coords=bc[:k * n].reshape(k, n).copy())


def mttkrp_coo(cidx, vals, U, mode, out=None):
    cidx = _cidx(cidx)
    n, nnz = cidx.shape
    R = U[0].shape[1]
    Uc = [np.ascontiguousarray(u, np.float64) for u in U]
    if out is None:
        out = np.zeros((Uc[mode].shape[0], R), np.float64)
    lib().orc_mttkrp_coo(n, nnz, cidx.ravel(), np.ascontiguousarray(vals, np.float64), R,
                         _ptrs(Uc), mode, out.reshape(-1))
    return out


class Alto:
    """Oracle ALTO tensor (create_alto) for CPU MTTKRP and CP-ALS."""

    def __init__(self, dims, cidx, vals, nprtn=None):
        self.dims = np.asarray(dims, dtype=np.uint64)
        cidx = _cidx(cidx)
        self.nmodes = cidx.shape[0]
        if nprtn is None:
            nprtn = os.cpu_count() or 1
        self.h = lib().orc_alto_create(self.nmodes, self.dims, cidx.shape[1], cidx.ravel(),
                                       np.ascontiguousarray(vals, np.float64), nprtn)
        if not self.h:
            raise ValueError("orc_alto_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_alto_free(self.h)
            self.h = None

    def mttkrp(self, U, mode, out=None):
        R = U[0].shape[1]
        Uc = [np.ascontiguousarray(u, np.float64) for u in U]
        if out is None:
            out = np.zeros((int(self.dims[mode]), R), np.float64)
        rc = lib().orc_alto_mttkrp(self.h, mode, R, _ptrs(Uc), out.reshape(-1))
        if rc != 0:
            raise ValueError("orc_alto_mttkrp failed")
        return out

    def cpd(self, U, max_iters, epsilon=0.0, verbose=False, pivot_eps=0.0):
        """Runs CP-ALS in place on copies of U; returns (U, lambda, fits, times, mttkrp_times).
        pivot_eps > 0 applies the GPU library's pivot floor R * pivot_eps * max V_ii
        and pseudo-inverse tolerance; 0 is the reference's solve rule."""
        R = U[0].shape[1]
        Uc = [np.array(u, dtype=np.float64, order="C", copy=True) for u in U]
        lam = np.zeros(R, np.float64)
        fits = np.zeros(max(max_iters, 1), np.float64)
        its = np.zeros(max(max_iters, 1), np.float64)
        mts = np.zeros(max(max_iters, 1), np.float64)
        n = lib().orc_cpd_alto_ex(self.h, R, _ptrs(Uc), lam, max_iters, epsilon, fits, its,
                                  mts, 1 if verbose else 0, float(pivot_eps))
        return Uc, lam, fits[:n].copy(), its[:n].copy(), mts[:n].copy()


def cholesky(V):
    V = np.ascontiguousarray(V, np.float64)
    R = V.shape[0]
    L = np.zeros_like(V)
    info = lib().orc_cholesky(R, V.ravel(), L.ravel())
    return info, L


def cholesky_solve_rows(L, B):
    L = np.ascontiguousarray(L, np.float64)
    B = np.array(B, dtype=np.float64, order="C", copy=True)
    lib().orc_cholesky_solve_rows(L.shape[0], L.ravel(), B.shape[0], B.reshape(-1))
    return B


def pinv_sym(V):
    V = np.ascontiguousarray(V, np.float64)
    P = np.zeros_like(V)
    lib().orc_pinv_sym(V.shape[0], V.ravel(), P.ravel())
    return P


def kruskal_norm(U, which):
    U = np.array(U, dtype=np.float64, order="C", copy=True)
    lam = np.zeros(U.shape[1], np.float64)
    lib().orc_kruskal_norm(U.shape[0], U.shape[1], U.reshape(-1), lam, which)
    return U, lam


def cpd_fit(U_last, mttkrp_last, lam, grams, normX2):
    grams = [np.ascontiguousarray(g, np.float64) for g in grams]
    return lib().orc_cpd_fit(len(grams), U_last.shape[0], U_last.shape[1],
                             np.ascontiguousarray(U_last, np.float64).ravel(),
                             np.ascontiguousarray(mttkrp_last, np.float64).ravel(),
                             np.ascontiguousarray(lam, np.float64), _ptrs(grams), normX2)


# ---- compiled reference (oracle/_ref), when present ----

def ref_alto_setup(dims, lit_bits=64):
    R = ref_lib()
    dims = np.asarray(dims, dtype=np.uint64)
    n = len(dims)
    lo = np.zeros(n, np.uint64); hi = np.zeros(n, np.uint64); pos = np.zeros(n, np.int32)
    R.ref_alto_setup(n, dims, lit_bits, lo, hi, pos)
    return lo, hi, pos


def ref_gen_alto(dims, cidx, vals, lit_bits=64):
    R = ref_lib()
    dims = np.asarray(dims, dtype=np.uint64)
    cidx = _cidx(cidx)
    nnz = cidx.shape[1]
    klo = np.zeros(nnz, np.uint64); khi = np.zeros(nnz, np.uint64); vo = np.zeros(nnz)
    R.ref_gen_alto(len(dims), dims, nnz, cidx.ravel(), np.ascontiguousarray(vals, np.float64),
                   lit_bits, klo, khi, vo)
    return klo, khi, vo


def ref_mttkrp(dims, cidx, vals, U, mode, nprtn):
    R = ref_lib()
    dims = np.asarray(dims, dtype=np.uint64)
    cidx = _cidx(cidx)
    rank = U[0].shape[1]
    Uc = [np.ascontiguousarray(u, np.float64) for u in U]
    out = np.zeros((int(dims[mode]), rank), np.float64)
    R.ref_mttkrp_alto_par(len(dims), dims, cidx.shape[1], cidx.ravel(),
                          np.ascontiguousarray(vals, np.float64), nprtn, rank,
                          _ptrs(Uc), mode, out.reshape(-1))
    return out


def ref_mttkrp_time(dims, cidx, vals, U, mode, nprtn, reps=2):
    """Mean seconds of the reference's own mttkrp_alto_par (compiled from its
    sources in oracle/_ref) on this tensor, construction excluded."""
    R = ref_lib()
    dims = np.ascontiguousarray(dims, np.uint64)
    cidx = np.ascontiguousarray(cidx, np.uint64)
    Uc = [np.ascontiguousarray(u, np.float64) for u in U]
    out = np.zeros((int(dims[mode]), Uc[0].shape[1]), np.float64)
    ptrs = (C.c_void_p * len(Uc))(*[u.ctypes.data for u in Uc])
    return R.ref_mttkrp_alto_par_time(len(dims), dims, cidx.shape[1], cidx.ravel(),
                                      np.ascontiguousarray(vals, np.float64), nprtn, Uc[0].shape[1], ptrs,
                                      mode, out.ravel(), reps)


def ref_kruskal_norm(U, which):
    R = ref_lib()
    U = np.array(U, dtype=np.float64, order="C", copy=True)
    lam = np.zeros(U.shape[1], np.float64)
    R.ref_kruskal_norm(U.shape[0], U.shape[1], U.reshape(-1), lam, which)
    return U, lam
