"""Python view of the BLCO CP-ALS library (lib/libblco.so) for tests and
benchmarks.  Everything here calls the C ABI declared in include/blco_host.h
and include/blco_hip.h; there is no Python or CPU fallback for any GPU entry
point: when the library or the GPU is missing, calls raise.

Names follow the reference C++ API (blco.hpp, cpd_gpu.hpp, kruskal_model.hpp):
  gen_blcotensor_host      -> encode()          (include/blco.hpp:196-331)
  gen_blcotensor_device    -> DeviceTensor      (src/blco.cu:12-58)
  mttkrp_alto_dev_onemode  -> mttkrp()          (src/alto_dev.cu:1954-2032)
  cpals_blco_dev           -> cpals(), CpAls    (src/cpd_gpu.cu:21-242)
  ata_gpu / pseudoinverse_gpu / columnwise_normscal / kruskal_fit_gpu
                           -> gram(), solve(), colnorm(), fit()
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libblco.so")

F64, F32, BF16 = 0, 1, 2
K_LVL1, K_LVL2, K_LVL3, K_AUTO, K_BATCHED, K_BATCHED2, K_ATOMIC, K_SORTED, K_SELECT = \
    1, 2, 3, 10, 11, 12, 13, 14, 20
MAX_MODES = 5
SEL_NFEAT, SEL_NCAND = 24, 15

_lib = None

u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
vp = C.c_void_p


class BlcoError(RuntimeError):
    pass


class GenSpec(C.Structure):
    _fields_ = [("nmodes", C.c_int), ("dims", C.c_uint64 * MAX_MODES),
                ("zipf_s", C.c_double * MAX_MODES), ("target_nnz", C.c_uint64),
                ("topup", C.c_int), ("gen_rank", C.c_int), ("value_kind", C.c_int),
                ("seed", C.c_uint64)]


class CpalsOpts(C.Structure):
    _fields_ = [("rank", C.c_int), ("max_iters", C.c_int), ("tolerance", C.c_double),
                ("kernel", C.c_int), ("precision", C.c_int), ("nprtn", C.c_int),
                ("thread_cf", C.c_int), ("mode_order", C.c_int), ("use_graph", C.c_int),
                ("verbose", C.c_int), ("solve_floor", C.c_int)]


def build():
    """Compiles lib/libblco.so (hipcc for gfx950 + g++)."""
    import subprocess
    subprocess.run(["make", "-C", HERE, "-j8"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise BlcoError("libblco.so is not built (run `make -C blocked-linearized-coordinate_amd`)")
        L = C.CDLL(LIB_PATH)
        sig = {
            "blco_last_error": (C.c_char_p, []),
            "blco_coo_create": (C.c_int, [C.c_int, u64p, C.c_uint64, u64p, f64p, C.POINTER(vp)]),
            "blco_coo_import": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(vp)]),
            "blco_coo_export": (C.c_int, [vp, C.c_char_p, C.c_int]),
            "blco_coo_info": (C.c_int, [vp, C.POINTER(C.c_int), u64p, C.POINTER(C.c_uint64)]),
            "blco_coo_cidx": (vp, [vp]),
            "blco_coo_vals": (vp, [vp]),
            "blco_coo_free": (None, [vp]),
            "blco_generate": (C.c_int, [C.POINTER(GenSpec), C.POINTER(vp)]),
            "blco_init_factor": (C.c_int, [C.c_uint64, C.c_int, C.c_uint64, C.c_int, f64p]),
            "blco_host_encode": (C.c_int, [C.c_int, u64p, C.c_uint64, u64p, f64p, C.c_uint64,
                                           C.POINTER(vp)]),
            "blco_host_encode_coo": (C.c_int, [vp, C.c_uint64, C.POINTER(vp)]),
            "blco_device_encode": (C.c_int, [C.c_int, u64p, C.c_uint64, u64p, f64p, C.c_uint64,
                                             C.POINTER(vp)]),
            "blco_host_info": (C.c_int, [vp, C.POINTER(C.c_int), u64p, C.POINTER(C.c_uint64),
                                         C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
            "blco_host_layout": (C.c_int, [vp, u64p, i32p, u64p, u64p]),
            "blco_host_blocks": (C.c_int, [vp, u64p, u64p, u64p]),
            "blco_host_coords": (vp, [vp]),
            "blco_host_values": (vp, [vp]),
            "blco_host_free": (None, [vp]),
            "blco_device_count": (C.c_int, []),
            "blco_malloc": (C.c_int, [C.POINTER(vp), C.c_uint64]),
            "blco_free": (C.c_int, [vp]),
            "blco_memcpy_htod": (C.c_int, [vp, vp, C.c_uint64]),
            "blco_memcpy_dtoh": (C.c_int, [vp, vp, C.c_uint64]),
            "blco_memcpy_dtod": (C.c_int, [vp, vp, C.c_uint64]),
            "blco_memset_zero": (C.c_int, [vp, C.c_uint64]),
            "blco_synchronize": (C.c_int, []),
            "blco_dev_create": (C.c_int, [vp, C.c_int, C.c_int, C.c_uint64, C.POINTER(vp)]),
            "blco_dev_create_arrays": (C.c_int, [C.c_int, u64p, u64p, i32p, C.c_uint64, u64p, f64p,
                                                 C.c_uint64, u64p, u64p, u64p, C.c_int, C.c_int,
                                                 C.c_uint64, C.POINTER(vp)]),
            "blco_dev_create_budget": (C.c_int, [vp, C.c_int, C.c_uint64, C.c_uint64, C.POINTER(vp)]),
            "blco_dev_residency": (C.c_int, [vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                             C.POINTER(C.c_uint64)]),
            "blco_dev_rows_plan": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_uint32),
                                             C.POINTER(C.c_int), C.POINTER(C.c_uint64)]),
            "blco_dev_free": (None, [vp]),
            "blco_generate_device": (C.c_int, [C.POINTER(GenSpec), C.c_uint64, C.c_int, C.POINTER(vp)]),
            "blco_dev_to_host": (C.c_int, [vp, C.POINTER(vp)]),
            "blco_dev_info": (C.c_int, [vp, C.POINTER(C.c_int), u64p, C.POINTER(C.c_uint64),
                                        C.POINTER(C.c_uint64)]),
            "blco_dev_sample": (C.c_int, [vp, C.c_uint64, C.POINTER(vp)]),
            "blco_dev_set_value_dtype": (C.c_int, [vp, C.c_int]),
            "blco_dev_row_copy": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, u64p, vp, vp, vp, vp]),
            "blco_dev_bytes": (C.c_uint64, [vp]),
            "blco_mttkrp": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.POINTER(vp), vp, C.c_int,
                                      C.c_int, C.c_int, vp]),
            "blco_auto_kernel": (C.c_int, [vp, C.c_int, C.c_int]),
            "blco_gram": (C.c_int, [vp, C.c_int, C.c_uint64, C.c_int, vp, vp]),
            "blco_hadamard": (C.c_int, [vp, vp, C.c_uint64, vp]),
            "blco_value_fill": (C.c_int, [vp, C.c_uint64, C.c_double, vp]),
            "blco_vecsum": (C.c_int, [vp, C.c_uint64, vp]),
            "blco_solve": (C.c_int, [C.POINTER(vp), C.c_int, C.c_int, C.c_int, vp, vp, C.c_int,
                                     C.c_uint64, C.POINTER(C.c_int), vp]),
            "blco_colnorm": (C.c_int, [vp, C.c_int, C.c_uint64, C.c_int, C.c_int, vp, vp]),
            "blco_fit": (C.c_int, [C.POINTER(vp), C.c_int, C.c_int, vp, vp, vp, C.c_int, C.c_uint64,
                                   C.c_double, C.POINTER(C.c_double), vp]),
            "blco_cpals_create": (C.c_int, [vp, C.POINTER(CpalsOpts), C.POINTER(vp), C.POINTER(vp)]),
            "blco_cpals_iterate": (C.c_int, [vp, C.c_int, f64p, C.POINTER(C.c_int)]),
            "blco_cpals_get": (C.c_int, [vp, C.POINTER(vp), f64p]),
            "blco_cpals_times": (C.c_int, [vp, f64p, C.POINTER(C.c_double), C.POINTER(C.c_int),
                                           i32p]),
            "blco_cpals_reset_times": (None, [vp]),
            "blco_cpals_free": (None, [vp]),
            "blco_cpals": (C.c_int, [vp, C.POINTER(CpalsOpts), C.POINTER(vp), f64p, f64p,
                                     C.POINTER(C.c_int), C.POINTER(C.c_double)]),
            "blco_mttkrp_time": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.POINTER(vp), vp, C.c_int,
                                           C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float),
                                           C.POINTER(C.c_int)]),
            "blco_mlp_create": (C.c_int, [C.c_int, i32p, C.c_uint64, C.POINTER(vp)]),
            "blco_mlp_free": (None, [vp]),
            "blco_mlp_nparams": (C.c_uint64, [vp]),
            "blco_mlp_get_params": (C.c_int, [vp, vp, vp]),
            "blco_mlp_set_params": (C.c_int, [vp, f32p]),
            "blco_mlp_forward": (C.c_int, [vp, f32p, C.c_int, f32p]),
            "blco_mlp_loss_grad": (C.c_int, [vp, f32p, f32p, f32p, C.c_int, C.POINTER(C.c_float)]),
            "blco_mlp_adam_step": (C.c_int, [vp, C.c_float, C.c_float, C.c_float, C.c_float]),
            "blco_mlp_train": (C.c_int, [vp, f32p, f32p, f32p, C.c_int, C.c_int, C.c_int, C.c_float,
                                         C.c_uint64, vp]),
            "blco_sel_candidate": (C.c_int, [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                             C.POINTER(C.c_int)]),
            "blco_sel_valid": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int]),
            "blco_mode_features": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, f32p]),
            "blco_selector_install": (C.c_int, [C.c_int, i32p, f32p, f32p, f32p]),
            "blco_selector_load": (C.c_int, [C.c_char_p]),
            "blco_selector_installed": (C.c_int, []),
            "blco_select_kernel": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int),
                                             C.POINTER(C.c_int), C.POINTER(C.c_int)]),
            "blco_batch_create": (C.c_int, [C.c_int, C.POINTER(vp), C.POINTER(vp)]),
            "blco_batch_free": (None, [vp]),
            "blco_batch_tune": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.POINTER(vp), C.POINTER(vp),
                                          C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                          C.POINTER(C.c_float)]),
            "blco_batch_mttkrp": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.POINTER(vp), C.POINTER(vp),
                                            C.c_int, C.c_int, C.c_int]),
            "blco_batch_cpals": (C.c_int, [vp, C.POINTER(CpalsOpts), C.POINTER(vp), f64p, f64p, i32p,
                                           C.POINTER(C.c_double)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        msg = lib().blco_last_error().decode(errors="replace")
        raise BlcoError("blco error %d: %s" % (rc, msg))


def last_error():
    return lib().blco_last_error().decode(errors="replace")


def device_count():
    return lib().blco_device_count()


def _ptr_array(ptrs):
    return (vp * len(ptrs))(*[vp(p) for p in ptrs])


def _dptr(t):
    """Device address of a DeviceArray (or an int address)."""
    if t is None:
        return None
    if isinstance(t, int):
        return t
    return t.data_ptr()


def _stream(stream):
    return stream


def synchronize():
    _check(lib().blco_synchronize())


def _to_bf16_bits(a):
    """float64 -> bfloat16 bit patterns, round to nearest even."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def _from_bf16_bits(b):
    return (b.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


_ESZ = {F64: 8, F32: 4, BF16: 2}


class DeviceArray:
    """A device buffer from blco_malloc holding f64, f32 or bf16 elements."""

    def __init__(self, shape, dtype=F64):
        self.shape = tuple(int(x) for x in shape)
        self.dtype = dtype
        self.size = int(np.prod(self.shape)) if self.shape else 1
        self.nbytes = self.size * _ESZ[dtype]
        p = vp()
        _check(lib().blco_malloc(C.byref(p), self.nbytes))
        self.ptr = p.value
        if self.nbytes:
            _check(lib().blco_memset_zero(self.ptr, self.nbytes))

    @classmethod
    def from_numpy(cls, a, dtype=F64):
        a = np.asarray(a, np.float64)
        d = cls(a.shape, dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.asarray(a, np.float64).reshape(-1)
        if self.dtype == F64:
            h = np.ascontiguousarray(a)
        elif self.dtype == F32:
            h = np.ascontiguousarray(a, np.float32)
        else:
            h = _to_bf16_bits(a)
        if self.nbytes:
            _check(lib().blco_memcpy_htod(self.ptr, h.ctypes.data, self.nbytes))

    def numpy(self):
        """Copy back as float64 (waits for all device work)."""
        h = np.zeros(self.size, {F64: np.float64, F32: np.float32, BF16: np.uint16}[self.dtype])
        if self.nbytes:
            _check(lib().blco_memcpy_dtoh(h.ctypes.data, self.ptr, self.nbytes))
        out = _from_bf16_bits(h) if self.dtype == BF16 else h.astype(np.float64)
        return out.reshape(self.shape)

    def data_ptr(self):
        return self.ptr

    def free(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.blco_free(self.ptr)
            self.ptr = None

    def __del__(self):
        self.free()


# ------------------------------------------------------------------ COO

class Coo:
    """Owned COO tensor (SparseTensor, include/sptensor.hpp:8-26)."""

    def __init__(self, handle):
        self.h = handle
        n = C.c_int(0)
        dims = np.zeros(MAX_MODES, np.uint64)
        nnz = C.c_uint64(0)
        _check(lib().blco_coo_info(self.h, C.byref(n), dims, C.byref(nnz)))
        self.nmodes = n.value
        self.dims = dims[:n.value].copy()
        self.nnz = nnz.value

    @classmethod
    def from_arrays(cls, dims, cidx, vals):
        dims = np.ascontiguousarray(dims, np.uint64)
        cidx = np.ascontiguousarray(cidx, np.uint64)
        vals = np.ascontiguousarray(vals, np.float64)
        h = vp()
        _check(lib().blco_coo_create(len(dims), dims, cidx.shape[1], cidx.reshape(-1), vals,
                                     C.byref(h)))
        return cls(h.value)

    @classmethod
    def generate(cls, dims, nnz, zipf=None, gen_rank=10, value_kind=0, seed=1, topup=False):
        s = _gen_spec(dims, nnz, zipf, gen_rank, value_kind, seed, topup)
        h = vp()
        _check(lib().blco_generate(C.byref(s), C.byref(h)))
        return cls(h.value)

    @classmethod
    def load(cls, path, fmt):
        h = vp()
        _check(lib().blco_coo_import(path.encode(), fmt, C.byref(h)))
        return cls(h.value)

    def save(self, path, fmt):
        _check(lib().blco_coo_export(self.h, path.encode(), fmt))

    @property
    def cidx(self):
        p = lib().blco_coo_cidx(self.h)
        n = self.nmodes * self.nnz
        if n == 0:
            return np.zeros((self.nmodes, 0), np.uint64)
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint64)), (n,)).reshape(
            self.nmodes, self.nnz).copy()

    @property
    def vals(self):
        p = lib().blco_coo_vals(self.h)
        if self.nnz == 0:
            return np.zeros(0, np.float64)
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_double)), (self.nnz,)).copy()

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.blco_coo_free(self.h)
            self.h = None


def init_factor(seed, mode, rows, rank):
    U = np.zeros((int(rows), int(rank)), np.float64)
    _check(lib().blco_init_factor(seed, mode, rows, rank, U.reshape(-1)))
    return U


def init_factors(seed, dims, rank):
    return [init_factor(seed, m, int(d), rank) for m, d in enumerate(dims)]


# -------------------------------------------------------------- BLCO host

class BlcoHost:
    """Host BLCO tensor (blcotensor, include/blco.hpp:23-78)."""

    def __init__(self, handle):
        self.h = handle
        n = C.c_int(0)
        dims = np.zeros(MAX_MODES, np.uint64)
        nnz = C.c_uint64(0); nb = C.c_uint64(0); mx = C.c_uint64(0)
        _check(lib().blco_host_info(self.h, C.byref(n), dims, C.byref(nnz), C.byref(nb),
                                    C.byref(mx)))
        self.nmodes = n.value
        self.dims = dims[:n.value].copy()
        self.nnz = nnz.value
        self.nblocks = nb.value
        self.max_block_nnz = mx.value
        k = self.nmodes
        self.masks = np.zeros(k, np.uint64); self.pos = np.zeros(k, np.int32)
        self.alto_lo = np.zeros(k, np.uint64); self.alto_hi = np.zeros(k, np.uint64)
        _check(lib().blco_host_layout(self.h, self.masks, self.pos, self.alto_lo, self.alto_hi))
        nbk = max(self.nblocks, 1)
        self.block_start = np.zeros(nbk, np.uint64); self.block_nnz = np.zeros(nbk, np.uint64)
        bc = np.zeros(nbk * k, np.uint64)
        _check(lib().blco_host_blocks(self.h, self.block_start, self.block_nnz, bc))
        self.block_start = self.block_start[:self.nblocks]
        self.block_nnz = self.block_nnz[:self.nblocks]
        self.block_coords = bc[:self.nblocks * k].reshape(self.nblocks, k)

    @property
    def coords(self):
        if self.nnz == 0:
            return np.zeros(0, np.uint64)
        p = lib().blco_host_coords(self.h)
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint64)), (self.nnz,)).copy()

    @property
    def values(self):
        if self.nnz == 0:
            return np.zeros(0, np.float64)
        p = lib().blco_host_values(self.h)
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_double)), (self.nnz,)).copy()

    def decode(self):
        """Mode indices of every nonzero in BLCO order (nmodes x nnz)."""
        c = self.coords
        out = np.zeros((self.nmodes, self.nnz), np.uint64)
        blk = np.repeat(np.arange(self.nblocks), self.block_nnz.astype(np.int64))
        for m in range(self.nmodes):
            out[m] = ((c >> np.uint64(self.pos[m])) & self.masks[m]) | self.block_coords[blk, m]
        return out

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.blco_host_free(self.h)
            self.h = None


def encode(coo_or_dims, cidx=None, vals=None, max_block_size=0):
    """gen_blcotensor_host: encodes a Coo (or dims + arrays) into BLCO."""
    h = vp()
    if isinstance(coo_or_dims, Coo):
        _check(lib().blco_host_encode_coo(coo_or_dims.h, max_block_size, C.byref(h)))
    else:
        dims = np.ascontiguousarray(coo_or_dims, np.uint64)
        cidx = np.ascontiguousarray(cidx, np.uint64)
        vals = np.ascontiguousarray(vals, np.float64)
        _check(lib().blco_host_encode(len(dims), dims, cidx.shape[1], cidx.reshape(-1), vals,
                                      max_block_size, C.byref(h)))
    return BlcoHost(h.value)


def encode_device(coo_or_dims, cidx=None, vals=None, max_block_size=0):
    """gen_blcotensor_host computed on the GPU (blco_device_encode): the same
    BlcoHost as encode(), bit for bit."""
    if isinstance(coo_or_dims, Coo):
        dims, cidx, vals = coo_or_dims.dims, coo_or_dims.cidx, coo_or_dims.vals
    else:
        dims = coo_or_dims
    dims = np.ascontiguousarray(dims, np.uint64)
    cidx = np.ascontiguousarray(cidx, np.uint64)
    vals = np.ascontiguousarray(vals, np.float64)
    h = vp()
    _check(lib().blco_device_encode(len(dims), dims, cidx.shape[1], cidx.reshape(-1), vals,
                                    max_block_size, C.byref(h)))
    return BlcoHost(h.value)


# ------------------------------------------------------------ device side

def _gen_spec(dims, nnz, zipf, gen_rank, value_kind, seed, topup):
    s = GenSpec()
    s.nmodes = len(dims)
    for i, d in enumerate(dims):
        s.dims[i] = int(d)
        s.zipf_s[i] = float(zipf[i]) if zipf is not None else 0.0
    s.target_nnz = int(nnz)
    s.topup = 1 if topup else 0
    s.gen_rank = gen_rank
    s.value_kind = value_kind
    s.seed = seed
    return s


class DeviceTensor:
    """BLCO tensor in HBM (or streamed), gen_blcotensor_device."""

    def __init__(self, host, value_dtype=F64, stream_data=False, slot_nnz=0, device_budget=0):
        """device_budget > 0: blco_dev_create_budget (resident when the
        nonzeros fit, else a resident prefix plus streaming through slots of
        slot_nnz nonzeros; stream_data is then ignored)."""
        h = vp()
        if device_budget:
            _check(lib().blco_dev_create_budget(host.h, value_dtype, int(device_budget), slot_nnz,
                                                C.byref(h)))
        else:
            _check(lib().blco_dev_create(host.h, value_dtype, 1 if stream_data else 0, slot_nnz,
                                         C.byref(h)))
        self.h = h.value
        self.nmodes = host.nmodes
        self.dims = host.dims.copy()
        self.nnz = host.nnz
        self.nblocks = host.nblocks
        self.value_dtype = value_dtype

    @classmethod
    def generate(cls, dims, nnz, zipf=None, gen_rank=10, value_kind=0, seed=1, topup=False,
                 max_block_size=0, value_dtype=F64):
        """The tensor of Coo.generate(...) encoded and made resident on the GPU
        without a host copy (blco_generate_device)."""
        s = _gen_spec(dims, nnz, zipf, gen_rank, value_kind, seed, topup)
        h = vp()
        _check(lib().blco_generate_device(C.byref(s), max_block_size, value_dtype, C.byref(h)))
        self = cls.__new__(cls)
        self.h = h.value
        n = C.c_int(0)
        dims = np.zeros(MAX_MODES, np.uint64)
        nnz = C.c_uint64(0); nb = C.c_uint64(0)
        _check(lib().blco_dev_info(self.h, C.byref(n), dims, C.byref(nnz), C.byref(nb)))
        self.nmodes, self.dims, self.nnz, self.nblocks = n.value, dims[:n.value].copy(), nnz.value, nb.value
        self.value_dtype = value_dtype
        return self

    def rows_plan(self, mode, rank, factor_dtype=F64):
        """(chunks per wave, interleaved, LDS hot rows) of kernel 14 on a mode
        after it has run there."""
        c, i, h = C.c_uint32(0), C.c_int(0), C.c_uint64(0)
        _check(lib().blco_dev_rows_plan(self.h, mode, rank, factor_dtype, C.byref(c), C.byref(i), C.byref(h)))
        return c.value, bool(i.value), h.value

    def residency(self):
        """(resident nonzeros, streamed nonzeros, device bytes)."""
        r, s_, b = C.c_uint64(0), C.c_uint64(0), C.c_uint64(0)
        _check(lib().blco_dev_residency(self.h, C.byref(r), C.byref(s_), C.byref(b)))
        return r.value, s_.value, b.value

    def to_host(self):
        """Host copy (BlcoHost, values as f64) of a resident tensor."""
        h = vp()
        _check(lib().blco_dev_to_host(self.h, C.byref(h)))
        return BlcoHost(h.value)

    def sample(self, step):
        """Every step-th nonzero in BLCO order, as a Coo."""
        h = vp()
        _check(lib().blco_dev_sample(self.h, int(step), C.byref(h)))
        return Coo(h.value)

    def row_copy(self, mode, rank, factor_dtype=F32):
        """Kernel 14's row-ordered copy of `mode` (built if needed): dict of
        tiles, inner, fibers, coords, vals, runs, groups (n x 4)."""
        info = np.zeros(5, np.uint64)
        _check(lib().blco_dev_row_copy(self.h, mode, rank, factor_dtype, info, None, None, None, None))
        n = int(self.nnz)
        coords = np.zeros(n, np.uint64)
        vals = np.zeros(n, np.float64 if self.value_dtype == F64 else np.float32)
        runs = np.zeros(int(info[4]), np.uint64)
        groups = np.zeros((int(info[3]), 4), np.uint64)
        _check(lib().blco_dev_row_copy(self.h, mode, rank, factor_dtype, info, coords.ctypes.data,
                                       vals.ctypes.data, runs.ctypes.data, groups.ctypes.data))
        inner = int(info[1])
        return dict(tiles=int(info[0]), inner=inner - (1 << 64) if inner >= (1 << 63) else inner,
                    fibers=bool(info[2]), coords=coords, vals=vals, runs=runs, groups=groups)

    def set_value_dtype(self, value_dtype):
        _check(lib().blco_dev_set_value_dtype(self.h, value_dtype))
        self.value_dtype = value_dtype

    def bytes(self):
        return lib().blco_dev_bytes(self.h)

    def features(self, mode, rank, factor_dtype=F32):
        """Selector features of (mode, rank, factor type), SEL_NFEAT floats."""
        f = np.zeros(SEL_NFEAT, np.float32)
        _check(lib().blco_mode_features(self.h, mode, rank, factor_dtype, f))
        return f

    def candidate_valid(self, mode, rank, factor_dtype, c):
        return bool(lib().blco_sel_valid(self.h, mode, rank, factor_dtype, c))

    def select_kernel(self, mode, rank, factor_dtype=F32):
        """(kernel id, lvl3 partitions, tiles per wave) that kernel id 20 runs
        for this mode."""
        k, n, t = C.c_int(0), C.c_int(0), C.c_int(0)
        _check(lib().blco_select_kernel(self.h, mode, rank, factor_dtype, C.byref(k), C.byref(n), C.byref(t)))
        return k.value, n.value, t.value

    def auto_kernel(self, mode, rank):
        return lib().blco_auto_kernel(self.h, mode, rank)

    def free(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.blco_dev_free(self.h)
            self.h = None

    def __del__(self):
        self.free()


def mttkrp(dev, mode, rank, factors, out, factor_dtype=F64, kernel=K_AUTO, nprtn=0,
           thread_cf=0, stream=None):
    """out += MTTKRP(mode); factors/out are DeviceArrays (or device addresses)."""
    ptrs = _ptr_array([_dptr(f) if f is not None else 0 for f in factors])
    _check(lib().blco_mttkrp(dev.h, mode, rank, factor_dtype, ptrs, _dptr(out), kernel, nprtn,
                             thread_cf, _stream(stream)))


def mttkrp_time(dev, mode, rank, factors, out, factor_dtype=F64, kernel=K_AUTO, nprtn=0, reps=3, thread_cf=0):
    """Mean event-timed ms of one MTTKRP launch sequence (after one untimed
    call), and the kernel id that ran."""
    ptrs = _ptr_array([_dptr(f) for f in factors])
    ms, used = C.c_float(0), C.c_int(0)
    _check(lib().blco_mttkrp_time(dev.h, mode, rank, factor_dtype, ptrs, _dptr(out), kernel, nprtn,
                                  thread_cf, reps, C.byref(ms), C.byref(used)))
    return ms.value, used.value


def sel_candidate(c):
    """Candidate c of the selector -> (kernel id, lvl3 partitions, tiles per
    wave; 0 = the kernel's default)."""
    k, n, t = C.c_int(0), C.c_int(0), C.c_int(0)
    _check(lib().blco_sel_candidate(c, C.byref(k), C.byref(n), C.byref(t)))
    return k.value, n.value, t.value


def selector_install(widths, params, xmean, xstd):
    """Installs a selector network (widths[0] = SEL_NFEAT, widths[-1] =
    SEL_NCAND); widths=None uninstalls."""
    if widths is None:
        _check(lib().blco_selector_install(0, np.zeros(1, np.int32), np.zeros(1, np.float32),
                                           np.zeros(1, np.float32), np.zeros(1, np.float32)))
        return
    w = np.ascontiguousarray(widths, np.int32)
    _check(lib().blco_selector_install(len(w) - 1, w, np.ascontiguousarray(params, np.float32),
                                       np.ascontiguousarray(xmean, np.float32),
                                       np.ascontiguousarray(xstd, np.float32)))


def selector_load(path):
    _check(lib().blco_selector_load(os.fsencode(path)))


def selector_installed():
    return bool(lib().blco_selector_installed())


def write_selector(path, widths, params, xmean, xstd):
    """Writes a selector file: b"BLCOSEL1", int32 L, int32 widths[L+1],
    f32 mean[NFEAT], f32 scale[NFEAT], f32 parameters."""
    w = np.asarray(widths, np.int32)
    with open(path, "wb") as f:
        f.write(b"BLCOSEL1")
        f.write(np.int32(len(w) - 1).tobytes())
        f.write(w.tobytes())
        f.write(np.asarray(xmean, np.float32).tobytes())
        f.write(np.asarray(xstd, np.float32).tobytes())
        f.write(np.asarray(params, np.float32).tobytes())


class Batch:
    """Many resident device tensors decomposed together (one persistent
    MTTKRP kernel per mode, graph-captured ALS sweeps)."""

    def __init__(self, devs):
        self.devs = list(devs)
        arr = (vp * len(self.devs))(*[vp(d.h) for d in self.devs])
        h = vp()
        _check(lib().blco_batch_create(len(self.devs), arr, C.byref(h)))
        self.h = h.value

    def _tables(self, factors, outs):
        F = [0] * (len(self.devs) * MAX_MODES)
        for t, fl in enumerate(factors):
            for m, f in enumerate(fl):
                F[t * MAX_MODES + m] = _dptr(f) or 0
        return _ptr_array(F), _ptr_array([_dptr(o) or 0 for o in outs])

    def tune(self, mode, rank, factors, outs, factor_dtype=F32):
        """(tpw, waves, workgroups per CU, ms) picked for (mode, rank, type)."""
        F, O = self._tables(factors, outs)
        a, b, c, ms = C.c_int(0), C.c_int(0), C.c_int(0), C.c_float(0)
        _check(lib().blco_batch_tune(self.h, mode, rank, factor_dtype, F, O, C.byref(a), C.byref(b),
                                     C.byref(c), C.byref(ms)))
        return a.value, b.value, c.value, ms.value

    def mttkrp(self, mode, rank, factors, outs, factor_dtype=F32, tpw=0, waves=0, wgs_per_cu=0):
        F, O = self._tables(factors, outs)
        _check(lib().blco_batch_mttkrp(self.h, mode, rank, factor_dtype, F, O, tpw, waves, wgs_per_cu))

    def cpals(self, U0s, rank, max_iters=20, tolerance=0.0, precision=F32):
        """U0s[t][m] host f64 initial factors.  Returns (U, lambda, fits,
        iters, ms per batch sweep after iteration 0)."""
        T = len(self.devs)
        o = _opts(rank, max_iters, tolerance, K_AUTO, precision, 0, 0, 0, False, False)
        U = [[np.array(u, dtype=np.float64, order="C", copy=True) for u in ul] for ul in U0s]
        ptrs = [0] * (T * MAX_MODES)
        for t, ul in enumerate(U):
            for m, u in enumerate(ul):
                ptrs[t * MAX_MODES + m] = u.ctypes.data
        lam = np.zeros(T * rank, np.float64)
        fits = np.zeros(T, np.float64)
        iters = np.zeros(T, np.int32)
        ms = C.c_double(0)
        _check(lib().blco_batch_cpals(self.h, C.byref(o), _ptr_array(ptrs), lam, fits, iters, C.byref(ms)))
        return U, lam.reshape(T, rank), fits, iters, ms.value

    def free(self):
        if self.h:
            lib().blco_batch_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Mlp:
    """Dense ReLU network run and trained by the library's HIP kernels."""

    def __init__(self, widths, seed=0):
        self.widths = [int(w) for w in widths]
        h = vp()
        _check(lib().blco_mlp_create(len(self.widths) - 1, np.asarray(self.widths, np.int32),
                                     seed, C.byref(h)))
        self.h = h.value
        self.nparams = int(lib().blco_mlp_nparams(self.h))

    def params(self):
        p = np.zeros(self.nparams, np.float32)
        _check(lib().blco_mlp_get_params(self.h, p.ctypes.data, None))
        return p

    def grads(self):
        g = np.zeros(self.nparams, np.float32)
        _check(lib().blco_mlp_get_params(self.h, None, g.ctypes.data))
        return g

    def set_params(self, p):
        _check(lib().blco_mlp_set_params(self.h, np.ascontiguousarray(p, np.float32)))

    def forward(self, X):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.zeros((X.shape[0], self.widths[-1]), np.float32)
        _check(lib().blco_mlp_forward(self.h, X, X.shape[0], Y))
        return Y

    def loss_grad(self, X, T, M):
        l = C.c_float(0)
        _check(lib().blco_mlp_loss_grad(self.h, np.ascontiguousarray(X, np.float32),
                                        np.ascontiguousarray(T, np.float32),
                                        np.ascontiguousarray(M, np.float32), len(X), C.byref(l)))
        return l.value

    def adam_step(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        _check(lib().blco_mlp_adam_step(self.h, lr, beta1, beta2, eps))

    def train(self, X, T, M, epochs, batch, lr, seed=0):
        trace = np.zeros(max(epochs, 1), np.float32)
        _check(lib().blco_mlp_train(self.h, np.ascontiguousarray(X, np.float32),
                                    np.ascontiguousarray(T, np.float32),
                                    np.ascontiguousarray(M, np.float32), len(X), epochs, batch, lr,
                                    seed, trace.ctypes.data))
        return trace[:epochs]

    def free(self):
        if self.h:
            lib().blco_mlp_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def gram(U, dtype, rows, rank, G, stream=None):
    _check(lib().blco_gram(_dptr(U), dtype, rows, rank, _dptr(G), _stream(stream)))


def hadamard(x, y, n, stream=None):
    _check(lib().blco_hadamard(_dptr(x), _dptr(y), n, _stream(stream)))


def value_fill(x, n, val, stream=None):
    _check(lib().blco_value_fill(_dptr(x), n, val, _stream(stream)))


def vecsum(x, n, stream=None):
    _check(lib().blco_vecsum(_dptr(x), n, _stream(stream)))


def solve(grams, mode, rank, M, U, dtype, rows, stream=None):
    """U = M V^{-1} (V^+ on fallback); returns 1 when the pseudo-inverse was used."""
    fb = C.c_int(0)
    _check(lib().blco_solve(_ptr_array([_dptr(g) for g in grams]), len(grams), mode, rank,
                            _dptr(M), _dptr(U), dtype, rows, C.byref(fb), _stream(stream)))
    return fb.value


def colnorm(U, dtype, rows, rank, which, lam, stream=None):
    _check(lib().blco_colnorm(_dptr(U), dtype, rows, rank, which, _dptr(lam), _stream(stream)))


def fit(grams, rank, lam, U_last, M_last, dtype, rows, normX2, stream=None):
    f = C.c_double(0.0)
    _check(lib().blco_fit(_ptr_array([_dptr(g) for g in grams]), len(grams), rank, _dptr(lam),
                          _dptr(U_last), _dptr(M_last), dtype, rows, normX2, C.byref(f),
                          _stream(stream)))
    return f.value


def _opts(rank, max_iters, tolerance, kernel, precision, nprtn, thread_cf, mode_order,
          use_graph, verbose, solve_floor=0):
    o = CpalsOpts()
    o.rank = rank; o.max_iters = max_iters; o.tolerance = tolerance; o.kernel = kernel
    o.precision = precision; o.nprtn = nprtn; o.thread_cf = thread_cf
    o.mode_order = mode_order; o.use_graph = 1 if use_graph else 0
    o.verbose = 1 if verbose else 0
    o.solve_floor = solve_floor
    return o


class CpAls:
    """Device-resident CP-ALS solver (cpals_blco_dev split into steps)."""

    def __init__(self, dev, U0, rank, precision=F64, kernel=K_AUTO, tolerance=0.0, nprtn=0,
                 thread_cf=0, mode_order=0, use_graph=False, verbose=False, solve_floor=0):
        self.dev = dev
        self.rank = rank
        self.nmodes = dev.nmodes
        self.o = _opts(rank, 0, tolerance, kernel, precision, nprtn, thread_cf, mode_order,
                       use_graph, verbose, solve_floor)
        U0 = [np.ascontiguousarray(u, np.float64) for u in U0]
        h = vp()
        _check(lib().blco_cpals_create(dev.h, C.byref(self.o),
                                       _ptr_array([u.ctypes.data for u in U0]), C.byref(h)))
        self.h = h.value

    def iterate(self, n):
        fits = np.zeros(max(n, 1), np.float64)
        done = C.c_int(0)
        _check(lib().blco_cpals_iterate(self.h, n, fits, C.byref(done)))
        return fits[:done.value].copy()

    def get(self):
        U = [np.zeros((int(d), self.rank), np.float64) for d in self.dev.dims]
        lam = np.zeros(self.rank, np.float64)
        _check(lib().blco_cpals_get(self.h, _ptr_array([u.ctypes.data for u in U]), lam))
        return U, lam

    def times(self):
        ms = np.zeros(MAX_MODES, np.float64)
        it = C.c_double(0.0)
        n = C.c_int(0)
        k = np.zeros(MAX_MODES, np.int32)
        _check(lib().blco_cpals_times(self.h, ms, C.byref(it), C.byref(n), k))
        return dict(mttkrp_ms=ms[:self.nmodes].copy(), iter_ms=it.value, iters=n.value,
                    kernels=k[:self.nmodes].copy())

    def reset_times(self):
        lib().blco_cpals_reset_times(self.h)

    def free(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.blco_cpals_free(self.h)
            self.h = None

    def __del__(self):
        self.free()


def cpals(dev, U0, rank, max_iters, tolerance=0.0, precision=F64, kernel=K_AUTO, nprtn=0,
          thread_cf=0, mode_order=0, use_graph=False, verbose=False):
    """cpals_blco_dev in one call; returns (U, lambda, fits)."""
    o = _opts(rank, max_iters, tolerance, kernel, precision, nprtn, thread_cf, mode_order,
              use_graph, verbose)
    U = [np.array(u, dtype=np.float64, order="C", copy=True) for u in U0]
    lam = np.zeros(rank, np.float64)
    fits = np.zeros(max(max_iters, 1), np.float64)
    iters = C.c_int(0)
    ff = C.c_double(0.0)
    _check(lib().blco_cpals(dev.h, C.byref(o), _ptr_array([u.ctypes.data for u in U]), lam, fits,
                            C.byref(iters), C.byref(ff)))
    return U, lam, fits[:iters.value].copy()
