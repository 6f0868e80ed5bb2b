"""The GPU generator (blco_generate_device) against the host generator and
host encoder: for the same spec the resident device tensor, copied back,
equals blco_generate + blco_host_encode bit for bit -- BLCO words, values,
blocks and block coordinates -- across uniform and Zipf modes, top-up on and
off, both value kinds, 64- and 65-bit keys, block caps, and f32 values."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    # dims, nnz, zipf, gen_rank, value_kind, seed, topup, max_block_size
    ((2000, 1500, 1000), 300_000, None, 10, 0, 1001, False, 0),
    ((300, 200, 100), 500_000, (1.3, 1.3, 1.3), 8, 1, 77, True, 0),  # near saturation, many rounds
    ((12092, 9184, 28818), 2_000_000, (1.1, 1.1, 0.0), 20, 1, 1002, True, 0),
    ((4_800_000, 1_800_000, 1_800_000), 1_500_000, (1.2, 1.2, 1.2), 20, 1, 1005, True, 0),  # 65 bits
    ((4_800_000, 1_800_000, 1_800_000), 400_000, (1.2, 1.2, 1.2), 20, 1, 1005, True, 50_000),
    ((3407, 2410, 6035, 33683), 27_128, None, 8, 0, 1003026, False, 0),  # order 4
    ((50, 40, 30, 20, 10), 100_000, (0.5, 0.0, 1.0, 0.0, 1.5), 3, 1, 9, True, 7_000),  # order 5
]


def _same(a, b):
    assert a.nnz == b.nnz and a.nblocks == b.nblocks
    assert np.array_equal(a.block_start, b.block_start)
    assert np.array_equal(a.block_nnz, b.block_nnz)
    assert np.array_equal(a.block_coords, b.block_coords)
    assert np.array_equal(a.masks, b.masks) and np.array_equal(a.pos, b.pos)
    assert np.array_equal(a.coords, b.coords)


@pytest.mark.parametrize("case", range(len(CASES)))
def test_device_generator_matches_host(B, gpu, case):
    dims, nnz, zipf, rg, vk, seed, topup, mbs = CASES[case]
    coo = B.Coo.generate(dims, nnz, zipf=zipf, gen_rank=rg, value_kind=vk, seed=seed, topup=topup)
    want = B.encode(coo, max_block_size=mbs)
    dev = B.DeviceTensor.generate(dims, nnz, zipf=zipf, gen_rank=rg, value_kind=vk, seed=seed, topup=topup,
                                  max_block_size=mbs)
    assert dev.nnz == coo.nnz
    got = dev.to_host()
    _same(got, want)
    assert np.array_equal(got.values, want.values), np.flatnonzero(got.values != want.values)[:10]
    dev.free()


def test_device_generator_f32_values_and_conversion(B, gpu):
    dims, nnz = (12092, 9184, 28818), 1_000_000
    kw = dict(zipf=(1.1, 1.1, 0.0), gen_rank=20, value_kind=1, seed=1002, topup=True)
    want = B.encode(B.Coo.generate(dims, nnz, **kw))
    d32 = B.DeviceTensor.generate(dims, nnz, value_dtype=B.F32, **kw)
    got = d32.to_host()
    _same(got, want)
    assert np.array_equal(got.values, want.values.astype(np.float32).astype(np.float64))
    d32.free()
    d64 = B.DeviceTensor.generate(dims, nnz, **kw)
    d64.set_value_dtype(B.F32)
    assert np.array_equal(d64.to_host().values, want.values.astype(np.float32).astype(np.float64))
    d64.free()


def test_device_tensor_sample(B, gpu):
    dims, nnz = (2000, 1500, 1000), 200_000
    coo = B.Coo.generate(dims, nnz, gen_rank=10, seed=5)
    h = B.encode(coo, max_block_size=30_000)
    dev = B.DeviceTensor(h)
    s = dev.sample(7)
    idx = np.arange(0, h.nnz, 7)
    assert s.nnz == len(idx)
    assert np.array_equal(s.cidx, h.decode()[:, idx])
    assert np.array_equal(s.vals, h.values[idx])
    dev.free()


def test_generated_tensor_runs_cpals_like_uploaded(B, gpu):
    """CP-ALS on the generated device tensor follows CP-ALS on the uploaded
    host-generated tensor (same tensor, same kernels; the MTTKRP's atomic
    additions may sum in another order, hence 1e-12 rather than equality)."""
    dims, nnz, R = (400, 300, 200), 60_000, 8
    kw = dict(zipf=(1.0, 0.5, 0.0), gen_rank=6, value_kind=1, seed=31, topup=True)
    U0 = B.init_factors(31, dims, R)
    a = B.DeviceTensor.generate(dims, nnz, **kw)
    b = B.DeviceTensor(B.encode(B.Coo.generate(dims, nnz, **kw)))
    fa = B.cpals(a, U0, R, 5, 0.0)[2]
    fb = B.cpals(b, U0, R, 5, 0.0)[2]
    assert np.allclose(fa, fb, rtol=1e-12, atol=0)
    a.free()
    b.free()


ROW_CASES = [
    # dims, nnz, zipf, seed, max_block_size, env, rank, value dtype (0 f64, 1 f32), factor dtype
    ((12092, 9184, 28818), 2_000_000, (1.1, 1.1, 0.0), 1002, 0, {}, 32, 1, "F32"),
    ((12092, 9184, 28818), 2_000_000, (1.1, 1.1, 0.0), 1002, 0, {"BLCO_XCD_TILES": "1"}, 32, 1, "F32"),
    ((12092, 9184, 28818), 1_000_000, (1.1, 1.1, 0.0), 1002, 0, {"BLCO_FIBERS": "1", "BLCO_XCD_TILES": "1"}, 16, 0,
     "F64"),
    ((4_800_000, 1_800_000, 1_800_000), 1_000_000, (1.2, 1.2, 1.2), 1005, 200_000, {"BLCO_XCD_TILES": "1"}, 32, 1,
     "F32"),
    ((3407, 2410, 6035, 33683), 200_000, None, 1003026, 30_000, {"BLCO_XCD_TILES": "1", "BLCO_FIBERS": "1"}, 16, 1,
     "BF16"),
    ((50, 40, 30, 20, 10), 100_000, (0.5, 0.0, 1.0, 0.0, 1.5), 9, 7_000, {}, 8, 0, "F64"),
]


@pytest.mark.parametrize("case", range(len(ROW_CASES)))
def test_row_copy_built_on_device_matches_host_build(B, gpu, case, monkeypatch):
    """Kernel 14's row-ordered copies built on the GPU equal the host build
    (download, host sort, upload) word for word: the order, the groups, the
    run starts, the tiling and fiber decisions."""
    dims, nnz, zipf, seed, mbs, env, R, vdt, fdt = ROW_CASES[case]
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    h = B.encode(B.Coo.generate(dims, nnz, zipf=zipf, gen_rank=5, value_kind=1, seed=seed, topup=True),
                 max_block_size=mbs)
    ftype = getattr(B, fdt)
    for mode in range(len(dims)):
        monkeypatch.setenv("BLCO_ROWORDER", "host")
        a = B.DeviceTensor(h, value_dtype=vdt)
        want = a.row_copy(mode, R, ftype)
        a.free()
        monkeypatch.delenv("BLCO_ROWORDER")
        b = B.DeviceTensor(h, value_dtype=vdt)
        got = b.row_copy(mode, R, ftype)
        b.free()
        for k in ("tiles", "inner", "fibers"):
            assert got[k] == want[k], (mode, k, got[k], want[k])
        for k in ("coords", "vals", "runs", "groups"):
            assert np.array_equal(got[k], want[k]), (mode, k)
