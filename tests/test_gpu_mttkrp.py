"""MTTKRP on the GPU through the C ABI, checked against the oracle's ALTO CPU
MTTKRP on the same seeded tensors.

Tolerances (relative Frobenius error against the FP64 oracle):
  FP64 factors/values: 1e-12 (the bar of the plain form)
  FP32 factors/values: 2e-6 (inputs rounded to f32, f32 accumulation)
  BF16 factors:        2e-2 (8-bit mantissa factors, f32 accumulation)
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = {0: 1e-12, 1: 2e-6, 2: 2e-2}


def rel(a, b):
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (n if n > 0 else 1.0)


def run_mttkrp(B, dev, U, mode, dtype, kernel, nprtn=0, tpw=0):
    R = U[0].shape[1]
    F = [B.DeviceArray.from_numpy(u, dtype) for u in U]
    out = B.DeviceArray((int(dev.dims[mode]), R), B.F64 if dtype == 0 else B.F32)
    B.mttkrp(dev, mode, R, F, out, factor_dtype=dtype, kernel=kernel, nprtn=nprtn, thread_cf=tpw)
    return out.numpy()


CASES = {
    # name: (dims, nnz, zipf, max_block_size)
    "c1shape": ((2000, 1500, 1000), 60000, None, 0),
    "c2shape": ((12092, 9184, 28818), 150000, (1.1, 1.1, 0.0), 0),
    "order4": ((40, 4000, 5, 600), 40000, None, 0),
    "order5": ((9, 80, 7, 60, 5), 30000, None, 0),
    "bits65": ((4800000, 1800000, 1800000), 40000, (1.2, 1.2, 1.2), 0),
    "manyblocks": ((3000, 2000, 1000), 50000, None, 7000),
    "order2": ((5000, 3000), 20000, None, 0),
}


@pytest.fixture(scope="module")
def tensors(B, O):
    out = {}
    for name, (dims, nnz, zipf, mbs) in CASES.items():
        t = B.Coo.generate(dims, nnz, zipf=zipf, gen_rank=5, value_kind=1, seed=77, topup=False)
        out[name] = (t, B.encode(t, max_block_size=mbs), O.Alto(dims, t.cidx, t.vals, nprtn=8))
    return out


def check(B, O, tensors, name, dtype, kernel, rank, nprtn=0, tpw=0, stream=False):
    t, h, at = tensors[name]
    dev = B.DeviceTensor(h, value_dtype=0 if dtype == 0 else 1, stream_data=stream,
                         slot_nnz=9000 if stream else 0)
    rng = np.random.default_rng(rank)
    U = [rng.random((int(d), rank)) for d in t.dims]
    print("case", name, "dtype", dtype, "kernel", kernel, "rank", rank, "nprtn", nprtn, flush=True)
    for m in range(t.nmodes):
        want = at.mttkrp(U, m)
        got = run_mttkrp(B, dev, U, m, dtype, kernel, nprtn, tpw)
        e = rel(got, want)
        assert e <= TOL[dtype], (name, dtype, kernel, rank, m, e)
    dev.free()


@pytest.mark.parametrize("name", list(CASES))
def test_lvl1_f64_all_shapes(B, O, gpu, tensors, name):
    check(B, O, tensors, name, 0, B.K_BATCHED, 32)


@pytest.mark.parametrize("rank", [8, 16, 20, 32, 64, 3])
def test_lvl1_f64_ranks(B, O, gpu, tensors, rank):
    check(B, O, tensors, "c1shape", 0, B.K_BATCHED, rank)


@pytest.mark.parametrize("kernel", ["K_LVL1", "K_LVL2", "K_BATCHED2", "K_ATOMIC", "K_SORTED", "K_AUTO"])
def test_kernels_f64(B, O, gpu, tensors, kernel):
    check(B, O, tensors, "manyblocks", 0, getattr(B, kernel), 16)
    check(B, O, tensors, "c2shape", 0, getattr(B, kernel), 32)


@pytest.mark.parametrize("dtype", [1, 2])
@pytest.mark.parametrize("rank", [16, 32, 64])
def test_precision_forms(B, O, gpu, tensors, dtype, rank):
    check(B, O, tensors, "c2shape", dtype, B.K_BATCHED, rank)
    check(B, O, tensors, "order4", dtype, B.K_AUTO, rank)
    check(B, O, tensors, "c2shape", dtype, B.K_SORTED, rank)


@pytest.mark.parametrize("name", list(CASES))
def test_row_ordered_all_shapes(B, O, gpu, tensors, name):
    """Kernel 14 on every shape (multi-block and 65-bit included), f64 and
    f32, and ranks that do not fill a lane group."""
    check(B, O, tensors, name, 0, B.K_SORTED, 32)
    check(B, O, tensors, name, 1, B.K_SORTED, 20)
    check(B, O, tensors, name, 0, B.K_SORTED, 3, tpw=1)


@pytest.mark.parametrize("nnz", [1_500_000, 3_000_000])
def test_row_ordered_default_chunk_lengths(B, O, gpu, nnz):
    """Kernel 14 at its default chunk length for tensors between 2^20 and
    2^22 nonzeros (4 and 8 tiles of 64 per wave; the small shapes above take
    2 and the 40M-nonzero tensor 16), FP64 at rank 32 and FP32 at rank 20,
    on a 4-mode tensor with Zipf-skewed modes."""
    dims = (3000, 2000, 500, 40)
    t = B.Coo.generate(dims, nnz, zipf=(1.1, 0.8, 0.0, 0.0), gen_rank=5, value_kind=1, seed=nnz % 9973,
                       topup=True)
    assert t.nnz == nnz
    at = O.Alto(dims, t.cidx, t.vals, nprtn=8)
    h = B.encode(t)
    for dtype, rank in ((0, 32), (1, 20)):
        dev = B.DeviceTensor(h, value_dtype=dtype)
        rng = np.random.default_rng(rank)
        U = [rng.random((d, rank)) for d in dims]
        print("default chunk length: nnz", nnz, "dtype", dtype, "rank", rank, flush=True)
        for m in range(len(dims)):
            e = rel(run_mttkrp(B, dev, U, m, dtype, B.K_SORTED), at.mttkrp(U, m))
            assert e <= TOL[dtype], (nnz, dtype, rank, m, e)
        dev.free()


@pytest.mark.parametrize("name", list(CASES))
def test_row_ordered_xcd_tiles(B, O, gpu, tensors, name, monkeypatch):
    """Kernel 14 over copies grouped by 8 tiles of the split mode (forced on
    these small tensors, whose tiles are unequal, so the padded chunk lists
    carry empty chunks)."""
    monkeypatch.setenv("BLCO_XCD_TILES", "1")
    check(B, O, tensors, name, 0, B.K_SORTED, 32)
    check(B, O, tensors, name, 1, B.K_SORTED, 20, tpw=1)


@pytest.mark.parametrize("name", [n for n in CASES if n != "order2"])
@pytest.mark.parametrize("tiles", ["0", "1"])
def test_fiber_kernel(B, O, gpu, tensors, name, tiles, monkeypatch):
    """The fiber kernel (forced on: these small tensors have short fibers)
    over fiber-ordered copies, untiled and with 8 forced tiles, for every
    precision and for ranks that do not fill a lane group."""
    monkeypatch.setenv("BLCO_FIBERS", "1")
    monkeypatch.setenv("BLCO_XCD_TILES", tiles)
    check(B, O, tensors, name, 0, B.K_SORTED, 32)
    check(B, O, tensors, name, 1, B.K_SORTED, 20, tpw=1)
    check(B, O, tensors, name, 2, B.K_SORTED, 16)
    check(B, O, tensors, name, 0, B.K_SORTED, 3)


def test_fiber_order_without_fiber_kernel(B, O, gpu, tensors, monkeypatch):
    """The row-ordered kernel over copies kept in ALTO order within rows."""
    monkeypatch.setenv("BLCO_FIBERS", "0")
    for name in ("c2shape", "order4", "bits65"):
        check(B, O, tensors, name, 0, B.K_SORTED, 32)
        check(B, O, tensors, name, 1, B.K_SORTED, 16)


@pytest.mark.parametrize("fibers", ["0", "1"])
def test_row_ordered_64bit_offsets(B, O, gpu, fibers, monkeypatch):
    """A rank-64 FP64 factor of 8.5M rows holds more than 4 GiB, so kernel 14
    (row-ordered and fiber forms) runs its 64-bit-offset variants, which
    gather through flat addresses instead of 32-bit buffer offsets."""
    monkeypatch.setenv("BLCO_FIBERS", fibers)
    dims, R = (8_500_000, 300, 200), 64
    assert dims[0] * R * 8 > 1 << 32
    t = B.Coo.generate(dims, 50000, gen_rank=4, value_kind=1, seed=64, topup=False)
    at = O.Alto(dims, t.cidx, t.vals, nprtn=8)
    dev = B.DeviceTensor(B.encode(t))
    rng = np.random.default_rng(64)
    U = [rng.random((d, R)) for d in dims]
    for m in range(3):
        print("64-bit offsets: fibers", fibers, "mode", m, flush=True)
        e = rel(run_mttkrp(B, dev, U, m, 0, B.K_SORTED), at.mttkrp(U, m))
        assert e <= TOL[0], (m, e)
    dev.free()


@pytest.fixture(scope="module")
def big_zipf(B, O):
    """40M unique nonzeros (over 32768 chunks of kernel 14's copies) with
    Zipf(1.2) indices on the two long modes (their 128 most frequent rows
    hold about half of the nonzeros)."""
    dims = (300_000, 200_000, 3000)
    t = B.Coo.generate(dims, 40_000_000, zipf=(1.2, 1.2, 0.0), gen_rank=4, value_kind=1, seed=36,
                       topup=True)
    assert t.nnz == 40_000_000
    return dims, t, B.encode(t), O.Alto(dims, t.cidx, t.vals)


@pytest.mark.parametrize("fibers,dtype", [("1", 1), ("0", 1), ("1", 0), ("0", 0)])
def test_row_ordered_hot_windows(B, O, gpu, big_zipf, fibers, dtype, monkeypatch):
    """Kernel 14 with its LDS windows of hot rows and several chunks per
    wave (a grid-interleaved walk in the fiber form): every mode against the
    oracle, FP32 at rank 32 and FP64 at rank 16 (128-byte rows), and the plan
    shows the windows in use."""
    monkeypatch.setenv("BLCO_FIBERS", fibers)
    dims, t, h, at = big_zipf
    R = 32 if dtype == 1 else 16
    dev = B.DeviceTensor(h, value_dtype=dtype)
    rng = np.random.default_rng(7)
    U = [rng.random((d, R)) for d in dims]
    windows = 0
    for m in range(3):
        print("hot windows: fibers", fibers, "dtype", dtype, "mode", m, flush=True)
        e = rel(run_mttkrp(B, dev, U, m, dtype, B.K_SORTED), at.mttkrp(U, m))
        assert e <= TOL[dtype], (m, e)
        cpw, inter, rows = dev.rows_plan(m, R, dtype)
        assert inter == (fibers == "1" and cpw > 1)
        windows += rows > 0
    assert windows >= 2
    dev.free()


@pytest.mark.parametrize("nprtn", [1, 4, 17])
def test_lvl3_short_modes(B, O, gpu, nprtn):
    dims = (60, 90, 5000)
    t = B.Coo.generate(dims, 80000, gen_rank=4, seed=3)
    h = B.encode(t)
    at = O.Alto(dims, t.cidx, t.vals, nprtn=8)
    for dtype in (0, 1):
        dev = B.DeviceTensor(h, value_dtype=dtype)
        rng = np.random.default_rng(5)
        U = [rng.random((d, 32)) for d in dims]
        for m in (0, 1):
            assert dev.auto_kernel(m, 32) == B.K_LVL3
            got = run_mttkrp(B, dev, U, m, dtype, B.K_LVL3, nprtn)
            assert rel(got, at.mttkrp(U, m)) <= TOL[dtype]
        # lvl3 asked for a long mode runs the segmented kernel instead
        got = run_mttkrp(B, dev, U, 2, dtype, B.K_LVL3, nprtn)
        assert rel(got, at.mttkrp(U, 2)) <= TOL[dtype]
        dev.free()


@pytest.mark.parametrize("tpw", [1, 2, 8, 64])
def test_thread_coarsening(B, O, gpu, tensors, tpw):
    check(B, O, tensors, "c2shape", 0, B.K_BATCHED, 32, tpw=tpw)


@pytest.mark.parametrize("kernel", ["K_LVL1", "K_BATCHED", "K_ATOMIC", "K_LVL3", "K_SORTED"])
def test_streamed_blocks(B, O, gpu, tensors, kernel):
    check(B, O, tensors, "manyblocks", 0, getattr(B, kernel), 16, stream=True)
    check(B, O, tensors, "bits65", 1, getattr(B, kernel), 32, stream=True)


@pytest.mark.parametrize("kernel", ["K_LVL1", "K_BATCHED", "K_BATCHED2", "K_ATOMIC", "K_LVL3", "K_AUTO"])
def test_device_budget_streams_the_rest(B, O, gpu, tensors, kernel):
    """A device budget below the tensor's size keeps the leading half of the
    nonzeros resident and streams the other half through two slots; every
    mode's MTTKRP matches the oracle, and (FP64) the resident tensor's to
    1e-10."""
    for name, dtype, rank in (("manyblocks", 0, 16), ("bits65", 1, 32)):
        t, h, at = tensors[name]
        vdt = 0 if dtype == 0 else 1
        per = 8 + (8 if dtype == 0 else 4)
        slot = 4096  # two slots hold less than a quarter of either tensor
        assert 4 * slot < h.nnz
        budget = 2 * slot * per + (h.nnz // 2) * per
        dev = B.DeviceTensor(h, value_dtype=vdt, slot_nnz=slot, device_budget=budget)
        full = B.DeviceTensor(h, value_dtype=vdt)
        res, streamed, _ = dev.residency()
        assert (res, streamed) == (h.nnz // 2, h.nnz - h.nnz // 2)
        assert full.residency()[:2] == (h.nnz, 0)
        rng = np.random.default_rng(rank)
        U = [rng.random((int(d), rank)) for d in t.dims]
        print("case", name, "budget", budget, "kernel", kernel, flush=True)
        for m in range(t.nmodes):
            got = run_mttkrp(B, dev, U, m, dtype, getattr(B, kernel))
            assert rel(got, at.mttkrp(U, m)) <= TOL[dtype], (name, kernel, m)
            if dtype == 0:
                ref = run_mttkrp(B, full, U, m, dtype, getattr(B, kernel))
                assert rel(got, ref) <= 1e-10, (name, kernel, m)
        dev.free()
        full.free()


def test_device_budget_limits(B, O, gpu, tensors):
    """A budget that holds the nonzeros keeps the tensor resident but no
    row-ordered copy beside it (the automatic kernel then runs the row
    window, still matching the oracle); a budget below the two slots is an
    error."""
    t, h, at = tensors["c2shape"]
    exact = h.nnz * 16
    dev = B.DeviceTensor(h, device_budget=exact)
    res, streamed, b0 = dev.residency()
    assert (res, streamed) == (h.nnz, 0)
    U = [np.random.default_rng(3).random((int(d), 32)) for d in t.dims]
    for m in range(t.nmodes):
        got = run_mttkrp(B, dev, U, m, 0, B.K_AUTO)
        assert rel(got, at.mttkrp(U, m)) <= TOL[0]
    assert dev.residency()[2] < b0 + h.nnz * 16  # no copy was built
    dev.free()
    with pytest.raises(B.BlcoError):
        B.DeviceTensor(h, slot_nnz=1 << 16, device_budget=(1 << 16) * 16)


def test_accumulates_into_output(B, O, gpu, tensors):
    """The output is accumulated into (the caller zeroes it)."""
    t, h, at = tensors["c1shape"]
    dev = B.DeviceTensor(h)
    rng = np.random.default_rng(9)
    U = [rng.random((int(d), 16)) for d in t.dims]
    F = [B.DeviceArray.from_numpy(u) for u in U]
    base = rng.random((int(t.dims[1]), 16))
    out = B.DeviceArray.from_numpy(base)
    B.mttkrp(dev, 1, 16, F, out, kernel=B.K_BATCHED)
    want = base + at.mttkrp(U, 1)
    assert rel(out.numpy(), want) <= 1e-12


def test_empty_tensor_is_a_no_op(B, gpu):
    h = B.encode((50, 40, 30), np.zeros((3, 0), np.uint64), np.zeros(0))
    dev = B.DeviceTensor(h)
    F = [B.DeviceArray.from_numpy(np.ones((d, 8))) for d in (50, 40, 30)]
    out = B.DeviceArray((50, 8))
    B.mttkrp(dev, 0, 8, F, out)
    assert np.abs(out.numpy()).sum() == 0.0


def test_bad_arguments_raise_before_launch(B, gpu, tensors):
    t, h, _ = tensors["c1shape"]
    dev = B.DeviceTensor(h)
    F = [B.DeviceArray((int(d), 16)) for d in t.dims]
    out = B.DeviceArray((int(t.dims[0]), 16))
    with pytest.raises(B.BlcoError):
        B.mttkrp(dev, 3, 16, F, out)  # mode out of range
    with pytest.raises(B.BlcoError):
        B.mttkrp(dev, 0, 16, F, out, kernel=99)  # unknown kernel id
    with pytest.raises(B.BlcoError):
        B.mttkrp(dev, 0, 16, F, out, factor_dtype=1)  # f32 factors on f64 values
    with pytest.raises(B.BlcoError):
        B.mttkrp(dev, 0, 16, [F[0], None, F[2]], out)  # missing factor


def test_row_ordered_copy_survives_window_scratch_growth(B, O, gpu, tensors):
    """Growing the row-window scratch releases the scratch only: the
    row-ordered copies built before stay valid (checked after reusing
    device memory)."""
    t, h, at = tensors["c2shape"]
    dev = B.DeviceTensor(h)
    rng = np.random.default_rng(21)
    U = [rng.random((int(d), 16)) for d in t.dims]
    want = at.mttkrp(U, 0)
    assert rel(run_mttkrp(B, dev, U, 0, 0, B.K_SORTED), want) <= 1e-12
    for rank in (8, 16, 32):
        Ur = [rng.random((int(d), rank)) for d in t.dims]
        for m in range(t.nmodes):
            run_mttkrp(B, dev, Ur, m, 0, B.K_BATCHED2)
    junk = [B.DeviceArray.from_numpy(np.full(1 << 20, 7.0)) for _ in range(16)]
    assert rel(run_mttkrp(B, dev, U, 0, 0, B.K_SORTED), want) <= 1e-12
    for j in junk:
        j.free()
    dev.free()
