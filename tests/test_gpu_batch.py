"""Many small tensors decomposed together: the persistent batch MTTKRP
against the oracle's ALTO MTTKRP of every tensor, the autotuner, and the
graph-captured batch CP-ALS against one blco_cpals run per tensor.

Tolerances: MTTKRP relative Frobenius error 1e-12 (FP64) and 2e-6 (FP32)
per tensor, as for the single-tensor kernels; the batch CP-ALS runs the same
iteration as the single-tensor driver with a different MTTKRP summation
order, so FP64 fits agree to 1e-9 and factors to 1e-7 relative after 6
iterations, and the stopping iteration of each tensor is the same.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [((40, 300, 200), 6000), ((500, 60), 4000), ((30, 20, 50, 40), 9000),
          ((9, 12, 7, 11, 8), 3000), ((800, 700, 600), 12000), ((64, 64, 64, 64), 7000)]


def rel(a, b):
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (n if n > 0 else 1.0)


@pytest.fixture(scope="module")
def small(B, O):
    out = []
    for k, (dims, nnz) in enumerate(SHAPES):
        t = B.Coo.generate(dims, nnz, gen_rank=4, value_kind=0, seed=300 + k)
        out.append((t, B.encode(t), O.Alto(dims, t.cidx, t.vals, nprtn=4)))
    return out


@pytest.mark.parametrize("dtype", [0, 1])
@pytest.mark.parametrize("rank", [16, 32, 7])
def test_batch_mttkrp_matches_oracle(B, O, gpu, small, dtype, rank):
    tol = 1e-12 if dtype == 0 else 2e-6
    devs = [B.DeviceTensor(h, value_dtype=dtype) for _, h, _ in small]
    bt = B.Batch(devs)
    rng = np.random.default_rng(rank)
    U = [[rng.random((d, rank)) for d in t.dims] for t, _, _ in small]
    F = [[B.DeviceArray.from_numpy(u, dtype) for u in ul] for ul in U]
    odt = B.F64 if dtype == 0 else B.F32
    for mode in range(5):
        outs = [B.DeviceArray((t.dims[mode], rank), odt) if t.nmodes > mode else None for t, _, _ in small]
        for shape in ((1, 1, 1), (4, 4, 4), (8, 2, 8), (2, 8, 1)):
            for o in outs:
                if o is not None:
                    o.upload(np.zeros(o.shape))
            print("batch mode", mode, "dtype", dtype, "rank", rank, "shape", shape, flush=True)
            bt.mttkrp(mode, rank, F, outs, dtype, *shape)
            for (t, _, at), o, u in zip(small, outs, U):
                if o is None:
                    continue
                e = rel(o.numpy(), at.mttkrp(u, mode))
                assert e <= tol, (t.dims, mode, shape, e)
    bt.free()
    for d in devs:
        d.free()


def test_batch_tuner_picks_a_listed_shape(B, gpu, small):
    devs = [B.DeviceTensor(h, value_dtype=1) for _, h, _ in small]
    bt = B.Batch(devs)
    F = [[B.DeviceArray.from_numpy(np.ones((d, 16)), B.F32) for d in t.dims] for t, _, _ in small]
    outs = [B.DeviceArray((t.dims[0], 16), B.F32) for t, _, _ in small]
    tpw, waves, wpc, ms = bt.tune(0, 16, F, outs, B.F32)
    assert tpw in (1, 2, 4, 8) and waves in (1, 2, 4, 8) and wpc in (1, 2, 4, 8)
    assert waves * wpc <= 32 and ms > 0
    assert bt.tune(0, 16, F, outs, B.F32) == (tpw, waves, wpc, ms)  # cached
    bt.free()
    for d in devs:
        d.free()


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_cpals_matches_single_runs(B, gpu, small, precision):
    R, iters = 8, 6
    devs = [B.DeviceTensor(h, value_dtype=0 if precision == 0 else 1) for _, h, _ in small]
    U0 = [B.init_factors(500 + k, t.dims, R) for k, (t, _, _) in enumerate(small)]
    singles = []
    for dev, u0 in zip(devs, U0):
        s = B.CpAls(dev, u0, R, precision=precision)
        fits = s.iterate(iters)
        U, lam = s.get()
        s.free()
        singles.append((U, lam, fits))
    bt = B.Batch(devs)
    U, lam, fits, its, ms = bt.cpals(U0, R, max_iters=iters, precision=precision)
    bt.free()
    ftol, utol = (1e-9, 1e-7) if precision == 0 else (1e-4, 1e-2)
    for k, (Us, lams, fs) in enumerate(singles):
        assert its[k] == iters
        assert abs(fits[k] - fs[-1]) <= ftol, (k, fits[k], fs[-1])
        for m in range(len(Us)):
            assert rel(U[k][m], Us[m]) <= utol, (k, m)
        assert rel(lam[k], lams) <= utol
    assert ms > 0
    for d in devs:
        d.free()


def test_batch_cpals_stops_each_tensor_on_its_own(B, gpu, small):
    R, tol = 6, 1e-3
    devs = [B.DeviceTensor(h) for _, h, _ in small]
    U0 = [B.init_factors(700 + k, t.dims, R) for k, (t, _, _) in enumerate(small)]
    want = []
    for dev, u0 in zip(devs, U0):
        _, _, f = B.cpals(dev, u0, R, max_iters=40, tolerance=tol)
        want.append(f)
    bt = B.Batch(devs)
    _, _, fits, its, _ = bt.cpals(U0, R, max_iters=40, tolerance=tol, precision=0)
    bt.free()
    for k, f in enumerate(want):
        assert its[k] == len(f), (k, its[k], len(f))
        assert abs(fits[k] - f[-1]) <= 1e-9
    for d in devs:
        d.free()


def test_batch_rejects_bad_input(B, gpu, small):
    t, h, _ = small[0]
    dev = B.DeviceTensor(h)
    streamed = B.DeviceTensor(h, stream_data=True)
    with pytest.raises(B.BlcoError):
        B.Batch([])
    with pytest.raises(B.BlcoError):
        B.Batch([dev, streamed])
    bt = B.Batch([dev])
    F = [[B.DeviceArray((d, 8)) for d in t.dims]]
    with pytest.raises(B.BlcoError):
        bt.mttkrp(3, 8, F, [B.DeviceArray((10, 8))], B.F64)  # no tensor has mode 3
    with pytest.raises(B.BlcoError):
        bt.mttkrp(0, 8, F, [B.DeviceArray((t.dims[0], 8))], B.F32)  # f32 factors on f64 values
    bt.free()
    dev.free()
    streamed.free()
