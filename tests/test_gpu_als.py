"""ALS step and CP-ALS driver on the GPU through the C ABI, checked against
the oracle's restatement of cpd_alto and against closed forms.

Tolerances: Gram / solve / norm / fit pieces in FP64 to 1e-11 relative; the
FP64 CP-ALS fit trajectory to 1e-8 absolute per iteration (the plain form's bar);
the FP32 / BF16 forms' final fit to 1e-4 / 5e-3 absolute of the FP64 oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def dev_array(B, a, dtype=0):
    return B.DeviceArray.from_numpy(a, dtype)


@pytest.mark.parametrize("rank", [1, 8, 16, 20, 32, 48, 64])
@pytest.mark.parametrize("rows", [1, 7, 1000, 28818])
def test_gram_f64(B, gpu, rank, rows):
    rng = np.random.default_rng(rank * 7 + rows)
    U = rng.random((rows, rank)) - 0.3
    G = B.DeviceArray((rank, rank))
    B.gram(dev_array(B, U), 0, rows, rank, G)
    want = U.T @ U
    got = G.numpy()
    assert np.abs(got - want).max() <= 1e-12 * max(1.0, np.abs(want).max())
    assert (got == got.T).all()


@pytest.mark.parametrize("dtype,tol", [(1, 2e-6), (2, 1e-5)])
def test_gram_f32_bf16(B, gpu, dtype, tol):
    rng = np.random.default_rng(3)
    U = rng.random((5000, 32))
    Ud = dev_array(B, U, dtype)
    Uq = Ud.numpy()  # what the device holds
    G = B.DeviceArray((32, 32))
    B.gram(Ud, dtype, 5000, 32, G)
    want = Uq.T @ Uq
    assert np.abs(G.numpy() - want).max() <= tol * np.abs(want).max()


def make_grams(rng, dims, R):
    U = [rng.random((d, R)) for d in dims]
    return U, [u.T @ u for u in U]


@pytest.mark.parametrize("rank", [8, 16, 32, 64, 20])
def test_solve_cholesky(B, gpu, rank):
    rng = np.random.default_rng(rank)
    dims = (300, 200, 100)
    _, grams = make_grams(rng, dims, rank)
    V = grams[0] * grams[2]
    M = rng.random((200, rank))
    Gd = [dev_array(B, g) for g in grams]
    Md = dev_array(B, M)
    Ud = B.DeviceArray(M.shape)
    fb = B.solve(Gd, 1, rank, Md, Ud, 0, 200)
    assert fb == 0
    want = np.linalg.solve(V, M.T).T
    got = Ud.numpy()
    assert np.linalg.norm(got - want) <= 1e-9 * np.linalg.norm(want) * max(1, np.linalg.cond(V) * 1e-6)


def test_solve_pinv_fallback(B, O, gpu):
    """Two identical factor columns make V singular: the Cholesky fails and
    the minimum-norm solution M V^+ is returned."""
    rng = np.random.default_rng(4)
    R = 8
    dims = (50, 40, 30)
    U = [rng.random((d, R)) for d in dims]
    for u in U:
        u[:, 3] = u[:, 1]
    grams = [u.T @ u for u in U]
    V = grams[1] * grams[2]
    M = rng.random((50, R))
    Gd = [dev_array(B, g) for g in grams]
    Md = dev_array(B, M)
    Ud = B.DeviceArray(M.shape)
    fb = B.solve(Gd, 0, R, Md, Ud, 0, 50)
    assert fb == 1
    P = O.pinv_sym(V)
    want = M @ P
    got = Ud.numpy()
    assert np.linalg.norm(got - want) <= 1e-8 * np.linalg.norm(want)
    assert np.allclose(P, np.linalg.pinv(V, hermitian=True), rtol=1e-6, atol=1e-9 * np.abs(P).max())


def test_solve_near_singular_uses_pinv(B, O, gpu):
    """A column equal to another up to 1e-10 makes V singular to working
    precision: its Cholesky pivot falls below R * eps * max V_ii, so the
    pseudo-inverse (not a factor of rounding noise) is applied."""
    rng = np.random.default_rng(9)
    R = 8
    dims = (60, 40, 30)
    U = [rng.random((d, R)) for d in dims]
    for u in U:
        u[:, 5] = u[:, 2] * (1 + 1e-10)
    grams = [u.T @ u for u in U]
    V = grams[1] * grams[2]
    M = rng.random((60, R))
    Gd = [dev_array(B, g) for g in grams]
    Md = dev_array(B, M)
    Ud = B.DeviceArray(M.shape)
    fb = B.solve(Gd, 0, R, Md, Ud, 0, 60)
    assert fb == 1
    want = M @ O.pinv_sym(V)
    assert np.linalg.norm(Ud.numpy() - want) <= 1e-6 * np.linalg.norm(want)


EPS64, EPS32 = 2.0 ** -52, 2.0 ** -23


def _degenerate_runs(B, precision, solve_floor):
    dims = (3407, 2410, 6035, 33683)
    t = B.Coo.generate(dims, 27128, gen_rank=8, value_kind=0, seed=1003026)
    h = B.encode(t)
    U0 = B.init_factors(1003026, dims, 16)
    runs = []
    for kernel in (B.K_AUTO, B.K_BATCHED, B.K_SELECT):
        print("degenerate tensor: precision", precision, "solve floor", solve_floor, "kernel", kernel, flush=True)
        dev = B.DeviceTensor(h, value_dtype=0 if precision == 0 else 1)
        s = B.CpAls(dev, U0, 16, precision=B.F64 if precision == 0 else B.F32, kernel=kernel,
                    solve_floor=solve_floor)
        fits = s.iterate(12)
        U, lam = s.get()
        s.free()
        dev.free()
        assert len(fits) == 12 and np.all(np.isfinite(fits)), (kernel, fits)
        assert all(np.all(np.isfinite(u)) for u in U) and np.all(np.isfinite(lam))
        runs.append(fits)
    return t, dims, U0, runs


def test_cpals_degenerate_tensor_fp64_matches_oracle(B, O, gpu):
    """A 4-mode tensor with 27k nonzeros over modes of up to 33683 rows
    (most rows empty) at rank 16: its normal equations are singular to
    working precision from iteration 4 on.  The plain form follows the
    oracle's CP-ALS (the same pivot floor, R * 2^-52 * max V_ii, and
    pseudo-inverse tolerance) to 1e-8 per iteration with every kernel."""
    t, dims, U0, runs = _degenerate_runs(B, 0, 0)
    at = O.Alto(dims, t.cidx, t.vals, nprtn=8)
    _, _, fo, _, _ = at.cpd(U0, 12, 0.0, pivot_eps=EPS64)
    for fits in runs:
        assert np.abs(fits - fo).max() <= 1e-8, (fits, fo)
    finals = [f[-1] for f in runs]
    assert max(finals) - min(finals) <= 1e-8, finals


def test_cpals_degenerate_tensor_fp32_default_floor(B, O, gpu):
    """The FP32 form with the default (f64) pivot floor, the rule that keeps
    it on the FP64 oracle's trajectory on regular tensors: on this
    degenerate tensor its f32 MTTKRP sums, amplified by the near-singular
    normal equations, make the trajectory depend on the kernel's summation
    order.  Every kernel stays finite and ends within 5e-3 of the FP64
    oracle's final fit (the three kernels' spread measured 1.3e-3)."""
    t, dims, U0, runs = _degenerate_runs(B, 1, 0)
    at = O.Alto(dims, t.cidx, t.vals, nprtn=8)
    _, _, fo, _, _ = at.cpd(U0, 12, 0.0, pivot_eps=EPS64)
    for fits in runs:
        assert abs(fits[-1] - fo[-1]) <= 5e-3, (fits, fo)


def test_cpals_degenerate_tensor_fp32_precision_floor(B, O, gpu):
    """The FP32 form with solve_floor = 1 (pivot floor and pseudo-inverse
    tolerance at f32 precision, that of its Grams): a V singular to f32
    precision takes the pseudo-inverse whatever the MTTKRP's summation
    order, so every kernel follows the oracle's CP-ALS under the same rule
    to 1e-5 per iteration and the kernels agree to 1e-5."""
    t, dims, U0, runs = _degenerate_runs(B, 1, 1)
    at = O.Alto(dims, t.cidx, t.vals, nprtn=8)
    _, _, fo, _, _ = at.cpd(U0, 12, 0.0, pivot_eps=EPS32)
    for fits in runs:
        assert np.abs(fits - fo).max() <= 1e-5, (fits, fo)
    finals = [f[-1] for f in runs]
    assert max(finals) - min(finals) <= 1e-5, finals


@pytest.mark.parametrize("which", [0, 1])
@pytest.mark.parametrize("dtype", [0, 1])
def test_colnorm(B, O, gpu, which, dtype):
    rng = np.random.default_rng(5)
    U = rng.random((3000, 32)) * 4 - 1
    U[:, 5] = 0.0  # zero column: lambda 0, column left at zero (L2)
    Ud = dev_array(B, U, dtype)
    held = Ud.numpy()  # the values as stored on the device, before the norm
    lam = B.DeviceArray((32,))
    B.colnorm(Ud, dtype, 3000, 32, which, lam)
    want, wl = O.kruskal_norm(held, which)
    tol = 1e-12 if dtype == 0 else 1e-6
    assert np.allclose(lam.numpy(), wl, rtol=tol)
    assert np.allclose(Ud.numpy(), want, rtol=tol, atol=tol)


def test_fit_matches_oracle(B, O, gpu):
    rng = np.random.default_rng(6)
    dims = (40, 30, 20)
    R = 16
    U = [rng.random((d, R)) for d in dims]
    grams = [u.T @ u for u in U]
    lam = rng.random(R) + 0.5
    M = rng.random((20, R))
    normX2 = 1234.5
    want = O.cpd_fit(U[2], M, lam, grams, normX2)
    got = B.fit([dev_array(B, g) for g in grams], R, dev_array(B, lam),
                dev_array(B, U[2]), dev_array(B, M), 0, 20, normX2)
    assert abs(got - want) <= 1e-12 * max(1.0, abs(want))


def test_hadamard_fill_vecsum(B, gpu):
    rng = np.random.default_rng(7)
    x = rng.random(1000)
    y = rng.random(1000)
    xd, yd = dev_array(B, x), dev_array(B, y)
    B.hadamard(xd, yd, 1000)
    assert np.allclose(xd.numpy(), x * y, rtol=1e-15)
    B.value_fill(xd, 1000, 2.5)
    assert (xd.numpy() == 2.5).all()
    B.vecsum(yd, 1000)
    assert abs(yd.numpy()[0] - y.sum()) <= 1e-12 * y.sum()


# ------------------------------------------------------------- CP-ALS


def c1_like(B, O, nnz=100000, seed=1001):
    dims = (2000, 1500, 1000)
    t = B.Coo.generate(dims, nnz, gen_rank=10, value_kind=0, seed=seed)
    return dims, t, B.encode(t), O.Alto(dims, t.cidx, t.vals, nprtn=8)


@pytest.fixture(scope="module")
def small(B, O):
    return c1_like(B, O)


def test_cpals_f64_trajectory_matches_oracle(B, O, gpu, small):
    dims, t, h, at = small
    R = 16
    U0 = B.init_factors(1001, dims, R)
    Uo, lo, fo, _, _ = at.cpd(U0, 20, 0.0)
    dev = B.DeviceTensor(h)
    U, lam, fits = B.cpals(dev, U0, R, 20, 0.0)
    assert len(fits) == 20
    assert np.abs(fits - fo).max() <= 1e-8, np.abs(fits - fo).max()
    for a, b in zip(U, Uo):
        assert np.linalg.norm(a - b) <= 1e-6 * np.linalg.norm(b)
    assert np.allclose(lam, lo, rtol=1e-6)


@pytest.mark.parametrize("kernel", ["K_LVL1", "K_ATOMIC", "K_AUTO"])
def test_cpals_kernels_agree(B, O, gpu, small, kernel):
    dims, t, h, at = small
    R = 8
    U0 = B.init_factors(5, dims, R)
    _, _, fo, _, _ = at.cpd(U0, 5, 0.0)
    dev = B.DeviceTensor(h)
    _, _, fits = B.cpals(dev, U0, R, 5, 0.0, kernel=getattr(B, kernel))
    assert np.abs(fits - fo).max() <= 1e-8


def test_cpals_four_modes_and_short_modes(B, O, gpu):
    dims = (60, 700, 50, 90)
    t = B.Coo.generate(dims, 80000, gen_rank=6, seed=9)
    at = O.Alto(dims, t.cidx, t.vals, nprtn=8)
    U0 = B.init_factors(9, dims, 16)
    _, _, fo, _, _ = at.cpd(U0, 10, 0.0)
    dev = B.DeviceTensor(B.encode(t))
    _, _, fits = B.cpals(dev, U0, 16, 10, 0.0)
    assert np.abs(fits - fo).max() <= 1e-8


@pytest.mark.parametrize("precision,tol", [(1, 1e-4), (2, 5e-3)])
def test_cpals_precision_forms(B, O, gpu, small, precision, tol):
    dims, t, h, at = small
    R = 16
    U0 = B.init_factors(1001, dims, R)
    _, _, fo, _, _ = at.cpd(U0, 20, 0.0)
    dev = B.DeviceTensor(h, value_dtype=1)
    _, _, fits = B.cpals(dev, U0, R, 20, 0.0, precision=precision)
    print("precision", precision, "final fit", fits[-1], "oracle", fo[-1], flush=True)
    assert abs(fits[-1] - fo[-1]) <= tol


def test_cpals_graph_replay_matches_eager(B, gpu, small):
    """A hipGraph-captured sweep replayed many times gives the eager fits."""
    dims, t, h, at = small
    R = 16
    U0 = B.init_factors(3, dims, R)
    dev = B.DeviceTensor(h)
    e = B.CpAls(dev, U0, R)
    fe = e.iterate(8)
    g = B.CpAls(dev, U0, R, use_graph=True)
    fg = np.concatenate([g.iterate(3), g.iterate(5)])
    assert np.abs(fe - fg).max() <= 1e-12
    Ue, le = e.get()
    Ug, lg = g.get()
    for a, b in zip(Ue, Ug):
        assert np.abs(a - b).max() <= 1e-9 * np.abs(a).max()


def test_cpals_tolerance_stops(B, O, gpu, small):
    dims, t, h, at = small
    U0 = B.init_factors(2, dims, 8)
    _, _, fo, _, _ = at.cpd(U0, 50, 1e-3)
    dev = B.DeviceTensor(h)
    _, _, fits = B.cpals(dev, U0, 8, 50, 1e-3)
    assert len(fits) == len(fo) < 50


def test_cpals_streamed_matches_resident(B, gpu, small):
    dims, t, h, at = small
    U0 = B.init_factors(4, dims, 16)
    a = B.DeviceTensor(h)
    b = B.DeviceTensor(h, stream_data=True, slot_nnz=20000)
    _, _, fa = B.cpals(a, U0, 16, 5, 0.0)
    _, _, fb = B.cpals(b, U0, 16, 5, 0.0)
    assert np.abs(fa - fb).max() <= 1e-10


def test_cpals_device_budget_matches_resident(B, gpu, small):
    """CP-ALS on a tensor of which a device budget holds about a third
    (the rest streamed) follows the resident trajectory to 1e-10."""
    dims, t, h, at = small
    U0 = B.init_factors(4, dims, 16)
    a = B.DeviceTensor(h)
    b = B.DeviceTensor(h, slot_nnz=20000, device_budget=(2 * 20000 + h.nnz // 3) * 16)
    assert 0 < b.residency()[1] < h.nnz
    _, _, fa = B.cpals(a, U0, 16, 5, 0.0)
    _, _, fb = B.cpals(b, U0, 16, 5, 0.0)
    assert np.abs(fa - fb).max() <= 1e-10


def test_cpals_mode_order_longest_first(B, O, gpu, small):
    dims, t, h, at = small
    U0 = B.init_factors(6, dims, 8)
    dev = B.DeviceTensor(h)
    _, _, f = B.cpals(dev, U0, 8, 5, 0.0, mode_order=1)
    assert np.isfinite(f).all() and f[-1] > f[0]
