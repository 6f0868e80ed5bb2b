"""The library's HIP network kernels (forward, masked MSE, backward, Adam)
against the float64 CPU oracle (oracle/mlp.py), and the learned kernel
selector's plumbing through the C ABI.

Tolerances: the kernels compute in float32 with fused multiply-adds over at
most 4096 products, so forward outputs, the loss and every gradient entry
are compared at 2e-5 relative to the largest magnitude of that quantity.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp as Morc  # noqa: E402

pytestmark = pytest.mark.gpu

WIDTHS = [24, 64, 64, 10]


def batch(widths, n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, widths[0])).astype(np.float32)
    T = rng.normal(size=(n, widths[-1])).astype(np.float32)
    Mk = (rng.random((n, widths[-1])) < 0.8).astype(np.float32)
    return X, T, Mk


def close(a, b, tol=2e-5):
    scale = max(1.0, float(np.abs(b).max()))
    return float(np.abs(np.asarray(a, np.float64) - b).max()) <= tol * scale


@pytest.mark.parametrize("widths", [WIDTHS, [7, 5], [24, 128, 10], [3, 256, 17, 9, 2]])
def test_forward_and_gradient_match_oracle(B, gpu, widths):
    net = B.Mlp(widths, seed=11)
    p = net.params().astype(np.float64)
    assert len(p) == Morc.nparams(widths)
    for n in (1, 37, 300):
        X, T, Mk = batch(widths, n, n)
        Y = net.forward(X)
        assert close(Y, Morc.forward(widths, p, X)[-1]), (widths, n)
        L = net.loss_grad(X, T, Mk)
        Lo, go = Morc.loss_grad(widths, p, X.astype(np.float64), T.astype(np.float64),
                                Mk.astype(np.float64))
        assert close(L, Lo), (widths, n, L, Lo)
        assert close(net.grads(), go), (widths, n)
    net.free()


def test_initialisation_is_he_uniform_and_seeded(B, gpu):
    a = B.Mlp(WIDTHS, seed=5).params()
    b = B.Mlp(WIDTHS, seed=5).params()
    c = B.Mlp(WIDTHS, seed=6).params()
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    for W, bias in Morc.split(WIDTHS, a):
        lim = np.sqrt(6.0 / W.shape[1])
        assert np.abs(W).max() <= lim and np.abs(W).max() > 0.8 * lim
        assert not bias.any()


def test_adam_steps_match_oracle(B, gpu):
    net = B.Mlp(WIDTHS, seed=3)
    p = net.params().astype(np.float64)
    opt = Morc.Adam(len(p))
    for k in range(4):
        X, T, Mk = batch(WIDTHS, 64, 100 + k)
        net.loss_grad(X, T, Mk)
        net.adam_step(1e-2)
        _, g = Morc.loss_grad(WIDTHS, p, X.astype(np.float64), T.astype(np.float64),
                              Mk.astype(np.float64))
        p = opt.step(p, g, 1e-2)
        assert close(net.params(), p, 1e-4), k


def test_training_fits_and_is_deterministic(B, gpu):
    rng = np.random.default_rng(0)
    X = rng.normal(size=(2000, 24)).astype(np.float32)
    A = rng.normal(size=(24, 10)) / 5
    T = (np.tanh(X @ A) + 0.1 * (X[:, :10] ** 2)).astype(np.float32)
    Mk = np.ones_like(T)
    runs = []
    for _ in range(2):
        net = B.Mlp(WIDTHS, seed=1)
        trace = net.train(X, T, Mk, epochs=40, batch=128, lr=3e-3, seed=2)
        runs.append((trace, net.params()))
        net.free()
    trace = runs[0][0]
    assert trace[-1] < 0.1 * trace[0], trace
    assert np.array_equal(runs[0][1], runs[1][1])


def test_selector_install_and_choice(B, gpu):
    dims = (60, 3000, 2000)
    t = B.Coo.generate(dims, 50000, seed=4)
    dev = B.DeviceTensor(B.encode(t), value_dtype=1)
    f = dev.features(1, 16, B.F32)
    assert f.shape == (B.SEL_NFEAT,)
    assert np.isclose(f[0], np.log2(3000)) and f[4] == 4 and f[5] == 3 and f[6] == 2
    assert abs(f[8:16].sum() - 1) < 1e-5 and abs(f[18:22].sum() - 1) < 1e-5
    assert f[23] == 0 and dev.features(0, 16, B.F32)[23] == 1
    # a network with zero weights predicts its last-layer biases; a feature
    # scale of 100 keeps every feature within the selector's size guard
    widths = [B.SEL_NFEAT, 8, B.SEL_NCAND]
    nparams = B.SEL_NFEAT * 8 + 8 + 8 * B.SEL_NCAND + B.SEL_NCAND
    xm, xs = np.zeros(B.SEL_NFEAT), np.full(B.SEL_NFEAT, 100.0)

    def install(pref, scale=xs, bias=None):
        p = np.zeros(nparams, np.float32)
        p[-B.SEL_NCAND:] = 5.0
        for rank, c in enumerate(pref):
            p[nparams - B.SEL_NCAND + c] = rank
        for c, v in (bias or {}).items():
            p[nparams - B.SEL_NCAND + c] = v
        B.selector_install(widths, p, xm, scale)

    try:
        install([6, 2])  # lvl3 with 8 partitions, then atomic
        assert dev.select_kernel(0, 16, B.F32) == (B.K_LVL3, 8, 0)
        assert dev.select_kernel(1, 16, B.F32) == (B.K_ATOMIC, 1, 0)  # lvl3 cannot hold 3000 rows
        install([1])
        assert dev.select_kernel(0, 16, B.F32) == (B.K_BATCHED, 1, 0)
        install([12])  # segmented lvl1 with 16 tiles per wave
        assert dev.select_kernel(0, 16, B.F32) == (B.K_BATCHED, 1, 16)
        # the rule's candidate (mode 0: lvl3 with 16 partitions) is kept
        # unless another is predicted 2^0.1 times faster
        install([], bias={6: 1.0, 7: 1.05})
        assert dev.select_kernel(0, 16, B.F32) == (B.K_LVL3, 16, 0)
        install([], bias={6: 1.0, 7: 1.3})
        assert dev.select_kernel(0, 16, B.F32) == (B.K_LVL3, 8, 0)
        # features beyond 4 training standard deviations (unit scale: log2
        # of 50000 nonzeros is 15.6) leave the mode to the rule
        install([6, 2], scale=np.ones(B.SEL_NFEAT))
        assert dev.select_kernel(0, 16, B.F32) == (dev.auto_kernel(0, 16), 0, 0)
        B.selector_install(None, None, None, None)
        assert dev.select_kernel(1, 16, B.F32) == (dev.auto_kernel(1, 16), 0, 0)
    finally:
        B.selector_install(None, None, None, None)
    dev.free()


def test_timed_mttkrp_reports_the_kernel(B, gpu):
    dims = (50, 400, 300)
    t = B.Coo.generate(dims, 20000, seed=8)
    dev = B.DeviceTensor(B.encode(t), value_dtype=1)
    F = [B.DeviceArray.from_numpy(np.ones((d, 16)), B.F32) for d in dims]
    out = B.DeviceArray((50, 16), B.F32)
    for c in range(B.SEL_NCAND):
        if not dev.candidate_valid(0, 16, B.F32, c):
            continue
        k, np_, tpw = B.sel_candidate(c)
        ms, used = B.mttkrp_time(dev, 0, 16, F, out, B.F32, k, np_, reps=2, thread_cf=tpw)
        assert ms > 0 and used == k
    dev.free()


def test_selector_training_from_committed_timings(B, gpu):
    """Trains the selector network with the library's kernels on the
    committed timings of the training tensors (0..399 and 500..529) and
    checks the library's decision (the rule's candidate unless the network
    predicts a clear gain, within the trained sizes) on the held-out ones
    (400..499 and 530..539): total-time regret against the exhaustive choice
    <= 10%, and below the -k 10 rule's."""
    import glob
    import importlib.util
    seldir = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                          "blocked-linearized-coordinate_amd", "selector")
    spec = importlib.util.spec_from_file_location("blco_selector_train", os.path.join(seldir, "train.py"))
    Tr = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(Tr)
    sel = os.path.join(seldir, "data")
    parts = [np.load(f) for f in sorted(glob.glob(os.path.join(sel, "*.npz")))]
    D = {k: np.concatenate([p[k] for p in parts]) for k in ("feat", "times", "tensor", "rule")}
    valid = ~np.isnan(D["times"])
    logt = np.where(valid, np.log2(np.where(valid, D["times"], 1.0)), 0.0).astype(np.float32)
    t = D["tensor"]
    tr = (t < 400) | ((t >= 500) & (t < 530))
    te = ~tr
    X = D["feat"].astype(np.float64)
    mu, sd = X[tr].mean(axis=0), X[tr].std(axis=0)
    sd = np.where(sd > 1e-6, sd, 1.0)
    Xn = ((X - mu) / sd).astype(np.float32)
    net = B.Mlp([B.SEL_NFEAT, 64, 64, B.SEL_NCAND], seed=7)
    trace = net.train(Xn[tr], logt[tr], valid[tr].astype(np.float32), epochs=300, batch=128, lr=2e-3, seed=7)
    assert trace[-1] < 0.05 * trace[0]
    pred = net.forward(Xn[te])
    net.free()
    T = D["times"][te]
    choice = Tr.choose_lib(pred, valid[te], D["rule"][te], X[te], Xn[te])
    best = np.nanmin(T, axis=1)
    learned = T[np.arange(len(T)), choice].sum() / best.sum() - 1
    rule = T[np.arange(len(T)), D["rule"][te]].sum() / best.sum() - 1
    assert learned <= 0.10 and learned < rule, (learned, rule)


def test_collect_retime_updates_only_the_listed_candidates(B, gpu, tmp_path):
    """collect.py --retime: the tensors recorded in a timing file are drawn
    again from their specs and only the listed candidates' columns are
    re-timed; a spec whose tensor no longer matches the recorded nonzero
    count is refused."""
    import importlib.util
    import json
    seldir = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                          "blocked-linearized-coordinate_amd", "selector")
    spec = importlib.util.spec_from_file_location("blco_selector_collect", os.path.join(seldir, "collect.py"))
    Col = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(Col)
    sp = dict(t=7, seed=123457, dims=[50, 60, 70], nnz=5000, zipf=[0.0, 0.5, 1.0], rank=8)
    dev = Col.make_tensor(B, sp)
    sp.update(nnz_unique=int(dev.nnz), blocks=int(dev.nblocks))
    dev.free()
    times = np.full((3, B.SEL_NCAND), 1000.0)
    path = str(tmp_path / "part.npz")

    def write(meta):
        np.savez(path, feat=np.zeros((3, B.SEL_NFEAT), np.float32), times=times, tensor=np.full(3, 7),
                 mode=np.arange(3), rank=np.full(3, 8), rule=np.full(3, 10), meta=np.array(json.dumps(meta)))

    write([sp])
    Col.retime(B, [path], [10, 13], False)
    with np.load(path) as z:
        T = z["times"]
        assert json.loads(str(z["meta"])) == [sp] and list(z["mode"]) == [0, 1, 2]
    for c in range(B.SEL_NCAND):
        if c in (10, 13):
            assert np.all((T[:, c] > 0) & (T[:, c] < 1000.0)), (c, T[:, c])
        else:
            assert np.all(T[:, c] == 1000.0), c
    write([dict(sp, nnz_unique=sp["nnz_unique"] + 1)])
    with pytest.raises(RuntimeError):
        Col.retime(B, [path], [10], False)
