"""Trains the learned MTTKRP kernel selector on timings from collect.py and
reports its regret.

The network (24 features -> 64 -> 64 -> 15 predicted log2 times) is trained
on the GPU by the library's own kernels (blco_mlp_train: forward, masked MSE,
backward, Adam) on the training tensors (is_train); the candidates that cannot
run for a mode are masked out of the loss.  On the held-out tensors it
reports, against the exhaustive choice (the fastest measured candidate of
each mode):

  regret_total = sum(time of the choice) / sum(fastest time) - 1
  regret_mean  = mean over modes of (time of the choice / fastest - 1)

for the library's learned choice (choose_lib), the network's plain argmin
and the -k 10 rule, and writes the selector file
that the library loads (selector.bin next to this script by default).

  python train.py data/part0.npz data/part1.npz ... --out selector.bin --report report.json
"""
import argparse
import importlib.util
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
HIDDEN = [64, 64]
N_TRAIN = 400


def load_pkg():
    spec = importlib.util.spec_from_file_location("blco", os.path.join(PKG, "__init__.py"),
                                                  submodule_search_locations=[PKG])
    B = importlib.util.module_from_spec(spec)
    sys.modules["blco"] = B
    spec.loader.exec_module(B)
    return B


def load(paths):
    parts = [np.load(p, allow_pickle=False) for p in paths]
    cat = {k: np.concatenate([p[k] for p in parts]) for k in ("feat", "times", "tensor", "mode", "rank", "rule")}
    # a tensor measured twice keeps its first rows
    seen, keep = set(), []
    for i, (t, m) in enumerate(zip(cat["tensor"], cat["mode"])):
        if (t, m) not in seen:
            seen.add((t, m))
            keep.append(i)
    return {k: v[keep] for k, v in cat.items()}


def regret(times, choice):
    best = np.nanmin(times, axis=1)
    got = times[np.arange(len(times)), choice]
    assert not np.isnan(got).any(), "a choice that cannot run"
    return dict(regret_total=float(got.sum() / best.sum() - 1.0),
                regret_mean=float(np.mean(got / best - 1.0)),
                regret_median=float(np.median(got / best - 1.0)),
                exact=float(np.mean(got == best)),
                time_ms=float(got.sum()), best_ms=float(best.sum()))


def is_train(t):
    """Training tensors: 0..399 of the 1e5-1e7-nonzero set and 500..529 of
    the 1e7-1e8-nonzero set; the rest are held out."""
    t = np.asarray(t)
    return (t < N_TRAIN) | ((t >= 500) & (t < 530))


def choose(pred, valid):
    p = np.where(valid, pred, np.inf)
    return np.argmin(p, axis=1)


# The library's decision (csrc/hip/selector.hip): the rule's candidate unless
# the network predicts 2^MARGIN times less time, and the rule outright for
# tensors of more than 2^MAX_LOG2_NNZ nonzeros and for modes whose size
# features lie beyond SIZE_Z training standard deviations.
MARGIN, MAX_LOG2_NNZ, SIZE_Z, SIZE_FEATS = 0.1, 23.25, 4.0, [0, 1, 2, 3, 4, 7, 16, 22]


def choose_lib(pred, valid, rule, X, Xn):
    best = choose(pred, valid)
    idx = np.arange(len(pred))
    inside = (X[:, 1] <= MAX_LOG2_NNZ) & (np.abs(Xn[:, SIZE_FEATS]) <= SIZE_Z).all(axis=1)
    rule_ok = valid[idx, rule]
    gain = np.where(rule_ok, pred[idx, rule] - pred[idx, best], np.inf)
    return np.where(inside & (gain > MARGIN), best, np.where(rule_ok, rule, best))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("data", nargs="+")
    ap.add_argument("--out", default=os.path.join(HERE, "selector.bin"))
    ap.add_argument("--report", default=None)
    ap.add_argument("--epochs", type=int, default=400)
    ap.add_argument("--batch", type=int, default=128)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=7)
    args = ap.parse_args()
    B = load_pkg()
    D = load(args.data)
    X = D["feat"].astype(np.float64)
    valid = ~np.isnan(D["times"])
    logt = np.where(valid, np.log2(np.where(valid, D["times"], 1.0)), 0.0)
    tr = is_train(D["tensor"])
    te = ~tr
    mu = X[tr].mean(axis=0)
    sd = X[tr].std(axis=0)
    sd = np.where(sd > 1e-6, sd, 1.0)
    Xn = ((X - mu) / sd).astype(np.float32)
    widths = [B.SEL_NFEAT] + HIDDEN + [B.SEL_NCAND]
    net = B.Mlp(widths, seed=args.seed)
    trace = net.train(Xn[tr], logt[tr].astype(np.float32), valid[tr].astype(np.float32),
                      epochs=args.epochs, batch=args.batch, lr=args.lr, seed=args.seed)
    pred = net.forward(Xn)
    params = net.params()
    net.free()
    rep = dict(widths=widths, epochs=args.epochs, batch=args.batch, lr=args.lr,
               n_train_modes=int(tr.sum()), n_test_modes=int(te.sum()),
               n_train_tensors=int(len(set(D["tensor"][tr].tolist()))),
               n_test_tensors=int(len(set(D["tensor"][te].tolist()))),
               loss_first=float(trace[0]), loss_last=float(trace[-1]))
    big = D["tensor"] >= 500
    for name, sel in (("train", tr), ("heldout", te), ("heldout_1e5_1e7", te & ~big), ("heldout_1e7_1e8", te & big)):
        if not sel.any():
            continue
        T = D["times"][sel]
        rep[name] = dict(learned=regret(T, choose_lib(pred[sel], valid[sel], D["rule"][sel], X[sel], Xn[sel])),
                         learned_argmin=regret(T, choose(pred[sel], valid[sel])),
                         rule_k10=regret(T, D["rule"][sel]),
                         always_window=regret(T, np.zeros(sel.sum(), int)),
                         always_lvl1=regret(T, np.ones(sel.sum(), int)),
                         always_row_ordered=regret(T, np.full(sel.sum(), 10)))
        resid = (pred[sel] - logt[sel])[valid[sel]]
        rep[name]["log2_time_rmse"] = float(np.sqrt(np.mean(resid ** 2)))
    B.write_selector(args.out, widths, params, mu, sd)
    if args.report:
        with open(args.report, "w") as f:
            json.dump(rep, f, indent=1)
    print(json.dumps(rep, indent=1))


if __name__ == "__main__":
    main()
