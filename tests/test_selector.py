"""The shipped kernel selector (blocked-linearized-coordinate_amd/selector/
selector.bin) against the committed timing data: its choices, computed with
the CPU oracle's forward pass (oracle/mlp.py) from the file's own feature
scaling and parameters, are within 10% of the exhaustive choice on the
held-out tensors (total time), and beat the -k 10 rule there.  Also the file
format and that the library loads it (no GPU needed to load)."""
import glob
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp as Morc  # noqa: E402

SEL = os.path.join(ROOT, "blocked-linearized-coordinate_amd", "selector")


def read_selector(path):
    raw = open(path, "rb").read()
    assert raw[:8] == b"BLCOSEL1"
    nl = int(np.frombuffer(raw, np.int32, 1, 8)[0])
    w = np.frombuffer(raw, np.int32, nl + 1, 12).tolist()
    off = 12 + 4 * (nl + 1)
    nf = w[0]
    mu = np.frombuffer(raw, np.float32, nf, off)
    sd = np.frombuffer(raw, np.float32, nf, off + 4 * nf)
    p = np.frombuffer(raw, np.float32, Morc.nparams(w), off + 8 * nf)
    assert off + 8 * nf + 4 * len(p) == len(raw)
    return w, mu.astype(np.float64), sd.astype(np.float64), p.astype(np.float64)


def load_data():
    parts = [np.load(f) for f in sorted(glob.glob(os.path.join(SEL, "data", "*.npz")))]
    return {k: np.concatenate([p[k] for p in parts]) for k in ("feat", "times", "tensor", "rule")}


def total_regret(times, choice):
    best = np.nanmin(times, axis=1)
    got = times[np.arange(len(times)), choice]
    assert not np.isnan(got).any()
    return got.sum() / best.sum() - 1.0


def test_selector_file_beats_rule_on_heldout():
    w, mu, sd, p = read_selector(os.path.join(SEL, "selector.bin"))
    D = load_data()
    te = D["tensor"] >= 400
    assert te.sum() >= 300 and (~te).sum() >= 1000
    X = (D["feat"][te].astype(np.float64) - mu) / sd
    pred = Morc.forward(w, p, X)[-1]
    valid = ~np.isnan(D["times"][te])
    choice = np.argmin(np.where(valid, pred, np.inf), axis=1)
    learned = total_regret(D["times"][te], choice)
    rule = total_regret(D["times"][te], D["rule"][te])
    assert learned <= 0.10, learned
    assert learned < rule, (learned, rule)


def test_library_loads_the_selector(B):
    B.selector_load(os.path.join(SEL, "selector.bin"))
    assert B.selector_installed()
    B.selector_install(None, None, None, None)
    assert not B.selector_installed()
