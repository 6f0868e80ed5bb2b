"""Timing data for the learned MTTKRP kernel selector.

Tensor t of the selector set is drawn from seed 1004 * 100000 + t: order
uniform in {3, 4, 5}; mode lengths log-uniform in [1e2, 1e7]; nonzeros
log-uniform in [nnz_min, nnz_max] (duplicates merged, no top-up); per-mode
Zipf exponent uniform in [0, 1.5]; rank from {8, 16, 32, 64}; FP32 values
(Poisson counts of a rank-8 model).  Tensors 0..499 take nonzeros in [1e5,
1e7] (0..399 training, 400..499 held out), tensors 500..539 in [1e7, 1e8]
(500..529 training, 530..539 held out).  Each tensor is generated, encoded
and made resident on the GPU (blco_generate_device, the same tensor as the
host generator's; keys wider than it takes are generated on the host).  For every mode, every candidate that can run (kernel,
lvl3 partitions, tiles per wave) is timed with HIP events
(blco_mttkrp_time) on factors of ones, and the features of the mode are
recorded.

  python collect.py --first 0 --count 100 --out data/part0.npz
  python collect.py --first 500 --count 40 --nnz-min 1e7 --nnz-max 1e8 --out data/large.npz

Writes the file after every tensor, so an interrupted run keeps what it
measured.  Prints one line per tensor.

After a change to a kernel, --retime re-times the listed candidates of every
tensor in existing files in place (the tensors are drawn again from the
specs the files record), leaving the other columns as measured:

  python collect.py --retime 10 --update data/part_000.npz data/part_001.npz
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)


def load_pkg():
    spec = importlib.util.spec_from_file_location("blco", os.path.join(PKG, "__init__.py"),
                                                  submodule_search_locations=[PKG])
    B = importlib.util.module_from_spec(spec)
    sys.modules["blco"] = B
    spec.loader.exec_module(B)
    return B


BASE_SEED = 1004 * 100000


def tensor_spec(t, nnz_min, nnz_max):
    rng = np.random.default_rng(BASE_SEED + t)
    order = int(rng.integers(3, 6))
    dims = [int(round(10 ** rng.uniform(2, 7))) for _ in range(order)]
    nnz = int(round(10 ** rng.uniform(np.log10(nnz_min), np.log10(nnz_max))))
    zipf = [float(rng.uniform(0, 1.5)) for _ in range(order)]
    rank = int(rng.choice([8, 16, 32, 64]))
    return dict(t=t, seed=BASE_SEED + t, dims=dims, nnz=nnz, zipf=zipf, rank=rank)


def ones(B, rows, rank):
    """f32 device factor of ones (uploaded without a float64 copy)."""
    a = B.DeviceArray((rows, rank), B.F32)
    h = np.ones((rows, rank), np.float32)
    B._check(B.lib().blco_memcpy_htod(a.ptr, h.ctypes.data, a.nbytes))
    return a


def rule_candidate(B, dev, mode, rank, nprtn_rule):
    k = dev.auto_kernel(mode, rank)
    for c in range(B.SEL_NCAND):
        kc, nc, tc = B.sel_candidate(c)
        if kc == k and tc == 0 and (k != B.K_LVL3 or nc == nprtn_rule):
            return c
    raise RuntimeError("rule kernel %d is not a candidate" % k)


def time_candidate(B, dev, mode, rank, F, out, c):
    k, np_, tpw = B.sel_candidate(c)
    ms, used = B.mttkrp_time(dev, mode, rank, F, out, B.F32, k, np_, reps=1, thread_cf=tpw)
    if used != k:
        raise RuntimeError("candidate %d ran kernel %d" % (c, used))
    if ms < 20.0:  # short: average more launches
        ms, _ = B.mttkrp_time(dev, mode, rank, F, out, B.F32, k, np_, reps=5, thread_cf=tpw)
    return ms


def make_tensor(B, sp):
    """Tensor of spec sp on the device with FP32 values: the GPU generator,
    or the host generator for keys wider than it takes."""
    try:
        return B.DeviceTensor.generate(sp["dims"], sp["nnz"], zipf=sp["zipf"], gen_rank=8, value_kind=0,
                                       seed=sp["seed"], topup=False, value_dtype=1)
    except B.BlcoError:
        coo = B.Coo.generate(sp["dims"], sp["nnz"], zipf=sp["zipf"], gen_rank=8, value_kind=0,
                             seed=sp["seed"], topup=False)
        return B.DeviceTensor(B.encode(coo), value_dtype=1)


def retime(B, paths, cands, verbose):
    """Re-times candidates `cands` of every (tensor, mode) row of the data
    files `paths` in place."""
    for path in paths:
        with np.load(path) as z:
            D = {k: z[k] for k in z.files}
        meta = json.loads(str(D["meta"]))
        times = D["times"].copy()
        for sp in meta:
            t0 = time.perf_counter()
            dev = make_tensor(B, sp)
            if int(dev.nnz) != sp["nnz_unique"]:
                got = int(dev.nnz)
                dev.free()
                raise RuntimeError("tensor %d: %d nonzeros, the file records %d" % (sp["t"], got, sp["nnz_unique"]))
            R = sp["rank"]
            F = [ones(B, d, R) for d in sp["dims"]]
            rows = np.nonzero(D["tensor"] == sp["t"])[0]
            for i in rows:
                m = int(D["mode"][i])
                out = B.DeviceArray((sp["dims"][m], R), B.F32)
                for c in cands:
                    old = times[i, c]
                    times[i, c] = time_candidate(B, dev, m, R, F, out, c) if dev.candidate_valid(m, R, B.F32, c) \
                        else np.nan
                    if verbose:
                        print("  tensor %d mode %d candidate %d: %.3f -> %.3f ms" % (sp["t"], m, c, old, times[i, c]),
                              file=sys.stderr, flush=True)
                out.free()
            for f in F:
                f.free()
            dev.free()
            D["times"] = times
            np.savez(path, **D)
            print("%s tensor %d: candidates %s re-timed in %.1f s" % (os.path.basename(path), sp["t"], cands,
                                                                     time.perf_counter() - t0), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--first", type=int, default=0)
    ap.add_argument("--count", type=int, default=500)
    ap.add_argument("--nnz-min", type=float, default=1e5)
    ap.add_argument("--nnz-max", type=float, default=1e7)
    ap.add_argument("--nprtn-rule", type=int, default=16,
                    help="lvl3 partitions of the -k 10 rule (the reference's default is the "
                         "OpenMP thread count)")
    ap.add_argument("--out")
    ap.add_argument("--verbose", action="store_true", help="one line per timed candidate (stderr)")
    ap.add_argument("--retime", default="", help="comma list of candidates to re-time in the --update files")
    ap.add_argument("--update", nargs="+", default=[], help="data files whose --retime columns are re-timed")
    args = ap.parse_args()
    B = load_pkg()
    if B.device_count() < 1:
        sys.exit("collect.py needs a GPU")
    if args.update:
        cands = [int(c) for c in args.retime.split(",") if c]
        if not cands or any(c < 0 or c >= B.SEL_NCAND for c in cands):
            sys.exit("--update needs --retime with candidates in [0, %d)" % B.SEL_NCAND)
        retime(B, args.update, cands, args.verbose)
        return
    if not args.out:
        sys.exit("--out is required")
    rows = dict(feat=[], times=[], tensor=[], mode=[], rank=[], rule=[])
    meta = []
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    for t in range(args.first, args.first + args.count):
        sp = tensor_spec(t, args.nnz_min, args.nnz_max)
        t0 = time.perf_counter()
        dev = make_tensor(B, sp)
        R = sp["rank"]
        F = [ones(B, d, R) for d in sp["dims"]]
        tgen = time.perf_counter() - t0
        if args.verbose:
            print("  tensor %d: set up in %.1f s" % (t, tgen), file=sys.stderr, flush=True)
        for m in range(len(sp["dims"])):
            out = B.DeviceArray((sp["dims"][m], R), B.F32)
            tm = np.full(B.SEL_NCAND, np.nan)
            for c in range(B.SEL_NCAND):
                if dev.candidate_valid(m, R, B.F32, c):
                    tc = time.perf_counter()
                    tm[c] = time_candidate(B, dev, m, R, F, out, c)
                    if args.verbose:
                        print("  tensor %d mode %d candidate %d: %.3f ms (%.1f s)"
                              % (t, m, c, tm[c], time.perf_counter() - tc), file=sys.stderr, flush=True)
            rows["feat"].append(dev.features(m, R, B.F32))
            rows["times"].append(tm)
            rows["tensor"].append(t)
            rows["mode"].append(m)
            rows["rank"].append(R)
            rows["rule"].append(rule_candidate(B, dev, m, R, args.nprtn_rule))
            out.free()
        sp.update(nnz_unique=int(dev.nnz), blocks=int(dev.nblocks))
        meta.append(sp)
        for f in F:
            f.free()
        dev.free()
        np.savez(args.out, feat=np.array(rows["feat"], np.float32), times=np.array(rows["times"]),
                 tensor=np.array(rows["tensor"]), mode=np.array(rows["mode"]),
                 rank=np.array(rows["rank"]), rule=np.array(rows["rule"]),
                 meta=np.array(json.dumps(meta)))
        best = [np.nanmin(x) for x in rows["times"][-len(sp["dims"]):]]
        print("tensor %d: order %d nnz %d blocks %d rank %d setup %.1f s total %.1f s best ms %s"
              % (t, len(sp["dims"]), dev.nnz, dev.nblocks, R, tgen, time.perf_counter() - t0,
                 np.round(best, 3).tolist()), flush=True)


if __name__ == "__main__":
    main()
