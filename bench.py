"""Benchmark: CP-ALS seconds per iteration and MTTKRP HBM GB/s on one MI355X,
FP32 form (f32 factors/values, f64 solve/lambda/fit) against the plain FP64
form, on the 12092 x 9184 x 28818 tensor with 77M nonzeros (modes 0,1
Zipf(1.1), mode 2 uniform, log1p Poisson values from a rank-20 Gamma(1,1)
model), rank 32.

A step is one CP-ALS iteration (MTTKRP + solve + normalisation + Gram for
every mode, then the fit).  The tensor is generated, encoded to BLCO and
uploaded before timing; W warm-up iterations run untimed, then K iterations
are timed with HIP events on the solver stream (and a host clock around
them).  MTTKRP kernel time per mode is event-timed on the same stream.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "CP-ALS s/iter + MTTKRP HBM GB/s: FP32+learned selector vs plain FP64, 1 GPU"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: dims, nnz, zipf, gen_rank, value_kind, rank, seed (1000 + config)
    "nell2-shape-77M": dict(dims=(12092, 9184, 28818), nnz=77_000_000, zipf=(1.1, 1.1, 0.0),
                            gen_rank=20, value_kind=1, rank=32, seed=1002),
    "small-1M": dict(dims=(2000, 1500, 1000), nnz=1_000_000, zipf=None, gen_rank=10,
                     value_kind=0, rank=16, seed=1001),
}


# The batched set: 64 tensors from seed 1003 * 1000 + t, order 3 or 4
# (50/50), mode lengths log-uniform in [100, 5e4], nonzeros log-uniform in
# [1e4, 2e6], uniform indices, Poisson values of a rank-8 Gamma(1,1) model;
# rank 16, FP32 factors with FP64 solve.
BATCH_SET = dict(count=64, seed=1003 * 1000, rank=16)


def batch_specs():
    out = []
    for t in range(BATCH_SET["count"]):
        rng = np.random.default_rng(BATCH_SET["seed"] + t)
        order = 3 if rng.random() < 0.5 else 4
        dims = tuple(int(round(10 ** rng.uniform(2, np.log10(5e4)))) for _ in range(order))
        nnz = int(round(10 ** rng.uniform(4, np.log10(2e6))))
        out.append((dims, nnz, BATCH_SET["seed"] + t))
    return out


def run_batched(B, args):
    """One batch CP-ALS (persistent MTTKRP per mode + captured sweep) against
    one plain CP-ALS per tensor, same initial factors, same iteration count."""
    R = BATCH_SET["rank"]
    t0 = time.perf_counter()
    tens = []
    for dims, nnz, seed in batch_specs():
        c = B.Coo.generate(dims, nnz, gen_rank=8, value_kind=0, seed=seed)
        tens.append((c, B.encode(c), B.init_factors(seed, dims, R)))
    total = sum(int(c.nnz) for c, _, _ in tens)
    log("batched set: %d tensors, %d nonzeros, generated in %.1f s" % (len(tens), total, time.perf_counter() - t0))
    devs = [B.DeviceTensor(h, value_dtype=1) for _, h, _ in tens]
    iters = args.warmup + args.steps
    # plain: one solver per tensor, eager launches, after the same warm-up
    plain_ms = 0.0
    plain_fits = []
    for dev, (_, _, U0) in zip(devs, tens):
        s = B.CpAls(dev, U0, R, precision=B.F32)
        s.iterate(args.warmup)
        s.reset_times()
        B.synchronize()
        fits = s.iterate(args.steps)
        plain_ms += s.times()["iter_ms"] / max(s.times()["iters"], 1)
        plain_fits.append(fits[-1])
        s.free()
    bt = B.Batch(devs)
    t0 = time.perf_counter()
    _, _, fits, its, ms = bt.cpals([u for _, _, u in tens], R, max_iters=iters, precision=B.F32)
    wall = time.perf_counter() - t0
    bt.free()
    for d in devs:
        d.free()
    fits, plain_fits = np.asarray(fits), np.asarray(plain_fits)
    for k in np.nonzero(~np.isfinite(fits) | ~np.isfinite(plain_fits))[0]:
        log("tensor %d %s: batch fit %r (%d iterations), plain fit %r"
            % (k, tens[k][0].dims, fits[k], its[k], plain_fits[k]))
    dfit = float(np.max(np.abs(fits - plain_fits)))
    log("batched: %.3f ms per sweep of all tensors (plain per-tensor sum %.3f ms), max fit diff %.2e"
        % (ms, plain_ms, dfit))
    out = {
        "metric": "CP-ALS seconds per iteration over the batched set (all tensors), 1 GPU",
        "value": round(ms / 1e3, 6), "unit": "s/iter", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 factors/values, f64 solve/lambda/fit",
        "data": "synthetic (seeded counter-based Poisson generator)",
        "config": {"workload": "batched-64", "tensors": len(tens), "nnz_total": total, "rank": R,
                   "parallelism": "single GPU"},
        "plain_s_per_iter": round(plain_ms / 1e3, 6),
        "speedup_vs_plain": round(plain_ms / ms, 3) if ms > 0 else None,
        "max_fit_diff_vs_plain": dfit,
        "host_wall_s": round(wall, 3),
    }
    print(json.dumps(out), flush=True)


# The large set: 4.8M x 1.8M x 1.8M (65 index bits: two BLCO key blocks),
# 1.7e9 unique nonzeros, Zipf(1.2) on every mode (redrawn until unique),
# log1p Poisson values of a rank-20 Gamma(1,1) model, rank 32, seed 1005.
LARGE = dict(dims=(4_800_000, 1_800_000, 1_800_000), nnz=1_700_000_000, zipf=(1.2, 1.2, 1.2),
             gen_rank=20, value_kind=1, rank=32, seed=1005)


def run_large(B, args):
    """FP32 + selector form and plain FP64 form, fully HBM-resident, then the
    FP32 form with the tensor streamed from pinned host memory (device memory
    held by the tensor: two slots of --slot-nnz nonzeros)."""
    os.environ.setdefault("BLCO_PROGRESS", "1")
    W, R = LARGE, LARGE["rank"]
    nnz_req = int(W["nnz"] * args.large_scale)
    t0 = time.perf_counter()
    coo = B.Coo.generate(W["dims"], nnz_req, zipf=W["zipf"], gen_rank=W["gen_rank"],
                         value_kind=W["value_kind"], seed=W["seed"], topup=True)
    nnz = int(coo.nnz)
    log("generated %d nonzeros in %.1f s" % (nnz, time.perf_counter() - t0))
    t0 = time.perf_counter()
    h = B.encode(coo)
    del coo
    log("BLCO encode %.1f s, %d block(s)" % (time.perf_counter() - t0, h.nblocks))
    U0 = B.init_factors(W["seed"], W["dims"], R)
    forms = [("fp32", B.F32, 1, False), ("fp64", B.F64, 0, False), ("bf16", B.BF16, 1, False),
             ("fp32_streamed", B.F32, 1, True)]
    if args.large_forms:
        forms = [f for f in forms if f[0] in args.large_forms.split(",")]
    res = {}
    for name, prec, vdt, stream in forms:
        t0 = time.perf_counter()
        dev = B.DeviceTensor(h, value_dtype=vdt, stream_data=stream, slot_nnz=args.slot_nnz if stream else 0)
        log("%s: device tensor in %.1f s, %.2f GB on the device" % (name, time.perf_counter() - t0,
                                                                     dev.bytes() / 1e9))
        kernel = B.K_SELECT if prec != B.F64 else B.K_AUTO
        r = run_form(B, dev, U0, R, prec, kernel, args.warmup, args.steps)
        dev_bytes = dev.bytes()
        dev.free()
        s_v = 8 if vdt == 0 else 4
        s_f = {B.F64: 8, B.F32: 4, B.BF16: 2}[prec]
        per_mode_ms = r["mttkrp_ms"] / max(r["iters"], 1)
        ab = [algorithmic_bytes(W["dims"], nnz, R, s_v, s_f, m) for m in range(3)]
        res[name] = dict(s_per_iter=r["iter_ms"] / max(r["iters"], 1) / 1e3,
                         mttkrp_ms_per_mode=[round(float(x), 3) for x in per_mode_ms],
                         mttkrp_gbs=round(sum(ab) / (sum(per_mode_ms) * 1e-3) / 1e9, 1),
                         kernels=r["kernels"].tolist(), fit=float(r["fits"][-1]),
                         device_gb=round(dev_bytes / 1e9, 2))
        log("%s: %.4f s/iter, MTTKRP ms/mode %s, kernels %s, fit %.9f" % (
            name, res[name]["s_per_iter"], res[name]["mttkrp_ms_per_mode"], res[name]["kernels"],
            res[name]["fit"]))
    main = res.get("fp32") or next(iter(res.values()))
    out = {
        "metric": "CP-ALS s/iter, 1.7B-nonzero tensor, FP32+learned selector vs plain FP64, 1 GPU",
        "value": round(main["s_per_iter"], 6), "unit": "s/iter", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(main["s_per_iter"] * 1e3, 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 factors/values, f64 solve/lambda/fit",
        "data": "synthetic (seeded counter-based Poisson generator)",
        "config": {"workload": "large-1.7B", "dims": list(W["dims"]), "nnz": nnz, "rank": R,
                   "zipf": list(W["zipf"]), "blocks": int(h.nblocks), "parallelism": "single GPU"},
        "forms": res,
    }
    for low in ("fp32", "bf16"):
        if low in res and "fp64" in res:
            out["speedup_%s_vs_fp64" % low] = round(res["fp64"]["s_per_iter"] / res[low]["s_per_iter"], 3)
            out["fit_diff_%s_vs_fp64" % low] = abs(res[low]["fit"] - res["fp64"]["fit"])
    print(json.dumps(out), flush=True)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def profiled_traffic(workload):
    """HBM bytes per MTTKRP launch for `workload` from the committed rocprofv3
    summary: FETCH_SIZE + WRITE_SIZE (scaled by the copy-kernel calibration it
    records), averaged over the MTTKRP launches of the profiled run (every
    k_mttkrp_* kernel, weighted by its dispatch count), or None when no
    summary covers it."""
    path = os.path.join(ROOT, "profiles", "mttkrp_row_ordered_%s_fp32.txt" % workload)
    if not os.path.exists(path):
        return None
    in_traffic = False
    n_tot, b_tot = 0, 0.0
    for line in open(path):
        if line.startswith("HBM traffic"):
            in_traffic = True
        elif in_traffic and "k_mttkrp_" in line:
            f = line.split()
            try:
                n = int(f[-4])
                b_tot += n * float(f[-1]) * 1e6  # total MB per dispatch
                n_tot += n
            except (ValueError, IndexError):
                return None
    return round(b_tot / n_tot) if n_tot else None


def algorithmic_bytes(dims, nnz, rank, s_v, s_f, mode):
    """Compulsory MTTKRP traffic of one mode: stream the tensor once (8-byte
    coordinate + value), read each input factor once, write the output once."""
    b = nnz * (8 + s_v)
    for m, d in enumerate(dims):
        b += d * rank * s_f
    return b


def run_form(B, dev, U0, rank, precision, kernel, warmup, steps):
    s = B.CpAls(dev, U0, rank, precision=precision, kernel=kernel)
    f0 = s.iterate(warmup) if warmup else np.zeros(0)
    s.reset_times()
    B.synchronize()
    t0 = time.perf_counter()
    fits = s.iterate(steps)
    B.synchronize()
    wall = time.perf_counter() - t0
    tm = s.times()
    s.free()
    return dict(fits=np.concatenate([f0, fits]), wall=wall, **tm)


def cpu_baseline(t, dims, rank, U0, iters, log_fn):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.perf_counter()
    at = O.Alto(dims, t.cidx, t.vals, nprtn=threads)
    log_fn("cpu baseline: ALTO construction %.1f s (%d threads)" % (time.perf_counter() - t0, threads))
    # 1 warm-up iteration then `iters` timed, from the same initial factors
    _, _, fits, its, mts = at.cpd(U0, 1 + iters, 0.0)
    return dict(s_per_iter=float(np.mean(its[1:])), mttkrp_s=float(np.mean(mts[1:])),
                threads=threads, fits=fits)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="nell2-shape-77M",
                    choices=sorted(WORKLOADS) + ["batched-64", "large-1.7B"])
    ap.add_argument("--large-scale", type=float, default=1.0,
                    help="large-1.7B: fraction of the 1.7e9 nonzeros to generate")
    ap.add_argument("--large-forms", default="", help="large-1.7B: comma list of fp32,fp64,bf16,fp32_streamed")
    ap.add_argument("--slot-nnz", type=int, default=1 << 26,
                    help="large-1.7B streamed form: nonzeros per device slot (two slots)")
    ap.add_argument("--cpu-iters", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="also run one device-to-device copy of BYTES (the reference for "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import __graft_entry__ as G
    B = G._import_pkg()
    if B.device_count() < 1:
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    if args.workload == "batched-64":
        run_batched(B, args)
        return
    if args.workload == "large-1.7B":
        run_large(B, args)
        return
    W = WORKLOADS[args.workload]
    dims, R = W["dims"], W["rank"]

    t0 = time.perf_counter()
    coo = B.Coo.generate(dims, W["nnz"], zipf=W["zipf"], gen_rank=W["gen_rank"],
                         value_kind=W["value_kind"], seed=W["seed"], topup=W["zipf"] is not None)
    log("generated %d nonzeros in %.1f s" % (coo.nnz, time.perf_counter() - t0))
    t0 = time.perf_counter()
    h = B.encode(coo)
    log("BLCO encode %.1f s, %d block(s)" % (time.perf_counter() - t0, h.nblocks))
    U0 = B.init_factors(W["seed"], dims, R)
    nnz = coo.nnz

    if args.copy_calibration:
        src = B.DeviceArray((args.copy_calibration // 8,), B.F64)
        dst = B.DeviceArray((args.copy_calibration // 8,), B.F64)
        B._check(B.lib().blco_memcpy_dtod(dst.ptr, src.ptr, src.nbytes))
        log("copy calibration: %d bytes device to device" % src.nbytes)
        src.free()
        dst.free()

    results = {}
    forms = [("fp32", B.F32, 1)] + ([] if args.no_fp64 else [("fp64", B.F64, 0)])
    for name, prec, vdt in forms:
        dev = B.DeviceTensor(h, value_dtype=vdt)
        kernel = B.K_SELECT if prec != B.F64 else B.K_AUTO
        r = run_form(B, dev, U0, R, prec, kernel, args.warmup, args.steps)
        dev.free()
        s_v = 8 if vdt == 0 else 4
        s_f = 8 if prec == B.F64 else 4
        per_mode_ms = r["mttkrp_ms"] / max(r["iters"], 1)
        abytes = [algorithmic_bytes(dims, nnz, R, s_v, s_f, m) for m in range(len(dims))]
        gbs = [ab / (ms * 1e-3) / 1e9 for ab, ms in zip(abytes, per_mode_ms)]
        r.update(per_mode_ms=per_mode_ms, abytes=abytes, mode_gbs=gbs,
                 s_per_iter=r["iter_ms"] / max(r["iters"], 1) / 1e3,
                 mttkrp_gbs=sum(abytes) / (sum(per_mode_ms) * 1e-3) / 1e9)
        results[name] = r
        log("%s: %.4f s/iter (host %.4f), MTTKRP ms/mode %s, GB/s %s, kernels %s, final fit %.9f" % (
            name, r["s_per_iter"], r["wall"] / max(args.steps, 1),
            np.round(per_mode_ms, 3).tolist(), np.round(gbs, 1).tolist(), r["kernels"].tolist(),
            r["fits"][-1]))

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(coo, dims, R, U0, args.cpu_iters, log)
        log("cpu baseline: %.3f s/iter (MTTKRP %.3f s)" % (cpu["s_per_iter"], cpu["mttkrp_s"]))

    main_r = results["fp32"]
    out = {
        "metric": METRIC,
        "value": round(main_r["s_per_iter"], 6),
        "unit": "s/iter",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_r["s_per_iter"] * 1e3, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 factors/values, f64 solve/lambda/fit",
        "data": "synthetic (seeded counter-based Poisson generator)",
        "config": {"workload": args.workload, "dims": list(dims), "nnz": int(nnz), "rank": R,
                   "zipf": list(W["zipf"]) if W["zipf"] else None,
                   "kernel_choice": ("learned selector" if B.selector_installed()
                                     else "-k 10 rule (no selector installed)"),
                   "kernels_per_mode_fp32": main_r["kernels"].tolist(),
                   "parallelism": "single GPU"},
        "mttkrp_gbs_fp32": round(main_r["mttkrp_gbs"], 1),
        "mttkrp_ms_per_mode_fp32": [round(x, 4) for x in main_r["per_mode_ms"]],
        "fit_fp32": float(main_r["fits"][-1]),
        "roofline": {"bound": "hbm", "achieved": round(main_r["mttkrp_gbs"], 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(main_r["mttkrp_gbs"] / HBM_PEAK_GBS, 4),
                     "traffic": (profiled_traffic(args.workload)
                                 if list(main_r["kernels"]) == [14] * len(dims) else None),
                     "traffic_source": "profiles/mttkrp_row_ordered_%s_fp32.txt (FETCH_SIZE x2.000 + "
                                       "WRITE_SIZE x1.000, calibrated on a 1 GiB copy; mean over the MTTKRP launches)" % args.workload,
                     "kernel": "MTTKRP launches of one FP32 iteration (kernel ids %s per mode); "
                               "algorithmic bytes nnz*(8+4) + sum_m I_m*R*4 per launch"
                               % main_r["kernels"].tolist()},
    }
    if "fp64" in results:
        r64 = results["fp64"]
        out.update({"s_per_iter_fp64": round(r64["s_per_iter"], 6),
                    "speedup_fp32_vs_fp64": round(r64["s_per_iter"] / main_r["s_per_iter"], 3),
                    "mttkrp_gbs_fp64": round(r64["mttkrp_gbs"], 1),
                    "mttkrp_ms_per_mode_fp64": [round(x, 4) for x in r64["per_mode_ms"]],
                    "fit_fp64": float(r64["fits"][-1]),
                    "fit_diff_fp32_vs_fp64": float(abs(main_r["fits"][-1] - r64["fits"][-1]))})
    if cpu is not None:
        out["cpu_baseline"] = {"value": round(cpu["s_per_iter"], 4), "unit": "s/iter",
                               "cores": cpu["threads"], "kind": "port",
                               "sample": "oracle ALTO CPU CP-ALS (OpenMP) on the full tensor, "
                                         "rank 32 FP64, mean of %d iterations after 1 warm-up"
                                         % args.cpu_iters,
                               "mttkrp_s_per_iter": round(cpu["mttkrp_s"], 4)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
