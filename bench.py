"""Benchmark: CP-ALS seconds per iteration and MTTKRP HBM GB/s on one MI355X,
the FP32 + learned-selector form (f32 factors/values, f64 solve, lambda and
fit) against the plain FP64 form.

Default workload: the large tensor, 4.8M x 1.8M x 1.8M (65 index bits: two
BLCO key blocks), 1.7e9 unique nonzeros with Zipf(1.2) indices on every
mode, log1p Poisson values of a rank-20 Gamma(1,1) model, seed 1005; rank-32
CP-ALS, fully HBM-resident.  The tensor is generated, encoded and made
resident on the GPU (blco_generate_device: the host generator's tensor,
without a host copy).  The FP64 form runs first; the values are then
converted to f32 in place and the FP32 form runs on the same tensor.

A step is one CP-ALS iteration (MTTKRP + solve + normalisation + Gram for
every mode, then the fit).  W warm-up iterations run untimed (the first
builds the row-ordered copies of kernel 14), then K iterations are timed
with HIP events on the solver stream (and a host clock around them).  The
MTTKRP time of each mode is event-timed on the same stream.

Other workloads: nell2-shape-77M (12092 x 9184 x 28818, 77M nonzeros, with
the per-mode MTTKRP error of the FP32 form against the FP64 oracle),
small-1M, batched-64 (64 small tensors in one persistent kernel),
selector-heldout (MTTKRP of the selector's 110 held-out tensors under the
learned selector against the -k 10 rule).

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import glob
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

# The large set: 4.8M x 1.8M x 1.8M, 1.7e9 unique nonzeros, Zipf(1.2) on
# every mode (redrawn until unique), log1p Poisson values of a rank-20
# Gamma(1,1) model, rank 32, seed 1005.
LARGE = dict(dims=(4_800_000, 1_800_000, 1_800_000), nnz=1_700_000_000, zipf=(1.2, 1.2, 1.2),
             gen_rank=20, value_kind=1, rank=32, seed=1005)

# The batched set: 64 tensors from seed 1003 * 1000 + t, order 3 or 4
# (50/50), mode lengths log-uniform in [100, 5e4], nonzeros log-uniform in
# [1e4, 2e6], uniform indices, Poisson values of a rank-8 Gamma(1,1) model;
# rank 16, FP32 factors with FP64 solve.
BATCH_SET = dict(count=64, seed=1003 * 1000, rank=16)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(dims, nnz, rank, s_v, s_f, mode):
    """Compulsory MTTKRP traffic of one mode: stream the tensor once (8-byte
    coordinate + value), read each input factor once, write the output once."""
    b = nnz * (8 + s_v)
    for d in dims:
        b += d * rank * s_f
    return b


def profiled_traffic(workload):
    """HBM bytes per MTTKRP launch for `workload` from the committed rocprofv3
    summary profiles/mttkrp_*_<workload>_fp32.txt: FETCH_SIZE + WRITE_SIZE
    (scaled by the copy-kernel calibration it records), averaged over the
    MTTKRP launches of the profiled run (every k_mttkrp_* kernel, weighted by
    its dispatch count); None when no summary covers it."""
    paths = sorted(glob.glob(os.path.join(ROOT, "profiles", "mttkrp_*_%s_fp32.txt" % workload)))
    if not paths:
        return None, None
    path = paths[0]
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
                return None, None
    return (round(b_tot / n_tot) if n_tot else None), os.path.relpath(path, ROOT)


def run_form(B, dev, U0, rank, precision, kernel, warmup, steps):
    """CP-ALS on `dev`: the first warm-up iteration timed on the host clock
    (it builds kernel 14's row-ordered copies), the rest untimed, then
    `steps` iterations timed with HIP events (per-mode MTTKRP and whole
    iterations) and a host clock."""
    s = B.CpAls(dev, U0, rank, precision=precision, kernel=kernel)
    t0 = time.perf_counter()
    f0 = s.iterate(1) if warmup else np.zeros(0)
    B.synchronize()
    first = time.perf_counter() - t0
    f1 = s.iterate(warmup - 1) if warmup > 1 else np.zeros(0)
    s.reset_times()
    B.synchronize()
    t0 = time.perf_counter()
    fits = s.iterate(steps)
    B.synchronize()
    wall = time.perf_counter() - t0
    tm = s.times()
    s.free()
    return dict(fits=np.concatenate([f0, f1, fits]), wall=wall, first_iter_s=first, **tm)


def summarize(r, dims, nnz, rank, s_v, s_f):
    per_mode_ms = r["mttkrp_ms"] / max(r["iters"], 1)
    abytes = [algorithmic_bytes(dims, nnz, rank, s_v, s_f, m) for m in range(len(dims))]
    r.update(per_mode_ms=per_mode_ms, abytes=abytes,
             mode_gbs=[ab / (ms * 1e-3) / 1e9 for ab, ms in zip(abytes, per_mode_ms)],
             s_per_iter=r["iter_ms"] / max(r["iters"], 1) / 1e3,
             mttkrp_gbs=sum(abytes) / (sum(per_mode_ms) * 1e-3) / 1e9)
    return r


def copy_calibration(B, nbytes):
    """One device-to-device copy of nbytes (the runtime's copy kernel): the
    reference for scaling rocprofv3's FETCH_SIZE / WRITE_SIZE to bytes."""
    src = B.DeviceArray((nbytes // 8,), B.F64)
    dst = B.DeviceArray((nbytes // 8,), B.F64)
    B._check(B.lib().blco_memcpy_dtod(dst.ptr, src.ptr, src.nbytes))
    B.synchronize()
    log("copy calibration: %d bytes device to device" % src.nbytes)
    src.free()
    dst.free()


def copy_ceiling(B, nbytes=1 << 32, reps=10):
    """Practical HBM ceiling beside the published 8 TB/s: a device-to-device
    copy of nbytes (the runtime's copy kernel), bytes read plus bytes
    written per second, host-timed over reps copies after one untimed."""
    src = B.DeviceArray((nbytes // 8,), B.F64)
    dst = B.DeviceArray((nbytes // 8,), B.F64)
    B._check(B.lib().blco_memcpy_dtod(dst.ptr, src.ptr, src.nbytes))
    t0 = time.perf_counter()
    for _ in range(reps):
        B._check(B.lib().blco_memcpy_dtod(dst.ptr, src.ptr, src.nbytes))
    t = (time.perf_counter() - t0) / reps
    src.free()
    dst.free()
    gbs = 2 * nbytes / t / 1e9
    log("HBM copy ceiling: %.0f GB/s (%d-byte device copy, read + write)" % (gbs, nbytes))
    return {"gbs": round(gbs, 1), "frac_of_peak": round(gbs / HBM_PEAK_GBS, 4),
            "what": "device-to-device copy of %d bytes (runtime copy kernel), bytes read + written per second, "
                    "host-timed mean of %d" % (nbytes, reps)}


def mttkrp_outputs(B, dev, U0, rank, ftype):
    """MTTKRP of every mode from the factors U0 (host f64, converted to
    ftype), copied back as f64."""
    facs = [B.DeviceArray.from_numpy(u, ftype) for u in U0]
    outs = []
    for m in range(len(U0)):
        o = B.DeviceArray((U0[m].shape[0], rank), B.F64 if ftype == B.F64 else B.F32)
        B.mttkrp(dev, m, rank, facs, o, factor_dtype=ftype, kernel=B.K_AUTO if ftype == B.F64 else B.K_SELECT)
        outs.append(o.numpy())
        o.free()
    for f in facs:
        f.free()
    return outs


def rel_err(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def quiet_stdout(fn, *a, **k):
    """Runs fn with file descriptor 1 pointed at stderr (the reference's CPU
    code prints its ALTO diagnostics on stdout, where only the JSON line may
    go), flushing C stdio before the descriptor is restored."""
    import ctypes
    libc = ctypes.CDLL(None)
    sys.stdout.flush()
    saved = os.dup(1)
    os.dup2(2, 1)
    try:
        return fn(*a, **k)
    finally:
        libc.fflush(None)
        os.dup2(saved, 1)
        os.close(saved)


def cpu_baseline(coo, dims, rank, U0, iters, scale, ref_reps, log_fn):
    """The oracle's OpenMP ALTO CP-ALS (a port of include/cpd.hpp) on `coo`:
    one warm-up iteration, then `iters` timed ones, each from U0; and, when
    oracle/_ref is built, the reference's own mttkrp_alto_par on every mode.
    On a sample (scale = full nonzeros / sample nonzeros > 1) the MTTKRP
    time is scaled by `scale` and the rest of the iteration (the ALS step on
    the full-size factors) is taken as measured."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.perf_counter()
    at = quiet_stdout(O.Alto, dims, coo.cidx, coo.vals, nprtn=threads)
    log_fn("cpu baseline: ALTO construction of %d nonzeros %.1f s (%d threads)"
           % (coo.nnz, time.perf_counter() - t0, threads))
    quiet_stdout(at.cpd, U0, 1, 0.0)
    log_fn("cpu baseline: warm-up iteration done")
    its, mts = [], []
    for k in range(iters):
        _, _, _, it, mt = quiet_stdout(at.cpd, U0, 1, 0.0)
        its.append(float(it[0]))
        mts.append(float(mt[0]))
        log_fn("cpu baseline: iteration %d: %.3f s (MTTKRP %.3f s)" % (k, its[-1], mts[-1]))
    it_s, mt_s = float(np.mean(its)), float(np.mean(mts))
    out = dict(s_per_iter=(it_s - mt_s) + mt_s * scale, mttkrp_s=mt_s * scale, sample_s_per_iter=it_s,
               sample_mttkrp_s=mt_s, threads=threads)
    del at
    ref = os.path.join(ROOT, "oracle", "_ref", "libaltoref.so")
    if ref_reps and os.path.exists(ref):
        per_mode = []
        for m in range(len(dims)):
            per_mode.append(quiet_stdout(O.ref_mttkrp_time, dims, coo.cidx, coo.vals, U0, m, threads,
                                         reps=ref_reps) * scale)
            log_fn("cpu baseline: reference mttkrp_alto_par mode %d: %.3f s (scaled)" % (m, per_mode[-1]))
        out["reference_mttkrp_s_per_mode"] = [round(x, 4) for x in per_mode]
    return out


def cpu_baseline_json(cpu, sample_desc):
    d = {"value": round(cpu["s_per_iter"], 4), "unit": "s/iter", "cores": cpu["threads"], "kind": "port",
         "sample": sample_desc, "mttkrp_s_per_iter": round(cpu["mttkrp_s"], 4)}
    if "reference_mttkrp_s_per_mode" in cpu:
        d["reference_mttkrp"] = {"kind": "reference", "s_per_mode": cpu["reference_mttkrp_s_per_mode"],
                                 "what": "the reference's own mttkrp_alto_par (include/alto.hpp:544-550) "
                                         "compiled from its sources into oracle/_ref, same sample, same scaling"}
    return d


def roofline_json(r, kernels, workload, s_v, s_f):
    traffic, src = profiled_traffic(workload)
    return {"bound": "hbm", "achieved": round(r["mttkrp_gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(r["mttkrp_gbs"] / HBM_PEAK_GBS, 4), "traffic": traffic,
            "traffic_source": ("%s (FETCH_SIZE + WRITE_SIZE per MTTKRP launch, calibrated on a device copy "
                               "of known size)" % src) if src else None,
            "kernel": "MTTKRP launches of one FP32 iteration (kernel ids %s per mode); algorithmic bytes "
                      "nnz*(8+%d) + sum_m I_m*R*%d per launch, / HIP-event time per launch" % (kernels, s_v, s_f)}


def run_large(B, args):
    os.environ.setdefault("BLCO_PROGRESS", "1")
    W, R = LARGE, LARGE["rank"]
    dims = W["dims"]
    nnz_req = int(W["nnz"] * args.large_scale)
    t0 = time.perf_counter()
    dev = B.DeviceTensor.generate(dims, nnz_req, zipf=W["zipf"], gen_rank=W["gen_rank"],
                                  value_kind=W["value_kind"], seed=W["seed"], topup=True, value_dtype=B.F64)
    gen_s = time.perf_counter() - t0
    nnz, nblocks = int(dev.nnz), int(dev.nblocks)
    log("generated, encoded and resident on the GPU: %d nonzeros, %d blocks, %.1f s, %.2f GB"
        % (nnz, dev.nblocks, gen_s, dev.bytes() / 1e9))
    U0 = B.init_factors(W["seed"], dims, R)
    if args.copy_calibration:
        copy_calibration(B, args.copy_calibration)
    # (left out of counter runs, whose scaling reads the calibration copy)
    ceiling = None if args.copy_calibration else copy_ceiling(B)
    forms = [f for f in (args.large_forms or "fp64,fp32").split(",") if f]
    res = {}
    out64 = None
    if "fp64" in forms:
        r = summarize(run_form(B, dev, U0, R, B.F64, B.K_AUTO, args.warmup, args.steps), dims, nnz, R, 8, 8)
        r["device_gb"] = dev.bytes() / 1e9
        res["fp64"] = r
        log("fp64: %.4f s/iter (host %.4f), first iteration %.1f s, MTTKRP ms/mode %s, kernels %s, fit %.9f"
            % (r["s_per_iter"], r["wall"] / max(args.steps, 1), r["first_iter_s"],
               np.round(r["per_mode_ms"], 2).tolist(), r["kernels"].tolist(), r["fits"][-1]))
        if not args.no_error:
            out64 = mttkrp_outputs(B, dev, U0, R, B.F64)
    dev.set_value_dtype(B.F32)
    for name, prec in (("fp32", B.F32), ("bf16", B.BF16)):
        if name not in forms:
            continue
        r = summarize(run_form(B, dev, U0, R, prec, B.K_SELECT, args.warmup, args.steps), dims, nnz, R, 4,
                      4 if prec == B.F32 else 2)
        r["device_gb"] = dev.bytes() / 1e9
        res[name] = r
        log("%s: %.4f s/iter (host %.4f), first iteration %.1f s, MTTKRP ms/mode %s, kernels %s, fit %.9f"
            % (name, r["s_per_iter"], r["wall"] / max(args.steps, 1), r["first_iter_s"],
               np.round(r["per_mode_ms"], 2).tolist(), r["kernels"].tolist(), r["fits"][-1]))
    sel = selector_vs_rule(B, dev, U0, R, args, dims, nnz) if "fp32" in res else None
    errs = None
    if out64 is not None and "fp32" in res:
        out32 = mttkrp_outputs(B, dev, U0, R, B.F32)
        errs = [rel_err(a, b) for a, b in zip(out32, out64)]
        log("fp32 MTTKRP relative error per mode vs fp64: %s" % ["%.2e" % e for e in errs])
        del out32
    del out64
    cpu = None
    sample_desc = None
    if not args.no_cpu_baseline:
        step = max(1, nnz // args.cpu_sample_nnz)
        t0 = time.perf_counter()
        sample = dev.sample(step)
        log("cpu baseline: sample of %d nonzeros (every %d-th in BLCO order) in %.1f s"
            % (sample.nnz, step, time.perf_counter() - t0))
        scale = nnz / max(sample.nnz, 1)
        cpu = cpu_baseline(sample, dims, R, U0, args.cpu_iters, scale, args.ref_reps, log)
        sample_desc = ("oracle ALTO CPU CP-ALS (OpenMP, rank 32, FP64) on every %d-th nonzero of the tensor "
                       "(%d nonzeros, same dims; at least 4 per row of every mode, so the ALTO MTTKRP takes "
                       "the partitioned path it takes on the full tensor), mean of %d iteration(s) after 1 "
                       "warm-up; MTTKRP time scaled by %.1f (full / sample nonzeros), ALS step on the full-size "
                       "factors as measured; the sample itself took %.3f s/iter (MTTKRP %.3f s)"
                       % (step, sample.nnz, args.cpu_iters, scale, cpu["sample_s_per_iter"],
                          cpu["sample_mttkrp_s"]))
        log("cpu baseline: %.3f s/iter scaled (MTTKRP %.3f s)" % (cpu["s_per_iter"], cpu["mttkrp_s"]))
        del sample
    streamed = None
    if "fp32_streamed" in forms:
        h = dev.to_host()
        dev.free()
        dev = None
        ds = B.DeviceTensor(h, value_dtype=B.F32, stream_data=True, slot_nnz=args.slot_nnz)
        r = summarize(run_form(B, ds, U0, R, B.F32, B.K_SELECT, 1, max(1, args.steps // 4)), dims, nnz, R, 4, 4)
        streamed = dict(s_per_iter=round(r["s_per_iter"], 4), device_gb=round(ds.bytes() / 1e9, 2),
                        mttkrp_ms_per_mode=[round(float(x), 2) for x in r["per_mode_ms"]])
        log("fp32 streamed: %.4f s/iter, %.2f GB on the device" % (r["s_per_iter"], ds.bytes() / 1e9))
        ds.free()
    if dev is not None:
        dev.free()
    main_name = "fp32" if "fp32" in res else next(iter(res))
    m = res[main_name]
    out = {
        "metric": METRIC, "value": round(m["s_per_iter"], 6), "unit": "s/iter", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(m["s_per_iter"] * 1e3, 4),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 factors/values, f64 solve/lambda/fit",
        "data": "synthetic (seeded counter-based Poisson generator, generated on the GPU)",
        "config": {"workload": "large-1.7B", "dims": list(dims), "nnz": nnz, "rank": R, "zipf": list(W["zipf"]),
                   "blocks": nblocks,
                   "kernel_choice": ("learned selector" if B.selector_installed()
                                     else "-k 10 rule (no selector installed)"),
                   "kernels_per_mode_fp32": m["kernels"].tolist(), "parallelism": "single GPU"},
        "generate_s": round(gen_s, 1),
        "mttkrp_gbs_fp32": round(m["mttkrp_gbs"], 1),
        "mttkrp_ms_per_mode_fp32": [round(float(x), 3) for x in m["per_mode_ms"]],
        "first_iter_s_fp32": round(m["first_iter_s"], 2),
        "device_gb_fp32": round(m["device_gb"], 2),
        "fit_fp32": float(m["fits"][-1]),
        # the committed traffic profile covers the full-size tensor only
        "roofline": roofline_json(m, m["kernels"].tolist(),
                                  "large-1.7B" if args.large_scale == 1.0 else "large-1.7B-scale%g" % args.large_scale,
                                  4, 4),
        "hbm_copy_ceiling": ceiling,
    }
    if "fp64" in res:
        r64 = res["fp64"]
        out.update({"s_per_iter_fp64": round(r64["s_per_iter"], 6),
                    "speedup_fp32_vs_fp64": round(r64["s_per_iter"] / m["s_per_iter"], 3),
                    "mttkrp_gbs_fp64": round(r64["mttkrp_gbs"], 1),
                    "mttkrp_ms_per_mode_fp64": [round(float(x), 3) for x in r64["per_mode_ms"]],
                    "first_iter_s_fp64": round(r64["first_iter_s"], 2),
                    "device_gb_fp64": round(r64["device_gb"], 2),
                    "fit_fp64": float(r64["fits"][-1]),
                    "fit_diff_fp32_vs_fp64": float(abs(m["fits"][-1] - r64["fits"][-1]))})
    if errs is not None:
        out["mttkrp_rel_err_fp32_vs_fp64_per_mode"] = [float("%.3e" % e) for e in errs]
    if "bf16" in res:
        out["s_per_iter_bf16"] = round(res["bf16"]["s_per_iter"], 6)
        out["fit_bf16"] = float(res["bf16"]["fits"][-1])
    if streamed:
        out["fp32_streamed"] = streamed
    if sel is not None:
        out["selector"] = sel
    if cpu is not None:
        out["cpu_baseline"] = cpu_baseline_json(cpu, sample_desc)
    print(json.dumps(out), flush=True)


def selector_vs_rule(B, dev, U0, R, args, dims, nnz):
    """The learned selector's choice per mode (kernel id, lvl3 partitions,
    tiles per wave) beside the -k 10 rule's; when they differ, the FP32 form
    is also timed under the rule."""
    if not B.selector_installed():
        return None
    choice = [list(dev.select_kernel(m, R, B.F32)) for m in range(len(dims))]
    rule = [[dev.auto_kernel(m, R), 0, 0] for m in range(len(dims))]
    out = {"choice_per_mode": choice, "rule_per_mode": rule}
    if any(c[0] != r[0] or c[2] != 0 or (c[0] == B.K_LVL3 and c[1] != 0) for c, r in zip(choice, rule)):
        r = summarize(run_form(B, dev, U0, R, B.F32, B.K_AUTO, args.warmup, args.steps), dims, nnz, R, 4, 4)
        out["s_per_iter_fp32_rule"] = round(r["s_per_iter"], 6)
        out["mttkrp_ms_per_mode_fp32_rule"] = [round(float(x), 3) for x in r["per_mode_ms"]]
        log("fp32 under the -k 10 rule: %.4f s/iter, MTTKRP ms/mode %s"
            % (r["s_per_iter"], np.round(r["per_mode_ms"], 2).tolist()))
    return out


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
    order = np.argsort(-np.abs(fits - plain_fits))
    top = [[int(k), float("%.3e" % abs(fits[k] - plain_fits[k]))] for k in order[:3]]
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
        "largest_fit_diffs": top,  # [tensor index, |batch fit - plain fit|]
        "host_wall_s": round(wall, 3),
    }
    print(json.dumps(out), flush=True)


def heldout_specs():
    """The held-out tensors of the selector's timing set (the training never
    saw them): 400..499 (1e5-1e7 nonzeros) and 530..539 (1e7-1e8), drawn as
    the selector's collection draws them."""
    import importlib.util
    path = os.path.join(ROOT, "blocked-linearized-coordinate_amd", "selector", "collect.py")
    spec = importlib.util.spec_from_file_location("blco_selector_collect", path)
    col = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(col)
    return ([col.tensor_spec(t, 1e5, 1e7) for t in range(400, 500)]
            + [col.tensor_spec(t, 1e7, 1e8) for t in range(530, 540)])


def run_selector_heldout(B, args):
    """MTTKRP time of every mode of the held-out selector tensors (FP32 values
    and factors) under the learned selector (-k 20) against the -k 10 rule
    (lvl3 with 16 partitions, the rule's set-up in the selector's timing
    data), each event-timed over --steps launches after one untimed call."""
    if not B.selector_installed():
        sys.exit("selector-heldout needs the selector file")
    reps = max(args.steps, 1)
    t_start = time.perf_counter()
    tot = {"all": [0.0, 0.0], "1e5-1e7": [0.0, 0.0], "1e7-1e8": [0.0, 0.0]}
    nmodes = changed = 0
    worst = []
    for sp in heldout_specs():
        try:
            dev = B.DeviceTensor.generate(sp["dims"], sp["nnz"], zipf=sp["zipf"], gen_rank=8, value_kind=0,
                                          seed=sp["seed"], topup=False, value_dtype=1)
        except B.BlcoError:  # keys wider than the GPU generator takes
            coo = B.Coo.generate(sp["dims"], sp["nnz"], zipf=sp["zipf"], gen_rank=8, value_kind=0,
                                 seed=sp["seed"], topup=False)
            dev = B.DeviceTensor(B.encode(coo), value_dtype=1)
            del coo
        R = sp["rank"]
        F = []
        for d in sp["dims"]:
            a = B.DeviceArray((d, R), B.F32)
            h = np.ones((d, R), np.float32)
            B._check(B.lib().blco_memcpy_htod(a.ptr, h.ctypes.data, a.nbytes))
            F.append(a)
        rng = "1e7-1e8" if sp["t"] >= 500 else "1e5-1e7"
        line = []
        for m in range(len(sp["dims"])):
            out = B.DeviceArray((sp["dims"][m], R), B.F32)
            ms_r, _ = B.mttkrp_time(dev, m, R, F, out, B.F32, B.K_AUTO, 16, reps=reps)
            ms_s, _ = B.mttkrp_time(dev, m, R, F, out, B.F32, B.K_SELECT, 16, reps=reps)
            ch = list(dev.select_kernel(m, R, B.F32))
            if ch[0] != dev.auto_kernel(m, R) or ch[2] != 0 or (ch[0] == B.K_LVL3 and ch[1] != 16):
                changed += 1
            for k in ("all", rng):
                tot[k][0] += ms_r
                tot[k][1] += ms_s
            worst.append((ms_s / ms_r, sp["t"], m, ch))
            line.append("%.3f/%.3f" % (ms_r, ms_s))
            nmodes += 1
            out.free()
        for a in F:
            a.free()
        log("tensor %d (%s, %d nonzeros, rank %d): rule/selector ms %s"
            % (sp["t"], rng, dev.nnz, R, " ".join(line)))
        dev.free()
    worst.sort(reverse=True)
    res = {k: {"rule_ms": round(v[0], 3), "selector_ms": round(v[1], 3),
               "gain": round(v[0] / v[1] - 1.0, 4) if v[1] > 0 else None} for k, v in tot.items()}
    out = {
        "metric": "MTTKRP ms summed over every mode of the selector's held-out tensors, 1 GPU",
        "value": res["all"]["selector_ms"], "unit": "ms", "n_gpus": 1, "steps": reps, "warmup": 1,
        "ms_per_step": res["all"]["selector_ms"], "higher_is_better": False, "scaling": "strong",
        "vs_baseline": None, "dtype": "f32 values and factors",
        "data": "synthetic (the selector's held-out tensors: seeds 100400400..100400499, 100400530..100400539)",
        "config": {"workload": "selector-heldout", "tensors": 110, "modes": nmodes,
                   "parallelism": "single GPU"},
        "rule_ms": res["all"]["rule_ms"], "gain_vs_rule": res["all"]["gain"], "per_range": res,
        "modes_changed_by_selector": changed,
        "worst_modes": [dict(tensor=t, mode=m, choice=c, ratio=round(r, 3)) for r, t, m, c in worst[:5]],
        "host_wall_s": round(time.perf_counter() - t_start, 1),
    }
    print(json.dumps(out), flush=True)


def run_workload(B, args):
    """nell2-shape-77M or small-1M: host-generated, both forms, the per-mode
    MTTKRP error of the FP32 form against the FP64 oracle, CPU baseline on
    the full tensor."""
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
        copy_calibration(B, args.copy_calibration)
    # (left out of counter runs, whose scaling reads the calibration copy)
    ceiling = None if args.copy_calibration else copy_ceiling(B)
    results = {}
    forms = [("fp32", B.F32, 1)] + ([] if args.no_fp64 else [("fp64", B.F64, 0)])
    errs = None
    sel = None
    for name, prec, vdt in forms:
        dev = B.DeviceTensor(h, value_dtype=vdt)
        kernel = B.K_SELECT if prec != B.F64 else B.K_AUTO
        r = summarize(run_form(B, dev, U0, R, prec, kernel, args.warmup, args.steps), dims, nnz, R,
                      8 if vdt == 0 else 4, 8 if prec == B.F64 else 4)
        if name == "fp32" and not args.no_error:
            out32 = mttkrp_outputs(B, dev, U0, R, B.F32)
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import oracle as O
            at = quiet_stdout(O.Alto, dims, coo.cidx, coo.vals,
                              nprtn=int(os.environ.get("OMP_NUM_THREADS", os.cpu_count())))
            errs = [rel_err(out32[m], quiet_stdout(at.mttkrp, U0, m)) for m in range(len(dims))]
            log("fp32 MTTKRP relative error per mode vs the FP64 oracle: %s" % ["%.2e" % e for e in errs])
        if name == "fp32":
            sel = selector_vs_rule(B, dev, U0, R, args, dims, nnz)
        dev.free()
        results[name] = r
        log("%s: %.4f s/iter (host %.4f), MTTKRP ms/mode %s, GB/s %s, kernels %s, final fit %.9f" % (
            name, r["s_per_iter"], r["wall"] / max(args.steps, 1),
            np.round(r["per_mode_ms"], 3).tolist(), np.round(r["mode_gbs"], 1).tolist(), r["kernels"].tolist(),
            r["fits"][-1]))
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(coo, dims, R, U0, args.cpu_iters, 1.0, args.ref_reps, log)
        log("cpu baseline: %.3f s/iter (MTTKRP %.3f s)" % (cpu["s_per_iter"], cpu["mttkrp_s"]))
    m = results["fp32"]
    out = {
        "metric": METRIC, "value": round(m["s_per_iter"], 6), "unit": "s/iter", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(m["s_per_iter"] * 1e3, 4),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 factors/values, f64 solve/lambda/fit",
        "data": "synthetic (seeded counter-based Poisson generator)",
        "config": {"workload": args.workload, "dims": list(dims), "nnz": int(nnz), "rank": R,
                   "zipf": list(W["zipf"]) if W["zipf"] else None,
                   "kernel_choice": ("learned selector" if B.selector_installed()
                                     else "-k 10 rule (no selector installed)"),
                   "kernels_per_mode_fp32": m["kernels"].tolist(), "parallelism": "single GPU"},
        "mttkrp_gbs_fp32": round(m["mttkrp_gbs"], 1),
        "mttkrp_ms_per_mode_fp32": [round(float(x), 4) for x in m["per_mode_ms"]],
        "fit_fp32": float(m["fits"][-1]),
        "roofline": roofline_json(m, m["kernels"].tolist(), args.workload, 4, 4),
        "hbm_copy_ceiling": ceiling,
    }
    if errs is not None:
        out["mttkrp_rel_err_fp32_vs_fp64_oracle_per_mode"] = [float("%.3e" % e) for e in errs]
    if "fp64" in results:
        r64 = results["fp64"]
        out.update({"s_per_iter_fp64": round(r64["s_per_iter"], 6),
                    "speedup_fp32_vs_fp64": round(r64["s_per_iter"] / m["s_per_iter"], 3),
                    "mttkrp_gbs_fp64": round(r64["mttkrp_gbs"], 1),
                    "mttkrp_ms_per_mode_fp64": [round(float(x), 4) for x in r64["per_mode_ms"]],
                    "fit_fp64": float(r64["fits"][-1]),
                    "fit_diff_fp32_vs_fp64": float(abs(m["fits"][-1] - r64["fits"][-1]))})
    if sel is not None:
        out["selector"] = sel
    if cpu is not None:
        out["cpu_baseline"] = cpu_baseline_json(
            cpu, "oracle ALTO CPU CP-ALS (OpenMP) on the full tensor, rank %d FP64, mean of %d iterations "
                 "after 1 warm-up" % (R, args.cpu_iters))
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="large-1.7B",
                    choices=["large-1.7B", "batched-64", "selector-heldout"] + sorted(WORKLOADS))
    ap.add_argument("--large-scale", type=float, default=1.0,
                    help="large-1.7B: fraction of the 1.7e9 nonzeros to generate")
    ap.add_argument("--large-forms", default="",
                    help="large-1.7B: comma list of fp64,fp32,bf16,fp32_streamed (default fp64,fp32)")
    ap.add_argument("--slot-nnz", type=int, default=0,
                    help="large-1.7B streamed form: nonzeros per device slot (two slots; 0: the "
                         "library's default, 2^24)")
    ap.add_argument("--cpu-iters", type=int, default=0,
                    help="timed CPU iterations (0: 1 on large-1.7B, 2 otherwise)")
    ap.add_argument("--cpu-sample-nnz", type=int, default=20_000_000,
                    help="large-1.7B: nonzeros of the CPU baseline's sample")
    ap.add_argument("--ref-reps", type=int, default=1,
                    help="timed calls of the reference's own mttkrp_alto_par per mode (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-error", action="store_true", help="skip the per-mode FP32 MTTKRP error")
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
    if not args.cpu_iters:
        args.cpu_iters = 1 if args.workload == "large-1.7B" else 2
    if args.workload == "batched-64":
        run_batched(B, args)
    elif args.workload == "selector-heldout":
        run_selector_heldout(B, args)
    elif args.workload == "large-1.7B":
        if args.no_fp64:
            args.large_forms = args.large_forms or "fp32"
        run_large(B, args)
    else:
        run_workload(B, args)


if __name__ == "__main__":
    main()
