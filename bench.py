"""Benchmark of the memoized CP-ALS sweep on one MI355X.

Workload (the configuration the metric is quoted on): 5-mode tensor
100000 x 50000 x 20000 x 5000 x 1000, 100M coordinates drawn per mode from
Zipf(1.1), lognormal(0,1) values, duplicates summed, rank 32.  A step is one
ALS sweep (N memoised MTTKRPs, Grams, (hadamard of Grams)^-1, column
normalisation, fit) captured in one hipGraph.

Printed: one JSON line whose value is ms per sweep of the learned plan (the
trained cost model's choice, adatm_amd/planner/costmodel.bin; the analytic
plan when there is no model) with FP32 factors and memo and FP64 solve/fit,
next to the plain analytic plan swept in FP64; the roofline of the dominant
MTTKRP launch (algorithmic bytes over its duration, measured here with HIP
events on the engine's stream); and the C/OpenMP FP64 oracle timed on the
host cores on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = ("ms/CP-ALS sweep (full MTTKRP seq), learned vs plain plan; "
          "MTTKRP HBM GB/s")
KC_NAMES = {0: "k_reduce (leaf SpTTM)", 1: "k_reduce (level Hadamard-reduce)",
            2: "k_scatter (memo reuse)", 3: "dense ALS steps",
            4: "k_fix_chunk/k_fix_multi (split-parent merge)"}
MTTKRP_CLASSES = (0, 1, 2, 4)
WORKLOAD_TEXT = {1: "uniform coordinates, U(0,1) values",
                 2: "Zipf(1.1) coordinates per mode, lognormal values",
                 3: "100000 Gaussian clusters sigma 20, U(-1,1) values"}


def launch_name(cls: int, info: int, leaf: bool = True) -> str:
    if cls in (0, 1):
        return (f"k_reduce<float, float, {'true' if cls == 0 else 'false'}, {info}> "
                f"({'leaf SpTTM' if cls == 0 else 'level Hadamard-reduce'}, "
                f"{info} level(s) fused)")
    if cls == 2:
        return f"k_scatter<float, float, {info}, ...> (memo reuse at level {info})"
    return KC_NAMES[cls]


def template_name(cls: int, info: int, nmodes: int) -> str:
    """The kernel's template instantiation as rocprofv3 names it (FP32)."""
    if cls in (0, 1):
        return f"k_reduce<float, float, {'true' if cls == 0 else 'false'}, {info}>"
    return f"k_scatter<float, float, {info}, {'true' if info == nmodes - 1 else 'false'}>"


def measured_traffic(config: int, kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary of this workload (FETCH_SIZE + WRITE_SIZE, copy-calibrated)."""
    path = os.path.join(ROOT, "profiles", f"c{config}_hbm_traffic.md")
    try:
        for line in open(path):
            if line.startswith("TRAFFIC_JSON "):
                t = json.loads(line[len("TRAFFIC_JSON "):]).get(kernel)
                return (t["fetch_bytes"] + t["write_bytes"]) if t else None
    except OSError:
        return None
    return None


def finite(x):
    """the JSON line holds no NaN/Inf (strict parsers reject them): a
    non-finite number becomes null"""
    if isinstance(x, float) and not math.isfinite(x):
        return None
    if isinstance(x, dict):
        return {k: finite(v) for k, v in x.items()}
    if isinstance(x, (list, tuple)):
        return [finite(v) for v in x]
    return x


PREC_NAME = {0: "fp64", 1: "fp32", 2: "bf16rows"}


def root_pass_name(flags: int, ngrp: int) -> str:
    """the root-pass form of every memo group"""
    import adatm_amd as A
    if flags & A.ENGINE_STAGED:
        return "staged (a pass per memo level)"
    return ", ".join(f"g{g}: " + ("staged" if flags & A.ENGINE_STAGED_GROUP(g) else "fused")
                     for g in range(ngrp))


def root_flags(choice: str, tiles) -> int:
    """engine creation flags of the root passes: the explicit choice, else
    the autotuner's (the fourth entry of its tuple)"""
    import adatm_amd as A
    if choice:
        return {"fused": 0, "staged": A.ENGINE_STAGED}[choice]
    return int(tiles[3]) if len(tiles) > 3 else 0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--rank", type=int, default=32)
    ap.add_argument("--ndraws", type=int, default=0, help="override draws (0 = full)")
    ap.add_argument("--cpu-draws", type=int, default=0,
                    help="time the CPU oracle on this many draws (0 = the whole workload)")
    ap.add_argument("--cpu1-draws", type=int, default=2_000_000,
                    help="draws of the single-thread CPU oracle run (0 = skip it)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--root-pass", default="", choices=["", "fused", "staged"],
                    help="root MTTKRP of every memo group: one leaf-to-root pass writing the "
                         "memo levels beside (fused) or passes stopping at every memo level "
                         "(staged); default: the autotuner's choice per group")
    ap.add_argument("--no-convert", action="store_true",
                    help="skip timing splatt_csf_convert on the workload")
    ap.add_argument("--model", default="", help="cost model file (default: the trained one)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="default work-item tiles instead of the autotuner's")
    ap.add_argument("--autotune", action="store_true",
                    help="run the autotuner instead of using its recorded choice for the "
                         "named workload (adatm_amd/planner/tuned_tiles.json)")
    ap.add_argument("--tiles", default="", metavar="RED,SCAT[,BATCH[,FLAGS]]",
                    help="use these work-item tiles (reduction batch, engine creation flags) "
                         "for every engine (no autotuning)")
    ap.add_argument("--instances", type=int, default=64, help="batched workload (--config 5)")
    ap.add_argument("--iters", type=int, default=50, help="batched workload sweeps")
    ap.add_argument("--no-persistent", action="store_true",
                    help="batched workload: forked branches only (no grouped or persistent form)")
    ap.add_argument("--no-sequential", action="store_true",
                    help="batched workload: skip timing the instances one after another")
    ap.add_argument("--hw-queues", type=int, default=0,
                    help="batched workload: GPU_MAX_HW_QUEUES for this process (hardware "
                         "queues the instances' graph branches spread over; 0 = the HIP "
                         "runtime's default, 4), at most 32")
    ap.add_argument("--bf16", action="store_true",
                    help="also sweep with BF16 factor-row storage (FP32 accumulation)")
    ap.add_argument("--copy-cal", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES (FP32 elementwise "
                         "kernel, 3 times): the reference point that scales the "
                         "FETCH_SIZE / WRITE_SIZE counters to bytes")
    args = ap.parse_args()
    if args.copy_cal:
        return copy_calibration(args.copy_cal)
    if args.config == 5:
        return batched_c5(args)

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not "
              "supported (multi-GPU sharding by root subtrees is designed, not built)",
              file=sys.stderr)
        return 2

    import numpy as np
    import adatm_amd as A

    t0 = time.time()
    spec = A.named_spec(args.config, args.ndraws)
    t = A.Tensor.synth(spec)
    ndraws = t.nnz
    t.canonicalize()
    log(f"tensor: {ndraws} draws -> {t.nnz} nonzeros, dims {t.dims} "
        f"({time.time() - t0:.1f}s)")
    R = args.rank
    t1 = time.time()
    cands, ch = A.plan_analytic(t, R, ref_quirks=True, device=0)
    plan = cands[ch]
    log(f"analytic plan {ch}: {plan.groups(t.nmodes)} ({time.time() - t1:.1f}s)")
    model_path = args.model or os.path.join(ROOT, "adatm_amd", "planner", "costmodel.bin")
    learned = None
    if os.path.exists(model_path):
        t1 = time.time()
        model = A.Mlp(device=0, path=model_path)
        lc, lch, pred = A.plan_learned(t, R, model, device=0)
        model.close()
        learned = (lch, lc[lch])
        log(f"learned plan {lch}: {lc[lch].groups(t.nmodes)} (predicted log10 ms "
            f"{[round(float(x), 3) for x in pred]}, {time.time() - t1:.1f}s)")
    else:
        log(f"no trained cost model at {model_path}: learned plan = analytic plan")
    lplan_idx, lplan = learned if learned else (ch, plan)

    runs = [("learned_fp32", lplan, A.FP32)]
    if args.bf16:
        runs.append(("learned_bf16", lplan, A.BF16ROWS))
    if lplan_idx != ch:
        runs.append(("analytic_fp32", plan, A.FP32))
    if not args.no_fp64:
        runs.append(("plain_fp64", plan, A.FP64))
    res = {}
    Mref = {}
    tile_src = {}
    tuned_table = {}
    try:
        with open(os.path.join(ROOT, "adatm_amd", "planner", "tuned_tiles.json")) as f:
            tuned_table = json.load(f)
    except OSError:
        pass
    for name, pl, prec in runs:
        t2 = time.time()
        tiles = (0, 0, 0)
        pidx = lplan_idx if pl is lplan else ch
        tuned = tuned_table.get(f"C{args.config}/R{R}/plan{pidx}/{PREC_NAME[prec]}")
        if args.tiles:
            tiles = tuple(int(x) for x in args.tiles.split(","))
            tile_src[name] = "--tiles"
        elif tuned and not args.autotune and not args.ndraws:
            tiles = tuple(tuned[:4])
            tile_src[name] = "adatm_amd/planner/tuned_tiles.json (autotuner, adatm_amd/tune_named.py)"
            log(f"{name}: tuned tiles {tiles} (autotuner's {tuned[4]:.3f} ms/sweep)")
        elif not args.no_autotune:
            tiles, tms = A.autotune_tiles(t, R, pl, prec, 0, 3)
            tile_src[name] = "autotuned in this run"
            log(f"{name}: autotuned tiles {tiles} ({tms:.3f} ms/sweep, {time.time() - t2:.1f}s)")
        e = A.Engine(t, R, pl, precision=prec, device=0, tiles=tiles[:3],
                     flags=root_flags(args.root_pass, tiles))
        tiles = (tuple(tiles[:3]) + (0, 0, 0))[:3] + (root_flags(args.root_pass, tiles),)
        e.init_factors(spec.seed + 1)
        log(f"{name} engine built ({time.time() - t2:.1f}s)")
        if args.warmup > 0:
            e.run_timed(1, args.warmup, graph=True)
        ms = e.run_timed(1, args.steps, graph=True)
        seq_ms, _ = e.time(0, 1, 5, per_kernel=False)
        _, pk = e.time(1, 0, 1, graph=False)
        launch_log = e.launches()
        F, seq_bytes, seq_flops, dev_bytes = e.stats()
        # the fit of a CP-ALS run of `steps` sweeps from the seeded factors and
        # the worst status of its solves (0 clean, 1 with a ridge, 2 failed)
        e.init_factors(spec.seed + 1)
        try:
            cfit = float(e.cpd(args.steps, tol=0.0, graph=True)[-1])
        except A.AdatmError as exc:
            log(f"{name}: CP-ALS from the seeded factors: {exc}")
            cfit = float("nan")
        status = e.solve_status()
        res[name] = dict(ms_per_sweep=ms / args.steps, fit_last=cfit, solve_status=status,
                         tiles=list(tiles),
                         root_pass=root_pass_name(tiles[3], pl.n_grp),
                         seq_ms=seq_ms, pk=pk, seq_bytes=seq_bytes,
                         seq_flops=seq_flops, dev_bytes=dev_bytes,
                         nfibs=F.tolist(), launches=launch_log)
        log(f"{name}: {ms / args.steps:.3f} ms/sweep, MTTKRP seq {seq_ms:.3f} ms, "
            f"fit {cfit:.6g} (solve status {status}), device {dev_bytes / 2**30:.1f} GiB")
        # MTTKRP of every mode from the same factors, to compare precisions
        e.init_factors(spec.seed + 1)
        Mref[name] = [e.mttkrp(m) for m in range(t.nmodes)]
        e.close()
    errs = errs_bf16 = None
    if "plain_fp64" in Mref:
        def rel(name):
            return [float(np.linalg.norm(a - b) / np.linalg.norm(b))
                    for a, b in zip(Mref[name], Mref["plain_fp64"])]
        errs = rel("learned_fp32")
        if "learned_bf16" in Mref:
            errs_bf16 = rel("learned_bf16")
    Mref.clear()

    r32 = res["learned_fp32"]
    r64 = res.get("plain_fp64")
    pk = r32["pk"]
    times = {c: pk[c] for c in range(5)}
    # dominant single launch of the MTTKRP kernels (one instrumented sweep)
    mk = [x for x in r32["launches"] if x[0] in (0, 1, 2)]
    dcls, dinfo, avg_ms, bytes_per_launch = max(mk, key=lambda x: x[2])
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    mttkrp_ms = sum(times[c] for c in MTTKRP_CLASSES)
    mttkrp_bytes = sum(pk[16 + c] for c in MTTKRP_CLASSES)
    traffic = measured_traffic(args.config, template_name(dcls, dinfo, t.nmodes))

    convert_s = None if args.no_convert else splatt_convert_seconds(A, t)
    if convert_s is not None:
        log(f"splatt_csf_convert (device plan and CSF build): {convert_s:.2f}s")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args, R, A, t)

    out = {
        "metric": METRIC,
        "value": round(r32["ms_per_sweep"], 4),
        "unit": "ms/sweep",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r32["ms_per_sweep"], 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (fp64 solve/fit)",
        "data": "synthetic (seeded Philox generator, no dataset)",
        "config": {
            "workload": (f"C{args.config}: {t.nmodes}-mode "
                         f"{'x'.join(str(d) for d in spec.dims[:t.nmodes])}, "
                         f"{ndraws} draws -> {t.nnz} nnz ({WORKLOAD_TEXT.get(args.config, '')}), "
                         f"R={R}, one hipGraph per ALS sweep"),
            "nnz": t.nnz, "rank": R,
            "planner": "learned (MLP cost model)" if learned else "analytic (no trained model)",
            "learned_plan": lplan.groups(t.nmodes), "learned_plan_index": lplan_idx,
            "analytic_plan": plan.groups(t.nmodes), "analytic_plan_index": ch,
            "tiles": {k: v["tiles"] for k, v in res.items()},
            "tiles_source": tile_src,
            "root_pass": {k: v["root_pass"] for k, v in res.items()},
        },
        "roofline": {
            "bound": "hbm", "kernel": launch_name(dcls, dinfo),
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": (f"profiles/c{args.config}_hbm_traffic.md: L2-miss (fabric) "
                               f"bytes, rocprofv3 --pmc FETCH_SIZE + WRITE_SIZE scaled by a "
                               f"copy kernel's; includes Infinity Cache hits, so an upper bound "
                               f"on HBM bytes" if traffic is not None else None),
            "bytes_per_launch": bytes_per_launch, "avg_launch_ms": avg_ms,
        },
        "cpu_baseline": cpu,
        "plain_fp64_ms_per_sweep": round(r64["ms_per_sweep"], 4) if r64 else None,
        "analytic_fp32_ms_per_sweep": (round(res["analytic_fp32"]["ms_per_sweep"], 4)
                                       if "analytic_fp32" in res else
                                       round(r32["ms_per_sweep"], 4)),
        "speedup_learned_fp32_vs_plain_fp64": (round(r64["ms_per_sweep"] / r32["ms_per_sweep"], 3)
                                               if r64 else None),
        "mttkrp_seq_ms": round(r32["seq_ms"], 4),
        "mttkrp_GBps": round(mttkrp_bytes / (mttkrp_ms * 1e-3) / 1e9, 1),
        "mttkrp_hbm_frac": round(mttkrp_bytes / (mttkrp_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "fit_learned_fp32": r32["fit_last"],
        "solve_status": {k: v["solve_status"] for k, v in res.items()},
        "fit_plain_fp64": r64["fit_last"] if r64 else None,
        "mttkrp_rel_err_fp32_vs_fp64": errs,
        "splatt_csf_convert_s": round(convert_s, 3) if convert_s is not None else None,
        "bf16_rows_ms_per_sweep": (round(res["learned_bf16"]["ms_per_sweep"], 4)
                                   if "learned_bf16" in res else None),
        "fit_bf16_rows": res["learned_bf16"]["fit_last"] if "learned_bf16" in res else None,
        "mttkrp_rel_err_bf16_rows_vs_fp64": errs_bf16,
        "kernel_ms": {KC_NAMES[c]: round(times[c], 4) for c in range(5)},
        "launches_ms": [[launch_name(c, i), round(ms, 4), round(b / 1e9, 4)]
                        for c, i, ms, b in r32["launches"] if c != 3],
    }
    print(json.dumps(finite(out), allow_nan=False), flush=True)
    return 0


def splatt_convert_seconds(A, t) -> float:
    """Wall time of the SPLATT API's splatt_csf_convert on the workload's
    coordinates (64-bit indices, default options: the analytic plan with
    device fiber counts, every group's CSF built on the device and handed
    back as SPLATT arrays)."""
    import ctypes as C
    import numpy as np

    L = A.lib()
    L.splatt_default_opts.restype = C.POINTER(C.c_double)
    L.splatt_csf_convert.argtypes = [C.c_uint64, C.c_uint64, C.POINTER(C.POINTER(C.c_uint64)),
                                     C.POINTER(C.c_double), C.POINTER(C.c_void_p),
                                     C.POINTER(C.c_double)]
    L.splatt_free_csf.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    L.splatt_free_opts.argtypes = [C.POINTER(C.c_double)]
    cols = [np.ascontiguousarray(t.ind(m), dtype=np.uint64) for m in range(t.nmodes)]
    ip = (C.POINTER(C.c_uint64) * t.nmodes)(
        *[c.ctypes.data_as(C.POINTER(C.c_uint64)) for c in cols])
    v = np.ascontiguousarray(t.vals, dtype=np.float64)
    opts = L.splatt_default_opts()
    out = C.c_void_p()
    t0 = time.perf_counter()
    rc = L.splatt_csf_convert(t.nmodes, t.nnz, ip, v.ctypes.data_as(C.POINTER(C.c_double)),
                              C.byref(out), opts)
    sec = time.perf_counter() - t0
    if rc != A.SPLATT_SUCCESS:
        log(f"splatt_csf_convert failed: {L.adatm_last_error().decode()}")
        sec = None
    else:
        L.splatt_free_csf(out, opts)
    L.splatt_free_opts(opts)
    return sec


C5_SEED = 0xADA7A005


def c5_spec(k: int):
    """Instance k of the batched workload: order U{4..6}, dims log-uniform
    1e2..1e5, draws log-uniform 1e4..1e6, uniform or Zipf(1.1) coordinates,
    U(0,1) values, rank in {8, 16, 32}; a pure function of k."""
    import math
    import numpy as np
    import adatm_amd as A

    rng = np.random.Generator(np.random.PCG64([C5_SEED, k]))
    N = int(rng.integers(4, 7))
    dims = np.exp(rng.uniform(math.log(1e2), math.log(1e5), N)).astype(np.int64)
    draws = int(np.exp(rng.uniform(math.log(1e4), math.log(1e6))))
    zipf = bool(rng.integers(0, 2))
    R = int((8, 16, 32)[rng.integers(0, 3)])
    spec = A.make_spec(list(dims), draws, dist=A.DIST_ZIPF if zipf else A.DIST_UNIFORM,
                       zipf_s=1.1, values=A.VALS_U01, seed=C5_SEED + k)
    return spec, R


def batched_c5(args) -> int:
    """64 uneven CP-ALS instances: analytic plans, autotuned tiles per
    instance, FP32 with FP64 solves, every instance's sweep in one hipGraph
    (one launch per batched sweep) whose rounds run all instances' MTTKRPs
    as grouped launches (one per depth and kind of stage: the value), beside
    the same batch with one persistent kernel launch per round and with one
    forked branch of launches per instance, the same engines run one
    instance after another, and FP64 fits of the same instances."""
    if args.hw_queues:
        # read by the HIP runtime when it starts: set before the library loads
        os.environ["GPU_MAX_HW_QUEUES"] = str(min(max(args.hw_queues, 1), 32))
    import numpy as np
    import adatm_amd as A

    n, iters = args.instances, args.iters
    t0 = time.time()
    eng, e64, tiles, seq_bytes = [], [], [], 0.0
    nnz_total = 0
    for k in range(n):
        spec, R = c5_spec(k)
        t = A.Tensor.synth(spec)
        t.canonicalize()
        cands, ch = A.plan_analytic(t, R, device=0)
        tl = (0, 0, 0) if args.no_autotune else A.autotune_tiles(t, R, cands[ch], A.FP32, 0, 2)[0]
        e = A.Engine(t, R, cands[ch], precision=A.FP32, tiles=tl)
        e.init_factors(spec.seed + 1)
        if not args.no_fp64:
            f = A.Engine(t, R, cands[ch], precision=A.FP64)
            f.init_factors(spec.seed + 1)
            e64.append(f)
        seq_bytes += e.stats()[1]
        nnz_total += t.nnz
        eng.append(e)
        tiles.append(tl)
    log(f"C5: {n} instances, {nnz_total} nnz in total, built and tuned in "
        f"{time.time() - t0:.1f}s")
    b = A.Batch(eng)

    def run(form):
        b.set_form(form)
        b.cpd(2)                               # warm-up (graph capture)
        for k, e in enumerate(eng):            # restart from the seeded factors
            e.init_factors(c5_spec(k)[0].seed + 1)
        return b.cpd(iters), b.form

    br_ms = br_same = pe_ms = pe_same = None
    if not args.no_persistent:
        # the forked branches and the persistent launch first, then the
        # grouped launches (the default form: the value)
        (br_ms, br_fits), _ = run(A.FORM_FORKED)
        (pe_ms, pe_fits), _ = run(A.FORM_PERSISTENT)
    (ms, fits), form = run(A.FORM_FORKED if args.no_persistent else A.FORM_GROUPED)
    if br_ms is not None:
        br_same = bool(np.array_equal(fits, br_fits))
        pe_same = bool(np.array_equal(fits, pe_fits))
        log(f"C5: grouped launches {ms / iters:.3f} ms/sweep (fits equal to the forked "
            f"branches': {br_same}, to the persistent launch's: {pe_same}), forked "
            f"branches {br_ms / iters:.3f}, persistent launch {pe_ms / iters:.3f} ms/sweep")
    seq_ms = None if args.no_sequential else b.sequential(2) / 2.0
    dfit = None
    if e64:
        fits64 = np.array([f.cpd(iters, tol=0.0, graph=True)[-1] for f in e64])
        dfit = float(np.max(np.abs(fits[:, -1] - fits64)))
    per = ms / iters
    out = {
        "metric": f"ms per batched CP-ALS sweep ({n} uneven instances, one hipGraph "
                  "launch); aggregate MTTKRP GB/s",
        "value": round(per, 4), "unit": "ms/sweep", "n_gpus": 1, "steps": iters,
        "warmup": 2, "ms_per_step": round(per, 4), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32 (fp64 solve/fit)",
        "data": "synthetic (seeded Philox generator, no dataset)",
        "config": {"workload": f"C5: {n} tensors, order 4-6, dims log-U[1e2,1e5], draws "
                               "log-U[1e4,1e6], uniform or Zipf(1.1), R in {8,16,32}, "
                               "analytic plans, autotuned tiles",
                   "nnz_total": nnz_total,
                   "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")),
                   "tiles": [list(x) for x in tiles]},
        "batched_form": {A.FORM_FORKED: "forked branches", A.FORM_PERSISTENT: "persistent launch",
                         A.FORM_GROUPED: "grouped launches"}[form],
        "forked_branches_ms_per_sweep": round(br_ms / iters, 4) if br_ms else None,
        "persistent_launch_ms_per_sweep": round(pe_ms / iters, 4) if pe_ms else None,
        "fits_equal_forked": br_same,
        "fits_equal_persistent": pe_same,
        "sequential_ms_per_sweep": round(seq_ms, 4) if seq_ms else None,
        "batching_speedup": round(seq_ms / per, 3) if seq_ms else None,
        "aggregate_mttkrp_GBps": round(seq_bytes / (per * 1e-3) / 1e9, 1),
        "max_abs_fit_diff_fp32_vs_fp64": dfit,
    }
    print(json.dumps(finite(out), allow_nan=False), flush=True)
    b.close()
    return 0


def copy_calibration(nbytes: int) -> int:
    """A plain read-once / write-once kernel of known size under the same
    counters: FETCH_SIZE and WRITE_SIZE of one dispatch, divided by nbytes,
    give the factor that turns the counters into bytes."""
    import torch

    n = nbytes // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 4, "bytes_written": n * 4,
                      "kernel": "torch elementwise mul (read n, write n)"}), flush=True)
    return 0


def cpu_baseline(args, R, A, t=None):
    """The oracle's C/OpenMP FP64 restatement of the reference CPU path, timed
    on this host's cores: the whole workload (the canonical tensor `t` the GPU
    legs used, or the first --cpu-draws draws of the same generator when that
    is set), one ALS sweep, median of 3 after 1 warm-up, on every core; and a
    single-thread run (the only setting in which the literal reference is
    race-free) on the first --cpu1-draws draws, 1 warm-up + 1 timed sweep."""
    from oracle import corc

    full = A.named_spec(args.config, args.ndraws)
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads = min(threads, int(os.environ.get("OMP_NUM_THREADS", threads)))

    def timed(spec, tensor, nthreads, nrep):
        s = tensor
        if s is None:
            s = A.Tensor.synth(spec)
            s.canonicalize()
        cands, ch = A.plan_analytic(s, R, ref_quirks=True, device=-1)
        t0 = time.time()
        seq = corc.OracleSeq(s.inds(), s.vals, s.dims, cands[ch].groups(s.nmodes), R)
        build_s = time.time() - t0
        U = [A.rand_matrix(spec.seed + 1, m, d, R) for m, d in enumerate(s.dims)]
        sec = seq.time(U, 1, nrep, nthreads)
        seq.close()
        return sec, s.nnz, ch, build_s

    sample = args.cpu_draws if 0 < args.cpu_draws < full.ndraws else 0
    spec = A.named_spec(args.config, sample) if sample else full
    sec, nnz, ch, build_s = timed(spec, None if sample else t, threads, 3)
    log(f"cpu baseline: {nnz} nnz, {sec * 1e3:.1f} ms/sweep on {threads} threads "
        f"(oracle CSF build {build_s:.1f}s)")
    what = (f"the whole workload ({nnz} nnz)" if not sample else
            f"first {sample} draws of the same generator ({nnz} nnz; not scaled)")
    out = {
        "value": round(sec * 1e3, 1), "unit": "ms/sweep",
        "cores": threads, "kind": "port",
        "sample": f"{what}, plan {ch}, one FP64 ALS sweep, median of 3 after 1 warm-up",
    }
    if args.cpu1_draws > 0:
        n1 = min(args.cpu1_draws, full.ndraws)
        sec1, nnz1, ch1, _ = timed(A.named_spec(args.config, n1), None, 1, 1)
        log(f"cpu baseline, 1 thread: {nnz1} nnz, {sec1 * 1e3:.1f} ms/sweep")
        out["single_thread"] = {
            "value": round(sec1 * 1e3, 1), "unit": "ms/sweep", "cores": 1,
            "sample": (f"first {n1} draws of the same generator ({nnz1} nnz), plan {ch1}, "
                       "one FP64 ALS sweep after 1 warm-up (not scaled)")}
    return out


if __name__ == "__main__":
    sys.exit(main())
