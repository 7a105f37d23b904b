#!/usr/bin/env python3
"""Benchmark of the CGFDM dynamic-rupture time step on one MI355X.

One step = one full RK4 time step of the whole job on the GPU: rupture
sensor, and per stage the fault trial (fault_pre, and for rate and state the
FP64 Newton solve fault_rs), the fused wave stage (interior / absorbing-band /
free-surface instantiations + PML auxiliary update), the split-node
derivatives (fault_deriv) and the fault RK (fault_post).  Inputs are
generated with the seeded synthetic generators of dynfd3dcuda_amd.configs and
are resident in HBM before timing starts.

Default workload: the large mixed-precision configuration, 1600 x 1000 x 500
points at 25 m (800 M points): topography, bent fault, 3-layer medium,
20-layer ADE-PML, rate-and-state aging law with the learned Newton initial
guess (the committed network trained on four seeds of the 400x300x200
rate-state workload, none of them the timed one), metric stored as bfloat16
deviations with FP32 media (MetricBF16=2), FP32 wavefield, FP64 friction and
slip/state accumulation.  Prints one JSON line:
  value          grid-point updates per second (pt-steps/s) of the whole step
  roofline       the dominant stage kernel (the stage-kernel class with the
                 largest GPU time in an untimed window of PROFILE_STEPS steps
                 with every class timed), timed with HIP events on the
                 solver's stream around its own launches over the timed
                 region; achieved = algorithmic bytes per launch
                 (points x B_alg/4; B_alg = 768 B/pt-step FP32, 696 B with
                 the deviation-form metric, 656 B with its media palette:
                 the reference RK4 structure's 144 wavefield words and the
                 coefficient words as stored) / average launch time (the
                 kernels move 18 wavefield words fewer, B_moved, since stage
                 1 derives stage 0's accumulator); traffic =
                 HBM bytes per launch from the committed rocprofv3
                 FETCH/WRITE summary of the same workload; stream ceilings =
                 a hand-written streaming kernel's GB/s for read:write mixes;
                 frac_with_aux counts the ADE-PML auxiliary words the kernel
                 reads and advances too (AUX_READ_BYTES, AUX_UPDATE_BYTES
                 per point and face)
  cpu_baseline   the FP64 oracle (C++/OpenMP) on the host cores, on the
                 100 x 1000 x 500 sub-slab around the fault of the same
                 workload (the full grid needs ~300 GB of host memory in FP64)
  fp32_metric    the same workload with the metric and media in FP32
                 (MetricBF16=0): plain vs mixed form on the same GPU
  side           the 400x300x200 slip-weakening workload: throughput and its
                 dominant kernel's roofline
  kernel_ms_per_step  every kernel class's GPU time per step, from the
                 untimed window
  rate_state     the rate-state fault step three ways on the 400x300x200
                 rate-state workload: the reference's Newton rule, plain
                 Newton, and the learned guess (committed network), timed
                 over a window from the nucleation ramp through the rupture
                 front's passage, with the share of solves whose wavefront
                 evaluated the network, the share that started from its
                 guess and the per-wavefront largest-iteration histogram
Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--no-cpu] [--no-rs] ...
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "grid-pt updates/s + HBM roofline %; RS fault step time MLP-Newton vs plain"
PEAK_HBM_GBS = 8000.0  # MI355X HBM3E peak
PROFILE_STEPS = 5      # untimed steps with every kernel class timed (the per-class breakdown)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "stage_traffic_{workload}.txt")
DEFAULT_WORKLOAD = "large_rs_mlp_bf16_1600x1000x500"
SIDE_WORKLOAD = "curved_sw_400x300x200"
CPU_SUB_NX = 100       # fault-normal planes of the CPU baseline's sub-slab of the large workload
NEWTON_MLP = os.path.join(ROOT, "dynfd3dcuda_amd", "data", "newton_guess_4x64x64x1.f32")


def b_alg(params, palette=False, skip=0):
    """Algorithmic bytes per grid-point update (one full RK4 step): 144
    wavefield words + 48 coefficient words (12 per stage: FP32; bfloat16;
    or 9 bfloat16 metric deviations and 3 exact FP32 media words, or with
    a media palette the deviations and a 16-bit palette entry; less the 4
    bytes of each 32-bit word of deviations that is 0 on the whole grid,
    which the stage kernels do not load: bits of skip)."""
    m = params.get("MetricBF16", 0)
    if m == 2:
        return 144 * 4 + 4 * (9 * 2 + (2 if palette else 3 * 4) - 4 * bin(skip).count("1"))
    return 144 * 4 + (48 * 4 if not m else 48 * 2)


# wavefield bytes per interior point-step the stage kernels do not move
# against B_alg's RK4 structure: the accumulator of stage 0 is derived in
# stage 1 from the step start and the stage-0 value, so stage 0 writes 9
# words fewer and stage 1 reads 9 fewer
DERIVED_T_SAVING = 18 * 4

# algorithmic bytes of the ADE-PML auxiliary fields per (point, face) pair
# and RK stage (pmlx.cu:126-210, rk4.cu:154-257): the 9 auxiliary words read
# for the -aux/beta term of the stage; the ODE step's step-start and
# accumulator reads (none in stage 0, 18 words in stages 1 and 2, 9 in stage
# 3: 11.25 on average) and its writes (auxiliary, step-start and
# accumulator words in stage 0, auxiliary and accumulator in 1 and 2, the
# auxiliary words in 3: 18 on average), its own read of the 9 auxiliary
# words being the one above; like B_alg these follow the reference's RK4
# structure (the kernels derive stage 0's accumulator in stage 1 here too,
# 18 words per pair-step fewer)
AUX_READ_BYTES = 36
AUX_UPDATE_BYTES = 117


def workload_name(name, p):
    what = {"flat_sw_100x200x100": "plain flat grid, slip weakening, Cerjan sponge",
            "curved_sw_400x300x200": "curved grid + topography, bent slip-weakening fault, 3-layer medium, "
                                     "20-layer ADE-PML",
            "curved_rs_400x300x200": "curved grid, rate-and-state aging law, plain Newton",
            "curved_rs_mlp_400x300x200": "curved grid, rate-and-state aging law, learned Newton guess",
            "large_rs_mlp_bf16_1600x1000x500": "curved grid + topography, bent fault, 3-layer medium, 20-layer "
                                               "ADE-PML, rate-and-state aging law with the learned Newton guess, "
                                               "metric as BF16 deviations, exact media"}
    return f"{name}: {p['NX']}x{p['NY']}x{p['NZ']} at {p['DH']:g} m, {what.get(name, '')}"


def read_traffic(workload, kernel):
    """HBM bytes of `kernel` per RK stage (one launch of the interior
    class) from the committed rocprofv3 summary."""
    path = TRAFFIC_FILE.format(workload=workload)
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if ":" in line and not line.lstrip().startswith("#"):
            k, v = line.split(":", 1)
            vals[k.strip()] = v.strip()
    v = vals.get(f"{kernel}_hbm_bytes_per_stage")
    try:
        return float(v) if v is not None else None
    except ValueError:
        return None


def hip_runtime():
    """The HIP runtime the solver library runs on (already mapped into this
    process; a second runtime in the same process would find no device)."""
    import ctypes
    for line in open("/proc/self/maps"):
        if "libamdhip64" in line:
            return ctypes.CDLL(line.split()[-1])
    return None


def copy_ceiling(nbytes=1 << 30, reps=10):
    """Device-to-device copy rate (GB/s of bytes read + written) of one
    buffer of `nbytes` with the runtime's copy, timed with HIP events."""
    import ctypes
    hip = hip_runtime()
    if hip is None:
        return None
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    hip.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
    hip.hipFree.argtypes = [vp]
    hip.hipMemset.argtypes = [vp, ctypes.c_int, sz]
    hip.hipMemcpyAsync.argtypes = [vp, vp, sz, ctypes.c_int, vp]
    hip.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
    hip.hipEventRecord.argtypes = [vp, vp]
    hip.hipEventSynchronize.argtypes = [vp]
    hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
    hip.hipEventDestroy.argtypes = [vp]
    a, b, e0, e1 = vp(), vp(), vp(), vp()
    if hip.hipMalloc(ctypes.byref(a), nbytes) or hip.hipMalloc(ctypes.byref(b), nbytes):
        return None
    try:
        hip.hipMemset(a, 1, nbytes)
        hip.hipEventCreate(ctypes.byref(e0))
        hip.hipEventCreate(ctypes.byref(e1))
        for _ in range(2):
            hip.hipMemcpyAsync(b, a, nbytes, 3, None)  # 3 = hipMemcpyDeviceToDevice
        hip.hipEventRecord(e0, None)
        for _ in range(reps):
            hip.hipMemcpyAsync(b, a, nbytes, 3, None)
        hip.hipEventRecord(e1, None)
        hip.hipEventSynchronize(e1)
        ms = ctypes.c_float()
        hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
        hip.hipEventDestroy(e0)
        hip.hipEventDestroy(e1)
    finally:
        hip.hipFree(a)
        hip.hipFree(b)
    return 2.0 * nbytes * reps / (ms.value * 1e-3) / 1e9 if ms.value > 0 else None


def stream_ceilings():
    """GB/s of the library's streaming probe kernel (non-temporal 16-byte
    loads and stores, 1 GiB arrays) for read:write array mixes."""
    import dynfd3dcuda_amd as dyn
    out = {}
    for nr, nw in ((1, 0), (1, 1), (2, 1), (3, 1)):
        out[f"{nr}:{nw}"] = dyn.stream_probe(nr, nw)
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cpu_baseline(case, budget_s, what):
    """Time the FP64 oracle on `case` for a bounded number of steps on the
    host cores.  The oracle is the checker of the test suite and is used here
    only as the reported CPU baseline."""
    from oracle.binding import Oracle
    p = case["params"]
    orc = Oracle(p)
    if "geom" in case:
        orc.set_geometry(*case["geom"])
    if "media1d" in case:
        orc.set_media1d(*case["media1d"])
    if case.get("topo") is not None:
        orc.set_topography(case["topo"])
    for k, v in case.get("stress", {}).items():
        orc.set_stress(k, v)
    t0 = time.perf_counter()
    orc.setup()
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    orc.step(0)  # warm-up
    t1 = time.perf_counter() - t0
    n = int(max(1, min(50, budget_s // max(t1, 1e-3))))
    t0 = time.perf_counter()
    for it in range(1, 1 + n):
        orc.step(it)
    dt = (time.perf_counter() - t0) / n
    npts = p["NX"] * p["NY"] * p["NZ"]
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": npts / dt, "unit": "grid-pt updates/s", "cores": cores, "kind": "port",
            "effective_GBps": npts / dt * b_alg(p) / 1e9,
            "sample": f"FP64 oracle (FP64 metric and wavefield), {what}: {p['NX']}x{p['NY']}x{p['NZ']} points, "
                      f"{n} steps after 1 warm-up ({dt:.2f} s/step; setup {t_setup:.1f} s not timed), "
                      f"{cores} threads, {cpu_model()}"}


def overrides(args):
    """params.json keys given with --set KEY=VALUE."""
    out = {}
    for kv in args.set:
        k, v = kv.split("=", 1)
        out[k] = float(v) if any(ch in v for ch in ".eE") else int(v)
    return out


def timed_run(case, steps, warmup, workload):
    """Set up `case`, run `warmup` steps, an untimed window of PROFILE_STEPS
    steps with every kernel class timed, then `steps` timed steps with HIP
    events around the dominant stage-kernel class only."""
    import numpy as np
    import dynfd3dcuda_amd as dyn
    from dynfd3dcuda_amd import configs
    p = case["params"]
    solver = dyn.Solver(p)
    configs.apply(case, solver)
    t0 = time.perf_counter()
    solver.setup()
    t_setup = time.perf_counter() - t0
    d = solver.dims()
    npts = d["ni"] * d["nj"] * d["nk"]
    print(f"bench: {workload}: setup of {npts} points in {t_setup:.1f} s", file=sys.stderr, flush=True)
    for it in range(warmup):
        solver.step(it)
    solver.synchronize()
    it0 = warmup
    solver.profile(True)
    for it in range(it0, it0 + PROFILE_STEPS):
        solver.step(it)
    solver.synchronize()
    prof_all = solver.profile_report()
    solver.profile(False)
    stage = {k: v for k, v in prof_all.items() if k.startswith("stage_") and v[1] > 0}
    dom = max(stage, key=lambda k: stage[k][0])
    it0 += PROFILE_STEPS
    rs = p.get("Friction_type", 0) >= 1 and p.get("Rupture", 1)
    if rs:
        solver.put("newton_counts", np.zeros(5 * d["nj"] * d["nk"], np.int32))
    solver.profile_classes([dom])
    solver.profile(True)
    t0 = time.perf_counter()
    for it in range(it0, it0 + steps):
        solver.step(it)
    solver.synchronize()
    t = time.perf_counter() - t0
    prof = solver.profile_report()
    solver.profile(False)
    solver.profile_classes(None)
    newton = None
    if rs:
        cnt = solver.get("newton_counts").reshape(5, -1).astype(np.float64).sum(axis=1)
        newton = {"solves": int(cnt[0]), "iters_per_solve": cnt[1] / max(cnt[0], 1), "fallbacks": int(cnt[2]),
                  "learned_start_share": cnt[3] / max(cnt[0], 1), "network_wave_share": cnt[4] / max(cnt[0], 1)}
    finite = bool(np.all(np.isfinite(solver.get("Vs1"))) and np.all(np.isfinite(solver.get("Ts1"))))
    aux = solver.profile_aux()
    palette = solver.media_palette_entries() > 0
    skip = solver.metric_skip_mask()
    solver.close()
    value = npts * steps / t
    ba = b_alg(p, palette, skip)
    ms, n, pts = prof[dom]
    avg_s = ms / n / 1e3
    achieved = pts * ba / 4.0 / avg_s / 1e9  # one launch = one RK stage of its points
    aux_bytes = sum(rd * AUX_READ_BYTES + up * AUX_UPDATE_BYTES for rd, up in aux.values())  # per RK stage
    rd, up = aux.get(dom, (0, 0))
    achieved_aux = (pts * ba / 4.0 + rd * AUX_READ_BYTES + up * AUX_UPDATE_BYTES) / avg_s / 1e9
    return {"value": value, "ms_per_step": 1e3 * t / steps, "npts": npts, "B_alg": ba, "setup_s": t_setup,
            "step_frac": value * ba / (PEAK_HBM_GBS * 1e9), "dom": dom, "achieved": achieved,
            "B_moved": ba - DERIVED_T_SAVING,
            "avg_launch_ms": avg_s * 1e3, "points_per_launch": pts, "traffic": read_traffic(workload, dom),
            "kernel_ms_per_step": {k: v[0] / PROFILE_STEPS for k, v in prof_all.items() if v[1] > 0},
            "stage_points_per_launch": {k: v[2] for k, v in stage.items()}, "newton": newton,
            "aux_pairs_per_launch": aux, "achieved_with_aux": achieved_aux, "palette": palette,
            "metric_words_skipped": bin(skip).count("1"),
            "step_frac_with_aux": (value * ba + 4 * aux_bytes * steps / t) / (PEAK_HBM_GBS * 1e9),
            "fault_outputs_finite": finite}


def rate_state_leg(args, mlp_path):
    """The rate-state solve three ways on the 400x300x200 rate-state
    workload: the reference's stopping rule (NewtonMode 0: |dx|/|x| < 1e-12,
    at most 100 iterations, trial_ratestate.cu:22-57), plain Newton on the
    residual rule (NewtonMode 1) and the learned guess (NewtonMode 2, same
    residual rule and fallback), each timed over the same window of steps."""
    import numpy as np
    import dynfd3dcuda_amd as dyn
    from dynfd3dcuda_amd import configs
    lead, window = args.rs_lead, args.rs_steps
    fault = ("fault_pre", "fault_rs", "fault_deriv", "fault_post")
    res = {}
    for mode in (0, 1, 2):
        case = configs.curved_rs_400x300x200(nt=lead + window + 2)
        case["params"].update({"NewtonMode": mode, "NewtonTol": 1e-10, "NewtonFallbackIter": 20})
        if mode == 2:
            case["params"]["NewtonMLP"] = mlp_path
        case["params"].update(overrides(args))
        s = dyn.Solver(case["params"])
        configs.apply(case, s)
        s.setup()
        d = s.dims()
        n2 = d["nj"] * d["nk"]
        for it in range(lead):
            s.step(it)
        s.synchronize()
        slide0 = int(np.sum(np.hypot(s.get("Vs1"), s.get("Vs2")) > 1e-3))
        s.put("newton_counts", np.zeros(5 * n2, np.int32))
        s.put("newton_wave_hist", np.zeros(32, np.int32))
        s.profile(True)
        for it in range(lead, lead + window):
            s.step(it)
        s.synchronize()
        prof = s.profile_report()
        s.profile(False)
        cnt = s.get("newton_counts").reshape(5, -1).astype(np.float64)
        hist = s.get("newton_wave_hist").astype(np.int64)
        nst = 4 * window
        per_node = cnt[1] / np.maximum(cnt[0], 1)
        solves = max(cnt[0].sum(), 1)
        nz = np.nonzero(hist)[0]
        res[mode] = {"newton_kernel_ms_per_stage": prof["fault_rs"][0] / nst,
                     "fault_ms_per_stage": sum(prof[k][0] for k in fault) / nst,
                     "iters_per_solve": cnt[1].sum() / solves,
                     "max_node_mean_iters": float(per_node.max()),
                     "fallbacks": int(cnt[2].sum()),
                     "learned_start_share": cnt[3].sum() / solves,
                     "network_wave_share": cnt[4].sum() / solves,
                     "wave_max_iters_hist": {int(b): int(hist[b]) for b in nz},
                     "sliding_nodes_first": slide0,
                     "sliding_nodes_last": int(np.sum(np.hypot(s.get("Vs1"), s.get("Vs2")) > 1e-3))}
        s.close()
    ref, pl, ml = res[0], res[1], res[2]
    share = pl["newton_kernel_ms_per_stage"] / pl["fault_ms_per_stage"]
    return {"workload": f"curved_rs_400x300x200 (seed 2): the curved 400x300x200 grid with the aging law, {n2} fault "
                        f"nodes; steps {lead}..{lead + window - 1}, from the nucleation ramp through the rupture "
                        f"front's passage",
            "network": os.path.relpath(mlp_path, ROOT),
            "reference_rule": ref, "plain": pl, "learned": ml,
            "newton_share_of_fault_step": share,
            "amdahl_bound_fault_step_speedup": 1.0 / (1.0 - share),
            "fault_step_speedup_vs_reference_rule": ref["fault_ms_per_stage"] / ml["fault_ms_per_stage"],
            "newton_kernel_speedup": pl["newton_kernel_ms_per_stage"] / ml["newton_kernel_ms_per_stage"],
            "fault_step_speedup": pl["fault_ms_per_stage"] / ml["fault_ms_per_stage"]}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default=DEFAULT_WORKLOAD,
                    help="synthetic workload, one of dynfd3dcuda_amd.configs.WORKLOADS")
    ap.add_argument("--seed", type=int, default=None)
    ap.add_argument("--mlp", default=NEWTON_MLP, help="learned-guess network (float32 blob) for NewtonMode 2")
    ap.add_argument("--set", action="append", default=[], metavar="KEY=VALUE",
                    help="override a params.json key of the workload (repeatable)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=30.0, help="CPU-baseline timing budget")
    ap.add_argument("--no-rs", action="store_true", help="skip the rate-state plain-vs-learned leg")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32-metric leg")
    ap.add_argument("--no-side", action="store_true", help="skip the 400x300x200 slip-weakening leg")
    ap.add_argument("--no-ceilings", action="store_true", help="skip the copy and streaming-kernel ceilings")
    ap.add_argument("--rs-lead", type=int, default=200, help="steps before the rate-state timing window")
    ap.add_argument("--rs-steps", type=int, default=800, help="rate-state timing window (steps)")
    ap.add_argument("--calibrate", action="store_true",
                    help="only run a 1 GiB device-to-device copy (counter calibration for rocprofv3)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: the solver runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the whole domain lives in one MI355X's HBM; see the sharding design notes)", file=sys.stderr)
        sys.exit(2)

    if args.calibrate:
        import torch
        n = 256 * 1024 * 1024
        a = torch.ones(n, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "copy", "bytes_read": 4 * n, "bytes_written": 4 * n, "copies": 3}))
        return

    from dynfd3dcuda_amd import configs

    if args.workload not in configs.WORKLOADS:
        ap.error(f"unknown workload {args.workload}; choose from {', '.join(configs.WORKLOADS)}")
    gen = configs.WORKLOADS[args.workload]
    kw = {"nt": args.steps + args.warmup + PROFILE_STEPS + 2}
    if args.seed is not None:
        kw["seed"] = args.seed
    case = gen(**kw)
    p = case["params"]
    p.update(overrides(args))
    if p.get("NewtonMode") == 2 and not p.get("NewtonMLP"):
        if not os.path.exists(args.mlp):
            raise SystemExit(f"bench.py: NewtonMode 2 needs the learned-guess network; {args.mlp} is missing")
        p["NewtonMLP"] = args.mlp
    r = timed_run(case, args.steps, args.warmup, args.workload)
    out = {
        "metric": METRIC, "value": r["value"], "unit": "grid-pt updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": r["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic (seeded generators, dynfd3dcuda_amd.configs)",
        "config": {"workload": workload_name(args.workload, p), "grid_points": r["npts"],
                   "metric_storage": {0: "fp32", 1: "bf16", 2: "bf16_deviation_media_palette" if r["palette"]
                                      else "bf16_deviation_fp32_media"}[p.get("MetricBF16", 0)],
                   "newton": {0: "reference rule", 1: "plain guess, residual rule",
                              2: "learned guess, residual rule"}.get(p.get("NewtonMode", 0), "")
                   if p.get("Friction_type", 0) >= 1 else None,
                   "network": os.path.relpath(p["NewtonMLP"], ROOT) if p.get("NewtonMLP") else None,
                   "B_alg_bytes_per_pt_step": r["B_alg"], "step_hbm_roofline_frac": r["step_frac"],
                   "B_moved_bytes_per_pt_step": r["B_moved"],
                   "metric_words_skipped": r["metric_words_skipped"],
                   "step_hbm_frac_of_bytes_moved": r["value"] * r["B_moved"] / (PEAK_HBM_GBS * 1e9),
                   "setup_s": round(r["setup_s"], 1), "stage_points_per_launch": r["stage_points_per_launch"],
                   "aux_pairs_per_launch": r["aux_pairs_per_launch"],
                   "aux_bytes_per_pair": {"read": AUX_READ_BYTES, "update": AUX_UPDATE_BYTES},
                   "step_hbm_roofline_frac_with_aux": r["step_frac_with_aux"],
                   "newton_in_timed_steps": r["newton"], "fault_outputs_finite": r["fault_outputs_finite"]},
        "roofline": {"bound": "hbm", "kernel": r["dom"], "achieved": r["achieved"], "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": r["achieved"] / PEAK_HBM_GBS, "traffic": r["traffic"],
                     "frac_with_aux": r["achieved_with_aux"] / PEAK_HBM_GBS,
                     "points_per_launch": r["points_per_launch"], "avg_launch_ms": r["avg_launch_ms"]},
        "kernel_ms_per_step": r["kernel_ms_per_step"],
        "cpu_baseline": None,
        "fp32_metric": None,
        "side": None,
        "rate_state": None,
    }
    print(f"bench: main line {r['ms_per_step']:.2f} ms/step", file=sys.stderr, flush=True)
    if not args.no_ceilings:
        ceil = copy_ceiling()
        streams = stream_ceilings()
        out["roofline"]["copy_ceiling_GBps"] = ceil
        out["roofline"]["stream_ceilings_GBps"] = streams
        out["roofline"]["frac_of_read2_write1_ceiling"] = r["achieved"] / streams["2:1"]
    if not args.no_fp32 and p.get("MetricBF16", 0):
        c32 = gen(**kw)
        c32["params"].update(overrides(args))
        c32["params"].update({"MetricBF16": 0, "NewtonMLP": p.get("NewtonMLP", "")})
        r32 = timed_run(c32, args.steps, args.warmup, args.workload + "_fp32")
        out["fp32_metric"] = {"metric_storage": "fp32", "value": r32["value"], "ms_per_step": r32["ms_per_step"],
                              "B_alg_bytes_per_pt_step": r32["B_alg"], "step_hbm_roofline_frac": r32["step_frac"],
                              "dominant_kernel": r32["dom"], "dominant_frac": r32["achieved"] / PEAK_HBM_GBS,
                              "mixed_speedup": r32["ms_per_step"] / r["ms_per_step"]}
        print(f"bench: fp32-metric leg {r32['ms_per_step']:.2f} ms/step", file=sys.stderr, flush=True)
    if not args.no_side and args.workload != SIDE_WORKLOAD:
        cs = configs.WORKLOADS[SIDE_WORKLOAD](nt=50 + PROFILE_STEPS + 10)
        rs_ = timed_run(cs, 50, 5, SIDE_WORKLOAD)
        out["side"] = {"workload": workload_name(SIDE_WORKLOAD, cs["params"]), "value": rs_["value"],
                       "ms_per_step": rs_["ms_per_step"], "steps": 50, "warmup": 5,
                       "step_hbm_roofline_frac": rs_["step_frac"],
                       "step_hbm_roofline_frac_with_aux": rs_["step_frac_with_aux"],
                       "roofline": {"kernel": rs_["dom"], "achieved": rs_["achieved"],
                                    "frac": rs_["achieved"] / PEAK_HBM_GBS, "traffic": rs_["traffic"],
                                    "frac_with_aux": rs_["achieved_with_aux"] / PEAK_HBM_GBS,
                                    "points_per_launch": rs_["points_per_launch"],
                                    "avg_launch_ms": rs_["avg_launch_ms"]},
                       "kernel_ms_per_step": rs_["kernel_ms_per_step"]}
        print(f"bench: side leg {rs_['ms_per_step']:.2f} ms/step", file=sys.stderr, flush=True)
    if not args.no_rs:
        if not os.path.exists(args.mlp):
            raise SystemExit(f"bench.py: the rate-state leg needs the learned-guess network; {args.mlp} is missing")
        out["rate_state"] = rate_state_leg(args, args.mlp)
        print("bench: rate-state leg done", file=sys.stderr, flush=True)
    if not args.no_cpu:
        if args.workload == DEFAULT_WORKLOAD:
            sub = configs.large_rs_mlp_bf16_1600x1000x500(sub_nx=CPU_SUB_NX, nt=60)
            what = (f"the {CPU_SUB_NX}x1000x500 sub-slab around the fault of the same workload (same fault, "
                    f"stress, media and topography rows; 50 M of its 800 M points)")
        else:
            sub = gen(**kw)
            sub["params"].update(overrides(args))
            what = "the full grid of the same workload"
        out["cpu_baseline"] = cpu_baseline(sub, args.cpu_seconds, what)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
