#!/usr/bin/env python3
"""Benchmark of the CGFDM dynamic-rupture time step on one MI355X.

One step = one full RK4 time step of the whole job on the GPU: rupture
sensor, and per stage the fault trial (fault_pre), the fused wave stage
(interior / absorbing-band / free-surface instantiations + PML auxiliary
update), the split-node derivatives (fault_deriv) and the fault RK
(fault_post).  Inputs are generated with the seeded synthetic generators of
dynfd3dcuda_amd.configs and are resident in HBM before timing starts.

Default workload: the curved-grid slip-weakening configuration, 400 x 300 x
200 points at 50 m (topography, bent fault, 3-layer medium, 20-layer ADE-PML,
FP32).  Prints one JSON line:
  value          grid-point updates per second (pt-steps/s) of the whole step
  roofline       the dominant stage kernel (the stage-kernel class with the
                 largest GPU time in an untimed window of PROFILE_STEPS steps
                 with every class timed), timed with HIP events on the
                 solver's stream around its own launches over the timed
                 region; achieved = algorithmic bytes per launch
                 (points x B_alg/4, B_alg = 768 B/pt-step FP32, 672 B with
                 BF16 metric) / average launch time; traffic = HBM bytes per
                 launch from the committed rocprofv3 FETCH/WRITE summary when
                 present
  cpu_baseline   the FP64 oracle (C++/OpenMP) on the host cores, same
                 workload, a few steps
  mixed          the same workload with the metric and media stored in BF16
  mixed_deviation  the same with the metric as BF16 deviations, media exact
                 (FP32 arithmetic): throughput and the step's fraction of the
                 672 B/pt-step roofline
  kernel_ms_per_step  every kernel class's GPU time per step, from the
                 untimed window
  rate_state     rate-and-state fault step: the reference's Newton rule,
                 plain Newton and the learned initial guess, same grid with
                 the aging law: the
                 plain run's Newton samples train the network on the GPU,
                 then both runs are timed over the same window of steps
                 inside the rupture (GPU time per RK stage of the Newton
                 kernel and of the whole fault part, iterations per solve)
Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload curved_sw_400x300x200] [--no-rs]
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


def b_alg(params):
    """Algorithmic bytes per grid-point update (one full RK4 step): 144
    wavefield words + 48 coefficient words (12 per stage: FP32; bfloat16;
    or 9 bfloat16 metric deviations and 3 exact FP32 media words)."""
    m = params.get("MetricBF16", 0)
    return 144 * 4 + (48 * 4 if not m else 48 * 2 if m == 1 else 4 * (9 * 2 + 3 * 4))


def workload_name(name, p):
    what = {"flat_sw_100x200x100": "plain flat grid, slip weakening, Cerjan sponge",
            "curved_sw_400x300x200": "curved grid + topography, bent slip-weakening fault, 3-layer medium, "
                                     "20-layer ADE-PML",
            "curved_rs_400x300x200": "curved grid, rate-and-state aging law, plain Newton",
            "curved_rs_mlp_400x300x200": "curved grid, rate-and-state aging law, learned Newton guess",
            "large_rs_mlp_bf16_1600x1000x500": "large grid, rate-and-state, BF16 metric/media"}
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
    buffer of `nbytes`, timed with HIP events: the practical ceiling that a
    streaming kernel on this GPU can reach."""
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


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cpu_baseline(case, budget_s):
    """Time the FP64 oracle on the same workload (same grid and inputs) for a
    few steps on the host cores.  The oracle is the checker of the test suite
    and is used here only as the reported CPU baseline."""
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
    n = int(max(1, min(20, budget_s // max(t1, 1e-3))))
    t0 = time.perf_counter()
    for it in range(1, 1 + n):
        orc.step(it)
    dt = (time.perf_counter() - t0) / n
    npts = p["NX"] * p["NY"] * p["NZ"]
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": npts / dt, "unit": "grid-pt updates/s", "cores": cores, "kind": "port",
            "effective_GBps": npts / dt * b_alg(p) / 1e9,
            "sample": f"FP64 oracle, full {p['NX']}x{p['NY']}x{p['NZ']} grid of the same workload, {n} steps after 1 "
                      f"warm-up ({dt:.2f} s/step; setup {t_setup:.1f} s not timed), {cpu_model()}"}


def mixed_leg(case, steps, warmup, mode=1):
    """Same workload with MetricBF16=mode (1: metric and media words in
    bfloat16; 2: metric words as bfloat16 deviations, media words exact):
    throughput of the whole step."""
    import dynfd3dcuda_amd as dyn
    from dynfd3dcuda_amd import configs
    p = dict(case["params"], MetricBF16=mode)
    s = dyn.Solver(p)
    configs.apply(case, s)
    s.setup()
    d = s.dims()
    for it in range(warmup):
        s.step(it)
    s.synchronize()
    t0 = time.perf_counter()
    for it in range(warmup, warmup + steps):
        s.step(it)
    s.synchronize()
    t = time.perf_counter() - t0
    s.close()
    v = d["ni"] * d["nj"] * d["nk"] * steps / t
    ba = b_alg(p)
    return {"metric_storage": "bf16" if mode == 1 else "bf16_deviation_exact_media", "value": v, "unit": "grid-pt updates/s", "ms_per_step": 1e3 * t / steps,
            "B_alg_bytes_per_pt_step": ba, "step_hbm_roofline_frac": v * ba / (PEAK_HBM_GBS * 1e9)}


def overrides(args):
    """params.json keys given with --set KEY=VALUE."""
    out = {}
    for kv in args.set:
        k, v = kv.split("=", 1)
        out[k] = float(v) if any(ch in v for ch in ".eE") else int(v)
    return out


def rate_state_leg(args):
    """The rate-state solve three ways on the rate-state configuration: the
    reference's stopping rule (NewtonMode 0: |dx|/|x| < 1e-12, at most 100
    iterations, trial_ratestate.cu:22-57), plain Newton on the residual rule
    (NewtonMode 1) and the learned guess (NewtonMode 2, same residual rule
    and fallback)."""
    import numpy as np
    import dynfd3dcuda_amd as dyn
    from dynfd3dcuda_amd import configs, mlp
    lead, window = args.rs_lead, args.rs_steps
    fault = ("fault_pre", "fault_rs", "fault_deriv", "fault_post")
    res = {}
    params = None
    for mode in (0, 1, 2):
        case = configs.curved_rs_400x300x200(nt=lead + window + 2)
        case["params"].update({"NewtonMode": mode, "NewtonTol": 1e-10, "NewtonFallbackIter": 20})
        case["params"].update(overrides(args))
        s = dyn.Solver(case["params"])
        configs.apply(case, s)
        if mode == 2:
            s.set_newton_mlp(params)
        s.setup()
        d = s.dims()
        t0 = time.perf_counter()
        if mode == 1:
            F, T = mlp.collect(s, 0, lead, every=args.rs_sample_every, keep=args.rs_samples)
        else:
            for it in range(lead):
                s.step(it)
        s.synchronize()
        t_lead = time.perf_counter() - t0
        s.put("newton_counts", np.zeros(3 * d["nj"] * d["nk"], np.int32))
        s.profile(True)
        for it in range(lead, lead + window):
            s.step(it)
        s.synchronize()
        prof = s.profile_report()
        s.profile(False)
        cnt = s.get("newton_counts").reshape(3, -1).astype(np.float64)
        nst = 4 * window
        per_node = cnt[1] / np.maximum(cnt[0], 1)
        res[mode] = {"newton_kernel_ms_per_stage": prof["fault_rs"][0] / nst,
                     "fault_ms_per_stage": sum(prof[k][0] for k in fault) / nst,
                     "iters_per_solve": cnt[1].sum() / max(cnt[0].sum(), 1),
                     "max_node_mean_iters": float(per_node.max()),
                     "fallbacks": int(cnt[2].sum()),
                     "sliding_nodes": int(np.sum(np.hypot(s.get("Vs1"), s.get("Vs2")) > 1e-3)),
                     "lead_seconds": round(t_lead, 1)}
        if mode == 1:
            t0 = time.perf_counter()
            params, hist = mlp.train(F, T, epochs=args.rs_epochs, batch=4096, lr=3e-3, seed=0)
            res["train"] = {"samples": int(len(T)), "epochs": args.rs_epochs, "seconds": round(time.perf_counter() - t0, 2),
                            "loss_first": float(hist[0]), "loss_last": float(hist[-1])}
            fault_nodes = d["nj"] * d["nk"]
        s.close()
    pl, ml = res[1], res[2]
    return {"workload": f"curved_rs_400x300x200: the curved 400x300x200 grid with the aging law, {fault_nodes} fault nodes; steps {lead}..{lead + window - 1}",
            "reference_rule": res[0], "plain": pl, "learned": ml, "train": res["train"],
            "fault_step_speedup_vs_reference_rule": res[0]["fault_ms_per_stage"] / ml["fault_ms_per_stage"],
            "newton_kernel_speedup": pl["newton_kernel_ms_per_stage"] / ml["newton_kernel_ms_per_stage"],
            "fault_step_speedup": pl["fault_ms_per_stage"] / ml["fault_ms_per_stage"]}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="curved_sw_400x300x200",
                    help="synthetic workload, one of dynfd3dcuda_amd.configs.WORKLOADS")
    ap.add_argument("--seed", type=int, default=None)
    ap.add_argument("--set", action="append", default=[], metavar="KEY=VALUE",
                    help="override a params.json key of the workload (repeatable)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="CPU-baseline timing budget")
    ap.add_argument("--no-rs", action="store_true", help="skip the rate-state plain-vs-learned leg")
    ap.add_argument("--no-mixed", action="store_true", help="skip the BF16-metric leg")
    ap.add_argument("--rs-lead", type=int, default=600, help="steps before the rate-state timing window")
    ap.add_argument("--rs-steps", type=int, default=100, help="rate-state timing window (steps)")
    ap.add_argument("--rs-sample-every", type=int, default=4, help="training-sample stride (steps)")
    ap.add_argument("--rs-samples", type=int, default=2_000_000, help="training samples kept")
    ap.add_argument("--rs-epochs", type=int, default=8, help="training epochs")
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

    import dynfd3dcuda_amd as dyn
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
    note = None
    if p.get("NewtonMode") == 2 and not p.get("NewtonMLP"):
        # no network given: the step runs with plain Newton (same residual rule)
        p["NewtonMode"] = 1
        note = "NewtonMode 1 (no learned-guess network given)"
    solver = dyn.Solver(p)
    configs.apply(case, solver)
    t_setup = time.perf_counter()
    solver.setup()
    d = solver.dims()
    npts = d["ni"] * d["nj"] * d["nk"]
    print(f"bench: setup of {npts} points in {time.perf_counter() - t_setup:.1f} s", file=sys.stderr, flush=True)

    for it in range(args.warmup):
        solver.step(it)
    solver.synchronize()
    # an untimed window with every kernel class timed: the per-class
    # breakdown, and the stage-kernel class with the largest GPU time
    nprof = PROFILE_STEPS
    it0 = args.warmup
    solver.profile(True)
    for it in range(it0, it0 + nprof):
        solver.step(it)
    solver.synchronize()
    prof_all = solver.profile_report()
    solver.profile(False)
    stage = {k: v for k, v in prof_all.items() if k.startswith("stage_") and v[1] > 0}
    dom = max(stage, key=lambda k: stage[k][0])
    # the timed region: HIP events around that class's launches only (events
    # around every launch would add their own packets to each step)
    it0 += nprof
    solver.profile_classes([dom])
    solver.profile(True)
    t0 = time.perf_counter()
    for it in range(it0, it0 + args.steps):
        solver.step(it)
    solver.synchronize()
    t = time.perf_counter() - t0
    prof = solver.profile_report()
    solver.profile(False)
    solver.profile_classes(None)

    value = npts * args.steps / t
    ba = b_alg(p)
    ms, n, pts = prof[dom]
    avg_s = ms / n / 1e3
    bytes_per_launch = pts * ba / 4.0  # one launch = one RK stage of its points
    achieved = bytes_per_launch / avg_s / 1e9
    traffic = read_traffic(args.workload, dom)
    out = {
        "metric": METRIC, "value": value, "unit": "grid-pt updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * t / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic (seeded generators, dynfd3dcuda_amd.configs)",
        "config": {"workload": workload_name(args.workload, p) + (f"; {note}" if note else ""),
                   "grid_points": npts, "B_alg_bytes_per_pt_step": ba,
                   "step_hbm_roofline_frac": value * ba / (PEAK_HBM_GBS * 1e9),
                   "stage_points_per_launch": {k: v[2] for k, v in stage.items()}},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": traffic,
                     "points_per_launch": pts, "avg_launch_ms": avg_s * 1e3},
        "kernel_ms_per_step": {k: v[0] / nprof for k, v in prof_all.items() if v[1] > 0},
        "cpu_baseline": None,
        "mixed": None,
        "mixed_deviation": None,
        "rate_state": None,
    }
    solver.close()
    ceil = copy_ceiling()
    out["roofline"]["copy_ceiling_GBps"] = ceil
    out["roofline"]["frac_of_copy_ceiling"] = achieved / ceil if ceil else None
    if not args.no_mixed:
        out["mixed"] = mixed_leg(case, args.steps, args.warmup)
        out["mixed_deviation"] = mixed_leg(case, args.steps, args.warmup, mode=2)
    if not args.no_rs:
        out["rate_state"] = rate_state_leg(args)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(case, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
