"""Summaries of rocprofv3 output for the solver's kernels.

  python -m dynfd3dcuda_amd.profiling stats  <trace-dir>
      per kernel class: launches, average / total GPU time (from a
      `rocprofv3 --kernel-trace --output-format csv` run)
  python -m dynfd3dcuda_amd.profiling traffic --fetch D1 --write D2 --cal-fetch D3 --cal-write D4 --bench-json F
      HBM bytes per RK stage of each kernel class (all its dispatches in one
      stage: the band class has two) from `--pmc FETCH_SIZE` and
      `--pmc WRITE_SIZE` runs, each scaled by what the same counter reads for
      a device-to-device copy of known size (`bench.py --calibrate`), next to
      the algorithmic bytes of the stage kernels (points per stage from the
      JSON line bench.py printed in the same configuration), and those with
      the ADE-PML auxiliary words each class reads and advances added

Kernel classes: stage_interior / stage_band / stage_surface (the
instantiations of the wave stage kernel and of the pipelined stage kernel,
whose BAND instantiations are band tiles; stage_surface includes the two
free-surface precompute passes), pml_aux, fault_pre, fault_rs, fault_deriv,
fault_post, rup_sensor, thermpress, setup_metric (the setup's metric
pass), setup_other (its other passes); everything else is listed under its
own name.
"""
import argparse
import collections
import csv
import glob
import json
import os
import re
import sys


STAGE_CLASSES = ["stage_interior", "stage_band", "stage_surface"]


def kernel_class(name):
    m = re.search(r"wave_stage_kernel<([^>]*)>", name)
    if m:  # template arguments FX, FY, FZ, metric type, KIND[, tile width]
        return STAGE_CLASSES[int(m.group(1).split(",")[4])]
    m = re.search(r"wave_stage_kernelI(?:Li-?\d+E|Lin1E){3}[ft]Li([012])E", name)  # mangled form
    if m:
        return STAGE_CLASSES[int(m.group(1))]
    m = re.search(r"wave_interior_kernel<([^>]*)>", name)
    if m:  # the pipelined kernel: FX, FY, FZ, metric type, tile width, threads, waves per SIMD[, BAND, RING, TOP]
        args = [x.strip() for x in m.group(1).split(",")]
        if len(args) > 9 and args[9] == "true":
            return "stage_surface"
        return "stage_band" if len(args) > 7 and args[7] == "true" else "stage_interior"
    if "wave_interior_kernel" in name:
        return "stage_interior"
    if "traction_top_kernel" in name or "image_top_kernel" in name:
        return "stage_surface"  # the free-surface precompute passes
    for k in ("pml_aux_kernel", "fault_pre_kernel", "fault_rs_kernel", "fault_deriv_kernel", "fault_post_kernel"):
        if k in name:
            return k[: -len("_kernel")]
    if "rup_sensor" in name:
        return "rup_sensor"
    if "thermpress_kernel" in name:
        return "thermpress"
    if "metric_media_kernel" in name:
        return "setup_metric"
    for k in ("coord_kernel", "coord_ghost_kernel", "hmin_kernel", "gather_kernel", "metric_range_kernel",
              "fault_coef_kernel", "wave_free_kernel", "media_palette_kernel", "metric_pair_or_kernel"):
        if k in name:
            return "setup_other"  # the other setup passes (coordinates, CFL range, coefficients)
    return name.split("(")[0].strip()


def _csv(dirpath, suffix):
    files = glob.glob(os.path.join(dirpath, "**", f"*{suffix}"), recursive=True)
    if not files:
        raise SystemExit(f"no *{suffix} under {dirpath}")
    rows = []
    for f in files:
        rows += list(csv.DictReader(open(f)))
    return rows


def stats(dirpath):
    agg = collections.defaultdict(list)
    for r in _csv(dirpath, "kernel_trace.csv"):
        dur = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
        agg[kernel_class(r["Kernel_Name"])].append(dur)
    tot = sum(sum(v) for v in agg.values())
    lines = [f"{'kernel class':40s} {'launches':>9s} {'avg_us':>10s} {'total_ms':>10s} {'share':>7s}"]
    for k, v in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
        lines.append(f"{k[:40]:40s} {len(v):9d} {sum(v) / len(v):10.1f} {sum(v) / 1e3:10.3f} {100 * sum(v) / tot:6.1f}%")
    return "\n".join(lines)


def _counter(dirpath, counter):
    agg = collections.defaultdict(list)
    for r in _csv(dirpath, "counter_collection.csv"):
        if r["Counter_Name"] == counter:
            agg[kernel_class(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    return agg


def traffic(a):
    # calibration: a 1 GiB device-to-device copy reads and writes 2^30 bytes
    cal_bytes = float(2 ** 30)
    cf = _counter(a.cal_fetch, "FETCH_SIZE")
    cw = _counter(a.cal_write, "WRITE_SIZE")
    copy_k = max(cf, key=lambda k: sum(cf[k]))
    f_fetch = cal_bytes / (max(cf[copy_k]) * 1024.0)
    f_write = cal_bytes / (max(cw[copy_k]) * 1024.0)
    fetch = _counter(a.fetch, "FETCH_SIZE")
    write = _counter(a.write, "WRITE_SIZE")
    lines = [ln for ln in open(a.bench_json).read().splitlines() if ln.lstrip().startswith("{")]
    bj = json.loads(lines[-1])
    pts = bj["config"].get("stage_points_per_launch", {})
    ba = bj["config"]["B_alg_bytes_per_pt_step"]
    aux = bj["config"].get("aux_pairs_per_launch", {})
    abyte = bj["config"].get("aux_bytes_per_pair", {})
    out = [f"# calibration kernel: {copy_k}; FETCH_SIZE x {f_fetch:.3f}, WRITE_SIZE x {f_write:.3f} (KiB -> bytes x factor)",
           f"fetch_scale: {f_fetch:.4f}", f"write_scale: {f_write:.4f}", f"B_alg_bytes_per_pt_step: {ba}"]
    # one fault_pre dispatch per RK stage (rupture runs); else the interior stage kernel
    nst = len(fetch.get("fault_pre", fetch.get("stage_interior", [1])))
    out.append(f"rk_stages_profiled: {nst}")
    for k in sorted(fetch):
        if k not in write or k.startswith("__") or k.startswith("setup_"):
            continue  # copies, fills and setup passes are not per-stage work
        fb = sum(fetch[k]) / nst * 1024.0 * f_fetch
        wb = sum(write[k]) / nst * 1024.0 * f_write
        out.append(f"{k}_fetch_bytes_per_stage: {fb:.6g}")
        out.append(f"{k}_write_bytes_per_stage: {wb:.6g}")
        out.append(f"{k}_hbm_bytes_per_stage: {fb + wb:.6g}")
        alg = pts.get(k, 0) * ba / 4.0
        if alg:
            out.append(f"{k}_alg_bytes_per_stage: {alg:.6g}")
            out.append(f"{k}_hbm_over_alg: {(fb + wb) / alg:.3f}")
        if k in aux and abyte:  # the ADE-PML auxiliary words the class reads and advances
            rd, up = aux[k]
            alg_aux = alg + rd * abyte["read"] + up * abyte["update"]
            out.append(f"{k}_alg_bytes_with_aux_per_stage: {alg_aux:.6g}")
            out.append(f"{k}_hbm_over_alg_with_aux: {(fb + wb) / alg_aux:.3f}")
    return "\n".join(out)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    t = sub.add_parser("traffic")
    for k in ("fetch", "write", "cal_fetch", "cal_write", "bench_json"):
        t.add_argument("--" + k.replace("_", "-"), dest=k, required=True)
    a = ap.parse_args(argv)
    print(stats(a.dir) if a.cmd == "stats" else traffic(a))


if __name__ == "__main__":
    main(sys.argv[1:])
