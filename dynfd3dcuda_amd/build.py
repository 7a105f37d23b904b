"""Build the HIP library (libcgfdm.so) and the cgfdm CLI for gfx950, in-tree.

Run as ``python -m dynfd3dcuda_amd.build`` or through __graft_entry__.build().
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)

SOURCES = ["params.cpp", "setup.cpp", "netcdf_classic.cpp", "kernels.hip", "wave_stage.hip", "wave_stage_f32.hip", "wave_stage_bf16.hip",
           "wave_stage_bf16dev.hip", "mlp_train.hip", "stream_probe.hip", "field_check.hip", "capi.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# packed FP32 instructions are disabled for all device code; multiply-adds
# are fused only within one source expression (-ffp-contract=on), so kernels
# that share an operator's code round it the same way whatever the register
# allocation and block structure around it
DEVFLAGS = ["--offload-arch=gfx950", "-ffp-contract=on", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-Wno-unused-result", "-I", os.path.join(ROOT, "include")]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(jobs=8):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    procs = []
    for s in SOURCES:
        src = os.path.join(CSRC, s)
        obj = os.path.join(objdir, s + ".o")
        lang = ["-x", "hip"] if s.endswith(".hip") or s == "capi.cpp" or s == "setup.cpp" else []
        cmd = [HIPCC] + COMMON + DEVFLAGS + lang + ["-c", src, "-o", obj]
        print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        objs.append(obj)
        if len(procs) >= jobs:
            for p in procs:
                if p.wait() != 0:
                    raise RuntimeError("compile failed")
            procs = []
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError("compile failed")
    lib = os.path.join(HERE, "libcgfdm.so")
    _run([HIPCC] + DEVFLAGS + ["-shared", "-fopenmp", "-o", lib] + objs)
    exe = os.path.join(HERE, "cgfdm")
    _run([HIPCC] + COMMON + DEVFLAGS + [os.path.join(CSRC, "cgfdm_main.cpp"), "-o", exe,
                                        "-L", HERE, "-lcgfdm", "-Wl,-rpath,$ORIGIN"])
    return lib


def build_oracle():
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-C", odir])
    return os.path.join(odir, "_build", "liboracle.so")


if __name__ == "__main__":
    build()
    if "--no-oracle" not in sys.argv:
        build_oracle()
