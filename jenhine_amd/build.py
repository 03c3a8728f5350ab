"""Builds the native libraries of jenhine_amd in-tree.

* ``libjhip.so``   -- the gfx950 HIP kernels behind the C-ABI of include/jhip.h
* ``libjhsynth.so`` -- the synthetic scan generator (host C, OpenMP)
* ``libjhhost.so`` / test binaries -- the C++ agent/pipe host (built by its own
  Makefile under host/)

Every HIP translation unit is compiled for gfx950 only, with packed-FP32
instruction selection disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)

HIP_SOURCES = ["jhip_core.hip", "jhip_bp.hip", "jhip_filter.hip", "jhip_composite.hip", "jhip_fp.hip",
               "jhip_graph.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-Wno-unused-command-line-argument",
]


def _run(cmd, cwd=None):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd, cwd=cwd)


def build_hip(verbose=True):
    out = os.path.join(HERE, "libjhip.so")
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = out + ".tmp"
    _run(["/opt/rocm/bin/hipcc", *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"), *srcs, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_probe():
    """gather_rate and valu_rate: the vector-memory and VALU issue-rate probes
    the roofline cites."""
    outdir = os.path.join(HERE, "build")
    os.makedirs(outdir, exist_ok=True)
    outs = []
    for name in ("gather_rate", "valu_rate"):
        out = os.path.join(outdir, name)
        _run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O3", "-std=c++17",
              "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument",
              os.path.join(CSRC, name + ".hip"), "-o", out + ".tmp"])
        os.replace(out + ".tmp", out)
        outs.append(out)
    return outs


def build_synth():
    out = os.path.join(HERE, "libjhsynth.so")
    tmp = out + ".tmp"
    _run(["gcc", "-O3", "-fPIC", "-shared", "-fopenmp", "-std=gnu11", os.path.join(CSRC, "synth.c"),
          "-o", tmp, "-lm"])
    os.replace(tmp, out)
    return out


def build_host():
    mk = os.path.join(HERE, "host", "Makefile")
    if os.path.exists(mk):
        _run(["make", "-s", "-j8", "-C", os.path.join(HERE, "host")])


def build_all():
    build_synth()
    build_hip()
    build_host()
    build_probe()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "hip": build_hip, "synth": build_synth, "host": build_host, "probe": build_probe}[w]()
