"""Build libfsx.so (the gfx950 HIP kernels + C ABI) in-tree with hipcc.

Each source compiles to its own object in parallel (every translation unit carries its own
gfx950 code object; no relocatable device code is needed), then one hipcc line links them.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfsx.so")
OBJ_DIR = os.path.join(HERE, "_build")
SOURCES = ["api.hip", "sift.hip", "match.hip", "geom.hip", "fmatch.hip", "mlp.hip", "verify.hip", "tracks.hip",
           "clahe.hip"]

# Strict IEEE arithmetic (the kernels reproduce numpy's float32/float64 rounding step by
# step): no contraction into FMA, correctly rounded float32 divide/sqrt, denormals kept.
# Packed-FP32 instruction selection is disabled for every GPU code object.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]

# Per-source additions.  fmatch.hip: MFMA accumulators in ordinary VGPRs rather than AGPRs, so
# the screen's per-tile gate reads them without v_accvgpr_read copies.
EXTRA_FLAGS = {"fmatch.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"]}


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    return "hipcc"


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
    return r.returncode


def build(verbose: bool = False, jobs: int | None = None) -> str:
    os.makedirs(OBJ_DIR, exist_ok=True)
    srcs = list(SOURCES)
    objs = [os.path.join(OBJ_DIR, os.path.splitext(s)[0] + ".o") for s in srcs]
    cmds = [[hipcc()] + HIPCC_FLAGS + EXTRA_FLAGS.get(s, []) + ["-I", CSRC, "-c", "-o", o, os.path.join(CSRC, s)]
            for s, o in zip(srcs, objs)]
    jobs = jobs or max(1, min(len(cmds), 16, os.cpu_count() or 1))
    with ThreadPoolExecutor(jobs) as ex:
        rcs = list(ex.map(lambda c: _run(c, verbose), cmds))
    if any(rcs):
        raise RuntimeError("hipcc failed compiling %s" % [s for s, rc in zip(srcs, rcs) if rc])
    tmp = LIB + ".tmp"
    if _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs, verbose):
        raise RuntimeError("hipcc failed linking libfsx.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
