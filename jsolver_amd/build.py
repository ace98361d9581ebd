"""Builds the HIP library of jsolver_amd in-tree for gfx950.

    python -m jsolver_amd.build          (or jsolver_amd.build.build())

Output: jsolver_amd/_lib/libjsolver_amd.so.  hipcc cross-compiles without a
GPU.  FMA contraction is disabled in the stencil code so the FP64 kernels
evaluate the reference's formulas term by term (IEEE), and packed FP32 ops are
disabled for all device code.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
ROOT = os.path.dirname(HERE)
ARCH = os.environ.get("JSV_OFFLOAD_ARCH", "gfx950")

LIBS = {
    "libjsolver_amd.so": ["jsv_api.hip", "jsv_unet.hip"],
}


def hipcc():
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    outs = []
    for lib, srcs in LIBS.items():
        out = os.path.join(OUT_DIR, lib)
        tmp = out + ".tmp"
        cmd = [hipcc(), "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-shared",
               "-ffp-contract=off",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-Wno-unused-command-line-argument",
               "-I", os.path.join(ROOT, "include"), "-I", CSRC,
               "-o", tmp] + [os.path.join(CSRC, s) for s in srcs]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"hipcc failed building {lib}")
        if verbose and r.stderr.strip():
            # hipcc's host pass does not know the device-only target feature
            lines = [l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l]
            if lines:
                sys.stderr.write("\n".join(lines) + "\n")
        os.replace(tmp, out)
        outs.append(out)
    return outs


if __name__ == "__main__":
    for o in build(verbose=True):
        print("built", o)
