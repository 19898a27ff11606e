"""Build libernn_hip.so in-tree with hipcc for gfx950.

    python -m ernn.build        (from elastodynamics-encoded-rnn_amd/)

The library is a plain C-ABI shared object (include/ernn.h); it links the HIP
runtime by soname, so inside a process that has imported torch it binds to the
same libamdhip64 that PyTorch-ROCm uses.
"""
from __future__ import annotations

import os
import re
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG_DIR)                      # elastodynamics-encoded-rnn_amd/
REPO = os.path.dirname(ROOT)
CSRC = os.path.join(ROOT, "csrc")
INCLUDE = os.path.join(REPO, "include")
LIB_NAME = "libernn_hip.so"
LIB_PATH = os.path.join(PKG_DIR, LIB_NAME)
SOURCES = ["ernn_kernels.hip", "ernn_plan.cpp", "ernn_fem.hip", "ernn_fem.cpp", "ernn_tv_adam.hip"]

# -ffp-contract=off: every fused multiply-add is an explicit fmaf in the
# source, so each cell's arithmetic is the same in every kernel shape and a
# recomputed state equals the forward's bit for bit; packed FP32 instructions
# are disabled for all device code.
HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (set HIPCC or install ROCm)")


def _spilling_kernels(remarks: str):
    """Kernels whose resource-usage remarks report scratch memory or AGPRs
    (register spills): [(kernel, what, amount)]."""
    bad, name = [], None
    for line in remarks.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
        for what, pat in (("scratch B/lane", r"ScratchSize \[bytes/lane\]: (\d+)"), ("AGPRs", r"AGPRs: (\d+)")):
            m = re.search(pat, line)
            if m and int(m.group(1)) > 0 and name:
                bad.append((name, what, int(m.group(1))))
    return bad


def build(verbose: bool = False) -> str:
    """Compile every source for gfx950 and link LIB_PATH; returns its path.

    Fails if any kernel spills registers (to AGPRs or scratch): every tile
    shape the library launches must keep its working set in VGPRs."""
    cc = hipcc()
    objs = []
    build_dir = os.path.join(ROOT, "_build")
    os.makedirs(build_dir, exist_ok=True)
    for src in SOURCES:
        obj = os.path.join(build_dir, src.replace(".", "_") + ".o")   # ernn_fem.hip and ernn_fem.cpp differ
        cmd = [cc, *HIPCC_FLAGS, "-I", INCLUDE, "-I", CSRC, "-x", "hip",
               "-Rpass-analysis=kernel-resource-usage",
               "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stderr)
            raise subprocess.CalledProcessError(r.returncode, cmd)
        bad = _spilling_kernels(r.stderr)
        if bad:
            raise RuntimeError("kernels spill registers: " + ", ".join(f"{n} ({a} {w})" for n, w, a in bad))
        objs.append(obj)
    tmp = LIB_PATH + ".tmp"
    cmd = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
