"""Compile libdsx.so (the HIP kernels + C ABI) for gfx950, in-tree.

``python -m datashader_amd.build`` or ``__graft_entry__.build()``.  hipcc cross-compiles
for gfx950 without a GPU present.  Every source is compiled to its own object in
parallel (each object carries its own gfx950 code object), then the objects are linked
into one shared library.
"""
from __future__ import annotations

import glob
import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB = os.path.join(PKG, "libdsx.so")

# -ffp-contract=off: the reference's projection rounds x*sx and +tx separately; keep
#   every float expression uncontracted (the kernels also use explicit __dmul_rn etc.).
# -munsafe-fp-atomics: float atomics compile to the hardware instructions (device
#   memory from hipMalloc is coarse grained), not CAS loops.
# -packed-fp32-ops off: no packed FP32 instructions are emitted.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build libdsx.so)")


def sources():
    return sorted(glob.glob(os.path.join(PKG, "csrc", "*.hip")))


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    return r.returncode, r.stdout + r.stderr


def build(verbose: bool = False, jobs: int | None = None) -> str:
    """Compile every HIP source into ``datashader_amd/libdsx.so``; returns its path."""
    cc = hipcc()
    inc = ["-I", os.path.join(ROOT, "include")]
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, min(os.cpu_count() or 1, 8)))
    with tempfile.TemporaryDirectory(prefix="dsx_build_") as tmpdir:
        objs = [os.path.join(tmpdir, os.path.basename(s) + ".o") for s in srcs]
        cmds = [[cc] + HIPCC_FLAGS + inc + ["-c", "-o", o, s] for s, o in zip(srcs, objs)]
        with ThreadPoolExecutor(max_workers=jobs) as pool:
            results = list(pool.map(lambda c: _run(c, verbose), cmds))
        bad = [(s, out) for s, (rc, out) in zip(srcs, results) if rc != 0]
        if bad:
            for s, out in bad:
                sys.stderr.write(f"--- {s}\n{out}")
            raise RuntimeError("hipcc failed building libdsx.so")
        tmp = LIB + ".tmp"
        rc, out = _run([cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs,
                       verbose)
        if rc != 0:
            sys.stderr.write(out)
            raise RuntimeError("hipcc failed linking libdsx.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
