"""Build the engine's shared library (``libskx.so``) for gfx950 with hipcc.

The library is built in-tree next to this file so that it travels with the
source tree; nothing is installed.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libskx.so"
SOURCES = ["skx_kernels.hip", "skx_runtime.hip", "skx_diag.hip", "skx_init.hip", "skx_observables.hip", "skx_io.cpp"] + [
    f"skx_stage_{th}_{p}.hip" for th in ("pion", "omega", "coulomb", "rho", "berger", "lightly") for p in ("f32", "f64")]
ARCH = os.environ.get("SKX_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed FP32 instructions stay off: all FP32 math is scalar per lane
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # a*b+c is fused exactly where one source expression says so, never by
    # the backend: every instantiation of a stage body (solo, batched) then
    # rounds identically, so a batched lattice is bit-identical to its solo run
    "-ffp-contract=on",
]

# Per-source additions.  The FP32 pion, Coulomb, Berger and lightly bound
# stage kernels take the ILP-first instruction scheduler: it spends their
# register headroom on more independent work with no spill, and no occupancy
# loss in the kernels their default launch shapes use (pion: stage 3 -5%,
# stage 4 -2%, whole step +0.9% at 128^3; Coulomb: step -1.4% at 128^3, -2.0%
# at 160^3; Berger: -1.4 to -2.0%; lightly bound: -0.1 to -1.5%; its energy
# pass on a uniform 32x8 tile drops from 3 to 2 waves per SIMD).  For omega
# (stage 3) and the rho curl pass it costs a wave, and FP64 pion measured 1.7%
# slower, so it is not applied there.  Scheduling only orders instructions:
# results are bit-identical either way.
_MAX_ILP = ["-mllvm", "-amdgpu-sched-strategy=max-ilp"]
SOURCE_FLAGS = {
    "skx_stage_pion_f32.hip": _MAX_ILP,
    "skx_stage_coulomb_f32.hip": _MAX_ILP,
    "skx_stage_berger_f32.hip": _MAX_ILP,
    "skx_stage_lightly_f32.hip": _MAX_ILP,
}


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the engine must be built with ROCm's hipcc")


def build(verbose: bool = True, jobs: int | None = None) -> Path:
    """Compile every HIP source and link ``libskx.so``; returns its path."""
    cc = hipcc()
    objdir = PKG / "_obj"
    objdir.mkdir(exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        obj = objdir / (Path(src).stem + ".o")
        objs.append(obj)
        cmd = [cc, *HIPCC_FLAGS, *SOURCE_FLAGS.get(src, []), "-I", str(PKG.parent / "include"), "-c", str(CSRC / src),
               "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((src, out))
        elif verbose and out.strip():
            print(out, flush=True)
    if failed:
        for src, out in failed:
            sys.stderr.write(f"--- {src}\n{out}\n")
        raise RuntimeError("hipcc failed for: " + ", ".join(s for s, _ in failed))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
