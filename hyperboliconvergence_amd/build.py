"""Build the in-tree HIP library libhcamd.so for gfx950 (hipcc, no JIT cache).

  python -m hyperboliconvergence_amd.build [--jobs N]

Every stage-kernel instantiation (law x solution basis x flux x precision) is its
own object so the build runs in parallel; FP64 objects and the host geometry are
compiled with -ffp-contract=off (bit-faithful arithmetic), the mixed-precision
objects with contraction allowed.
"""
import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(ROOT, "_build", "hcamd")
LIB = os.path.join(PKG, "libhcamd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = ["-std=c++17", "-O3", f"--offload-arch={ARCH}", "-fPIC", "-I", CSRC, "-I", INCLUDE,
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result"]

KERNELS = [(0, 0, 0), (0, 0, 1)] + [(1, b, f) for b in (0, 1) for f in range(5)]


def _jobs():
    objs = []
    for (law, basis, flux) in KERNELS:
        for prec in (64, 32):
            contract = "-ffp-contract=off" if prec == 64 else "-ffp-contract=fast"
            out = os.path.join(BUILD, f"stage_L{law}_B{basis}_F{flux}_P{prec}.o")
            cmd = [HIPCC, *COMMON, contract, f"-DHC_LAW={law}", f"-DHC_BASIS={basis}", f"-DHC_FLUX={flux}",
                   f"-DHC_PREC={prec}", "-x", "hip", "-c", os.path.join(CSRC, "hc_kernel_inst.hip"), "-o", out]
            objs.append((out, cmd))
    for src in ("hc_geom.cpp", "hc_capi.cpp"):
        out = os.path.join(BUILD, src.replace(".cpp", ".o"))
        cmd = [HIPCC, *COMMON, "-ffp-contract=off", "-x", "hip", "-c", os.path.join(CSRC, src), "-o", out]
        objs.append((out, cmd))
    return objs


def _run(job):
    out, cmd = job
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {os.path.basename(out)}:\n{' '.join(cmd)}\n{r.stderr[-4000:]}")
    return out


def build(jobs=None, verbose=False):
    """Compile every HIP source and link libhcamd.so; always recompiles."""
    os.makedirs(BUILD, exist_ok=True)
    work = _jobs()
    n = jobs or min(8, os.cpu_count() or 4)
    with ThreadPoolExecutor(max_workers=n) as ex:
        objs = list(ex.map(_run, work))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(objs)} objects")
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    a = ap.parse_args()
    build(a.jobs, verbose=True)
    sys.exit(0)
