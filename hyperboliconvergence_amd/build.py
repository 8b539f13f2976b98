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

KERNELS = [(0, 0, 0), (0, 0, 1)] + [(1, b, f) for b in (0, 1) for f in range(7)]


def _mixed_wide(law, flux):
    """The mixed path keeps FP64 face states for these (hc_point.hpp mixed_wide)."""
    return law == 1 and flux in (0, 1, 2, 5, 6)


def _jobs(mixed_contract="fast", build_dir=None, waves_per_eu=(0, 0), mixed_fast_div="auto"):
    objs = []
    bdir = build_dir or BUILD
    for (law, basis, flux) in KERNELS:
        for prec in (64, 32):
            # FP64 objects: no contraction (bit-faithful); mixed objects: contraction and the
            # hardware FP32 reciprocal / square root instead of the correctly rounded sequences,
            # with FP32 denormals flushed (no range reduction around them; the mixed path's FP32
            # quantities are jumps, slopes and dissipation terms far above FLT_MIN)
            fast_div = {"all": True, "none": False}.get(mixed_fast_div, True)
            fpflags = (["-ffp-contract=off"] if prec == 64 else
                       [f"-ffp-contract={mixed_contract}"] +
                       (["-fno-hip-fp32-correctly-rounded-divide-sqrt", "-fgpu-flush-denormals-to-zero"]
                        if fast_div else []))
            wpe = waves_per_eu[0] if prec == 64 else waves_per_eu[1]
            if wpe:
                fpflags.append(f"-DHC_STAGE_WAVES_PER_EU={wpe}")
            out = os.path.join(bdir, f"stage_L{law}_B{basis}_F{flux}_P{prec}.o")
            cmd = [HIPCC, *COMMON, *fpflags, f"-DHC_LAW={law}", f"-DHC_BASIS={basis}", f"-DHC_FLUX={flux}",
                   f"-DHC_PREC={prec}", "-x", "hip", "-c", os.path.join(CSRC, "hc_kernel_inst.hip"), "-o", out]
            objs.append((out, cmd))
    for src in ("hc_geom.cpp", "hc_capi.cpp", "hc_adjoint.cpp", "hc_adjoint_inst.hip", "hc_spatial.hip"):
        out = os.path.join(bdir, src.replace(".cpp", ".o").replace(".hip", ".o"))
        cmd = [HIPCC, *COMMON, "-ffp-contract=off", "-x", "hip", "-c", os.path.join(CSRC, src), "-o", out]
        objs.append((out, cmd))
    return objs


def _run(job):
    out, cmd = job
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {os.path.basename(out)}:\n{' '.join(cmd)}\n{r.stderr[-4000:]}")
    return out


def build(jobs=None, verbose=False, lib=LIB, mixed_contract="fast", build_dir=None, waves_per_eu=(0, 0),
          mixed_fast_div="auto"):
    """Compile every HIP source and link libhcamd.so; always recompiles.
    waves_per_eu: occupancy hint (waves per SIMD) of the FP64 / mixed stage kernels, 0 = none.
    mixed_fast_div: hardware FP32 reciprocal / square root in the mixed objects ("auto" = "all"),
    or "none" for the correctly rounded sequences."""
    bdir = build_dir or BUILD
    os.makedirs(bdir, exist_ok=True)
    work = _jobs(mixed_contract, bdir, waves_per_eu, mixed_fast_div)
    n = jobs or min(8, os.cpu_count() or 4)
    with ThreadPoolExecutor(max_workers=n) as ex:
        objs = list(ex.map(_run, work))
    tmp = lib + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, lib)
    if verbose:
        print(f"built {lib} from {len(objs)} objects")
    return lib


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--out", default=LIB)
    ap.add_argument("--mixed-contract", default="fast", choices=["fast", "on", "off"])
    ap.add_argument("--waves-per-eu", type=int, nargs=2, default=(0, 0), metavar=("FP64", "MIXED"))
    ap.add_argument("--mixed-fast-div", default="auto", choices=["auto", "all", "none"])
    a = ap.parse_args()
    build(a.jobs, verbose=True, lib=a.out, mixed_contract=a.mixed_contract,
          build_dir=None if a.out == LIB else os.path.join(BUILD, "alt"), waves_per_eu=tuple(a.waves_per_eu),
          mixed_fast_div=a.mixed_fast_div)
    sys.exit(0)
