"""The C-ABI library (libhcamd.so) loads without a GPU and exports exactly the entry points
declared in include/hcamd.h; the ctypes layout of hc_case_desc matches the C struct."""
import ctypes as C
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hcamd.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hc_[a-z_0-9]+)\s*\(", text)))


def test_header_declarations_match_python_binding():
    import hyperboliconvergence_amd as hca
    assert declared_symbols() == sorted(hca.C_ABI_SYMBOLS)


def test_library_exports_every_declared_symbol():
    import hyperboliconvergence_amd as hca
    assert os.path.exists(hca.LIBPATH), "build the library first (python -m hyperboliconvergence_amd.build)"
    out = subprocess.run(["nm", "-D", "--defined-only", hca.LIBPATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s+(hc_[a-z_0-9]+)$", out, flags=re.M))
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing


def test_library_loads_without_gpu_and_reports_version():
    import hyperboliconvergence_amd as hca
    L = hca.library()
    assert b"gfx950" in L.hc_version()


def test_case_struct_layout_matches_c():
    """Compile a tiny C program that prints sizeof/offsetof of hc_case_desc and compare with ctypes."""
    import hyperboliconvergence_amd as hca
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "hcamd.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(hc_case_desc), offsetof(hc_case_desc, alpha),
         offsetof(hc_case_desc, beta), offsetof(hc_case_desc, cfl), offsetof(hc_case_desc, period_j),
         offsetof(hc_case_desc, sp));
  return 0; }
'''
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "t")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        vals = list(map(int, subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()))
    D = hca.CaseDesc
    assert vals == [C.sizeof(D), D.alpha.offset, D.beta.offset, D.cfl.offset, D.period_j.offset, D.sp.offset]


def test_oracle_case_layout_matches_c_abi():
    import pyoracle as po
    import hyperboliconvergence_amd as hca
    assert C.sizeof(po.Case) == C.sizeof(hca.CaseDesc)
    for (n1, _), (n2, _) in zip(po.Case._fields_, hca.CaseDesc._fields_):
        assert n1 == n2 and getattr(po.Case, n1).offset == getattr(hca.CaseDesc, n2).offset


def test_context_requires_built_library_no_fallback(monkeypatch, tmp_path):
    import hyperboliconvergence_amd as hca
    monkeypatch.setattr(hca, "LIBPATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(hca, "_LIB", None)
    with pytest.raises(hca.HCError):
        hca.Context(0)


def test_every_entry_point_rejects_a_null_context():
    """The C ABI's error contract without a GPU: a null context handle returns HC_ERR_ARG from
    every entry point that takes one, hc_destroy(NULL) is a no-op (run in a child process, so a
    crash would fail the test instead of ending the run)."""
    code = r'''
import ctypes as C, sys
sys.path.insert(0, %r)
import hyperboliconvergence_amd as hca
L = hca.library()
bad = []
for name in hca.C_ABI_SYMBOLS:
    f = getattr(L, name)
    args = f.argtypes or []
    if name in ("hc_create", "hc_version") or not args or args[0] is not C.c_void_p:
        continue
    vals = []
    for t in args:
        vals.append(None if t is C.c_void_p else t(0) if t is not C.c_double else 0.0)
    vals[0] = None
    rc = f(*vals)
    if name == "hc_last_error":
        continue
    if name == "hc_destroy":   # like free(NULL): nothing to release
        if rc != 0:
            bad.append((name, rc))
        continue
    if rc != -1:
        bad.append((name, rc))
print("BAD", bad)
sys.exit(1 if bad else 0)
''' % os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
