"""Every `file:line` citation of the reference in the sources points inside the cited file.

The oracle, the kernels, the C-ABI header and the drop-in layer cite the reference's files
(relative to its cpp/inc/, cpp/ or repository root) for each function they restate.  This test
resolves each citation against the reference tree and fails when a cited line lies outside the
file, and checks that the citations of the quirks that shape the benchmark cases land on the
reference lines that hold them.  Skipped where the reference tree is absent.
"""
import collections
import os
import re

import pytest

REF = "/root/reference"
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "cpp", "inc")),
                                reason="reference tree not present")

SOURCES = ["oracle", "include", "hyperboliconvergence_amd", "tests", "bench.py", "__graft_entry__.py", "DESIGN.md",
           "INTEGRATION.md", "README.md"]
EXT = (".cpp", ".hpp", ".h", ".hip", ".py", ".md")
CITE = re.compile(r"((?:[\w.+-]+/)*[\w.+-]+\.(?:h|ipp|hpp|cpp|py|gnp))((?::\d+(?:-\d+)?)(?:,\s?\d+(?:-\d+)?)*)")


def _source_files():
    for s in SOURCES:
        p = os.path.join(ROOT, s)
        if os.path.isfile(p):
            yield p
            continue
        for dp, _, fns in os.walk(p):
            for fn in fns:
                if fn.endswith(EXT) and fn != os.path.basename(__file__):
                    yield os.path.join(dp, fn)


def _reference_index():
    idx = collections.defaultdict(list)
    for dp, dn, fns in os.walk(REF):
        dn[:] = [d for d in dn if d != ".git"]
        for fn in fns:
            idx[fn].append(os.path.join(dp, fn))
    return idx


def _resolve(path, idx):
    for root in (os.path.join(REF, "cpp", "inc"), os.path.join(REF, "cpp"), REF):
        q = os.path.join(root, path)
        if os.path.isfile(q):
            return [q]
    return [c for c in idx.get(os.path.basename(path), []) if c.endswith("/" + path) or c.endswith(path)]


def _lines(path, cache={}):
    if path not in cache:
        with open(path, errors="replace") as f:
            cache[path] = f.read().split("\n")
    return cache[path]


def _citations():
    idx = _reference_index()
    out = []
    for fn in _source_files():
        with open(fn, errors="replace") as f:
            for ln, line in enumerate(f, 1):
                for m in CITE.finditer(line):
                    cands = _resolve(m.group(1), idx)
                    if not cands:
                        continue
                    ranges = []
                    for part in m.group(2)[1:].split(","):
                        a, _, b = part.strip().partition("-")
                        ranges.append((int(a), int(b) if b else int(a)))
                    out.append((os.path.relpath(fn, ROOT), ln, m.group(0), cands, ranges))
    return out


def test_every_citation_lies_inside_the_cited_file():
    cites = _citations()
    assert len(cites) > 200   # the sources cite the reference throughout
    bad = []
    for fn, ln, text, cands, ranges in cites:
        if not any(all(1 <= a <= b <= len(_lines(c)) for a, b in ranges) for c in cands):
            bad.append(f"{fn}:{ln}: {text} ({[len(_lines(c)) for c in cands]} lines)")
    assert not bad, "citations outside the cited file:\n" + "\n".join(bad)


# citation -> text that the first cited range must contain
ANCHORS = {
    "base/fluxes/roeFlux.ipp:40-54": ("roeAverage", "rotateToFace", "rotateFromFace"),
    "euler/fluxes/roe.ipp:50-78": ("eps=0.05",),
    "euler/fluxes/slau.ipp:4-161": ("sqrt( 0.5*( sl.velocity2()", "-ma*( betal + betar - 1. )"),
    "euler/fluxes/ausmPlusUP.ipp:6-171": ("Ku     = 0.75*2.", "Kp     = 0.25", "lmd= 0.5*(m0_p+1)/fa_p"),
    "euler/fluxes/machSplittings.ipp:8-43": ("numeric_limits<Real>::max()",),
    "limiter.ipp:81-89": ("VanLeer2::limit",),
    "limiter.ipp:26-37": ("VanAlbada2::limit",),
    "limiter.ipp:39-47": ("MinMod2::limit",),
    "euler/boundary/entropy.ipp:18-88": ("Entropy", "pow( pb/sb, 1./gam )"),
    "euler/boundary/riemann.ipp:18-113": ("Riemann",),
    "euler/boundary/inviscidWall.ipp:54-57": ("qinterior + direction*dq",),
    "spatial/lsqMetrics.h:181-213": ("qmetrics",),
    "spatial/lsqMetrics.h:104-163": ("boundary_xmetrics", "boundary_point"),
    "spatial/eulerForwardUpdate.h:155-160": ("cfl/res.lambda",),
    "spatial/eulerForwardUpdate.h:76-85": ("dt/cell.volume",),
    "spatial/spectralRadius.h:19-48": ("numeric_limits<Real>::min()",),
    "geometry/volume.ipp:25-66": ("fem::dn2",),
    "mesh/generate/twoD.h:139-179": ("make_cylinder_mesh",),
    "spatial/boundary/periodicLoop.h:129-176": ("boundaryId==0", "boundaryId==2"),
    "lsq/factor_solve.ipp:71-100": ("bias_solves",),
}


@pytest.mark.parametrize("cite", sorted(ANCHORS))
def test_quirk_citations_land_on_the_reference_lines(cite):
    path, _, rng = cite.rpartition(":")
    a, _, b = rng.partition("-")
    (f,) = _resolve(path, _reference_index())
    text = "\n".join(_lines(f)[int(a) - 1:int(b)])
    for needle in ANCHORS[cite]:
        assert needle in text, f"{cite} does not hold {needle!r}"
    used = [c for c in _citations() if cite in c[2] or c[2].endswith(cite)]
    assert used, f"{cite} is not cited anywhere"
