"""The suffixed LIGGGHTS-INL styles that bind the C-ABI (liggghts-inl_amd/
styles: verlet/mi355x, gran/mi355x, wall/gran/mi355x, mesh/surface/mi355x,
move/mesh/mi355x, nve/sphere/mi355x, setforce/mi355x, aveforce/mi355x)
compile against the reference's own headers: every reference member they
use exists with the type they use, and every C-ABI call matches
include/liggghts_dem.h.  Each header also expands to its style macro, as
`Make.sh style` collects it into style_*.h.  CPU only; skipped where the
reference sources are absent.
"""
import glob
import os
import re
import subprocess

import pytest

from conftest import ROOT

REF_SRC = "/root/reference/src"
STYLES = os.path.join(ROOT, "liggghts-inl_amd", "styles")

pytestmark = pytest.mark.skipif(not os.path.isdir(REF_SRC), reason="reference sources not present")


def gxx(path, extra=()):
    cmd = ["g++", "-std=c++11", "-Wall", "-Werror", "-fsyntax-only", "-I", REF_SRC, "-I",
           os.path.join(REF_SRC, "STUBS"), "-I", os.path.join(ROOT, "include"), "-I", STYLES, *extra, path]
    return subprocess.run(cmd, capture_output=True, text=True)


@pytest.mark.parametrize("src", sorted(os.path.basename(p) for p in glob.glob(os.path.join(STYLES, "*.cpp"))))
def test_style_compiles_against_reference_headers(src):
    r = gxx(os.path.join(STYLES, src))
    assert r.returncode == 0, r.stderr[-3000:]


MACROS = {"IntegrateStyle": "INTEGRATE_CLASS", "PairStyle": "PAIR_CLASS", "FixStyle": "FIX_CLASS"}


@pytest.mark.parametrize("hdr", sorted(os.path.basename(p) for p in glob.glob(os.path.join(STYLES, "*.h"))))
def test_style_macro(hdr, tmp_path):
    text = open(os.path.join(STYLES, hdr)).read()
    found = re.findall(r"^(IntegrateStyle|PairStyle|FixStyle)\(([^,]+),(\w+)\)", text, re.M)
    assert found, "no style macro"
    for macro, key, cls in found:
        assert key.endswith("/mi355x")
        # the class the macro names is defined in the header's other branch
        assert re.search(r"class %s\b" % cls, text)
        probe = tmp_path / "probe.cpp"
        # one variable per macro line: a class may serve several styles, as
        # FixMeshSurface serves mesh/surface and mesh/surface/planar
        probe.write_text('#define %s\n#define CAT2(a,b) a##b\n#define CAT(a,b) CAT2(a,b)\n'
                         '#define %s(key,Class) static const char *CAT(k_##Class, __LINE__) = #key;\n'
                         '#include "%s"\n' % (MACROS[macro], macro, hdr))
        r = gxx(str(probe), extra=("-Wno-unused-variable",))
        assert r.returncode == 0, r.stderr[-2000:]
