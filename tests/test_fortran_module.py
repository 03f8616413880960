"""The Fortran module cosmokdtree_gpu compiles for every combination of the
reference's compile-time switches (src/Makefile:6-31: PERIODIC,
DOUBLEPRECISION, LONGINT; DIMEN 1, 3 and 6), and exports the public names
of the reference's module cosmokdtree (src/cosmokdtree.f90:4-6).  CPU only:
syntax and semantics checks, no run."""
import itertools
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "cosmokdtree_amd", "fortran", "cosmokdtree_gpu.F90")
REF_PUBLIC = {"build_kdtree", "deallocate_kdtree", "knn_search", "ball_search", "box_search", "KDTreeNode",
              "KDTreeResult", "deallocate_query"}


def test_public_names():
    text = open(SRC).read()
    m = re.search(r"public ::(.*?)\n\n", text, re.S)
    names = {x.strip() for x in m.group(1).replace("&", ",").split(",") if x.strip()}
    assert REF_PUBLIC <= names


@pytest.mark.skipif(shutil.which("amdflang") is None, reason="amdflang not installed")
@pytest.mark.parametrize("periodic,dp,longint,dimen",
                         [(p, d, li, dim) for p, d, li in itertools.product((0, 1), repeat=3) for dim in (3,)]
                         + [(0, 1, 0, 1), (1, 0, 1, 6)])
def test_module_compiles(tmp_path, periodic, dp, longint, dimen):
    r = subprocess.run(["amdflang", "-cpp", "-fsyntax-only", f"-Dperiodic={periodic}", f"-Ddimen={dimen}",
                        f"-Ddoubleprecision={dp}", f"-Dlongint={longint}", "-module-dir", str(tmp_path), SRC],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
