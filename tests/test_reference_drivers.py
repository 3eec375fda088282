"""The reference's own benchmark drivers (src/benchmarks/rings.cpp and
tensile.cpp) and its main program (src/mfree_iwf.cpp: the stepping loop,
contact calls and the results writer every 10th step), compiled unchanged
against include/mfree and linked with the library
(tests/cpp/build_reference_drivers.sh, in a temporary directory: nothing of
the reference enters the repository or travels with it).  A probe program
constructs the drivers' bodies on the CPU and reports which library method
each would step with; the reference's main, which steps forever, is linked
only."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SRC = "/root/reference/src"

pytestmark = pytest.mark.skipif(not os.path.isdir(os.path.join(REF_SRC, "benchmarks")),
                                reason="reference tree not present")


def test_reference_drivers_compile_link_and_map_to_pipelines(tmp_path):
    lib = os.path.join(ROOT, "mfree_iwf_amd", "libmfree_iwf_amd.so")
    if not os.path.exists(lib):
        pytest.skip("library not built")
    subprocess.run([os.path.join(ROOT, "tests", "cpp", "build_reference_drivers.sh"), str(tmp_path), REF_SRC],
                   check=True, capture_output=True, text=True)
    # the reference's main program links unchanged against the surface
    assert os.access(tmp_path / "reference_main", os.X_OK)
    out = subprocess.run([str(tmp_path / "reference_drivers")], check=True, capture_output=True, text=True).stdout
    got = {}
    for line in out.splitlines():
        name, *bodies = line.split()
        got[name] = [tuple(int(v) for v in b.split(":")) for b in bodies]
    UL, TL, GD, CR, PT = 0, 1, 2, 3, 4
    # the library's pipelines
    assert [m for m, _ in got["rings_monaghan"]] == [UL] and got["rings_monaghan"][0][1] > 0
    assert [m for m, _ in got["tensile_tl"]] == [TL]
    assert [m for m, _ in got["rings_godunov"]] == [GD] and got["rings_godunov"][0][1] > 0
    # two rings as separate bodies (their contact runs in the main loop,
    # contact.h)
    assert [m for m, _ in got["rings_monaghan_contact"]] == [UL, UL]
    assert [m for m, _ in got["rings_godunov_contact"]] == [GD, GD]
    assert [m for m, _ in got["tensile_corot"]] == [CR]
    assert [m for m, _ in got["tensile_potential"]] == [PT]
    assert [m for m, _ in got["rings_corot_contact"]] == [CR, CR]
    assert [m for m, _ in got["rings_potential_contact"]] == [PT, PT]
    # methods outside the library are refused, not run on the CPU
    for name in ("rings_tl_contact",
                 # UL with boundary conditions
                 "tensile_monaghan", "tensile_godunov"):
        assert all(m == -1 for m, _ in got[name]), name
