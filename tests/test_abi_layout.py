"""The ctypes mirrors of the C ABI's structures (mfree_iwf_amd/_capi.py)
against the C header itself: a small C program compiled with gcc against
include/mfree_iwf_amd.h prints sizeof and offsetof of every field, which must
equal the ctypes layout (a field added on one side only would shift every
later field on the other)."""
import os
import shutil
import subprocess

import pytest

from mfree_iwf_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# C typedef -> ctypes structure
PAIRS = [
    ("mfree_params", _capi.Params),
    ("mfree_options", _capi.Options),
    ("mfree_particles", _capi.Particles),
    ("mfree_rings_spec", _capi.RingsSpec),
    ("mfree_plate_spec", _capi.PlateSpec),
    ("mfree_surrogate_options", _capi.SurrogateOptions),
    ("mfree_judge_stats", _capi.JudgeStats),
]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="no C compiler")
def test_ctypes_structures_match_the_header(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "mfree_iwf_amd.h"', "int main(void) {"]
    for cname, cls in PAIRS:
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in cls._fields_:
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for line in out:
        if line:
            name, field, val = line.split()
            got[(name, field)] = int(val)
    for cname, cls in PAIRS:
        assert got[(cname, "size")] == _capi.C.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert got[(cname, f)] == getattr(cls, f).offset, (cname, f)
