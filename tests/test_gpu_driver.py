"""QWS's own test driver (main.cc) linked unchanged against libqws_mi355x.so.

oracle/Makefile.ref links the reference driver's object (compiled from main.cc where
it lies) against this library instead of the reference objects
(oracle/_ref/qws_main_mi355x).  Run with data/CASE0's arguments it initialises QWS
through qws_init_, fills glud/glus/clvd/clvs through the exported pointers, and prints
every norm that check.sh compares: the even/odd family, the DD operators, the clover
multiplies, the even/odd BiCGStab and the mixed-precision DD solve
(main.cc:200-312,584-795).  Each printed value must match data/CASE0 to check.sh's
tolerances (check.sh:17-26) and in the same order and section.  The binary exists only
where the reference sources were present when the repository was built.
"""
import os
import re
import subprocess

import pytest

from tests import casekit as K

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "qws_main_mi355x")


def _norm_lines(text):
    """(name, value, section) of the printed norms, sectioned as data/CASE* is."""
    pat = re.compile(r" : +[0-9]|2 += +[0-9]")
    out, section = [], ""
    for line in text.splitlines():
        if line.startswith("LDDHMC solver"):
            section = "HP" if "Half Precision" in line else "DD"
        elif line.startswith("test bicgstab"):
            section = "EO"
        if "git commit hash" in line or not pat.search(line):
            continue
        out.append((line.split()[0], float(line.split()[-1]), section))
    return out


@pytest.mark.skipif(not os.path.exists(BIN), reason="oracle/_ref/qws_main_mi355x not built (no reference sources)")
def test_reference_driver_on_the_library_reproduces_case0():
    argv = [str(a) for a in K.CASES["CASE0"]["argv"]]
    print("running", BIN, " ".join(argv), flush=True)
    p = subprocess.run([BIN] + argv, capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    got = _norm_lines(p.stdout)
    want = [tuple(w) for w in K.CASES["CASE0"]["norms"]]
    assert [(n, s) for n, _, s in got] == [(n, s) for n, _, s in want], p.stdout[-3000:]
    for (name, v, sec), (_, w, _) in zip(got, want):
        assert K.check_value(name, v, w), (name, sec, v, w)
