"""Shared helpers: the reference driver's probes on the CASE inputs, check.sh
tolerances, and small random fields for parity tests."""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

with open(os.path.join(HERE, "golden", "case_norms.json")) as _f:
    CASES = json.load(_f)


def case_geometry(case):
    a = CASES[case]["argv"]
    nloc, npe = a[0:4], a[4:8]
    L = [nloc[m] * npe[m] for m in range(4)]
    B = [nloc[0] // 2, nloc[1], nloc[2], nloc[3]]
    return nloc, npe, L, B


def check_value(name, got, want):
    """check.sh:17-26: rnorm^2 absolute < 1e-15; *_s_/_s relative < 3e-6; else 1e-14."""
    if name == "rnorm^2":
        return got < 1e-15
    tol = 3e-6 if (name.endswith("_s_") or name.endswith("_s")) else 1e-14
    return abs(got - want) / abs(want) < tol


def expected(case, name, section=""):
    for n, v, sec in CASES[case]["norms"]:
        if n == name and (not section or sec == section):
            return v
    raise KeyError(name)


def xdomain0_mask(L, nloc):
    V = L[0] * L[1] * L[2] * L[3]
    x = np.arange(V) % L[0]
    return (x % nloc[0]) < nloc[0] // 2


def random_su3_like(V, rng):
    """Random complex 3x3 links (not unitary) and a Hermitian-ish clover, for
    operator parity tests: U (4,V,3,3,2), C (V,2,36) with diagonal ~1."""
    U = rng.standard_normal((4, V, 3, 3, 2)) * 0.5
    C = rng.standard_normal((V, 2, 36)) * 0.1
    C[:, :, :6] += 0.5
    return U, C
