"""The FP16-stored SAP method near the critical kappa.

With SAP iterates, links and clover stored in FP16 (the method of bicgstab_dd_mix2_hf_,
bicgstab_dd_mix2_hf.cc:110-378), the inner BiCGStab can end with a residual larger
than its source when D M_SAP has eigenvalues comparable to the FP16 rounding of M_SAP:
on the CPU oracle (8^4, eps 0.25, csw 1.1, kappa 0.1548) the FP16 inner solves end at
1.3-6.4 times their source norm and the FP64 outer iteration does not converge, while
the FP32 SAP converges.  The solver then switches that solve to the FP32 SAP with the
same number of sweeps (sap_escalated) and reaches the FP64 true residual; far from
the critical kappa it stays on FP16.  Every result is checked with the oracle's FP64 operator and against an
all-FP64 solve (plain BiCGStab on D, bicgstab_dd_d.cc:64)."""
import numpy as np
import pytest

import qws_amd as Q
from oracle import oracle as O

pytestmark = pytest.mark.gpu

HP = dict(nsap=3, nm=6, tol=1e-11, maxiter=40, tol_s=1e-3, maxiter_s=100, true_tol=1e-10, max_restarts=6,
          sap_precision=Q.F16)


def _setup(L, kappa):
    lat = Q.Lattice(L, [4, 4, 4, 4], -1, 12, kappa=kappa)
    lat.generate_gauge(0.25, 3)
    lat.compute_clover(1.1)
    return lat, lat.gaussian(Q.F64, 4)


def _check(lat, x, b):
    orc = O.Oracle(lat.L, [4, 4, 4, 4], -1, lat.kappa, lat.gauge(), lat.clover_field())
    bb, xx = b.download(), x.download()
    res = float(np.sqrt(O.norm2(bb - orc.dslash(xx)) / O.norm2(bb)))
    # all-FP64 reference solve, started from the mixed solution
    xr = lat.spinor(Q.F64, xx)
    lat.bicgstab_dd(xr, b, 4000, 1e-14)
    ref = xr.download()
    return res, float(np.linalg.norm(xx - ref) / np.linalg.norm(ref))


@pytest.mark.parametrize("L,kappa,escalates", [([8, 8, 8, 8], 0.1545, True), ([16, 16, 16, 16], 0.144, False)])
def test_fp16_sap_method_reaches_1e10(L, kappa, escalates):
    print(f"[precision] {L} kappa={kappa}", flush=True)
    lat, b = _setup(L, kappa)
    x = lat.spinor(Q.F64)
    st = lat.solve(x, b, **HP)
    print(f"[precision] {st}", flush=True)
    assert st["sap_escalated"] == (1 if escalates else 0), st
    assert st["true_residual"] <= 1e-10, st
    res, err = _check(lat, x, b)
    assert res <= 1e-10, res
    assert err <= 1e-8, err
    lat.close()
