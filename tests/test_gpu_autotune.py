"""The solver's autotuner (qwsg_autotune_solve): every candidate (nsap, nm) pair is timed
on a real solve of the lattice, the fastest pair that meets the FP64 true-residual
target is returned, and a solve with the returned settings meets it too."""
import numpy as np
import pytest

import qws_amd as Q

pytestmark = pytest.mark.gpu


def test_autotune_solve_picks_the_fastest_candidate_that_converges():
    lat = Q.Lattice([8, 8, 8, 16], [4, 4, 4, 4], -1, 12, kappa=0.13)
    lat.generate_gauge(0.3, 21)
    lat.compute_clover(1.0)
    b = lat.gaussian(Q.F64, 5)
    x = lat.spinor(Q.F64)
    prm = dict(tol=1e-11, maxiter=40, tol_s=1e-3, maxiter_s=80, true_tol=1e-10, max_restarts=4)
    best, sec = lat.autotune_solve(x, b, [1, 2, 3], [2, 4, 6], **prm)
    s = np.array(sec)
    print("[autotune] seconds", s.round(4).tolist(), "best", best["nsap"], best["nm"], flush=True)
    assert s.shape == (3, 3) and (s > 0).all()
    i, j = np.unravel_index(np.argmin(s), s.shape)
    assert (best["nsap"], best["nm"]) == ([1, 2, 3][i], [2, 4, 6][j])
    x.zero()
    st = lat.solve(x, b, **best)
    assert st["true_residual"] <= 1e-10
    # a candidate outside QWS's limits (qws.cc:798-799) is refused
    with pytest.raises(Q.QwsError):
        lat.autotune_solve(x, b, [11], [2], **prm)
    lat.close()
