"""The solver's autotuners: qwsg_autotune_solve times every candidate (nsap, nm) pair
on a real solve of the lattice, returns the fastest pair that meets the FP64
true-residual target, and a solve with the returned settings meets it too;
qwsg_autotune_blocks does the same over SAP block shapes."""
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


def test_autotune_blocks_solves_the_same_problem_on_every_shape():
    """qwsg_autotune_blocks: the candidate lattices get this lattice's fields and source
    through the plain site order, so each admissible shape solves the same system; shapes
    that do not tile the extents are marked, and the returned shape is the fastest."""
    L = [16, 8, 8, 16]
    lat = Q.Lattice(L, [4, 4, 4, 4], -1, 12, kappa=0.13)
    lat.generate_gauge(0.3, 22)
    lat.compute_clover(1.0)
    b = lat.gaussian(Q.F64, 6)
    prm = dict(tol=1e-11, maxiter=40, tol_s=1e-3, maxiter_s=80, true_tol=1e-10, max_restarts=4)
    shapes = [[4, 4, 4, 4], [8, 4, 4, 2], [8, 4, 2, 4], [4, 4, 4, 8], [8, 4, 4, 4], [3, 4, 4, 4], [16, 4, 4, 4]]
    best, sec = lat.autotune_blocks(b, shapes, **prm)
    print("[autotune_blocks] seconds", [round(s, 4) for s in sec], "best", best, flush=True)
    assert all(s > 0 for s in sec[:5])
    assert sec[5] < 0 and sec[6] < 0  # 3 does not divide 16; one block in x is refused
    assert best == shapes[int(np.argmax([-s if s > 0 else -1e300 for s in sec]))]
    # the chosen shape solves the original lattice's system to the same solution
    x0 = lat.spinor(Q.F64)
    assert lat.solve(x0, b, **prm)["true_residual"] <= 1e-10
    ref = x0.download()
    c = Q.Lattice(L, best, -1, 12, kappa=0.13)
    c.load_gauge(lat.gauge())
    c.load_clover(lat.clover_field())
    bc = c.spinor(Q.F64, b.download())
    xc = c.spinor(Q.F64)
    assert c.solve(xc, bc, **prm)["true_residual"] <= 1e-10
    err = np.linalg.norm(xc.download() - ref) / np.linalg.norm(ref)
    print("[autotune_blocks] relative difference to the 4^4 solution", err, flush=True)
    assert err < 1e-6  # |r|/|b| <= 1e-10 times the operator's condition number
    c.close()
    lat.close()


def test_autotune_blocks_with_18_real_links_and_bad_candidates():
    """With 18-real FP32 links the 512-site shapes run the unfused SAP (the fused kernel's
    LDS holds 12-real links at that size): still admissible and converging.  An empty
    candidate list and a lattice without FP64 fields are refused."""
    L = [16, 8, 8, 8]
    lat = Q.Lattice(L, [4, 4, 4, 4], -1, 18, kappa=0.12)
    lat.generate_gauge(0.3, 23)
    lat.compute_clover(1.0)
    b = lat.gaussian(Q.F64, 7)
    prm = dict(tol=1e-11, maxiter=40, tol_s=1e-3, maxiter_s=80, true_tol=1e-10, max_restarts=4)
    best, sec = lat.autotune_blocks(b, [[4, 4, 4, 4], [8, 4, 4, 4]], **prm)
    print("[autotune_blocks 18-real] seconds", [round(s, 4) for s in sec], "best", best, flush=True)
    assert all(s > 0 for s in sec)
    with pytest.raises(Q.QwsError):
        lat.autotune_blocks(b, [], **prm)
    bare = Q.Lattice(L, [4, 4, 4, 4], -1, 18, kappa=0.12)
    bare.load_gauge_f32(lat.gauge().astype(np.float32))
    with pytest.raises(Q.QwsError):
        bare.autotune_blocks(bare.gaussian(Q.F64, 7), [[4, 4, 4, 4]], **prm)
    bare.close()
    lat.close()
