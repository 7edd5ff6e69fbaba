"""Solves over several lattices at once (qwsg_solve_multi): lattices of different
volumes, fields and kappas advance in lockstep, every SAP half-sweep one launch over
the blocks of all of them.  Each right-hand side must come out bit for bit as the
lattice's own solve gives it (same kernels per block, per-lattice reductions), with
the same iteration counts, and meet its FP64 true residual, recomputed by the CPU
oracle's operator on the downloaded solution (the gate of data/CASE0:61-63)."""
import numpy as np
import pytest

import qws_amd as Q
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SOLVE = dict(nsap=2, nm=6, tol=1e-11, maxiter=40, tol_s=1e-3, maxiter_s=80, true_tol=1e-10, max_restarts=4)


def _lattice(L, kappa, eps, seed, links=12):
    lat = Q.Lattice(L, [4, 4, 4, 4], -1, links, kappa=kappa)
    lat.generate_gauge(eps, seed)
    lat.compute_clover(1.0)
    return lat


def _residual(lat, x, b):
    orc = O.Oracle(lat.L, [4, 4, 4, 4], -1, lat.kappa, lat.gauge(), lat.clover_field())
    bb = b.download()
    return float(np.sqrt(O.norm2(bb - orc.dslash(x.download())) / O.norm2(bb)))


@pytest.mark.parametrize("shapes", [
    # (extents, kappa, RHS count): uneven volumes, one with several RHS
    [([8, 8, 8, 8], 0.128, 1), ([8, 8, 8, 16], 0.134, 3), ([16, 8, 8, 8], 0.131, 2)],
    [([16, 16, 16, 16], 0.1365, 1), ([8, 8, 8, 8], 0.125, 2)],
])
def test_fused_multi_lattice_solve_is_each_lattice_alone(shapes):
    lats, parts = [], []
    for i, (L, kappa, n) in enumerate(shapes):
        print(f"[multi] lattice {L} kappa {kappa} nrhs {n}", flush=True)
        lat = _lattice(L, kappa, 0.3, 20 + i)
        bs = [lat.gaussian(Q.F64, 100 * i + k) for k in range(n)]
        xs = [lat.spinor(Q.F64) for _ in range(n)]
        lats.append(lat)
        parts.append((lat, xs, bs))
    fused = Q.solve_multi(parts, **SOLVE)
    for (lat, xs, bs), st in zip(parts, fused):
        alone_x = [lat.spinor(Q.F64) for _ in xs]
        alone = lat.solve_batch(alone_x, bs, **SOLVE)
        for k in range(len(xs)):
            a, f = alone_x[k].download(), xs[k].download()
            assert np.array_equal(a, f), (lat.L, k)
            assert st[k]["inner_iters"] == alone[k]["inner_iters"] and st[k]["outer_iters"] == alone[k]["outer_iters"]
            assert st[k]["true_residual"] <= 1e-10
        assert _residual(lat, xs[0], bs[0]) <= 1e-10
    for lat in lats:
        lat.close()


def test_fused_multi_lattice_solve_rejects_bad_input():
    a = _lattice([8, 8, 8, 8], 0.12, 0.3, 3)
    x, b = a.spinor(Q.F64), a.gaussian(Q.F64, 1)
    with pytest.raises(Q.QwsError):
        Q.solve_multi([(a, [x], [b]), (a, [a.spinor(Q.F64)], [b])], **SOLVE)  # a lattice twice
    with pytest.raises(Q.QwsError):
        Q.solve_multi([(a, [x], [b])], **dict(SOLVE, nsap=11))
    a.close()
