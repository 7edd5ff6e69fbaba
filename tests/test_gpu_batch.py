"""Batched solves: several right-hand sides in one set of launches.

Each right-hand side keeps its own recurrence and its own deterministic
reductions, so a batched solve must reproduce each single solve bit for bit
("the same alone and in the mix"), and every solution must satisfy the FP64 true
residual target (checked again with the CPU oracle's operator)."""
import numpy as np
import pytest

import qws_amd as Q
from oracle import oracle as O

pytestmark = pytest.mark.gpu

PRM = dict(nsap=4, nm=2, tol=1e-11, maxiter=30, tol_s=1e-3, maxiter_s=30, true_tol=1e-10, max_restarts=3)


@pytest.fixture(scope="module")
def lat():
    L, B = [8, 8, 8, 16], [4, 4, 4, 4]
    lat = Q.Lattice(L, B, -1, 12, kappa=0.13)
    lat.generate_gauge(0.3, 11)
    lat.compute_clover(1.0)
    return lat


def _sources(lat, n):
    srcs = [lat.gaussian(Q.F64, 100 + k) for k in range(n - 2)]
    srcs.append(lat.point(Q.F64, (0, 0, 0, 0), 0, 0))
    srcs.append(lat.point(Q.F64, (3, 1, 2, 5), 2, 1))
    return srcs


def test_batch_equals_single_bitwise(lat):
    n = 5
    bs = _sources(lat, n)
    xs = [lat.spinor(Q.F64) for _ in range(n)]
    st = lat.solve_batch(xs, bs, **PRM)
    for k in range(n):
        x1 = lat.spinor(Q.F64)
        s1 = lat.solve(x1, bs[k], **PRM)
        assert np.array_equal(x1.download(), xs[k].download()), k
        assert s1["outer_iters"] == st[k]["outer_iters"] and s1["inner_iters"] == st[k]["inner_iters"]
        assert st[k]["true_residual"] <= 1e-10


def test_batch_with_initial_guess_and_oracle_residual(lat):
    n = 3
    bs = _sources(lat, n)
    xs = [lat.gaussian(Q.F64, 500 + k) for k in range(n)]
    st = lat.solve_batch(xs, bs, **PRM)
    U, C = lat.gauge(), lat.clover_field()
    orc = O.Oracle(lat.L, lat.B, -1, 0.13, U, C)
    for k in range(n):
        b = bs[k].download()
        r = b - orc.dslash(xs[k].download())
        assert np.sqrt(O.norm2(r) / O.norm2(b)) <= 1e-10
        assert st[k]["true_residual"] <= 1e-10


def test_twelve_point_sources(lat):
    """The 12 spin-colour point sources at the origin in one batch: every solution's
    residual recomputed with the oracle's operator, and distinct sources give distinct
    solutions."""
    bs = [lat.point(Q.F64, (0, 0, 0, 0), s, c) for c in range(3) for s in range(4)]
    xs = [lat.spinor(Q.F64) for _ in bs]
    st = lat.solve_batch(xs, bs, **PRM)
    assert all(s["true_residual"] <= 1e-10 for s in st)
    orc = O.Oracle(lat.L, lat.B, -1, 0.13, lat.gauge(), lat.clover_field())
    sols = [x.download() for x in xs]
    for k in range(12):
        b = bs[k].download()
        assert np.sqrt(O.norm2(b - orc.dslash(sols[k])) / O.norm2(b)) <= 1e-10, k
    assert not np.allclose(sols[0], sols[1])


def test_zero_source_in_a_batch(lat):
    """A zero source has the exact solution 0: it takes no iterations and leaves the
    other right-hand sides of the batch as they are alone (bit for bit)."""
    bs = _sources(lat, 3)
    zero = lat.spinor(Q.F64, np.zeros((lat.V, 3, 4, 2)))
    xs = [lat.spinor(Q.F64) for _ in range(4)]
    xs[1] = lat.spinor(Q.F64, np.ones((lat.V, 3, 4, 2)))  # content that must be replaced by 0
    st = lat.solve_batch(xs, [bs[0], zero, bs[1], bs[2]], **PRM)
    assert not np.any(xs[1].download())
    assert st[1]["outer_iters"] == 0 and st[1]["inner_iters"] == 0 and st[1]["true_residual"] == 0
    for k, kb in ((0, 0), (2, 1), (3, 2)):
        x1 = lat.spinor(Q.F64)
        lat.solve(x1, bs[kb], **PRM)
        assert np.array_equal(x1.download(), xs[k].download()), k
