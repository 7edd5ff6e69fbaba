"""The benchmark configurations at their full sizes, checked against the CPU oracle.

32^3x64 plain solve (links exp(0.3 i H), Philox seed 2, kappa = 0.1365, c_sw = 1.0,
antiperiodic t, 4^4 SAP blocks, 12-real FP32 links): the FP64 and FP32 operators and
the fused SAP preconditioner against the oracle on the whole lattice, then the mixed
solve to an FP64 true residual of 1e-10, recomputed by the oracle's FP64 operator on
the downloaded solution (gate: data/CASE0:61-63 and check.sh:17-26 check the final
solution of the reference's solve the same way).

The 12 spin-colour point sources at the origin on the same field, one batched solve:
every solution's residual recomputed by the oracle, and two of them equal to their
single solves bit for bit.

The 48^3x96 precision variant (FP16-stored SAP at kappa = 0.137): its true residual and
its error against an FP64-refined continuation.
"""
import numpy as np
import pytest

import qws_amd as Q
from oracle import oracle as O

pytestmark = pytest.mark.gpu

C2 = dict(L=[32, 32, 32, 64], B=[4, 4, 4, 4], eps=0.3, seed=2, kappa=0.1365, csw=1.0)
SOLVE = dict(nsap=2, nm=6, tol=1e-11, maxiter=40, tol_s=1e-3, maxiter_s=80, true_tol=1e-10, max_restarts=4)


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def oracle_residual(orc, x, b):
    return float(np.sqrt(O.norm2(b - orc.dslash(x)) / O.norm2(b)))


@pytest.fixture(scope="module")
def c2():
    print("[configs] 32^3x64 fields", flush=True)
    lat = Q.Lattice(C2["L"], C2["B"], -1, 12, kappa=C2["kappa"])
    lat.generate_gauge(C2["eps"], C2["seed"])
    lat.compute_clover(C2["csw"])
    orc = O.Oracle(C2["L"], C2["B"], -1, C2["kappa"], lat.gauge(), lat.clover_field())
    yield lat, orc
    lat.close()


def test_c2_operators_full_lattice(c2):
    lat, orc = c2
    b = lat.gaussian(Q.F64, 2).download()
    for prec, dt, tol in ((Q.F64, np.float64, 1e-13), (Q.F32, np.float32, 1e-6)):
        src = b.astype(dt)
        out = lat.spinor(prec)
        lat.dslash(out, lat.spinor(prec, src))
        assert rel(out.download(), orc.dslash(src)) < tol, prec
    # the fused FP32 SAP preconditioner of the solve (prec_ddd_s_ with nsap = 2, nm = 6)
    src = b.astype(np.float32)
    out = lat.spinor(Q.F32)
    lat.prec_ddd(out, lat.spinor(Q.F32, src), SOLVE["nsap"], SOLVE["nm"])
    assert rel(out.download(), orc.prec_ddd_s(src, SOLVE["nsap"], SOLVE["nm"])) < 5e-6


def test_c2_solve_to_1e10(c2):
    lat, orc = c2
    b = lat.gaussian(Q.F64, C2["seed"])
    x = lat.spinor(Q.F64)
    st = lat.solve(x, b, **SOLVE)
    assert st["true_residual"] <= 1e-10 and st["outer_iters"] >= 1
    res = oracle_residual(orc, x.download(), b.download())
    assert res <= 1e-10, res
    # the library's own FP64 residual agrees with the oracle's
    assert abs(res - st["true_residual"]) <= 1e-3 * res + 1e-15


def test_c4_twelve_point_sources_one_batch(c2):
    lat, orc = c2
    bs = [lat.point(Q.F64, (0, 0, 0, 0), sp, c) for c in range(3) for sp in range(4)]
    xs = [lat.spinor(Q.F64) for _ in bs]
    st = lat.solve_batch(xs, bs, **SOLVE)
    sols = []
    for k in range(12):
        x, b = xs[k].download(), bs[k].download()
        assert st[k]["true_residual"] <= 1e-10, k
        res = oracle_residual(orc, x, b)
        assert res <= 1e-10, (k, res)
        sols.append(x)
    # a point source's solution is a column of D^-1: the 12 are linearly independent
    # (Gram matrix of full rank), and two of them equal their single solves bit for bit
    S = np.stack([a.ravel() for a in sols])
    w = np.linalg.eigvalsh(S @ S.T)
    del S
    assert w.min() > 1e-8 * w.max()
    for k in (0, 7):
        x1 = lat.spinor(Q.F64)
        s1 = lat.solve(x1, bs[k], **SOLVE)
        assert np.array_equal(x1.download(), sols[k]), k
        assert s1["inner_iters"] == st[k]["inner_iters"]


# ---- the uneven batch and the served mix, at their full volumes
SERVE_VOLUMES = ([8, 8, 8, 8], [16, 16, 16, 32], [24, 24, 24, 48], [32, 32, 32, 64])
SERVE_FIELDS = [(L, [4, 4, 4, 4], 0.3, 8 + i) for i, L in enumerate(SERVE_VOLUMES)]


@pytest.fixture(scope="module")
def server():
    print("[configs] server with the four resident fields", flush=True)
    s = Q.Server(SERVE_FIELDS, links32=12, slots=4, **SOLVE)
    yield s
    s.close()


def check_request_with_oracle(req, res, nrhs_checked):
    """Re-solve a served request on a lattice of its own (same links, clover for its
    kappa, sources seed*1000+k): the XOR of the solutions' digests must equal the
    request's digest (the served solutions are these, bit for bit), and the oracle's
    FP64 residual of the first `nrhs_checked` solutions must be <= 1e-10."""
    L, B, eps, fseed = SERVE_FIELDS[req["field"]]
    lat = Q.Lattice(L, B, -1, 12, kappa=req["kappa"])
    lat.generate_gauge(eps, fseed)
    lat.compute_clover(1.0)
    bs = [lat.gaussian(Q.F64, req["seed"] * 1000 + k) for k in range(req["nrhs"])]
    xs = [lat.spinor(Q.F64) for _ in bs]
    st = lat.solve_batch(xs, bs, **SOLVE)
    dig = 0
    for x in xs:
        dig ^= x.digest()
    assert dig == res["digest"], req
    assert sum(s["inner_iters"] for s in st) == res["inner_iters"]
    orc = O.Oracle(L, B, -1, req["kappa"], lat.gauge(), lat.clover_field())
    for k in range(min(nrhs_checked, req["nrhs"])):
        r = oracle_residual(orc, xs[k].download(), bs[k].download())
        assert r <= 1e-10, (req, k, r)
    del orc
    lat.close()


def test_c4_uneven_batch(server):
    """16^3x32, 24^3x48 and 32^3x64 with one Gaussian source each (seeds 4-6),
    kappa in [0.130, 0.137], submitted together: every solution meets 1e-10 (oracle)."""
    rng = np.random.default_rng(4)
    reqs = [dict(field=i, kappa=float(rng.uniform(0.130, 0.137)), nrhs=1, seed=4 + k, arrival=0.0)
            for k, i in enumerate((1, 2, 3))]
    res = server.run(reqs)
    for q, r in zip(reqs, res):
        print(f"[configs] uneven: field {q['field']} kappa {q['kappa']:.4f}", flush=True)
        assert r["true_residual_max"] <= 1e-10
        check_request_with_oracle(q, r, 1)


def test_c5_served_mix_at_full_volumes(server):
    """A Poisson trace over the four volumes, kappa in [0.125, 0.137], 1-12 RHS: each
    request's result is the same alone and in the mix (digests), and one request per
    volume is checked against the oracle."""
    rng = np.random.default_rng(7)
    fields = [0, 1, 2, 3, 1, 0, 2]
    nrhs = [12, 3, 2, 1, 6, 5, 1]
    t = np.cumsum(rng.exponential(1.0 / 20.0, len(fields)))
    tr = [dict(field=f, kappa=float(rng.uniform(0.125, 0.137)), nrhs=n, seed=300 + i, arrival=float(t[i]))
          for i, (f, n) in enumerate(zip(fields, nrhs))]
    mix = server.run(tr)
    for i, q in enumerate(tr):
        assert mix[i]["true_residual_max"] <= 1e-10, q
        alone = server.run([dict(q, arrival=0.0)])[0]
        assert alone["digest"] == mix[i]["digest"], q
        assert alone["inner_iters"] == mix[i]["inner_iters"], q
    iv = sorted((m["admit"], m["finish"]) for m in mix)
    assert any(iv[k + 1][0] < iv[k][1] for k in range(len(iv) - 1))  # some ran at the same time
    seen = set()
    for i, q in enumerate(tr):
        if q["field"] in seen:
            continue
        seen.add(q["field"])
        print(f"[configs] served: field {q['field']} nrhs {q['nrhs']} kappa {q['kappa']:.4f}", flush=True)
        check_request_with_oracle(q, mix[i], 2)


def test_c3_precision_variant_full_size():
    """48^3x96 (links exp(0.25 i H), seed 3, kappa = 0.137, c_sw = 1.1) with the FP16-stored
    SAP: near the critical kappa an FP16-SAP inner solve diverges, the solve switches to
    the FP32 SAP (sap_escalated) and still reaches an FP64 true residual of 1e-10; its
    error against an FP64-refined continuation (the FP32-SAP solve from the result, driven
    to 1e-13) stays at the level of that residual."""
    print("[configs] 48^3x96 fields", flush=True)
    lat = Q.Lattice([48, 48, 48, 96], [4, 4, 4, 4], -1, 12, kappa=0.137)
    try:
        lat.generate_gauge(0.25, 3)
        lat.compute_clover(1.1)
        b = lat.gaussian(Q.F64, 3)
        x = lat.spinor(Q.F64)
        print("[configs] 48^3x96 FP16-SAP solve", flush=True)
        st = lat.solve(x, b, nsap=4, nm=6, tol=1e-11, maxiter=5, tol_s=1e-3, maxiter_s=40, true_tol=1e-10,
                       max_restarts=60, sap_precision=Q.F16)
        print("[configs] %s" % st, flush=True)
        assert st["true_residual"] <= 1e-10 and st["sap_escalated"] == 1, st
        xr = lat.spinor(Q.F64, x.download())
        print("[configs] 48^3x96 FP64-refined continuation", flush=True)
        ref = lat.solve(xr, b, nsap=3, nm=6, tol=1e-14, maxiter=5, tol_s=1e-3, maxiter_s=40, true_tol=1e-13,
                        max_restarts=60)
        assert ref["true_residual"] <= 1e-13, ref
        err = rel(x.download(), xr.download())
        print("[configs] error vs the FP64-refined continuation %.3e" % err, flush=True)
        assert err <= 1e-9, err
    finally:
        lat.close()
