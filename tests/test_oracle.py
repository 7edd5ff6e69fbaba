"""CPU oracle checks (no GPU).

The oracle is pinned three ways:
  * golden norms of the reference's own test driver (data/CASE*, check.sh tolerances);
  * full-vector agreement with the reference library compiled into oracle/_ref
    (skipped when it is not built);
  * closed forms: free-field plane waves, Clifford algebra of the gamma matrices
    implied by the projector tables, dense clover construction, Neumann series
    of the Jacobi block solve, textbook multiplicative Schwarz, and
    prec_ddd_s(in) == D prec_s(in).
"""
import numpy as np
import pytest

from oracle import bqcd as BQ
from oracle import oracle as O
from oracle import ref_worker as RW
from tests import casekit as K

OPERATOR_CASES = ["CASE0", "CASE_4111_1111", "CASE1", "CASE2", "CASE3", "CASE4", "CASE5", "CASE6",
                  "CASE7", "CASE8", "CASE9"]


def _probe_norms(case):
    """The driver's DD-operator probes (main.cc:584-795) on the CASE inputs."""
    nloc, npe, L, B = K.case_geometry(case)
    _, U, C, b, _ = O.case_inputs(nloc, npe)
    orc = O.Oracle(L, B, -1, 0.05, U, C)
    dom0 = K.xdomain0_mask(L, nloc)
    bs = b.astype(np.float32)
    got = {"input": O.norm2(b), "input_s_": O.norm2(bs)}
    cl = orc.clover(b)
    cl[~dom0] = 0
    got["clvd_vm_"] = got["clvd_vm2_"] = O.norm2(cl)
    cls = cl.astype(np.float32)  # clv_s_ = same map in FP32
    got["clv_s_"] = O.norm2(cls)
    face = np.zeros_like(b)
    orc.face(face, b, -0.05, -1)
    face[~dom0] = 0
    got["ddd_out_pre|pos_d_"] = O.norm2(face)
    din = orc.dslash(b, O.HOPS_INSIDE)
    din[~dom0] = 0
    got["ddd_in_d_"] = O.norm2(din)
    got["ddd_d_"] = O.norm2(orc.dslash(b))
    fs = np.zeros_like(bs)
    orc.face_qws_uniform_s(fs, bs, nloc, npe, 0, -0.05)
    got["ddd_out_pre|pos_s_"] = O.norm2(fs)
    dins = orc.dslash(bs, O.HOPS_INSIDE)
    dins[~dom0] = 0
    got["ddd_in_s_"] = O.norm2(dins)
    got["ddd_s_"] = O.norm2(orc.dslash(bs))
    return got


@pytest.mark.parametrize("case", OPERATOR_CASES)
def test_case_operator_norms(case):
    got = _probe_norms(case)
    for name, value in got.items():
        want = K.expected(case, name)
        assert K.check_value(name, value, want), (case, name, value, want)


@pytest.mark.parametrize("case", ["CASE0", "CASE_4111_1111", "CASE1"])
def test_case_dd_mixed_solve(case):
    """bicgstab_dd_mix_ with nsap=4, nm=2, 5 outer x 2 x 50 inner (main.cc:262-312)."""
    nloc, npe, L, B = K.case_geometry(case)
    _, U, C, b, x0 = O.case_inputs(nloc, npe)
    orc = O.Oracle(L, B, -1, 0.05, U, C)
    x, it, inner = orc.bicgstab_dd_mix(x0, b, -1.0, 6, -1.0, 50, 4, 2)
    assert it == 6 and inner == 500
    r = b - orc.dslash(x)
    assert K.check_value("bnorm^2", O.norm2(b), K.expected(case, "bnorm^2", "DD"))
    assert O.norm2(r) < 1e-15
    assert K.check_value("xnorm^2", O.norm2(x), K.expected(case, "xnorm^2", "DD"))


# ---------------------------------------------------------------- reference library
needs_ref = pytest.mark.skipif(not RW.available(), reason="oracle/_ref/libqws_ref.so not built")


def _rel(a, b):
    return np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel())


@needs_ref
@pytest.mark.parametrize("nloc,kappa", [([32, 6, 4, 3], 0.05), ([32, 4, 2, 4], 0.11)])
def test_against_reference_library(nloc, kappa):
    """Full vectors of every hot-path function vs the reference compiled from its sources."""
    L, U, C, b, x0 = O.case_inputs(nloc)
    if kappa != 0.05:  # a second field set: random but with the same structure
        rng = np.random.default_rng(7)
        U, C = K.random_su3_like(b.shape[0], rng)
    B = [nloc[0] // 2] + nloc[1:]
    orc = O.Oracle(L, B, -1, kappa, U, C)
    ops = ["ddd_d", "ddd_s", "ddd_in_s0", "ddd_in_s1", "ddd_out_s0", "ddd_out_s1", "jinv0", "jinv1",
           "prec_ddd_s", "prec_s", "bicgstab_precdd_s", "bicgstab_dd_mix"]
    ref = RW.call(nloc, kappa, U, C, b, ops, x0=x0, maxiter_s=4, maxiter=3)
    bs = b.astype(np.float32)
    assert _rel(orc.dslash(b), ref["ddd_d"]) < 1e-14
    assert _rel(orc.dslash(bs), ref["ddd_s"]) < 1e-6
    for d in (0, 1):
        o = np.zeros_like(bs)
        orc.dslash(bs, O.HOPS_INSIDE, d, out=o)
        assert _rel(o, ref["ddd_in_s%d" % d]) < 1e-6
        o = np.zeros_like(bs)
        orc.face(o, bs, -kappa, d)
        assert _rel(o, ref["ddd_out_s%d" % d]) < 1e-6
        o = np.zeros_like(bs)
        orc.jinv_s(bs, d, 2, x=o)
        assert _rel(o, ref["jinv%d" % d]) < 1e-6
    assert _rel(orc.prec_ddd_s(bs, 4, 2), ref["prec_ddd_s"]) < 5e-6
    assert _rel(orc.prec_s(bs, 4, 2), ref["prec_s"]) < 5e-6
    x, _ = orc.bicgstab_precdd_s(bs, -1, 4, 4, 2)
    assert _rel(x, ref["bicgstab_precdd_s"]) < 1e-4
    x, it, _ = orc.bicgstab_dd_mix(x0, b, -1, 3, -1, 4, 4, 2)
    assert int(ref["bicgstab_dd_mix_iter"]) == it
    assert _rel(x, ref["bicgstab_dd_mix"]) < 1e-4


# ---------------------------------------------------------------- closed forms
def gammas():
    """Euclidean gamma matrices implied by the projector tables: gamma_4 =
    diag(1,1,-1,-1), gamma_k = [[0, i sigma_k], [-i sigma_k, 0]] (wilson_d.h:540-657)."""
    s = [np.array([[0, 1], [1, 0]], complex), np.array([[0, -1j], [1j, 0]]), np.array([[1, 0], [0, -1]], complex)]
    z = np.zeros((2, 2), complex)
    g = [np.block([[z, 1j * sk], [-1j * sk, z]]) for sk in s]
    g.append(np.diag([1, 1, -1, -1]).astype(complex))
    return g


def test_gamma_clifford():
    g = gammas()
    for m in range(4):
        for n in range(4):
            ac = g[m] @ g[n] + g[n] @ g[m]
            assert np.allclose(ac, 2 * np.eye(4) * (m == n))


def _to_c(psi):
    return psi[..., 0] + 1j * psi[..., 1]  # (V, 3, 4) complex


def _from_c(z):
    return np.stack([z.real, z.imag], axis=-1)


def test_free_field_plane_wave():
    """U = 1, T^-1 = 1: D e^{ipx} chi = D(p) e^{ipx} chi with
    D(p) = 1 - kappa sum_mu [(1-g_mu) e^{ip_mu} + (1+g_mu) e^{-ip_mu}], antiperiodic t."""
    L = [4, 6, 2, 8]
    V = int(np.prod(L))
    U = np.zeros((4, V, 3, 3, 2))
    for c in range(3):
        U[:, :, c, c, 0] = 1.0
    C = np.zeros((V, 2, 36))
    C[:, :, :6] = 0.5
    kappa = 0.13
    orc = O.Oracle(L, [2, 3, 2, 4], -1, kappa, U, C)
    g = gammas()
    rng = np.random.default_rng(1)
    n = [1, 2, 1, 3]
    p = [2 * np.pi * n[0] / L[0], 2 * np.pi * n[1] / L[1], 2 * np.pi * n[2] / L[2], (2 * n[3] + 1) * np.pi / L[3]]
    chi = rng.standard_normal((3, 4)) + 1j * rng.standard_normal((3, 4))
    idx = np.arange(V)
    x = [idx % L[0], (idx // L[0]) % L[1], (idx // (L[0] * L[1])) % L[2], idx // (L[0] * L[1] * L[2])]
    phase = np.exp(1j * sum(p[m] * x[m] for m in range(4)))
    psi = phase[:, None, None] * chi[None]
    Dp = np.eye(4, dtype=complex)
    for m in range(4):
        Dp -= kappa * ((np.eye(4) - g[m]) * np.exp(1j * p[m]) + (np.eye(4) + g[m]) * np.exp(-1j * p[m]))
    want = phase[:, None, None] * (chi @ Dp.T)[None]
    got = _to_c(orc.dslash(_from_c(psi)))
    assert np.abs(got - want).max() < 1e-12


def test_clover_dense_and_identity():
    """Packed clover multiply == dense 12x12 built from the packing rules of the
    chirality blocks (rotation to (s0+-s2, s1+-s3)); A = 1/2 gives the identity."""
    V = 5
    rng = np.random.default_rng(2)
    C = rng.standard_normal((V, 2, 36))
    U = np.zeros((4, V, 3, 3, 2))
    orc = O.Oracle([5, 1, 1, 1], [5, 1, 1, 1], 1, 0.1, U, C)
    psi = rng.standard_normal((V, 3, 4, 2))
    got = _to_c(orc.clover(psi))
    pairs = [(r, c) for r in range(6) for c in range(r + 1, 6)]
    for s in range(V):
        M = np.zeros((12, 12), complex)  # basis index = colour*4 + spin
        for i, sg in enumerate((1, -1)):
            a = C[s, i]
            A = np.diag(a[:6]).astype(complex)
            for k, (r, c) in enumerate(pairs):
                A[r, c] = a[6 + 2 * k] + 1j * a[7 + 2 * k]
                A[c, r] = np.conj(A[r, c])
            # x_i = R_i psi, out += R_i^T A x
            R = np.zeros((6, 12))
            for c in range(3):
                R[c, c * 4 + 0] = 1
                R[c, c * 4 + 2] = sg
                R[3 + c, c * 4 + 1] = 1
                R[3 + c, c * 4 + 3] = sg
            M += R.T @ A @ R
        want = M @ _to_c(psi[s]).reshape(12)
        assert np.abs(got[s].reshape(12) - want).max() < 1e-12
    C[:] = 0
    C[:, :, :6] = 0.5
    orc2 = O.Oracle([5, 1, 1, 1], [5, 1, 1, 1], 1, 0.1, U, C)
    assert np.abs(orc2.clover(psi) - psi).max() < 1e-15


def _small_problem(L, B, seed=3, kappa=0.12):
    V = int(np.prod(L))
    rng = np.random.default_rng(seed)
    U, C = K.random_su3_like(V, rng)
    psi = rng.standard_normal((V, 3, 4, 2))
    return O.Oracle(L, B, -1, kappa, U, C), psi


def test_inside_plus_face_is_full():
    orc, psi = _small_problem([8, 4, 4, 4], [4, 2, 2, 2])
    full = orc.dslash(psi)
    inside = orc.dslash(psi, O.HOPS_INSIDE)
    face = np.zeros_like(psi)
    orc.face(face, psi, -orc.kappa, -1)
    assert np.abs(inside + face - full).max() < 1e-13


def test_jinv_neumann_series():
    """x = sum_{k=0}^{nm} (1 - D_in)^k b on the blocks of one colour."""
    orc, psi = _small_problem([8, 4, 4, 4], [4, 2, 2, 2])
    psi = psi.astype(np.float32)
    col = 1
    mask = orc.dslash(np.zeros_like(psi), O.HOPS_INSIDE, col, out=np.full_like(psi, np.nan))
    sel = ~np.isnan(mask[:, 0, 0, 0])
    for nm in (1, 3):
        x = orc.jinv_s(psi, col, nm)
        term = psi.astype(np.float64).copy()
        term[~sel] = 0
        acc = term.copy()
        for _ in range(nm):
            din = orc.dslash(term, O.HOPS_INSIDE, col, out=np.zeros_like(term))
            term = term - din
            term[~sel] = 0
            acc += term
        assert np.abs(x[sel] - acc[sel]).max() < 1e-5


def _schwarz_textbook(orc, b, nsap, nm):
    """Multiplicative Schwarz x_c += A_c (b - D x)_c, nsap+1 sweeps over colours
    0,1, with A_c the Jacobi block inverse (proper colouring)."""
    x = np.zeros_like(b)
    for _ in range(nsap + 1):
        for c in (0, 1):
            r = b - orc.dslash(x)
            dx = orc.jinv_s(r.astype(np.float32), c, nm).astype(np.float64)
            sel = _colour_mask(orc, c)
            x[sel] += dx[sel]
    return x


def _colour_mask(orc, c):
    m = orc.dslash(np.zeros((orc.V, 3, 4, 2)), O.HOPS_INSIDE, c, out=np.full((orc.V, 3, 4, 2), np.nan))
    return ~np.isnan(m[:, 0, 0, 0])


def test_sap_is_multiplicative_schwarz_proper():
    orc, psi = _small_problem([8, 4, 4, 8], [2, 2, 2, 2], kappa=0.1)
    assert not orc.improper
    got = orc.prec_s(psi.astype(np.float32), 3, 2)
    want = _schwarz_textbook(orc, psi, 3, 2)
    assert _rel(got, want) < 2e-5


@pytest.mark.parametrize("L,B", [([8, 4, 4, 8], [2, 2, 2, 2]), ([32, 6, 4, 3], [16, 6, 4, 3]),
                                 ([8, 4, 2, 4], [4, 2, 2, 4])])
def test_prec_ddd_equals_d_prec(L, B):
    """prec_ddd_s_(in) == ddd_s_(prec_s_(in)) (test_qws.F90:168,181-182), both branches."""
    orc, psi = _small_problem(L, B, kappa=0.1)
    ps = psi.astype(np.float32)
    a = orc.prec_ddd_s(ps, 4, 2)
    bb = orc.dslash(orc.prec_s(ps, 4, 2))
    assert _rel(a, bb) < 2e-5


# ---------------------------------------------------------------- FP16 storage (prec_ddd_s_h.h)
def test_to_half_is_ieee_binary16():
    """The oracle's binary16 rounding equals numpy's IEEE float16 conversion
    (round to nearest even, subnormals, overflow to infinity)."""
    rng = np.random.default_rng(1)
    x = rng.standard_normal(50000) * 10.0 ** rng.uniform(-9, 5, 50000)
    edges = [65504, 65519.9, 65520, 1e-8, 2.0 ** -24, 3 * 2.0 ** -25, 2.0 ** -25, 0.0, -0.0, 1 + 2.0 ** -11,
             1 + 3 * 2.0 ** -11, 6.1e-5, -6.2e-5]
    x = np.concatenate([x, edges]).astype(np.float32)
    with np.errstate(over="ignore"):
        want = x.astype(np.float16).astype(np.float32)
    assert np.array_equal(O.to_half(x), want)


@pytest.mark.parametrize("L,B", [([8, 4, 4, 8], [2, 2, 2, 2]), ([8, 2, 2, 4], [4, 2, 2, 2])])
def test_msap_h_is_fp16_prec_s_with_one_sweep_less(L, B):
    """assign_mult_mSAP_s_h_(nsap) is prec_s_ with nsap-1 sweeps (prec_ddd_s_h.h:133 vs
    qws.cc:955): with FP16 storage it stays within FP16 rounding of the FP32
    prec_s_(nsap-1) and well away from prec_s_(nsap); the scaling to unit RMS makes
    it homogeneous (M(a b) = a M(b) up to rounding)."""
    V = int(np.prod(L))
    rng = np.random.default_rng(3)
    U, C = K.random_su3_like(V, rng)
    orc = O.Oracle(L, B, -1, 0.12, U, C)
    b = rng.standard_normal((V, 3, 4, 2)).astype(np.float32)
    rel = lambda a, c: float(np.linalg.norm(a - c) / np.linalg.norm(c))
    for nsap, nm in [(4, 2), (2, 1), (1, 3)]:
        h = orc.msap_h(b, nsap, nm)
        assert 1e-5 < rel(h, orc.prec_s(b, nsap - 1, nm)) < 3e-3
        assert rel(h, orc.prec_s(b, nsap, nm)) > 1e-2
    h1 = orc.msap_h(b, 3, 2)
    assert np.array_equal(orc.msap_h(b * np.float32(1024.0), 3, 2), h1 * np.float32(1024.0))


# ---------------------------------------------------------------- even/odd family (main.cc:584-790)
def _eo_case0(vlen, dtype):
    """data/CASE0 inputs read as the even/odd family reads them: the same host arrays
    (glud/glus, clvd/clvs, b/bs) taken as two parity halves of x-compressed sites."""
    nloc = [32, 6, 4, 3]
    _, U, C, b, _ = O.case_inputs(nloc)
    vol = int(np.prod(nloc)) // 2 // vlen
    Ue = RW.eo_links_from_raw(RW.links_to_aosoa(U.astype(dtype), nloc, vlen), nloc, vlen)
    Ce = RW.eo_clover_from_raw(RW.clover_to_aosoa(C.astype(dtype), nloc, vlen), nloc, vlen)
    raw = RW.spinor_to_aosoa(b.astype(dtype), nloc, vlen).reshape(2 * vol, -1)
    inp = RW.eo_spinor_from_raw(raw[:vol], nloc, vlen)
    orc = O.Oracle(nloc, [16, 6, 4, 3], -1, 0.05, U, C)
    return (lambda *a, **k: orc.deo(Ue, Ce, *a, **k)), inp


def test_even_odd_family_case0_fp64():
    """Every FP64 even/odd operator of the reference driver (main.cc:584-700) reproduces
    data/CASE0:4-24 (check.sh: 1e-14 relative).  Called as the driver calls them:
    out parity po = 1 from the parity pe = 0 input, fac = -kappa^2; mtilde with
    pce = 0, pco = 1 (qws.cc:551-576)."""
    D, inp = _eo_case0(8, np.float64)
    fac = -0.05 ** 2
    got = {
        "deo_in_vm_": D(1, inp, set=1),
        "dee_deo_in_vm_": D(1, inp, set=1, clv=True),
        "deo_dag_in_vm_": D(1, inp, set=1, dag=True),
        "dee_deo_dag_in_vm_": D(1, inp, set=1, dag=True, clv=True),
        "one_minus_dee_deo_in_vm_": D(1, inp, set=1, clv=True, mode=2, fac=fac, in0=inp),
        "one_minus_deo_dag_in_vm_": D(1, inp, set=1, dag=True, mode=2, fac=fac, in0=inp),
        "deo_out_pre|pos_vm_": D(1, inp, set=2, mode=1, fac=fac),
        "deo_dag_out_pre|pos_vm_": D(1, inp, set=2, dag=True, mode=1, fac=fac),
        "dee_deo_out_pre|pos_vm_": D(1, inp, set=2, clv=True, mode=1, fac=fac),
        "dee_deo_dag_out_pre|pos_vm_": D(1, inp, set=2, dag=True, clv=True, mode=1, fac=fac),
        "clvd_vm2_": D(0, inp, set=3, clv=True),
        "deo_vm_": D(1, inp),
        "dee_deo_vm_": D(1, inp, clv=True),
        "deo_dag_vm_": D(1, inp, dag=True),
        "dee_deo_dag_vm_": D(1, inp, dag=True, clv=True),
        "one_minus_dee_deo_vm_": D(1, inp, clv=True, mode=2, fac=fac, in0=inp),
        "one_minus_deo_dag_vm_": D(1, inp, dag=True, mode=2, fac=fac, in0=inp),
    }
    mt = D(0, D(1, inp, clv=True), clv=True, mode=2, fac=fac, in0=inp)
    got["mtilde_vm_"] = mt
    mtdag = lambda v: D(0, D(1, D(0, v, set=3, clv=True), dag=True, clv=True), dag=True, mode=2, fac=fac, in0=v)
    got["mtilde_dag_vm_"] = mtdag(inp)
    got["mtdagmt_vm_"] = mtdag(mt)
    for name, v in got.items():
        assert K.check_value(name, O.norm2(v), K.expected("CASE0", name)), name
    # in + out = all hops, and the dagger is the adjoint: <a, H b> = <H^+ a, b>
    assert np.allclose(D(1, inp, set=1) + D(1, inp, set=2), D(1, inp), rtol=0, atol=1e-12)
    rng = np.random.default_rng(4)
    a = rng.standard_normal(inp.shape)
    lhs = np.vdot(a.view(np.complex128), D(1, inp).view(np.complex128))
    rhs = np.vdot(D(0, a, dag=True).view(np.complex128), inp.view(np.complex128))
    assert abs(lhs - rhs) < 1e-9 * abs(lhs)


def test_even_odd_family_case0_fp32():
    """FP32 even/odd operators (main.cc:724-760) against data/CASE0:30-39 (3e-6)."""
    S, inp = _eo_case0(16, np.float32)
    fs = np.float32(-0.05 ** 2)
    got = {
        "deo_in_s_": S(1, inp, set=1),
        "dee_deo_in_s_": S(1, inp, set=1, clv=True),
        "one_minus_dee_deo_in_s_": S(1, inp, set=1, clv=True, mode=2, fac=fs, in0=inp),
        "deo_out_pre|pos_s_": S(1, inp, set=2, mode=1, fac=fs),
        "dee_deo_out_pre|pos_s_": S(1, inp, set=2, clv=True, mode=1, fac=fs),
        "clv_s_": S(0, inp, set=3, clv=True),
        "deo_s_": S(1, inp),
        "dee_deo_s_": S(1, inp, clv=True),
        "one_minus_dee_deo_s_": S(1, inp, clv=True, mode=2, fac=fs, in0=inp),
        "mtilde_s_": S(0, S(1, inp, clv=True), clv=True, mode=2, fac=fs, in0=inp),
    }
    for name, v in got.items():
        assert K.check_value(name, O.norm2(v), K.expected("CASE0", name)), name


# ---------------------------------------------------------------- even/odd Krylov solvers
def _spd(n, seed, cond=50.0):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return (q * np.geomspace(1.0, cond, n)) @ q.T, rng.standard_normal(n)


def test_eo_cg_restatement_solves_spd_systems():
    """cg reaches the linear-algebra solution from zero and from a guess; a plausible
    slip (a sign in the residual update, beta from the wrong norms) fails this."""
    from oracle import eo_solvers as ES
    A, b = _spd(48, 1)
    want = np.linalg.solve(A, b)
    for x0 in (np.zeros(48), np.random.default_rng(2).standard_normal(48)):
        x, it = ES.cg(lambda v: A @ v, b, x0, 200)
        assert 1 < it < 200
        assert np.linalg.norm(x - want) < 1e-11 * np.linalg.norm(want)
    x, it = ES.cg(lambda v: A @ v, b, np.zeros(48), 3)  # maxiter bounds the iterations
    assert it == 3


def test_eo_mcg_restatement_solves_every_shift():
    """mcg's x_k solve (A + sigma_k) x = b: the sign and placement of the shift in the
    xi recursion are fixed by comparing with the dense solutions."""
    from oracle import eo_solvers as ES
    A, b = _spd(48, 3)
    sigma = [0.0, 0.05, 0.7, 3.0]
    xs, it = ES.mcg(lambda v: A @ v, b, sigma, 300)
    assert 1 < it < 300
    for s, x in zip(sigma, xs):
        want = np.linalg.solve(A + s * np.eye(48), b)
        assert np.linalg.norm(x - want) < 1e-10 * np.linalg.norm(want), s
    x0, _ = ES.cg(lambda v: A @ v, b, np.zeros(48), 300)
    assert np.linalg.norm(xs[0] - x0) < 1e-11 * np.linalg.norm(x0)


@needs_ref
def test_eo_solvers_match_reference_library():
    """The reference's own cg_vm_ and mcg_vm_ (oracle/_ref) on data/CASE0's fields
    (kappa 0.05, x-domain 0 as the parity field) agree with the restatement run on the
    reference's mtdagmt_vm_.  The normal equations take ~1100 iterations here.  A
    shift whose scale factor xi underflows before the unshifted residual converges
    ends as NaN in the reference (mcg.cc:160 divides by it), and in the restatement."""
    nloc = [32, 6, 4, 3]
    _, U, C, b, _ = O.case_inputs(nloc)
    res = RW.call(nloc, 0.05, U, C, b, ["cg_vm", "mcg_vm"], maxiter=3000, sigma=[0.0, 0.01, 0.05, 0.5])
    rel = lambda a, w: np.linalg.norm(a - w) / np.linalg.norm(w)
    assert rel(res["cg_vm_oracle"], res["cg_vm"]) < 1e-12
    assert 500 < int(res["cg_vm_oracle_iter"]) < 3000
    for k in range(3):
        assert rel(res["mcg_vm_oracle"][k], res["mcg_vm"][k]) < 1e-12, k
    assert rel(res["mcg_vm"][0], res["cg_vm"]) < 1e-12
    assert np.isnan(res["mcg_vm"][3]).any() and np.isnan(res["mcg_vm_oracle"][3]).any()


def test_bqcd_layouts():
    """oracle/bqcd.py against the loaders' own statements.  The clover is packed from
    the dense Hermitian blocks in row pairs; QWS's loader takes BQCD's packing apart
    with an index table (qws_bqcd.cc:301-337): QWS[0,1] <- B[0,1], QWS[2,3] <-
    B[20,21], QWS[4,5] <- B[32,33], QWS[6..23] <- B[2..19], QWS[24..33] <- B[22..31],
    QWS[34,35] <- B[34,35]; both must give the same numbers.  Links follow
    u[ri + 2 col1 + 6 col2 + 18 (i-1) + 18 volh eo + 36 volh mu] -> c[col2][col1][ri]
    (qws_bqcd.cc:140-151), spinors h[ri + 2 s + 8 c + 24 (i-1)] (qws_bqcd.cc:512-528);
    the site numbering is a bijection onto 1..volh within each parity."""
    nloc = [4, 2, 2, 6]
    V = int(np.prod(nloc))
    rng = np.random.default_rng(3)
    i, eo = BQ.site_numbers(nloc)
    for p in (0, 1):
        assert sorted(i[eo == p]) == list(range(1, V // 2 + 1))
    C = rng.standard_normal((V, 2, 36))
    B = BQ.clover(C, nloc).reshape(2, V // 2, 2, 36)[eo, i - 1]
    perm = [0, 1, 20, 21, 32, 33] + list(range(2, 20)) + list(range(22, 32)) + [34, 35]
    assert np.array_equal(B[..., perm], C)
    U = rng.standard_normal((4, V, 3, 3, 2))
    u = BQ.links(U, nloc)
    for _ in range(50):
        mu, s, c1, c2, ri = (int(rng.integers(n)) for n in (4, V, 3, 3, 2))
        k = ri + 2 * c1 + 6 * c2 + 18 * (i[s] - 1) + 18 * (V // 2) * eo[s] + 36 * (V // 2) * mu
        assert u[k] == U[mu, s, c2, c1, ri]
    psi = rng.standard_normal((V, 3, 4, 2))
    he, ho = BQ.spinor(psi, nloc)
    for _ in range(50):
        s, c, sp, ri = (int(rng.integers(n)) for n in (V, 3, 4, 2))
        h = he if eo[s] == 0 else ho
        assert h[ri + 2 * sp + 8 * c + 24 * (i[s] - 1)] == psi[s, c, sp, ri]
    assert np.array_equal(BQ.spinor_from(he, ho, nloc), psi)


def test_clover_constant_abelian_field_closed_form():
    """The oracle's clover-from-links on a constant x-y field strength: T's spectrum is
    1 -+ kappa csw sin(phi q) by spin, known in closed form (tests/clover_field.py), so a
    wrong factor of F (e.g. (Q - Q^+)/4) or a flipped sign of F or sigma fails."""
    from oracle import fieldgen as FG
    from tests import clover_field as CF

    L = [4, 4, 2, 2]
    for n, kappa, csw in ((1, 0.13, 2.0), (3, 0.12, 1.5)):
        U, phi = CF.abelian_links(L, n)
        packed, off, T = FG.clover_inverse(L, U, kappa, csw)
        want = CF.expected_packed_clover(int(np.prod(L)), phi, kappa, csw)
        assert off < 1e-15
        assert np.abs(packed - want).max() < 1e-14
        # the factor and the sign matter: both alternatives are far off
        assert np.abs(CF.expected_packed_clover(int(np.prod(L)), phi, kappa, 2 * csw) - packed).max() > 1e-2
        assert np.abs(CF.expected_packed_clover(int(np.prod(L)), -phi, kappa, csw) - packed).max() > 1e-2
