"""FP16-storage SAP (assign_mult_mSAP_s_h_, prec_ddd_s_h.h:63-231) and the solvers
built on it (bicgstab_precdd_s_h.cc:83-370, bicgstab_dd_mix2_hf.cc:110-378).

The library stores the SAP iterates, links and clover inverse in binary16 and
computes in FP32.  The oracle restates that scheme with the same storage points.
FP32 summation order differs between the two, which flips an occasional binary16
rounding, and each flip moves the neighbours' values by more than FP32 noise, so
the two results differ at the FP16 noise level itself rather than bit for bit.
The tests therefore ask: the GPU's M_SAP is within 2e-3 relative of the oracle's
and of the all-FP32 M_SAP with one sweep less (the FP16 SAP's definition), far from
the FP32 M_SAP with nsap sweeps, and its deviation from FP32 is the oracle's within
a factor 3 (so the rounding really happens, at the same points).  Solves carry no
FP16 error: they reach the FP64 true residual and the reference's golden solution
norms."""
import ctypes

import numpy as np
import pytest

import qws_amd as Q
from oracle import oracle as O
from oracle import ref_worker as RW
from tests import casekit as K
from tests.test_gpu_parity import Problem, rel

pytestmark = pytest.mark.gpu

TOL_H = 2e-3

GEOMS = [
    ("case0", [32, 6, 4, 3], [16, 6, 4, 3]),   # improper colouring (QWS single rank)
    ("mixed", [8, 4, 2, 4], [4, 2, 2, 4]),     # z, t single block: improper
    ("b4", [16, 8, 8, 8], [4, 4, 4, 4]),       # 4^4 blocks, proper colouring
    ("b512", [16, 8, 8, 8], [8, 4, 4, 4]),     # 512-site blocks (fused kernel)
]


@pytest.fixture(scope="module", params=GEOMS, ids=[g[0] for g in GEOMS])
def prob(request):
    _, L, B = request.param
    return Problem(L, B)


@pytest.mark.parametrize("nsap,nm", [(4, 2), (1, 3), (2, 1)])
def test_msap_h_matches_oracle(prob, nsap, nm):
    p, lat = prob, prob.lat
    b = p.b.astype(np.float32)
    out = lat.spinor(Q.F32)
    lat.msap_h(out, lat.spinor(Q.F32, b), nsap, nm)
    got = out.download()
    want = p.orc.msap_h(b, nsap, nm)
    full32 = p.orc.prec_s(b, max(nsap - 1, 0), nm)
    assert rel(got, want) < TOL_H
    assert rel(got, full32) < TOL_H
    ratio = rel(got, full32) / rel(want, full32)
    assert 1 / 3 < ratio < 3
    if nsap > 1:
        assert rel(got, p.orc.prec_s(b, nsap, nm)) > 5 * TOL_H


@pytest.mark.parametrize("L,B,su3", [([16, 8, 8, 8], [4, 4, 4, 4], True), ([16, 8, 8, 8], [4, 4, 4, 4], False),
                                     ([16, 8, 8, 8], [8, 4, 4, 4], True)])
def test_msap_h_face_buffer_is_bitwise_the_gather(L, B, su3):
    """The FP16 SAP's face updates through the face buffer equal the gathered ones
    bit for bit (the same FP32 hop arithmetic on the same FP16 fields)."""
    p = Problem(L, B, su3=su3)
    b = p.b.astype(np.float32)
    res = {}
    for fb in (True, False):
        p.lat.set_sap_face_buffer(fb)
        out = p.lat.spinor(Q.F32)
        p.lat.msap_h(out, p.lat.spinor(Q.F32, b), 3, 2)
        res[fb] = out.download()
    p.lat.set_sap_face_buffer(True)
    assert np.array_equal(res[True], res[False])
    assert rel(res[True], p.orc.msap_h(b, 3, 2)) < 2 * TOL_H


def test_msap_h_links12():
    """12-real links: the third row is rebuilt in FP32 from the FP16 rows 0 and 1,
    so the oracle (which rounds all 18 reals) agrees to a few FP16 roundings."""
    p = Problem([16, 8, 8, 8], [4, 4, 4, 4], su3=True)
    b = p.b.astype(np.float32)
    out = p.lat.spinor(Q.F32)
    p.lat.msap_h(out, p.lat.spinor(Q.F32, b), 4, 2)
    assert rel(out.download(), p.orc.msap_h(b, 4, 2)) < 2 * TOL_H


def test_msap_h_batch_scale_invariance(prob):
    """The FP16 copy is scaled to unit RMS component, so scaling the input by a
    power of two scales the output exactly."""
    lat = prob.lat
    b = prob.b.astype(np.float32)
    o1, o2 = lat.spinor(Q.F32), lat.spinor(Q.F32)
    lat.msap_h(o1, lat.spinor(Q.F32, b), 3, 2)
    lat.msap_h(o2, lat.spinor(Q.F32, b * np.float32(2.0 ** -20)), 3, 2)
    assert np.array_equal(o2.download(), o1.download() * np.float32(2.0 ** -20))


def test_inner_hp_solve_true_residual():
    """bicgstab_precdd_s_h_: x accumulates M_SAP p directly, so the FP32 true
    residual of D x = b follows the recursive one."""
    p = Problem([16, 8, 8, 8], [4, 4, 4, 4], kappa=0.12, seed=9)
    b = p.b.astype(np.float32)
    b /= np.sqrt(O.norm2(b))
    x = p.lat.spinor(Q.F32)
    it = p.lat.bicgstab_precdd_h(x, p.lat.spinor(Q.F32, b), 1e-5, 100, 4, 2)
    assert 0 < it < 100
    r = b - p.orc.dslash(x.download())
    assert np.sqrt(O.norm2(r)) < 3e-5


def test_hp_solve_matches_fp32_sap_solve():
    """The FP16-SAP mixed solve reaches the FP64 true residual, and its solution
    agrees with the FP32-SAP solve's (the solution of D x = b is unique)."""
    p = Problem([8, 8, 8, 16], [4, 4, 4, 4], kappa=0.12, seed=13)
    prm = dict(nsap=4, nm=2, tol=1e-12, maxiter=40, tol_s=1e-3, maxiter_s=40, true_tol=1e-10, max_restarts=3)
    bb = p.lat.spinor(Q.F64, p.b)
    xh, xf = p.lat.spinor(Q.F64), p.lat.spinor(Q.F64)
    sh = p.lat.solve(xh, bb, sap_precision=Q.F16, **prm)
    sf = p.lat.solve(xf, bb, sap_precision=Q.F32, **prm)
    assert sh["true_residual"] <= 1e-10 and sf["true_residual"] <= 1e-10
    r = p.b - p.orc.dslash(xh.download())
    assert np.sqrt(O.norm2(r) / O.norm2(p.b)) <= 1e-10
    assert rel(xh.download(), xf.download()) < 1e-8


def test_hp_solve_golden_4111():
    """data/CASE_4111_1111_HP: the half-precision mixed solve with the driver's
    settings (6 outer iterations of bicgstab_dd_mix2_hf_, 50 inner) reaches the
    reference's solution, |x|^2 = 1.39578596651524e+02 under check.sh's 1e-14."""
    case = "CASE_4111_1111"
    nloc, npe, L, B = K.case_geometry(case)
    _, U, C, b, x0 = O.case_inputs(nloc, npe)
    lat = Q.Lattice(L, B, -1, 18, kappa=0.05)
    lat.load_gauge(U)
    lat.load_clover(C)
    x = lat.spinor(Q.F64, x0)
    # qwsg's maxiter counts like bicgstab_dd_mix_ (iterations 1..maxiter-1)
    st = lat.solve(x, lat.spinor(Q.F64, b), nsap=4, nm=2, tol=-1, maxiter=7, tol_s=-1, maxiter_s=50,
                   sap_precision=Q.F16)
    assert st["outer_iters"] == 7
    xs = x.download()
    orc = O.Oracle(L, B, -1, 0.05, U, C)
    assert O.norm2(b - orc.dslash(xs)) < 1e-15
    assert K.check_value("xnorm^2", O.norm2(xs), K.expected("CASE_4111_1111_HP", "xnorm^2", "HP"))


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def test_dropin_hp_case0():
    """The half-precision entry points on QWS's host layout: assign_mult_mSAP_s_h_
    against the oracle, assign_mult_wd_s_ against ddd_s_, and bicgstab_dd_mix2_hf_
    with the driver's settings (main.cc:337) reaching data/CASE0's solution."""
    L = Q.lib()
    nloc = [32, 6, 4, 3]
    _, U, C, b, x0 = O.case_inputs(nloc)
    ci = lambda v: ctypes.byref(ctypes.c_int(v))
    cd = lambda v: ctypes.byref(ctypes.c_double(v))
    keep = [(ctypes.c_int * 4)(1, 1, 1, 1), (ctypes.c_int * 4)(1, 1, 1, -1), ctypes.c_int(0), ctypes.c_int(1),
            (ctypes.c_int * 4)(1, 2, 2, 2)]
    L.qws_init_(ci(32), ci(6), ci(4), ci(3), keep[0], keep[1], ctypes.byref(keep[2]), ctypes.byref(keep[3]), keep[4])
    vold = ctypes.c_int.in_dll(L, "vold").value
    vols = ctypes.c_int.in_dll(L, "vols").value

    def glob(name, ct, n):
        return np.ctypeslib.as_array((ct * n).from_address(ctypes.c_void_p.in_dll(L, name).value))

    glob("glud", ctypes.c_double, vold * 8 * 144)[:] = RW.links_to_aosoa(U, nloc, 8).ravel()
    glob("glus", ctypes.c_float, vols * 8 * 288)[:] = RW.links_to_aosoa(U.astype(np.float32), nloc, 16).ravel()
    glob("clvd", ctypes.c_double, vold * 2 * 576)[:] = RW.clover_to_aosoa(C, nloc, 8).ravel()
    glob("clvs", ctypes.c_float, vols * 2 * 1152)[:] = RW.clover_to_aosoa(C.astype(np.float32), nloc, 16).ravel()
    ctypes.c_double.in_dll(L, "kappa").value = 0.05
    ctypes.c_double.in_dll(L, "mkappa").value = -0.05
    ctypes.c_double.in_dll(L, "kappa2").value = 0.0025
    L.qws_mi355x_fields_modified_()
    L.qws_h_init_(ctypes.c_void_p.in_dll(L, "glus"), ctypes.c_void_p.in_dll(L, "clvs"))

    b32 = b.astype(np.float32)
    bs = RW.spinor_to_aosoa(b32, nloc, 16)
    out = np.zeros_like(bs)
    L.assign_mult_mSAP_s_h_(_ptr(out), _ptr(bs), ci(4), ci(2))
    orc = O.Oracle([32, 6, 4, 3], [16, 6, 4, 3], -1, 0.05, U, C)
    want = RW.spinor_to_aosoa(orc.msap_h(b32, 4, 2), nloc, 16)
    assert rel(out, want) < TOL_H
    w1, w2 = np.zeros_like(bs), np.zeros_like(bs)
    L.assign_mult_wd_s_(_ptr(w1), _ptr(bs))
    L.ddd_s_(_ptr(w2), _ptr(bs))
    assert np.array_equal(w1, w2)

    bd = RW.spinor_to_aosoa(b, nloc, 8)
    x = RW.spinor_to_aosoa(x0, nloc, 8)
    it = ctypes.c_int(-7)
    L.bicgstab_dd_mix2_hf_(_ptr(x), _ptr(bd), cd(-1.0), ctypes.byref(it), ci(6), cd(-1.0), ci(50), ci(4), ci(2))
    assert it.value == -7  # written only on convergence (bicgstab_dd_mix2_hf.cc:243,323)
    r = np.zeros_like(bd)
    L.ddd_d_(_ptr(r), _ptr(x))
    assert O.norm2(bd - r) < 1e-15
    assert K.check_value("xnorm^2", O.norm2(x), K.expected("CASE0", "xnorm^2", "DD"))
    # the status variant: the same solution bit for bit; a fixed iteration count (tol < 0)
    # is success; a tolerance two outer iterations cannot reach returns 1
    L.bicgstab_dd_mix2_hf_status_.restype = ctypes.c_int
    xs = RW.spinor_to_aosoa(x0, nloc, 8)
    tr = ctypes.c_double(0)
    rc = L.bicgstab_dd_mix2_hf_status_(_ptr(xs), _ptr(bd), cd(-1.0), ctypes.byref(it), ci(6), cd(-1.0), ci(50),
                                       ci(4), ci(2), ctypes.byref(tr))
    assert rc == 0 and np.array_equal(xs, x)
    assert 0 <= tr.value < 1e-9 and abs(tr.value - np.sqrt(O.norm2(bd - r) / O.norm2(bd))) < 1e-12
    xs = RW.spinor_to_aosoa(x0, nloc, 8)
    # (inner solves to 0.1 only: two outer iterations cannot reach 1e-14)
    rc = L.bicgstab_dd_mix2_hf_status_(_ptr(xs), _ptr(bd), cd(1e-14), ctypes.byref(it), ci(2), cd(0.1), ci(50),
                                       ci(4), ci(2), ctypes.byref(tr))
    assert rc == 1 and tr.value > 1e-10, (rc, it.value, tr.value)
    L.qws_finalize_()
