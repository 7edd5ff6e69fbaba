"""Even/odd operator family on the GPU (deo_*, dee_deo_*, one_minus_*, clv, mtilde).

Parity: the QWS entry points reproduce data/CASE0's even/odd golden norms
(main.cc:584-760; check.sh: FP64 1e-14, FP32 3e-6 relative), and the device-resident
interface matches the CPU oracle's restatement on random fields for every hop set,
dagger, clover and mode (FP64 1e-13, FP32 1e-6 relative).  On even extents the even/odd
fields gathered from the lattice's own fields give mtilde = the Schur complement of
the full operator D, checked against D itself."""
import ctypes

import numpy as np
import pytest

import qws_amd as Q
from oracle import oracle as O
from oracle import ref_worker as RW
from tests import casekit as K
from tests.test_gpu_parity import rel

pytestmark = pytest.mark.gpu


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def test_dropin_even_odd_case0():
    L = Q.lib()
    nloc = [32, 6, 4, 3]
    _, U, C, b, _ = O.case_inputs(nloc)
    ci = lambda v: ctypes.byref(ctypes.c_int(v))
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
    for n in ("one_minus_dee_deo_in_vm_", "one_minus_deo_dag_in_vm_"):
        getattr(L, n).argtypes = [ctypes.c_void_p] * 5 + [ctypes.c_double]
    for n in ("deo_out_pos_vm_", "deo_dag_out_pos_vm_", "dee_deo_out_pos_vm_", "dee_deo_dag_out_pos_vm_"):
        getattr(L, n).argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_double]
    for n in ("one_minus_dee_deo_in_s_",):
        getattr(L, n).argtypes = [ctypes.c_void_p] * 5 + [ctypes.c_float]
    for n in ("deo_out_pos_s_", "dee_deo_out_pos_s_"):
        getattr(L, n).argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_float]

    pe, po = ctypes.c_int(0), ctypes.c_int(1)
    P = (ctypes.byref(po), ctypes.byref(pe))  # the driver's argument order: out parity po, in parity pe
    fac = -0.0025
    # FP64 (main.cc:584-700): in = the first vold vectors of b
    bd = RW.spinor_to_aosoa(b, nloc, 8).reshape(2 * vold, -1)
    inp = np.ascontiguousarray(bd[:vold])
    out = np.zeros_like(inp)

    def run(name, *args):
        out[:] = 0
        getattr(L, name)(*P, _ptr(out), _ptr(inp), *args)
        return O.norm2(out)

    got = {
        "deo_in_vm_": run("deo_in_vm_"),
        "dee_deo_in_vm_": run("dee_deo_in_vm_"),
        "deo_dag_in_vm_": run("deo_dag_in_vm_"),
        "dee_deo_dag_in_vm_": run("dee_deo_dag_in_vm_"),
        "one_minus_dee_deo_in_vm_": run("one_minus_dee_deo_in_vm_", _ptr(inp), fac),
        "one_minus_deo_dag_in_vm_": run("one_minus_deo_dag_in_vm_", _ptr(inp), fac),
        "deo_vm_": run("deo_vm_"),
        "dee_deo_vm_": run("dee_deo_vm_"),
        "one_minus_dee_deo_vm_": run("one_minus_dee_deo_vm_", _ptr(inp)),
        "deo_dag_vm_": run("deo_dag_vm_"),
        "dee_deo_dag_vm_": run("dee_deo_dag_vm_"),
        "one_minus_deo_dag_vm_": run("one_minus_deo_dag_vm_", _ptr(inp)),
    }
    for pre, pos, name in (("deo_out_pre_vm_", "deo_out_pos_vm_", "deo_out_pre|pos_vm_"),
                           ("deo_dag_out_pre_vm_", "deo_dag_out_pos_vm_", "deo_dag_out_pre|pos_vm_"),
                           ("deo_out_pre_vm_", "dee_deo_out_pos_vm_", "dee_deo_out_pre|pos_vm_"),
                           ("deo_dag_out_pre_vm_", "dee_deo_dag_out_pos_vm_", "dee_deo_dag_out_pre|pos_vm_")):
        out[:] = 0
        getattr(L, pre)(*P, _ptr(out), _ptr(inp))
        getattr(L, pos)(*P, _ptr(out), _ptr(inp), fac)
        got[name] = O.norm2(out)
    for name, fn in (("mtilde_vm_", L.mtilde_vm_), ("mtilde_dag_vm_", L.mtilde_dag_vm_),
                     ("mtdagmt_vm_", L.mtdagmt_vm_)):
        out[:] = 0
        fn(_ptr(out), _ptr(inp))
        got[name] = O.norm2(out)
    # FP32 (main.cc:724-760): in = the first vols vectors of bs
    bs = RW.spinor_to_aosoa(b.astype(np.float32), nloc, 16).reshape(2 * vols, -1)
    ins = np.ascontiguousarray(bs[:vols])
    outs = np.zeros_like(ins)

    def run_s(name, *args):
        outs[:] = 0
        getattr(L, name)(*P, _ptr(outs), _ptr(ins), *args)
        return O.norm2(outs)

    got.update({
        "deo_in_s_": run_s("deo_in_s_"),
        "dee_deo_in_s_": run_s("dee_deo_in_s_"),
        "one_minus_dee_deo_in_s_": run_s("one_minus_dee_deo_in_s_", _ptr(ins), np.float32(fac)),
        "deo_s_": run_s("deo_s_"),
        "dee_deo_s_": run_s("dee_deo_s_"),
        "one_minus_dee_deo_s_": run_s("one_minus_dee_deo_s_", _ptr(ins)),
    })
    for pos, name in (("deo_out_pos_s_", "deo_out_pre|pos_s_"), ("dee_deo_out_pos_s_", "dee_deo_out_pre|pos_s_")):
        outs[:] = 0
        L.deo_out_pre_s_(*P, _ptr(outs), _ptr(ins))
        getattr(L, pos)(*P, _ptr(outs), _ptr(ins), np.float32(fac))
        got[name] = O.norm2(outs)
    outs[:] = ins
    L.clv_s_(ctypes.byref(pe), _ptr(outs))
    got["clv_s_"] = O.norm2(outs)
    outs[:] = 0
    L.mtilde_s_(_ptr(outs), _ptr(ins))
    got["mtilde_s_"] = O.norm2(outs)
    L.qws_finalize_()
    for name, v in got.items():
        assert K.check_value(name, v, K.expected("CASE0", name)), (name, v)


GEOMS = [([8, 4, 4, 8], -1), ([32, 6, 4, 3], -1), ([16, 8, 2, 4], 1)]


@pytest.fixture(scope="module", params=GEOMS, ids=["even", "case0", "periodic"])
def eo(request):
    L, tbc = request.param
    V = int(np.prod(L))
    rng = np.random.default_rng(17)
    U, C = K.random_su3_like(V, rng)
    Ue = rng.standard_normal((2, 4, V // 2, 18))
    Ce = rng.standard_normal((2, V // 2, 72)) * 0.1
    Ce[:, :, :6] += 0.5
    Ce[:, :, 36:42] += 0.5
    lat = Q.Lattice(L, [L[0] // 2, L[1], L[2], L[3]], tbc, 18, kappa=0.11)
    lat.eo_load_gauge(Ue)
    lat.eo_load_clover(Ce)
    orc = O.Oracle(L, [L[0] // 2, L[1], L[2], L[3]], tbc, 0.11, U, C)
    return lat, orc, Ue, Ce, rng


@pytest.mark.parametrize("set_", [0, 1, 2, 3])
@pytest.mark.parametrize("dag", [False, True])
@pytest.mark.parametrize("clv", [False, True])
def test_deo_matches_oracle(eo, set_, dag, clv):
    lat, orc, Ue, Ce, rng = eo
    vh = lat.V // 2
    a = rng.standard_normal((vh, 24))
    a0 = rng.standard_normal((vh, 24))
    o0 = rng.standard_normal((vh, 24))
    for pe in (0, 1):
        for mode in (0, 1, 2):
            for prec, tol in ((Q.F64, 1e-13), (Q.F32, 1e-6)):
                dt = np.float64 if prec == Q.F64 else np.float32
                out = lat.eo_spinor(prec, o0)
                lat.deo(out, lat.eo_spinor(prec, a), pe, set_, dag, clv, mode, -0.37, lat.eo_spinor(prec, a0))
                want = orc.deo(Ue.astype(dt), Ce.astype(dt), pe, a.astype(dt), set=set_, dag=dag, clv=clv,
                               mode=mode, fac=dt(-0.37), out=o0.astype(dt).copy(), in0=a0.astype(dt))
                assert rel(out.download(), want) < tol, (pe, mode, prec)


def test_mtilde_is_schur_complement_of_d():
    """Even extents, fields gathered from the lattice's own: with y_e = x and
    y_o = kappa T_oo^-1 H_oe x, the full D gives (D y)_o = 0 and (D y)_e = mtilde x
    (qws.cc:551-558 is 1 - kappa^2 T_ee^-1 H_eo T_oo^-1 H_oe)."""
    L, B = [8, 4, 4, 8], [4, 2, 2, 4]
    V = int(np.prod(L))
    rng = np.random.default_rng(23)
    U, C = K.random_su3_like(V, rng)
    lat = Q.Lattice(L, B, -1, 18, kappa=0.12)
    lat.load_gauge(U)
    lat.load_clover(C)
    lat.eo_from_lattice()
    orc = O.Oracle(L, B, -1, 0.12, U, C)
    x = rng.standard_normal((V // 2, 24))
    # parity-p site h <-> lattice site (2 xh + (p + y + z + t) % 2, y, z, t)
    t, z, y, xh = np.indices((L[3], L[2], L[1], L[0] // 2)).reshape(4, -1)
    site = lambda p: 2 * xh + (p + y + z + t) % 2 + L[0] * (y + L[1] * (z + L[2] * t))
    yo = lat.eo_spinor(Q.F64)
    lat.deo(yo, lat.eo_spinor(Q.F64, x), 1, clv=True, mode=0)
    full = np.zeros((V, 24))
    full[site(0)] = x
    full[site(1)] = 0.12 * yo.download()
    d = orc.dslash(full.reshape(V, 3, 4, 2)).reshape(V, 24)
    mt = lat.eo_spinor(Q.F64)
    lat.mtilde(mt, lat.eo_spinor(Q.F64, x), 0, 0)
    assert np.abs(d[site(1)]).max() < 1e-13 * np.abs(d).max()
    assert rel(mt.download(), d[site(0)]) < 1e-13
    # mtilde^+ is the adjoint of mtilde, and which=2 is their product
    w = rng.standard_normal((V // 2, 24))
    md = lat.eo_spinor(Q.F64)
    lat.mtilde(md, lat.eo_spinor(Q.F64, w), 0, 1)
    cx = lambda a: a.view(np.complex128)
    lhs = np.vdot(cx(w), cx(mt.download()))
    rhs = np.vdot(cx(md.download()), cx(x))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    mm = lat.eo_spinor(Q.F64)
    lat.mtilde(mm, lat.eo_spinor(Q.F64, x), 0, 2)
    lat.mtilde(md, mt, 0, 1)
    assert rel(mm.download(), md.download()) < 1e-14


def test_dropin_bicgstab_vm_case0():
    """bicgstab_vm_ on mtilde_vm_ with the driver's settings (main.cc:200-245: x from
    zero, maxiter 100) reaches data/CASE0's solution (xnorm^2 under check.sh's 1e-14)."""
    L = Q.lib()
    nloc = [32, 6, 4, 3]
    _, U, C, b, _ = O.case_inputs(nloc)
    ci = lambda v: ctypes.byref(ctypes.c_int(v))
    keep = [(ctypes.c_int * 4)(1, 1, 1, 1), (ctypes.c_int * 4)(1, 1, 1, -1), ctypes.c_int(0), ctypes.c_int(1),
            (ctypes.c_int * 4)(1, 2, 2, 2)]
    L.qws_init_(ci(32), ci(6), ci(4), ci(3), keep[0], keep[1], ctypes.byref(keep[2]), ctypes.byref(keep[3]), keep[4])
    vold = ctypes.c_int.in_dll(L, "vold").value
    glud = np.ctypeslib.as_array((ctypes.c_double * (vold * 8 * 144)).from_address(
        ctypes.c_void_p.in_dll(L, "glud").value))
    clvd = np.ctypeslib.as_array((ctypes.c_double * (vold * 2 * 576)).from_address(
        ctypes.c_void_p.in_dll(L, "clvd").value))
    glud[:] = RW.links_to_aosoa(U, nloc, 8).ravel()
    clvd[:] = RW.clover_to_aosoa(C, nloc, 8).ravel()
    ctypes.c_double.in_dll(L, "kappa").value = 0.05
    ctypes.c_double.in_dll(L, "kappa2").value = 0.0025
    L.qws_mi355x_fields_modified_()
    bd = np.ascontiguousarray(RW.spinor_to_aosoa(b, nloc, 8).reshape(2 * vold, -1)[:vold])
    x = np.zeros_like(bd)
    it = ctypes.c_int(0)
    L.bicgstab_vm_(_ptr(x), _ptr(bd), ctypes.byref(it), ci(100))
    r = np.zeros_like(bd)
    L.mtilde_vm_(_ptr(r), _ptr(x))
    L.qws_finalize_()
    assert 1 < it.value < 100
    assert K.check_value("bnorm^2", O.norm2(bd), K.expected("CASE0", "bnorm^2", "EO"))
    assert O.norm2(bd - r) < 1e-22
    assert K.check_value("xnorm^2", O.norm2(x), K.expected("CASE0", "xnorm^2", "EO"))


def _case0_dropin_setup(L):
    nloc = [32, 6, 4, 3]
    _, U, C, b, _ = O.case_inputs(nloc)
    ci = lambda v: ctypes.byref(ctypes.c_int(v))
    keep = [(ctypes.c_int * 4)(1, 1, 1, 1), (ctypes.c_int * 4)(1, 1, 1, -1), ctypes.c_int(0), ctypes.c_int(1),
            (ctypes.c_int * 4)(1, 2, 2, 2)]
    L.qws_init_(ci(32), ci(6), ci(4), ci(3), keep[0], keep[1], ctypes.byref(keep[2]), ctypes.byref(keep[3]), keep[4])
    vold = ctypes.c_int.in_dll(L, "vold").value
    glud = np.ctypeslib.as_array((ctypes.c_double * (vold * 8 * 144)).from_address(
        ctypes.c_void_p.in_dll(L, "glud").value))
    clvd = np.ctypeslib.as_array((ctypes.c_double * (vold * 2 * 576)).from_address(
        ctypes.c_void_p.in_dll(L, "clvd").value))
    glud[:] = RW.links_to_aosoa(U, nloc, 8).ravel()
    clvd[:] = RW.clover_to_aosoa(C, nloc, 8).ravel()
    ctypes.c_double.in_dll(L, "kappa").value = 0.05
    ctypes.c_double.in_dll(L, "kappa2").value = 0.0025
    L.qws_mi355x_fields_modified_()
    bd = np.ascontiguousarray(RW.spinor_to_aosoa(b, nloc, 8).reshape(2 * vold, -1)[:vold])
    return nloc, U, C, b, bd, keep


@pytest.mark.skipif(not RW.available(), reason="oracle/_ref/libqws_ref.so not built")
def test_dropin_cg_mcg_vm_case0():
    """cg_vm_ and mcg_vm_ through the QWS ABI on data/CASE0's fields (kappa 0.05) agree
    with the reference's own cg_vm_/mcg_vm_ (oracle/_ref) and meet the residual
    (mtdagmt_vm_ + sigma) x = b; a shift whose scale factor underflows ends as NaN in
    both, as in the reference."""
    L = Q.lib()
    nloc, U, C, b, bd, keep = _case0_dropin_setup(L)
    sigma = [0.0, 0.01, 0.05, 0.5]
    ci = lambda v: ctypes.byref(ctypes.c_int(v))
    x = np.zeros_like(bd)
    it = ctypes.c_int(0)
    L.cg_vm_(_ptr(x), _ptr(bd), ctypes.byref(it), ci(3000))
    xx = np.zeros((len(sigma),) + bd.shape)
    sg = np.array(sigma)
    L.mcg_vm_(_ptr(xx), _ptr(bd), ci(len(sigma)), sg.ctypes.data_as(ctypes.c_void_p), ci(3000))
    res = []
    for v in [x] + [xx[k] for k in range(3)]:
        r = np.zeros_like(bd)
        L.mtdagmt_vm_(_ptr(r), _ptr(np.ascontiguousarray(v)))
        res.append(r)
    L.qws_finalize_()
    assert 500 < it.value < 3000
    assert O.norm2(bd - res[0]) < 1e-24 * O.norm2(bd)
    for k in range(3):
        assert O.norm2(bd - res[k + 1] - sigma[k] * xx[k]) < 1e-24 * O.norm2(bd), k
    assert np.isnan(xx[3]).any()
    # both stop at |r|/|b| < 1e-14 on an operator of condition ~1e4-1e5, after ~1100
    # iterations with differently ordered sums: the solutions agree to ~cond x 1e-14
    ref = RW.call(nloc, 0.05, U, C, b, ["cg_vm", "mcg_vm"], maxiter=3000, sigma=sigma)
    assert rel(x, ref["cg_vm"]) < 1e-8
    for k in range(3):
        assert rel(xx[k], ref["mcg_vm"][k]) < 1e-8, k


def test_eo_cg_mcg_device():
    """The device-resident CG and multi-shift CG on random fields (even extents, fields
    gathered from the lattice) meet their residuals, and the zero shift equals CG."""
    L, B = [8, 4, 4, 8], [4, 2, 2, 4]
    V = int(np.prod(L))
    rng = np.random.default_rng(31)
    U, C = K.random_su3_like(V, rng)
    lat = Q.Lattice(L, B, -1, 18, kappa=0.12)
    lat.load_gauge(U)
    lat.load_clover(C)
    lat.eo_from_lattice()
    bh = rng.standard_normal((V // 2, 24))
    b = lat.eo_spinor(Q.F64, bh)
    x = lat.eo_spinor(Q.F64, np.zeros_like(bh))
    it = lat.eo_cg(x, b, maxiter=5000, tol=1e-12)
    assert 1 < it < 5000
    sigma = [0.0, 0.1, 1.0]
    xs, it2 = lat.eo_mcg(b, sigma, maxiter=5000, tol=1e-12)
    assert abs(it2 - it) <= 1
    q = lat.eo_spinor(Q.F64)
    for s, xk in zip(sigma, xs):
        lat.mtilde(q, xk, 0, 2)
        r = bh - q.download() - s * xk.download()
        assert O.norm2(r) < 1e-22 * O.norm2(bh), s
    assert rel(xs[0].download(), x.download()) < 1e-10
