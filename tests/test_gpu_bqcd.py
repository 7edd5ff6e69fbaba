"""QWS's BQCD-layout loaders and storers (qws_bqcd.cc:66-647) through the C ABI.

BQCD-side arrays come from oracle/bqcd.py with the site numbering of the tests'
stand-in for BQCD's std_xyzt2i_/e_o_ (oracle/libbqcd_sites.so, loaded into the
global symbol scope, where the library finds them at call time).

* DD loaders (qws_bqcd.cc:125-380,500-642): the host arrays they fill equal the
  reference's AoSoA layout of the same fields bit for bit (the layout the
  reference-library parity tests use), and the storers give back what was loaded.
* Even/odd loaders (qws_bqcd.cc:66-498): the host arrays equal the even/odd
  reading (parity-p site at compressed x index xh has x = 2 xh + (p+y+z+t)%2,
  deo_in_d.cc:80-85) bit for bit, and mtilde_vm_ on the fields they load is the
  Schur complement of the ddd_d_ operator on the fields the DD loaders load from
  the same BQCD arrays: with y_e = x and y_o = -(D (x, 0))_o, (D y)_o = 0 and
  (D y)_e = mtilde x (qws.cc:551-558), to 1e-12 relative.
"""
import ctypes
import os

import numpy as np
import pytest

import qws_amd as Q
from oracle import bqcd as BQ
from oracle import oracle as O
from oracle import ref_worker as RW
from tests.test_gpu_parity import rel

pytestmark = pytest.mark.gpu

SITES = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "libbqcd_sites.so")
NLOC = [32, 4, 4, 6]
KAPPA = 0.12


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _ci(v):
    return ctypes.byref(ctypes.c_int(v))


@pytest.fixture(scope="module")
def qws():
    sites = ctypes.CDLL(SITES, mode=ctypes.RTLD_GLOBAL)
    sites.bqcd_sites_set_((ctypes.c_int * 4)(*NLOC))
    L = Q.lib()
    keep = [(ctypes.c_int * 4)(1, 1, 1, 1), (ctypes.c_int * 4)(1, 1, 1, -1), ctypes.c_int(0), ctypes.c_int(1),
            (ctypes.c_int * 4)(1, 2, 2, 2)]
    L.qws_init_(_ci(NLOC[0]), _ci(NLOC[1]), _ci(NLOC[2]), _ci(NLOC[3]), keep[0], keep[1], ctypes.byref(keep[2]),
                ctypes.byref(keep[3]), keep[4])
    vold = ctypes.c_int.in_dll(L, "vold").value
    vols = ctypes.c_int.in_dll(L, "vols").value

    def glob(name, ct, n):
        return np.ctypeslib.as_array((ct * n).from_address(ctypes.c_void_p.in_dll(L, name).value))

    arrays = {"glud": glob("glud", ctypes.c_double, vold * 8 * 144),
              "glus": glob("glus", ctypes.c_float, vols * 8 * 288),
              "clvd": glob("clvd", ctypes.c_double, vold * 2 * 576),
              "clvs": glob("clvs", ctypes.c_float, vols * 2 * 1152)}
    _, U, C, b, _ = O.case_inputs(NLOC)
    yield L, arrays, U, C, b, vold, vols, keep, sites
    L.qws_finalize_()


def test_dd_loaders_fill_the_reference_layout(qws):
    L, A, U, C, b, vold, vols, keep, _ = qws
    V = int(np.prod(NLOC))
    L.qws_loadg_dd_(_ptr(BQ.links(U, NLOC)), _ci(V // 2), ctypes.byref(ctypes.c_double(KAPPA)))
    L.qws_loadc_dd_(_ptr(BQ.clover(C, NLOC)), _ci(V // 2))
    assert ctypes.c_double.in_dll(L, "kappa").value == KAPPA
    assert ctypes.c_double.in_dll(L, "mkappa").value == -KAPPA
    assert np.array_equal(A["glud"], RW.links_to_aosoa(U, NLOC, 8).ravel())
    assert np.array_equal(A["glus"], RW.links_to_aosoa(U.astype(np.float32), NLOC, 16).ravel())
    assert np.array_equal(A["clvd"], RW.clover_to_aosoa(C, NLOC, 8).ravel())
    assert np.array_equal(A["clvs"], RW.clover_to_aosoa(C.astype(np.float32), NLOC, 16).ravel())
    he, ho = BQ.spinor(b, NLOC)
    sd = np.zeros((2 * vold, 3, 4, 2, 8))
    L.qws_loadfd_dd_(_ptr(sd), _ptr(he), _ptr(ho))
    assert np.array_equal(sd.ravel(), RW.spinor_to_aosoa(b, NLOC, 8).ravel())
    ss = np.zeros((2 * vols, 3, 4, 2, 16), dtype=np.float32)
    L.qws_loadfs_dd_(_ptr(ss), _ptr(he), _ptr(ho))
    assert np.array_equal(ss.ravel(), RW.spinor_to_aosoa(b.astype(np.float32), NLOC, 16).ravel())
    oe, oo = np.zeros_like(he), np.zeros_like(ho)
    L.qws_storefd_dd_(_ptr(oe), _ptr(oo), _ptr(sd))
    assert np.array_equal(oe, he) and np.array_equal(oo, ho)
    L.qws_storefs_dd_(_ptr(oe), _ptr(oo), _ptr(ss))
    assert np.array_equal(oe, he.astype(np.float32)) and np.array_equal(oo, ho.astype(np.float32))


def _eo_sites():
    """(2, V/2) lexicographic index of the parity-p site at compressed index h."""
    nx, ny, nz, nt = NLOC
    t, z, y, xh = np.indices((nt, nz, ny, nx // 2)).reshape(4, -1)
    return np.stack([2 * xh + (p + y + z + t) % 2 + nx * (y + ny * (z + nz * t)) for p in (0, 1)])


def test_eo_loaders_fill_the_parity_layout_and_give_the_schur_complement(qws):
    L, A, U, C, b, vold, vols, keep, _ = qws
    V = int(np.prod(NLOC))
    rng = np.random.default_rng(5)
    x = rng.standard_normal((V // 2, 24))
    site = _eo_sites()
    full = np.zeros((V, 24))
    full[site[0]] = x
    he, ho = BQ.spinor(full.reshape(V, 3, 4, 2), NLOC)
    assert not ho.any()
    # the full operator on the DD-loaded fields
    L.qws_loadg_dd_(_ptr(BQ.links(U, NLOC)), _ci(V // 2), ctypes.byref(ctypes.c_double(KAPPA)))
    L.qws_loadc_dd_(_ptr(BQ.clover(C, NLOC)), _ci(V // 2))
    yd = np.zeros((2 * vold, 3, 4, 2, 8))
    w = np.zeros_like(yd)
    we, wo = np.zeros_like(he), np.zeros_like(ho)
    L.qws_loadfd_dd_(_ptr(yd), _ptr(he), _ptr(ho))
    L.ddd_d_(_ptr(w), _ptr(yd))
    L.qws_storefd_dd_(_ptr(we), _ptr(wo), _ptr(w))
    yo = -wo
    L.qws_loadfd_dd_(_ptr(yd), _ptr(he), _ptr(yo))
    L.ddd_d_(_ptr(w), _ptr(yd))
    L.qws_storefd_dd_(_ptr(we), _ptr(wo), _ptr(w))
    assert np.abs(wo).max() < 1e-13 * np.abs(we).max()
    # mtilde on the even/odd-loaded fields
    L.qws_loadg_bqcd_(_ptr(BQ.links(U, NLOC)), _ci(V // 2), ctypes.byref(ctypes.c_double(KAPPA)))
    L.qws_loadc_bqcd_(_ptr(BQ.clover(C, NLOC)), _ci(V // 2))
    eo_links = RW.eo_links_from_raw(A["glud"], NLOC, 8)
    eo_clover = RW.eo_clover_from_raw(A["clvd"], NLOC, 8)
    for p in (0, 1):
        assert np.array_equal(eo_links[p], U.reshape(4, V, 18)[:, site[p]]), p
        assert np.array_equal(eo_clover[p], C.reshape(V, 72)[site[p]]), p
    assert np.array_equal(RW.eo_links_from_raw(A["glus"], NLOC, 16),
                          eo_links.astype(np.float32))
    xin = np.zeros((vold, 3, 4, 2, 8))
    L.qws_loadfd_bqcd_(_ptr(xin), _ptr(he))
    assert np.array_equal(RW.eo_spinor_from_raw(xin, NLOC, 8), x)
    xs = np.zeros((vols, 3, 4, 2, 16), dtype=np.float32)
    L.qws_loadfs_bqcd_(_ptr(xs), _ptr(he))
    assert np.array_equal(RW.eo_spinor_from_raw(xs, NLOC, 16), x.astype(np.float32))
    m = np.zeros_like(xin)
    L.mtilde_vm_(_ptr(m), _ptr(xin))
    me = np.zeros_like(he)
    L.qws_storefd_bqcd_(_ptr(me), _ptr(m))
    assert rel(me, we) < 1e-12
    back = np.zeros_like(he)
    L.qws_storefs_bqcd_(_ptr(back), _ptr(xs))
    assert np.array_equal(back, he.astype(np.float32))


def test_bqcd_qws_h_init_is_qws_h_init_of_the_fp32_fields(qws):
    """bqcd_qws_h_init_() (qws_bqcd.cc:647) = qws_h_init_(glus, clvs): the FP16 SAP
    gives the same bits after either."""
    L, A, U, C, b, vold, vols, keep, _ = qws
    V = int(np.prod(NLOC))
    L.qws_loadg_dd_(_ptr(BQ.links(U, NLOC)), _ci(V // 2), ctypes.byref(ctypes.c_double(KAPPA)))
    L.qws_loadc_dd_(_ptr(BQ.clover(C, NLOC)), _ci(V // 2))
    bs = np.ascontiguousarray(RW.spinor_to_aosoa(b.astype(np.float32), NLOC, 16))
    outs = []
    for init in (lambda: L.bqcd_qws_h_init_(),
                 lambda: L.qws_h_init_(ctypes.c_void_p.in_dll(L, "glus"), ctypes.c_void_p.in_dll(L, "clvs"))):
        init()
        out = np.zeros_like(bs)
        L.assign_mult_mSAP_s_h_(_ptr(out), _ptr(bs), _ci(3), _ci(2))
        outs.append(out)
    assert np.isfinite(outs[0]).all() and np.abs(outs[0]).max() > 0
    assert np.array_equal(outs[0], outs[1])
