"""CPU tests of the library's host-side code: the BQCD loaders and storers
(qws_bqcd.cc:66-642), the FP64<->FP32 reorders (reordering.cc:59-204) and the driver
helpers (qws.cc:445-500).  They touch only host memory, so they run without a GPU:
the geometry globals are set directly and the field pointers point at numpy buffers
for the duration of a test (restored afterwards, so qws_init_ never frees them).

Expected values: the reference's AoSoA layouts as oracle/ref_worker.py states them (the
layouts that the reference-library parity tests use), BQCD's layouts from oracle/bqcd.py
with the tests' stand-in site numbering (oracle/libbqcd_sites.so)."""
import ctypes
import os

import numpy as np
import pytest

import qws_amd as Q
from oracle import bqcd as BQ
from oracle import ref_worker as RW

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SITES = os.path.join(ROOT, "oracle", "libbqcd_sites.so")
NLOC = [32, 2, 2, 4]


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _ci(v):
    return ctypes.byref(ctypes.c_int(v))


@pytest.fixture
def host():
    """The library with NLOC's geometry globals and numpy-backed field arrays."""
    L = Q.lib()
    nx, ny, nz, nt = NLOC
    g = {n: ctypes.c_int.in_dll(L, n) for n in ("nx", "ny", "nz", "nt", "nxh", "nxd", "nxs", "vold", "vols")}
    old_g = {n: v.value for n, v in g.items()}
    ptrs = {n: ctypes.c_void_p.in_dll(L, n) for n in ("glud", "glus", "clvd", "clvs")}
    old_p = {n: v.value for n, v in ptrs.items()}
    vals = dict(nx=nx, ny=ny, nz=nz, nt=nt, nxh=nx // 2, nxd=nx // 2 // 8, nxs=nx // 2 // 16)
    vals["vold"] = vals["nxd"] * ny * nz * nt
    vals["vols"] = vals["nxs"] * ny * nz * nt
    for n, v in vals.items():
        g[n].value = v
    arrays = {"glud": np.zeros(vals["vold"] * 8 * 144), "glus": np.zeros(vals["vols"] * 8 * 288, np.float32),
              "clvd": np.zeros(vals["vold"] * 2 * 576), "clvs": np.zeros(vals["vols"] * 2 * 1152, np.float32)}
    for n, a in arrays.items():
        ptrs[n].value = a.ctypes.data
    try:
        yield L, arrays, vals
    finally:
        for n, v in old_p.items():
            ptrs[n].value = v
        for n, v in old_g.items():
            g[n].value = v


@pytest.fixture(scope="module")
def sites():
    lib = ctypes.CDLL(SITES, mode=ctypes.RTLD_GLOBAL)
    lib.bqcd_sites_set_((ctypes.c_int * 4)(*NLOC))
    return lib


def _fields(seed):
    V = int(np.prod(NLOC))
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((4, V, 3, 3, 2)), rng.standard_normal((V, 2, 36)),
            rng.standard_normal((V, 3, 4, 2)))


def _eo_sites():
    nx, ny, nz, nt = NLOC
    t, z, y, xh = np.indices((nt, nz, ny, nx // 2)).reshape(4, -1)
    return np.stack([2 * xh + (p + y + z + t) % 2 + nx * (y + ny * (z + nz * t)) for p in (0, 1)])


def test_bqcd_dd_loaders_and_storers(host, sites):
    L, A, vals = host
    U, C, b = _fields(1)
    V = int(np.prod(NLOC))
    L.qws_loadg_dd_(_ptr(BQ.links(U, NLOC)), _ci(V // 2), ctypes.byref(ctypes.c_double(0.125)))
    L.qws_loadc_dd_(_ptr(BQ.clover(C, NLOC)), _ci(V // 2))
    assert ctypes.c_double.in_dll(L, "kappa").value == 0.125
    assert ctypes.c_double.in_dll(L, "kappa2").value == 0.125 ** 2
    assert np.array_equal(A["glud"], RW.links_to_aosoa(U, NLOC, 8).ravel())
    assert np.array_equal(A["glus"], RW.links_to_aosoa(U.astype(np.float32), NLOC, 16).ravel())
    assert np.array_equal(A["clvd"], RW.clover_to_aosoa(C, NLOC, 8).ravel())
    assert np.array_equal(A["clvs"], RW.clover_to_aosoa(C.astype(np.float32), NLOC, 16).ravel())
    he, ho = BQ.spinor(b, NLOC)
    sd = np.zeros(2 * vals["vold"] * 192)
    L.qws_loadfd_dd_(_ptr(sd), _ptr(he), _ptr(ho))
    assert np.array_equal(sd, RW.spinor_to_aosoa(b, NLOC, 8).ravel())
    ss = np.zeros(2 * vals["vols"] * 384, np.float32)
    L.qws_loadfs_dd_(_ptr(ss), _ptr(he), _ptr(ho))
    assert np.array_equal(ss, RW.spinor_to_aosoa(b.astype(np.float32), NLOC, 16).ravel())
    oe, oo = np.zeros_like(he), np.zeros_like(ho)
    L.qws_storefd_dd_(_ptr(oe), _ptr(oo), _ptr(sd))
    assert np.array_equal(oe, he) and np.array_equal(oo, ho)
    L.qws_storefs_dd_(_ptr(oe), _ptr(oo), _ptr(ss))
    assert np.array_equal(BQ.spinor_from(oe, oo, NLOC), b.astype(np.float32))


def test_bqcd_even_odd_loaders_and_storers(host, sites):
    L, A, vals = host
    U, C, b = _fields(2)
    V = int(np.prod(NLOC))
    site = _eo_sites()
    L.qws_loadg_bqcd_(_ptr(BQ.links(U, NLOC)), _ci(V // 2), ctypes.byref(ctypes.c_double(0.1)))
    L.qws_loadc_bqcd_(_ptr(BQ.clover(C, NLOC)), _ci(V // 2))
    links = RW.eo_links_from_raw(A["glud"], NLOC, 8)
    clover = RW.eo_clover_from_raw(A["clvd"], NLOC, 8)
    for p in (0, 1):
        assert np.array_equal(links[p], U.reshape(4, V, 18)[:, site[p]]), p
        assert np.array_equal(clover[p], C.reshape(V, 72)[site[p]]), p
    assert np.array_equal(RW.eo_links_from_raw(A["glus"], NLOC, 16), links.astype(np.float32))
    assert np.array_equal(RW.eo_clover_from_raw(A["clvs"], NLOC, 16), clover.astype(np.float32))
    he, _ = BQ.spinor(b, NLOC)
    xd = np.zeros(vals["vold"] * 192)
    L.qws_loadfd_bqcd_(_ptr(xd), _ptr(he))
    assert np.array_equal(RW.eo_spinor_from_raw(xd, NLOC, 8), b.reshape(V, 24)[site[0]])
    xs = np.zeros(vals["vols"] * 384, np.float32)
    L.qws_loadfs_bqcd_(_ptr(xs), _ptr(he))
    assert np.array_equal(RW.eo_spinor_from_raw(xs, NLOC, 16), b.reshape(V, 24)[site[0]].astype(np.float32))
    back = np.zeros_like(he)
    L.qws_storefd_bqcd_(_ptr(back), _ptr(xd))
    assert np.array_equal(back, he)
    L.qws_storefs_bqcd_(_ptr(back), _ptr(xs))
    assert np.array_equal(back, he.astype(np.float32))


def test_reorders(host):
    L, A, vals = host
    _, _, b = _fields(3)
    bd = RW.spinor_to_aosoa(b, NLOC, 8)
    s32 = np.zeros(2 * vals["vols"] * 384, np.float32)
    L.fermi_reorder_d2s_dd_(_ptr(s32), _ptr(bd))
    assert np.array_equal(s32, RW.spinor_to_aosoa(b.astype(np.float32), NLOC, 16).ravel())
    d64 = np.zeros(2 * vals["vold"] * 192)
    L.fermi_reorder_s2d_dd_(_ptr(d64), _ptr(s32))
    assert np.array_equal(d64, RW.spinor_to_aosoa(b.astype(np.float32).astype(np.float64), NLOC, 8).ravel())
    # renormalising forms: out = in / |in| (dnorm = |in|), out = in * dnorm
    L.fermi_reorder_d2s_dd_renorm_.argtypes = [ctypes.c_void_p] * 3
    L.fermi_reorder_s2d_dd_renorm_.argtypes = [ctypes.c_void_p] * 3
    dn = ctypes.c_double(0)
    L.fermi_reorder_d2s_dd_renorm_(_ptr(s32), _ptr(bd), ctypes.byref(dn))
    nb = np.sqrt(np.sum(b * b))
    assert abs(dn.value - nb) < 1e-14 * nb
    assert np.abs(RW.spinor_from_aosoa(s32.reshape(2 * vals["vols"], 3, 4, 2, 16), NLOC, 16) - b / nb).max() < 1e-7
    L.fermi_reorder_s2d_dd_renorm_(_ptr(d64), _ptr(s32), ctypes.byref(dn))
    assert np.array_equal(RW.spinor_from_aosoa(d64.reshape(2 * vals["vold"], 3, 4, 2, 8), NLOC, 8),
                          RW.spinor_from_aosoa(s32.reshape(2 * vals["vols"], 3, 4, 2, 16), NLOC, 16)
                          .astype(np.float64) * dn.value)


def test_driver_helpers(capfd):
    L = Q.lib()
    rng = np.random.default_rng(4)
    a = rng.standard_normal(3 * 192)
    f = rng.standard_normal(2 * 384).astype(np.float32)
    L.print_scdnorm2.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int]
    L.print_scsnorm2.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int]
    capfd.readouterr()
    L.print_scdnorm2(b"ddd_d_", _ptr(a), 3)
    L.print_scsnorm2(b"ddd_s_", _ptr(f), 2)
    out = capfd.readouterr().out.splitlines()
    # "%30s : %24.14e" (qws.cc:482): the name right-aligned in 30 columns
    assert out[0][:30] == "%30s" % "ddd_d_" and out[0][30:33] == " : " and len(out[0]) == 33 + 24
    na, nf = np.sum(a * a), np.sum(f.astype(np.float64) ** 2)
    assert abs(float(out[0][33:]) - na) < 1e-13 * na
    assert abs(float(out[1][33:]) - nf) < 1e-6 * nf
    L.zero_scd_field.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.zero_scd_field(_ptr(a), 2)
    assert not a[:384].any() and a[384:].all()
    L.get_clock.restype = ctypes.c_double
    t0 = L.get_clock()
    assert t0 > 1.5e9 and L.get_clock() >= t0
