"""CPU tests of the library's host-side code: the FP64<->FP32 reorders
(reordering.cc:59-204) and the driver helpers (qws.cc:445-500).  They touch only host
memory, so they run without a GPU: the geometry globals are set directly and the field
pointers point at numpy buffers for the duration of a test (restored afterwards, so
qws_init_ never frees them).  (The BQCD loaders and storers permute on the GPU; their
tests are in test_gpu_bqcd.py.)

Expected values: the reference's AoSoA layouts as oracle/ref_worker.py states them (the
layouts that the reference-library parity tests use)."""
import ctypes
import os

import numpy as np
import pytest

import qws_amd as Q
from oracle import ref_worker as RW

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
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


def _fields(seed):
    V = int(np.prod(NLOC))
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((4, V, 3, 3, 2)), rng.standard_normal((V, 2, 36)),
            rng.standard_normal((V, 3, 4, 2)))


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
