"""Drop-in spinor transfers through the pinned staging chunks at production shape:
every host spinor a QWS entry point takes or returns crosses two pinned chunks
alternately (the host copy of one overlapping the DMA of the other).  With the chunk
size forced small (QWS_MI355X_PIN_CHUNK, read when qws_init_ creates the lattice)
the FP64 and FP32 spinors of a 32x8x8x8 lattice take 3 and 2 chunks with a partial
last one, and each chunk's host copy is large enough to run on several threads:
at kappa = 0 the operator is the identity, so ddd_d_/ddd_s_ must return their input
bit for bit; at kappa = 0.12 ddd_d_ must match the CPU oracle."""
import ctypes

import numpy as np
import pytest

import qws_amd as Q
from oracle import oracle as O
from oracle import ref_worker as RW
from tests import casekit as K

pytestmark = pytest.mark.gpu

NLOC = [32, 8, 8, 8]
CHUNK = 5 * 256 * 1024  # 1.25 MiB: FP64 spinor 3 MiB -> 3 chunks, FP32 1.5 MiB -> 2


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def test_multi_chunk_roundtrip_and_operator(monkeypatch):
    monkeypatch.setenv("QWS_MI355X_PIN_CHUNK", str(CHUNK))
    L = Q.lib()
    V = int(np.prod(NLOC))
    assert V * 192 > 2 * CHUNK and (V * 192) % CHUNK != 0 and V * 96 > CHUNK and (V * 96) % CHUNK != 0
    rng = np.random.default_rng(21)
    U, C = K.random_su3_like(V, rng)
    b = rng.standard_normal((V, 3, 4, 2))
    ci = lambda v: ctypes.byref(ctypes.c_int(v))
    keep = [(ctypes.c_int * 4)(1, 1, 1, 1), (ctypes.c_int * 4)(1, 1, 1, -1), ctypes.c_int(0), ctypes.c_int(1),
            (ctypes.c_int * 4)(1, 2, 2, 2)]
    L.qws_init_(*(ci(n) for n in NLOC), keep[0], keep[1], ctypes.byref(keep[2]), ctypes.byref(keep[3]), keep[4])
    try:
        vold = ctypes.c_int.in_dll(L, "vold").value
        vols = ctypes.c_int.in_dll(L, "vols").value

        def glob(name, ct, n):
            return np.ctypeslib.as_array((ct * n).from_address(ctypes.c_void_p.in_dll(L, name).value))

        glob("glud", ctypes.c_double, vold * 8 * 144)[:] = RW.links_to_aosoa(U, NLOC, 8).ravel()
        glob("glus", ctypes.c_float, vols * 8 * 288)[:] = RW.links_to_aosoa(U.astype(np.float32), NLOC, 16).ravel()
        glob("clvd", ctypes.c_double, vold * 2 * 576)[:] = RW.clover_to_aosoa(C, NLOC, 8).ravel()
        glob("clvs", ctypes.c_float, vols * 2 * 1152)[:] = RW.clover_to_aosoa(C.astype(np.float32), NLOC, 16).ravel()

        def set_kappa(k):
            ctypes.c_double.in_dll(L, "kappa").value = k
            ctypes.c_double.in_dll(L, "mkappa").value = -k
            ctypes.c_double.in_dll(L, "kappa2").value = k * k
            L.qws_mi355x_fields_modified_()

        bd = RW.spinor_to_aosoa(b, NLOC, 8)
        bs = RW.spinor_to_aosoa(b.astype(np.float32), NLOC, 16)
        set_kappa(0.0)
        out = np.full_like(bd, np.nan)
        L.ddd_d_(_ptr(out), _ptr(bd))
        assert np.array_equal(out, bd)
        outs = np.full_like(bs, np.nan)
        L.ddd_s_(_ptr(outs), _ptr(bs))
        assert np.array_equal(outs, bs)
        set_kappa(0.12)
        L.ddd_d_(_ptr(out), _ptr(bd))
        got = RW.spinor_from_aosoa(out, NLOC, 8)
        want = O.Oracle(NLOC, [NLOC[0] // 2] + NLOC[1:], -1, 0.12, U, C).dslash(b)
        assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-13
    finally:
        L.qws_finalize_()
