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


def test_gpu_sync_fields_uploads_the_host_arrays():
    """Fields written through QWS's global pointers after a first operator call reach
    the GPU only when the caller says so: without a call the operator keeps the old
    fields; after qws_gpu_sync_fields_() ddd_d_ matches the oracle on the new ones."""
    L = Q.lib()
    nloc = [32, 4, 4, 4]  # QWS's FP32 layout needs nx/2 to be a multiple of 16
    V = int(np.prod(nloc))
    rng = np.random.default_rng(33)
    U1, C1 = K.random_su3_like(V, rng)
    U2, C2 = K.random_su3_like(V, rng)
    b = rng.standard_normal((V, 3, 4, 2))
    ci = lambda v: ctypes.byref(ctypes.c_int(v))
    keep = [(ctypes.c_int * 4)(1, 1, 1, 1), (ctypes.c_int * 4)(1, 1, 1, -1), ctypes.c_int(0), ctypes.c_int(1),
            (ctypes.c_int * 4)(1, 2, 2, 2)]
    L.qws_init_(*(ci(n) for n in nloc), keep[0], keep[1], ctypes.byref(keep[2]), ctypes.byref(keep[3]), keep[4])
    try:
        vold = ctypes.c_int.in_dll(L, "vold").value

        def glob(name, ct, n):
            return np.ctypeslib.as_array((ct * n).from_address(ctypes.c_void_p.in_dll(L, name).value))

        def write(U, C):
            glob("glud", ctypes.c_double, vold * 8 * 144)[:] = RW.links_to_aosoa(U, nloc, 8).ravel()
            glob("clvd", ctypes.c_double, vold * 2 * 576)[:] = RW.clover_to_aosoa(C, nloc, 8).ravel()

        kap = 0.11
        ctypes.c_double.in_dll(L, "kappa").value = kap
        ctypes.c_double.in_dll(L, "mkappa").value = -kap
        ctypes.c_double.in_dll(L, "kappa2").value = kap * kap
        blocks = [nloc[0] // 2] + nloc[1:]
        bd = RW.spinor_to_aosoa(b, nloc, 8)
        out = np.zeros_like(bd)
        write(U1, C1)
        L.qws_gpu_sync_fields_()
        L.ddd_d_(_ptr(out), _ptr(bd))
        want1 = O.Oracle(nloc, blocks, -1, kap, U1, C1).dslash(b)
        got = RW.spinor_from_aosoa(out, nloc, 8)
        assert np.linalg.norm(got - want1) / np.linalg.norm(want1) < 1e-13
        write(U2, C2)  # not yet announced: the GPU still holds U1, C1
        L.ddd_d_(_ptr(out), _ptr(bd))
        assert np.array_equal(RW.spinor_from_aosoa(out, nloc, 8), got)
        L.qws_gpu_sync_fields_()
        L.ddd_d_(_ptr(out), _ptr(bd))
        want2 = O.Oracle(nloc, blocks, -1, kap, U2, C2).dslash(b)
        got2 = RW.spinor_from_aosoa(out, nloc, 8)
        assert np.linalg.norm(got2 - want2) / np.linalg.norm(want2) < 1e-13
    finally:
        L.qws_finalize_()
