"""qwsg_prec_ddd_noalloc, the SAP preconditioner application that allocates nothing,
against qwsg_prec_ddd on the same lattice: bit-identical on consecutive calls whose
lattices shrink (its fixed scratch array then holds the previous call's larger
contents), and a lattice too large for the array is refused."""
import ctypes

import numpy as np
import pytest

import qws_amd as Q

pytestmark = pytest.mark.gpu


def _lattice(L, B, links, kappa, seed):
    lat = Q.Lattice(L, B, -1, links, kappa=kappa)
    lat.generate_gauge(0.3, seed)
    lat.compute_clover(1.0)
    src = np.random.default_rng(seed).standard_normal((lat.V, 3, 4, 2)).astype(np.float32)
    return lat, lat.spinor(Q.F32, src)


# largest first: each call runs on scratch the previous (larger) call left behind
CALLS = [
    ([16, 16, 16, 16], [4, 4, 4, 4], 12, 0.1365, 2, 6, 11),  # fused SAP with the face buffer
    ([8, 8, 8, 16], [4, 4, 4, 4], 12, 0.13, 3, 2, 12),  # smaller, fused
    ([8, 8, 8, 8], [4, 4, 4, 4], 18, 0.12, 0, 4, 13),  # nsap = 0: the final sequence only
    ([16, 6, 4, 3], [8, 6, 4, 3], 18, 0.12, 2, 3, 14),  # 144-site blocks: the step-by-step path
]


def test_noalloc_prec_ddd_is_bitwise_prec_ddd_on_shrinking_calls():
    for L, B, links, kappa, nsap, nm, seed in CALLS:
        print(f"noalloc prec_ddd L={L} B={B} links={links} nsap={nsap} nm={nm}", flush=True)
        lat, src = _lattice(L, B, links, kappa, seed)
        want = lat.spinor(Q.F32)
        lat.prec_ddd(want, src, nsap, nm)
        for rep in range(2):  # twice in a row on the same lattice as well
            got = lat.spinor(Q.F32)
            lat.prec_ddd_noalloc(got, src, nsap, nm)
            w, g = want.download(), got.download()
            assert np.array_equal(g, w), (L, rep, np.abs(g - w).max())
        assert np.linalg.norm(w) > 0


def test_noalloc_refuses_a_lattice_beyond_its_scratch():
    lat = Q.Lattice([32, 32, 32, 64], [4, 4, 4, 4], -1, 12, kappa=0.1365)  # 288 B/site: 604 MB
    a, b = lat.spinor(Q.F32), lat.spinor(Q.F32)
    rc = Q.lib().qwsg_prec_ddd_noalloc(lat.h, a.ptr, b.ptr, 2, 6)
    assert rc == 1
    msg = Q.lib().qwsg_last_error().decode()
    assert "scratch" in msg and "32^3x64 does not" in msg, msg
    # nsap/nm outside QWS's limits are refused as by prec_ddd_s_ (qws.cc:798-799)
    small, src = _lattice([8, 8, 8, 8], [4, 4, 4, 4], 12, 0.12, 3)
    out = small.spinor(Q.F32)
    assert Q.lib().qwsg_prec_ddd_noalloc(small.h, out.ptr, src.ptr, 11, 2) == 1
