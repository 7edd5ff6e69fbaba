"""BQCD's host layouts, as QWS's loaders read them (qws_bqcd.cc:66-642).

TEST INFRASTRUCTURE ONLY: tests/ compare the product's results with it; the product package
``qws_amd`` never imports it.

BQCD is not part of QWS: it passes its own arrays and numbers the sites of each
parity itself (std_xyzt2i_, 1-based, and e_o_).  ``site_numbers`` is the numbering
of the tests' stand-in library oracle/libbqcd_sites.so (bqcd_sites.c): parity
(x+y+z+t) % 2 and index volh - lex/2 with lex = x + Lx (y + Ly (z + Lz t)).

Plain fields are the oracle's: links (4, V, 3, 3, 2) in QWS storage c = W^T (the
hop multiplies by the link matrix W[a][b] = c[b][a], wilson_d.h:110-135), clover
inverse (V, 2, 36) in QWS's packing (clover_def.h:103-126), spinors (V, 3, 4, 2);
V in lexicographic order.

* links: u[ri + 2a + 6b + 18 (i-1) + 18 volh eo + 36 volh mu] = W[a][b] -- BQCD's
  complex u(3, 3, volh, 2, 4) in Fortran order (qws_bqcd.cc:140-151);
* clover inverse: 72 reals per site at 72 (i-1) + 72 volh eo, chirality 0 then 1;
  each chirality's Hermitian 6x6 block A packed in row pairs (0,1), (2,3), (4,5):
  A[2k][2k], A[2k+1][2k+1], then the entries of row 2k right of its diagonal, then
  those of row 2k+1, as (re, im).  Built here from the dense matrix, not from QWS's
  index permutation (qws_bqcd.cc:301-337), so that a wrong entry of that
  permutation shows;
* spinors: h[ri + 2 s + 8 c + 24 (i-1)], one array per parity (qws_bqcd.cc:512-528).
"""
import numpy as np


def site_numbers(nloc):
    """(index 1..volh, parity) of every lexicographic site."""
    nx, ny, nz, nt = nloc
    t, z, y, x = np.meshgrid(np.arange(nt), np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    lex = (x + nx * (y + ny * (z + nz * t))).ravel()
    volh = nx * ny * nz * nt // 2
    return volh - lex // 2, ((x + y + z + t) % 2).ravel()


def links(U, nloc):
    """(4, V, 3, 3, 2) QWS storage -> BQCD u, flat."""
    i, eo = site_numbers(nloc)
    volh = len(i) // 2
    W = np.swapaxes(np.asarray(U, dtype=np.float64), 2, 3)  # W[mu, v, a, b]
    u = np.zeros((4, 2, volh, 3, 3, 2))  # [mu, eo, i-1, b, a, ri]: a fastest
    u[:, eo, i - 1] = np.swapaxes(W, 2, 3)
    return u.ravel()


def _qws_dense(C):
    """(V, 2, 36) QWS packing -> (V, 2, 6, 6) Hermitian blocks."""
    V = C.shape[0]
    A = np.zeros((V, 2, 6, 6), dtype=np.complex128)
    for r in range(6):
        A[:, :, r, r] = C[:, :, r]
    k = 6
    for r in range(6):
        for c in range(r + 1, 6):
            A[:, :, r, c] = C[:, :, k] + 1j * C[:, :, k + 1]
            A[:, :, c, r] = np.conj(A[:, :, r, c])
            k += 2
    return A


def clover(C, nloc):
    """(V, 2, 36) QWS packing -> BQCD clover inverse, flat."""
    i, eo = site_numbers(nloc)
    volh = len(i) // 2
    A = _qws_dense(np.asarray(C, dtype=np.float64))
    packed = []
    for p in range(3):
        r0, r1 = 2 * p, 2 * p + 1
        packed += [A[:, :, r0, r0].real, A[:, :, r1, r1].real]
        for r in (r0, r1):
            for c in range(r + 1, 6):
                packed += [A[:, :, r, c].real, A[:, :, r, c].imag]
    blk = np.stack(packed, axis=-1)  # (V, 2, 36)
    assert blk.shape[-1] == 36
    out = np.zeros((2, volh, 2, 36))
    out[eo, i - 1] = blk
    return out.ravel()


def spinor(psi, nloc):
    """(V, 3, 4, 2) -> (even, odd) BQCD arrays, each flat (volh * 24)."""
    i, eo = site_numbers(nloc)
    volh = len(i) // 2
    psi = np.asarray(psi, dtype=np.float64).reshape(-1, 24)
    out = np.zeros((2, volh, 24))
    out[eo, i - 1] = psi
    return out[0].ravel(), out[1].ravel()


def spinor_from(he, ho, nloc):
    """Inverse of ``spinor``."""
    i, eo = site_numbers(nloc)
    h = np.stack([np.asarray(he).reshape(-1, 24), np.asarray(ho).reshape(-1, 24)])
    return h[eo, i - 1].reshape(-1, 3, 4, 2)
