"""GPU parity of the generated inputs (SU(3) links, clover inverse from links,
Gaussian and point sources) against the numpy oracle in oracle/fieldgen.py.
These inputs have no reference counterpart; they are pinned by the oracle (an
independent restatement: eigendecomposition vs Taylor exponential, dense 12x12
vs per-chirality 6x6 inversion) and by closed forms: unitarity, det U = 1, the
unit gauge giving T = 1, and gauge covariance of D with its clover term."""
import numpy as np
import pytest

import qws_amd as Q
from oracle import fieldgen as FG
from oracle import oracle as O

pytestmark = pytest.mark.gpu

L, B = [4, 4, 4, 8], [2, 2, 2, 4]


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


@pytest.fixture(scope="module")
def gen():
    lat = Q.Lattice(L, B, -1, 12, kappa=0.1365)
    lat.generate_gauge(0.3, 2)
    lat.compute_clover(1.0)
    return lat


def test_links_match_oracle(gen):
    U = gen.gauge()
    want = FG.su3_links(L, 0.3, 2)
    assert np.abs(U - want).max() < 1e-13
    c = U[..., 0] + 1j * U[..., 1]
    assert np.abs(np.einsum("mnij,mnkj->mnik", c, c.conj()) - np.eye(3)).max() < 1e-13
    assert np.abs(np.linalg.det(c) - 1).max() < 1e-13


def test_clover_matches_oracle(gen):
    C = gen.clover_field()
    want, off, T = FG.clover_inverse(L, FG.su3_links(L, 0.3, 2), 0.1365, 1.0)
    assert off < 1e-13
    assert np.abs(C - want).max() < 1e-12


def test_unit_gauge_clover_is_identity():
    lat = Q.Lattice(L, B, -1, 18, kappa=0.13)
    U = np.zeros((4, int(np.prod(L)), 3, 3, 2))
    for a in range(3):
        U[:, :, a, a, 0] = 1
    lat.load_gauge(U)
    lat.compute_clover(1.7)
    C = lat.clover_field()
    assert np.abs(C[:, :, :6] - 0.5).max() < 1e-15 and np.abs(C[:, :, 6:]).max() < 1e-15


def test_sources(gen):
    s = gen.gaussian(Q.F64, 9).download()
    assert np.abs(s - FG.gaussian_source(L, 9)).max() < 1e-13
    p = gen.point(Q.F64, (1, 2, 3, 5), 2, 1).download()
    assert np.array_equal(p, FG.point_source(L, 2, 1, (1, 2, 3, 5)))


def test_gauge_covariance():
    """D[U^G] (G psi) = G (D[U] psi) with the clover term rebuilt from U^G."""
    V = int(np.prod(L))
    lat = Q.Lattice(L, B, -1, 18, kappa=0.13)
    lat.generate_gauge(0.4, 5)
    U = lat.gauge()
    lat.compute_clover(1.3)
    rng = np.random.default_rng(3)
    z = rng.standard_normal((V, 3, 3)) + 1j * rng.standard_normal((V, 3, 3))
    G, _ = np.linalg.qr(z)  # unitary per site
    W = np.swapaxes(U[..., 0] + 1j * U[..., 1], -1, -2)  # physical W = c^T
    idx = np.arange(V)
    x = [idx % L[0], (idx // L[0]) % L[1], (idx // (L[0] * L[1])) % L[2], idx // (L[0] * L[1] * L[2])]
    Wg = np.empty_like(W)
    for mu in range(4):
        xp = [c.copy() for c in x]
        xp[mu] = (xp[mu] + 1) % L[mu]
        nb = xp[0] + L[0] * (xp[1] + L[1] * (xp[2] + L[2] * xp[3]))
        Wg[mu] = G @ W[mu] @ np.conj(np.swapaxes(G[nb], -1, -2))
    cg = np.swapaxes(Wg, -1, -2)
    Ug = np.stack([cg.real, cg.imag], axis=-1)
    latg = Q.Lattice(L, B, -1, 18, kappa=0.13)
    latg.load_gauge(Ug)
    latg.compute_clover(1.3)
    psi = rng.standard_normal((V, 3, 4, 2))
    pc = psi[..., 0] + 1j * psi[..., 1]
    gpsi = np.einsum("nab,nbs->nas", G, pc)
    out = lat.spinor(Q.F64)
    lat.dslash(out, lat.spinor(Q.F64, psi))
    outg = latg.spinor(Q.F64)
    latg.dslash(outg, latg.spinor(Q.F64, np.stack([gpsi.real, gpsi.imag], -1)))
    d = out.download()
    want = np.einsum("nab,nbs->nas", G, d[..., 0] + 1j * d[..., 1])
    got = outg.download()
    assert np.abs(got[..., 0] + 1j * got[..., 1] - want).max() < 1e-12


def test_generated_fields_drive_the_operator(gen):
    """FP32 D with 12-real links on the generated fields equals the oracle's D."""
    U, C = gen.gauge(), gen.clover_field()
    orc = O.Oracle(L, B, -1, 0.1365, U, C)
    b = FG.gaussian_source(L, 4)
    out = gen.spinor(Q.F32)
    gen.dslash(out, gen.spinor(Q.F32, b))
    assert rel(out.download(), orc.dslash(b.astype(np.float32))) < 1e-6


@pytest.mark.parametrize("n,kappa,csw", [(1, 0.13, 2.0), (3, 0.12, 1.5)])
def test_clover_constant_abelian_field_closed_form(n, kappa, csw):
    """The GPU's clover inverse on a constant x-y field strength has the spectrum
    1 -+ kappa csw sin(phi q) by spin (closed form, tests/clover_field.py): pins the
    normalisation and sign of the clover term, which gauge covariance and the unit
    gauge leave open."""
    from tests import clover_field as CF

    Lg, Bg = [4, 4, 4, 8], [2, 2, 2, 4]
    U, phi = CF.abelian_links(Lg, n)
    lat = Q.Lattice(Lg, Bg, -1, 18, kappa=kappa)
    lat.load_gauge(U)
    lat.compute_clover(csw)
    C = lat.clover_field()
    want = CF.expected_packed_clover(int(np.prod(Lg)), phi, kappa, csw)
    assert np.abs(C - want).max() < 1e-14
