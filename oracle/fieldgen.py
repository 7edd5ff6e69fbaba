"""Oracle for the generated inputs of the GPU workloads (no reference counterpart).

TEST INFRASTRUCTURE ONLY.  QWS receives its links and clover inverse from the
caller (qws_bqcd.cc:125-380) or from its driver's random numbers (main.cc:406-551);
it has no link generator, no clover construction and no point sources.  The GPU
library generates them; this module restates each step independently in numpy,
for small lattices:

* random numbers: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), pinned by its
  published known-answer vectors; uniform doubles u = ((hi:lo >> 11) + 1/2) 2^-53,
  Gaussians by Box-Muller;
* links: U = exp(i eps H), H = (A + A^+)/2 - tr/3 with A complex Gaussian (re, im
  ~ N(0,1)); the exponential through the eigendecomposition of H (the GPU uses a
  Taylor series with scaling and squaring); stored as QWS's c = U^T (the hop
  multiplies by W[i][j] = c[j][i], wilson_d.h:110-135);
* clover: T = 1 + i kappa csw sum_{mu<nu} sigma_{mu nu} F_{mu nu}, sigma_{mu nu} =
  (i/2)[gamma_mu, gamma_nu] with the gamma matrices of QWS's projectors, F the
  clover-leaf field strength (Q - Q^+)/8 (Luscher, Sint, Sommer, Weisz 1996);
  inverted as a dense 12x12 (the GPU inverts the two 6x6 chirality blocks) and
  packed as QWS stores it: per chirality 1/2 T^-1 in the (s0+-s2, s1+-s3) basis
  (clover_def.h:88-284).
"""
import numpy as np

M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
W0, W1 = np.uint32(0x9E3779B9), np.uint32(0xBB67AE85)
MASK = np.uint64(0xFFFFFFFF)

TAG_LINKS, TAG_SOURCE = 1, 2


def philox4x32(ctr, key):
    """ctr: (..., 4) uint32, key: (2,) uint32 -> (..., 4) uint32 (10 rounds)."""
    c = [np.asarray(ctr[..., i], dtype=np.uint64) for i in range(4)]
    k0, k1 = np.uint32(key[0]), np.uint32(key[1])
    for _ in range(10):
        p0 = M0 * c[0]
        p1 = M1 * c[2]
        n0 = ((p1 >> np.uint64(32)) ^ c[1] ^ np.uint64(k0)) & MASK
        n1 = p1 & MASK
        n2 = ((p0 >> np.uint64(32)) ^ c[3] ^ np.uint64(k1)) & MASK
        n3 = p0 & MASK
        c = [n0, n1, n2, n3]
        k0 = np.uint32((int(k0) + int(W0)) & 0xFFFFFFFF)
        k1 = np.uint32((int(k1) + int(W1)) & 0xFFFFFFFF)
    return np.stack([x.astype(np.uint32) for x in c], axis=-1)


def seed_key(seed):
    return np.array([seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF], dtype=np.uint32)


def gauss_complex(ctr, key):
    """Complex Gaussian (re, im ~ N(0,1)) from one Philox block per counter."""
    r = philox4x32(ctr, key).astype(np.uint64)
    a = ((r[..., 0] << np.uint64(32)) | r[..., 1]) >> np.uint64(11)
    b = ((r[..., 2] << np.uint64(32)) | r[..., 3]) >> np.uint64(11)
    u1 = (a.astype(np.float64) + 0.5) * 2.0 ** -53
    u2 = (b.astype(np.float64) + 0.5) * 2.0 ** -53
    rad = np.sqrt(-2.0 * np.log(u1))
    return rad * np.cos(2 * np.pi * u2) + 1j * rad * np.sin(2 * np.pi * u2)


def _counters(V, n2, tag):
    lex = np.arange(V, dtype=np.uint32)
    ctr = np.zeros((V, len(n2), 4), dtype=np.uint32)
    ctr[..., 0] = lex[:, None]
    ctr[..., 1] = np.asarray(n2, dtype=np.uint32)[None, :]
    ctr[..., 2] = tag
    return ctr


def su3_links(L, eps, seed):
    """Links (4, V, 3, 3, 2) in QWS storage (c = U^T), U = exp(i eps H)."""
    V = int(np.prod(L))
    key = seed_key(seed)
    out = np.zeros((4, V, 3, 3, 2))
    for mu in range(4):
        ctr = _counters(V, [(mu << 8) | k for k in range(9)], TAG_LINKS)
        A = gauss_complex(ctr, key).reshape(V, 3, 3)
        H = 0.5 * (A + np.conj(np.swapaxes(A, 1, 2)))
        tr = np.trace(H, axis1=1, axis2=2).real / 3.0
        H = H - tr[:, None, None] * np.eye(3)[None]
        w, v = np.linalg.eigh(H)
        U = np.einsum("nij,nj,nkj->nik", v, np.exp(1j * eps * w), np.conj(v))
        c = np.swapaxes(U, 1, 2)  # c = U^T
        out[mu, ..., 0] = c.real
        out[mu, ..., 1] = c.imag
    return out


def gaussian_source(L, seed):
    """Spinor (V, 3, 4, 2) with re, im ~ N(0,1) per component k = colour*4 + spin."""
    V = int(np.prod(L))
    z = gauss_complex(_counters(V, list(range(12)), TAG_SOURCE), seed_key(seed)).reshape(V, 3, 4)
    return np.stack([z.real, z.imag], axis=-1)


def point_source(L, spin, colour, site=(0, 0, 0, 0)):
    V = int(np.prod(L))
    psi = np.zeros((V, 3, 4, 2))
    x, y, z, t = site
    psi[x + L[0] * (y + L[1] * (z + L[2] * t)), colour, spin, 0] = 1.0
    return psi


def gammas():
    """Euclidean gamma matrices of QWS's projector tables (wilson_d.h:540-657)."""
    s = [np.array([[0, 1], [1, 0]], complex), np.array([[0, -1j], [1j, 0]]), np.array([[1, 0], [0, -1]], complex)]
    z = np.zeros((2, 2), complex)
    g = [np.block([[z, 1j * sk], [-1j * sk, z]]) for sk in s]
    g.append(np.diag([1, 1, -1, -1]).astype(complex))
    return g


def _W(U):
    """Physical link matrices W[mu] (V,3,3) from QWS storage c: W = c^T."""
    c = U[..., 0] + 1j * U[..., 1]
    return np.swapaxes(c, -1, -2)


def field_strength(L, U):
    """F_{mu nu}(x) = (Q - Q^+)/8 with Q the four clover leaves; (6, V, 3, 3) for
    (mu,nu) in (0,1),(0,2),(0,3),(1,2),(1,3),(2,3)."""
    Lx, Ly, Lz, Lt = L
    W = _W(U).reshape(4, Lt, Lz, Ly, Lx, 3, 3)
    axis = {0: 3, 1: 2, 2: 1, 3: 0}  # lattice direction -> array axis of (t,z,y,x)

    def sh(a, mu, n):  # value at x + n*mu
        return np.roll(a, -n, axis=axis[mu])

    dag = lambda a: np.conj(np.swapaxes(a, -1, -2))
    mm = lambda *ms: _chain(ms)
    F = []
    for mu in range(4):
        for nu in range(mu + 1, 4):
            Wm, Wn = W[mu], W[nu]
            l1 = mm(Wm, sh(Wn, mu, 1), dag(sh(Wm, nu, 1)), dag(Wn))
            l2 = mm(Wn, dag(sh(sh(Wm, nu, 1), mu, -1)), dag(sh(Wn, mu, -1)), sh(Wm, mu, -1))
            l3 = mm(dag(sh(Wm, mu, -1)), dag(sh(sh(Wn, mu, -1), nu, -1)), sh(sh(Wm, mu, -1), nu, -1), sh(Wn, nu, -1))
            l4 = mm(dag(sh(Wn, nu, -1)), sh(Wm, nu, -1), sh(sh(Wn, nu, -1), mu, 1), dag(Wm))
            Q = l1 + l2 + l3 + l4
            F.append(((Q - dag(Q)) / 8.0).reshape(-1, 3, 3))
    return np.array(F)


def _chain(ms):
    r = ms[0]
    for m in ms[1:]:
        r = r @ m
    return r


def clover_T(L, U, kappa, csw):
    """Dense clover term T (V, 12, 12), basis index colour*4 + spin."""
    g = gammas()
    F = field_strength(L, U)
    V = F.shape[1]
    T = np.tile(np.eye(12, dtype=complex), (V, 1, 1))
    k = 0
    for mu in range(4):
        for nu in range(mu + 1, 4):
            sig = 0.5j * (g[mu] @ g[nu] - g[nu] @ g[mu])
            T += 1j * kappa * csw * np.einsum("nab,st->nasbt", F[k], sig).reshape(V, 12, 12)
            k += 1
    return T


def chiral_rotation():
    """12x12 orthogonal map to the chirality basis used by QWS: row (i, p, c) for
    chirality i (s0 +- s2, s1 +- s3)/sqrt2, spin pair p, colour c."""
    Q = np.zeros((12, 12))
    r = 1 / np.sqrt(2)
    for i, sg in enumerate((1, -1)):
        for p in range(2):
            for c in range(3):
                row = 6 * i + 3 * p + c
                Q[row, c * 4 + p] = r
                Q[row, c * 4 + p + 2] = sg * r
    return Q


def pack_clover_inverse(Tinv):
    """(V,12,12) T^-1 -> QWS packing (V, 2, 36) of 1/2 T^-1 per chirality; also
    returns the largest off-chirality element (must vanish)."""
    Q = chiral_rotation()
    Tc = np.einsum("ab,nbc,dc->nad", Q, Tinv, Q)
    off = max(np.abs(Tc[:, :6, 6:]).max(), np.abs(Tc[:, 6:, :6]).max())
    V = Tinv.shape[0]
    out = np.zeros((V, 2, 36))
    pairs = [(r, c) for r in range(6) for c in range(r + 1, 6)]
    for i in range(2):
        A = 0.5 * Tc[:, 6 * i:6 * i + 6, 6 * i:6 * i + 6]
        out[:, i, :6] = np.real(np.diagonal(A, axis1=1, axis2=2))
        for k, (r, c) in enumerate(pairs):
            out[:, i, 6 + 2 * k] = A[:, r, c].real
            out[:, i, 7 + 2 * k] = A[:, r, c].imag
    return out, off


def clover_inverse(L, U, kappa, csw):
    T = clover_T(L, U, kappa, csw)
    packed, off = pack_clover_inverse(np.linalg.inv(T))
    return packed, off, T
