"""A constant abelian field strength in the x-y plane, for the closed-form check of
the clover term's normalisation and sign.

Links (physical W, QWS stores c = W^T; W is diagonal here so c = W), colour charges
q = (1, -1, 0) (an element of SU(3)'s Cartan subgroup, det W = 1):
    W_x(x, y, z, t) = diag exp(i phi q y)
    W_y(x, y, z, t) = diag exp(-i phi q Ly x)  on y = Ly - 1, else 1
    W_z = W_t = 1,        phi = 2 pi n / (Lx Ly)  (flux quantised on the torus)
Every x-y plaquette W_x(x) W_y(x+x) W_x(x+y)^+ W_y(x)^+ equals exp(-i phi q), on the
wrapped rows too; all other plaquettes are 1.  So the four clover leaves are equal,
Q = 4 exp(-i phi q), and the clover-leaf field strength F = (Q - Q^+)/8 =
-i sin(phi q).  With T = 1 + i kappa csw sum_{mu<nu} sigma_{mu nu} F_{mu nu} and
sigma_{mu nu} = (i/2)[gamma_mu, gamma_nu] (Luscher, Sint, Sommer, Weisz 1996, in the
hopping-parameter normalisation), T = 1 + kappa csw sin(phi q) sigma_01.  In QWS's
spin basis (gamma_k = [[0, i s_k], [-i s_k, 0]], gamma_4 = diag(1, 1, -1, -1), the
matrices its projectors imply) gamma_0 gamma_1 = i diag(s_3, s_3), so
sigma_01 = -diag(s_3, s_3) = diag(-1, 1, -1, 1).  T is diagonal: spins 0 and 2 of
colour c get 1 - kappa csw sin(phi q_c), spins 1 and 3 get 1 + kappa csw sin(phi q_c).
In QWS's packing (per chirality 1/2 T^-1 in the basis (s0 +- s2, s1 +- s3), real
diagonal entry 3 p + c for spin pair p) both chiralities hold 0.5 / (1 -+ a_c) on the
diagonal and nothing off it.
"""
import numpy as np

CHARGES = (1, -1, 0)


def abelian_links(L, n=1):
    """(4, V, 3, 3, 2) links of the constant x-y field with n flux quanta, and phi."""
    Lx, Ly, Lz, Lt = L
    phi = 2 * np.pi * n / (Lx * Ly)
    V = Lx * Ly * Lz * Lt
    s = np.arange(V)
    x, y = s % Lx, (s // Lx) % Ly
    q = np.array(CHARGES, dtype=float)
    ph = np.zeros((4, V, 3))
    ph[0] = phi * np.outer(y, q)
    ph[1] = np.where((y == Ly - 1)[:, None], -phi * Ly * np.outer(x, q), 0.0)
    U = np.zeros((4, V, 3, 3, 2))
    for c in range(3):
        U[:, :, c, c, 0] = np.cos(ph[..., c])
        U[:, :, c, c, 1] = np.sin(ph[..., c])
    return U, phi


def expected_packed_clover(V, phi, kappa, csw):
    """(V, 2, 36): 0.5 / (1 - a_c) for spin pair 0, 0.5 / (1 + a_c) for pair 1."""
    out = np.zeros((V, 2, 36))
    for c, q in enumerate(CHARGES):
        a = kappa * csw * np.sin(phi * q)
        out[:, :, c] = 0.5 / (1 - a)
        out[:, :, 3 + c] = 0.5 / (1 + a)
    return out
