"""Restatement of QWS's even/odd Krylov solvers on a Hermitian positive operator.

TEST INFRASTRUCTURE ONLY: tests/ compare the product's results with it; the product package
``qws_amd`` never imports it.

Both solvers work on real vectors (a parity field as its 24*V/2 reals) with an
``apply(v) -> A v`` callable, A = mtilde^+ mtilde in QWS; the inner product that
QWS forms (sum of re*re + im*im over the complex components) is the real dot
product of those reals.

* ``cg`` follows cg_vm_ (cg.cc:65-172): q = A x, r = b - q, p = r,
  rho0 = |r|^2 (cg.cc:75-96); for iter = 1 .. maxiter-1: q = A p,
  alpha = rho0 / <p, q> (cg.cc:100-116), x = x + alpha p, r = r - alpha q
  (cg.cc:118-131), stop when sqrt(|r|^2 / |b|^2) < tol (cg.cc:148),
  beta = |r|^2 / rho0, p = r + beta p (cg.cc:153-167).
* ``mcg`` follows mcg_vm_ (mcg.cc:65-219): one Krylov space for every shift
  sigma_k, solving (A + sigma_k) x_k = b from zero.  alpha is the negated CG step
  (mcg.cc:131), r = r + alpha q (mcg.cc:134-140); per shift the scale
  xi2 = xi1 xi0 alpha0 / (alpha beta (xi0 - xi1) + xi0 alpha0 (1 - sigma alpha))
  and x_k = x_k - (alpha xi2 / xi1) pp_k (mcg.cc:158-172), then the stop test on
  the unshifted residual (mcg.cc:176), p = r + beta p (mcg.cc:180-195) and
  pp_k = beta (xi2/xi1)^2 pp_k + xi2 r (mcg.cc:197-212).

The update forms follow the reference's fused multiply-adds where numpy can:
numpy rounds a*x + y in two steps, so results agree with the reference to
rounding, not bit for bit.
"""
import numpy as np


def cg(apply, b, x0, maxiter, tol=1e-14):
    """x solving A x = b from x0; returns (x, iterations run)."""
    x = np.array(x0, dtype=np.float64, copy=True)
    b = np.asarray(b, dtype=np.float64)
    q = apply(x)
    r = b - q
    p = r.copy()
    bnorm = float(b @ b)
    rho0 = float(r @ r)
    it = 1
    while it < maxiter:
        q = apply(p)
        alpha = rho0 / float(p @ q)
        x = x + alpha * p
        r = r - alpha * q
        rnorm = float(r @ r)
        if np.sqrt(rnorm / bnorm) < tol:
            break
        beta = rnorm / rho0
        rho0 = rnorm
        p = r + beta * p
        it += 1
    return x, it


def mcg(apply, b, sigma, maxiter, tol=1e-14):
    """[x_k] solving (A + sigma_k) x_k = b from zero; returns (xs, iterations run)."""
    b = np.asarray(b, dtype=np.float64)
    n = len(sigma)
    r = b.copy()
    p = b.copy()
    bnorm = float(b @ b)
    rho0 = bnorm
    xs = [np.zeros_like(b) for _ in range(n)]
    pp = [b.copy() for _ in range(n)]
    xi0, xi1, xi2 = [1.0] * n, [1.0] * n, [0.0] * n
    alpha0, beta = 1.0, 0.0
    it = 1
    while it < maxiter:
        q = apply(p)
        alpha = -rho0 / float(p @ q)
        r = r + alpha * q
        rnorm = float(r @ r)
        for k in range(n):
            xi2[k] = xi1[k] * xi0[k] * alpha0 / (alpha * beta * (xi0[k] - xi1[k])
                                                 + xi0[k] * alpha0 * (1 - sigma[k] * alpha))
            xs[k] = xs[k] - (alpha * xi2[k] / xi1[k]) * pp[k]
        if np.sqrt(rnorm / bnorm) < tol:
            break
        beta = rnorm / rho0
        rho0 = rnorm
        p = r + beta * p
        for k in range(n):
            pp[k] = beta * (xi2[k] / xi1[k]) ** 2 * pp[k] + xi2[k] * r
            xi0[k] = xi1[k]
            xi1[k] = xi2[k]
        alpha0 = alpha
        it += 1
    return xs, it
