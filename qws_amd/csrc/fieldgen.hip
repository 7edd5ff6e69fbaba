// fieldgen.hip -- generated inputs for the GPU workloads (no reference
// counterpart: QWS receives its fields from the caller, qws_bqcd.cc:125-380).
//
// * Philox4x32-10 counter-based random numbers (Salmon et al., SC'11), one block
//   per (site, direction/component, purpose): fully parallel and independent of
//   the launch shape.  Counter = {lexicographic site, (mu<<8)|k, tag, 0}, key = seed.
// * SU(3) links U = exp(i eps H), H = (A + A^+)/2 - tr/3 with complex Gaussian A;
//   the exponential by Taylor series with scaling and squaring in FP64; stored in
//   QWS's convention c = U^T.
// * Clover term T = 1 + i kappa csw sum_{mu<nu} sigma_{mu nu} F_{mu nu} from the
//   clover-leaf field strength, built directly in the chirality basis of QWS's
//   packing, each 6x6 block inverted in FP64 (Gauss-Jordan, partial pivoting),
//   and stored as 1/2 T^-1 per chirality (clover_def.h:88-284).
#include "kernels.h"
#include "fieldgen.h"

namespace qws {

// ------------------------------------------------------------------ Philox4x32-10
__device__ __forceinline__ uint4 philox(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t lo0 = 0xD2511F53u * c.x, hi0 = __umulhi(0xD2511F53u, c.x);
    const uint32_t lo1 = 0xCD9E8D57u * c.z, hi1 = __umulhi(0xCD9E8D57u, c.z);
    c = make_uint4(hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0);
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

__device__ __forceinline__ Cx<double> gauss(uint32_t lex, uint32_t sub, uint32_t tag, uint64_t seed) {
  uint4 r = philox(make_uint4(lex, sub, tag, 0u), (uint32_t)seed, (uint32_t)(seed >> 32));
  const uint64_t a = ((((uint64_t)r.x) << 32) | r.y) >> 11;
  const uint64_t b = ((((uint64_t)r.z) << 32) | r.w) >> 11;
  const double u1 = ((double)a + 0.5) * 0x1.0p-53;
  const double u2 = ((double)b + 0.5) * 0x1.0p-53;
  const double rad = sqrt(-2.0 * log(u1));
  double sn, cs;
  sincos(2.0 * M_PI * u2, &sn, &cs);
  return {rad * cs, rad * sn};
}

__device__ __forceinline__ long long lex_of(const Geo& g, long long s, int x[4]) {
  SiteC c;
  site_coords(g, s, c);
#pragma unroll
  for (int m = 0; m < 4; ++m) x[m] = c.b[m] * g.B[m] + c.l[m];
  return x[0] + (long long)g.L[0] * (x[1] + (long long)g.L[1] * (x[2] + (long long)g.L[2] * x[3]));
}

// ------------------------------------------------------------------ 3x3 complex
struct M3 { Cx<double> a[9]; };

__device__ __forceinline__ M3 mmul(const M3& x, const M3& y) {
  M3 r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      Cx<double> s = {0, 0};
#pragma unroll
      for (int k = 0; k < 3; ++k) s = cadd(s, cmul(x.a[i * 3 + k], y.a[k * 3 + j]));
      r.a[i * 3 + j] = s;
    }
  return r;
}
__device__ __forceinline__ M3 mdag(const M3& x) {
  M3 r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) r.a[i * 3 + j] = cconj(x.a[j * 3 + i]);
  return r;
}

// exp(X) for X = i eps H: scale by 2^-s until the max row sum is <= 1/4, Taylor to
// order 14, square s times.
__device__ M3 mexp(M3 X) {
  double nrm = 0;
  for (int i = 0; i < 3; ++i) {
    double rs = 0;
    for (int j = 0; j < 3; ++j) rs += hypot(X.a[i * 3 + j].re, X.a[i * 3 + j].im);
    nrm = fmax(nrm, rs);
  }
  int s = 0;
  while (nrm > 0.25) {
    nrm *= 0.5;
    ++s;
  }
  const double sc = ldexp(1.0, -s);
  for (int k = 0; k < 9; ++k) X.a[k] = cscale(sc, X.a[k]);
  M3 E, term;
  for (int k = 0; k < 9; ++k) E.a[k] = term.a[k] = {(k % 4 == 0) ? 1.0 : 0.0, 0.0};
  for (int n = 1; n <= 14; ++n) {
    term = mmul(term, X);
    for (int k = 0; k < 9; ++k) {
      term.a[k] = cscale(1.0 / n, term.a[k]);
      E.a[k] = cadd(E.a[k], term.a[k]);
    }
  }
  for (int i = 0; i < s; ++i) E = mmul(E, E);
  return E;
}

__global__ void k_gen_links(Geo g, double eps, uint64_t seed, double* Ud) {
  const long long s = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (s >= g.V) return;
  int x[4];
  const uint32_t lex = (uint32_t)lex_of(g, s, x);
  double2* U = reinterpret_cast<double2*>(Ud);
  for (int mu = 0; mu < 4; ++mu) {
    M3 A, X;
    for (int k = 0; k < 9; ++k) A.a[k] = gauss(lex, (uint32_t)((mu << 8) | k), TAG_LINKS, seed);
    // H = (A + A^+)/2 - tr/3 ; X = i eps H
    Cx<double> H[9];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) H[i * 3 + j] = cscale(0.5, cadd(A.a[i * 3 + j], cconj(A.a[j * 3 + i])));
    const double tr = (H[0].re + H[4].re + H[8].re) / 3.0;
    for (int i = 0; i < 3; ++i) H[i * 4].re -= tr;
    for (int k = 0; k < 9; ++k) X.a[k] = cscale(eps, cimul(H[k]));
    M3 W = mexp(X);
    // QWS storage c[i][j] = W[j][i]
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        Cx<double> v = W.a[j * 3 + i];
        U[((long long)mu * 9 + i * 3 + j) * g.V + s] = make_double2(v.re, v.im);
      }
  }
}

// FP32 links (NL reals) from the FP64 links
__global__ void k_links_d2s(Geo g, int nl, const double* Ud, float* Us) {
  const long long s = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (s >= g.V) return;
  const int nc = nl / 2;
  const double2* src = reinterpret_cast<const double2*>(Ud);
  float2* dst = reinterpret_cast<float2*>(Us);
  for (int mu = 0; mu < 4; ++mu)
    for (int k = 0; k < nc; ++k) {
      double2 v = src[((long long)mu * 9 + k) * g.V + s];
      dst[((long long)mu * nc + k) * g.V + s] = make_float2((float)v.x, (float)v.y);
    }
}

// ------------------------------------------------------------------ clover
struct I4 { int v[4]; };

__device__ __forceinline__ M3 load_W(const Geo& g, const double* Ud, int mu, I4 x) {
  // physical W = c^T of the link at site x (periodic gauge field)
  SiteC c;
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const int xm = (x.v[m] % g.L[m] + g.L[m]) % g.L[m];
    c.b[m] = xm / g.B[m];
    c.l[m] = xm % g.B[m];
  }
  const long long s = site_index(g, c.b, c.l);
  const double2* U = reinterpret_cast<const double2*>(Ud);
  M3 W;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double2 v = U[((long long)mu * 9 + j * 3 + i) * g.V + s];
      W.a[i * 3 + j] = {v.x, v.y};
    }
  return W;
}

__device__ M3 field_strength(const Geo& g, const double* Ud, const int x[4], int mu, int nu) {
  auto at = [&](int dm, int dn) {
    I4 p = {{x[0], x[1], x[2], x[3]}};
    p.v[mu] += dm;
    p.v[nu] += dn;
    return p;
  };
  M3 Q;
  // leaf 1: W_mu(x) W_nu(x+mu) W_mu(x+nu)^+ W_nu(x)^+
  M3 l = mmul(mmul(load_W(g, Ud, mu, at(0, 0)), load_W(g, Ud, nu, at(1, 0))),
              mmul(mdag(load_W(g, Ud, mu, at(0, 1))), mdag(load_W(g, Ud, nu, at(0, 0)))));
  Q = l;
  // leaf 2: W_nu(x) W_mu(x+nu-mu)^+ W_nu(x-mu)^+ W_mu(x-mu)
  l = mmul(mmul(load_W(g, Ud, nu, at(0, 0)), mdag(load_W(g, Ud, mu, at(-1, 1)))),
           mmul(mdag(load_W(g, Ud, nu, at(-1, 0))), load_W(g, Ud, mu, at(-1, 0))));
  for (int k = 0; k < 9; ++k) Q.a[k] = cadd(Q.a[k], l.a[k]);
  // leaf 3: W_mu(x-mu)^+ W_nu(x-mu-nu)^+ W_mu(x-mu-nu) W_nu(x-nu)
  l = mmul(mmul(mdag(load_W(g, Ud, mu, at(-1, 0))), mdag(load_W(g, Ud, nu, at(-1, -1)))),
           mmul(load_W(g, Ud, mu, at(-1, -1)), load_W(g, Ud, nu, at(0, -1))));
  for (int k = 0; k < 9; ++k) Q.a[k] = cadd(Q.a[k], l.a[k]);
  // leaf 4: W_nu(x-nu)^+ W_mu(x-nu) W_nu(x-nu+mu) W_mu(x)^+
  l = mmul(mmul(mdag(load_W(g, Ud, nu, at(0, -1))), load_W(g, Ud, mu, at(0, -1))),
           mmul(load_W(g, Ud, nu, at(1, -1)), mdag(load_W(g, Ud, mu, at(0, 0)))));
  for (int k = 0; k < 9; ++k) Q.a[k] = cadd(Q.a[k], l.a[k]);
  M3 F;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) F.a[i * 3 + j] = cscale(0.125, csub(Q.a[i * 3 + j], cconj(Q.a[j * 3 + i])));
  return F;
}

// In-place inverse of a 6x6 complex matrix (Gauss-Jordan, partial pivoting).
__device__ void inv6(Cx<double> m[36]) {
  Cx<double> inv[36];
  for (int k = 0; k < 36; ++k) inv[k] = {(k % 7 == 0) ? 1.0 : 0.0, 0.0};
  for (int c = 0; c < 6; ++c) {
    int piv = c;
    double best = hypot(m[c * 6 + c].re, m[c * 6 + c].im);
    for (int r = c + 1; r < 6; ++r) {
      double v = hypot(m[r * 6 + c].re, m[r * 6 + c].im);
      if (v > best) { best = v; piv = r; }
    }
    if (piv != c)
      for (int k = 0; k < 6; ++k) {
        Cx<double> t = m[c * 6 + k]; m[c * 6 + k] = m[piv * 6 + k]; m[piv * 6 + k] = t;
        t = inv[c * 6 + k]; inv[c * 6 + k] = inv[piv * 6 + k]; inv[piv * 6 + k] = t;
      }
    const Cx<double> d = m[c * 6 + c];
    const double dn = d.re * d.re + d.im * d.im;
    const Cx<double> rd = {d.re / dn, -d.im / dn};
    for (int k = 0; k < 6; ++k) {
      m[c * 6 + k] = cmul(m[c * 6 + k], rd);
      inv[c * 6 + k] = cmul(inv[c * 6 + k], rd);
    }
    for (int r = 0; r < 6; ++r) {
      if (r == c) continue;
      const Cx<double> f = m[r * 6 + c];
      for (int k = 0; k < 6; ++k) {
        m[r * 6 + k] = csub(m[r * 6 + k], cmul(f, m[c * 6 + k]));
        inv[r * 6 + k] = csub(inv[r * 6 + k], cmul(f, inv[c * 6 + k]));
      }
    }
  }
  for (int k = 0; k < 36; ++k) m[k] = inv[k];
}

__global__ void __launch_bounds__(64) k_clover(Geo g, const double* Ud, double kcsw, SigmaChiral sig, double* Cd) {
  const long long s = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (s >= g.V) return;
  int x[4];
  lex_of(g, s, x);
  M3 F[6];
  int k = 0;
  for (int mu = 0; mu < 4; ++mu)
    for (int nu = mu + 1; nu < 4; ++nu) F[k++] = field_strength(g, Ud, x, mu, nu);
  double2* C = reinterpret_cast<double2*>(Cd);
  for (int i = 0; i < 2; ++i) {
    // T_i[(p,c),(p',c')] = delta + i kappa csw sum_k sig_k^(i)[p][p'] F_k[c][c'], index p*3 + c
    Cx<double> T[36];
    for (int r = 0; r < 6; ++r)
      for (int q = 0; q < 6; ++q) {
        const int p = r / 3, c = r % 3, pp = q / 3, cc = q % 3;
        Cx<double> acc = {0, 0};
        for (int kk = 0; kk < 6; ++kk) {
          const Cx<double> sg = {sig.v[kk][i][p][pp][0], sig.v[kk][i][p][pp][1]};
          acc = cadd(acc, cmul(sg, F[kk].a[c * 3 + cc]));
        }
        acc = cscale(kcsw, cimul(acc));
        if (r == q) acc.re += 1.0;
        T[r * 6 + q] = acc;
      }
    inv6(T);
    // pack 1/2 T^-1: diagonal reals, then the upper triangle row by row
    double a[36];
    for (int r = 0; r < 6; ++r) a[r] = 0.5 * T[r * 6 + r].re;
    int pk = 6;
    for (int r = 0; r < 6; ++r)
      for (int q = r + 1; q < 6; ++q) {
        a[pk++] = 0.5 * T[r * 6 + q].re;
        a[pk++] = 0.5 * T[r * 6 + q].im;
      }
    for (int f = 0; f < 18; ++f) C[(long long)(18 * i + f) * g.V + s] = make_double2(a[2 * f], a[2 * f + 1]);
  }
}

// FP32 clover (float4 [18][V]) from FP64 (double2 [36][V])
__global__ void k_clover_d2s(Geo g, const double* Cd, float* Cs) {
  const long long s = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (s >= g.V) return;
  const double2* src = reinterpret_cast<const double2*>(Cd);
  float4* dst = reinterpret_cast<float4*>(Cs);
  for (int f = 0; f < 18; ++f) {
    double2 a = src[(long long)(2 * f) * g.V + s], b = src[(long long)(2 * f + 1) * g.V + s];
    dst[(long long)f * g.V + s] = make_float4((float)a.x, (float)a.y, (float)b.x, (float)b.y);
  }
}

// ------------------------------------------------------------------ sources
__global__ void k_gaussian_spinor(Geo g, uint64_t seed, double* out) {
  const long long s = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (s >= g.V) return;
  int x[4];
  const uint32_t lex = (uint32_t)lex_of(g, s, x);
  Cx<double> v[12];
  for (int k = 0; k < 12; ++k) v[k] = gauss(lex, (uint32_t)k, TAG_SOURCE, seed);
  SpinorIO<double>::store(out, s, g.V, v);
}

__global__ void k_point_spinor(Geo g, int px, int py, int pz, int pt, int spin, int colour, double* out) {
  const long long s = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (s >= g.V) return;
  int x[4];
  lex_of(g, s, x);
  Cx<double> v[12];
  for (int k = 0; k < 12; ++k) v[k] = {0, 0};
  if (x[0] == px && x[1] == py && x[2] == pz && x[3] == pt) v[colour * 4 + spin] = {1, 0};
  SpinorIO<double>::store(out, s, g.V, v);
}

// ------------------------------------------------------------------ launchers
static dim3 grid_for(long long V, int T) { return dim3((unsigned)((V + T - 1) / T)); }

void launch_gen_links(const Geo& g, double eps, uint64_t seed, double* Ud, hipStream_t st) {
  hipLaunchKernelGGL(k_gen_links, grid_for(g.V, 128), dim3(128), 0, st, g, eps, seed, Ud);
}
void launch_links_d2s(const Geo& g, int nl, const double* Ud, float* Us, hipStream_t st) {
  hipLaunchKernelGGL(k_links_d2s, grid_for(g.V, 256), dim3(256), 0, st, g, nl, Ud, Us);
}
void launch_clover(const Geo& g, const double* Ud, double kcsw, const SigmaChiral& sig, double* Cd, hipStream_t st) {
  hipLaunchKernelGGL(k_clover, grid_for(g.V, 64), dim3(64), 0, st, g, Ud, kcsw, sig, Cd);
}
void launch_clover_d2s(const Geo& g, const double* Cd, float* Cs, hipStream_t st) {
  hipLaunchKernelGGL(k_clover_d2s, grid_for(g.V, 256), dim3(256), 0, st, g, Cd, Cs);
}
void launch_gaussian_spinor(const Geo& g, uint64_t seed, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_gaussian_spinor, grid_for(g.V, 128), dim3(128), 0, st, g, seed, out);
}
void launch_point_spinor(const Geo& g, const int x[4], int spin, int colour, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_point_spinor, grid_for(g.V, 256), dim3(256), 0, st, g, x[0], x[1], x[2], x[3], spin, colour,
                     out);
}

}  // namespace qws
