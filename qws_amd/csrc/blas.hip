// blas.hip -- fused BiCGStab vector updates and reductions, batched over RHS.
//
// Vectors are flat arrays of interleaved (re,im) pairs; a batch of nrhs vectors
// sits at base + r * stride.  The same kernels serve the FP32 inner solver
// (bicgstab_precdd_s.cc:85-559) and the FP64 outer solver (bicgstab_dd_mix.cc:105-388).
// blockIdx.y selects the right-hand side; RHS whose active flag is 0 are skipped.
// Every reduction accumulates in FP64 and writes one partial per workgroup;
// k_finalize sums the partials in a fixed order (bit-reproducible, and independent
// of what else is in the batch) and updates that RHS's scalars in device memory,
// so an iteration needs no host round trip unless the caller asks for residuals.
#include <math.h>

#include "kernels.h"

namespace qws {

template <typename R> struct V2;
template <> struct V2<float> { using T = float2; };
template <> struct V2<double> { using T = double2; };

__device__ __forceinline__ bool rhs_off(const int* active) { return active && !active[blockIdx.y]; }

// Block-wide sum of NV doubles; thread 0 writes part[(r*4 + v)*RED_BLOCKS + blockIdx.x].
template <int NV>
__device__ __forceinline__ void block_reduce_store(double (&v)[NV], double* part) {
  __shared__ double sh[NV][RED_THREADS / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double x = v[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    if (lane == 0) sh[k][wv] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double x = 0;
      for (int w = 0; w < RED_THREADS / 64; ++w) x += sh[k][w];
      part[((long long)blockIdx.y * 4 + k) * RED_BLOCKS + blockIdx.x] = x;
    }
  }
}

#define GRID_STRIDE(i, n) \
  for (long long i = blockIdx.x * (long long)RED_THREADS + threadIdx.x; i < (n); i += (long long)RED_BLOCKS * RED_THREADS)
#define RHS(ptr, T) (reinterpret_cast<T*>(ptr) + (long long)blockIdx.y * (stride / 2))

template <typename R>
__global__ void __launch_bounds__(RED_THREADS)
k_init_inner(long long n, long long stride, const int* active, R* x, R* r, R* p, R* r0, const R* b, double* part) {
  using T = typename V2<R>::T;
  if (rhs_off(active)) return;
  T *X = RHS(x, T), *Rr = RHS(r, T), *P = RHS(p, T), *R0 = RHS(r0, T);
  const T* B = RHS(b, const T);
  double acc[1] = {0};
  GRID_STRIDE(i, n) {
    T bv = B[i];
    T z;
    z.x = 0;
    z.y = 0;
    X[i] = z;
    Rr[i] = bv;
    P[i] = bv;
    R0[i] = bv;
    acc[0] += (double)bv.x * bv.x + (double)bv.y * bv.y;
  }
  block_reduce_store<1>(acc, part);
}

template <typename R>
__global__ void __launch_bounds__(RED_THREADS)
k_init_outer(long long n, long long stride, const int* active, R* r, R* p, R* r0, const R* b, const R* q, double* part) {
  using T = typename V2<R>::T;
  if (rhs_off(active)) return;
  T *Rr = RHS(r, T), *P = RHS(p, T), *R0 = RHS(r0, T);
  const T *B = RHS(b, const T), *Qv = RHS(q, const T);
  double acc[2] = {0, 0};
  GRID_STRIDE(i, n) {
    T bv = B[i], qv = Qv[i], rv;
    rv.x = bv.x - qv.x;
    rv.y = bv.y - qv.y;
    Rr[i] = rv;
    P[i] = rv;
    R0[i] = rv;
    acc[0] += (double)bv.x * bv.x + (double)bv.y * bv.y;
    acc[1] += (double)rv.x * rv.x + (double)rv.y * rv.y;
  }
  block_reduce_store<2>(acc, part);
}

// <a,b> = sum conj(a) b  (bicgstab_dd_mix.cc:180-181)
template <typename R>
__global__ void __launch_bounds__(RED_THREADS)
k_dot(long long n, long long stride, const int* active, const R* a, const R* b, double* part) {
  using T = typename V2<R>::T;
  if (rhs_off(active)) return;
  const T *A = RHS(a, const T), *B = RHS(b, const T);
  double acc[2] = {0, 0};
  GRID_STRIDE(i, n) {
    T av = A[i], bv = B[i];
    acc[0] += (double)av.x * bv.x + (double)av.y * bv.y;
    acc[1] += (double)av.x * bv.y - (double)av.y * bv.x;
  }
  block_reduce_store<2>(acc, part);
}

template <typename R>
__global__ void __launch_bounds__(RED_THREADS)
k_norm_dot(long long n, long long stride, const int* active, const R* t, const R* r, double* part) {
  using T = typename V2<R>::T;
  if (rhs_off(active)) return;
  const T *Tt = RHS(t, const T), *Rr = RHS(r, const T);
  double acc[3] = {0, 0, 0};
  GRID_STRIDE(i, n) {
    T tv = Tt[i], rv = Rr[i];
    acc[0] += (double)tv.x * tv.x + (double)tv.y * tv.y;
    acc[1] += (double)tv.x * rv.x + (double)tv.y * rv.y;
    acc[2] += (double)tv.x * rv.y - (double)tv.y * rv.x;
  }
  block_reduce_store<3>(acc, part);
}

template <typename R>
__global__ void __launch_bounds__(RED_THREADS)
k_norm(long long n, long long stride, const int* active, const R* a, double* part) {
  using T = typename V2<R>::T;
  if (rhs_off(active)) return;
  const T* A = RHS(a, const T);
  double acc[1] = {0};
  GRID_STRIDE(i, n) {
    T av = A[i];
    acc[0] += (double)av.x * av.x + (double)av.y * av.y;
  }
  block_reduce_store<1>(acc, part);
}

// x += c v1 ; r -= c v2 ; |r|^2 [, <r0,r>]   (bicgstab_precdd_s.cc:281-295,427-440)
// c = alpha (coef 0) or omega (coef 1) of this RHS.  MODE (XR_*): XR_FULL as above;
// XR_R_ONLY leaves x alone (x and v1 unused); XR_DEFERRED first adds alpha v0 to x, the
// x update of the previous half-step, then c v1 (the same roundings as two passes).
template <typename R, bool WITH_R0, int MODE>
__global__ void __launch_bounds__(RED_THREADS)
k_update_xr(long long n, long long stride, const int* active, R* x, const R* v0, const R* v1, R* r, const R* v2,
            const Scalars* sc, int coef, const R* r0, double* part) {
  using T = typename V2<R>::T;
  if (rhs_off(active)) return;
  const Scalars& S = sc[blockIdx.y];
  const double* c = coef == 0 ? S.alpha : S.omega;
  const R cr = (R)c[0], ci = (R)c[1];
  const R ar = (R)S.alpha[0], ai = (R)S.alpha[1];
  T *X = MODE == XR_R_ONLY ? nullptr : RHS(x, T), *Rr = RHS(r, T);
  const T *A = MODE == XR_R_ONLY ? nullptr : RHS(v1, const T), *Qv = RHS(v2, const T);
  const T *P0 = MODE == XR_DEFERRED ? RHS(v0, const T) : nullptr, *Z = WITH_R0 ? RHS(r0, const T) : nullptr;
  double acc[WITH_R0 ? 3 : 1];
#pragma unroll
  for (int k = 0; k < (WITH_R0 ? 3 : 1); ++k) acc[k] = 0;
  GRID_STRIDE(i, n) {
    T rv = Rr[i], q = Qv[i];
    if constexpr (MODE != XR_R_ONLY) {
      T xv = X[i], a = A[i];
      if constexpr (MODE == XR_DEFERRED) {
        const T a0 = P0[i];
        xv.x = xv.x + ar * a0.x - ai * a0.y;
        xv.y = xv.y + ar * a0.y + ai * a0.x;
      }
      xv.x = xv.x + cr * a.x - ci * a.y;
      xv.y = xv.y + cr * a.y + ci * a.x;
      X[i] = xv;
    }
    rv.x = rv.x - cr * q.x + ci * q.y;
    rv.y = rv.y - cr * q.y - ci * q.x;
    Rr[i] = rv;
    acc[0] += (double)rv.x * rv.x + (double)rv.y * rv.y;
    if (WITH_R0) {
      T z = Z[i];
      acc[1] += (double)z.x * rv.x + (double)z.y * rv.y;
      acc[2] += (double)z.x * rv.y - (double)z.y * rv.x;
    }
  }
  block_reduce_store<WITH_R0 ? 3 : 1>(acc, part);
}

// x += alpha v0: the x update of a half-step whose residual met the tolerance, where
// that update is otherwise deferred to the next half-step (XR_DEFERRED)
template <typename R>
__global__ void __launch_bounds__(RED_THREADS)
k_update_x(long long n, long long stride, const int* active, R* x, const R* v0, const Scalars* sc) {
  using T = typename V2<R>::T;
  if (rhs_off(active)) return;
  const Scalars& S = sc[blockIdx.y];
  const R ar = (R)S.alpha[0], ai = (R)S.alpha[1];
  T* X = RHS(x, T);
  const T* P0 = RHS(v0, const T);
  GRID_STRIDE(i, n) {
    T xv = X[i];
    const T a0 = P0[i];
    xv.x = xv.x + ar * a0.x - ai * a0.y;
    xv.y = xv.y + ar * a0.y + ai * a0.x;
    X[i] = xv;
  }
}

// p = r + beta (p - omega q)   (bicgstab_precdd_s.cc:529-543)
template <typename R>
__global__ void __launch_bounds__(RED_THREADS)
k_update_p(long long n, long long stride, const int* active, R* p, const R* r, const R* q, const Scalars* sc) {
  using T = typename V2<R>::T;
  if (rhs_off(active)) return;
  const Scalars& S = sc[blockIdx.y];
  const R ar = (R)S.omega[0], ai = (R)S.omega[1], br = (R)S.beta[0], bi = (R)S.beta[1];
  T* P = RHS(p, T);
  const T *Qv = RHS(q, const T), *Rr = RHS(r, const T);
  GRID_STRIDE(i, n) {
    T pv = P[i];
    const T qv = Qv[i], rv = Rr[i];
    bicg_p_update<R>(pv.x, pv.y, qv.x, qv.y, rv.x, rv.y, ar, ai, br, bi);
    P[i] = pv;
  }
}

__device__ __forceinline__ void cdiv(double ar, double ai, double br, double bi, double* o) {
  // a / b with scaling (std::complex semantics up to rounding)
  if (fabs(br) >= fabs(bi)) {
    double t = bi / br, d = br + bi * t;
    o[0] = (ar + ai * t) / d;
    o[1] = (ai - ar * t) / d;
  } else {
    double t = br / bi, d = bi + br * t;
    o[0] = (ar * t + ai) / d;
    o[1] = (ai * t - ar) / d;
  }
}

// Deterministic sum of the partials of RHS blockIdx.x, then the scalar
// recurrences of BiCGStab with the reference's breakdown guards
// (bicgstab_dd_mix.cc:192,279,359).
__global__ void __launch_bounds__(RED_THREADS) k_finalize(const double* part, int nval, Scalars* scs, int mode,
                                                          const int* active) {
  const int r = blockIdx.x;
  if (active && !active[r]) return;
  __shared__ double sh[4][RED_THREADS];
  for (int v = 0; v < nval; ++v) {
    double x = 0;
    for (int i = threadIdx.x; i < RED_BLOCKS; i += RED_THREADS) x += part[((long long)r * 4 + v) * RED_BLOCKS + i];
    sh[v][threadIdx.x] = x;
  }
  __syncthreads();
  for (int w = RED_THREADS / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int v = 0; v < nval; ++v) sh[v][threadIdx.x] += sh[v][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  Scalars* sc = scs + r;
  double red[4] = {0, 0, 0, 0};
  for (int v = 0; v < nval; ++v) red[v] = sh[v][0];
  for (int v = 0; v < 4; ++v) sc->red[v] = red[v];
  switch (mode) {
    case FIN_INIT_INNER:
      sc->bnorm = red[0];
      sc->rnorm = red[0];
      sc->rho0[0] = red[0];
      sc->rho0[1] = 0;
      break;
    case FIN_INIT_OUTER:
      sc->bnorm = red[0];
      sc->rnorm = red[1];
      sc->rho0[0] = red[1];
      sc->rho0[1] = 0;
      break;
    case FIN_ALPHA:
      if (red[0] == 0 && red[1] == 0) { sc->alpha[0] = 0; sc->alpha[1] = 0; }
      else cdiv(sc->rho0[0], sc->rho0[1], red[0], red[1], sc->alpha);
      break;
    case FIN_RNORM:
      sc->rnorm = red[0];
      break;
    case FIN_OMEGA:
      sc->tt = red[0];
      if (red[0] == 0) { sc->omega[0] = 0; sc->omega[1] = 0; }
      else { sc->omega[0] = red[1] / red[0]; sc->omega[1] = red[2] / red[0]; }
      break;
    case FIN_RHO: {
      sc->rnorm = red[0];
      sc->rho[0] = red[1];
      sc->rho[1] = red[2];
      // beta = alpha * rho / (rho0 * omega)
      double nr = sc->alpha[0] * red[1] - sc->alpha[1] * red[2];
      double ni = sc->alpha[0] * red[2] + sc->alpha[1] * red[1];
      double dr = sc->rho0[0] * sc->omega[0] - sc->rho0[1] * sc->omega[1];
      double di = sc->rho0[0] * sc->omega[1] + sc->rho0[1] * sc->omega[0];
      double bt[2] = {0, 0};
      if (dr != 0 || di != 0) cdiv(nr, ni, dr, di, bt);
      sc->rho0[0] = red[1];
      sc->rho0[1] = red[2];
      bool om0 = (sc->omega[0] == 0 && sc->omega[1] == 0);
      bool rh0 = (red[1] == 0 && red[2] == 0);
      if (om0 || rh0) { bt[0] = 0; bt[1] = 0; }
      sc->beta[0] = bt[0];
      sc->beta[1] = bt[1];
      break;
    }
    case FIN_NRM:
      sc->nrm = sqrt(red[0]);
      break;
    default:
      break;
  }
}

// ---- per-site precision conversion (the two precisions have different layouts)
template <typename S>
__global__ void __launch_bounds__(RED_THREADS)
k_d2s_scaled(long long V, long long so, long long si, const int* active, S* out, const double* in, const Scalars* sc) {
  using R = typename Store<S>::R;
  if (rhs_off(active)) return;
  const double inv = 1.0 / sc[blockIdx.y].nrm;
  S* O = out + blockIdx.y * so;
  const double* I = in + blockIdx.y * si;
  GRID_STRIDE(s, V) {
    Cx<double> v[12];
    Cx<R> w[12];
    SpinorIO<double>::load(I, s, V, v);
#pragma unroll
    for (int k = 0; k < 12; ++k) w[k] = {(R)(v[k].re * inv), (R)(v[k].im * inv)};
    SpinorIO<S>::store(O, s, V, w);
  }
}

template <typename S>
__global__ void __launch_bounds__(RED_THREADS)
k_s2d_scaled(long long V, long long so, long long si, const int* active, double* out, const S* in, const Scalars* sc) {
  using R = typename Store<S>::R;
  if (rhs_off(active)) return;
  const double nrm = sc[blockIdx.y].nrm;
  double* O = out + blockIdx.y * so;
  const S* I = in + blockIdx.y * si;
  GRID_STRIDE(s, V) {
    Cx<R> w[12];
    Cx<double> v[12];
    SpinorIO<S>::load(I, s, V, w);
#pragma unroll
    for (int k = 0; k < 12; ++k) v[k] = {(double)w[k].re * nrm, (double)w[k].im * nrm};
    SpinorIO<double>::store(O, s, V, v);
  }
}

template <typename SO, typename SI>
__global__ void __launch_bounds__(RED_THREADS) k_convert(long long V, SO* out, const SI* in) {
  using RI = typename Store<SI>::R;
  using RO = typename Store<SO>::R;
  GRID_STRIDE(s, V) {
    Cx<RI> w[12];
    Cx<RO> v[12];
    SpinorIO<SI>::load(in, s, V, w);
#pragma unroll
    for (int k = 0; k < 12; ++k) v[k] = {(RO)w[k].re, (RO)w[k].im};
    SpinorIO<SO>::store(out, s, V, v);
  }
}

// elementwise real conversion of a field (the FP32 and FP16 field layouts share their order)
template <typename SO, typename SI>
__global__ void __launch_bounds__(RED_THREADS) k_convert_reals(long long n, SO* out, const SI* in) {
  GRID_STRIDE(i, n) out[i] = (SO)(float)in[i];
}

// Scaling between an FP32 vector and its FP16 copy with unit RMS component:
// sigma = sqrt(24 V / |in|^2) from sc[r].red[0] = |in_r|^2 (sigma = 0 for a zero
// vector).  dir > 0: out = sigma * in ; dir < 0: out = in / sigma.
template <typename SO, typename SI>
__global__ void __launch_bounds__(RED_THREADS)
k_scale_rms(long long V, long long so, long long si, const int* active, SO* out, const SI* in, const Scalars* sc,
            int dir) {
  if (rhs_off(active)) return;
  const double n2 = sc[blockIdx.y].red[0];
  const double sigma = n2 > 0 ? sqrt(24.0 * (double)V / n2) : 0.0;
  const float f = dir > 0 ? (float)sigma : (sigma > 0 ? (float)(1.0 / sigma) : 0.0f);
  SO* O = out + blockIdx.y * so;
  const SI* I = in + blockIdx.y * si;
  GRID_STRIDE(s, V) {
    Cx<float> v[12];
    SpinorIO<SI>::load(I, s, V, v);
#pragma unroll
    for (int k = 0; k < 12; ++k) v[k] = {v[k].re * f, v[k].im * f};
    SpinorIO<SO>::store(O, s, V, v);
  }
}

// Order-independent 64-bit digest of n 8-byte words: XOR over i of a SplitMix64
// finalisation of (w[i] + (i + 1) * golden ratio).  Equal for bit-identical data
// whatever the scheduling; used to compare solutions without copying them out.
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
__global__ void __launch_bounds__(RED_THREADS) k_digest(long long n, const unsigned long long* w,
                                                        unsigned long long* out) {
  unsigned long long h = 0;
  GRID_STRIDE(i, n) h ^= mix64(w[i] + 0x9E3779B97F4A7C15ULL * (unsigned long long)(i + 1));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) h ^= __shfl_xor(h, off, 64);
  if ((threadIdx.x & 63) == 0) atomicXor(out, h);
}

template <typename R>
__global__ void __launch_bounds__(RED_THREADS) k_sub(long long n, R* out, const R* a, const R* b) {
  GRID_STRIDE(i, n) out[i] = a[i] - b[i];
}

// y = a x + y as one FMA (mode 0), or y = b y + a x with the product a x rounded first
// (mode 1): the two update forms of QWS's CG solvers (fmadd_d / fnmadd_d, cg.cc, mcg.cc)
template <typename R>
__global__ void __launch_bounds__(RED_THREADS) k_lincomb(long long n, R a, const R* x, R b, R* y, int mode) {
  if (mode == 0) {
    GRID_STRIDE(i, n) y[i] = fma(a, x[i], y[i]);
  } else {
    GRID_STRIDE(i, n) y[i] = fma(b, y[i], a * x[i]);
  }
}

// ------------------------------------------------------------------ launchers
#define LAUNCH_B(kern, nb, ...) \
  hipLaunchKernelGGL(kern, dim3(RED_BLOCKS, (nb)), dim3(RED_THREADS), 0, st, __VA_ARGS__)

template <typename R>
void launch_bicg_init_inner(const Batch& B, long long n, R* x, R* r, R* p, R* r0, const R* b, double* part,
                            hipStream_t st) {
  LAUNCH_B((k_init_inner<R>), B.n, n, B.stride, B.active, x, r, p, r0, b, part);
}
template <typename R>
void launch_bicg_init_outer(const Batch& B, long long n, R* r, R* p, R* r0, const R* b, const R* q, double* part,
                            hipStream_t st) {
  LAUNCH_B((k_init_outer<R>), B.n, n, B.stride, B.active, r, p, r0, b, q, part);
}
template <typename R>
void launch_dot(const Batch& B, long long n, const R* a, const R* b, double* part, hipStream_t st) {
  LAUNCH_B((k_dot<R>), B.n, n, B.stride, B.active, a, b, part);
}
template <typename R>
void launch_norm_dot(const Batch& B, long long n, const R* t, const R* r, double* part, hipStream_t st) {
  LAUNCH_B((k_norm_dot<R>), B.n, n, B.stride, B.active, t, r, part);
}
template <typename R>
void launch_update_xr(const Batch& B, long long n, R* x, const R* v1, R* r, const R* v2, const Scalars* sc, int coef,
                      const R* r0, double* part, hipStream_t st, int mode, const R* v0) {
  if (mode == XR_DEFERRED && (!r0 || !v0)) throw std::invalid_argument("update_xr: the deferred update needs r0 and v0");
  if (mode == XR_R_ONLY && r0) throw std::invalid_argument("update_xr: the r-only update takes no r0");
  if (mode == XR_DEFERRED)
    LAUNCH_B((k_update_xr<R, true, XR_DEFERRED>), B.n, n, B.stride, B.active, x, v0, v1, r, v2, sc, coef, r0, part);
  else if (mode == XR_R_ONLY)
    LAUNCH_B((k_update_xr<R, false, XR_R_ONLY>), B.n, n, B.stride, B.active, x, v0, v1, r, v2, sc, coef, r0, part);
  else if (r0)
    LAUNCH_B((k_update_xr<R, true, XR_FULL>), B.n, n, B.stride, B.active, x, v0, v1, r, v2, sc, coef, r0, part);
  else LAUNCH_B((k_update_xr<R, false, XR_FULL>), B.n, n, B.stride, B.active, x, v0, v1, r, v2, sc, coef, r0, part);
}
template <typename R>
void launch_update_x(const Batch& B, long long n, R* x, const R* v0, const Scalars* sc, hipStream_t st) {
  LAUNCH_B((k_update_x<R>), B.n, n, B.stride, B.active, x, v0, sc);
}
template <typename R>
void launch_update_p(const Batch& B, long long n, R* p, const R* r, const R* q, const Scalars* sc, hipStream_t st) {
  LAUNCH_B((k_update_p<R>), B.n, n, B.stride, B.active, p, r, q, sc);
}
template <typename R>
void launch_norm(const Batch& B, long long n, const R* a, double* part, hipStream_t st) {
  LAUNCH_B((k_norm<R>), B.n, n, B.stride, B.active, a, part);
}
void launch_finalize(const Batch& B, const double* part, int nval, Scalars* sc, int mode, hipStream_t st) {
  hipLaunchKernelGGL(k_finalize, dim3(B.n), dim3(RED_THREADS), 0, st, part, nval, sc, mode, B.active);
}
template <typename S>
void launch_d2s_scaled(const Batch& B, long long V, long long so, long long si, S* out, const double* in,
                       const Scalars* sc, hipStream_t st) {
  LAUNCH_B((k_d2s_scaled<S>), B.n, V, so, si, B.active, out, in, sc);
}
template <typename S>
void launch_s2d_scaled(const Batch& B, long long V, long long so, long long si, double* out, const S* in,
                       const Scalars* sc, hipStream_t st) {
  LAUNCH_B((k_s2d_scaled<S>), B.n, V, so, si, B.active, out, in, sc);
}
template <typename SO, typename SI> void launch_convert(long long V, SO* out, const SI* in, hipStream_t st) {
  LAUNCH_B((k_convert<SO, SI>), 1, V, out, in);
}
template <typename SO, typename SI> void launch_convert_reals(long long n, SO* out, const SI* in, hipStream_t st) {
  LAUNCH_B((k_convert_reals<SO, SI>), 1, n, out, in);
}
template <typename SO, typename SI>
void launch_scale_rms(const Batch& B, long long V, long long so, long long si, SO* out, const SI* in, const Scalars* sc,
                      int dir, hipStream_t st) {
  LAUNCH_B((k_scale_rms<SO, SI>), B.n, V, so, si, B.active, out, in, sc, dir);
}
void launch_digest(long long nwords, const void* data, unsigned long long* out, hipStream_t st) {
  LAUNCH_B(k_digest, 1, nwords, (const unsigned long long*)data, out);
}
template <typename R> void launch_sub(long long n, R* out, const R* a, const R* b, hipStream_t st) {
  LAUNCH_B((k_sub<R>), 1, n, out, a, b);
}
template <typename R> void launch_lincomb(long long n, R a, const R* x, R b, R* y, int mode, hipStream_t st) {
  LAUNCH_B((k_lincomb<R>), 1, n, a, x, b, y, mode);
}

// HBM probe: out = the sum of nin input arrays of n4 float4 each (nin reads per write;
// nin = 1 is a copy), coalesced 128-bit accesses.  Each thread handles U elements a
// grid apart, with all U x nin loads issued before the sums; grid-stride beyond that.
template <int U>
__global__ void __launch_bounds__(256) k_stream(long long n4, int nin, const float4* __restrict__ in,
                                                float4* __restrict__ out) {
  const long long step = (long long)gridDim.x * 256;
  for (long long i0 = blockIdx.x * 256LL + threadIdx.x; i0 < n4; i0 += step * U) {
    float4 s[U];
#pragma unroll
    for (int u = 0; u < U; ++u) s[u] = i0 + u * step < n4 ? in[i0 + u * step] : make_float4(0, 0, 0, 0);
    for (int k = 1; k < nin; ++k) {
      float4 v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) v[u] = i0 + u * step < n4 ? in[k * n4 + i0 + u * step] : make_float4(0, 0, 0, 0);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        s[u].x += v[u].x;
        s[u].y += v[u].y;
        s[u].z += v[u].z;
        s[u].w += v[u].w;
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (i0 + u * step < n4) out[i0 + u * step] = s[u];
  }
}

// One input stream: output chunk c (256 float4) is the sum of the nin consecutive
// 256-float4 chunks c*nin .. c*nin+nin-1 of `in`, so that every workgroup reads one
// contiguous region (a read-mostly kernel over a single array).
__global__ void __launch_bounds__(256) k_stream_one(long long nchunk, int nin, const float4* __restrict__ in,
                                                    float4* __restrict__ out) {
  for (long long c = blockIdx.x; c < nchunk; c += gridDim.x) {
    const float4* p = in + c * nin * 256 + threadIdx.x;
    float4 s = p[0];
    for (int k = 1; k < nin; ++k) {
      const float4 v = p[k * 256];
      s.x += v.x;
      s.y += v.y;
      s.z += v.z;
      s.w += v.w;
    }
    out[c * 256 + threadIdx.x] = s;
  }
}

// cfg 0: one element per thread; 1: 4 per thread; 2: grid-stride over 2048 workgroups;
// 3, 4: one input stream (k_stream_one; n4 a multiple of 256), one workgroup per output
// chunk or grid-stride over 2048 workgroups
void launch_stream(int cfg, long long n4, int nin, const float4* in, float4* out, hipStream_t st) {
  if (cfg >= 3) {
    const long long nchunk = n4 / 256;
    const unsigned grid = cfg == 3 ? (unsigned)nchunk : 2048u;
    hipLaunchKernelGGL(k_stream_one, dim3(grid), dim3(256), 0, st, nchunk, nin, in, out);
  } else if (cfg == 0) {
    hipLaunchKernelGGL(k_stream<1>, dim3((unsigned)((n4 + 255) / 256)), dim3(256), 0, st, n4, nin, in, out);
  } else if (cfg == 1) {
    hipLaunchKernelGGL(k_stream<4>, dim3((unsigned)((n4 + 1023) / 1024)), dim3(256), 0, st, n4, nin, in, out);
  } else {
    hipLaunchKernelGGL(k_stream<1>, dim3(2048), dim3(256), 0, st, n4, nin, in, out);
  }
}

#define INST(R)                                                                                                       \
  template void launch_bicg_init_inner<R>(const Batch&, long long, R*, R*, R*, R*, const R*, double*, hipStream_t);   \
  template void launch_bicg_init_outer<R>(const Batch&, long long, R*, R*, R*, const R*, const R*, double*,           \
                                          hipStream_t);                                                               \
  template void launch_dot<R>(const Batch&, long long, const R*, const R*, double*, hipStream_t);                     \
  template void launch_norm_dot<R>(const Batch&, long long, const R*, const R*, double*, hipStream_t);                \
  template void launch_update_xr<R>(const Batch&, long long, R*, const R*, R*, const R*, const Scalars*, int,         \
                                    const R*, double*, hipStream_t, int, const R*);                                   \
  template void launch_update_x<R>(const Batch&, long long, R*, const R*, const Scalars*, hipStream_t);              \
  template void launch_update_p<R>(const Batch&, long long, R*, const R*, const R*, const Scalars*, hipStream_t);     \
  template void launch_norm<R>(const Batch&, long long, const R*, double*, hipStream_t);                              \
  template void launch_sub<R>(long long, R*, const R*, const R*, hipStream_t);                                      \
  template void launch_lincomb<R>(long long, R, const R*, R, R*, int, hipStream_t);
INST(float)
INST(double)
template void launch_d2s_scaled<float>(const Batch&, long long, long long, long long, float*, const double*,
                                       const Scalars*, hipStream_t);
template void launch_d2s_scaled<__half>(const Batch&, long long, long long, long long, __half*, const double*,
                                        const Scalars*, hipStream_t);
template void launch_s2d_scaled<float>(const Batch&, long long, long long, long long, double*, const float*,
                                       const Scalars*, hipStream_t);
template void launch_s2d_scaled<__half>(const Batch&, long long, long long, long long, double*, const __half*,
                                        const Scalars*, hipStream_t);
template void launch_convert<float, double>(long long, float*, const double*, hipStream_t);
template void launch_convert<double, float>(long long, double*, const float*, hipStream_t);
template void launch_convert<__half, float>(long long, __half*, const float*, hipStream_t);
template void launch_convert<float, __half>(long long, float*, const __half*, hipStream_t);
template void launch_convert_reals<__half, float>(long long, __half*, const float*, hipStream_t);
template void launch_scale_rms<__half, float>(const Batch&, long long, long long, long long, __half*, const float*,
                                              const Scalars*, int, hipStream_t);
template void launch_scale_rms<float, __half>(const Batch&, long long, long long, long long, float*, const __half*,
                                              const Scalars*, int, hipStream_t);

}  // namespace qws
