// sap.hip -- fused SAP half-sweep: one workgroup per SAP block of one colour.
//
// One launch does, for every block of colour c (multiplicative Schwarz step of
// qws.cc:809-841, with the face update of the previous half-sweep moved in front):
//   prologue   s_c += kappa T^-1 H_face q            (ddd_out_pre_s_/ddd_out_pos_s_)
//   block solve x = A_c s_c: x_0 = s_c, x_{k+1} = s_c + kappa T^-1 H_in x_k, nm times
//              (jinv_ddd_in_s_, static_solver.cc:65-140)
//   epilogue   s_c += in_c - x + kappa T^-1 H_in x,  q_c = x   (qws.cc:814-823)
//              or the final sequences of prec_ddd_s_ / prec_s_ (qws.cc:842-925, 967-972)
// The block's forward links and the iterate x live in LDS, the site's clover
// inverse in registers, so links and clover cross HBM once per half-sweep instead
// of once per operator application.  Requires proper colouring (face neighbours
// belong to the other colour) and Vb == SAP_THREADS sites per block.  Instantiated
// for FP32 storage and for FP16 storage with FP32 arithmetic.
#include <type_traits>

#include "kernels.h"
#include "sap.h"

namespace qws {

// LDS element type: the storage type (FP16 links and iterate stay FP16 in LDS)
__device__ __forceinline__ float lds_get(const float* p, int i) { return p[i]; }
__device__ __forceinline__ float lds_get(const __half* p, int i) { return __half2float(p[i]); }
__device__ __forceinline__ void lds_put(float* p, int i, float v) { p[i] = v; }
__device__ __forceinline__ void lds_put(__half* p, int i, float v) { p[i] = __float2half_rn(v); }
// value as stored: identity for FP32, binary16 rounding for FP16
__device__ __forceinline__ float stored(float v, const float*) { return v; }
__device__ __forceinline__ float stored(float v, const __half*) { return __half2float(__float2half_rn(v)); }

template <typename S, int NL>
__device__ __forceinline__ void load_link_lds(const S* lU, int mu, int t, Cx<float> u[9]) {
  constexpr int NC = NL / 2;
#pragma unroll
  for (int k = 0; k < NC; ++k)
    u[k] = {lds_get(lU, ((mu * NC + k) * 2 + 0) * SAP_THREADS + t), lds_get(lU, ((mu * NC + k) * 2 + 1) * SAP_THREADS + t)};
  if (NL == 12) {
    Cx<float> a = cmul(u[1], u[5]), b = cmul(u[2], u[4]);
    u[6] = cconj(csub(a, b));
    a = cmul(u[2], u[3]); b = cmul(u[0], u[5]);
    u[7] = cconj(csub(a, b));
    a = cmul(u[0], u[4]); b = cmul(u[1], u[3]);
    u[8] = cconj(csub(a, b));
  }
}

template <typename S>
__device__ __forceinline__ void lds_load_spinor(const S* lx, int t, Cx<float> v[12]) {
#pragma unroll
  for (int k = 0; k < 12; ++k) v[k] = {lds_get(lx, (2 * k) * SAP_THREADS + t), lds_get(lx, (2 * k + 1) * SAP_THREADS + t)};
}
template <typename S>
__device__ __forceinline__ void lds_store_spinor(S* lx, int t, const Cx<float> v[12]) {
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    lds_put(lx, (2 * k) * SAP_THREADS + t, v[k].re);
    lds_put(lx, (2 * k + 1) * SAP_THREADS + t, v[k].im);
  }
}

// in-block hop along MU, DIR from LDS
template <typename S, int NL, int MU, int DIR>
__device__ __forceinline__ void lds_hop(const Geo& g, const SiteC& c, int t, const S* lx, const S* lU,
                                        Cx<float> acc[12]) {
  const int l = c.l[MU] + DIR;
  if (l < 0 || l >= g.B[MU]) return;
  const int tn = t + DIR * g.sB[MU];
  Cx<float> v[12], a[6], b[6], u[9];
  lds_load_spinor(lx, tn, v);
  project<float, MU, DIR>(v, a);
  load_link_lds<S, NL>(lU, MU, DIR > 0 ? t : tn, u);
  link_mult<float, DIR>(u, a, b);
  reconstruct<float, MU, DIR>(b, 2.0f, acc);
}

template <typename S, int NL>
__device__ __forceinline__ void lds_hop_sum(const Geo& g, const SiteC& c, int t, const S* lx, const S* lU,
                                            Cx<float> acc[12]) {
#pragma unroll
  for (int k = 0; k < 12; ++k) acc[k] = {0, 0};
  lds_hop<S, NL, 0, +1>(g, c, t, lx, lU, acc);
  lds_hop<S, NL, 0, -1>(g, c, t, lx, lU, acc);
  lds_hop<S, NL, 1, +1>(g, c, t, lx, lU, acc);
  lds_hop<S, NL, 1, -1>(g, c, t, lx, lU, acc);
  lds_hop<S, NL, 2, +1>(g, c, t, lx, lU, acc);
  lds_hop<S, NL, 2, -1>(g, c, t, lx, lU, acc);
  lds_hop<S, NL, 3, +1>(g, c, t, lx, lU, acc);
  lds_hop<S, NL, 3, -1>(g, c, t, lx, lU, acc);
}

// y = A x for the packed Hermitian 6x6 whose reals k sit in a(k)
template <typename A>
__device__ __forceinline__ void herm6_mult_f(A a, const Cx<float> x[6], Cx<float> y[6]) {
#pragma unroll
  for (int r = 0; r < 6; ++r) y[r] = cscale(a(r), x[r]);
  int p = 6;
#pragma unroll
  for (int r = 0; r < 6; ++r)
#pragma unroll
    for (int c = r + 1; c < 6; ++c) {
      Cx<float> m = {a(p), a(p + 1)};
      y[r] = cadd(y[r], cmul(m, x[c]));
      y[c] = cadd(y[c], cmulc(m, x[r]));
      p += 2;
    }
}

// The site's clover inverse held in registers for the whole launch: 72 FP32 reals,
// or for FP16 storage 36 packed pairs (converted to FP32 at each use).
template <typename S> struct CloverRegs;
template <> struct CloverRegs<float> {
  float c[72];
  __device__ __forceinline__ void load(const float* C, long long s, long long V) {
    CloverIO<float>::load_chir(C, s, V, 0, c);
    CloverIO<float>::load_chir(C, s, V, 1, c + 36);
  }
  __device__ __forceinline__ float get(int k) const { return c[k]; }
};
template <> struct CloverRegs<__half> {
  __half2 c[36];
  __device__ __forceinline__ void load(const __half* C, long long s, long long V) {
    const Half4* p = reinterpret_cast<const Half4*>(C);
#pragma unroll
    for (int f = 0; f < 18; ++f) {
      const Half4 h = p[(long long)f * V + s];
      c[2 * f] = h.a;
      c[2 * f + 1] = h.b;
    }
  }
  __device__ __forceinline__ float get(int k) const {
    const float2 f = __half22float2(c[k >> 1]);
    return (k & 1) ? f.y : f.x;
  }
};

template <typename S>
__device__ __forceinline__ void clover_regs(const CloverRegs<S>& cl, Cx<float> v[12]) {
  Cx<float> x[6], y0[6], y1[6];
#pragma unroll
  for (int c = 0; c < 3; ++c) { x[c] = cadd(v[c * 4 + 0], v[c * 4 + 2]); x[3 + c] = cadd(v[c * 4 + 1], v[c * 4 + 3]); }
  herm6_mult_f([&](int k) { return cl.get(k); }, x, y0);
#pragma unroll
  for (int c = 0; c < 3; ++c) { x[c] = csub(v[c * 4 + 0], v[c * 4 + 2]); x[3 + c] = csub(v[c * 4 + 1], v[c * 4 + 3]); }
  herm6_mult_f([&](int k) { return cl.get(36 + k); }, x, y1);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    v[c * 4 + 0] = cadd(y0[c], y1[c]);
    v[c * 4 + 2] = csub(y0[c], y1[c]);
    v[c * 4 + 1] = cadd(y0[3 + c], y1[3 + c]);
    v[c * 4 + 3] = csub(y0[3 + c], y1[3 + c]);
  }
}

// Storage S = float: everything FP32.  S = __half: spinors, links, clover and the
// LDS copies in FP16, arithmetic in FP32, every stored value rounded to FP16 where
// the unfused FP16 path stores it (face-updated s, each Jacobi iterate, s, q).
template <typename S, int NL>
__global__ void __launch_bounds__(SAP_THREADS, 1)
k_sap_block(Geo g, const int* __restrict__ list, int nrhs, long long stride, const int* __restrict__ active, float kappa,
            int nm, int flags, S* s, const S* __restrict__ in, const S* qface, S* qout, S* out,
            const S* __restrict__ U, const S* __restrict__ C) {
  constexpr int NC = NL / 2;
  __shared__ S lU[4 * NC * 2 * SAP_THREADS];
  __shared__ S lx[24 * SAP_THREADS];
  const int t = threadIdx.x;
  const int blk = list[xcd_remap(g, blockIdx.x, gridDim.x)];
  const long long site = (long long)blk * SAP_THREADS + t;
  SiteC c;
  site_coords(g, site, c);

  // forward links of this site into LDS (coalesced [mu][k][V] reads), once for all RHS
#pragma unroll
  for (int mu = 0; mu < 4; ++mu)
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      const Cx<float> v = load_cx<S>(U, ((long long)mu * NC + k) * g.V + site);
      lds_put(lU, ((mu * NC + k) * 2 + 0) * SAP_THREADS + t, v.re);
      lds_put(lU, ((mu * NC + k) * 2 + 1) * SAP_THREADS + t, v.im);
    }
  CloverRegs<S> cl;
  cl.load(C, site, g.V);
  const int epi = flags & SAP_EPI_MASK;

  for (int r = 0; r < nrhs; ++r) {
    if (active && !active[r]) continue;  // uniform over the workgroup
    const long long off = r * stride;
    Cx<float> sv[12], acc[12];
    SpinorIO<S>::load(((flags & SAP_SRC_IN) ? in : s) + off, site, g.V, sv);

    if (flags & SAP_FACE_PRO) {
      Cx<float> f[12];
      hop_sum<S, NL, HOPS_FACE>(g, site, c, qface + off, opaque(U), f);
      clover_regs(cl, f);
#pragma unroll
      for (int k = 0; k < 12; ++k) sv[k] = {stored(sv[k].re + kappa * f[k].re, lx), stored(sv[k].im + kappa * f[k].im, lx)};
    }

    // block solve: x_0 = s_c ; x_{k+1} = s_c + kappa T^-1 H_in x_k
    Cx<float> x[12];
    __syncthreads();  // the previous RHS's readers of lx are done (and lU is complete)
    lds_store_spinor(lx, t, sv);
    __syncthreads();
    for (int it = 0; it < nm; ++it) {
      lds_hop_sum<S, NL>(g, c, t, lx, lU, acc);
      clover_regs(cl, acc);
#pragma unroll
      for (int k = 0; k < 12; ++k) x[k] = cadd(sv[k], cscale(kappa, acc[k]));
      __syncthreads();
      lds_store_spinor(lx, t, x);
      __syncthreads();
    }
    lds_load_spinor(lx, t, x);  // x as stored
    SpinorIO<S>::store(qout + off, site, g.V, x);

    if (epi == SAP_EPI_NONE) continue;
    // acc = T^-1 H_in x
    lds_hop_sum<S, NL>(g, c, t, lx, lU, acc);
    clover_regs(cl, acc);
    Cx<float> o[12];
    if (epi == SAP_EPI_SUPDATE) {
      // s_c = s_c + in_c - x + kappa T^-1 H_in x
      Cx<float> iv[12];
      SpinorIO<S>::load(in + off, site, g.V, iv);
#pragma unroll
      for (int k = 0; k < 12; ++k) o[k] = cadd(csub(cadd(sv[k], iv[k]), x[k]), cscale(kappa, acc[k]));
      SpinorIO<S>::store(s + off, site, g.V, o);
    } else {
      // out_c = [- kappa T^-1 H_face q] + x - kappa T^-1 H_in x; the face term is
      // s_c - sv (the prologue added kappa T^-1 H_face q to s_c)
#pragma unroll
      for (int k = 0; k < 12; ++k) o[k] = csub(x[k], cscale(kappa, acc[k]));
      if (epi == SAP_EPI_OUT_FACE) {
        Cx<float> s0[12];
        SpinorIO<S>::load(((flags & SAP_SRC_IN) ? in : s) + off, site, g.V, s0);
#pragma unroll
        for (int k = 0; k < 12; ++k) o[k] = cadd(o[k], csub(s0[k], sv[k]));
      }
      SpinorIO<S>::store(out + off, site, g.V, o);
    }
  }
}

// ---------------------------------------------------------------- FP16 storage kernel
// FP16 operands with FP32 accumulation: the link and clover products use
// v_dot2_f32_f16 (two FP16 products summed into an FP32 accumulator), the spin
// projections packed FP16 adds, the hop sums, the clover results and the iterate
// update FP32.  The block's links sit in LDS as 18-real FP16 (row 2 rebuilt once at
// staging), the iterate as FP16.
typedef _Float16 h2v __attribute__((ext_vector_type(2)));

__device__ __forceinline__ h2v swp(h2v a) { return __builtin_shufflevector(a, a, 1, 0); }
__device__ __forceinline__ h2v neg_hi(h2v a) { return h2v{a.x, -a.y}; }
// c + a.x b.x + a.y b.y (v_dot2_f32_f16).  Only the compiler's own dot instructions
// are used, so that its hazard checks see every dependent pair.
__device__ __forceinline__ float dot2(h2v a, h2v b, float c) { return __builtin_amdgcn_fdot2(a, b, c, false); }
__device__ __forceinline__ h2v to_h2(Cx<float> v) { return h2v{(_Float16)v.re, (_Float16)v.im}; }
__device__ __forceinline__ Cx<float> from_h2(h2v v) { return {(float)v.x, (float)v.y}; }

// FP16 spin projection of the hop along MU, DIR (the table of project<>, packed)
template <int MU, int DIR>
__device__ __forceinline__ void project_h2(const h2v v[12], h2v a[6]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const h2v s0 = v[c * 4 + 0], s1 = v[c * 4 + 1], s2 = v[c * 4 + 2], s3 = v[c * 4 + 3];
    h2v a0, a1;
    // i z = (-z.y, z.x), -i z = (z.y, -z.x)
    if (MU == 0) {
      if (DIR > 0) { a0 = s0 + h2v{s3.y, -s3.x}; a1 = s1 + h2v{s2.y, -s2.x}; }
      else         { a0 = s0 + h2v{-s3.y, s3.x}; a1 = s1 + h2v{-s2.y, s2.x}; }
    } else if (MU == 1) {
      if (DIR > 0) { a0 = s0 - s3; a1 = s1 + s2; } else { a0 = s0 + s3; a1 = s1 - s2; }
    } else if (MU == 2) {
      if (DIR > 0) { a0 = s0 + h2v{s2.y, -s2.x}; a1 = s1 + h2v{-s3.y, s3.x}; }
      else         { a0 = s0 + h2v{-s2.y, s2.x}; a1 = s1 + h2v{s3.y, -s3.x}; }
    } else {
      if (DIR > 0) { a0 = s2; a1 = s3; } else { a0 = s0; a1 = s1; }
    }
    a[c * 2 + 0] = a0;
    a[c * 2 + 1] = a1;
  }
}

// b = W a (forward) or W^+ a (backward), W[i][j] = c[j][i] as in link_mult<>.
// With A = (a.x, a.y): u a = (u.A', u.As), A' = (a.x, -a.y), As = (a.y, a.x);
// conj(u) a = (u.A, u.As'), As' = (a.y, -a.x).
template <int DIR>
__device__ __forceinline__ void link_mult_h2(const h2v w[9], const h2v a[6], Cx<float> b[6]) {
  h2v p[6], q[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    p[k] = DIR > 0 ? neg_hi(a[k]) : a[k];
    q[k] = DIR > 0 ? swp(a[k]) : neg_hi(swp(a[k]));
  }
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      float re = 0.f, im = 0.f;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const h2v u = DIR > 0 ? w[i * 3 + j] : w[j * 3 + i];
        re = dot2(u, p[i * 2 + h], re);
        im = dot2(u, q[i * 2 + h], im);
      }
      b[j * 2 + h] = {re, im};
    }
}

template <int MU, int DIR>
__device__ __forceinline__ void lds_hop_h2(const Geo& g, const SiteC& c, int t, const h2v* lx, const h2v* lU,
                                           Cx<float> acc[12]) {
  const int l = c.l[MU] + DIR;
  if (l < 0 || l >= g.B[MU]) return;
  const int tn = t + DIR * g.sB[MU];
  h2v v[12], a[6], w[9];
  Cx<float> b[6];
#pragma unroll
  for (int k = 0; k < 12; ++k) v[k] = lx[k * SAP_THREADS + tn];
  project_h2<MU, DIR>(v, a);
  const int tl = DIR > 0 ? t : tn;
#pragma unroll
  for (int k = 0; k < 9; ++k) w[k] = lU[(MU * 9 + k) * SAP_THREADS + tl];
  link_mult_h2<DIR>(w, a, b);
  reconstruct<float, MU, DIR>(b, 2.0f, acc);
}

__device__ __forceinline__ void lds_hop_sum_h2(const Geo& g, const SiteC& c, int t, const h2v* lx, const h2v* lU,
                                               Cx<float> acc[12]) {
#pragma unroll
  for (int k = 0; k < 12; ++k) acc[k] = {0, 0};
  lds_hop_h2<0, +1>(g, c, t, lx, lU, acc);
  lds_hop_h2<0, -1>(g, c, t, lx, lU, acc);
  lds_hop_h2<1, +1>(g, c, t, lx, lU, acc);
  lds_hop_h2<1, -1>(g, c, t, lx, lU, acc);
  lds_hop_h2<2, +1>(g, c, t, lx, lU, acc);
  lds_hop_h2<2, -1>(g, c, t, lx, lU, acc);
  lds_hop_h2<3, +1>(g, c, t, lx, lU, acc);
  lds_hop_h2<3, -1>(g, c, t, lx, lU, acc);
}

// y = A x for one chirality: diagonal in FP32, off-diagonal entries (FP16 pairs
// cl[3..17] of the chirality) by dot2 on the FP16-rounded x
__device__ __forceinline__ void herm6_h2(const h2v* cl, const Cx<float> x[6], Cx<float> y[6]) {
  h2v X[6], Xn[6], XS[6], XSn[6];
#pragma unroll
  for (int r = 0; r < 6; ++r) {
    X[r] = to_h2(x[r]);
    Xn[r] = neg_hi(X[r]);
    XS[r] = swp(X[r]);
    XSn[r] = neg_hi(XS[r]);
  }
#pragma unroll
  for (int r = 0; r < 6; ++r) {
    const float d = (float)((r & 1) ? cl[r >> 1].y : cl[r >> 1].x);
    y[r] = cscale(d, x[r]);
  }
  int p = 3;
#pragma unroll
  for (int r = 0; r < 6; ++r)
#pragma unroll
    for (int c = r + 1; c < 6; ++c) {
      const h2v m = cl[p++];
      y[r].re = dot2(m, Xn[c], y[r].re);  // m x_c
      y[r].im = dot2(m, XS[c], y[r].im);
      y[c].re = dot2(m, X[r], y[c].re);   // conj(m) x_r
      y[c].im = dot2(m, XSn[r], y[c].im);
    }
}

__device__ __forceinline__ void clover_h2(const h2v cl[36], Cx<float> v[12]) {
  Cx<float> x[6], y0[6], y1[6];
#pragma unroll
  for (int c = 0; c < 3; ++c) { x[c] = cadd(v[c * 4 + 0], v[c * 4 + 2]); x[3 + c] = cadd(v[c * 4 + 1], v[c * 4 + 3]); }
  herm6_h2(cl, x, y0);
#pragma unroll
  for (int c = 0; c < 3; ++c) { x[c] = csub(v[c * 4 + 0], v[c * 4 + 2]); x[3 + c] = csub(v[c * 4 + 1], v[c * 4 + 3]); }
  herm6_h2(cl + 18, x, y1);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    v[c * 4 + 0] = cadd(y0[c], y1[c]);
    v[c * 4 + 2] = csub(y0[c], y1[c]);
    v[c * 4 + 1] = cadd(y0[3 + c], y1[3 + c]);
    v[c * 4 + 3] = csub(y0[3 + c], y1[3 + c]);
  }
}

__device__ __forceinline__ void lds_store_h2(h2v* lx, int t, const Cx<float> v[12]) {
#pragma unroll
  for (int k = 0; k < 12; ++k) lx[k * SAP_THREADS + t] = to_h2(v[k]);
}
__device__ __forceinline__ void lds_load_h2(const h2v* lx, int t, Cx<float> v[12]) {
#pragma unroll
  for (int k = 0; k < 12; ++k) v[k] = from_h2(lx[k * SAP_THREADS + t]);
}

template <int NL>
__global__ void __launch_bounds__(SAP_THREADS, 2)
k_sap_block_h(Geo g, const int* __restrict__ list, int nrhs, long long stride, const int* __restrict__ active,
              float kappa, int nm, int flags, __half* s, const __half* __restrict__ in, const __half* qface,
              __half* qout, __half* out, const __half* __restrict__ U, const __half* __restrict__ C) {
  constexpr int NC = NL / 2;
  __shared__ h2v lU[4 * 9 * SAP_THREADS];
  __shared__ h2v lx[12 * SAP_THREADS];
  const int t = threadIdx.x;
  const int blk = list[xcd_remap(g, blockIdx.x, gridDim.x)];
  const long long site = (long long)blk * SAP_THREADS + t;
  SiteC c;
  site_coords(g, site, c);
#pragma unroll
  for (int mu = 0; mu < 4; ++mu) {
    Cx<float> u[9];
    load_link<__half, NL>(U, mu, site, g.V, u);  // row 2 rebuilt in FP32 for NL = 12
#pragma unroll
    for (int k = 0; k < 9; ++k) lU[(mu * 9 + k) * SAP_THREADS + t] = to_h2(u[k]);
  }
  h2v cl[36];
  {
    const Half4* p = reinterpret_cast<const Half4*>(C);
#pragma unroll
    for (int f = 0; f < 18; ++f) {
      const Half4 h = p[(long long)f * g.V + site];
      cl[2 * f] = __builtin_bit_cast(h2v, h.a);
      cl[2 * f + 1] = __builtin_bit_cast(h2v, h.b);
    }
  }
  const int epi = flags & SAP_EPI_MASK;

  for (int r = 0; r < nrhs; ++r) {
    if (active && !active[r]) continue;
    const long long off = r * stride;
    Cx<float> sv[12], acc[12];
    SpinorIO<__half>::load(((flags & SAP_SRC_IN) ? in : s) + off, site, g.V, sv);
    if (flags & SAP_FACE_PRO) {
      Cx<float> f[12];
      hop_sum<__half, NL, HOPS_FACE>(g, site, c, qface + off, opaque(U), f);
      clover_h2(cl, f);
#pragma unroll
      for (int k = 0; k < 12; ++k) sv[k] = from_h2(to_h2(cadd(sv[k], cscale(kappa, f[k]))));
    }
    Cx<float> x[12];
    __syncthreads();
    lds_store_h2(lx, t, sv);
    __syncthreads();
    for (int it = 0; it < nm; ++it) {
      lds_hop_sum_h2(g, c, t, lx, lU, acc);
      clover_h2(cl, acc);
#pragma unroll
      for (int k = 0; k < 12; ++k) x[k] = cadd(sv[k], cscale(kappa, acc[k]));
      __syncthreads();
      lds_store_h2(lx, t, x);
      __syncthreads();
    }
    lds_load_h2(lx, t, x);
    SpinorIO<__half>::store(qout + off, site, g.V, x);
    if (epi == SAP_EPI_NONE) continue;
    lds_hop_sum_h2(g, c, t, lx, lU, acc);
    clover_h2(cl, acc);
    Cx<float> o[12];
    if (epi == SAP_EPI_SUPDATE) {
      Cx<float> iv[12];
      SpinorIO<__half>::load(in + off, site, g.V, iv);
#pragma unroll
      for (int k = 0; k < 12; ++k) o[k] = cadd(csub(cadd(sv[k], iv[k]), x[k]), cscale(kappa, acc[k]));
      SpinorIO<__half>::store(s + off, site, g.V, o);
    } else {
#pragma unroll
      for (int k = 0; k < 12; ++k) o[k] = csub(x[k], cscale(kappa, acc[k]));
      if (epi == SAP_EPI_OUT_FACE) {
        Cx<float> s0[12];
        SpinorIO<__half>::load(((flags & SAP_SRC_IN) ? in : s) + off, site, g.V, s0);
#pragma unroll
        for (int k = 0; k < 12; ++k) o[k] = cadd(o[k], csub(s0[k], sv[k]));
      }
      SpinorIO<__half>::store(out + off, site, g.V, o);
    }
  }
}

bool sap_fusable(const Geo& g) { return g.Vb == SAP_THREADS; }

template <typename S>
void launch_sap_block(const Geo& g, BlockSel sel, const Batch& B, int nl, float kappa, int nm, int flags, S* s,
                      const S* in, const S* qface, S* qout, S* out, const S* U, const S* C, hipStream_t st) {
  if (sel.n == 0) return;
  if (nm < 1) nm = 1;  // the first application is unconditional (static_solver.cc:77-101)
#define SAP_ARGS g, sel.list, B.n, B.stride, B.active, kappa, nm, flags, s, in, qface, qout, out, U, C
  if constexpr (std::is_same<S, __half>::value) {
    if (nl == 12)
      hipLaunchKernelGGL((k_sap_block_h<12>), dim3(sel.n), dim3(SAP_THREADS), 0, st, SAP_ARGS);
    else
      hipLaunchKernelGGL((k_sap_block_h<18>), dim3(sel.n), dim3(SAP_THREADS), 0, st, SAP_ARGS);
  } else {
    if (nl == 12)
      hipLaunchKernelGGL((k_sap_block<S, 12>), dim3(sel.n), dim3(SAP_THREADS), 0, st, SAP_ARGS);
    else
      hipLaunchKernelGGL((k_sap_block<S, 18>), dim3(sel.n), dim3(SAP_THREADS), 0, st, SAP_ARGS);
  }
#undef SAP_ARGS
}

template void launch_sap_block<float>(const Geo&, BlockSel, const Batch&, int, float, int, int, float*, const float*,
                                      const float*, float*, float*, const float*, const float*, hipStream_t);
template void launch_sap_block<__half>(const Geo&, BlockSel, const Batch&, int, float, int, int, __half*,
                                       const __half*, const __half*, __half*, __half*, const __half*, const __half*,
                                       hipStream_t);

}  // namespace qws
