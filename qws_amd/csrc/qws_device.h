// qws_device.h -- device-side data layout and site algebra for the MI355X
// Wilson-clover DD solver.
//
// Device layout (site-major SoA, blocked site order):
//   A lattice of extents L[4] is tiled by SAP blocks of extents B[4]; the device
//   site index is s = blk*Vb + loc with blk = bx + NB0*(by + NB1*(bz + NB2*bt))
//   and loc = lx + B0*(ly + B1*(lz + B2*lt)).  A workgroup therefore reads a block
//   as one contiguous range of every component array.
//   spinor FP32  : float4  [6][V]   float4 j = complex components 2j, 2j+1,
//                                    component k = colour*4 + spin (qws.h:170-177 order)
//   spinor FP64  : double2 [12][V]
//   spinor FP16  : Half4   [6][V]   same component order as FP32 (storage only;
//                                    arithmetic in FP32)
//   link         : Vec2    [4][NL/2][V]  complex c[i][j] at k = i*3 + j, link from s to
//                                    s+mu; NL = 18 stores 3 rows, NL = 12 stores rows 0,1
//                                    and rebuilds row 2 = conj(row0 x row1)
//   clover (T^-1): FP32 float4 [18][V] (72 reals), FP64 double2 [36][V],
//                  FP16 Half4 [18][V]; per site chirality i at reals 36i..36i+35 in
//                  QWS's packing (clover_def.h:88-284)
// Every kernel is written for 64-wide wavefronts; one thread owns one site.
#pragma once

#include <type_traits>

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qws {

// ------------------------------------------------------------------ geometry
struct Geo {
  int L[4];       // lattice extents
  int B[4];       // block extents
  int NB[4];      // blocks per dimension
  int sB[4];      // strides inside a block: 1, B0, B0*B1, B0*B1*B2
  int sN[4];      // strides between blocks: 1, NB0, NB0*NB1, NB0*NB1*NB2
  int Vb;         // sites per block
  int nblk;       // number of blocks
  long long V;    // sites
  int tbc;        // +1 periodic / -1 antiperiodic in t
  int hopmask;    // bit 2*mu + (dir < 0): hop taken by the full operator (0xFF); being a
                  // runtime value it keeps each hop in its own basic block, so the
                  // compiler cannot hoist all eight neighbour loads (256 -> ~146 VGPRs)
  int xcd_remap;  // 1: map workgroups to blocks XCD by XCD (xcd_remap), 0: identity
  int dslash_kernel;  // FP32 full operator on 256-site blocks: 0 global gathers (k_dslash),
                      // 3/4: block spinor staged in LDS, compiled for that many waves per SIMD
  int dslash_kernel_d;  // the same for FP64 (18-real links): 0 gathers, 2/3 LDS-staged
};

// MI355X dispatches the workgroups of a launch round-robin over its 8 XCDs, each
// with its own L2.  Give XCD x a contiguous range of logical workgroup ids (blocks
// that share faces then share an L2): a bijection of [0, nwg).
constexpr int NUM_XCD = 8;
__device__ __forceinline__ int xcd_remap_n(int on, int wg, int nwg) {
  if (!on) return wg;
  const int q = nwg / NUM_XCD, r = nwg % NUM_XCD;
  const int x = wg % NUM_XCD, i = wg / NUM_XCD;
  return x * q + (x < r ? x : r) + i;
}
__device__ __forceinline__ int xcd_remap(const Geo& g, int wg, int nwg) { return xcd_remap_n(g.xcd_remap, wg, nwg); }

// p, hidden from the optimiser.  Loads through the result taken inside a loop over
// right-hand sides are not loop-invariant to the compiler, so it does not hoist a
// site's links and clover out of the loop and keep them all in registers (which
// costs occupancy; the re-reads hit the cache).
template <typename T> __device__ __forceinline__ const T* opaque(const T* p) {
  asm volatile("" : "+s"(p));
  return p;
}
// the same for a per-thread value (everything derived from it is recomputed)
__device__ __forceinline__ long long opaque_v(long long x) {
  asm volatile("" : "+v"(x));
  return x;
}

struct SiteC {
  int b[4];  // block coordinates
  int l[4];  // coordinates inside the block
};

__device__ __forceinline__ void site_coords(const Geo& g, long long s, SiteC& c) {
  int blk = (int)(s / g.Vb);
  int loc = (int)(s - (long long)blk * g.Vb);
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    c.l[m] = loc % g.B[m];
    loc /= g.B[m];
    c.b[m] = blk % g.NB[m];
    blk /= g.NB[m];
  }
  c.l[3] = loc;
  c.b[3] = blk;
}

__device__ __forceinline__ long long site_index(const Geo& g, const int b[4], const int l[4]) {
  long long blk = b[0] + g.sN[1] * b[1] + g.sN[2] * b[2] + (long long)g.sN[3] * b[3];
  int loc = l[0] + g.sB[1] * l[1] + g.sB[2] * l[2] + g.sB[3] * l[3];
  return blk * g.Vb + loc;
}

__device__ __forceinline__ int block_colour(const SiteC& c) {
  return (c.b[0] + c.b[1] + c.b[2] + c.b[3]) & 1;
}

// Neighbour of a site one step along mu (dir = +1/-1).  inside = the hop stays in
// the block without wrapping (ddd_in_d.cc:95-188); every other hop is a face term.
__device__ __forceinline__ long long neighbour(const Geo& g, long long s, const SiteC& c, int mu,
                                               int dir, bool& inside) {
  int l = c.l[mu] + dir;
  if (l >= 0 && l < g.B[mu]) {
    inside = true;
    return s + dir * g.sB[mu];
  }
  inside = false;
  int b[4] = {c.b[0], c.b[1], c.b[2], c.b[3]};
  int ll[4] = {c.l[0], c.l[1], c.l[2], c.l[3]};
  b[mu] = (b[mu] + dir + g.NB[mu]) % g.NB[mu];
  ll[mu] = (l + g.B[mu]) % g.B[mu];
  return site_index(g, b, ll);
}

// ------------------------------------------------------------------ complex
template <typename R> struct Cx { R re, im; };

template <typename R> __device__ __forceinline__ Cx<R> cadd(Cx<R> a, Cx<R> b) { return {a.re + b.re, a.im + b.im}; }
template <typename R> __device__ __forceinline__ Cx<R> csub(Cx<R> a, Cx<R> b) { return {a.re - b.re, a.im - b.im}; }
// a b and conj(a) b with the fused multiply-add written out, so that every kernel
// rounds them the same way (left to the compiler, which product it fuses depends on
// the surrounding code)
__device__ __forceinline__ float fmar(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmar(double a, double b, double c) { return __builtin_fma(a, b, c); }
template <typename R> __device__ __forceinline__ Cx<R> cmul(Cx<R> a, Cx<R> b) {
  return {fmar(a.re, b.re, -(a.im * b.im)), fmar(a.re, b.im, a.im * b.re)};
}
// conj(a) * b
template <typename R> __device__ __forceinline__ Cx<R> cmulc(Cx<R> a, Cx<R> b) {
  return {fmar(a.re, b.re, a.im * b.im), fmar(a.re, b.im, -(a.im * b.re))};
}
template <typename R> __device__ __forceinline__ Cx<R> cconj(Cx<R> a) { return {a.re, -a.im}; }
template <typename R> __device__ __forceinline__ Cx<R> cimul(Cx<R> a) { return {-a.im, a.re}; }   // i*a
template <typename R> __device__ __forceinline__ Cx<R> cmimul(Cx<R> a) { return {a.im, -a.re}; }  // -i*a
template <typename R> __device__ __forceinline__ Cx<R> cscale(R s, Cx<R> a) { return {s * a.re, s * a.im}; }

// Complex multiply-accumulate as chains of fused multiply-adds (4 FMAs; written as
// c + a*b through cmul and cadd it costs 2 multiplies, 2 FMAs and 2 adds, since the
// compiler may not reassociate the sum).
// c + a b
template <typename R> __device__ __forceinline__ Cx<R> cfma(Cx<R> a, Cx<R> b, Cx<R> c) {
  return {fmar(-a.im, b.im, fmar(a.re, b.re, c.re)), fmar(a.im, b.re, fmar(a.re, b.im, c.im))};
}
// c + conj(a) b
template <typename R> __device__ __forceinline__ Cx<R> cfmac(Cx<R> a, Cx<R> b, Cx<R> c) {
  return {fmar(a.im, b.im, fmar(a.re, b.re, c.re)), fmar(-a.im, b.re, fmar(a.re, b.im, c.im))};
}
// c - a b
template <typename R> __device__ __forceinline__ Cx<R> cfnma(Cx<R> a, Cx<R> b, Cx<R> c) {
  return {fmar(a.im, b.im, fmar(-a.re, b.re, c.re)), fmar(-a.im, b.re, fmar(-a.re, b.im, c.im))};
}

// ------------------------------------------------------------------ storage types
// Storage precision S, arithmetic precision R.  (float,float), (double,double) and
// (__half,float) are used.
template <typename S> struct Store;
template <> struct Store<float> { using R = float; };
template <> struct Store<double> { using R = double; };
template <> struct Store<__half> { using R = float; };

// ---- spinors
template <typename S> struct SpinorIO;

template <> struct SpinorIO<float> {
  static __device__ __forceinline__ void load(const float* base, long long s, long long V, Cx<float> v[12]) {
    const float4* p = reinterpret_cast<const float4*>(base);
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      float4 q = p[j * V + s];
      v[2 * j] = {q.x, q.y};
      v[2 * j + 1] = {q.z, q.w};
    }
  }
  static __device__ __forceinline__ void store(float* base, long long s, long long V, const Cx<float> v[12]) {
    float4* p = reinterpret_cast<float4*>(base);
#pragma unroll
    for (int j = 0; j < 6; ++j) p[j * V + s] = make_float4(v[2 * j].re, v[2 * j].im, v[2 * j + 1].re, v[2 * j + 1].im);
  }
};

template <> struct SpinorIO<double> {
  static __device__ __forceinline__ void load(const double* base, long long s, long long V, Cx<double> v[12]) {
    const double2* p = reinterpret_cast<const double2*>(base);
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      double2 q = p[k * V + s];
      v[k] = {q.x, q.y};
    }
  }
  static __device__ __forceinline__ void store(double* base, long long s, long long V, const Cx<double> v[12]) {
    double2* p = reinterpret_cast<double2*>(base);
#pragma unroll
    for (int k = 0; k < 12; ++k) p[k * V + s] = make_double2(v[k].re, v[k].im);
  }
};

// four FP16 values in 8 bytes, ordered like the components of a float4
struct alignas(8) Half4 { __half2 a, b; };
__device__ __forceinline__ float4 half4_load(const __half* base, long long i) {
  const Half4 h = reinterpret_cast<const Half4*>(base)[i];
  const float2 x = __half22float2(h.a), y = __half22float2(h.b);
  return make_float4(x.x, x.y, y.x, y.y);
}
__device__ __forceinline__ void half4_store(__half* base, long long i, float4 f) {
  Half4 h;
  h.a = __floats2half2_rn(f.x, f.y);
  h.b = __floats2half2_rn(f.z, f.w);
  reinterpret_cast<Half4*>(base)[i] = h;
}

template <> struct SpinorIO<__half> {
  static __device__ __forceinline__ void load(const __half* base, long long s, long long V, Cx<float> v[12]) {
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      float4 q = half4_load(base, j * V + s);
      v[2 * j] = {q.x, q.y};
      v[2 * j + 1] = {q.z, q.w};
    }
  }
  static __device__ __forceinline__ void store(__half* base, long long s, long long V, const Cx<float> v[12]) {
#pragma unroll
    for (int j = 0; j < 6; ++j)
      half4_store(base, j * V + s, make_float4(v[2 * j].re, v[2 * j].im, v[2 * j + 1].re, v[2 * j + 1].im));
  }
};

// ---- streaming loads for the SAP block kernel.  Its links and clover are read once
// per block and half-sweep and never re-read from L2 within the launch; loading them
// nontemporally leaves L2 to the spinors of neighbouring blocks.  (The full operator
// measured slower with them and keeps ordinary loads.)
typedef float f2v __attribute__((ext_vector_type(2)));
typedef float f4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ f2v ld_nt(const f2v* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ f4v ld_nt(const f4v* p) { return __builtin_nontemporal_load(p); }

// ---- links
template <typename S> __device__ __forceinline__ Cx<typename Store<S>::R> load_cx(const S* base, long long i);
template <> __device__ __forceinline__ Cx<float> load_cx<float>(const float* base, long long i) {
  float2 q = reinterpret_cast<const float2*>(base)[i];
  return {q.x, q.y};
}
template <> __device__ __forceinline__ Cx<double> load_cx<double>(const double* base, long long i) {
  double2 q = reinterpret_cast<const double2*>(base)[i];
  return {q.x, q.y};
}
template <> __device__ __forceinline__ Cx<float> load_cx<__half>(const __half* base, long long i) {
  float2 q = __half22float2(reinterpret_cast<const __half2*>(base)[i]);
  return {q.x, q.y};
}
template <typename S> __device__ __forceinline__ Cx<float> load_cx_stream(const S* base, long long i) {
  return load_cx<S>(base, i);
}
template <> __device__ __forceinline__ Cx<float> load_cx_stream<float>(const float* base, long long i) {
  const f2v q = ld_nt(reinterpret_cast<const f2v*>(base) + i);
  return {q.x, q.y};
}

// row 2 of a 12-real link = conj(row0 x row1) (wilson_d.h:160-171)
template <typename R> __device__ __forceinline__ void link_row2(Cx<R> u[9]) {
  u[6] = cconj(cfnma(u[2], u[4], cmul(u[1], u[5])));
  u[7] = cconj(cfnma(u[0], u[5], cmul(u[2], u[3])));
  u[8] = cconj(cfnma(u[1], u[3], cmul(u[0], u[4])));
}

// u[i*3+j] = QWS c[i][j] of the link at site s, direction mu.
template <typename S, int NL>
__device__ __forceinline__ void load_link(const S* U, int mu, long long s, long long V, Cx<typename Store<S>::R> u[9]) {
  constexpr int NC = NL / 2;
  const long long base = (long long)mu * NC * V + s;
#pragma unroll
  for (int k = 0; k < NC; ++k) u[k] = load_cx<S>(U, base + k * V);
  if (NL == 12) link_row2(u);
}

// ------------------------------------------------------------------ site algebra
// Spin projection for the hop along mu, dir (forward: 1 - gamma_mu, backward:
// 1 + gamma_mu) in the Dirac basis of wilson_d.h:540-657.  v[c*4+s].
template <typename R, int MU, int DIR>
__device__ __forceinline__ void project(const Cx<R> v[12], Cx<R> a[6]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const Cx<R> s0 = v[c * 4 + 0], s1 = v[c * 4 + 1], s2 = v[c * 4 + 2], s3 = v[c * 4 + 3];
    Cx<R> a0, a1;
    if (MU == 0) {
      if (DIR > 0) { a0 = cadd(s0, cmimul(s3)); a1 = cadd(s1, cmimul(s2)); }
      else         { a0 = cadd(s0, cimul(s3));  a1 = cadd(s1, cimul(s2)); }
    } else if (MU == 1) {
      if (DIR > 0) { a0 = csub(s0, s3); a1 = cadd(s1, s2); }
      else         { a0 = cadd(s0, s3); a1 = csub(s1, s2); }
    } else if (MU == 2) {
      if (DIR > 0) { a0 = cadd(s0, cmimul(s2)); a1 = cadd(s1, cimul(s3)); }
      else         { a0 = cadd(s0, cimul(s2));  a1 = cadd(s1, cmimul(s3)); }
    } else {
      if (DIR > 0) { a0 = s2; a1 = s3; } else { a0 = s0; a1 = s1; }
    }
    a[c * 2 + 0] = a0;
    a[c * 2 + 1] = a1;
  }
}

// b = W a (forward) or W^+ a (backward) with W[i][j] = c[j][i] (wilson_d.h:110-135,176-187)
template <typename R, int DIR>
__device__ __forceinline__ void link_mult(const Cx<R> u[9], const Cx<R> a[6], Cx<R> b[6]) {
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      Cx<R> acc = DIR > 0 ? cmul(u[j], a[h]) : cmulc(u[j * 3], a[h]);
#pragma unroll
      for (int i = 1; i < 3; ++i)
        acc = DIR > 0 ? cfma(u[i * 3 + j], a[i * 2 + h], acc) : cfmac(u[j * 3 + i], a[i * 2 + h], acc);
      b[j * 2 + h] = acc;
    }
}

// Accumulate the reconstructed 4-spinor of a hop (wilson_d.h:362-537); tf is the
// factor of the t hops (2, or 2*tbc across the t edge).
template <typename R, int MU, int DIR>
__device__ __forceinline__ void reconstruct(const Cx<R> b[6], R tf, Cx<R> o[12]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const Cx<R> b0 = b[c * 2 + 0], b1 = b[c * 2 + 1];
    Cx<R>* s = o + c * 4;
    if (MU == 0) {
      s[0] = cadd(s[0], b0); s[1] = cadd(s[1], b1);
      if (DIR > 0) { s[2] = cadd(s[2], cimul(b1)); s[3] = cadd(s[3], cimul(b0)); }
      else         { s[2] = csub(s[2], cimul(b1)); s[3] = csub(s[3], cimul(b0)); }
    } else if (MU == 1) {
      s[0] = cadd(s[0], b0); s[1] = cadd(s[1], b1);
      if (DIR > 0) { s[2] = cadd(s[2], b1); s[3] = csub(s[3], b0); }
      else         { s[2] = csub(s[2], b1); s[3] = cadd(s[3], b0); }
    } else if (MU == 2) {
      s[0] = cadd(s[0], b0); s[1] = cadd(s[1], b1);
      if (DIR > 0) { s[2] = cadd(s[2], cimul(b0)); s[3] = csub(s[3], cimul(b1)); }
      else         { s[2] = csub(s[2], cimul(b0)); s[3] = cadd(s[3], cimul(b1)); }
    } else {
      auto ax = [tf](Cx<R> y, Cx<R> b) { return Cx<R>{fmar(tf, b.re, y.re), fmar(tf, b.im, y.im)}; };
      if (DIR > 0) { s[2] = ax(s[2], b0); s[3] = ax(s[3], b1); }
      else         { s[0] = ax(s[0], b0); s[1] = ax(s[1], b1); }
    }
  }
}

// ---- clover inverse: v <- T^-1 v (clover_def.h:88-284)
template <typename S> struct CloverIO;

template <> struct CloverIO<float> {
  // reals k of chirality i: float4 index (36i+k)/4
  template <bool STREAM = false>
  static __device__ __forceinline__ void load_chir(const float* C, long long s, long long V, int i, float a[36]) {
    const f4v* p = reinterpret_cast<const f4v*>(C);
#pragma unroll
    for (int f = 0; f < 9; ++f) {
      const f4v* q = p + (long long)(9 * i + f) * V + s;
      const f4v v = STREAM ? ld_nt(q) : *q;
      a[4 * f] = v.x; a[4 * f + 1] = v.y; a[4 * f + 2] = v.z; a[4 * f + 3] = v.w;
    }
  }
};
template <> struct CloverIO<double> {
  static __device__ __forceinline__ void load_chir(const double* C, long long s, long long V, int i, double a[36]) {
    const double2* p = reinterpret_cast<const double2*>(C);
#pragma unroll
    for (int f = 0; f < 18; ++f) {
      double2 q = p[(long long)(18 * i + f) * V + s];
      a[2 * f] = q.x; a[2 * f + 1] = q.y;
    }
  }
};
template <> struct CloverIO<__half> {
  static __device__ __forceinline__ void load_chir(const __half* C, long long s, long long V, int i, float a[36]) {
#pragma unroll
    for (int f = 0; f < 9; ++f) {
      float4 q = half4_load(C, (long long)(9 * i + f) * V + s);
      a[4 * f] = q.x; a[4 * f + 1] = q.y; a[4 * f + 2] = q.z; a[4 * f + 3] = q.w;
    }
  }
};

// y = A x for one packed Hermitian 6x6 (diagonal a[0..5], upper triangle row by
// row from a[6] as (re,im)).
template <typename R>
__device__ __forceinline__ void herm6_mult(const R a[36], const Cx<R> x[6], Cx<R> y[6]) {
#pragma unroll
  for (int r = 0; r < 6; ++r) y[r] = cscale(a[r], x[r]);
  int p = 6;
#pragma unroll
  for (int r = 0; r < 6; ++r)
#pragma unroll
    for (int c = r + 1; c < 6; ++c) {
      Cx<R> m = {a[p], a[p + 1]};
      y[r] = cfma(m, x[c], y[r]);
      y[c] = cfmac(m, x[r], y[c]);
      p += 2;
    }
}

template <typename S>
__device__ __forceinline__ void clover_apply(const S* C, long long s, long long V, Cx<typename Store<S>::R> v[12]) {
  using R = typename Store<S>::R;
  Cx<R> x[6], y0[6], y1[6];
  R a[36];
  // chirality 0: x = (s0+s2, s1+s3)
#pragma unroll
  for (int c = 0; c < 3; ++c) { x[c] = cadd(v[c * 4 + 0], v[c * 4 + 2]); x[3 + c] = cadd(v[c * 4 + 1], v[c * 4 + 3]); }
  CloverIO<S>::load_chir(C, s, V, 0, a);
  herm6_mult(a, x, y0);
#pragma unroll
  for (int c = 0; c < 3; ++c) { x[c] = csub(v[c * 4 + 0], v[c * 4 + 2]); x[3 + c] = csub(v[c * 4 + 1], v[c * 4 + 3]); }
  CloverIO<S>::load_chir(C, s, V, 1, a);
  herm6_mult(a, x, y1);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    v[c * 4 + 0] = cadd(y0[c], y1[c]);
    v[c * 4 + 2] = csub(y0[c], y1[c]);
    v[c * 4 + 1] = cadd(y0[3 + c], y1[3 + c]);
    v[c * 4 + 3] = csub(y0[3 + c], y1[3 + c]);
  }
}

// ------------------------------------------------------------------ hopping sum
enum HopSet { HOPS_ALL = 0, HOPS_INSIDE = 1, HOPS_FACE = 2, HOPS_NONE = 3 };  // NONE: acc = in (clover only)

template <int SET>
__device__ __forceinline__ bool take_hop(bool inside) {
  return SET == HOPS_ALL || (SET == HOPS_INSIDE ? inside : !inside);
}

template <typename S, int NL, int SET, int MU, int DIR>
__device__ __forceinline__ void one_hop(const Geo& g, long long s, const SiteC& c, const S* in, const S* U,
                                        Cx<typename Store<S>::R> acc[12]) {
  using R = typename Store<S>::R;
  bool inside;
  long long n = neighbour(g, s, c, MU, DIR, inside);
  if (!take_hop<SET>(inside)) return;
  if (SET == HOPS_ALL && !(g.hopmask & (1 << (2 * MU + (DIR > 0 ? 0 : 1))))) return;
  Cx<R> v[12], a[6], b[6], u[9];
  SpinorIO<S>::load(in, n, g.V, v);
  project<R, MU, DIR>(v, a);
  load_link<S, NL>(U, MU, DIR > 0 ? s : n, g.V, u);
  link_mult<R, DIR>(u, a, b);
  R tf = R(2);
  if (MU == 3) {
    const int t = c.b[3] * g.B[3] + c.l[3];
    if ((DIR > 0 && t == g.L[3] - 1) || (DIR < 0 && t == 0)) tf = R(2 * g.tbc);
  }
  reconstruct<R, MU, DIR>(b, tf, acc);
}

// acc = H_SET in at site s
template <typename S, int NL, int SET>
__device__ __forceinline__ void hop_sum(const Geo& g, long long s, const SiteC& c, const S* in, const S* U,
                                        Cx<typename Store<S>::R> acc[12]) {
  if (SET == HOPS_NONE) {
    SpinorIO<S>::load(in, s, g.V, acc);
    return;
  }
#pragma unroll
  for (int k = 0; k < 12; ++k) acc[k] = {0, 0};
  one_hop<S, NL, SET, 0, +1>(g, s, c, in, U, acc);
  one_hop<S, NL, SET, 0, -1>(g, s, c, in, U, acc);
  one_hop<S, NL, SET, 1, +1>(g, s, c, in, U, acc);
  one_hop<S, NL, SET, 1, -1>(g, s, c, in, U, acc);
  one_hop<S, NL, SET, 2, +1>(g, s, c, in, U, acc);
  one_hop<S, NL, SET, 2, -1>(g, s, c, in, U, acc);
  one_hop<S, NL, SET, 3, +1>(g, s, c, in, U, acc);
  one_hop<S, NL, SET, 3, -1>(g, s, c, in, U, acc);
}

}  // namespace qws
