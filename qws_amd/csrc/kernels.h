// kernels.h -- host-side declarations of the HIP kernel launchers.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "qws_device.h"

namespace qws {

// Blocks to run over: all blocks (list == nullptr) or the blocks of one colour.
struct BlockSel {
  const int* list;  // device array of block ids, or nullptr for every block
  int n;            // number of blocks
};

// out = a + b - c + gamma * T^-1 H_set(in)  on the selected blocks.
// Any of a, b, c may be nullptr (term absent).  out may alias a, b or c, never in.
// skip_interior: for set == HOPS_FACE with out == a, sites without a face hop
// are left untouched.
struct Batch;
template <typename S, int NL>
void launch_hop(const Geo& g, BlockSel sel, const Batch& B, int set, typename Store<S>::R gamma, S* out, const S* a,
                const S* b, const S* c, const S* in, const S* U, const S* C, hipStream_t st);

// out = 0 on the selected blocks
template <typename S>
void launch_zero_blocks(const Geo& g, BlockSel sel, const Batch& B, S* out, hipStream_t st);

// ---- BiCGStab vector kernels (full lattice, n = number of complex numbers)
struct Scalars {
  double bnorm, rnorm, nrm, tt;
  double rho0[2], rho[2], alpha[2], omega[2], beta[2];
  double red[4];
};

// BiCGStab's direction update of one complex component, p = r + beta (p - omega q)
// (bicgstab_precdd_s.cc:518-545), with the multiply-adds spelled out so that every
// kernel that applies it rounds alike.
template <typename R>
__device__ __forceinline__ void bicg_p_update(R& px, R& py, R qx, R qy, R rx, R ry, R wr, R wi, R br, R bi) {
  const R cr = fmar(wi, qy, fmar(-wr, qx, px));
  const R ci = fmar(-wi, qx, fmar(-wr, qy, py));
  px = fmar(br, cr, fmar(-bi, ci, rx));
  py = fmar(br, ci, fmar(bi, cr, ry));
}

constexpr int RED_BLOCKS = 1024;  // partial sums per reduction (fixed => deterministic)
constexpr int RED_THREADS = 256;

enum Finalize {
  FIN_INIT_INNER = 0,  // bnorm = rho0 = red0
  FIN_INIT_OUTER,      // bnorm = red0, rho0 = red1
  FIN_ALPHA,           // alpha = rho0 / (red0 + i red1)
  FIN_RNORM,           // rnorm = red0
  FIN_OMEGA,           // omega = (red1 + i red2) / red0
  FIN_RHO,             // rnorm = red0, rho = red1 + i red2, beta, rho0 = rho
  FIN_NRM,             // nrm = sqrt(red0)
  FIN_RAW              // red[] only
};

// A batch of right-hand sides: vector r of a buffer sits at base + r * stride
// (stride in elements of the buffer's type); RHS with active[r] == 0 are skipped.
struct Batch {
  int n;
  long long stride;
  const int* active;  // device flags, or nullptr (all active)
};

// n = complex numbers per vector; part = partials [nrhs][4][RED_BLOCKS]; sc = Scalars[nrhs]
template <typename R>
void launch_bicg_init_inner(const Batch& B, long long n, R* x, R* r, R* p, R* r0, const R* b, double* part,
                            hipStream_t st);
template <typename R>
void launch_bicg_init_outer(const Batch& B, long long n, R* r, R* p, R* r0, const R* b, const R* q, double* part,
                            hipStream_t st);
template <typename R> void launch_dot(const Batch& B, long long n, const R* a, const R* b, double* part, hipStream_t st);
template <typename R>
void launch_norm_dot(const Batch& B, long long n, const R* t, const R* r, double* part, hipStream_t st);
// x += c v1 ; r -= c v2 ; red |r|^2 (and <r0,r> if r0); c = alpha (coef 0) or omega (coef 1).
// mode XR_R_ONLY: r only (r0 must be null); XR_DEFERRED: x += alpha v0 first (r0 required)
enum { XR_FULL = 0, XR_R_ONLY = 1, XR_DEFERRED = 2 };
template <typename R>
void launch_update_xr(const Batch& B, long long n, R* x, const R* v1, R* r, const R* v2, const Scalars* sc, int coef,
                      const R* r0, double* part, hipStream_t st, int mode = XR_FULL, const R* v0 = nullptr);
// x += alpha v0 (alpha of each RHS)
template <typename R>
void launch_update_x(const Batch& B, long long n, R* x, const R* v0, const Scalars* sc, hipStream_t st);
// p = r + beta (p - omega q)
template <typename R>
void launch_update_p(const Batch& B, long long n, R* p, const R* r, const R* q, const Scalars* sc, hipStream_t st);
template <typename R> void launch_norm(const Batch& B, long long n, const R* a, double* part, hipStream_t st);
void launch_finalize(const Batch& B, const double* part, int nval, Scalars* sc, int mode, hipStream_t st);
// Per-site precision conversion of spinors (V sites; the storage layouts of the
// precisions differ): out = in / nrm (to lower precision), out = in * nrm (back).
// so/si: RHS strides of out/in in elements.
template <typename S>
void launch_d2s_scaled(const Batch& B, long long V, long long so, long long si, S* out, const double* in,
                       const Scalars* sc, hipStream_t st);
template <typename S>
void launch_s2d_scaled(const Batch& B, long long V, long long so, long long si, double* out, const S* in,
                       const Scalars* sc, hipStream_t st);
template <typename SO, typename SI> void launch_convert(long long V, SO* out, const SI* in, hipStream_t st);
// elementwise conversion of n reals
template <typename SO, typename SI> void launch_convert_reals(long long n, SO* out, const SI* in, hipStream_t st);
// FP16 copy with unit RMS component (dir > 0) and back (dir < 0); sc[r].red[0] = |in_r|^2
template <typename SO, typename SI>
void launch_scale_rms(const Batch& B, long long V, long long so, long long si, SO* out, const SI* in, const Scalars* sc,
                      int dir, hipStream_t st);
// *out ^= order-independent digest of nwords 8-byte words (clear *out first)
void launch_digest(long long nwords, const void* data, unsigned long long* out, hipStream_t st);
// HBM probe: out = sum of the nin consecutive arrays of n4 float4 in `in` (nin = 1: copy);
// cfg 0/1/2 = launch shape (1 or 4 elements per thread, or grid-stride); cfg 3/4: out
// chunk c (256 float4) = sum of the consecutive input chunks c*nin .. c*nin+nin-1
// (one input stream; n4 a multiple of 256), one workgroup per chunk or grid-stride
void launch_stream(int cfg, long long n4, int nin, const float4* in, float4* out, hipStream_t st);
// out = a - b (real arrays of length n)
template <typename R> void launch_sub(long long n, R* out, const R* a, const R* b, hipStream_t st);
// real arrays of length n: mode 0 y = fma(a, x, y); mode 1 y = fma(b, y, a * x)
template <typename R> void launch_lincomb(long long n, R a, const R* x, R b, R* y, int mode, hipStream_t st);

// ---- layout conversion (device site order <-> host layouts); host data staged in device memory
// kind: 0 = spinor (24 reals), 1 = links (4 x 18 reals), 2 = clover (72 reals)
// layout: 0 = plain lexicographic, 1 = QWS AoSoA with lane width vlen (single rank, 2 x-domains)
template <typename S>
void launch_import(const Geo& g, int kind, int layout, int vlen, int nl, const void* src, int src_is_double, S* dst,
                   hipStream_t st);
template <typename S>
void launch_export(const Geo& g, int kind, int layout, int vlen, const S* src, void* dst, int dst_is_double,
                   hipStream_t st);
// FP64 links (kind 1) or clover (kind 2) to the plain host layout (dst in device memory)
void launch_export_field(const Geo& g, int kind, const double* src, double* dst, hipStream_t st);

// ---- BQCD layouts (qws_bqcd.cc:66-647) <-> QWS's host AoSoA images in device memory
struct BqcdGeo {
  int L[4];
  long long volh;  // sites per parity in BQCD's arrays
  int eo_layout;   // 1: QWS's even/odd reading (x compressed), 0: the DD reading (x halves)
};
// QWS clover real k (0..35 of one chirality) <- BQCD's position (qws_bqcd.cc:301-337)
__host__ __device__ inline int bqcd_clover_index(int k) {
  if (k < 2) return k;
  if (k < 4) return 20 + (k - 2);
  if (k < 6) return 32 + (k - 4);
  if (k < 24) return 2 + (k - 6);
  if (k < 34) return 22 + (k - 24);
  return k;
}
// kind 0 spinor, 1 links, 2 clover; load 1: bq (BQCD) -> img (QWS AoSoA, lane width
// vlen), load 0: img -> bq; map: 2 (i - 1) + eo per lexicographic site (device)
template <typename T>
void launch_bqcd(const BqcdGeo& b, int kind, int load, const int* map, double* bq, T* img, int vlen,
                 hipStream_t st);

// ---- even/odd family on x-compressed parity fields (eo.hip)
// out (parity pe) from in (parity 1-pe): set 0 all hops, 1 no wrap, 2 wrap only, 3 none;
// dag: H^+; clv: T^-1 of parity pe; mode 0 out = acc, 1 out += fac acc, 2 out = in0 + fac acc
template <typename S>
void launch_deo(const Geo& g, int nl, int pe, int set, int dag, int clv, int mode, typename Store<S>::R fac, S* out,
                const S* in, const S* in0, const S* U, const S* C, hipStream_t st);
// kind 0 spinor (one parity), 1 links, 2 clover (both parities); layout 0 plain, 1 QWS AoSoA
template <typename S>
void launch_eo_import(const Geo& g, int kind, int layout, int vlen, int nl, const void* src, int src_is_double, S* dst,
                      hipStream_t st);
template <typename S>
void launch_eo_export(const Geo& g, int layout, int vlen, const S* src, void* dst, int dst_is_double, hipStream_t st);
template <typename S>
void launch_eo_gather(const Geo& g, int nl, const S* Ud, const S* Cd, S* Ue, S* Ce, hipStream_t st);

}  // namespace qws
