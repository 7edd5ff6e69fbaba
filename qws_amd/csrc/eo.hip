// eo.hip -- the even/odd operator family (deo_*, dee_deo_*, one_minus_*, mtilde) on
// x-compressed parity fields.
//
// A parity-p field holds Vh = V/2 sites h = xh + (Lx/2)(y + Ly(z + Lz t)); its row
// parity is ick = (p + y + z + t) % 2 (deo_in_d.cc:80-85).  Neighbours follow the
// reference's index arithmetic: x-forward of h is xh (ick 0) or xh + 1 (ick 1) of
// the other parity, x-backward xh - 1 (ick 0) or xh (ick 1); y/z/t keep xh; every
// index wraps.  With even extents this is the physical lattice with x = 2 xh + ick.
// Device layout per parity (site-major SoA like the blocked layout):
//   spinor  FP32 float4 [6][Vh], FP64 double2 [12][Vh]
//   links   [2][4][NL/2][Vh] Vec2  (link from site h of parity p)
//   clover  [2][18][Vh] float4 / [2][36][Vh] double2
// One thread per site of the output parity.
#include <type_traits>

#include "kernels.h"

namespace qws {

constexpr int EO_THREADS = 256;

struct EoSite {
  int c[4];  // xh, y, z, t
  int ick;
};

__device__ __forceinline__ void eo_coords(const Geo& g, int pe, long long h, EoSite& e) {
  const int nxh = g.L[0] >> 1;
  long long r = h;
  e.c[0] = (int)(r % nxh); r /= nxh;
  e.c[1] = (int)(r % g.L[1]); r /= g.L[1];
  e.c[2] = (int)(r % g.L[2]);
  e.c[3] = (int)(r / g.L[2]);
  e.ick = (pe + e.c[1] + e.c[2] + e.c[3]) & 1;
}

// One hop of H (or of H^+: projector and reconstruction of the opposite direction,
// deo_dag_in.cc) into acc; set 1 skips hops that wrap round the lattice ("_in_"),
// set 2 keeps only those ("_out_pre/_pos"), whose t edge carries 2*tbc.
template <typename S, int NL, int MU, int DIR, bool DAG>
__device__ __forceinline__ void eo_hop(const Geo& g, int pe, long long h, const EoSite& e, int set, const S* in,
                                       const S* U, long long Vh, Cx<typename Store<S>::R> acc[12]) {
  using R = typename Store<S>::R;
  const int nxh = g.L[0] >> 1;
  int n[4] = {e.c[0], e.c[1], e.c[2], e.c[3]};
  bool wraps;
  if (MU == 0) {
    const int v = e.c[0] + (DIR > 0 ? e.ick : e.ick - 1);
    wraps = v < 0 || v >= nxh;
    n[0] = v < 0 ? nxh - 1 : (v >= nxh ? 0 : v);
  } else {
    const int ext = g.L[MU];
    const int v = e.c[MU] + DIR;
    wraps = v < 0 || v >= ext;
    n[MU] = v < 0 ? ext - 1 : (v >= ext ? 0 : v);
  }
  if ((set == 1 && wraps) || (set == 2 && !wraps)) return;
  const long long hn = n[0] + (long long)nxh * (n[1] + (long long)g.L[1] * (n[2] + (long long)g.L[2] * n[3]));
  constexpr int PD = DAG ? -DIR : DIR;
  Cx<R> v[12], a[6], b[6], u[9];
  SpinorIO<S>::load(in, hn, Vh, v);
  project<R, MU, PD>(v, a);
  // forward: the link of h (parity pe); backward: the link of the neighbour (parity 1-pe)
  const S* Ub = U + (long long)(DIR > 0 ? pe : 1 - pe) * 4 * NL * Vh;
  load_link<S, NL>(Ub, MU, DIR > 0 ? h : hn, Vh, u);
  link_mult<R, DIR>(u, a, b);
  const R tf = (MU == 3 && wraps) ? (R)(2 * g.tbc) : (R)2;
  reconstruct<R, MU, PD>(b, tf, acc);
}

// For every site h of parity pe: acc = sum of the hops of `set` (0 all, 1 no wrap,
// 2 wrap only, 3 none: acc = in[h]) from in; clv: acc = T^-1 acc; then
// mode 0: out = acc, 1: out += fac acc, 2: out = in0 + fac acc.
template <typename S, int NL, bool DAG>
__global__ void __launch_bounds__(EO_THREADS)
k_deo(Geo g, int pe, int set, int clv, int mode, typename Store<S>::R fac, S* out, const S* __restrict__ in,
      const S* in0, const S* __restrict__ U, const S* __restrict__ C) {
  using R = typename Store<S>::R;
  const long long Vh = g.V >> 1;
  // XCD-contiguous ranges of sites: the z and t neighbours a workgroup reads were fetched
  // by workgroups on the same XCD (its L2), not by ones spread round-robin over the XCDs
  const long long h = xcd_remap(g, blockIdx.x, gridDim.x) * (long long)EO_THREADS + threadIdx.x;
  if (h >= Vh) return;
  EoSite e;
  eo_coords(g, pe, h, e);
  Cx<R> acc[12];
  if (set == 3) {
    SpinorIO<S>::load(in, h, Vh, acc);
  } else {
#pragma unroll
    for (int k = 0; k < 12; ++k) acc[k] = {0, 0};
    eo_hop<S, NL, 0, +1, DAG>(g, pe, h, e, set, in, U, Vh, acc);
    eo_hop<S, NL, 0, -1, DAG>(g, pe, h, e, set, in, U, Vh, acc);
    eo_hop<S, NL, 1, +1, DAG>(g, pe, h, e, set, in, U, Vh, acc);
    eo_hop<S, NL, 1, -1, DAG>(g, pe, h, e, set, in, U, Vh, acc);
    eo_hop<S, NL, 2, +1, DAG>(g, pe, h, e, set, in, U, Vh, acc);
    eo_hop<S, NL, 2, -1, DAG>(g, pe, h, e, set, in, U, Vh, acc);
    eo_hop<S, NL, 3, +1, DAG>(g, pe, h, e, set, in, U, Vh, acc);
    eo_hop<S, NL, 3, -1, DAG>(g, pe, h, e, set, in, U, Vh, acc);
  }
  if (clv) clover_apply<S>(C + (long long)pe * 72 * Vh, h, Vh, acc);
  if (mode != 0) {
    Cx<R> o[12];
    SpinorIO<S>::load(mode == 1 ? out : in0, h, Vh, o);
#pragma unroll
    for (int k = 0; k < 12; ++k) acc[k] = cadd(o[k], cscale(fac, acc[k]));
  }
  SpinorIO<S>::store(out, h, Vh, acc);
}

template <typename S>
void launch_deo(const Geo& g, int nl, int pe, int set, int dag, int clv, int mode, typename Store<S>::R fac, S* out,
                const S* in, const S* in0, const S* U, const S* C, hipStream_t st) {
  const long long Vh = g.V >> 1;
  dim3 grid((unsigned)((Vh + EO_THREADS - 1) / EO_THREADS)), block(EO_THREADS);
#define DEO_ARGS g, pe, set, clv, mode, fac, out, in, in0, U, C
  if constexpr (std::is_same<S, double>::value) {
    if (dag) hipLaunchKernelGGL((k_deo<S, 18, true>), grid, block, 0, st, DEO_ARGS);
    else hipLaunchKernelGGL((k_deo<S, 18, false>), grid, block, 0, st, DEO_ARGS);
  } else {
    if (nl == 12) {
      if (dag) hipLaunchKernelGGL((k_deo<S, 12, true>), grid, block, 0, st, DEO_ARGS);
      else hipLaunchKernelGGL((k_deo<S, 12, false>), grid, block, 0, st, DEO_ARGS);
    } else {
      if (dag) hipLaunchKernelGGL((k_deo<S, 18, true>), grid, block, 0, st, DEO_ARGS);
      else hipLaunchKernelGGL((k_deo<S, 18, false>), grid, block, 0, st, DEO_ARGS);
    }
  }
#undef DEO_ARGS
}

// ---------------------------------------------------------------- host layouts
// layout 0 (plain, x-compressed): spinor [Vh][24], links [2][4][Vh][18], clover [2][Vh][72].
// layout 1 (QWS AoSoA read as parity halves, qws.h:111-192 with deo_in_d.cc:65-90):
// vector i = xh/vlen + nxv (y + ny (z + nz t)), lane j = xh % vlen; spinor
// [(i*24 + k)*vlen + j], links glud[((i + vol*mu + 4*vol*p)*18 + k)*vlen + j], clover
// clvd[((i + vol*p)*72 + k)*vlen + j], vol = Vh / vlen.
__device__ __forceinline__ long long eo_host_offset(const Geo& g, int layout, int vlen, int kind, int p, int mu,
                                                    long long h, int k) {
  const long long Vh = g.V >> 1;
  if (layout == 0) {
    if (kind == 0) return h * 24 + k;
    if (kind == 1) return (((long long)p * 4 + mu) * Vh + h) * 18 + k;
    return ((long long)p * Vh + h) * 72 + k;
  }
  const long long vol = Vh / vlen;
  const long long i = h / vlen, j = h % vlen;  // xh = h % nxh and nxh % vlen == 0, so this splits xh
  if (kind == 0) return (i * 24 + k) * vlen + j;
  if (kind == 1) return ((i + vol * mu + 4 * vol * p) * 18 + k) * vlen + j;
  return ((i + vol * p) * 72 + k) * vlen + j;
}

// device offset (in reals) of real k: spinor/clover in 4-real (FP32) or 2-real
// (FP64) vectors, links in complex pairs
template <typename S>
__device__ __forceinline__ long long eo_dev_offset(const Geo& g, int kind, int nl, int p, int mu, long long h, int k) {
  const long long Vh = g.V >> 1;
  const int w = sizeof(S) == 8 ? 2 : 4;
  if (kind == 0) return ((long long)(k / w) * Vh + h) * w + (k % w);
  if (kind == 1) return ((((long long)p * 4 + mu) * (nl / 2) + (k >> 1)) * Vh + h) * 2 + (k & 1);
  return (((long long)p * (72 / w) + k / w) * Vh + h) * w + (k % w);
}

template <typename S>
__global__ void k_eo_import(Geo g, int kind, int layout, int vlen, int nl, const void* src, int src_is_double, S* dst) {
  const long long Vh = g.V >> 1;
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const int np = kind == 0 ? 1 : 2;
  if (t >= Vh * np) return;
  const int p = (int)(t / Vh);
  const long long h = t - p * Vh;
  const int nmu = kind == 1 ? 4 : 1;
  const int nrec = kind == 0 ? 24 : (kind == 1 ? nl : 72);
  for (int mu = 0; mu < nmu; ++mu)
    for (int k = 0; k < nrec; ++k) {
      const long long o = eo_host_offset(g, layout, vlen, kind, p, mu, h, k);
      const double v = src_is_double ? ((const double*)src)[o] : (double)((const float*)src)[o];
      dst[eo_dev_offset<S>(g, kind, nl, p, mu, h, k)] = (S)v;
    }
}

template <typename S>
__global__ void k_eo_export(Geo g, int layout, int vlen, const S* src, void* dst, int dst_is_double) {
  const long long Vh = g.V >> 1;
  const long long h = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (h >= Vh) return;
  for (int k = 0; k < 24; ++k) {
    const long long o = eo_host_offset(g, layout, vlen, 0, 0, 0, h, k);
    const double v = (double)src[eo_dev_offset<S>(g, 0, 0, 0, 0, h, k)];
    if (dst_is_double) ((double*)dst)[o] = v;
    else ((float*)dst)[o] = (float)v;
  }
}

// Even/odd links and clover from the blocked-layout fields: the parity-p site h is
// the lattice site (2 xh + (p + y + z + t) % 2, y, z, t) (even extents only).
template <typename S>
__global__ void k_eo_gather(Geo g, int nl, const S* Ud, const S* Cd, S* Ue, S* Ce) {
  const long long Vh = g.V >> 1;
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= 2 * Vh) return;
  const int p = (int)(t / Vh);
  const long long h = t - p * Vh;
  EoSite e;
  eo_coords(g, p, h, e);
  const int x[4] = {2 * e.c[0] + e.ick, e.c[1], e.c[2], e.c[3]};
  int bb[4], ll[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    bb[m] = x[m] / g.B[m];
    ll[m] = x[m] - bb[m] * g.B[m];
  }
  const long long s = site_index(g, bb, ll);
  const int nc = nl / 2;
  for (int mu = 0; mu < 4; ++mu)
    for (int k = 0; k < nc; ++k)
      for (int ri = 0; ri < 2; ++ri)
        Ue[((((long long)p * 4 + mu) * nc + k) * Vh + h) * 2 + ri] = Ud[(((long long)mu * nc + k) * g.V + s) * 2 + ri];
  const int w = sizeof(S) == 8 ? 2 : 4;
  for (int f = 0; f < 72 / w; ++f)
    for (int q = 0; q < w; ++q)
      Ce[(((long long)p * (72 / w) + f) * Vh + h) * w + q] = Cd[((long long)f * g.V + s) * w + q];
}

template <typename S>
void launch_eo_import(const Geo& g, int kind, int layout, int vlen, int nl, const void* src, int src_is_double, S* dst,
                      hipStream_t st) {
  const long long n = (g.V >> 1) * (kind == 0 ? 1 : 2);
  hipLaunchKernelGGL((k_eo_import<S>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g, kind, layout, vlen, nl,
                     src, src_is_double, dst);
}
template <typename S>
void launch_eo_export(const Geo& g, int layout, int vlen, const S* src, void* dst, int dst_is_double, hipStream_t st) {
  const long long n = g.V >> 1;
  hipLaunchKernelGGL((k_eo_export<S>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g, layout, vlen, src, dst,
                     dst_is_double);
}
template <typename S>
void launch_eo_gather(const Geo& g, int nl, const S* Ud, const S* Cd, S* Ue, S* Ce, hipStream_t st) {
  const long long n = g.V;
  hipLaunchKernelGGL((k_eo_gather<S>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g, nl, Ud, Cd, Ue, Ce);
}

#define INST_EO(S)                                                                                                  \
  template void launch_deo<S>(const Geo&, int, int, int, int, int, int, typename Store<S>::R, S*, const S*,          \
                              const S*, const S*, const S*, hipStream_t);                                            \
  template void launch_eo_import<S>(const Geo&, int, int, int, int, const void*, int, S*, hipStream_t);              \
  template void launch_eo_export<S>(const Geo&, int, int, const S*, void*, int, hipStream_t);                        \
  template void launch_eo_gather<S>(const Geo&, int, const S*, const S*, S*, S*, hipStream_t);
INST_EO(float)
INST_EO(double)

}  // namespace qws
