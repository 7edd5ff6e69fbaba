// dslash.hip -- hopping-term kernels: D, block-local D_in, face terms and the
// fused SAP updates, all as  out = a + b - c + gamma * T^-1 H_set(in).
//
// Reference semantics: ddd_in_s.cc:72-374 / ddd_in_d.cc:66-196 (hops inside a
// domain, then T^-1, then out = in - kappa * (...)), ddd_out_s.cc:112-626 and
// ddd_out_d.cc:84-235 (face hops of a domain, T^-1, out += factor * (...)),
// qws.cc:625-644 (full operator).  On one GPU the faces are read in place from
// the neighbouring block instead of being packed and exchanged.
#include "kernels.h"

namespace qws {

constexpr int HOP_THREADS = 256;

template <typename S, int NL, int SET>
__global__ void __launch_bounds__(HOP_THREADS)
k_hop(Geo g, const int* __restrict__ list, int wg_per_blk, int nrhs, long long stride, const int* __restrict__ active,
      typename Store<S>::R gamma, S* out, const S* a, const S* b, const S* c, const S* __restrict__ in,
      const S* __restrict__ U, const S* __restrict__ C, int skip_interior) {
  using R = typename Store<S>::R;
  const int wg = xcd_remap(g, blockIdx.x, gridDim.x);
  const int bi = wg / wg_per_blk;
  const int blk = list ? list[bi] : bi;
  const int loc = (wg - bi * wg_per_blk) * HOP_THREADS + threadIdx.x;
  if (loc >= g.Vb) return;
  const long long s = (long long)blk * g.Vb + loc;
  SiteC sc;
  site_coords(g, s, sc);
  if (SET == HOPS_FACE && skip_interior) {
    bool face = false;
#pragma unroll
    for (int m = 0; m < 4; ++m) face |= (sc.l[m] == 0) | (sc.l[m] == g.B[m] - 1);
    if (!face) return;
  }
  // right-hand sides share this thread's links and clover (cache-resident after the first)
  for (int r = 0; r < nrhs; ++r) {
    if (active && !active[r]) continue;
    const long long off = r * stride;
    const long long sr = opaque_v(s);
    SiteC scr;
    site_coords(g, sr, scr);
    Cx<R> acc[12];
    hop_sum<S, NL, SET>(g, sr, scr, in + off, opaque(U), acc);
    clover_apply<S>(opaque(C), sr, g.V, acc);
    Cx<R> v[12];
    if (a) {
      SpinorIO<S>::load(a + off, s, g.V, v);
    } else {
#pragma unroll
      for (int k = 0; k < 12; ++k) v[k] = {0, 0};
    }
    if (b) {
      Cx<R> w[12];
      SpinorIO<S>::load(b + off, s, g.V, w);
#pragma unroll
      for (int k = 0; k < 12; ++k) v[k] = cadd(v[k], w[k]);
    }
    if (c) {
      Cx<R> w[12];
      SpinorIO<S>::load(c + off, s, g.V, w);
#pragma unroll
      for (int k = 0; k < 12; ++k) v[k] = csub(v[k], w[k]);
    }
#pragma unroll
    for (int k = 0; k < 12; ++k) v[k] = cadd(v[k], cscale(gamma, acc[k]));
    SpinorIO<S>::store(out + off, s, g.V, v);
  }
}

// Full operator out = in + gamma T^-1 H in for one right-hand side over every
// site: the kernel for D itself (ddd_d_, ddd_s_ and the outer solver's D).  Written
// without the generality of k_hop (RHS loop, a/b/c terms, block lists) it keeps the
// live state of a thread small (about 90 VGPRs in FP32).  Residency is capped at 3
// waves per SIMD: with 6 waves the neighbour gathers of more sites are in flight at
// once and evict each other from L2 (FP32 32^3x64: 1.3x more HBM fetch, 2% slower).
template <typename S, int NL>
__global__ void __launch_bounds__(HOP_THREADS) __attribute__((amdgpu_waves_per_eu(1, 3)))
k_dslash(Geo g, const int* __restrict__ order, int wg_per_blk, const int* __restrict__ active,
         typename Store<S>::R gamma, S* __restrict__ out, const S* __restrict__ in, const S* __restrict__ U,
         const S* __restrict__ C) {
  using R = typename Store<S>::R;
  if (active && !*active) return;
  const int wg = xcd_remap(g, blockIdx.x, gridDim.x);
  const int bi = wg / wg_per_blk;
  const int loc = (wg - bi * wg_per_blk) * HOP_THREADS + threadIdx.x;
  if (loc >= g.Vb) return;
  const long long s = (long long)order[bi] * g.Vb + loc;
  SiteC sc;
  site_coords(g, s, sc);
  Cx<R> acc[12];
  hop_sum<S, NL, HOPS_ALL>(g, s, sc, in, U, acc);
  clover_apply<S>(C, s, g.V, acc);
  Cx<R> v[12];
  SpinorIO<S>::load(in, s, g.V, v);
#pragma unroll
  for (int k = 0; k < 12; ++k) v[k] = cadd(v[k], cscale(gamma, acc[k]));
  SpinorIO<S>::store(out, s, g.V, v);
}

template <typename S>
__global__ void __launch_bounds__(HOP_THREADS)
k_zero_blocks(Geo g, const int* list, int wg_per_blk, int nrhs, long long stride, const int* active, S* out) {
  using R = typename Store<S>::R;
  const int bi = blockIdx.x / wg_per_blk;
  const int blk = list ? list[bi] : bi;
  const int loc = (blockIdx.x - bi * wg_per_blk) * HOP_THREADS + threadIdx.x;
  if (loc >= g.Vb) return;
  const long long s = (long long)blk * g.Vb + loc;
  Cx<R> v[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) v[k] = {0, 0};
  for (int r = 0; r < nrhs; ++r)
    if (!active || active[r]) SpinorIO<S>::store(out + r * stride, s, g.V, v);
}

template <typename S, int NL>
void launch_hop(const Geo& g, BlockSel sel, const Batch& B, int set, typename Store<S>::R gamma, S* out, const S* a,
                const S* b, const S* c, const S* in, const S* U, const S* C, hipStream_t st) {
  if (set == HOPS_ALL && sel.n == g.nblk && a == in && b == nullptr && c == nullptr && out != in) {
    // the full operator: one lean launch per right-hand side
    const int wpb = (g.Vb + HOP_THREADS - 1) / HOP_THREADS;
    for (int r = 0; r < B.n; ++r)
      hipLaunchKernelGGL((k_dslash<S, NL>), dim3(sel.n * wpb), dim3(HOP_THREADS), 0, st, g, sel.list, wpb,
                         B.active ? B.active + r : nullptr, gamma, out + r * B.stride, in + r * B.stride, U, C);
    return;
  }
  const int wpb = (g.Vb + HOP_THREADS - 1) / HOP_THREADS;
  dim3 grid(sel.n * wpb), block(HOP_THREADS);
  const int skip = (set == HOPS_FACE && out == a && b == nullptr && c == nullptr) ? 1 : 0;
#define HOP_ARGS g, sel.list, wpb, B.n, B.stride, B.active, gamma, out, a, b, c, in, U, C
  switch (set) {
    case HOPS_ALL:
      hipLaunchKernelGGL((k_hop<S, NL, HOPS_ALL>), grid, block, 0, st, HOP_ARGS, skip);
      break;
    case HOPS_INSIDE:
      hipLaunchKernelGGL((k_hop<S, NL, HOPS_INSIDE>), grid, block, 0, st, HOP_ARGS, skip);
      break;
    case HOPS_FACE:
      hipLaunchKernelGGL((k_hop<S, NL, HOPS_FACE>), grid, block, 0, st, HOP_ARGS, skip);
      break;
    default:
      hipLaunchKernelGGL((k_hop<S, NL, HOPS_NONE>), grid, block, 0, st, HOP_ARGS, 0);
  }
#undef HOP_ARGS
}

template <typename S>
void launch_zero_blocks(const Geo& g, BlockSel sel, const Batch& B, S* out, hipStream_t st) {
  const int wpb = (g.Vb + HOP_THREADS - 1) / HOP_THREADS;
  hipLaunchKernelGGL((k_zero_blocks<S>), dim3(sel.n * wpb), dim3(HOP_THREADS), 0, st, g, sel.list, wpb, B.n, B.stride,
                     B.active, out);
}

#define INST_HOP(S, NL)                                                                                           \
  template void launch_hop<S, NL>(const Geo&, BlockSel, const Batch&, int, typename Store<S>::R, S*, const S*,     \
                                  const S*, const S*, const S*, const S*, const S*, hipStream_t);
INST_HOP(float, 18)
INST_HOP(float, 12)
INST_HOP(double, 18)
INST_HOP(__half, 18)
INST_HOP(__half, 12)
template void launch_zero_blocks<float>(const Geo&, BlockSel, const Batch&, float*, hipStream_t);
template void launch_zero_blocks<double>(const Geo&, BlockSel, const Batch&, double*, hipStream_t);
template void launch_zero_blocks<__half>(const Geo&, BlockSel, const Batch&, __half*, hipStream_t);

}  // namespace qws
