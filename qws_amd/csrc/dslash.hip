// dslash.hip -- hopping-term kernels: D, block-local D_in, face terms and the
// fused SAP updates, all as  out = a + b - c + gamma * T^-1 H_set(in).
//
// Reference semantics: ddd_in_s.cc:72-374 / ddd_in_d.cc:66-196 (hops inside a
// domain, then T^-1, then out = in - kappa * (...)), ddd_out_s.cc:112-626 and
// ddd_out_d.cc:84-235 (face hops of a domain, T^-1, out += factor * (...)),
// qws.cc:625-644 (full operator).  On one GPU the faces are read in place from
// the neighbouring block instead of being packed and exchanged.
#include <type_traits>

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
    for (int k = 0; k < 12; ++k) v[k] = {fmar(gamma, acc[k].re, v[k].re), fmar(gamma, acc[k].im, v[k].im)};
    SpinorIO<S>::store(out + off, s, g.V, v);
  }
}

// Full operator out = in + gamma T^-1 H in for one right-hand side over every
// site: the kernel for D itself (ddd_d_, ddd_s_ and the outer solver's D).  Written
// without the generality of k_hop (RHS loop, a/b/c terms, block lists) it keeps the
// live state of a thread small (about 90 VGPRs in FP32).  It is compiled for at most
// 3 waves per SIMD (amdgpu_waves_per_eu): the schedule that results orders the
// neighbour gathers so that FP32 32^3x64 fetches 1.3x less from HBM than the default
// schedule and runs 2% faster (the register count still admits 5 waves per SIMD).
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
  for (int k = 0; k < 12; ++k) v[k] = {fmar(gamma, acc[k].re, v[k].re), fmar(gamma, acc[k].im, v[k].im)};
  SpinorIO<S>::store(out, s, g.V, v);
}

// One hop of the LDS-staged full operator: the neighbour's spinor from LDS when it
// lies in the block, from global memory across a block face.
template <int NL, int MU, int DIR>
__device__ __forceinline__ void hop_staged(const Geo& g, long long s, const SiteC& c, int t, const float4* lx,
                                           const float* in, const float* U, Cx<float> acc[12]) {
  if (!(g.hopmask & (1 << (2 * MU + (DIR > 0 ? 0 : 1))))) return;
  bool inside;
  const long long n = neighbour(g, s, c, MU, DIR, inside);
  Cx<float> v[12], a[6], b[6], u[9];
  if (inside) {
    const int tn = t + DIR * g.sB[MU];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      const float4 q = lx[j * HOP_THREADS + tn];
      v[2 * j] = {q.x, q.y};
      v[2 * j + 1] = {q.z, q.w};
    }
  } else {
    SpinorIO<float>::load(in, n, g.V, v);
  }
  project<float, MU, DIR>(v, a);
  load_link<float, NL>(U, MU, DIR > 0 ? s : n, g.V, u);
  link_mult<float, DIR>(u, a, b);
  float tf = 2.0f;
  if (MU == 3) {
    const int tt = c.b[3] * g.B[3] + c.l[3];
    if ((DIR > 0 && tt == g.L[3] - 1) || (DIR < 0 && tt == 0)) tf = 2.0f * g.tbc;
  }
  reconstruct<float, MU, DIR>(b, tf, acc);
}

// Full operator, FP32, one workgroup per 256-site block: the block's own spinor is
// loaded once with coalesced 128-bit loads and staged in LDS; the hops inside the
// block read it there, only the hops across a block face gather from global memory.
// Each block's spinor thus crosses L2 once for its own hops instead of up to nine
// times, and the neighbour requests left are the faces (FP32 32^3x64: 1.39x the
// algorithmic HBM bytes against 1.48x for k_dslash, 6% faster).  W is the number of
// waves per SIMD the compiler schedules for (amdgpu_waves_per_eu); at about 104
// VGPRs 4 waves are resident either way.
// (Nontemporal link and clover loads were measured slower here, with more L2 misses.)
template <int NL, int W>
__global__ void __launch_bounds__(HOP_THREADS) __attribute__((amdgpu_waves_per_eu(1, W)))
k_dslash_lds(Geo g, const int* __restrict__ order, const int* __restrict__ active, float gamma, float* __restrict__ out,
             const float* __restrict__ in, const float* __restrict__ U, const float* __restrict__ C) {
  __shared__ float4 lx[6 * HOP_THREADS];
  if (active && !*active) return;
  const int t = threadIdx.x;
  const long long s = (long long)order[xcd_remap(g, blockIdx.x, gridDim.x)] * HOP_THREADS + t;
  const float4* iv = reinterpret_cast<const float4*>(in);
#pragma unroll
  for (int j = 0; j < 6; ++j) lx[j * HOP_THREADS + t] = iv[(long long)j * g.V + s];
  __syncthreads();
  SiteC sc;
  site_coords(g, s, sc);
  Cx<float> acc[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) acc[k] = {0, 0};
  hop_staged<NL, 0, +1>(g, s, sc, t, lx, in, U, acc);
  hop_staged<NL, 0, -1>(g, s, sc, t, lx, in, U, acc);
  hop_staged<NL, 1, +1>(g, s, sc, t, lx, in, U, acc);
  hop_staged<NL, 1, -1>(g, s, sc, t, lx, in, U, acc);
  hop_staged<NL, 2, +1>(g, s, sc, t, lx, in, U, acc);
  hop_staged<NL, 2, -1>(g, s, sc, t, lx, in, U, acc);
  hop_staged<NL, 3, +1>(g, s, sc, t, lx, in, U, acc);
  hop_staged<NL, 3, -1>(g, s, sc, t, lx, in, U, acc);
  clover_apply<float>(C, s, g.V, acc);
  float4* ov = reinterpret_cast<float4*>(out);
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    const float4 o = lx[j * HOP_THREADS + t];
    ov[(long long)j * g.V + s] = make_float4(fmar(gamma, acc[2 * j].re, o.x), fmar(gamma, acc[2 * j].im, o.y),
                                             fmar(gamma, acc[2 * j + 1].re, o.z), fmar(gamma, acc[2 * j + 1].im, o.w));
  }
}

template <int NL>
void launch_dslash_lds(const Geo& g, BlockSel sel, const int* active, float gamma, float* out, const float* in,
                       const float* U, const float* C, hipStream_t st) {
#define DSL_LDS(W) \
  hipLaunchKernelGGL((k_dslash_lds<NL, W>), dim3(sel.n), dim3(HOP_THREADS), 0, st, g, sel.list, active, gamma, out, in, U, C)
  if (g.dslash_kernel == 3) DSL_LDS(3);
  else DSL_LDS(4);
#undef DSL_LDS
}

// One FP64 hop of the LDS-staged full operator: the arithmetic of one_hop<double, 18>,
// the neighbour's spinor from LDS when it lies in the block.
template <int MU, int DIR>
__device__ __forceinline__ void hop_staged_d(const Geo& g, long long s, const SiteC& c, int t, const double2* lx,
                                             const double* in, const double* U, Cx<double> acc[12]) {
  if (!(g.hopmask & (1 << (2 * MU + (DIR > 0 ? 0 : 1))))) return;
  bool inside;
  const long long n = neighbour(g, s, c, MU, DIR, inside);
  Cx<double> v[12], a[6], b[6], u[9];
  if (inside) {
    const int tn = t + DIR * g.sB[MU];
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      const double2 q = lx[k * HOP_THREADS + tn];
      v[k] = {q.x, q.y};
    }
  } else {
    SpinorIO<double>::load(in, n, g.V, v);
  }
  project<double, MU, DIR>(v, a);
  load_link<double, 18>(U, MU, DIR > 0 ? s : n, g.V, u);
  link_mult<double, DIR>(u, a, b);
  double tf = 2.0;
  if (MU == 3) {
    const int tt = c.b[3] * g.B[3] + c.l[3];
    if ((DIR > 0 && tt == g.L[3] - 1) || (DIR < 0 && tt == 0)) tf = double(2 * g.tbc);
  }
  reconstruct<double, MU, DIR>(b, tf, acc);
}

// FP64 full operator (the outer solver's D, ddd_d_) on 256-site blocks with the block's
// spinor staged in LDS (12 double2 planes, 48 KB): bit for bit k_dslash<double, 18>.
template <int W>
__global__ void __launch_bounds__(HOP_THREADS) __attribute__((amdgpu_waves_per_eu(1, W)))
k_dslash_lds_d(Geo g, const int* __restrict__ order, const int* __restrict__ active, double gamma,
               double* __restrict__ out, const double* __restrict__ in, const double* __restrict__ U,
               const double* __restrict__ C) {
  __shared__ double2 lx[12 * HOP_THREADS];
  if (active && !*active) return;
  const int t = threadIdx.x;
  const long long s = (long long)order[xcd_remap(g, blockIdx.x, gridDim.x)] * HOP_THREADS + t;
  const double2* iv = reinterpret_cast<const double2*>(in);
#pragma unroll
  for (int k = 0; k < 12; ++k) lx[k * HOP_THREADS + t] = iv[(long long)k * g.V + s];
  __syncthreads();
  SiteC sc;
  site_coords(g, s, sc);
  Cx<double> acc[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) acc[k] = {0, 0};
  hop_staged_d<0, +1>(g, s, sc, t, lx, in, U, acc);
  hop_staged_d<0, -1>(g, s, sc, t, lx, in, U, acc);
  hop_staged_d<1, +1>(g, s, sc, t, lx, in, U, acc);
  hop_staged_d<1, -1>(g, s, sc, t, lx, in, U, acc);
  hop_staged_d<2, +1>(g, s, sc, t, lx, in, U, acc);
  hop_staged_d<2, -1>(g, s, sc, t, lx, in, U, acc);
  hop_staged_d<3, +1>(g, s, sc, t, lx, in, U, acc);
  hop_staged_d<3, -1>(g, s, sc, t, lx, in, U, acc);
  clover_apply<double>(C, s, g.V, acc);
  double2* ov = reinterpret_cast<double2*>(out);
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    const double2 o = lx[k * HOP_THREADS + t];
    ov[(long long)k * g.V + s] = make_double2(fmar(gamma, acc[k].re, o.x), fmar(gamma, acc[k].im, o.y));
  }
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
    if constexpr (std::is_same<S, float>::value) {
      if (g.Vb == HOP_THREADS && g.dslash_kernel != 0) {
        for (int r = 0; r < B.n; ++r)
          launch_dslash_lds<NL>(g, sel, B.active ? B.active + r : nullptr, gamma, out + r * B.stride,
                                in + r * B.stride, U, C, st);
        return;
      }
    }
    if constexpr (std::is_same<S, double>::value && NL == 18) {
      if (g.Vb == HOP_THREADS && g.dslash_kernel_d != 0) {
        for (int r = 0; r < B.n; ++r) {
          const int* act = B.active ? B.active + r : nullptr;
          if (g.dslash_kernel_d == 2)
            hipLaunchKernelGGL((k_dslash_lds_d<2>), dim3(sel.n), dim3(HOP_THREADS), 0, st, g, sel.list, act, gamma,
                               out + r * B.stride, in + r * B.stride, U, C);
          else
            hipLaunchKernelGGL((k_dslash_lds_d<3>), dim3(sel.n), dim3(HOP_THREADS), 0, st, g, sel.list, act, gamma,
                               out + r * B.stride, in + r * B.stride, U, C);
        }
        return;
      }
    }
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
