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
// belong to the other colour) and Vb == SAP_THREADS sites per block.
#include "kernels.h"
#include "sap.h"

namespace qws {

template <int NL>
__device__ __forceinline__ void load_link_lds(const float* lU, int mu, int t, Cx<float> u[9]) {
  constexpr int NC = NL / 2;
#pragma unroll
  for (int k = 0; k < NC; ++k) u[k] = {lU[((mu * NC + k) * 2 + 0) * SAP_THREADS + t], lU[((mu * NC + k) * 2 + 1) * SAP_THREADS + t]};
  if (NL == 12) {
    Cx<float> a = cmul(u[1], u[5]), b = cmul(u[2], u[4]);
    u[6] = cconj(csub(a, b));
    a = cmul(u[2], u[3]); b = cmul(u[0], u[5]);
    u[7] = cconj(csub(a, b));
    a = cmul(u[0], u[4]); b = cmul(u[1], u[3]);
    u[8] = cconj(csub(a, b));
  }
}

__device__ __forceinline__ void lds_load_spinor(const float* lx, int t, Cx<float> v[12]) {
#pragma unroll
  for (int k = 0; k < 12; ++k) v[k] = {lx[(2 * k) * SAP_THREADS + t], lx[(2 * k + 1) * SAP_THREADS + t]};
}
__device__ __forceinline__ void lds_store_spinor(float* lx, int t, const Cx<float> v[12]) {
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    lx[(2 * k) * SAP_THREADS + t] = v[k].re;
    lx[(2 * k + 1) * SAP_THREADS + t] = v[k].im;
  }
}

// in-block hop along MU, DIR from LDS
template <int NL, int MU, int DIR>
__device__ __forceinline__ void lds_hop(const Geo& g, const SiteC& c, int t, const float* lx, const float* lU,
                                        Cx<float> acc[12]) {
  const int l = c.l[MU] + DIR;
  if (l < 0 || l >= g.B[MU]) return;
  const int tn = t + DIR * g.sB[MU];
  Cx<float> v[12], a[6], b[6], u[9];
  lds_load_spinor(lx, tn, v);
  project<float, MU, DIR>(v, a);
  load_link_lds<NL>(lU, MU, DIR > 0 ? t : tn, u);
  link_mult<float, DIR>(u, a, b);
  reconstruct<float, MU, DIR>(b, 2.0f, acc);
}

template <int NL>
__device__ __forceinline__ void lds_hop_sum(const Geo& g, const SiteC& c, int t, const float* lx, const float* lU,
                                            Cx<float> acc[12]) {
#pragma unroll
  for (int k = 0; k < 12; ++k) acc[k] = {0, 0};
  lds_hop<NL, 0, +1>(g, c, t, lx, lU, acc);
  lds_hop<NL, 0, -1>(g, c, t, lx, lU, acc);
  lds_hop<NL, 1, +1>(g, c, t, lx, lU, acc);
  lds_hop<NL, 1, -1>(g, c, t, lx, lU, acc);
  lds_hop<NL, 2, +1>(g, c, t, lx, lU, acc);
  lds_hop<NL, 2, -1>(g, c, t, lx, lU, acc);
  lds_hop<NL, 3, +1>(g, c, t, lx, lU, acc);
  lds_hop<NL, 3, -1>(g, c, t, lx, lU, acc);
}

// clover inverse held in registers: a[0..35] chirality 0, a[36..71] chirality 1
__device__ __forceinline__ void clover_regs(const float cl[72], Cx<float> v[12]) {
  Cx<float> x[6], y0[6], y1[6];
#pragma unroll
  for (int c = 0; c < 3; ++c) { x[c] = cadd(v[c * 4 + 0], v[c * 4 + 2]); x[3 + c] = cadd(v[c * 4 + 1], v[c * 4 + 3]); }
  herm6_mult(cl, x, y0);
#pragma unroll
  for (int c = 0; c < 3; ++c) { x[c] = csub(v[c * 4 + 0], v[c * 4 + 2]); x[3 + c] = csub(v[c * 4 + 1], v[c * 4 + 3]); }
  herm6_mult(cl + 36, x, y1);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    v[c * 4 + 0] = cadd(y0[c], y1[c]);
    v[c * 4 + 2] = csub(y0[c], y1[c]);
    v[c * 4 + 1] = cadd(y0[3 + c], y1[3 + c]);
    v[c * 4 + 3] = csub(y0[3 + c], y1[3 + c]);
  }
}

template <int NL>
__global__ void __launch_bounds__(SAP_THREADS, 1)
k_sap_block(Geo g, const int* __restrict__ list, int nrhs, long long stride, const int* __restrict__ active, float kappa,
            int nm, int flags, float* s, const float* __restrict__ in, const float* qface, float* qout, float* out,
            const float* __restrict__ U, const float* __restrict__ C) {
  constexpr int NC = NL / 2;
  __shared__ float lU[4 * NC * 2 * SAP_THREADS];
  __shared__ float lx[24 * SAP_THREADS];
  const int t = threadIdx.x;
  const int blk = list[xcd_remap(g, blockIdx.x, gridDim.x)];
  const long long site = (long long)blk * SAP_THREADS + t;
  SiteC c;
  site_coords(g, site, c);

  // forward links of this site into LDS (coalesced [mu][k][V] reads), once for all RHS
  {
    const float2* Ug = reinterpret_cast<const float2*>(U);
#pragma unroll
    for (int mu = 0; mu < 4; ++mu)
#pragma unroll
      for (int k = 0; k < NC; ++k) {
        float2 v = Ug[((long long)mu * NC + k) * g.V + site];
        lU[((mu * NC + k) * 2 + 0) * SAP_THREADS + t] = v.x;
        lU[((mu * NC + k) * 2 + 1) * SAP_THREADS + t] = v.y;
      }
  }
  float cl[72];
  CloverIO<float>::load_chir(C, site, g.V, 0, cl);
  CloverIO<float>::load_chir(C, site, g.V, 1, cl + 36);
  const int epi = flags & SAP_EPI_MASK;

  for (int r = 0; r < nrhs; ++r) {
    if (active && !active[r]) continue;  // uniform over the workgroup
    const long long off = r * stride;
    Cx<float> sv[12], acc[12], f[12];
    SpinorIO<float>::load(((flags & SAP_SRC_IN) ? in : s) + off, site, g.V, sv);

    if (flags & SAP_FACE_PRO) {
      hop_sum<float, NL, HOPS_FACE>(g, site, c, qface + off, opaque(U), f);
      clover_regs(cl, f);
#pragma unroll
      for (int k = 0; k < 12; ++k) sv[k] = cadd(sv[k], cscale(kappa, f[k]));
    }

    // block solve: x_0 = s_c ; x_{k+1} = s_c + kappa T^-1 H_in x_k
    Cx<float> x[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) x[k] = sv[k];
    __syncthreads();  // the previous RHS's readers of lx are done (and lU is complete)
    lds_store_spinor(lx, t, x);
    __syncthreads();
    for (int it = 0; it < nm; ++it) {
      lds_hop_sum<NL>(g, c, t, lx, lU, acc);
      clover_regs(cl, acc);
#pragma unroll
      for (int k = 0; k < 12; ++k) x[k] = cadd(sv[k], cscale(kappa, acc[k]));
      __syncthreads();
      lds_store_spinor(lx, t, x);
      __syncthreads();
    }
    SpinorIO<float>::store(qout + off, site, g.V, x);

    if (epi == SAP_EPI_NONE) continue;
    // acc = T^-1 H_in x
    lds_hop_sum<NL>(g, c, t, lx, lU, acc);
    clover_regs(cl, acc);
    Cx<float> o[12];
    if (epi == SAP_EPI_SUPDATE) {
      // s_c = s_c + in_c - x + kappa T^-1 H_in x
      Cx<float> iv[12];
      SpinorIO<float>::load(in + off, site, g.V, iv);
#pragma unroll
      for (int k = 0; k < 12; ++k) o[k] = cadd(csub(cadd(sv[k], iv[k]), x[k]), cscale(kappa, acc[k]));
      SpinorIO<float>::store(s + off, site, g.V, o);
    } else {
      // out_c = [- kappa f] + x - kappa T^-1 H_in x
#pragma unroll
      for (int k = 0; k < 12; ++k) o[k] = csub(x[k], cscale(kappa, acc[k]));
      if (epi == SAP_EPI_OUT_FACE) {
#pragma unroll
        for (int k = 0; k < 12; ++k) o[k] = csub(o[k], cscale(kappa, f[k]));
      }
      SpinorIO<float>::store(out + off, site, g.V, o);
    }
  }
}

bool sap_fusable(const Geo& g) { return g.Vb == SAP_THREADS; }

void launch_sap_block(const Geo& g, BlockSel sel, const Batch& B, int nl, float kappa, int nm, int flags, float* s,
                      const float* in, const float* qface, float* qout, float* out, const float* U, const float* C,
                      hipStream_t st) {
  if (sel.n == 0) return;
  if (nm < 1) nm = 1;  // the first application is unconditional (static_solver.cc:77-101)
#define SAP_ARGS g, sel.list, B.n, B.stride, B.active, kappa, nm, flags, s, in, qface, qout, out, U, C
  if (nl == 12)
    hipLaunchKernelGGL((k_sap_block<12>), dim3(sel.n), dim3(SAP_THREADS), 0, st, SAP_ARGS);
  else
    hipLaunchKernelGGL((k_sap_block<18>), dim3(sel.n), dim3(SAP_THREADS), 0, st, SAP_ARGS);
#undef SAP_ARGS
}

}  // namespace qws
