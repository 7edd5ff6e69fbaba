// mfma_trial.hip -- the multi-RHS clover product on the matrix cores against the
// vector ALUs, measured in isolation (qwsg_time_clover_rhs).
//
// The clover inverse is applied per site as two Hermitian 6x6 blocks, one per
// chirality, in QWS's packing (6 real diagonal + 15 complex upper entries,
// clover_def.h:88-284).  With R right-hand sides sharing the site's matrix the product
// is a small GEMM: in real form y = A x with the 12x12 A = [[Re, -Im], [Im, Re]] of the
// block and x a 12 x R matrix.  Both kernels here compute y_{n+1} = fp16(A y_n) nrep
// times from FP16 storage with FP32 accumulation, on data they keep in registers (the
// SAP's Jacobi steps apply the clover to data in registers and LDS, so the trial is
// arithmetic-bound as the SAP kernel is):
//   k_clover_rhs_valu: one lane per (site, RHS), 64 lanes = 4 sites x 16 RHS, FP32 FMAs;
//   k_clover_rhs_mfma: one wave per site at a time (a grid-stride loop over the sites),
//     v_mfma_f32_16x16x16f16 with A = the site's
//     block (12x12 padded to 16x16), B = its 16 RHS (K = 12 of 16): lane l holds
//     A[l % 16][4 (l / 16) .. +3] and B[4 (l / 16) .. +3][l % 16], and receives
//     D[4 (l / 16) .. +3][l % 16] -- the B layout of the next product.
// Layouts: clover [site][2][36] FP16; spinors [site][RHS 16][2 chiralities][12] FP16,
// the 12 reals of a chirality block being Re x_0..x_5, Im x_0..x_5.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>

#include "lattice.h"

namespace qws {

namespace {

typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int NR = 16;  // right-hand sides

// packed position of the Hermitian entry (i, j) of one 6x6 block and whether it is
// conjugated: diagonal i -> i (real), upper (i < j) -> 6 + 2 p(i, j) (re, im)
__host__ __device__ inline int herm_pair(int i, int j) {
  // p = index of (i, j), i < j, in the order (0,1) (0,2) .. (0,5) (1,2) .. (4,5)
  return i * 5 - i * (i - 1) / 2 + (j - i - 1);
}

// The A operand entry this lane supplies, A[row][k]: packed position and sign, or -1
// for the zero padding; the same for every site.
__device__ __forceinline__ void real_form_entry(int row, int k, int& pos, float& sign) {
  pos = -1, sign = 1.0f;
  if (row >= 12 || k >= 12) return;
  const int i = row % 6, j = k % 6;
  const bool im_row = row >= 6, im_col = k >= 6;
  const bool re_part = im_row == im_col;  // [[Re, -Im], [Im, Re]]
  if (i == j) {
    if (re_part) pos = i;  // the diagonal is real: its imaginary blocks vanish
    return;
  }
  const int p = 6 + 2 * (i < j ? herm_pair(i, j) : herm_pair(j, i));
  pos = re_part ? p : p + 1;
  // imaginary part of A[i][j]: +im for i < j, -im below the diagonal (conjugate); the
  // upper-right block carries a minus sign
  float s_im = i < j ? 1.0f : -1.0f;
  if (!re_part && !im_row) s_im = -s_im;
  sign = re_part ? 1.0f : s_im;
}

// One wave per site at a time, looping over the sites (grid-stride): the lanes' A
// entries are mapped once, then per site 8 half loads for A, one 8-byte load per
// chirality for B, 2 MFMAs per repetition, one 8-byte store per chirality.
__global__ void __launch_bounds__(256) k_clover_rhs_mfma(long long V, int nrep, const __half* __restrict__ C,
                                                         const __half* __restrict__ x, __half* __restrict__ y) {
  const int lane = threadIdx.x & 63;
  const int r = lane & 15, g = lane >> 4;
  int pos[4];
  float sgn[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) real_form_entry(r, 4 * g + j, pos[j], sgn[j]);
  const long long nw = (long long)gridDim.x * (blockDim.x / 64);
  for (long long s = (long long)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6); s < V; s += nw) {
    h4 a[2], b[2];
#pragma unroll
    for (int chi = 0; chi < 2; ++chi) {
      const __half* c = C + (s * 2 + chi) * 36;
#pragma unroll
      for (int j = 0; j < 4; ++j) a[chi][j] = pos[j] >= 0 ? (_Float16)(sgn[j] * __half2float(c[pos[j]])) : (_Float16)0.0f;
      if (g < 3) b[chi] = *reinterpret_cast<const h4*>(x + ((s * NR + r) * 2 + chi) * 12 + 4 * g);
      else b[chi] = h4{0, 0, 0, 0};
    }
    for (int n = 0; n < nrep; ++n) {
#pragma unroll
      for (int chi = 0; chi < 2; ++chi) {
        const f4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
        const f4 d = __builtin_amdgcn_mfma_f32_16x16x16f16(a[chi], b[chi], zero, 0, 0, 0);
#pragma unroll
        for (int j = 0; j < 4; ++j) b[chi][j] = (_Float16)d[j];
      }
    }
    if (g < 3) {
#pragma unroll
      for (int chi = 0; chi < 2; ++chi) *reinterpret_cast<h4*>(y + ((s * NR + r) * 2 + chi) * 12 + 4 * g) = b[chi];
    }
  }
}

__global__ void __launch_bounds__(256) k_clover_rhs_valu(long long V, int nrep, const __half* __restrict__ C,
                                                         const __half* __restrict__ x, __half* __restrict__ y) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long s = t / NR;
  const int r = (int)(t % NR);
  if (s >= V) return;
  float cre[2][6][6], cim[2][6][6];  // the blocks, unpacked once (registers)
#pragma unroll
  for (int chi = 0; chi < 2; ++chi) {
    const __half* c = C + (s * 2 + chi) * 36;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      cre[chi][i][i] = __half2float(c[i]);
      cim[chi][i][i] = 0.0f;
#pragma unroll
      for (int j = i + 1; j < 6; ++j) {
        const int p = 6 + 2 * herm_pair(i, j);
        cre[chi][i][j] = cre[chi][j][i] = __half2float(c[p]);
        cim[chi][i][j] = __half2float(c[p + 1]);
        cim[chi][j][i] = -cim[chi][i][j];
      }
    }
  }
  float vr[2][6], vi[2][6];
#pragma unroll
  for (int chi = 0; chi < 2; ++chi) {
    const __half* xs = x + ((s * NR + r) * 2 + chi) * 12;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      vr[chi][i] = __half2float(xs[i]);
      vi[chi][i] = __half2float(xs[6 + i]);
    }
  }
  for (int n = 0; n < nrep; ++n) {
#pragma unroll
    for (int chi = 0; chi < 2; ++chi) {
      float wr[6], wi[6];
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        float ar = 0.0f, ai = 0.0f;
#pragma unroll
        for (int j = 0; j < 6; ++j) {
          ar = __builtin_fmaf(cre[chi][i][j], vr[chi][j], ar);
          ar = __builtin_fmaf(-cim[chi][i][j], vi[chi][j], ar);
          ai = __builtin_fmaf(cim[chi][i][j], vr[chi][j], ai);
          ai = __builtin_fmaf(cre[chi][i][j], vi[chi][j], ai);
        }
        wr[i] = ar, wi[i] = ai;
      }
#pragma unroll
      for (int i = 0; i < 6; ++i) {  // stored in FP16, as the MFMA kernel's operands
        vr[chi][i] = __half2float(__float2half(wr[i]));
        vi[chi][i] = __half2float(__float2half(wi[i]));
      }
    }
  }
#pragma unroll
  for (int chi = 0; chi < 2; ++chi) {
    __half* ys = y + ((s * NR + r) * 2 + chi) * 12;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      ys[i] = __float2half(vr[chi][i]);
      ys[6 + i] = __float2half(vi[chi][i]);
    }
  }
}

// inputs: a Hermitian block 0.5 * (1 + small) per chirality (the size of QWS's 1/2 T^-1,
// so that repeated products stay in range), spinors of order 1
__global__ void k_clover_rhs_init(long long V, __half* C, __half* x) {
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= V) return;
  auto hash = [](unsigned long long v) {
    v ^= v >> 33;
    v *= 0xff51afd7ed558ccdULL;
    v ^= v >> 33;
    v *= 0xc4ceb9fe1a85ec53ULL;
    v ^= v >> 33;
    return (float)(v & 0xffffff) / 16777216.0f - 0.5f;  // [-0.5, 0.5)
  };
  for (int k = 0; k < 72; ++k) {
    const int kk = k % 36;
    const float v = kk < 6 ? 0.5f + 0.02f * hash(s * 72 + k) : 0.02f * hash(s * 72 + k);
    C[s * 72 + k] = __float2half(v);
  }
  for (int k = 0; k < NR * 24; ++k) x[s * NR * 24 + k] = __float2half(2.0f * hash(0x9e3779b97f4a7c15ULL + s * 384 + k));
}

__global__ void k_max_diff(long long n, const __half* a, const __half* b, float* out) {
  float m = 0.0f, s = 0.0f;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float u = __half2float(a[i]), v = __half2float(b[i]);
    m = fmaxf(m, fabsf(u - v));
    s = fmaxf(s, fabsf(v));
  }
  atomicMax(reinterpret_cast<int*>(out), __float_as_int(m));
  atomicMax(reinterpret_cast<int*>(out) + 1, __float_as_int(s));
}

}  // namespace

// Times `reps` launches of each kernel on V sites x 16 RHS (nrep products per launch);
// ms[0] VALU, ms[1] MFMA per launch; *maxdiff = max |y_valu - y_mfma| / max |y_valu|.
void time_clover_rhs(long long V, int nrep, int reps, double ms[2], double* maxdiff) {
  if (V < 1 || V > (1LL << 24) || nrep < 1 || reps < 1) throw std::invalid_argument("time_clover_rhs: bad sizes");
  hipStream_t st;
  QWS_HIP(hipStreamCreate(&st));
  __half *C = nullptr, *x = nullptr, *y0 = nullptr, *y1 = nullptr;
  float* d = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto cleanup = [&] {
    for (void* p : {(void*)C, (void*)x, (void*)y0, (void*)y1, (void*)d})
      if (p) (void)hipFree(p);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
  };
  try {
    const size_t nx = (size_t)V * NR * 24;
    QWS_HIP(hipMalloc((void**)&C, sizeof(__half) * V * 72));
    QWS_HIP(hipMalloc((void**)&x, sizeof(__half) * nx));
    QWS_HIP(hipMalloc((void**)&y0, sizeof(__half) * nx));
    QWS_HIP(hipMalloc((void**)&y1, sizeof(__half) * nx));
    QWS_HIP(hipMalloc((void**)&d, 2 * sizeof(float)));
    QWS_HIP(hipEventCreate(&e0));
    QWS_HIP(hipEventCreate(&e1));
    hipLaunchKernelGGL(k_clover_rhs_init, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, st, V, C, x);
    const dim3 gv((unsigned)((V * NR + 255) / 256)), gm((unsigned)std::min<long long>((V + 3) / 4, 4096));
    for (int which = 0; which < 2; ++which) {
      __half* y = which ? y1 : y0;
      auto launch = [&] {
        if (which) hipLaunchKernelGGL(k_clover_rhs_mfma, gm, dim3(256), 0, st, V, nrep, C, x, y);
        else hipLaunchKernelGGL(k_clover_rhs_valu, gv, dim3(256), 0, st, V, nrep, C, x, y);
      };
      launch();  // warm-up
      QWS_HIP(hipEventRecord(e0, st));
      for (int i = 0; i < reps; ++i) launch();
      QWS_HIP(hipEventRecord(e1, st));
      QWS_HIP(hipEventSynchronize(e1));
      float t = 0;
      QWS_HIP(hipEventElapsedTime(&t, e0, e1));
      ms[which] = t / reps;
    }
    QWS_HIP(hipMemsetAsync(d, 0, 2 * sizeof(float), st));
    hipLaunchKernelGGL(k_max_diff, dim3(1024), dim3(256), 0, st, (long long)nx, y1, y0, d);
    float h[2];
    QWS_HIP(hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, st));
    QWS_HIP(hipStreamSynchronize(st));
    *maxdiff = h[1] > 0 ? h[0] / h[1] : h[0];
  } catch (...) {
    cleanup();
    throw;
  }
  cleanup();
}

}  // namespace qws
