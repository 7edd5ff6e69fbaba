// layout.hip -- conversion between host field layouts (staged in device memory)
// and the device's blocked site-major SoA layout.
//
// Host layouts:
//   plain : lexicographic site index x + Lx*(y + Ly*(z + Lz*t)); per site 24
//           (spinor), 4 x 18 (links, [mu][V][18]) or 72 (clover) reals
//   QWS   : single-rank AoSoA of qws.h:111-192 with lane width vlen: site ->
//           x-domain d = x / (Lx/2), vector i = (x%nxh)/vlen + nxv*(y + ny*(z + nz*t)),
//           lane j = (x%nxh)%vlen (main.cc:426-429,485-488, reordering.cc:66-69);
//           spinor scd_t/scs_t[i + vol*d], links glud[i + vol*mu + 4*vol*d],
//           clover clvd[i + vol*d]
#include "kernels.h"

namespace qws {

struct HostMap {
  int layout, vlen, nxh, nxv;
  long long volv;  // vectors per x-domain
};

__device__ __forceinline__ void global_coords(const Geo& g, long long s, int x[4]) {
  SiteC c;
  site_coords(g, s, c);
#pragma unroll
  for (int m = 0; m < 4; ++m) x[m] = c.b[m] * g.B[m] + c.l[m];
}

// offset of real k of the per-site record (nrec reals) for direction/record r
__device__ __forceinline__ long long host_offset(const Geo& g, const HostMap& h, const int x[4], int kind, int mu,
                                                 int k) {
  if (h.layout == 0) {
    long long lex = x[0] + (long long)g.L[0] * (x[1] + (long long)g.L[1] * (x[2] + (long long)g.L[2] * x[3]));
    if (kind == 0) return lex * 24 + k;
    if (kind == 1) return ((long long)mu * g.V + lex) * 18 + k;
    return lex * 72 + k;
  }
  const int d = x[0] / h.nxh, xr = x[0] % h.nxh;
  const long long i = xr / h.vlen + (long long)h.nxv * (x[1] + (long long)g.L[1] * (x[2] + (long long)g.L[2] * x[3]));
  const int j = xr % h.vlen;
  if (kind == 0) return ((i + h.volv * d) * 24 + k) * h.vlen + j;
  if (kind == 1) return ((i + h.volv * mu + 4 * h.volv * d) * 18 + k) * h.vlen + j;
  return ((i + h.volv * d) * 72 + k) * h.vlen + j;
}

template <typename S> __device__ __forceinline__ void store_real(S* p, long long i, double v);
template <> __device__ __forceinline__ void store_real<float>(float* p, long long i, double v) { p[i] = (float)v; }
template <> __device__ __forceinline__ void store_real<double>(double* p, long long i, double v) { p[i] = v; }
template <> __device__ __forceinline__ void store_real<__half>(__half* p, long long i, double v) { p[i] = __float2half_rn((float)v); }

template <typename S> __device__ __forceinline__ double load_real(const S* p, long long i);
template <> __device__ __forceinline__ double load_real<float>(const float* p, long long i) { return p[i]; }
template <> __device__ __forceinline__ double load_real<double>(const double* p, long long i) { return p[i]; }
template <> __device__ __forceinline__ double load_real<__half>(const __half* p, long long i) { return __half2float(p[i]); }

// device offset of real k of a site record
template <typename S>
__device__ __forceinline__ long long dev_offset(const Geo& g, int kind, int nl, long long s, int mu, int k) {
  if (kind == 0) {
    if (sizeof(S) != 8) return ((long long)(k >> 2) * g.V + s) * 4 + (k & 3);  // float4 / Half4 [6][V]
    return ((long long)(k >> 1) * g.V + s) * 2 + (k & 1);                     // double2 [12][V]
  }
  if (kind == 1) return (((long long)mu * (nl / 2) + (k >> 1)) * g.V + s) * 2 + (k & 1);
  if (sizeof(S) != 8) return ((long long)(k >> 2) * g.V + s) * 4 + (k & 3);  // float4 / Half4 [18][V]
  return ((long long)(k >> 1) * g.V + s) * 2 + (k & 1);                       // double2 [36][V]
}

template <typename S>
__global__ void k_import(Geo g, HostMap h, int kind, int nl, const void* src, int src_is_double, S* dst) {
  const long long s = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (s >= g.V) return;
  int x[4];
  global_coords(g, s, x);
  const int nmu = kind == 1 ? 4 : 1;
  const int nrec = kind == 0 ? 24 : (kind == 1 ? nl : 72);
  for (int mu = 0; mu < nmu; ++mu)
    for (int k = 0; k < nrec; ++k) {
      long long o = host_offset(g, h, x, kind, mu, k);
      double v = src_is_double ? ((const double*)src)[o] : (double)((const float*)src)[o];
      store_real<S>(dst, dev_offset<S>(g, kind, nl, s, mu, k), v);
    }
}

template <typename S>
__global__ void k_export(Geo g, HostMap h, const S* src, void* dst, int dst_is_double) {
  const long long s = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (s >= g.V) return;
  int x[4];
  global_coords(g, s, x);
  for (int k = 0; k < 24; ++k) {
    double v = load_real<S>(src, dev_offset<S>(g, 0, 0, s, 0, k));
    long long o = host_offset(g, h, x, 0, 0, k);
    if (dst_is_double) ((double*)dst)[o] = v;
    else ((float*)dst)[o] = (float)v;
  }
}

static HostMap make_map(const Geo& g, int layout, int vlen) {
  HostMap h;
  h.layout = layout;
  h.vlen = vlen > 0 ? vlen : 1;
  h.nxh = g.L[0] / 2;
  h.nxv = h.nxh / h.vlen;
  h.volv = (long long)h.nxv * g.L[1] * g.L[2] * g.L[3];
  return h;
}

template <typename S>
void launch_import(const Geo& g, int kind, int layout, int vlen, int nl, const void* src, int src_is_double, S* dst,
                   hipStream_t st) {
  HostMap h = make_map(g, layout, vlen);
  const int T = 128;
  hipLaunchKernelGGL((k_import<S>), dim3((unsigned)((g.V + T - 1) / T)), dim3(T), 0, st, g, h, kind, nl, src,
                     src_is_double, dst);
}

template <typename S>
void launch_export(const Geo& g, int kind, int layout, int vlen, const S* src, void* dst, int dst_is_double,
                   hipStream_t st) {
  (void)kind;
  HostMap h = make_map(g, layout, vlen);
  const int T = 128;
  hipLaunchKernelGGL((k_export<S>), dim3((unsigned)((g.V + T - 1) / T)), dim3(T), 0, st, g, h, src, dst,
                     dst_is_double);
}

template void launch_import<float>(const Geo&, int, int, int, int, const void*, int, float*, hipStream_t);
template void launch_import<double>(const Geo&, int, int, int, int, const void*, int, double*, hipStream_t);
template void launch_import<__half>(const Geo&, int, int, int, int, const void*, int, __half*, hipStream_t);
template void launch_export<float>(const Geo&, int, int, int, const float*, void*, int, hipStream_t);
template void launch_export<double>(const Geo&, int, int, int, const double*, void*, int, hipStream_t);
template void launch_export<__half>(const Geo&, int, int, int, const __half*, void*, int, hipStream_t);

}  // namespace qws

namespace qws {

// FP64 links (kind 1, 18 reals) or clover (kind 2) -> plain host layout (staged)
__global__ void k_export_field(Geo g, HostMap h, int kind, const double* src, double* dst) {
  const long long s = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (s >= g.V) return;
  int x[4];
  global_coords(g, s, x);
  const int nmu = kind == 1 ? 4 : 1;
  const int nrec = kind == 1 ? 18 : 72;
  for (int mu = 0; mu < nmu; ++mu)
    for (int k = 0; k < nrec; ++k)
      dst[host_offset(g, h, x, kind, mu, k)] = src[dev_offset<double>(g, kind, 18, s, mu, k)];
}

void launch_export_field(const Geo& g, int kind, const double* src, double* dst, hipStream_t st) {
  HostMap h = make_map(g, 0, 1);
  const int T = 128;
  hipLaunchKernelGGL(k_export_field, dim3((unsigned)((g.V + T - 1) / T)), dim3(T), 0, st, g, h, kind, src, dst);
}

}  // namespace qws

namespace qws {

// BQCD's arrays <-> QWS's single-rank host AoSoA images held in device memory
// (qws_bqcd.cc:66-647), one thread per lexicographic site lex = x + Lx (y + Ly (z +
// Lz t)).  map[lex] = 2 (i - 1) + eo, with i BQCD's 1-based index of the site within
// its parity eo (std_xyzt2i_, e_o_), tabulated once on the host.  BQCD arrays:
// links [mu][eo][i][18], clover [eo][i][72] (QWS real k of chirality ud at BQCD
// position clover_perm[k] + 36 ud, qws_bqcd.cc:301-337), spinors [eo][i][24]; every
// record but the clover's is in QWS's own element order.  The QWS image of lane width
// vlen: DD reading (x split into halves, slot = x / (Lx/2)) or, with eo_layout, the
// even/odd reading (x compressed by 2, slot = the parity; spinors: even sites only).

template <typename T>
__global__ void k_bqcd(BqcdGeo b, int kind, int load, const int* __restrict__ map, double* bq, T* img, int vlen) {
  const long long lex = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long V = (long long)b.L[0] * b.L[1] * b.L[2] * b.L[3];
  if (lex >= V) return;
  long long r = lex;
  int x[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    x[m] = (int)(r % b.L[m]);
    r /= b.L[m];
  }
  const int code = map[lex];
  const long long i0 = code >> 1;
  const int eo = code & 1;
  if (kind == 0 && b.eo_layout && eo) return;  // the even/odd spinor loaders move the even sites only
  const int nxh = b.L[0] / 2, nxv = nxh / vlen;
  const int xx = b.eo_layout ? x[0] / 2 : x[0] % nxh;
  const int slot = b.eo_layout ? eo : x[0] / nxh;
  const long long vv = (long long)nxv * b.L[1] * b.L[2] * b.L[3];  // vectors per domain slot
  const long long iv = xx / vlen + (long long)nxv * (x[1] + (long long)b.L[1] * (x[2] + (long long)b.L[2] * x[3]));
  const int lane = xx % vlen;
  const int n = kind == 0 ? 24 : (kind == 1 ? 18 : 72);
  const int nmu = kind == 1 ? 4 : 1;
  for (int mu = 0; mu < nmu; ++mu) {
    const long long bqb = kind == 1 ? 18 * (i0 + b.volh * eo) + 36 * b.volh * mu : n * (i0 + b.volh * eo);
    const long long qb = kind == 1 ? (iv + vv * mu + 4 * vv * slot) * 18 : (iv + vv * slot) * n;
    for (int k = 0; k < n; ++k) {
      const int kb = kind == 2 ? bqcd_clover_index(k % 36) + 36 * (k / 36) : k;
      const long long qi = (qb + k) * vlen + lane;
      if (load) img[qi] = (T)bq[bqb + kb];
      else bq[bqb + kb] = (double)img[qi];
    }
  }
}

template <typename T>
void launch_bqcd(const BqcdGeo& b, int kind, int load, const int* map, double* bq, T* img, int vlen,
                 hipStream_t st) {
  const long long V = (long long)b.L[0] * b.L[1] * b.L[2] * b.L[3];
  const int NT = 128;
  hipLaunchKernelGGL((k_bqcd<T>), dim3((unsigned)((V + NT - 1) / NT)), dim3(NT), 0, st, b, kind, load, map, bq, img,
                     vlen);
}

template void launch_bqcd<double>(const BqcdGeo&, int, int, const int*, double*, double*, int, hipStream_t);
template void launch_bqcd<float>(const BqcdGeo&, int, int, const int*, double*, float*, int, hipStream_t);

}  // namespace qws
