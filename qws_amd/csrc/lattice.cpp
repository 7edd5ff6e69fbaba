// lattice.cpp -- host side of the device solver: field loading, the SAP step
// sequences, the FP32 inner and FP64 outer BiCGStab (batched over right-hand
// sides), true residual.
#include "lattice.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <type_traits>

#include "fieldgen.h"
#include "hostcopy.h"
#include "multi.h"
#include "sap.h"

namespace qws {

namespace {
template <typename T> T* dalloc(size_t n) {
  void* p = nullptr;
  QWS_HIP(hipMalloc(&p, n * sizeof(T)));
  return (T*)p;
}
Batch single() { return Batch{1, 0, nullptr}; }

}  // namespace

// QWS_MI355X_TRACE=1: print the residual history to stderr the way the reference's
// solvers do -- "@D <rhs> <outer half-step> <sqrt(|r|^2/|b|^2)>" for the FP64 outer
// solve (bicgstab_dd_mix.cc:162,244,331) and "@C <rhs> <iterations> <residual>" at the
// end of each inner solve (bicgstab_precdd_s.cc:594).  Reading the residuals costs a
// host synchronisation per half-step; off by default.
bool trace_on() {
  static const bool on = [] {
    const char* t = std::getenv("QWS_MI355X_TRACE");
    return t && t[0] == '1';
  }();
  return on;
}

Lattice::Lattice(const int L[4], const int B[4], int tbc, int links32) {
  if (links32 != 18 && links32 != 12) throw std::invalid_argument("links32 must be 18 or 12");
  if (tbc != 1 && tbc != -1) throw std::invalid_argument("tbc must be +1 or -1");
  long long V = 1;
  int Vb = 1, nblk = 1;
  improper_ = false;
  for (int m = 0; m < 4; ++m) {
    if (L[m] <= 0 || B[m] <= 0 || L[m] % B[m] != 0)
      throw std::invalid_argument("block extents must divide the lattice extents");
    g_.L[m] = L[m];
    g_.B[m] = B[m];
    g_.NB[m] = L[m] / B[m];
    V *= L[m];
    Vb *= B[m];
    nblk *= g_.NB[m];
    if (m == 0 && g_.NB[m] % 2 != 0) throw std::invalid_argument("the x direction needs an even number of blocks");
    if (m > 0 && g_.NB[m] != 1 && g_.NB[m] % 2 != 0)
      throw std::invalid_argument("y/z/t need an even number of blocks or a single block");
    if (m > 0 && g_.NB[m] == 1) improper_ = true;  // QWS's npe[mu]==1 branch (qws.cc:805)
  }
  if (V >= (1LL << 31)) throw std::invalid_argument("lattice too large (site index must fit in 31 bits)");
  g_.sB[0] = 1;
  g_.sN[0] = 1;
  for (int m = 1; m < 4; ++m) {
    g_.sB[m] = g_.sB[m - 1] * g_.B[m - 1];
    g_.sN[m] = g_.sN[m - 1] * g_.NB[m - 1];
  }
  g_.Vb = Vb;
  g_.nblk = nblk;
  g_.V = V;
  g_.tbc = tbc;
  g_.hopmask = 0xFF;
  // workgroup placement: XCD-contiguous block ranges unless QWS_MI355X_XCD_REMAP=0
  const char* xr = std::getenv("QWS_MI355X_XCD_REMAP");
  g_.xcd_remap = (xr && xr[0] == '0') ? 0 : 1;
  // full-operator kernel (FP32, 256-site blocks): LDS-staged, compiled for 4 waves per
  // SIMD, unless QWS_MI355X_DSLASH names another (0, 3, 4); autotune_dslash() times them
  const char* fbv = std::getenv("QWS_MI355X_SAP_FB");
  use_fb_ = !(fbv && fbv[0] == '0');
  // the inner BiCGStab's direction update inside the first SAP sweep unless QWS_MI355X_SAP_PUPD=0
  const char* pu = std::getenv("QWS_MI355X_SAP_PUPD");
  fold_pupd_ = !(pu && pu[0] == '0');
  const char* dk = std::getenv("QWS_MI355X_DSLASH");
  g_.dslash_kernel = dk ? std::atoi(dk) : 4;
  if (g_.dslash_kernel != 0 && g_.dslash_kernel != 3) g_.dslash_kernel = 4;
  // FP64 full operator (18-real links, 256-site blocks): LDS-staged for 2 waves per SIMD
  // (32^3x64: 0.82 ms against 0.96-1.02 ms with global gathers, bit-identical), unless
  // QWS_MI355X_DSLASH_D names another (0 gathers, 3 LDS-staged for 3 waves)
  const char* dkd = std::getenv("QWS_MI355X_DSLASH_D");
  g_.dslash_kernel_d = dkd ? std::atoi(dkd) : 2;
  if (g_.dslash_kernel_d != 0 && g_.dslash_kernel_d != 3) g_.dslash_kernel_d = 2;
  nl32_ = links32;
  pin_chunk_ = pin_chunk_from_env();
  QWS_HIP(hipStreamCreate(&st_));
  own_st_ = st_;
  // Launch order of the blocks: index order.  (A Morton order of the block
  // coordinates was measured slower for D on 32^3x64: consecutive workgroups then
  // read scattered memory.)
  std::vector<int> order(nblk);
  for (int b = 0; b < nblk; ++b) order[b] = b;
  d_order_ = dalloc<int>(nblk + 1);
  QWS_HIP(hipMemcpy(d_order_, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice));
  {
    // XCD x runs the workgroups [x q + min(x, r), ...) of a launch over every block
    // (xcd_remap); inside each such range sort the blocks with t fastest, then z, x, y
    std::vector<int> pen(order);
    const int q = nblk / NUM_XCD, r = nblk % NUM_XCD;
    auto key = [&](int b) {
      int c[4], rem = b;
      for (int m = 0; m < 4; ++m) c[m] = rem % g_.NB[m], rem /= g_.NB[m];
      return (((long long)c[1] * g_.NB[0] + c[0]) * g_.NB[2] + c[2]) * g_.NB[3] + c[3];
    };
    for (int x = 0, lo = 0; x < NUM_XCD; ++x) {
      const int n = q + (x < r ? 1 : 0);
      std::stable_sort(pen.begin() + lo, pen.begin() + lo + n, [&](int a, int b) { return key(a) < key(b); });
      lo += n;
    }
    d_order_pencil_ = dalloc<int>(nblk + 1);
    QWS_HIP(hipMemcpy(d_order_pencil_, pen.data(), pen.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  // block lists per colour (colour = (bx+by+bz+bt)%2, qws.cc:275-276)
  std::vector<int> lists[2];
  for (int b : order) {
    int r = b, sum = 0;
    for (int m = 0; m < 4; ++m) {
      sum += r % g_.NB[m];
      r /= g_.NB[m];
    }
    lists[sum & 1].push_back(b);
  }
  for (int c = 0; c < 2; ++c) {
    h_blocks_[c] = lists[c];
    nblk_colour_[c] = (int)lists[c].size();
    d_blocks_[c] = dalloc<int>(lists[c].size() + 1);
    QWS_HIP(hipMemcpy(d_blocks_[c], lists[c].data(), lists[c].size() * sizeof(int), hipMemcpyHostToDevice));
  }
  sc_one_ = dalloc<Scalars>(1);
  QWS_HIP(hipMemset(sc_one_, 0, sizeof(Scalars)));
  d_inner_mask_ = dalloc<int>(MAX_BATCH);
  QWS_HIP(hipHostMalloc((void**)&sc_host_, sizeof(Scalars) * MAX_BATCH, hipHostMallocDefault));
}

Lattice::~Lattice() {
  (void)hipStreamSynchronize(st_);
  batch_.reset();
  for (auto* p : {(void*)Ueo_d_, (void*)Ceo_d_, (void*)Ueo_s_, (void*)Ceo_s_})
    if (p) (void)hipFree(p);
  if (!owns_gauge_) Ud_ = nullptr, Us_ = nullptr;
  for (auto* p : {(void*)Ud_, (void*)Cd_, (void*)Us_, (void*)Cs_, (void*)Uh_, (void*)Ch_, (void*)part_, (void*)sc_in_,
                  (void*)sc_one_, (void*)sc_h_, (void*)sc_eo_, (void*)d_inner_mask_, (void*)d_blocks_[0], (void*)d_blocks_[1],
                  (void*)d_order_, (void*)d_order_pencil_})
    if (p) (void)hipFree(p);
  for (auto& w : work_) (void)hipFree(w.second.first);
  if (sc_host_) (void)hipHostFree(sc_host_);
  for (int k = 0; k < 2; ++k) {
    if (pin_[k]) (void)hipHostFree(pin_[k]);
    if (pin_ev_[k]) (void)hipEventDestroy(pin_ev_[k]);
  }
  if (own_st_) (void)hipStreamDestroy(own_st_);
}



// Caller memory is pageable: it goes through two pinned chunks, the host copy of one
// chunk overlapping the DMA of the other.  Both return with the copy complete.
void Lattice::copy_to_device(void* dev, const void* host, size_t bytes) {
  for (int k = 0; k < 2; ++k) {
    if (!pin_[k]) QWS_HIP(hipHostMalloc(&pin_[k], pin_chunk_, hipHostMallocDefault));
    if (!pin_ev_[k]) QWS_HIP(hipEventCreateWithFlags(&pin_ev_[k], hipEventDisableTiming));
  }
  int k = 0;
  for (size_t off = 0; off < bytes; off += pin_chunk_, k ^= 1) {
    const size_t n = std::min(pin_chunk_, bytes - off);
    QWS_HIP(hipEventSynchronize(pin_ev_[k]));  // the DMA that last used this chunk is done
    par_memcpy(pin_[k], (const char*)host + off, n);
    QWS_HIP(hipMemcpyAsync((char*)dev + off, pin_[k], n, hipMemcpyHostToDevice, st_));
    QWS_HIP(hipEventRecord(pin_ev_[k], st_));
  }
  QWS_HIP(hipStreamSynchronize(st_));
}

void Lattice::copy_to_host(void* host, const void* dev, size_t bytes) {
  for (int k = 0; k < 2; ++k) {
    if (!pin_[k]) QWS_HIP(hipHostMalloc(&pin_[k], pin_chunk_, hipHostMallocDefault));
    if (!pin_ev_[k]) QWS_HIP(hipEventCreateWithFlags(&pin_ev_[k], hipEventDisableTiming));
  }
  const size_t nchunk = (bytes + pin_chunk_ - 1) / pin_chunk_;
  // DMA of chunk i+1 runs while chunk i is copied out of its pinned buffer
  auto dma = [&](size_t i) {
    const size_t off = i * pin_chunk_, n = std::min(pin_chunk_, bytes - off);
    QWS_HIP(hipMemcpyAsync(pin_[i & 1], (const char*)dev + off, n, hipMemcpyDeviceToHost, st_));
    QWS_HIP(hipEventRecord(pin_ev_[i & 1], st_));
  };
  if (nchunk) dma(0);
  for (size_t i = 0; i < nchunk; ++i) {
    QWS_HIP(hipEventSynchronize(pin_ev_[i & 1]));
    if (i + 1 < nchunk) dma(i + 1);
    const size_t off = i * pin_chunk_, n = std::min(pin_chunk_, bytes - off);
    par_memcpy((char*)host + off, pin_[i & 1], n);
  }
}

void* Lattice::work(const std::string& name, size_t bytes) {
  if (arena_) {  // carved from caller-provided memory: nothing is allocated
    const size_t off = (arena_off_ + 255) & ~(size_t)255;
    if (off + bytes > arena_cap_)
      throw std::invalid_argument("work buffer '" + name + "' does not fit the provided scratch memory");
    arena_off_ = off + bytes;
    return arena_ + off;
  }
  for (auto& w : work_)
    if (w.first == name) {
      if (w.second.second >= bytes) return w.second.first;
      QWS_HIP(hipStreamSynchronize(st_));
      QWS_HIP(hipFree(w.second.first));
      w.second.first = nullptr;
      QWS_HIP(hipMalloc(&w.second.first, bytes));
      w.second.second = bytes;
      return w.second.first;
    }
  void* p = nullptr;
  QWS_HIP(hipMalloc(&p, bytes));
  work_.push_back({name, {p, bytes}});
  return p;
}

double* Lattice::reduction_partials(int nrhs) {
  if (nrhs > part_rhs_) {
    QWS_HIP(hipStreamSynchronize(st_));
    if (part_) QWS_HIP(hipFree(part_));
    part_ = dalloc<double>((size_t)nrhs * 4 * RED_BLOCKS);
    part_rhs_ = nrhs;
  }
  return part_;
}

size_t Lattice::spinor_bytes(int prec) const {
  const size_t es = prec == P64 ? 8 : (prec == P32 ? 4 : 2);
  return (size_t)g_.V * 24 * es;
}

size_t Lattice::field_bytes() const {
  size_t b = 0;
  if (Ud_) b += (size_t)g_.V * 72 * 8;
  if (Cd_) b += (size_t)g_.V * 72 * 8;
  if (Us_) b += (size_t)g_.V * 4 * nl32_ * 4;
  if (Cs_) b += (size_t)g_.V * 72 * 4;
  if (Uh_) b += (size_t)g_.V * 4 * nl32_ * 2;
  if (Ch_) b += (size_t)g_.V * 72 * 2;
  return b;
}

void* Lattice::alloc_spinor(int prec) {
  void* p = nullptr;
  QWS_HIP(hipMalloc(&p, spinor_bytes(prec)));
  QWS_HIP(hipMemsetAsync(p, 0, spinor_bytes(prec), st_));
  return p;
}

// ------------------------------------------------------------------ field loading
void Lattice::load_gauge(const void* U, int layout, int is_double, int target, bool on_device) {
  if (!owns_gauge_) throw std::runtime_error("load_gauge: this lattice borrows its links");
  const size_t es = is_double ? 8 : 4;
  const size_t bytes = (size_t)4 * g_.V * 18 * es;
  const void* stage = U;
  if (!on_device) {
    void* st = work("stage", bytes);
    QWS_HIP(hipMemcpyAsync(st, U, bytes, hipMemcpyHostToDevice, st_));
    stage = st;
  }
  const int vlen = layout == 1 ? (is_double ? 8 : 16) : 1;
  if (target & 1) {
    if (!Ud_) Ud_ = dalloc<double>((size_t)4 * 18 * g_.V);
    launch_import<double>(g_, 1, layout, vlen, 18, stage, is_double, Ud_, st_);
  }
  if (target & 2) {
    if (!Us_) Us_ = dalloc<float>((size_t)4 * nl32_ * g_.V);
    launch_import<float>(g_, 1, layout, vlen, nl32_, stage, is_double, Us_, st_);
    half_valid_ = false;
  }
  QWS_HIP(hipStreamSynchronize(st_));
}

void Lattice::load_clover(const void* C, int layout, int is_double, int target, bool on_device) {
  const size_t es = is_double ? 8 : 4;
  const size_t bytes = (size_t)g_.V * 72 * es;
  const void* stage = C;
  if (!on_device) {
    void* st = work("stage", bytes);
    QWS_HIP(hipMemcpyAsync(st, C, bytes, hipMemcpyHostToDevice, st_));
    stage = st;
  }
  const int vlen = layout == 1 ? (is_double ? 8 : 16) : 1;
  if (target & 1) {
    if (!Cd_) Cd_ = dalloc<double>((size_t)72 * g_.V);
    launch_import<double>(g_, 2, layout, vlen, 0, stage, is_double, Cd_, st_);
  }
  if (target & 2) {
    if (!Cs_) Cs_ = dalloc<float>((size_t)72 * g_.V);
    launch_import<float>(g_, 2, layout, vlen, 0, stage, is_double, Cs_, st_);
    half_valid_ = false;
  }
  QWS_HIP(hipStreamSynchronize(st_));
}

void Lattice::generate_gauge(double eps, unsigned long long seed) {
  if (!owns_gauge_) throw std::runtime_error("generate_gauge: this lattice borrows its links");
  if (!Ud_) Ud_ = dalloc<double>((size_t)4 * 18 * g_.V);
  if (!Us_) Us_ = dalloc<float>((size_t)4 * nl32_ * g_.V);
  launch_gen_links(g_, eps, (uint64_t)seed, Ud_, st_);
  launch_links_d2s(g_, nl32_, Ud_, Us_, st_);
  half_valid_ = false;
  QWS_HIP(hipStreamSynchronize(st_));
}

void Lattice::share_gauge(const Lattice& src) {
  for (int m = 0; m < 4; ++m)
    if (src.g_.L[m] != g_.L[m]) throw std::invalid_argument("share_gauge: lattice extents differ");
  if (src.nl32_ != nl32_ || src.g_.tbc != g_.tbc) throw std::invalid_argument("share_gauge: link storage differs");
  if (!src.Ud_) throw std::invalid_argument("share_gauge: the source has no links");
  if (src.g_.Vb != g_.Vb || src.g_.B[0] != g_.B[0] || src.g_.B[1] != g_.B[1] || src.g_.B[2] != g_.B[2])
    throw std::invalid_argument("share_gauge: block shapes differ (the device site order follows the blocks)");
  if (owns_gauge_) {
    if (Ud_) QWS_HIP(hipFree(Ud_));
    if (Us_) QWS_HIP(hipFree(Us_));
  }
  Ud_ = src.Ud_;
  Us_ = src.Us_;
  owns_gauge_ = false;
  half_valid_ = false;
}

void Lattice::compute_clover(double csw) {
  if (!Ud_) throw std::runtime_error("compute_clover: no FP64 links loaded");
  if (!Cd_) Cd_ = dalloc<double>((size_t)72 * g_.V);
  if (!Cs_) Cs_ = dalloc<float>((size_t)72 * g_.V);
  launch_clover(g_, Ud_, kappa_ * csw, sigma_chiral_blocks(), Cd_, st_);
  launch_clover_d2s(g_, Cd_, Cs_, st_);
  half_valid_ = false;
  QWS_HIP(hipStreamSynchronize(st_));
}

void Lattice::export_field(int kind, double* host) {
  const double* src = kind == 1 ? Ud_ : Cd_;
  if (!src) throw std::runtime_error("export_field: field not present");
  const size_t n = (size_t)g_.V * 72;  // links 4 x 18, clover 72 reals per site
  double* stage = (double*)work("stage", n * sizeof(double));
  launch_export_field(g_, kind, src, stage, st_);
  QWS_HIP(hipMemcpyAsync(host, stage, n * sizeof(double), hipMemcpyDeviceToHost, st_));
  QWS_HIP(hipStreamSynchronize(st_));
}

void Lattice::gaussian_spinor(int prec, void* dev, unsigned long long seed) {
  if (prec == P64) {
    launch_gaussian_spinor(g_, (uint64_t)seed, (double*)dev, st_);
  } else if (prec == P32) {
    double* tmp = (double*)work("gen_tmp", spinor_bytes(P64));
    launch_gaussian_spinor(g_, (uint64_t)seed, tmp, st_);
    launch_convert<float, double>(g_.V, (float*)dev, tmp, st_);
  } else {
    throw std::invalid_argument("gaussian_spinor: FP64 or FP32");
  }
}

void Lattice::point_spinor(int prec, void* dev, const int x[4], int spin, int colour) {
  for (int m = 0; m < 4; ++m)
    if (x[m] < 0 || x[m] >= g_.L[m]) throw std::invalid_argument("point_spinor: site outside the lattice");
  if (spin < 0 || spin > 3 || colour < 0 || colour > 2) throw std::invalid_argument("point_spinor: spin/colour");
  if (prec == P64) {
    launch_point_spinor(g_, x, spin, colour, (double*)dev, st_);
  } else if (prec == P32) {
    double* tmp = (double*)work("gen_tmp", spinor_bytes(P64));
    launch_point_spinor(g_, x, spin, colour, tmp, st_);
    launch_convert<float, double>(g_.V, (float*)dev, tmp, st_);
  } else {
    throw std::invalid_argument("point_spinor: FP64 or FP32");
  }
}

// FP16 copies of the FP32 links and clover inverse (assign_u_s2h, assign_clv_s2h,
// qws_h.cc:230-269), made on first use after the FP32 fields change.
void Lattice::ensure_half_fields() {
  if (half_valid_) return;
  if (!Us_ || !Cs_) throw std::runtime_error("FP16 fields are made from the FP32 fields, which are not loaded");
  const long long nu = 4LL * nl32_ * g_.V, nc = 72LL * g_.V;
  if (!Uh_) Uh_ = dalloc<__half>((size_t)nu);
  if (!Ch_) Ch_ = dalloc<__half>((size_t)nc);
  launch_convert_reals<__half, float>(nu, Uh_, Us_, st_);
  launch_convert_reals<__half, float>(nc, Ch_, Cs_, st_);
  half_valid_ = true;
}

void Lattice::upload_spinor(int prec, void* dev, const void* host, int layout, int host_is_double) {
  const size_t bytes = (size_t)g_.V * 24 * (host_is_double ? 8 : 4);
  void* stage = work("stage_spinor", bytes);
  copy_to_device(stage, host, bytes);
  const int vlen = layout == 1 ? (host_is_double ? 8 : 16) : 1;
  if (prec == P64) launch_import<double>(g_, 0, layout, vlen, 0, stage, host_is_double, (double*)dev, st_);
  else if (prec == P32) launch_import<float>(g_, 0, layout, vlen, 0, stage, host_is_double, (float*)dev, st_);
  else launch_import<__half>(g_, 0, layout, vlen, 0, stage, host_is_double, (__half*)dev, st_);
  QWS_HIP(hipStreamSynchronize(st_));
}

void Lattice::download_spinor(int prec, const void* dev, void* host, int layout, int host_is_double) {
  const size_t bytes = (size_t)g_.V * 24 * (host_is_double ? 8 : 4);
  void* stage = work("stage_spinor", bytes);
  const int vlen = layout == 1 ? (host_is_double ? 8 : 16) : 1;
  if (prec == P64) launch_export<double>(g_, 0, layout, vlen, (const double*)dev, stage, host_is_double, st_);
  else if (prec == P32) launch_export<float>(g_, 0, layout, vlen, (const float*)dev, stage, host_is_double, st_);
  else launch_export<__half>(g_, 0, layout, vlen, (const __half*)dev, stage, host_is_double, st_);
  copy_to_host(host, stage, bytes);
}

// ------------------------------------------------------------------ operators
template <> Lattice::FieldSet<double> Lattice::fields<double>() const {
  if (!Ud_ || !Cd_) throw std::runtime_error("FP64 gauge/clover fields are not loaded");
  return {Ud_, Cd_, 18};
}
template <> Lattice::FieldSet<float> Lattice::fields<float>() const {
  if (!Us_ || !Cs_) throw std::runtime_error("FP32 gauge/clover fields are not loaded");
  return {Us_, Cs_, nl32_};
}
template <> Lattice::FieldSet<__half> Lattice::fields<__half>() const {
  if (!Uh_ || !Ch_ || !half_valid_) throw std::runtime_error("FP16 gauge/clover fields are not available");
  return {Uh_, Ch_, nl32_};
}

BlockSel Lattice::sel(int colour) const {
  if (colour < 0) return {dslash_order_ == 1 ? d_order_pencil_ : d_order_, g_.nblk};
  return {d_blocks_[colour & 1], nblk_colour_[colour & 1]};
}

template <typename S>
void Lattice::hop(BlockSel s, const Batch& B, int set, typename Store<S>::R gamma, S* out, const S* a, const S* b,
                  const S* c, const S* in) {
  if (s.n == 0) return;
  FieldSet<S> f = fields<S>();
  if (f.nl == 18) launch_hop<S, 18>(g_, s, B, set, gamma, out, a, b, c, in, f.U, f.C, st_);
  else if constexpr (!std::is_same<S, double>::value)
    launch_hop<S, 12>(g_, s, B, set, gamma, out, a, b, c, in, f.U, f.C, st_);
}

#define DISPATCH_PREC(prec, CALL)                                  \
  do {                                                             \
    if ((prec) == P64) { using S = double; CALL; }                 \
    else if ((prec) == P32) { using S = float; CALL; }             \
    else { ensure_half_fields(); using S = __half; CALL; }         \
  } while (0)

void Lattice::dslash(int prec, void* out, const void* in, int hops, int colour) {
  DISPATCH_PREC(prec, (hop<S>(sel(colour), single(), hops, (typename Store<S>::R)(-kappa_), (S*)out, (const S*)in,
                              nullptr, nullptr, (const S*)in)));
}

void Lattice::dslash_batch(const Batch& B, double* out, const double* in) {
  hop<double>(sel(-1), B, HOPS_ALL, -kappa_, out, in, nullptr, nullptr, in);
}

void Lattice::face(int prec, void* out, const void* in, double factor, int colour) {
  DISPATCH_PREC(prec, (hop<S>(sel(colour), single(), HOPS_FACE, (typename Store<S>::R)factor, (S*)out, (const S*)out,
                              nullptr, nullptr, (const S*)in)));
}

void Lattice::clover(int prec, void* out, const void* in) {
  DISPATCH_PREC(prec, (hop<S>(sel(-1), single(), HOPS_NONE, (typename Store<S>::R)1, (S*)out, nullptr, nullptr,
                              nullptr, (const S*)in)));
}

// Jacobi block solve (static_solver.cc:65-140): x_1 = 2b - D_in b, then
// x_{k+1} = x_k + b - D_in x_k, i.e. x_{k+1} = b + kappa T^-1 H_in x_k with x_0 = b.
// Iterates alternate between x and tmp so that the last one lands in x.
template <typename S>
void Lattice::jinv_t(const Batch& B, int colour, S* x, const S* b, S* tmp, int nm) {
  using R = typename Store<S>::R;
  if (nm < 1) nm = 1;  // the first application is unconditional (static_solver.cc:77-101)
  const S* prev = b;
  for (int k = 1; k <= nm; ++k) {
    S* dst = ((nm - k) % 2 == 0) ? x : tmp;
    hop<S>(sel(colour), B, HOPS_INSIDE, (R)kappa_, dst, b, nullptr, nullptr, prev);
    prev = dst;
  }
}

void Lattice::jinv(int prec, void* x, const void* b, int colour, int nm) {
  DISPATCH_PREC(prec, ({
    S* tmp = (S*)work("jinv_tmp", spinor_bytes(prec));
    jinv_t<S>(single(), colour, (S*)x, (const S*)b, tmp, nm);
  }));
}

// SAP sweeps (qws.cc:809-841): for each colour c (0 = domain_e, 1 = domain_o):
//   q_c = A_c s_c ;  s_c += in_c - D_in q_c ;  s_o += kappa T^-1 H_face q
template <typename S>
void Lattice::sap_sweeps(const Batch& B, const S* in, S* s, S* q, S* tmp, int nsap, int nm) {
  using R = typename Store<S>::R;
  const R k = (R)kappa_;
  const int pr = sizeof(S) == 8 ? P64 : sizeof(S) == 4 ? P32 : P16;
  QWS_HIP(hipMemcpyAsync(s, in, spinor_bytes(pr) * B.n, hipMemcpyDeviceToDevice, st_));
  if (improper_) launch_zero_blocks<S>(g_, sel(1), B, q, st_);
  for (int isap = 0; isap < nsap; ++isap)
    for (int c = 0; c < 2; ++c) {
      jinv_t<S>(B, c, q, s, tmp, nm);
      hop<S>(sel(c), B, HOPS_INSIDE, k, s, s, in, q, q);  // s_c = s_c + in_c - q_c + kappa T^-1 H_in q_c
      hop<S>(sel(1 - c), B, HOPS_FACE, k, s, s, nullptr, nullptr, q);
    }
}

// Fused path: one launch per half-sweep (sap.hip), for FP32 or FP16 storage, proper
// colouring and 256-site blocks.
template <typename S> bool Lattice::fused_ok() const {
  if (std::is_same<S, float>::value && g_.Vb == 512 && nl32_ == 18) return false;  // LDS
  return !std::is_same<S, double>::value && use_fused_ && !improper_ && sap_fusable(g_);
}

// the nsap sweeps of qws.cc:809-841 with the face update of each half-sweep
// done at the start of the next one
template <typename S>
void Lattice::sap_sweeps_fused(const Batch& B, const S* in, S* s, S* q, int nsap, int nm, float* fb, long long fbs,
                               const SapSrcUpdate* upd) {
  const FieldSet<S> f = fields<S>();
  // the first sweep's two launches are the first readers of in_c for their colour c:
  // with upd, each applies the direction update to its own sites and stores it
  const SapSrcUpdate none;
  for (int isap = 0; isap < nsap; ++isap)
    for (int c = 0; c < 2; ++c) {
      int flags = SAP_EPI_SUPDATE;
      if (isap == 0) flags |= SAP_SRC_IN;
      if (isap == 0 && upd) flags |= SAP_SRC_PUPD;
      if (isap > 0 || c == 1) flags |= SAP_FACE_PRO;
      launch_sap_block<S>(g_, sel(c), B, f.nl, (float)kappa_, nm, fb_flags(flags, fb, true), s, in, q, q, nullptr, f.U,
                          f.C, fb, fbs, st_, (flags & SAP_SRC_PUPD) ? *upd : none);
    }
}

// The face buffer of a fused sequence (FP32): every launch but the last writes it,
// every face update reads it instead of gathering the other colour's q.
template <typename S> float* Lattice::sap_face_buffer(const Batch& B, long long& fbs) {
  fbs = 0;
  if constexpr (std::is_same<S, double>::value) return nullptr;
  fbs = sap_face_buffer_floats(g_);
  if (!fbs || !use_fb_) {
    fbs = 0;
    return nullptr;
  }
  return (float*)work("sap_fb", sizeof(float) * (size_t)fbs * B.n);
}

int Lattice::fb_flags(int flags, const float* fb, bool produce) {
  if (!fb) return flags;
  if (produce) flags |= SAP_FACE_OUT;
  if (flags & SAP_FACE_PRO) flags |= SAP_FACE_FB;
  return flags;
}

// The fused launches of prec_ddd_s_ (qws.cc:776-926) or prec_s_ (qws.cc:929-975) in
// FP32: nsap sweeps (qws.cc:809-841, each half-sweep's face update done at the start of
// the next launch), then the final sequence.  The first sweep's two launches are the
// first readers of in_c for their colour c: with pupd each applies the direction update
// to its own sites and stores it.
std::vector<SapStep> Lattice::sap_plan(const Batch& B, bool ddd, float* out, const float* in, int nsap, bool pupd,
                                       float*& fb, long long& fbs) {
  float* s = (float*)work("sap_s", spinor_bytes(P32) * B.n);
  float* q = (float*)work("sap_q", spinor_bytes(P32) * B.n);
  fb = sap_face_buffer<float>(B, fbs);
  std::vector<SapStep> v;
  auto step = [&](int colour, int flags, bool produce, const float* qface, float* qout, float* o, bool pu) {
    SapStep st;
    st.colour = colour, st.flags = fb_flags(flags, fb, produce), st.s = s, st.in = in, st.qface = qface;
    st.qout = qout, st.out = o, st.pupd = pu;
    v.push_back(st);
  };
  for (int isap = 0; isap < nsap; ++isap)
    for (int c = 0; c < 2; ++c) {
      int flags = SAP_EPI_SUPDATE;
      if (isap == 0) flags |= SAP_SRC_IN;
      if (isap > 0 || c == 1) flags |= SAP_FACE_PRO;
      step(c, flags, true, q, q, nullptr, isap == 0 && pupd);
    }
  const int src = nsap == 0 ? SAP_SRC_IN : 0;
  if (ddd) {
    // final e: q_e = A_e s_e, out_e = D_in q_e
    step(0, SAP_EPI_OUT | src | (nsap > 0 ? SAP_FACE_PRO : 0), true, q, q, out, false);
    // final o: s_o += kappa T^-1 H_face q, q_o = A_o s_o, out_o = D_oe q_e + D_in q_o
    // (and, with the face buffer, the face data of q_o for the last step)
    step(1, SAP_EPI_OUT_FACE | src | SAP_FACE_PRO, true, q, q, out, false);
    // out_e += D_eo q_o
    SapStep f;
    f.face = true, f.colour = 0, f.qface = q, f.out = out;
    v.push_back(f);
  } else {
    step(0, SAP_EPI_NONE | src | (nsap > 0 ? SAP_FACE_PRO : 0), true, q, out, nullptr, false);
    step(1, SAP_EPI_NONE | src | SAP_FACE_PRO, false, out, out, nullptr, false);
  }
  return v;
}

void Lattice::run_sap_plan(const Batch& B, const std::vector<SapStep>& plan, float* fb, long long fbs, int nm,
                           const SapSrcUpdate* upd) {
  const FieldSet<float> f = fields<float>();
  const float k = (float)kappa_;
  const SapSrcUpdate none;
  for (const SapStep& st : plan) {
    if (st.face) {
      if (fb) launch_sap_face(g_, sel(st.colour), B, f.nl, -k, st.out, fb, fbs, f.U, f.C, st_);
      else hop<float>(sel(st.colour), B, HOPS_FACE, -k, st.out, st.out, nullptr, nullptr, st.qface);
      continue;
    }
    launch_sap_block<float>(g_, sel(st.colour), B, f.nl, k, nm, st.flags | (st.pupd ? SAP_SRC_PUPD : 0), st.s, st.in,
                            st.qface, st.qout, st.out, f.U, f.C, fb, fbs, st_, st.pupd ? *upd : none);
  }
}

// prec_ddd_s_ (qws.cc:776-926): out = D M_SAP in, with the final sequence of the
// improper (npe[mu]==1) and of the proper colouring.
template <typename S>
void Lattice::prec_ddd_t(const Batch& B, S* out, const S* in, int nsap, int nm, const SapSrcUpdate* upd) {
  using R = typename Store<S>::R;
  const int pr = sizeof(S) == 8 ? P64 : sizeof(S) == 4 ? P32 : P16;
  S* s = (S*)work("sap_s", spinor_bytes(pr) * B.n);
  S* q = (S*)work("sap_q", spinor_bytes(pr) * B.n);
  const R k = (R)kappa_;
  if constexpr (std::is_same<S, float>::value) {
    if (upd && (const void*)upd->p != (const void*)in) throw std::logic_error("prec_ddd_t: the update must be of `in`");
    // the first sweep applies the update when there is one; otherwise a separate pass
    const bool fold = upd && fused_ok<S>() && nsap > 0;
    if (upd && !fold) launch_update_p<float>(B, g_.V * 12, upd->p, upd->r, upd->q, upd->sc, st_);
    if (fused_ok<S>()) {
      float* fb;
      long long fbs;
      const std::vector<SapStep> plan = sap_plan(B, true, out, in, nsap, fold, fb, fbs);
      run_sap_plan(B, plan, fb, fbs, nm, upd);
      return;
    }
  }
  if (upd && !std::is_same<S, float>::value) throw std::logic_error("prec_ddd_t: the direction update is FP32");
  S* tmp = (S*)work("sap_tmp", spinor_bytes(pr) * B.n);
  sap_sweeps<S>(B, in, s, q, tmp, nsap, nm);
  if (improper_) launch_zero_blocks<S>(g_, sel(1), B, q, st_);
  jinv_t<S>(B, 0, q, s, tmp, nm);
  hop<S>(sel(0), B, HOPS_INSIDE, -k, out, q, nullptr, nullptr, q);  // out_e = D_in q_e
  if (improper_) {
    hop<S>(sel(1), B, HOPS_FACE, k, s, s, nullptr, nullptr, q);      // s_o -= D_face q (q_o = 0)
    jinv_t<S>(B, 1, q, s, tmp, nm);
    hop<S>(sel(1), B, HOPS_INSIDE, -k, out, q, nullptr, nullptr, q);  // out_o = D_in q_o
    hop<S>(sel(0), B, HOPS_FACE, -k, out, out, nullptr, nullptr, q);  // out_e += D_face q
    hop<S>(sel(1), B, HOPS_FACE, -k, out, out, nullptr, nullptr, q);  // out_o += D_face q
  } else {
    hop<S>(sel(1), B, HOPS_FACE, -k, out, nullptr, nullptr, nullptr, q);  // out_o = D_oe q_e
    hop<S>(sel(1), B, HOPS_FACE, k, s, s, nullptr, nullptr, q);           // s_o -= out_o
    jinv_t<S>(B, 1, q, s, tmp, nm);
    hop<S>(sel(1), B, HOPS_INSIDE, -k, out, out, q, nullptr, q);          // out_o += D_in q_o
    hop<S>(sel(0), B, HOPS_FACE, -k, out, out, nullptr, nullptr, q);      // out_e += D_eo q_o
  }
}

// prec_s_ (qws.cc:929-975): out = M_SAP in
template <typename S>
void Lattice::prec_s_t(const Batch& B, S* out, const S* in, int nsap, int nm) {
  using R = typename Store<S>::R;
  const int pr = sizeof(S) == 8 ? P64 : sizeof(S) == 4 ? P32 : P16;
  S* s = (S*)work("sap_s", spinor_bytes(pr) * B.n);
  S* q = (S*)work("sap_q", spinor_bytes(pr) * B.n);
  if constexpr (std::is_same<S, float>::value) {
    if (fused_ok<S>()) {
      float* fb;
      long long fbs;
      const std::vector<SapStep> plan = sap_plan(B, false, out, in, nsap, false, fb, fbs);
      run_sap_plan(B, plan, fb, fbs, nm, nullptr);
      return;
    }
  } else if constexpr (!std::is_same<S, double>::value) {
    if (fused_ok<S>()) {
      const FieldSet<S> f = fields<S>();
      const int src = nsap == 0 ? SAP_SRC_IN : 0;
      long long fbs;
      float* fb = sap_face_buffer<S>(B, fbs);
      sap_sweeps_fused<S>(B, in, s, q, nsap, nm, fb, fbs);
      launch_sap_block<S>(g_, sel(0), B, f.nl, (float)kappa_, nm,
                          fb_flags(SAP_EPI_NONE | src | (nsap > 0 ? SAP_FACE_PRO : 0), fb, true), s, in, q, out, nullptr,
                          f.U, f.C, fb, fbs, st_);
      launch_sap_block<S>(g_, sel(1), B, f.nl, (float)kappa_, nm, fb_flags(SAP_EPI_NONE | src | SAP_FACE_PRO, fb, false),
                          s, in, out, out, nullptr, f.U, f.C, fb, fbs, st_);
      return;
    }
  }
  S* tmp = (S*)work("sap_tmp", spinor_bytes(pr) * B.n);
  sap_sweeps<S>(B, in, s, q, tmp, nsap, nm);
  if (improper_) launch_zero_blocks<S>(g_, sel(1), B, out, st_);
  jinv_t<S>(B, 0, out, s, tmp, nm);
  hop<S>(sel(1), B, HOPS_FACE, (R)kappa_, s, s, nullptr, nullptr, out);
  jinv_t<S>(B, 1, out, s, tmp, nm);
}

// assign_mult_mSAP_s_h_ (prec_ddd_s_h.h:63-231): out = M_SAP in with the SAP's
// iterates, links and clover inverse stored in FP16 and FP32 arithmetic.  It runs
// nsap - 1 colour sweeps and the final block solves (the sweep structure of
// prec_s_) on an FP16 copy of `in` scaled to unit RMS component; the reference
// scales to unit norm instead, which leaves large lattices' components in FP16's
// subnormal range.
void Lattice::msap_h_t(const Batch& B, float* out, const float* in, int nsap, int nm) {
  if (nsap > 10 || nm > 10) throw std::invalid_argument("nsap and nm must be <= 10 (prec_ddd_s_h.h:130-131)");
  ensure_half_fields();
  const int nb = B.n;
  if (!sc_h_) sc_h_ = dalloc<Scalars>(MAX_BATCH);
  __half* ih = (__half*)work("hp_in", spinor_bytes(P16) * nb);
  __half* oh = (__half*)work("hp_out", spinor_bytes(P16) * nb);
  double* part = reduction_partials(nb);
  launch_norm<float>(B, g_.V * 12, in, part, st_);
  launch_finalize(B, part, 1, sc_h_, FIN_RAW, st_);
  launch_scale_rms<__half, float>(B, g_.V, B.stride, B.stride, ih, in, sc_h_, +1, st_);
  prec_s_t<__half>(B, oh, ih, nsap > 1 ? nsap - 1 : 0, nm);
  launch_scale_rms<float, __half>(B, g_.V, B.stride, B.stride, out, oh, sc_h_, -1, st_);
}

void Lattice::msap_h(float* out, const float* in, int nsap, int nm) { msap_h_t(single(), out, in, nsap, nm); }

void Lattice::prec_ddd(int prec, void* out, const void* in, int nsap, int nm) {
  if (nsap > 10 || nm > 10) throw std::invalid_argument("prec_ddd_s: nsap and nm must be <= 10 (qws.cc:798-799)");
  DISPATCH_PREC(prec, (prec_ddd_t<S>(single(), (S*)out, (const S*)in, nsap, nm)));
}

size_t Lattice::prec_ddd_scratch_bytes() const {
  auto a256 = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t sp = spinor_bytes(P32);
  size_t third = sp;  // the step-by-step path's third spinor
  if (fused_ok<float>()) third = use_fb_ ? sizeof(float) * (size_t)sap_face_buffer_floats(g_) : 0;  // face buffer
  return a256(sp) + a256(sp) + third;  // s, q, then the third buffer, each at a 256-byte boundary
}

void Lattice::prec_ddd_scratch(void* out, const void* in, int nsap, int nm, void* scratch, size_t bytes) {
  if (nsap > 10 || nm > 10) throw std::invalid_argument("prec_ddd_s: nsap and nm must be <= 10 (qws.cc:798-799)");
  const size_t need = prec_ddd_scratch_bytes();
  if (need > bytes)
    throw std::invalid_argument("prec_ddd on this lattice needs " + std::to_string(need) +
                                " bytes of scratch memory, more than the " + std::to_string(bytes) + " provided");
  if (((uintptr_t)scratch & 255) != 0) throw std::invalid_argument("scratch memory must be 256-byte aligned");
  struct Restore {
    Lattice* L;
    ~Restore() { L->arena_ = nullptr, L->arena_cap_ = 0, L->arena_off_ = 0; }
  } restore{this};
  arena_ = (char*)scratch, arena_cap_ = bytes, arena_off_ = 0;
  prec_ddd_t<float>(single(), (float*)out, (const float*)in, nsap, nm);
}

void Lattice::prec_s(int prec, void* out, const void* in, int nsap, int nm) {
  if (nsap > 10 || nm > 10) throw std::invalid_argument("prec_s: nsap and nm must be <= 10");
  DISPATCH_PREC(prec, (prec_s_t<S>(single(), (S*)out, (const S*)in, nsap, nm)));
}

double Lattice::read_scalar(const Scalars* sc, size_t offset) {
  QWS_HIP(hipMemcpyAsync((char*)sc_host_ + offset, (const char*)sc + offset, sizeof(double), hipMemcpyDeviceToHost, st_));
  QWS_HIP(hipStreamSynchronize(st_));
  double v;
  std::memcpy(&v, (char*)sc_host_ + offset, sizeof(double));
  return v;
}

void Lattice::read_scalars(const Scalars* dev, int n) {
  QWS_HIP(hipMemcpyAsync(sc_host_, dev, sizeof(Scalars) * n, hipMemcpyDeviceToHost, st_));
  QWS_HIP(hipStreamSynchronize(st_));
}

double Lattice::norm2(int prec, const void* a) {
  const long long n = g_.V * 12;
  double* part = reduction_partials(1);
  if (prec == P64) launch_norm<double>(single(), n, (const double*)a, part, st_);
  else if (prec == P32) launch_norm<float>(single(), n, (const float*)a, part, st_);
  else throw std::invalid_argument("norm2: FP16 vectors are not reduced directly");
  launch_finalize(single(), part, 1, sc_one_, FIN_RAW, st_);
  return read_scalar(sc_one_, offsetof(Scalars, red));
}

unsigned long long Lattice::digest(int prec, const void* a) {
  const size_t words = spinor_bytes(prec) / 8;
  unsigned long long* d = (unsigned long long*)work("digest", sizeof(unsigned long long));
  QWS_HIP(hipMemsetAsync(d, 0, sizeof(unsigned long long), st_));
  launch_digest((long long)words, a, d, st_);
  unsigned long long h = 0;
  QWS_HIP(hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, st_));
  QWS_HIP(hipStreamSynchronize(st_));
  return h;
}

// ------------------------------------------------------------------ even/odd family
void Lattice::eo_load_gauge(const void* U, int layout, int vlen, int is_double, int target, bool on_device) {
  const size_t bytes = (size_t)4 * g_.V * 18 * (is_double ? 8 : 4);
  const void* stage = U;
  if (!on_device) {
    void* st = work("stage", bytes);
    QWS_HIP(hipMemcpyAsync(st, U, bytes, hipMemcpyHostToDevice, st_));
    stage = st;
  }
  if (target & 1) {
    if (!Ueo_d_) Ueo_d_ = dalloc<double>((size_t)4 * 18 * g_.V);
    launch_eo_import<double>(g_, 1, layout, vlen, 18, stage, is_double, Ueo_d_, st_);
  }
  if (target & 2) {
    if (!Ueo_s_) Ueo_s_ = dalloc<float>((size_t)4 * nl32_ * g_.V);
    launch_eo_import<float>(g_, 1, layout, vlen, nl32_, stage, is_double, Ueo_s_, st_);
  }
  QWS_HIP(hipStreamSynchronize(st_));
}

void Lattice::eo_load_clover(const void* C, int layout, int vlen, int is_double, int target, bool on_device) {
  const size_t bytes = (size_t)g_.V * 72 * (is_double ? 8 : 4);
  const void* stage = C;
  if (!on_device) {
    void* st = work("stage", bytes);
    QWS_HIP(hipMemcpyAsync(st, C, bytes, hipMemcpyHostToDevice, st_));
    stage = st;
  }
  if (target & 1) {
    if (!Ceo_d_) Ceo_d_ = dalloc<double>((size_t)72 * g_.V);
    launch_eo_import<double>(g_, 2, layout, vlen, 0, stage, is_double, Ceo_d_, st_);
  }
  if (target & 2) {
    if (!Ceo_s_) Ceo_s_ = dalloc<float>((size_t)72 * g_.V);
    launch_eo_import<float>(g_, 2, layout, vlen, 0, stage, is_double, Ceo_s_, st_);
  }
  QWS_HIP(hipStreamSynchronize(st_));
}

void Lattice::eo_fields_from_lattice() {
  for (int m = 0; m < 4; ++m)
    if (g_.L[m] % 2) throw std::invalid_argument("even/odd fields from the lattice need even extents");
  if (Ud_ && Cd_) {
    if (!Ueo_d_) Ueo_d_ = dalloc<double>((size_t)4 * 18 * g_.V);
    if (!Ceo_d_) Ceo_d_ = dalloc<double>((size_t)72 * g_.V);
    launch_eo_gather<double>(g_, 18, Ud_, Cd_, Ueo_d_, Ceo_d_, st_);
  }
  if (Us_ && Cs_) {
    if (!Ueo_s_) Ueo_s_ = dalloc<float>((size_t)4 * nl32_ * g_.V);
    if (!Ceo_s_) Ceo_s_ = dalloc<float>((size_t)72 * g_.V);
    launch_eo_gather<float>(g_, nl32_, Us_, Cs_, Ueo_s_, Ceo_s_, st_);
  }
  QWS_HIP(hipStreamSynchronize(st_));
}

void Lattice::eo_upload_spinor(int prec, void* dev, const void* host, int layout, int vlen, int host_is_double) {
  const size_t bytes = (size_t)(g_.V / 2) * 24 * (host_is_double ? 8 : 4);
  void* stage = work("stage_spinor", bytes);
  QWS_HIP(hipMemcpyAsync(stage, host, bytes, hipMemcpyHostToDevice, st_));
  if (prec == P64) launch_eo_import<double>(g_, 0, layout, vlen, 0, stage, host_is_double, (double*)dev, st_);
  else if (prec == P32) launch_eo_import<float>(g_, 0, layout, vlen, 0, stage, host_is_double, (float*)dev, st_);
  else throw std::invalid_argument("even/odd spinors are FP64 or FP32");
  QWS_HIP(hipStreamSynchronize(st_));
}

void Lattice::eo_download_spinor(int prec, const void* dev, void* host, int layout, int vlen, int host_is_double) {
  const size_t bytes = (size_t)(g_.V / 2) * 24 * (host_is_double ? 8 : 4);
  void* stage = work("stage_spinor", bytes);
  if (prec == P64) launch_eo_export<double>(g_, layout, vlen, (const double*)dev, stage, host_is_double, st_);
  else if (prec == P32) launch_eo_export<float>(g_, layout, vlen, (const float*)dev, stage, host_is_double, st_);
  else throw std::invalid_argument("even/odd spinors are FP64 or FP32");
  QWS_HIP(hipMemcpyAsync(host, stage, bytes, hipMemcpyDeviceToHost, st_));
  QWS_HIP(hipStreamSynchronize(st_));
}

void Lattice::deo(int prec, int pe, int set, int dag, int clv, int mode, double fac, void* out, const void* in,
                  const void* in0) {
  if (pe != 0 && pe != 1) throw std::invalid_argument("parity must be 0 or 1");
  if (set < 0 || set > 3 || mode < 0 || mode > 2) throw std::invalid_argument("deo: set in 0..3, mode in 0..2");
  if (mode == 2 && !in0) throw std::invalid_argument("deo: mode 2 needs in0");
  if (prec == P64) {
    if (!Ueo_d_ || !Ceo_d_) throw std::runtime_error("FP64 even/odd fields are not loaded");
    launch_deo<double>(g_, 18, pe, set, dag, clv, mode, fac, (double*)out, (const double*)in, (const double*)in0,
                       Ueo_d_, Ceo_d_, st_);
  } else if (prec == P32) {
    if (!Ueo_s_ || !Ceo_s_) throw std::runtime_error("FP32 even/odd fields are not loaded");
    launch_deo<float>(g_, nl32_, pe, set, dag, clv, mode, (float)fac, (float*)out, (const float*)in,
                      (const float*)in0, Ueo_s_, Ceo_s_, st_);
  } else {
    throw std::invalid_argument("even/odd operators are FP64 or FP32");
  }
}

void Lattice::mtilde(int prec, int pce, int which, double fac, void* out, const void* in) {
  if (which < 0 || which > 2) throw std::invalid_argument("mtilde: which in 0..2");
  if (pce != 0 && pce != 1) throw std::invalid_argument("parity must be 0 or 1");
  const int pco = 1 - pce;
  const size_t b = eo_spinor_bytes(prec);
  void* t0 = work("eo_t0", b);
  void* t1 = work("eo_t1", b);
  void* t2 = work("eo_t2", b);
  auto mt = [&](void* o, const void* i) {  // qws.cc:551-558
    deo(prec, pco, 0, 0, 1, 0, 0, t0, i, nullptr);
    deo(prec, pce, 0, 0, 1, 2, fac, o, t0, i);
  };
  auto mtdag = [&](void* o, const void* i) {  // qws.cc:559-568
    deo(prec, pce, 3, 0, 1, 0, 0, t0, i, nullptr);
    deo(prec, pco, 0, 1, 1, 0, 0, t1, t0, nullptr);
    deo(prec, pce, 0, 1, 0, 2, fac, o, t1, i);
  };
  if (which == 0) mt(out, in);
  else if (which == 1) mtdag(out, in);
  else {
    mt(t2, in);
    mtdag(out, t2);
  }
}

int Lattice::bicgstab_eo(double* x, const double* b, int maxiter, double tol, int pce, double fac) {
  const long long n = g_.V / 2 * 12;  // complex components of a parity field
  const size_t bytes = eo_spinor_bytes(P64);
  double* q = (double*)work("eob_q", bytes);
  double* r = (double*)work("eob_r", bytes);
  double* p = (double*)work("eob_p", bytes);
  double* t = (double*)work("eob_t", bytes);
  double* r0 = (double*)work("eob_r0", bytes);
  if (!sc_eo_) sc_eo_ = dalloc<Scalars>(1);
  double* part = reduction_partials(1);
  const Batch B = single();
  mtilde(P64, pce, 0, fac, q, x);
  launch_bicg_init_outer<double>(B, n, r, p, r0, b, q, part, st_);
  launch_finalize(B, part, 2, sc_eo_, FIN_INIT_OUTER, st_);
  auto converged = [&] {
    read_scalars(sc_eo_, 1);
    return std::sqrt(sc_host_[0].rnorm / sc_host_[0].bnorm) < tol;
  };
  int iter;
  for (iter = 1; iter < maxiter; ++iter) {
    mtilde(P64, pce, 0, fac, q, p);
    launch_dot<double>(B, n, r0, q, part, st_);
    launch_finalize(B, part, 2, sc_eo_, FIN_ALPHA, st_);
    launch_update_xr<double>(B, n, x, p, r, q, sc_eo_, 0, nullptr, part, st_);
    launch_finalize(B, part, 1, sc_eo_, FIN_RNORM, st_);
    if (converged()) break;
    mtilde(P64, pce, 0, fac, t, r);
    launch_norm_dot<double>(B, n, t, r, part, st_);
    launch_finalize(B, part, 3, sc_eo_, FIN_OMEGA, st_);
    launch_update_xr<double>(B, n, x, r, r, t, sc_eo_, 1, r0, part, st_);
    launch_finalize(B, part, 3, sc_eo_, FIN_RHO, st_);
    if (converged()) break;
    launch_update_p<double>(B, n, p, r, q, sc_eo_, st_);
  }
  sync();
  return iter;
}

// bicgstab_dd_s.cc:67-330 (FP32) and bicgstab_dd_d.cc:64-265 (FP64): the same
// iteration as bicgstab_eo on D = 1 - kappa T^-1 H; the reductions are FP64 here
template <typename S>
int Lattice::bicgstab_dd_t(S* x, const S* b, int maxiter, double tol) {
  using R = typename Store<S>::R;
  const int pr = sizeof(S) == 8 ? P64 : P32;
  const long long n = g_.V * 12;
  const size_t bytes = spinor_bytes(pr);
  S* q = (S*)work("bdd_q", bytes);
  S* r = (S*)work("bdd_r", bytes);
  S* p = (S*)work("bdd_p", bytes);
  S* t = (S*)work("bdd_t", bytes);
  S* r0 = (S*)work("bdd_r0", bytes);
  if (!sc_eo_) sc_eo_ = dalloc<Scalars>(1);
  double* part = reduction_partials(1);
  const Batch B = single();
  const R mk = (R)(-kappa_);
  auto D = [&](S* out, const S* in) { hop<S>(sel(-1), B, HOPS_ALL, mk, out, in, nullptr, nullptr, in); };
  D(q, x);
  launch_bicg_init_outer<S>(B, n, r, p, r0, b, q, part, st_);
  launch_finalize(B, part, 2, sc_eo_, FIN_INIT_OUTER, st_);
  auto converged = [&] {
    read_scalars(sc_eo_, 1);
    return std::sqrt(sc_host_[0].rnorm / sc_host_[0].bnorm) < tol;
  };
  int iter;
  for (iter = 1; iter < maxiter; ++iter) {
    D(q, p);
    launch_dot<S>(B, n, r0, q, part, st_);
    launch_finalize(B, part, 2, sc_eo_, FIN_ALPHA, st_);
    launch_update_xr<S>(B, n, x, p, r, q, sc_eo_, 0, nullptr, part, st_);
    launch_finalize(B, part, 1, sc_eo_, FIN_RNORM, st_);
    if (converged()) break;
    D(t, r);
    launch_norm_dot<S>(B, n, t, r, part, st_);
    launch_finalize(B, part, 3, sc_eo_, FIN_OMEGA, st_);
    launch_update_xr<S>(B, n, x, r, r, t, sc_eo_, 1, r0, part, st_);
    launch_finalize(B, part, 3, sc_eo_, FIN_RHO, st_);
    if (converged()) break;
    launch_update_p<S>(B, n, p, r, q, sc_eo_, st_);
  }
  sync();
  return iter;
}

int Lattice::bicgstab_dd(int prec, void* x, const void* b, int maxiter, double tol) {
  if (prec == P64) return bicgstab_dd_t<double>((double*)x, (const double*)b, maxiter, tol);
  if (prec == P32) return bicgstab_dd_t<float>((float*)x, (const float*)b, maxiter, tol);
  throw std::invalid_argument("bicgstab_dd: FP32 or FP64 vectors");
}

double Lattice::reduce_real(long long n, const double* a, const double* b) {
  if (n % 2) throw std::invalid_argument("reduce_real: an even number of reals");
  if (!sc_eo_) sc_eo_ = dalloc<Scalars>(1);
  double* part = reduction_partials(1);
  const Batch B = single();
  // complex view: |a|^2, or Re <a, b> = sum a_i b_i over the reals
  if (b) launch_dot<double>(B, n / 2, a, b, part, st_);
  else launch_norm<double>(B, n / 2, a, part, st_);
  launch_finalize(B, part, 1, sc_eo_, FIN_RAW, st_);
  read_scalars(sc_eo_, 1);
  return sc_host_[0].red[0];
}

// cg_vm_ (cg.cc:65-172) on the parity fields: q = A x, r = b - q, p = r; per iteration
// q = A p, alpha = |r|^2 / <p,q>, x += alpha p, r -= alpha q, stop when |r|/|b| < tol,
// p = r + beta p.  A = mtilde^+ mtilde.
int Lattice::cg_eo(double* x, const double* b, int maxiter, double tol, int pce, double fac) {
  const long long nr = g_.V / 2 * 24;  // reals of a parity field
  const size_t bytes = eo_spinor_bytes(P64);
  double* q = (double*)work("eoc_q", bytes);
  double* r = (double*)work("eoc_r", bytes);
  double* p = (double*)work("eoc_p", bytes);
  mtilde(P64, pce, 2, fac, q, x);
  launch_sub<double>(nr, r, b, q, st_);
  QWS_HIP(hipMemcpyAsync(p, r, bytes, hipMemcpyDeviceToDevice, st_));
  const double bnorm = reduce_real(nr, b, nullptr);
  double rho0 = reduce_real(nr, r, nullptr);
  int iter;
  for (iter = 1; iter < maxiter; ++iter) {
    mtilde(P64, pce, 2, fac, q, p);
    const double alpha = rho0 / reduce_real(nr, p, q);
    launch_lincomb<double>(nr, alpha, p, 1.0, x, 0, st_);
    launch_lincomb<double>(nr, -alpha, q, 1.0, r, 0, st_);
    const double rnorm = reduce_real(nr, r, nullptr);
    if (std::sqrt(rnorm / bnorm) < tol) break;
    const double beta = rnorm / rho0;
    rho0 = rnorm;
    launch_lincomb<double>(nr, 1.0, r, beta, p, 1, st_);
  }
  sync();
  return iter;
}

// mcg_vm_ (mcg.cc:65-219): one Krylov space for every shift; with alpha the negated
// CG step, xi the shift's scale factors (sxi0/1/2) and pp_k its search direction:
// r += alpha q; xs_k -= alpha xi2/xi1 pp_k; p = r + beta p; pp_k = beta (xi2/xi1)^2 pp_k + xi2 r.
int Lattice::mcg_eo(double* const* xs, const double* b, int ns, const double* sigma, int maxiter, double tol, int pce,
                    double fac) {
  if (ns < 1 || ns > 64) throw std::invalid_argument("mcg: 1..64 shifts");
  const long long nr = g_.V / 2 * 24;
  const size_t bytes = eo_spinor_bytes(P64);
  double* q = (double*)work("eom_q", bytes);
  double* r = (double*)work("eom_r", bytes);
  double* p = (double*)work("eom_p", bytes);
  double* pp = (double*)work("eom_pp", bytes * ns);
  std::vector<double> x0(ns, 1.0), x1(ns, 1.0), x2(ns, 0.0);
  double alpha0 = 1, beta = 0;
  QWS_HIP(hipMemcpyAsync(r, b, bytes, hipMemcpyDeviceToDevice, st_));
  QWS_HIP(hipMemcpyAsync(p, b, bytes, hipMemcpyDeviceToDevice, st_));
  const double bnorm = reduce_real(nr, b, nullptr);
  double rho0 = bnorm;
  for (int k = 0; k < ns; ++k) {
    QWS_HIP(hipMemsetAsync(xs[k], 0, bytes, st_));
    QWS_HIP(hipMemcpyAsync(pp + k * nr, r, bytes, hipMemcpyDeviceToDevice, st_));
  }
  int iter;
  for (iter = 1; iter < maxiter; ++iter) {
    mtilde(P64, pce, 2, fac, q, p);
    const double alpha = -rho0 / reduce_real(nr, p, q);
    launch_lincomb<double>(nr, alpha, q, 1.0, r, 0, st_);
    const double rnorm = reduce_real(nr, r, nullptr);
    for (int k = 0; k < ns; ++k) {
      x2[k] = x1[k] * x0[k] * alpha0 / (alpha * beta * (x0[k] - x1[k]) + x0[k] * alpha0 * (1 - sigma[k] * alpha));
      launch_lincomb<double>(nr, -(alpha * x2[k] / x1[k]), pp + k * nr, 1.0, xs[k], 0, st_);
    }
    if (std::sqrt(rnorm / bnorm) < tol) break;
    beta = rnorm / rho0;
    rho0 = rnorm;
    launch_lincomb<double>(nr, 1.0, r, beta, p, 1, st_);
    for (int k = 0; k < ns; ++k) {
      const double c = beta * (x2[k] / x1[k]) * (x2[k] / x1[k]);
      launch_lincomb<double>(nr, x2[k], r, c, pp + k * nr, 1, st_);
      x0[k] = x1[k];
      x1[k] = x2[k];
    }
    alpha0 = alpha;
  }
  sync();
  return iter;
}

// ------------------------------------------------------------------ inner BiCGStab
// Right-preconditioned BiCGStab for (D M_SAP) x = b, x0 = 0 (bicgstab_precdd_s.cc:85-559),
// for every RHS k with live[k] != 0.  B.active must point at d_inner_mask_; an RHS
// leaves the batch when its recursive residual meets tol (tol < 0: fixed count).
//
// Without hp it runs as Fused::bicg on this lattice alone.  hp: the method of
// bicgstab_precdd_s_h.cc:83-370 instead -- w = M_SAP p with the FP16 SAP, q = D w
// with the FP32 operator, and x accumulates w directly (x is the solution of D x = b,
// no final M_SAP).
template <typename S>
void Lattice::bicg_t(const Batch& Bin, S* x, const S* b, double tol, int maxiter, int nsap, int nm,
                     std::vector<int>& iters, bool hp) {
  static_assert(sizeof(S) == 4, "the inner solver runs on FP32 vectors");
  if (!hp) {
    std::vector<Fused::InnerJob> one{{this, Bin, (float*)x, (const float*)b, &iters}};
    Fused::bicg(one, tol, maxiter, nsap, nm, nullptr);
    return;
  }
  const int nb = Bin.n;
  const long long n = g_.V * 12;
  const size_t bytes = spinor_bytes(P32) * nb;
  S* q = (S*)work("bicg_q", bytes);
  S* r = (S*)work("bicg_r", bytes);
  S* p = (S*)work("bicg_p", bytes);
  S* t = (S*)work("bicg_t", bytes);
  S* r0 = (S*)work("bicg_r0", bytes);
  S* w = (S*)work("bicg_w", bytes);
  S* w2 = (S*)work("bicg_w2", bytes);
  const float mk = (float)(-kappa_);
  if (nb > sc_in_n_) {
    if (sc_in_) QWS_HIP(hipFree(sc_in_));
    sc_in_ = dalloc<Scalars>(nb);
    sc_in_n_ = nb;
  }
  Scalars* sc = sc_in_;
  double* part = reduction_partials(nb);
  std::vector<int> live(nb);
  for (int k = 0; k < nb; ++k) live[k] = iters[k] >= 0 ? 1 : 0;  // iters[k] < 0 marks RHS outside the batch
  for (int k = 0; k < nb; ++k) iters[k] = 0;
  auto upload = [&] {
    QWS_HIP(hipMemcpyAsync(d_inner_mask_, live.data(), sizeof(int) * nb, hipMemcpyHostToDevice, st_));
  };
  upload();
  const Batch B{nb, g_.V * 24, d_inner_mask_};
  launch_bicg_init_inner<S>(B, n, x, r, p, r0, b, part, st_);
  launch_finalize(B, part, 1, sc, FIN_INIT_INNER, st_);
  std::vector<double> bnorm(nb, 1.0);
  if (tol >= 0) {
    read_scalars(sc, nb);
    for (int k = 0; k < nb; ++k) bnorm[k] = sc_host_[k].bnorm;
  }
  auto converge = [&]() {  // drop converged RHS from the batch; true if any remain
    if (tol < 0) return true;
    read_scalars(sc, nb);
    bool changed = false, any = false;
    for (int k = 0; k < nb; ++k) {
      if (!live[k]) continue;
      // a non-finite residual (breakdown) ends the RHS's iteration as well
      if (std::sqrt(sc_host_[k].rnorm / bnorm[k]) < tol || !std::isfinite(sc_host_[k].rnorm)) {
        live[k] = 0;
        changed = true;
      } else {
        any = true;
      }
    }
    if (changed) upload();
    return any;
  };
  auto any_live = [&] {
    for (int v : live)
      if (v) return true;
    return false;
  };
  for (int iter = 0; iter < maxiter && any_live(); ++iter) {
    for (int k = 0; k < nb; ++k) iters[k] += live[k];
    msap_h_t(B, w, p, nsap, nm);
    hop<S>(sel(-1), B, HOPS_ALL, mk, q, w, nullptr, nullptr, w);
    launch_dot<S>(B, n, r0, q, part, st_);
    launch_finalize(B, part, 2, sc, FIN_ALPHA, st_);
    // r -= alpha q; x += alpha M p is deferred to the second half-step's update, or done
    // here for the RHS that leave after this half-step
    launch_update_xr<S>(B, n, x, w, r, q, sc, 0, nullptr, part, st_, XR_R_ONLY);
    launch_finalize(B, part, 1, sc, FIN_RNORM, st_);
    const std::vector<int> before(live);
    const bool more = converge();
    fix_.assign(nb, 0);
    bool anyfix = false;
    for (int k = 0; k < nb; ++k) anyfix |= (fix_[k] = before[k] && !live[k]) != 0;
    if (anyfix) {
      int* dfix = (int*)work("bicg_fix", sizeof(int) * nb);
      QWS_HIP(hipMemcpyAsync(dfix, fix_.data(), sizeof(int) * nb, hipMemcpyHostToDevice, st_));
      launch_update_x<S>(Batch{nb, B.stride, dfix}, n, x, w, sc, st_);
    }
    if (!more) break;
    msap_h_t(B, w2, r, nsap, nm);
    hop<S>(sel(-1), B, HOPS_ALL, mk, t, w2, nullptr, nullptr, w2);
    launch_norm_dot<S>(B, n, t, r, part, st_);
    launch_finalize(B, part, 3, sc, FIN_OMEGA, st_);
    launch_update_xr<S>(B, n, x, w2, r, t, sc, 1, r0, part, st_, XR_DEFERRED, w);
    launch_finalize(B, part, 3, sc, FIN_RHO, st_);
    if (!converge()) break;
    launch_update_p<S>(B, n, p, r, q, sc, st_);  // p = r + beta (p - omega q)
  }
  if (trace_on()) {
    read_scalars(sc, nb);
    for (int k = 0; k < nb; ++k)
      if (iters[k] > 0)
        std::fprintf(stderr, "@C %d %5d %12.7e\n", k, iters[k], std::sqrt(sc_host_[k].rnorm / bnorm[k]));
  }
}

int Lattice::bicgstab_precdd(int prec, void* x, const void* b, double tol, int maxiter, int nsap, int nm) {
  if (prec != P32) throw std::invalid_argument("bicgstab_precdd: FP32 vectors only");
  std::vector<int> it(1, 0);
  bicg_t<float>(Batch{1, g_.V * 24, d_inner_mask_}, (float*)x, (const float*)b, tol, maxiter, nsap, nm, it, false);
  return it[0];
}

int Lattice::bicgstab_precdd_h(float* x, const float* b, double tol, int maxiter, int nsap, int nm) {
  std::vector<int> it(1, 0);
  bicg_t<float>(Batch{1, g_.V * 24, d_inner_mask_}, x, b, tol, maxiter, nsap, nm, it, true);
  return it[0];
}

// ------------------------------------------------------------------ outer mixed BiCGStab
// u = M p: normalise, FP32 inner solve, prec_s_, rescale (bicgstab_dd_mix.cc:76-102,
// reordering.cc:108-204), for the RHS of B with live[k] != 0 (B.active holds the
// same flags on the device); sc = the outer scalars (nrm is set here); the inner
// iterations are added to inner[k].
bool Lattice::approx_inv(const Batch& B, Scalars* sc, const std::vector<int>& live, double* u, const double* p,
                         const SolveParams& prm, std::vector<int>& inner) {
  const long long n = g_.V * 12;
  const int nb = B.n;
  const size_t bytes = spinor_bytes(P32) * nb;
  float* bs = (float*)work("ai_b", bytes);
  float* ts = (float*)work("ai_t", bytes);
  float* xs = (float*)work("ai_x", bytes);
  double* part = reduction_partials(nb);
  launch_norm<double>(B, n, p, part, st_);
  launch_finalize(B, part, 1, sc, FIN_NRM, st_);
  launch_d2s_scaled<float>(B, g_.V, B.stride, B.stride, bs, p, sc, st_);
  std::vector<int> it(nb);
  for (int k = 0; k < nb; ++k) it[k] = live[k] ? 0 : -1;
  if (prm.sap_prec == P16) {
    // approx_inv_dirac_op_hf (bicgstab_dd_mix2_hf.cc:88-106): the inner solution is x itself
    bicg_t<float>(Batch{nb, g_.V * 24, d_inner_mask_}, xs, bs, prm.tol_s, prm.maxiter_s, prm.nsap, prm.nm, it, true);
    for (int k = 0; k < nb; ++k)
      if (live[k]) inner[k] += it[k];
    if (prm.escalate && prm.tol_s > 0) {
      // the inner solve's last residual relative to its (unit) source
      read_scalars(sc_in_, nb);
      for (int k = 0; k < nb; ++k)
        if (live[k]) {
          const double rel = std::sqrt(sc_host_[k].rnorm / sc_host_[k].bnorm);
          if (!(rel < 1.0)) return true;  // failed: no u; the caller switches to the FP32 SAP
        }
    }
  } else {
    bicg_t<float>(Batch{nb, g_.V * 24, d_inner_mask_}, ts, bs, prm.tol_s, prm.maxiter_s, prm.nsap, prm.nm, it, false);
    prec_s_t<float>(B, xs, ts, prm.nsap, prm.nm);
    for (int k = 0; k < nb; ++k)
      if (live[k]) inner[k] += it[k];
  }
  launch_s2d_scaled<float>(B, g_.V, B.stride, B.stride, u, xs, sc, st_);
  return false;
}

void check_sap_params(int nsap, int nm) {
  if (nsap < 0 || nsap > 10 || nm < 0 || nm > 10)
    throw std::invalid_argument("nsap and nm must be in 0..10 (qws.cc:798-799)");
}

void OuterBatch::set_params(const SolveParams& p) {
  check_sap_params(p.nsap, p.nm);
  prm_ = p;
  escalated_ = false;
}

OuterBatch::OuterBatch(Lattice& L, int capacity, const SolveParams& p) : L_(L), cap_(capacity), prm_(p) {
  if (cap_ < 1 || cap_ > MAX_BATCH) throw std::invalid_argument("batch capacity must be in 1..64");
  check_sap_params(p.nsap, p.nm);
  stride_ = L.volume() * 24;
  const size_t bytes = sizeof(double) * (size_t)stride_ * cap_;
  double** bufs[] = {&x_, &b_, &q_, &r_, &p_, &t_, &u_, &r0_};
  for (double** b : bufs) {
    *b = nullptr;
    QWS_HIP(hipMalloc((void**)b, bytes));
  }
  QWS_HIP(hipMalloc((void**)&sc_, sizeof(Scalars) * cap_));
  QWS_HIP(hipMemset(sc_, 0, sizeof(Scalars) * cap_));
  QWS_HIP(hipMalloc((void**)&d_mask_, sizeof(int) * cap_));
  state_.assign(cap_, EMPTY);
  iters_.assign(cap_, 0);
  inner_.assign(cap_, 0);
}

OuterBatch::~OuterBatch() {
  (void)hipStreamSynchronize(L_.stream());
  for (double* b : {x_, b_, q_, r_, p_, t_, u_, r0_})
    if (b) (void)hipFree(b);
  if (sc_) (void)hipFree(sc_);
  if (d_mask_) (void)hipFree(d_mask_);
}

int OuterBatch::active_count() const {
  int n = 0;
  for (int s : state_) n += s == RUNNING;
  return n;
}

void OuterBatch::upload_mask(const std::vector<int>& m) {
  QWS_HIP(hipMemcpyAsync(d_mask_, m.data(), sizeof(int) * cap_, hipMemcpyHostToDevice, L_.stream()));
}

// start slot k: x = x0 (or 0), r = b - D x, p = r0 = r (bicgstab_dd_mix.cc:136-158)
void OuterBatch::admit(int k, const double* b, const double* x0) {
  if (k < 0 || k >= cap_) throw std::invalid_argument("admit: slot out of range");
  hipStream_t st = L_.stream();
  const size_t bytes = sizeof(double) * stride_;
  if (b != this->b(k)) QWS_HIP(hipMemcpyAsync(this->b(k), b, bytes, hipMemcpyDeviceToDevice, st));
  if (!x0) QWS_HIP(hipMemsetAsync(x(k), 0, bytes, st));
  else if (x0 != x(k)) QWS_HIP(hipMemcpyAsync(x(k), x0, bytes, hipMemcpyDeviceToDevice, st));
  std::vector<int> only(cap_, 0);
  only[k] = 1;
  upload_mask(only);
  const Batch B{cap_, stride_, d_mask_};
  double* part = L_.reduction_partials(cap_);
  // q = D x; for x = 0 that is 0, and r = b - q = b exactly without the operator
  if (x0) L_.dslash_batch(B, q_, x_);
  else QWS_HIP(hipMemsetAsync(q_ + (size_t)k * stride_, 0, bytes, st));
  launch_bicg_init_outer<double>(B, L_.volume() * 12, r_, p_, r0_, b_, q_, part, st);
  launch_finalize(B, part, 2, sc_, FIN_INIT_OUTER, st);
  state_[k] = RUNNING;
  iters_[k] = 0;
  inner_[k] = 0;
}

// Slots whose recursive residual sqrt(|r|^2/|b|^2) met tol finish
// (bicgstab_dd_mix.cc:249,336).
void OuterBatch::check(std::vector<int>& live, int half) {
  if (prm_.tol < 0 && !trace_on()) return;
  const Scalars* h = L_.host_scalars();
  if (trace_on())  // outer half-step 2 iter - 1 (after alpha) or 2 iter (after omega)
    for (int k = 0; k < cap_; ++k)
      if (live[k])
        std::fprintf(stderr, "@D %d %5d %24.15e\n", k, 2 * iters_[k] - 1 + half, std::sqrt(h[k].rnorm / h[k].bnorm));
  if (prm_.tol < 0) return;
  bool changed = false;
  for (int k = 0; k < cap_; ++k)
    if (live[k] && (std::sqrt(h[k].rnorm / h[k].bnorm) < prm_.tol || !std::isfinite(h[k].rnorm))) {
      live[k] = 0;
      state_[k] = FINISHED;
      changed = true;
    }
  if (changed) upload_mask(live);
}

int OuterBatch::step() {
  std::vector<OuterBatch*> one{this};
  return Fused::outer_step(one, nullptr);
}

int Lattice::bicgstab_dd_mix(double* x, const double* b, const SolveParams& prm, int* inner_total) {
  if (prm.maxiter <= 1) {
    if (inner_total) *inner_total = 0;
    return prm.maxiter;
  }
  // the solver's vectors persist between calls (eight spinors: allocating and freeing
  // them per call costs more than moving x and b)
  if (!batch_ || batch_->capacity() != 1) {
    batch_.reset();
    batch_.reset(new OuterBatch(*this, 1, prm));
  }
  OuterBatch& ob = *batch_;
  ob.set_params(prm);
  ob.admit(0, b, x);
  while (ob.step() > 0) {
  }
  QWS_HIP(hipMemcpyAsync(x, ob.x(0), spinor_bytes(P64), hipMemcpyDeviceToDevice, st_));
  sync();
  if (inner_total) *inner_total = ob.inner_iterations(0);
  return ob.iterations(0);
}

// |b - D x| / |b| in FP64 on the GPU
double Lattice::true_residual(const double* x, const double* b) {
  const long long n = g_.V * 12;
  double* q = (double*)work("res_q", spinor_bytes(P64));
  dslash(P64, q, x, HOPS_ALL, -1);
  launch_sub<double>(2 * n, q, b, q, st_);
  double rr = norm2(P64, q);
  double bb = norm2(P64, b);
  return std::sqrt(rr / bb);
}

std::vector<SolveStats> Lattice::solve_batch(double* const* x, const double* const* b, int nrhs,
                                             const SolveParams& prm) {
  if (nrhs <= 0) return std::vector<SolveStats>(nrhs > 0 ? nrhs : 0);
  std::vector<std::vector<double*>> xs{std::vector<double*>(x, x + nrhs)};
  std::vector<std::vector<const double*>> bs{std::vector<const double*>(b, b + nrhs)};
  return Fused::solve({this}, xs, bs, prm, nullptr)[0];
}

SolveStats Lattice::solve(double* x, const double* b, const SolveParams& prm) {
  double* xs[1] = {x};
  const double* bs[1] = {b};
  return solve_batch(xs, bs, 1, prm)[0];
}

double Lattice::time_dslash(int prec, int reps) {
  void* in = work("time_in", spinor_bytes(prec));
  void* out = work("time_out", spinor_bytes(prec));
  QWS_HIP(hipMemsetAsync(in, 0, spinor_bytes(prec), st_));
  dslash(prec, out, in, HOPS_ALL, -1);  // warm up
  hipEvent_t e0, e1;
  QWS_HIP(hipEventCreate(&e0));
  QWS_HIP(hipEventCreate(&e1));
  QWS_HIP(hipEventRecord(e0, st_));
  for (int i = 0; i < reps; ++i) dslash(prec, out, in, HOPS_ALL, -1);
  QWS_HIP(hipEventRecord(e1, st_));
  QWS_HIP(hipEventSynchronize(e1));
  float ms = 0;
  QWS_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms / reps;
}

template <typename S>
double Lattice::time_sap_t(int reps, int nsap, int nm) {
  if (!fused_ok<S>()) throw std::invalid_argument("time_sap: the fused SAP kernel does not apply to this lattice");
  const int pr = sizeof(S) == 4 ? P32 : P16;
  S* in = (S*)work("time_in", spinor_bytes(pr));
  S* s = (S*)work("sap_s", spinor_bytes(pr));
  S* q = (S*)work("sap_q", spinor_bytes(pr));
  S* out = (S*)work("time_out", spinor_bytes(pr));
  if constexpr (std::is_same<S, float>::value) {
    gaussian_spinor(P32, in, 11);
  } else {
    float* tmp = (float*)work("time_tmp", spinor_bytes(P32));
    gaussian_spinor(P32, tmp, 11);
    launch_convert<__half, float>(g_.V, in, tmp, st_);
  }
  const Batch B = single();
  const FieldSet<S> f = fields<S>();
  long long fbs;
  float* fb = sap_face_buffer<S>(B, fbs);
  auto apply = [&] {
    sap_sweeps_fused<S>(B, in, s, q, nsap, nm, fb, fbs);
    launch_sap_block<S>(g_, sel(0), B, f.nl, (float)kappa_, nm, fb_flags(SAP_EPI_OUT | SAP_FACE_PRO, fb, true), s, in, q,
                        q, out, f.U, f.C, fb, fbs, st_);
    launch_sap_block<S>(g_, sel(1), B, f.nl, (float)kappa_, nm, fb_flags(SAP_EPI_OUT_FACE | SAP_FACE_PRO, fb, false), s,
                        in, q, q, out, f.U, f.C, fb, fbs, st_);
  };
  apply();  // warm up
  hipEvent_t e0, e1;
  QWS_HIP(hipEventCreate(&e0));
  QWS_HIP(hipEventCreate(&e1));
  QWS_HIP(hipEventRecord(e0, st_));
  for (int i = 0; i < reps; ++i) apply();
  QWS_HIP(hipEventRecord(e1, st_));
  QWS_HIP(hipEventSynchronize(e1));
  float ms = 0;
  QWS_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms / (reps * (2.0 * nsap + 2));
}

double measure_stream(long long bytes, int nin, int reps, bool one) {
  if (bytes < (1 << 20) || nin < 1 || nin > 16 || reps < 1) throw std::invalid_argument("measure_stream: bad arguments");
  const long long n4 = bytes / 16 / 256 * 256;
  float4 *in = nullptr, *out = nullptr;
  QWS_HIP(hipMalloc((void**)&in, 16 * n4 * nin));
  if (hipMalloc((void**)&out, 16 * n4) != hipSuccess) {
    (void)hipFree(in);
    throw std::runtime_error("measure_stream: out of device memory");
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float ms = 0;
  try {
    QWS_HIP(hipMemset(in, 0, 16 * n4 * nin));
    QWS_HIP(hipEventCreate(&e0));
    QWS_HIP(hipEventCreate(&e1));
    ms = 1e30f;
    for (int cfg = one ? 3 : 0; cfg < (one ? 5 : 3); ++cfg) {  // the fastest launch shape
      float t = 0;
      launch_stream(cfg, n4, nin, in, out, nullptr);  // warm up
      QWS_HIP(hipEventRecord(e0, nullptr));
      for (int r = 0; r < reps; ++r) launch_stream(cfg, n4, nin, in, out, nullptr);
      QWS_HIP(hipEventRecord(e1, nullptr));
      QWS_HIP(hipEventSynchronize(e1));
      QWS_HIP(hipEventElapsedTime(&t, e0, e1));
      ms = t < ms ? t : ms;
    }
  } catch (...) {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(in);
    (void)hipFree(out);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  QWS_HIP(hipFree(in));
  QWS_HIP(hipFree(out));
  return 16.0 * n4 * (nin + 1) * reps / (ms * 1e-3) / 1e9;
}

void Lattice::set_dslash_kernel(int k) {
  if (k != 0 && k != 3 && k != 4) throw std::invalid_argument("dslash kernel: 0 (global gathers), 3 or 4 (LDS-staged)");
  g_.dslash_kernel = k;
}

void Lattice::set_dslash_order(int order) {
  if (order != 0 && order != 1) throw std::invalid_argument("dslash order: 0 (index) or 1 (t-z pencils per XCD)");
  dslash_order_ = order;
}

int Lattice::autotune_dslash(int reps) {
  if (reps < 1) throw std::invalid_argument("autotune_dslash: reps must be positive");
  int best = g_.dslash_kernel, best_order = dslash_order_;
  double best_ms = 1e30;
  for (int order : {0, 1})
    for (int k : {0, 3, 4}) {
      g_.dslash_kernel = k;
      dslash_order_ = order;
      const double ms = time_dslash(P32, reps);
      if (ms < best_ms) {
        best_ms = ms;
        best = k;
        best_order = order;
      }
    }
  g_.dslash_kernel = best;
  dslash_order_ = best_order;
  return best;
}

SolveParams Lattice::autotune_solve(double* x, const double* b, const SolveParams& prm, const std::vector<int>& nsaps,
                                    const std::vector<int>& nms, std::vector<double>& seconds) {
  if (nsaps.empty() || nms.empty()) throw std::invalid_argument("autotune_solve: empty candidate lists");
  for (int v : nsaps) check_sap_params(v, 1);
  for (int v : nms) check_sap_params(1, v);
  autotune_dslash(20);
  seconds.assign(nsaps.size() * nms.size(), -1.0);
  SolveParams best = prm;
  double best_s = 1e300;
  double* xs[1] = {x};
  const double* bs[1] = {b};
  for (size_t i = 0; i < nsaps.size(); ++i)
    for (size_t j = 0; j < nms.size(); ++j) {
      SolveParams p = prm;
      p.nsap = nsaps[i], p.nm = nms[j];
      QWS_HIP(hipMemsetAsync(x, 0, spinor_bytes(P64), st_));
      sync();
      const auto t0 = std::chrono::steady_clock::now();
      const SolveStats st = solve_batch(xs, bs, 1, p)[0];
      const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      if (prm.true_tol > 0 && !(st.true_residual <= prm.true_tol)) continue;
      seconds[i * nms.size() + j] = sec;
      if (sec < best_s) best_s = sec, best = p;
    }
  if (best_s == 1e300) throw std::runtime_error("autotune_solve: no candidate met the true-residual target");
  return best;
}

int Lattice::autotune_blocks(const double* b, const SolveParams& prm, const std::vector<int>& shapes,
                             std::vector<double>& seconds) {
  if (shapes.empty() || shapes.size() % 4 != 0) throw std::invalid_argument("autotune_blocks: shapes holds 4 ints each");
  if (!Ud_ || !Cd_) throw std::runtime_error("autotune_blocks: FP64 gauge/clover fields are not loaded");
  // the fields and b leave the device in the plain site order, which every block shape reads
  std::vector<double> U((size_t)g_.V * 72), C((size_t)g_.V * 72), bh((size_t)g_.V * 24);
  export_field(1, U.data());
  export_field(2, C.data());
  download_spinor(P64, b, bh.data(), 0, 1);
  sync();
  const size_t n = shapes.size() / 4;
  seconds.assign(n, -1.0);
  int best = -1;
  for (size_t i = 0; i < n; ++i) {
    std::unique_ptr<Lattice> c;
    try {
      c.reset(new Lattice(g_.L, &shapes[4 * i], g_.tbc, nl32_));
    } catch (const std::invalid_argument&) {
      continue;  // not a tiling this solver admits
    }
    c->set_kappa(kappa_);
    c->load_gauge(U.data(), 0, 1, 3);
    c->load_clover(C.data(), 0, 1, 3);
    c->autotune_dslash(20);
    double* x = (double*)c->alloc_spinor(P64);
    double* bc = (double*)c->alloc_spinor(P64);
    c->upload_spinor(P64, bc, bh.data(), 0, 1);
    double* xs[1] = {x};
    const double* bs[1] = {bc};
    c->solve_batch(xs, bs, 1, prm);  // untimed: the work buffers are allocated here
    QWS_HIP(hipMemsetAsync(x, 0, c->spinor_bytes(P64), c->st_));
    c->sync();
    const auto t0 = std::chrono::steady_clock::now();
    const SolveStats st = c->solve_batch(xs, bs, 1, prm)[0];
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    QWS_HIP(hipFree(x));
    QWS_HIP(hipFree(bc));
    if (prm.true_tol > 0 && !(st.true_residual <= prm.true_tol)) continue;
    seconds[i] = sec;
    if (best < 0 || sec < seconds[best]) best = (int)i;
  }
  if (best < 0) throw std::runtime_error("autotune_blocks: no admissible shape met the true-residual target");
  return best;
}

double Lattice::time_sap(int prec, int reps, int nsap, int nm) {
  if (reps < 1 || nsap < 1 || nm < 1) throw std::invalid_argument("time_sap: reps, nsap and nm must be positive");
  if (prec == P32) return time_sap_t<float>(reps, nsap, nm);
  if (prec == P16) {
    ensure_half_fields();
    return time_sap_t<__half>(reps, nsap, nm);
  }
  throw std::invalid_argument("time_sap: FP32 or FP16 storage");
}

}  // namespace qws

namespace qws {
// used by the fused multi-lattice solves (multi.cpp)
template bool Lattice::fused_ok<float>() const;
template void Lattice::prec_ddd_t<float>(const Batch&, float*, const float*, int, int, const SapSrcUpdate*);
template void Lattice::prec_s_t<float>(const Batch&, float*, const float*, int, int);
}  // namespace qws
