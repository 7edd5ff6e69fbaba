// lattice.h -- host-side state of one lattice on one GPU: geometry, device
// fields, work buffers and the solvers built from the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "kernels.h"

namespace qws {

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define QWS_HIP(call)                                                                                   \
  do {                                                                                                  \
    hipError_t e_ = (call);                                                                             \
    if (e_ != hipSuccess)                                                                               \
      throw ::qws::HipError(std::string(#call) + ": " + hipGetErrorString(e_) + " (" __FILE__ ":" +    \
                            std::to_string(__LINE__) + ")");                                            \
  } while (0)

enum Prec { P64 = 0, P32 = 1, P16 = 2 };

struct SolveParams {
  int nsap = 4, nm = 2;
  double tol = -1;
  int maxiter = 6;
  double tol_s = -1;
  int maxiter_s = 50;
  double true_tol = -1;
  int max_restarts = 0;
  int sap_prec = P32;
};

struct SolveStats {
  int outer_iters = 0, inner_iters = 0, restarts = 0;
  double true_residual = -1, seconds = 0;
};

class Lattice {
 public:
  Lattice(const int L[4], const int B[4], int tbc, int links32);
  ~Lattice();
  Lattice(const Lattice&) = delete;
  Lattice& operator=(const Lattice&) = delete;

  const Geo& geo() const { return g_; }
  long long volume() const { return g_.V; }
  bool improper() const { return improper_; }
  double kappa() const { return kappa_; }
  void set_kappa(double k) { kappa_ = k; }
  hipStream_t stream() const { return st_; }
  int links32() const { return nl32_; }
  // fused SAP half-sweep kernels (default on where applicable)
  void set_fused_sap(bool on) { use_fused_ = on; }

  // fields from host arrays (layout 0 plain, 1 QWS AoSoA)
  void load_gauge(const void* U, int layout, int is_double, int target);  // target: bit0 FP64, bit1 FP32
  void load_clover(const void* C, int layout, int is_double, int target);
  void ensure_half_fields();
  // generated inputs: SU(3) links exp(i eps H), clover inverse from the links
  void generate_gauge(double eps, unsigned long long seed);
  void compute_clover(double csw);
  void export_field(int kind, double* host_plain);  // kind 1 links, 2 clover (FP64)
  void gaussian_spinor(int prec, void* dev, unsigned long long seed);
  void point_spinor(int prec, void* dev, const int x[4], int spin, int colour);

  // spinor I/O between host arrays and device buffers
  void* alloc_spinor(int prec);
  void upload_spinor(int prec, void* dev, const void* host, int layout, int host_is_double);
  void download_spinor(int prec, const void* dev, void* host, int layout, int host_is_double);
  size_t spinor_bytes(int prec) const;

  // operators (prec selects the field set)
  void dslash(int prec, void* out, const void* in, int hops, int colour);
  void face(int prec, void* out, const void* in, double factor, int colour);
  void clover(int prec, void* out, const void* in);
  void jinv(int prec, void* x, const void* b, int colour, int nm);
  void prec_ddd(int prec, void* out, const void* in, int nsap, int nm);
  void prec_s(int prec, void* out, const void* in, int nsap, int nm);
  int bicgstab_precdd(int prec, void* x, const void* b, double tol, int maxiter, int nsap, int nm);
  int bicgstab_dd_mix(double* x, const double* b, const SolveParams& p, int* inner_total);
  SolveStats solve(double* x, const double* b, const SolveParams& p);
  double norm2(int prec, const void* a);
  double true_residual(const double* x, const double* b);
  double time_dslash(int prec, int reps);
  void sync() { QWS_HIP(hipStreamSynchronize(st_)); }

 private:
  template <typename S> struct FieldSet { const S* U; const S* C; int nl; };
  template <typename S> FieldSet<S> fields() const;
  BlockSel sel(int colour) const;
  template <typename S>
  void hop(BlockSel s, int set, typename Store<S>::R gamma, S* out, const S* a, const S* b, const S* c,
           const S* in);
  template <typename S> void jinv_t(int colour, S* x, const S* b, S* tmp, int nm);
  template <typename S> void sap_sweeps(const S* in, S* s, S* q, S* tmp, int nsap, int nm);
  template <typename S> bool fused_ok() const;
  void sap_sweeps_fused(const float* in, float* s, float* q, int nsap, int nm);
  template <typename S> void prec_ddd_t(S* out, const S* in, int nsap, int nm);
  template <typename S> void prec_s_t(S* out, const S* in, int nsap, int nm);
  template <typename S> int bicg_t(S* x, const S* b, double tol, int maxiter, int nsap, int nm);
  void approx_inv(double* u, const double* p, const SolveParams& prm, int* inner);
  double read_scalar(const Scalars* sc, size_t offset);
  void* work(const std::string& name, size_t bytes);

  Geo g_;
  int nl32_;
  bool improper_;
  double kappa_ = 0;
  hipStream_t st_ = nullptr;
  int* d_blocks_[2] = {nullptr, nullptr};
  int nblk_colour_[2] = {0, 0};
  // fields
  double *Ud_ = nullptr, *Cd_ = nullptr;
  float *Us_ = nullptr, *Cs_ = nullptr;
  __half *Uh_ = nullptr, *Ch_ = nullptr;
  bool half_valid_ = false;
  bool use_fused_ = true;
  // reductions
  double* part_ = nullptr;
  Scalars* sc_in_ = nullptr;
  Scalars* sc_out_ = nullptr;
  Scalars* sc_host_ = nullptr;  // pinned
  std::vector<std::pair<std::string, std::pair<void*, size_t>>> work_;
};

}  // namespace qws
