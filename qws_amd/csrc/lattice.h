// lattice.h -- host-side state of one lattice on one GPU: geometry, device
// fields, work buffers and the solvers built from the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "kernels.h"

namespace qws {
struct SapSrcUpdate;
struct Fused;
class SapFusion;

// One launch of a fused FP32 SAP sequence on one lattice: the blocks of `colour`, the
// launch's flags and buffers (s, in, qface, qout, out of launch_sap_block), whether it
// applies the BiCGStab direction update (pupd); face: instead the final face term
// out_e += -kappa T^-1 H_eo q_o of prec_ddd_s_ (from the face buffer, or gathered from
// qface without one).
struct SapStep {
  int colour = 0, flags = 0;
  float* s = nullptr;
  const float* in = nullptr;
  const float* qface = nullptr;
  float* qout = nullptr;
  float* out = nullptr;
  bool pupd = false, face = false;
};
}

namespace qws {

// GB/s (reads + writes) of a streaming kernel summing nin arrays of `bytes` into one
// (nin = 1: copy): the fastest of three launch shapes, each averaged over reps
// launches; allocates and frees its buffers.  one: the nin inputs are consecutive
// chunks of one array, read region by region (a single input stream).
double measure_stream(long long bytes, int nin, int reps, bool one = false);
// the multi-RHS clover product on the vector ALUs and on the matrix cores, timed
// (mfma_trial.hip; measurement only): ms[0] VALU, ms[1] MFMA per launch
void time_clover_rhs(long long V, int nrep, int reps, double ms[2], double* maxdiff);

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
  // with sap_prec = P16 and tol_s > 0: switch to the FP32 SAP (with nsap - 1, the same
  // number of sweeps) when an inner solve ends with a relative residual >= 1 (near the
  // critical kappa the FP16-stored SAP can make the inner BiCGStab diverge); 0 keeps
  // the FP16 method throughout, as the reference
  int escalate = 1;
};

// nsap and nm of QWS's SAP must not exceed 10 (prec_ddd_s_ exits otherwise,
// qws.cc:798-799) nor be negative; throws std::invalid_argument
void check_sap_params(int nsap, int nm);

struct SolveStats {
  int outer_iters = 0, inner_iters = 0, restarts = 0;
  double true_residual = -1, seconds = 0;
  // the FP16-stored SAP method (sap_prec = P16) switched to the FP32 SAP during the solve
  // because an inner solve ended with a residual no smaller than its source's
  int sap_escalated = 0;
};

class Lattice;

// A batch of FP64 outer BiCGStab solves (bicgstab_dd_mix.cc:105-388) advanced one
// outer iteration at a time.  Each slot holds one right-hand side with its own
// recurrence; a slot can be (re)filled between iterations while the others run
// on (continuous batching), and every kernel processes all running slots in one
// launch, so the links and clover are read once per batch.
class OuterBatch {
 public:
  OuterBatch(Lattice& L, int capacity, const SolveParams& p);
  ~OuterBatch();
  int capacity() const { return cap_; }
  // rejects nsap or nm outside 0..10 (qws.cc:798-799), as the constructor does
  void set_params(const SolveParams& p);
  // start slot k on source b (device FP64 spinor) with initial guess x0 (nullptr: zero)
  void admit(int k, const double* b, const double* x0);
  // one outer iteration of every running slot; returns the number still running
  int step();
  bool running(int k) const { return state_[k] == RUNNING; }
  bool finished(int k) const { return state_[k] == FINISHED; }
  int iterations(int k) const { return iters_[k]; }
  int inner_iterations(int k) const { return inner_[k]; }
  double* x(int k) const { return x_ + (size_t)k * stride_; }
  double* b(int k) const { return b_ + (size_t)k * stride_; }
  void release(int k) { state_[k] = EMPTY; }
  // the batch switched from the FP16-stored SAP to the FP32 SAP (set_params resets it)
  bool escalated() const { return escalated_; }
  int active_count() const;

 private:
  enum { EMPTY = 0, RUNNING = 1, FINISHED = 2 };
  void upload_mask(const std::vector<int>& m);
  friend struct Fused;
  // converged slots leave `live`, from the scalars last read into the lattice's host
  // copy; half: 0 after the alpha update, 1 after the omega update
  void check(std::vector<int>& live, int half);
  Lattice& L_;
  int cap_;
  SolveParams prm_;
  long long stride_;
  double *x_, *b_, *q_, *r_, *p_, *t_, *u_, *r0_;
  Scalars* sc_;
  int* d_mask_;
  std::vector<int> state_, iters_, inner_;
  bool escalated_ = false;
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
  // run this lattice's work on another stream (nullptr: its own again); the lattices of
  // a solve fused over several lattices share one stream
  void set_stream(hipStream_t s) { st_ = s ? s : own_st_; }
  int links32() const { return nl32_; }
  // fused SAP half-sweep kernels (default on where applicable)
  void set_fused_sap(bool on) { use_fused_ = on; }

  // fields from host arrays (layout 0 plain, 1 QWS AoSoA)
  // target: bit0 FP64, bit1 FP32; on_device: U/C are already in device memory
  void load_gauge(const void* U, int layout, int is_double, int target, bool on_device = false);
  void load_clover(const void* C, int layout, int is_double, int target, bool on_device = false);
  // caller (pageable) host memory <-> device memory through the pinned chunks; synchronous
  void upload_bytes(void* dev, const void* host, size_t bytes) { copy_to_device(dev, host, bytes); }
  void download_bytes(void* host, const void* dev, size_t bytes) { copy_to_host(host, dev, bytes); }
  // generated inputs: SU(3) links exp(i eps H), clover inverse from the links
  void generate_gauge(double eps, unsigned long long seed);
  // use another lattice's links (same extents and link storage); they stay owned by src,
  // which must outlive this lattice and not change its links meanwhile
  void share_gauge(const Lattice& src);
  void compute_clover(double csw);
  void export_field(int kind, double* host_plain);  // kind 1 links, 2 clover (FP64)
  void gaussian_spinor(int prec, void* dev, unsigned long long seed);
  void point_spinor(int prec, void* dev, const int x[4], int spin, int colour);
  size_t field_bytes() const;

  // spinor I/O between host arrays and device buffers
  void* alloc_spinor(int prec);
  void upload_spinor(int prec, void* dev, const void* host, int layout, int host_is_double);
  void download_spinor(int prec, const void* dev, void* host, int layout, int host_is_double);
  size_t spinor_bytes(int prec) const;
  long long spinor_elems() const { return g_.V * 24; }

  // operators on one spinor (prec selects the field set)
  void dslash(int prec, void* out, const void* in, int hops, int colour);
  void face(int prec, void* out, const void* in, double factor, int colour);
  void clover(int prec, void* out, const void* in);
  void jinv(int prec, void* x, const void* b, int colour, int nm);
  void prec_ddd(int prec, void* out, const void* in, int nsap, int nm);
  // prec_ddd in FP32 with its work buffers carved from `scratch` (device memory of
  // `bytes`) instead of allocated: nothing is allocated; prec_ddd_scratch_bytes() is
  // what it needs.  Every buffer it uses is written before it is read.
  size_t prec_ddd_scratch_bytes() const;
  void prec_ddd_scratch(void* out, const void* in, int nsap, int nm, void* scratch, size_t bytes);
  void prec_s(int prec, void* out, const void* in, int nsap, int nm);
  int bicgstab_precdd(int prec, void* x, const void* b, double tol, int maxiter, int nsap, int nm);
  // FP16-storage SAP (assign_mult_mSAP_s_h_) and the inner solver built on it
  // (bicgstab_precdd_s_h_); FP32 vectors in and out
  void msap_h(float* out, const float* in, int nsap, int nm);
  int bicgstab_precdd_h(float* x, const float* b, double tol, int maxiter, int nsap, int nm);
  int bicgstab_dd_mix(double* x, const double* b, const SolveParams& p, int* inner_total);
  SolveStats solve(double* x, const double* b, const SolveParams& p);
  // n independent right-hand sides solved together (x[k] holds the initial guess)
  std::vector<SolveStats> solve_batch(double* const* x, const double* const* b, int n, const SolveParams& p);
  double norm2(int prec, const void* a);
  // order-independent digest of a device spinor's bytes (launch_digest): equal for
  // bit-identical spinors, whatever produced them
  unsigned long long digest(int prec, const void* a);
  // FP64 reductions over n reals to the host: |a|^2 (b == nullptr) or sum a_i b_i
  double reduce_real(long long n, const double* a, const double* b);
  double true_residual(const double* x, const double* b);
  double time_dslash(int prec, int reps);
  // average device time (ms) of one fused SAP launch: reps times the 2*nsap + 2
  // launches of one preconditioner application (prec_ddd_s_), single RHS
  double time_sap(int prec, int reps, int nsap, int nm);
  // FP32 full-operator kernel on 256-site blocks (0 global gathers, 3/4 LDS-staged);
  // the variants give the same results, autotune_dslash keeps the fastest on this
  // lattice (timed over `reps` launches each) and returns it
  // FP32 fused SAP: face updates through the face buffer (default) or by gathering
  // the other colour's spinors; both give the same results bit for bit
  void set_sap_face_buffer(bool on) { use_fb_ = on; }
  // FP32 inner BiCGStab with the fused SAP: its direction update p = r + beta (p - omega q)
  // runs inside the first SAP sweep of the next preconditioner application (default)
  // or as a separate pass; both give the same results bit for bit
  void set_sap_source_update(bool on) { fold_pupd_ = on; }
  void set_dslash_kernel(int k);
  // Block order of the launches over every block (the full operator D among them):
  // 0 = index order (x fastest); 1 = inside each XCD's contiguous range of workgroups,
  // the blocks sorted with t fastest, then z, x, y -- so that the t and z neighbours of
  // a block are read while they are still in that XCD's L2.  Results do not depend on it.
  void set_dslash_order(int order);
  int dslash_order() const { return dslash_order_; }
  int dslash_kernel() const { return g_.dslash_kernel; }
  int autotune_dslash(int reps);
  // The solver's SAP settings tuned on this lattice: one solve of b from x = 0 (x is
  // overwritten) for every (nsap, nm) of the candidate lists with prm's other settings,
  // timed on the host clock; returns prm with the fastest pair whose FP64 true residual
  // meets prm.true_tol (all of them when true_tol <= 0), seconds[i * nms + j] per pair
  // (negative: did not meet it).  Also picks the full-operator kernel (autotune_dslash).
  SolveParams autotune_solve(double* x, const double* b, const SolveParams& prm, const std::vector<int>& nsaps,
                             const std::vector<int>& nms, std::vector<double>& seconds);
  // The SAP block shape tuned for this lattice's extents, FP64 fields and kappa: for each
  // candidate (shapes[4 i .. 4 i + 3], x y z t) a lattice of that shape is built with
  // copies of the fields, solves b (FP64, this lattice's layout) once untimed and once
  // timed on the host clock with prm; seconds[i] is negative where the shape is not
  // admissible (does not tile the extents, odd block counts) or the timed solve missed
  // prm.true_tol.  Returns the index of the fastest shape.  This lattice is not changed.
  int autotune_blocks(const double* b, const SolveParams& prm, const std::vector<int>& shapes,
                      std::vector<double>& seconds);
  void sync() { QWS_HIP(hipStreamSynchronize(st_)); }

  // ---- even/odd family (deo_*, mtilde) on x-compressed parity fields of V/2 sites
  size_t eo_spinor_bytes(int prec) const { return spinor_bytes(prec) / 2; }
  // host fields in layout 0 (plain: links [2][4][V/2][18], clover [2][V/2][72]) or 1 (QWS
  // AoSoA with lane width vlen, read as parity halves); target bit0 FP64, bit1 FP32
  void eo_load_gauge(const void* U, int layout, int vlen, int is_double, int target, bool on_device = false);
  void eo_load_clover(const void* C, int layout, int vlen, int is_double, int target, bool on_device = false);
  // even/odd fields from the lattice's own fields (all extents even)
  void eo_fields_from_lattice();
  void eo_upload_spinor(int prec, void* dev, const void* host, int layout, int vlen, int host_is_double);
  void eo_download_spinor(int prec, const void* dev, void* host, int layout, int vlen, int host_is_double);
  // out (parity pe) from in (parity 1-pe); see launch_deo
  void deo(int prec, int pe, int set, int dag, int clv, int mode, double fac, void* out, const void* in,
           const void* in0);
  // which 0: mtilde = 1 + fac T^-1 H T^-1 H on parity pce (qws.cc:551-558; fac = -kappa^2),
  // 1: its adjoint (qws.cc:559-568), 2: mtilde^+ mtilde (qws.cc:569-574)
  void mtilde(int prec, int pce, int which, double fac, void* out, const void* in);
  // bicgstab_vm_ (bicgstab.cc:65-290): FP64 BiCGStab on mtilde (parity pce, fac =
  // -kappa^2) from the initial guess in x, iterations 1..maxiter-1, stopping when
  // sqrt(|r|^2/|b|^2) < tol (the reference's 1e-14); returns the last iteration
  int bicgstab_eo(double* x, const double* b, int maxiter, double tol, int pce, double fac);
  // plain BiCGStab on the full operator D (FP32 or FP64 vectors, FP64 reductions): x
  // is the initial guess; stops at sqrt(|r|^2/|b|^2) < tol or after maxiter-1 iterations
  // and returns the last iteration
  int bicgstab_dd(int prec, void* x, const void* b, int maxiter, double tol);
  // cg.cc:65 -- CG on mtilde^+ mtilde from the initial guess in x; returns the iterations
  int cg_eo(double* x, const double* b, int maxiter, double tol, int pce, double fac);
  // mcg.cc:65 -- multi-shift CG: the parity field xs[k] solves
  // (mtilde^+ mtilde + sigma[k]) xs[k] = b, from zero; returns the iterations
  int mcg_eo(double* const* xs, const double* b, int ns, const double* sigma, int maxiter, double tol, int pce,
             double fac);

  // ---- batched building blocks (used by OuterBatch)
  void dslash_batch(const Batch& B, double* out, const double* in);
  // returns true when the FP16-storage method's inner solve failed (some RHS ended with
  // a relative residual >= 1 or not finite): the caller switches to the FP32 SAP
  bool approx_inv(const Batch& B, Scalars* sc, const std::vector<int>& live, double* u, const double* p,
                  const SolveParams& prm, std::vector<int>& inner);
  Scalars* host_scalars() { return sc_host_; }
  void read_scalars(const Scalars* dev, int n);
  void* work(const std::string& name, size_t bytes);
  double* reduction_partials(int nrhs);

 private:
  template <typename S> struct FieldSet { const S* U; const S* C; int nl; };
  template <typename S> FieldSet<S> fields() const;
  BlockSel sel(int colour) const;
  template <typename S>
  void hop(BlockSel s, const Batch& B, int set, typename Store<S>::R gamma, S* out, const S* a, const S* b,
           const S* c, const S* in);
  template <typename S> void jinv_t(const Batch& B, int colour, S* x, const S* b, S* tmp, int nm);
  template <typename S> void sap_sweeps(const Batch& B, const S* in, S* s, S* q, S* tmp, int nsap, int nm);
  template <typename S> bool fused_ok() const;
  template <typename S> double time_sap_t(int reps, int nsap, int nm);
  template <typename S>
  void sap_sweeps_fused(const Batch& B, const S* in, S* s, S* q, int nsap, int nm, float* fb, long long fbs,
                        const SapSrcUpdate* upd = nullptr);
  template <typename S> float* sap_face_buffer(const Batch& B, long long& fbs);
  static int fb_flags(int flags, const float* fb, bool produce);
  bool use_fb_ = true;
  bool fold_pupd_ = true;
  // upd: apply BiCGStab's direction update to `in` (== upd->p) first
  template <typename S>
  void prec_ddd_t(const Batch& B, S* out, const S* in, int nsap, int nm, const SapSrcUpdate* upd = nullptr);
  template <typename S> void prec_s_t(const Batch& B, S* out, const S* in, int nsap, int nm);
  template <typename S> int bicgstab_dd_t(S* x, const S* b, int maxiter, double tol);
  void msap_h_t(const Batch& B, float* out, const float* in, int nsap, int nm);
  template <typename S>
  void bicg_t(const Batch& B, S* x, const S* b, double tol, int maxiter, int nsap, int nm, std::vector<int>& iters,
              bool hp);
  double read_scalar(const Scalars* sc, size_t offset);

  Geo g_;
  int nl32_;
  bool improper_;
  double kappa_ = 0;
  hipStream_t st_ = nullptr;
  int* d_blocks_[2] = {nullptr, nullptr};
  int* d_order_ = nullptr;  // every block, in index order
  int* d_order_pencil_ = nullptr;  // every block, t-z pencils inside each XCD range
  int dslash_order_ = 0;
  int nblk_colour_[2] = {0, 0};
  // fields
  double *Ud_ = nullptr, *Cd_ = nullptr;
  float *Us_ = nullptr, *Cs_ = nullptr;
  __half *Uh_ = nullptr, *Ch_ = nullptr;
  double *Ueo_d_ = nullptr, *Ceo_d_ = nullptr;  // even/odd fields
  float *Ueo_s_ = nullptr, *Ceo_s_ = nullptr;
  bool half_valid_ = false;
  bool owns_gauge_ = true;
  bool use_fused_ = true;
  // reductions
  double* part_ = nullptr;
  int part_rhs_ = 0;
  Scalars* sc_in_ = nullptr;
  int sc_in_n_ = 0;
  std::vector<int> fix_;  // the FP16-SAP inner solve: RHS leaving after a first half-step
  Scalars* sc_one_ = nullptr;
  Scalars* sc_host_ = nullptr;  // pinned, room for MAX_BATCH
  // host <-> device copies of caller memory (pageable) through two pinned chunks
  void* pin_[2] = {nullptr, nullptr};
  size_t pin_chunk_ = (size_t)32 << 20;  // bytes per chunk (QWS_MI355X_PIN_CHUNK)
  hipEvent_t pin_ev_[2] = {nullptr, nullptr};
  void copy_to_device(void* dev, const void* host, size_t bytes);
  void copy_to_host(void* host, const void* dev, size_t bytes);
  Scalars* sc_h_ = nullptr;     // FP16 SAP input norms, room for MAX_BATCH
  Scalars* sc_eo_ = nullptr;    // even/odd solver
  int* d_inner_mask_ = nullptr;
  std::vector<std::pair<std::string, std::pair<void*, size_t>>> work_;
  // while set, work() carves its buffers from this memory instead (prec_ddd_scratch)
  char* arena_ = nullptr;
  size_t arena_cap_ = 0, arena_off_ = 0;

  std::unique_ptr<OuterBatch> batch_;  // reused by solve/solve_batch (same capacity)

  friend class OuterBatch;
  friend struct Fused;
  void ensure_half_fields();
  hipStream_t own_st_ = nullptr;
  std::vector<int> h_blocks_[2];  // host copy of the block lists per colour
  // The launches of one fused FP32 prec_ddd_s_ (ddd) or prec_s_ application on the
  // blocks of this lattice (fused_ok<float>() must hold); pupd: the first sweep applies
  // BiCGStab's direction update.  fb/fbs: the face buffer it uses (nullptr: none).
  std::vector<SapStep> sap_plan(const Batch& B, bool ddd, float* out, const float* in, int nsap, bool pupd, float*& fb,
                                long long& fbs);
  void run_sap_plan(const Batch& B, const std::vector<SapStep>& plan, float* fb, long long fbs, int nm,
                    const SapSrcUpdate* upd);
};

constexpr int MAX_BATCH = 64;

}  // namespace qws
