// api.cpp -- the C ABI of libqws_mi355x.so: the device-resident qwsg_*
// interface and the QWS drop-in entry points with their host mirrors.
#include <hip/hip_runtime.h>

#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/qws_mi355x.h"
#include "handle.h"
#include "lattice.h"
#include "multi.h"
#include "serve.h"

using qws::Lattice;

namespace qws {
std::string& last_error() {
  static thread_local std::string err;
  return err;
}
}  // namespace qws

using qws::guarded;

extern "C" {

const char* qwsg_last_error(void) { return qws::last_error().c_str(); }

int qwsg_device_count(int* n) {
  return guarded([&] { QWS_HIP(hipGetDeviceCount(n)); });
}
int qwsg_time_clover_rhs(long long V, int nrep, int reps, double* ms, double* maxdiff) {
  return guarded([&] { qws::time_clover_rhs(V, nrep, reps, ms, maxdiff); });
}
int qwsg_stream_bandwidth(long long bytes, int nin, int reps, double* gbs) {
  return guarded([&] { *gbs = qws::measure_stream(bytes, nin, reps); });
}
int qwsg_stream_bandwidth_one(long long bytes, int nin, int reps, double* gbs) {
  return guarded([&] { *gbs = qws::measure_stream(bytes, nin, reps, true); });
}

int qwsg_create(const int* L, const int* B, int tbc, int links32, qwsg_lattice** out) {
  return guarded([&] {
    Lattice* l = new Lattice(L, B, tbc, links32);
    *out = new qwsg_lattice{l};
  });
}
int qwsg_destroy(qwsg_lattice* h) {
  return guarded([&] {
    if (h) {
      delete h->lat;
      delete h;
    }
  });
}
int qwsg_set_sap_fused(qwsg_lattice* h, int on) {
  return guarded([&] { qws::lat_of(h)->set_fused_sap(on != 0); });
}
int qwsg_set_sap_face_buffer(qwsg_lattice* h, int on) {
  return guarded([&] { qws::lat_of(h)->set_sap_face_buffer(on != 0); });
}
int qwsg_set_sap_source_update(qwsg_lattice* h, int on) {
  return guarded([&] { qws::lat_of(h)->set_sap_source_update(on != 0); });
}
int qwsg_set_kappa(qwsg_lattice* h, double k) {
  return guarded([&] { qws::lat_of(h)->set_kappa(k); });
}
int qwsg_volume(qwsg_lattice* h, long long* V) {
  return guarded([&] { *V = qws::lat_of(h)->volume(); });
}
int qwsg_load_gauge(qwsg_lattice* h, const double* U) {
  return guarded([&] { qws::lat_of(h)->load_gauge(U, 0, 1, 3); });
}
int qwsg_load_clover(qwsg_lattice* h, const double* C) {
  return guarded([&] { qws::lat_of(h)->load_clover(C, 0, 1, 3); });
}
int qwsg_load_gauge_f32(qwsg_lattice* h, const float* U) {
  return guarded([&] { qws::lat_of(h)->load_gauge(U, 0, 0, 2); });
}
int qwsg_load_clover_f32(qwsg_lattice* h, const float* C) {
  return guarded([&] { qws::lat_of(h)->load_clover(C, 0, 0, 2); });
}
int qwsg_generate_gauge(qwsg_lattice* h, double eps, unsigned long long seed) {
  return guarded([&] { qws::lat_of(h)->generate_gauge(eps, seed); });
}
int qwsg_compute_clover(qwsg_lattice* h, double csw) {
  return guarded([&] { qws::lat_of(h)->compute_clover(csw); });
}
int qwsg_gauge_download(qwsg_lattice* h, double* U) {
  return guarded([&] { qws::lat_of(h)->export_field(1, U); });
}
int qwsg_clover_download(qwsg_lattice* h, double* C) {
  return guarded([&] { qws::lat_of(h)->export_field(2, C); });
}
int qwsg_spinor_gaussian(qwsg_lattice* h, int prec, void* dev, unsigned long long seed) {
  return guarded([&] { qws::need("qwsg_spinor_gaussian", dev); qws::lat_of(h)->gaussian_spinor(prec, dev, seed); });
}
int qwsg_spinor_point(qwsg_lattice* h, int prec, void* dev, const int* site, int spin, int colour) {
  return guarded([&] {
    qws::need("qwsg_spinor_point", dev);
    qws::lat_of(h)->point_spinor(prec, dev, site, spin, colour);
  });
}
int qwsg_spinor_alloc(qwsg_lattice* h, int prec, void** dev) {
  return guarded([&] { qws::need("qwsg_spinor_alloc", dev); *dev = qws::lat_of(h)->alloc_spinor(prec); });
}
int qwsg_spinor_free(void* dev) {
  return guarded([&] { QWS_HIP(hipFree(dev)); });
}
int qwsg_spinor_upload(qwsg_lattice* h, int prec, void* dev, const double* host) {
  return guarded([&] {
    qws::need("qwsg_spinor_upload", dev, host);
    qws::lat_of(h)->upload_spinor(prec, dev, host, 0, 1);
  });
}
int qwsg_spinor_download(qwsg_lattice* h, int prec, const void* dev, double* host) {
  return guarded([&] {
    qws::need("qwsg_spinor_download", dev, host);
    qws::lat_of(h)->download_spinor(prec, dev, host, 0, 1);
  });
}
int qwsg_spinor_zero(qwsg_lattice* h, int prec, void* dev) {
  return guarded([&] { qws::need("qwsg_spinor_zero", dev);
    QWS_HIP(hipMemsetAsync(dev, 0, qws::lat_of(h)->spinor_bytes(prec), qws::lat_of(h)->stream()));
  });
}
int qwsg_dslash(qwsg_lattice* h, int prec, void* out, const void* in, int hops, int colour) {
  return guarded([&] { qws::need("qwsg_dslash", out, in); qws::lat_of(h)->dslash(prec, out, in, hops, colour); });
}
int qwsg_face(qwsg_lattice* h, int prec, void* out, const void* in, double factor, int colour) {
  return guarded([&] { qws::need("qwsg_face", out, in); qws::lat_of(h)->face(prec, out, in, factor, colour); });
}
int qwsg_clover(qwsg_lattice* h, int prec, void* out, const void* in) {
  return guarded([&] { qws::need("qwsg_clover", out, in); qws::lat_of(h)->clover(prec, out, in); });
}
int qwsg_jinv(qwsg_lattice* h, int prec, void* x, const void* b, int colour, int nm) {
  return guarded([&] { qws::need("qwsg_jinv", x, b); qws::lat_of(h)->jinv(prec, x, b, colour, nm); });
}
int qwsg_prec_ddd(qwsg_lattice* h, int prec, void* out, const void* in, int nsap, int nm) {
  return guarded([&] { qws::need("qwsg_prec_ddd", out, in); qws::lat_of(h)->prec_ddd(prec, out, in, nsap, nm); });
}
int qwsg_prec_s(qwsg_lattice* h, int prec, void* out, const void* in, int nsap, int nm) {
  return guarded([&] { qws::need("qwsg_prec_s", out, in); qws::lat_of(h)->prec_s(prec, out, in, nsap, nm); });
}
int qwsg_bicgstab_precdd(qwsg_lattice* h, int prec, void* x, const void* b, double tol, int maxiter, int nsap, int nm,
                         int* iters) {
  return guarded([&] { qws::need("qwsg_bicgstab_precdd", x, b);
    int it = qws::lat_of(h)->bicgstab_precdd(prec, x, b, tol, maxiter, nsap, nm);
    if (iters) *iters = it;
  });
}
int qwsg_msap_h(qwsg_lattice* h, void* out, const void* in, int nsap, int nm) {
  return guarded([&] {
    qws::need("qwsg_msap_h", out, in);
    qws::lat_of(h)->msap_h((float*)out, (const float*)in, nsap, nm);
  });
}
int qwsg_bicgstab_precdd_h(qwsg_lattice* h, void* x, const void* b, double tol, int maxiter, int nsap, int nm,
                           int* iters) {
  return guarded([&] { qws::need("qwsg_bicgstab_precdd_h", x, b);
    int it = qws::lat_of(h)->bicgstab_precdd_h((float*)x, (const float*)b, tol, maxiter, nsap, nm);
    if (iters) *iters = it;
  });
}
int qwsg_bicgstab_dd(qwsg_lattice* h, int prec, void* x, const void* b, int maxiter, double tol, int* iters) {
  return guarded([&] { qws::need("qwsg_bicgstab_dd", x, b);
    if (prec != qws::P64 && prec != qws::P32) throw std::invalid_argument("bicgstab_dd: FP64 or FP32 vectors");
    const int it = qws::lat_of(h)->bicgstab_dd(prec, x, b, maxiter, tol);
    if (iters) *iters = it;
  });
}
}  // extern "C"
namespace {
qws::SolveParams solve_params(const qwsg_solve_params* p) {
  if (!p) throw std::invalid_argument("no solve parameters");
  qws::SolveParams sp;
  sp.nsap = p->nsap;
  sp.nm = p->nm;
  sp.tol = p->tol;
  sp.maxiter = p->maxiter;
  sp.tol_s = p->tol_s;
  sp.maxiter_s = p->maxiter_s;
  sp.true_tol = p->true_tol;
  sp.max_restarts = p->max_restarts;
  sp.sap_prec = p->sap_precision;
  qws::check_sap_params(sp.nsap, sp.nm);
  return sp;
}
void put_stats(const qws::SolveStats& s, qwsg_solve_stats* st) {
  st->outer_iters = s.outer_iters;
  st->inner_iters = s.inner_iters;
  st->restarts = s.restarts;
  st->true_residual = s.true_residual;
  st->seconds = s.seconds;
  st->sap_escalated = s.sap_escalated;
}
}  // namespace
extern "C" {

int qwsg_solve(qwsg_lattice* h, void* x, const void* b, const qwsg_solve_params* p, qwsg_solve_stats* st) {
  return guarded([&] { qws::need("qwsg_solve", x, b);
    const qws::SolveParams sp = solve_params(p);
    qws::SolveStats s = qws::lat_of(h)->solve((double*)x, (const double*)b, sp);
    if (st) put_stats(s, st);
  });
}
int qwsg_solve_batch(qwsg_lattice* h, int n, void* const* x, const void* const* b, const qwsg_solve_params* p,
                     qwsg_solve_stats* st) {
  return guarded([&] { qws::need("qwsg_solve_batch", x, b);
    if (n < 1 || n > qws::MAX_BATCH) throw std::invalid_argument("batch size must be in 1..64");
    const qws::SolveParams sp = solve_params(p);
    std::vector<qws::SolveStats> s =
        qws::lat_of(h)->solve_batch((double* const*)x, (const double* const*)b, n, sp);
    for (int k = 0; k < n && st; ++k) put_stats(s[k], st + k);
  });
}
int qwsg_solve_multi(int nlat, qwsg_lattice* const* lats, const int* nrhs, void* const* x, const void* const* b,
                     const qwsg_solve_params* p, qwsg_solve_stats* st) {
  return guarded([&] { qws::need("qwsg_solve_multi", x, b);
    if (nlat < 1 || nlat > 64) throw std::invalid_argument("solve_multi: 1..64 lattices");
    const qws::SolveParams sp = solve_params(p);
    std::vector<Lattice*> L(nlat);
    std::vector<std::vector<double*>> xs(nlat);
    std::vector<std::vector<const double*>> bs(nlat);
    int at = 0;
    for (int i = 0; i < nlat; ++i) {
      if (!lats[i] || !lats[i]->lat) throw std::invalid_argument("solve_multi: no lattice");
      if (nrhs[i] < 1 || nrhs[i] > qws::MAX_BATCH) throw std::invalid_argument("solve_multi: 1..64 RHS per lattice");
      L[i] = lats[i]->lat;
      for (int j = 0; j < i; ++j)
        if (L[j] == L[i]) throw std::invalid_argument("solve_multi: a lattice appears twice");
      for (int k = 0; k < nrhs[i]; ++k, ++at) {
        xs[i].push_back((double*)x[at]);
        bs[i].push_back((const double*)b[at]);
      }
    }
    // every lattice works on the first one's stream for the call
    hipStream_t s0 = L[0]->stream();
    for (Lattice* l : L) l->sync();
    struct Restore {
      std::vector<Lattice*>& L;
      ~Restore() {
        (void)hipStreamSynchronize(L[0]->stream());
        for (size_t i = 1; i < L.size(); ++i) L[i]->set_stream(nullptr);
      }
    } restore{L};
    for (Lattice* l : L) l->set_stream(s0);
    qws::SapFusion fz(s0);
    const auto res = qws::Fused::solve(L, xs, bs, sp, &fz);
    at = 0;
    for (int i = 0; i < nlat; ++i)
      for (int k = 0; k < nrhs[i]; ++k, ++at)
        if (st) put_stats(res[i][k], st + at);
  });
}
int qwsg_eo_load_gauge(qwsg_lattice* h, const double* U) {
  return guarded([&] { qws::lat_of(h)->eo_load_gauge(U, 0, 1, 1, 3); });
}
int qwsg_eo_load_clover(qwsg_lattice* h, const double* C) {
  return guarded([&] { qws::lat_of(h)->eo_load_clover(C, 0, 1, 1, 3); });
}
int qwsg_eo_from_lattice(qwsg_lattice* h) {
  return guarded([&] { qws::lat_of(h)->eo_fields_from_lattice(); });
}
int qwsg_eo_spinor_alloc(qwsg_lattice* h, int prec, void** dev) {
  return guarded([&] { qws::need("qwsg_eo_spinor_alloc", dev);
    if (prec != qws::P64 && prec != qws::P32) throw std::invalid_argument("even/odd spinors are FP64 or FP32");
    *dev = nullptr;
    QWS_HIP(hipMalloc(dev, qws::lat_of(h)->eo_spinor_bytes(prec)));
    QWS_HIP(hipMemset(*dev, 0, qws::lat_of(h)->eo_spinor_bytes(prec)));
  });
}
int qwsg_eo_spinor_upload(qwsg_lattice* h, int prec, void* dev, const double* host) {
  return guarded([&] {
    qws::need("qwsg_eo_spinor_upload", dev, host);
    qws::lat_of(h)->eo_upload_spinor(prec, dev, host, 0, 1, 1);
  });
}
int qwsg_eo_spinor_download(qwsg_lattice* h, int prec, const void* dev, double* host) {
  return guarded([&] {
    qws::need("qwsg_eo_spinor_download", dev, host);
    qws::lat_of(h)->eo_download_spinor(prec, dev, host, 0, 1, 1);
  });
}
int qwsg_deo(qwsg_lattice* h, int prec, int pe, int set, int dag, int clv, int mode, double fac, void* out,
             const void* in, const void* in0) {
  return guarded([&] {
    qws::need("qwsg_deo", out, in);
    qws::lat_of(h)->deo(prec, pe, set, dag, clv, mode, fac, out, in, in0);
  });
}
int qwsg_mtilde(qwsg_lattice* h, int prec, int pce, int which, void* out, const void* in) {
  return guarded([&] { qws::need("qwsg_mtilde", out, in);
    const double k = qws::lat_of(h)->kappa();
    qws::lat_of(h)->mtilde(prec, pce, which, -k * k, out, in);
  });
}
struct qwsg_server {
  qws::Server* srv;
};
int qwsg_server_create(const qwsg_field_spec* fields, int nfields, int links32, int tbc, double csw,
                       const qwsg_solve_params* p, int slots, double mem_budget, qwsg_server** out) {
  return guarded([&] {
    if (nfields < 1) throw std::invalid_argument("a server needs at least one field");
    std::vector<qws::FieldSpec> fs(nfields);
    for (int i = 0; i < nfields; ++i) {
      for (int m = 0; m < 4; ++m) {
        fs[i].L[m] = fields[i].L[m];
        fs[i].B[m] = fields[i].B[m];
      }
      fs[i].eps = fields[i].eps;
      fs[i].seed = fields[i].seed;
    }
    qws::ServeConfig cfg;
    cfg.links32 = links32;
    cfg.tbc = tbc;
    cfg.csw = csw;
    cfg.slots = slots;
    cfg.mem_budget = mem_budget;
    cfg.prm.nsap = p->nsap;
    cfg.prm.nm = p->nm;
    cfg.prm.tol = p->tol;
    cfg.prm.maxiter = p->maxiter;
    cfg.prm.tol_s = p->tol_s;
    cfg.prm.maxiter_s = p->maxiter_s;
    cfg.prm.true_tol = p->true_tol;
    cfg.prm.max_restarts = p->max_restarts;
    cfg.prm.sap_prec = p->sap_precision;
    qws::check_sap_params(cfg.prm.nsap, cfg.prm.nm);
    *out = new qwsg_server{new qws::Server(fs, cfg)};
  });
}
int qwsg_server_run(qwsg_server* h, const qwsg_request* reqs, int nreq, qwsg_request_result* out) {
  return guarded([&] {
    if (!h || !h->srv) throw std::invalid_argument("null server handle");
    if (nreq < 0 || (nreq > 0 && (!reqs || !out))) throw std::invalid_argument("server_run: bad request array");
    std::vector<qws::Request> tr(nreq);
    for (int i = 0; i < nreq; ++i) tr[i] = {reqs[i].field, reqs[i].kappa, reqs[i].nrhs, reqs[i].seed, reqs[i].arrival};
    std::vector<qws::RequestResult> r = h->srv->run(tr);
    for (int i = 0; i < nreq; ++i) {
      out[i].arrival = r[i].arrival;
      out[i].admit = r[i].admit;
      out[i].finish = r[i].finish;
      out[i].outer_iters = r[i].outer_iters;
      out[i].inner_iters = r[i].inner_iters;
      out[i].restarts = r[i].restarts;
      out[i].true_residual_max = r[i].true_residual_max;
      out[i].bytes = r[i].bytes;
      out[i].digest = r[i].digest;
    }
  });
}
int qwsg_server_request_bytes(long long V, int nrhs, double* bytes) {
  return guarded([&] {
    if (V < 1 || nrhs < 1 || nrhs > qws::MAX_BATCH) throw std::invalid_argument("server_request_bytes: V >= 1, 1..64 RHS");
    *bytes = qws::Server::request_bytes(V, qws::Server::slot_capacity(nrhs));
  });
}
int qwsg_server_destroy(qwsg_server* h) {
  return guarded([&] {
    if (h) {
      delete h->srv;
      delete h;
    }
  });
}
int qwsg_eo_bicgstab(qwsg_lattice* h, void* x, const void* b, int maxiter, double tol, int pce, int* iters) {
  return guarded([&] { qws::need("qwsg_eo_bicgstab", x, b);
    const double k = qws::lat_of(h)->kappa();
    const int it = qws::lat_of(h)->bicgstab_eo((double*)x, (const double*)b, maxiter, tol, pce, -k * k);
    if (iters) *iters = it;
  });
}
int qwsg_eo_cg(qwsg_lattice* h, void* x, const void* b, int maxiter, double tol, int pce, int* iters) {
  return guarded([&] { qws::need("qwsg_eo_cg", x, b);
    const double k = qws::lat_of(h)->kappa();
    const int it = qws::lat_of(h)->cg_eo((double*)x, (const double*)b, maxiter, tol, pce, -k * k);
    if (iters) *iters = it;
  });
}
int qwsg_eo_mcg(qwsg_lattice* h, void* const* xs, const void* b, int nshift, const double* sigma, int maxiter,
                double tol, int pce, int* iters) {
  return guarded([&] { qws::need("qwsg_eo_mcg", xs, b);
    const double k = qws::lat_of(h)->kappa();
    const int it =
        qws::lat_of(h)->mcg_eo((double* const*)xs, (const double*)b, nshift, sigma, maxiter, tol, pce, -k * k);
    if (iters) *iters = it;
  });
}
int qwsg_norm2(qwsg_lattice* h, int prec, const void* a, double* out) {
  return guarded([&] { qws::need("qwsg_norm2", a); *out = qws::lat_of(h)->norm2(prec, a); });
}
int qwsg_spinor_digest(qwsg_lattice* h, int prec, const void* a, unsigned long long* out) {
  return guarded([&] { qws::need("qwsg_spinor_digest", a); *out = qws::lat_of(h)->digest(prec, a); });
}
int qwsg_synchronize(qwsg_lattice* h) {
  return guarded([&] { qws::lat_of(h)->sync(); });
}
int qwsg_time_dslash(qwsg_lattice* h, int prec, int reps, double* ms) {
  return guarded([&] { *ms = qws::lat_of(h)->time_dslash(prec, reps); });
}
int qwsg_set_dslash_kernel(qwsg_lattice* h, int k) {
  return guarded([&] { qws::lat_of(h)->set_dslash_kernel(k); });
}
int qwsg_autotune_solve(qwsg_lattice* h, void* x, const void* b, const qwsg_solve_params* p, const int* nsap,
                        int nnsap, const int* nm, int nnm, qwsg_solve_params* best, double* seconds) {
  return guarded([&] { qws::need("qwsg_autotune_solve", x, b);
    if (!p || !best || !nsap || !nm || nnsap < 1 || nnm < 1) throw std::invalid_argument("autotune_solve: bad arguments");
    const qws::SolveParams sp = solve_params(p);
    std::vector<double> sec;
    const qws::SolveParams b2 = qws::lat_of(h)->autotune_solve((double*)x, (const double*)b, sp,
                                                       std::vector<int>(nsap, nsap + nnsap),
                                                       std::vector<int>(nm, nm + nnm), sec);
    *best = *p;
    best->nsap = b2.nsap;
    best->nm = b2.nm;
    if (seconds) std::copy(sec.begin(), sec.end(), seconds);
  });
}
int qwsg_autotune_blocks(qwsg_lattice* h, const void* b, const qwsg_solve_params* p, const int* shapes, int nshape,
                         int* best, double* seconds) {
  return guarded([&] { qws::need("qwsg_autotune_blocks", b, p, shapes, best);
    if (nshape < 1) throw std::invalid_argument("autotune_blocks: no candidate shapes");
    std::vector<double> sec;
    *best = qws::lat_of(h)->autotune_blocks((const double*)b, solve_params(p),
                                            std::vector<int>(shapes, shapes + 4 * (size_t)nshape), sec);
    if (seconds) std::copy(sec.begin(), sec.end(), seconds);
  });
}
int qwsg_set_dslash_order(qwsg_lattice* h, int order) {
  return guarded([&] { qws::lat_of(h)->set_dslash_order(order); });
}
int qwsg_get_dslash_order(qwsg_lattice* h, int* order) {
  return guarded([&] { *order = qws::lat_of(h)->dslash_order(); });
}
int qwsg_autotune_dslash(qwsg_lattice* h, int reps, int* chosen) {
  return guarded([&] {
    const int k = qws::lat_of(h)->autotune_dslash(reps);
    if (chosen) *chosen = k;
  });
}
int qwsg_time_sap(qwsg_lattice* h, int prec, int reps, int nsap, int nm, double* ms) {
  return guarded([&] { *ms = qws::lat_of(h)->time_sap(prec, reps, nsap, nm); });
}

// ============================================================================ QWS drop-in
// Host mirrors of QWS's globals (qws.cc:74-124).
glud_t* glud = nullptr;
glus_t* glus = nullptr;
clvd_t* clvd = nullptr;
clvs_t* clvs = nullptr;
double kappa = 0, kappa2 = 0, mkappa = 0;
int nx = 0, ny = 0, nz = 0, nt = 0, nxh = 0, nxd = 0, nxs = 0, vold = 0, vols = 0;
int npe[4] = {1, 1, 1, 1};
double fbc[4][2] = {{2, 2}, {2, 2}, {2, 2}, {2, 2}};
int rank = 0, domain_e = 0, domain_o = 1;
// qws.cc:124 -- per-rank output file of QWS's drivers (opened there only for per-rank
// timing output; never opened here)
FILE* para_outputfile = nullptr;

// BQCD provides these (qws_bqcd.cc:56-57); weak so that the library loads without BQCD.
int std_xyzt2i_(int*) __attribute__((weak));
int e_o_(int*) __attribute__((weak));

}  // extern "C"

namespace {

struct DropIn {
  Lattice* lat = nullptr;
  // the host arrays changed since the device fields were built from them: links and
  // clover, for the DD reading (dirty_*) and the even/odd reading (eo_dirty_*)
  bool dirty_g = true, dirty_c = true;
  bool eo_dirty_g = true, eo_dirty_c = true;
  int* bqcd_map = nullptr;  // device: 2 (i - 1) + eo per lexicographic site (BQCD's numbering)
  bool default_blocks = true;
  int tbc = -1;
  int blocks[4] = {0, 0, 0, 0};  // the SAP blocks of the device lattice
  int links32 = 18;              // FP32 link storage on the device (12: row 2 rebuilt)
  int* pce = nullptr;
  int* pco = nullptr;
  void* dev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};  // d0 d1 s0 s1 pre_d pre_s
  void* eo[2][4] = {{nullptr, nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr, nullptr}};  // [P64/P32][in out in0 pre]
  void* eo_multi = nullptr;  // mcg_vm_'s solutions (one FP64 parity field per shift)
  size_t eo_multi_bytes = 0;
  std::vector<char> host_tmp[2];
};
DropIn g_qws;

[[noreturn]] void fatal(const char* where, const std::string& what) {
  std::printf("ERROR at %s : %s\n", where, what.c_str());
  std::fflush(stdout);
  std::exit(1);
}

template <typename F> void drop_in(const char* where, F&& f) {
  try {
    f();
  } catch (const std::exception& e) {
    fatal(where, e.what());
  }
}

void mark_fields_modified() {
  g_qws.dirty_g = g_qws.dirty_c = true;
  g_qws.eo_dirty_g = g_qws.eo_dirty_c = true;
}

Lattice& lat_sync(const char* where) {
  if (!g_qws.lat) fatal(where, "qws_init_ has not been called");
  if (g_qws.dirty_g) {
    g_qws.lat->load_gauge(glud, 1, 1, 1);
    g_qws.lat->load_gauge(glus, 1, 0, 2);
    g_qws.dirty_g = false;
  }
  if (g_qws.dirty_c) {
    g_qws.lat->load_clover(clvd, 1, 1, 1);
    g_qws.lat->load_clover(clvs, 1, 0, 2);
    g_qws.dirty_c = false;
  }
  g_qws.lat->set_kappa(kappa);  // the caller may set kappa directly (main.cc:173-175)
  return *g_qws.lat;
}

void* dbuf(int i, int prec) {
  if (!g_qws.dev[i]) g_qws.dev[i] = g_qws.lat->alloc_spinor(prec);
  return g_qws.dev[i];
}

size_t dom_bytes_d() { return sizeof(scd_t) * (size_t)vold; }
size_t dom_bytes_s() { return sizeof(scs_t) * (size_t)vols; }

// a full-lattice host AoSoA array holding `part` (one domain) at domain d, zero elsewhere
const void* embed(int slot, const void* part, size_t dom_bytes, int d) {
  auto& buf = g_qws.host_tmp[slot];
  buf.assign(2 * dom_bytes, 0);
  std::memcpy(buf.data() + d * dom_bytes, part, dom_bytes);
  return buf.data();
}

void require_domain_blocks(const char* where) {
  if (!g_qws.default_blocks) fatal(where, "per-domain entry points need the default SAP domains (x-halves)");
}

void release_drop_in() {
  // (frees on the way out: a failure here has nothing left to report to)
  for (auto& p : g_qws.dev)
    if (p) {
      (void)hipFree(p);
      p = nullptr;
    }
  for (auto& row : g_qws.eo)
    for (auto& p : row)
      if (p) {
        (void)hipFree(p);
        p = nullptr;
      }
  if (g_qws.eo_multi) (void)hipFree(g_qws.eo_multi);
  g_qws.eo_multi = nullptr;
  if (g_qws.bqcd_map) (void)hipFree(g_qws.bqcd_map);
  g_qws.bqcd_map = nullptr;
  g_qws.eo_multi_bytes = 0;
  delete g_qws.lat;
  g_qws.lat = nullptr;
}

void create_drop_in_lattice(const int B[4]) {
  int L[4] = {nx, ny, nz, nt};
  release_drop_in();
  for (int i = 0; i < 4; ++i) g_qws.blocks[i] = B[i];
  g_qws.lat = new Lattice(L, B, g_qws.tbc, g_qws.links32);
  mark_fields_modified();
}

// The even/odd family reads glud/glus and clvd/clvs as two parity halves of
// x-compressed sites (qws_bqcd.cc:82-85); upload that reading when they changed.
Lattice& lat_sync_eo(const char* where) {
  Lattice& L = lat_sync(where);
  if (g_qws.eo_dirty_g) {
    L.eo_load_gauge(glud, 1, QWS_VLEND, 1, 1);
    L.eo_load_gauge(glus, 1, QWS_VLENS, 0, 2);
    g_qws.eo_dirty_g = false;
  }
  if (g_qws.eo_dirty_c) {
    L.eo_load_clover(clvd, 1, QWS_VLEND, 1, 1);
    L.eo_load_clover(clvs, 1, QWS_VLENS, 0, 2);
    g_qws.eo_dirty_c = false;
  }
  return L;
}

void* eobuf(int prec, int i) {
  void*& p = g_qws.eo[prec == qws::P64 ? 0 : 1][i];
  if (!p) {
    if (hipMalloc(&p, g_qws.lat->eo_spinor_bytes(prec)) != hipSuccess) throw std::runtime_error("hipMalloc failed");
  }
  return p;
}

// n consecutive FP64 parity fields (grown on demand, kept until qws_finalize_)
void* eobuf_n(int prec, int n) {
  const size_t need = g_qws.lat->eo_spinor_bytes(prec) * (size_t)n;
  if (g_qws.eo_multi_bytes < need) {
    if (g_qws.eo_multi) (void)hipFree(g_qws.eo_multi);
    g_qws.eo_multi = nullptr;
    g_qws.eo_multi_bytes = 0;
    if (hipMalloc(&g_qws.eo_multi, need) != hipSuccess) throw std::runtime_error("hipMalloc failed");
    g_qws.eo_multi_bytes = need;
  }
  return g_qws.eo_multi;
}

// out (parity *pe, one parity field of the host layout) from in (parity 1 - *pe):
// set/dag/clv/mode as Lattice::deo; mode 1 adds to out, mode 2 starts from in0;
// set 2 takes its input from the snapshot recorded by the matching *_out_pre_*.
void eo_call(const char* where, int prec, const int* pe, const int* po, int set, int dag, int clv, int mode, double fac,
             void* out, const void* in, const void* in0) {
  if ((*pe != 0 && *pe != 1) || *po != 1 - *pe) fatal(where, "parities must be 0 and 1");
  drop_in(where, [&] {
    Lattice& L = lat_sync_eo(where);
    const int vlen = prec == qws::P64 ? QWS_VLEND : QWS_VLENS;
    const int hd = prec == qws::P64 ? 1 : 0;
    void* bi = eobuf(prec, 0);
    void* bo = eobuf(prec, 1);
    void* b0 = nullptr;
    if (set == 2) {
      bi = eobuf(prec, 3);
    } else {
      L.eo_upload_spinor(prec, bi, in, 1, vlen, hd);
    }
    if (mode == 1) L.eo_upload_spinor(prec, bo, out, 1, vlen, hd);
    if (mode == 2) {
      b0 = eobuf(prec, 2);
      L.eo_upload_spinor(prec, b0, in0, 1, vlen, hd);
    }
    L.deo(prec, *pe, set, dag, clv, mode, fac, bo, bi, b0);
    L.eo_download_spinor(prec, bo, out, 1, vlen, hd);
  });
}

void eo_pre(const char* where, int prec, const void* in) {
  drop_in(where, [&] {
    Lattice& L = lat_sync_eo(where);
    L.eo_upload_spinor(prec, eobuf(prec, 3), in, 1, prec == qws::P64 ? QWS_VLEND : QWS_VLENS, prec == qws::P64);
  });
}

void eo_mtilde(const char* where, int prec, int which, double fac, void* out, const void* in) {
  if (!g_qws.pce || (*g_qws.pce != 0 && *g_qws.pce != 1)) fatal(where, "pce must point at 0 or 1");
  drop_in(where, [&] {
    Lattice& L = lat_sync_eo(where);
    const int vlen = prec == qws::P64 ? QWS_VLEND : QWS_VLENS;
    const int hd = prec == qws::P64 ? 1 : 0;
    L.eo_upload_spinor(prec, eobuf(prec, 0), in, 1, vlen, hd);
    L.mtilde(prec, *g_qws.pce, which, fac, eobuf(prec, 1), eobuf(prec, 0));
    L.eo_download_spinor(prec, eobuf(prec, 1), out, 1, vlen, hd);
  });
}

}  // namespace

extern "C" {

void qws_init_(int* lx, int* ly, int* lz, int* lt, int* npe_f, int* fbc_f, int* pce_f, int* pco_f, int* block_size) {
  (void)block_size;
  g_qws.pce = pce_f;  // kept, as the reference does (qws.cc:134-135)
  g_qws.pco = pco_f;
  for (int i = 0; i < 4; ++i) npe[i] = npe_f[i];
  if (npe[0] * npe[1] * npe[2] * npe[3] != 1)
    fatal(__func__, "this library runs one rank on one GPU: npe must be (1,1,1,1)");
  nx = *lx;
  ny = *ly;
  nz = *lz;
  nt = *lt;
  nxh = nx / 2;
  nxd = nx / 2 / QWS_VLEND;
  nxs = nx / 2 / QWS_VLENS;
  if (nxd == 0) { std::printf("ERROR at %s : nx/2/VLEND must be >=1\n", __func__); std::exit(1); }
  if (nxs == 0) { std::printf("ERROR at %s : nx/2/VLENS must be >=1\n", __func__); std::exit(1); }
  if (nxh % QWS_VLENS != 0) fatal(__func__, "nx/2 must be a multiple of 16");
  vold = nxd * ny * nz * nt;
  vols = nxs * ny * nz * nt;
  rank = 0;
  domain_e = 0;  // (py+pz+pt)%2 on the single rank (qws.cc:275-276)
  domain_o = 1;
  // t boundary (qws.cc:298-310, single rank: pt == 0 == npe[3]-1)
  if (fbc_f[3] != 1) {
    fbc[3][0] = fbc[3][1] = (double)(fbc_f[3] * 2);
  } else {
    fbc[3][0] = fbc[3][1] = 2;
  }
  if (fbc_f[3] != 1 && fbc_f[3] != -1) fatal(__func__, "fbc_f[3] must be +1 or -1");
  g_qws.tbc = fbc_f[3];
  auto alloc = [](size_t bytes) {
    void* p = nullptr;
    if (posix_memalign(&p, 256, bytes) != 0) fatal("qws_init_", "host allocation failed");
    std::memset(p, 0, bytes);
    return p;
  };
  std::free(glud);
  std::free(glus);
  std::free(clvd);
  std::free(clvs);
  glud = (glud_t*)alloc(sizeof(glud_t) * (size_t)vold * 4 * 2);
  clvd = (clvd_t*)alloc(sizeof(clvd_t) * (size_t)vold * 2);
  glus = (glus_t*)alloc(sizeof(glus_t) * (size_t)vols * 4 * 2);
  clvs = (clvs_t*)alloc(sizeof(clvs_t) * (size_t)vols * 2);
  int B[4] = {nxh, ny, nz, nt};
  g_qws.default_blocks = true;
  g_qws.links32 = 18;
  drop_in(__func__, [&] { create_drop_in_lattice(B); });
}

void qws_finalize_(void) {
  release_drop_in();
  std::free(glud);
  std::free(glus);
  std::free(clvd);
  std::free(clvs);
  glud = nullptr;
  glus = nullptr;
  clvd = nullptr;
  clvs = nullptr;
}

void qws_mi355x_fields_modified_(void) {
  mark_fields_modified();
}

void qws_gpu_sync_fields_(void) {
  drop_in(__func__, [&] {
    mark_fields_modified();
    lat_sync("qws_gpu_sync_fields_");
    QWS_HIP(hipStreamSynchronize(g_qws.lat->stream()));
  });
}

void qws_mi355x_set_sap_blocks_(int* b) {
  if (!g_qws.lat) fatal(__func__, "qws_init_ has not been called");
  int B[4] = {b[0], b[1], b[2], b[3]};
  g_qws.default_blocks = (B[0] == nxh && B[1] == ny && B[2] == nz && B[3] == nt);
  drop_in(__func__, [&] { create_drop_in_lattice(B); });
}

void qws_mi355x_set_links12_(int* on) {
  if (!g_qws.lat) fatal(__func__, "qws_init_ has not been called");
  g_qws.links32 = (on && *on) ? 12 : 18;
  const int B[4] = {g_qws.blocks[0], g_qws.blocks[1], g_qws.blocks[2], g_qws.blocks[3]};
  drop_in(__func__, [&] { create_drop_in_lattice(B); });
}

// ---- operators on host AoSoA arrays
void ddd_d_(scd_t* out, scd_t* in) {
  drop_in(__func__, [&] {
    Lattice& L = lat_sync("ddd_d_");
    void* a = dbuf(0, qws::P64);
    void* b = dbuf(1, qws::P64);
    L.upload_spinor(qws::P64, a, in, 1, 1);
    L.dslash(qws::P64, b, a, qws::HOPS_ALL, -1);
    L.download_spinor(qws::P64, b, out, 1, 1);
  });
}

void ddd_s_(scs_t* out, scs_t* in) {
  drop_in(__func__, [&] {
    Lattice& L = lat_sync("ddd_s_");
    void* a = dbuf(2, qws::P32);
    void* b = dbuf(3, qws::P32);
    L.upload_spinor(qws::P32, a, in, 1, 0);
    L.dslash(qws::P32, b, a, qws::HOPS_ALL, -1);
    L.download_spinor(qws::P32, b, out, 1, 0);
  });
}

// domain-local operator: out, in point at domain d's data (ddd_in_s.cc:72-84)
static void ddd_in_generic(int prec, void* out, const void* in, int d, const char* where) {
  require_domain_blocks(where);
  Lattice& L = lat_sync(where);
  const bool dbl = prec == qws::P64;
  const size_t db = dbl ? dom_bytes_d() : dom_bytes_s();
  void* a = dbuf(dbl ? 0 : 2, prec);
  void* b = dbuf(dbl ? 1 : 3, prec);
  L.upload_spinor(prec, a, embed(0, in, db, d), 1, dbl);
  L.dslash(prec, b, a, qws::HOPS_INSIDE, d);
  auto& full = g_qws.host_tmp[1];
  full.resize(2 * db);
  L.download_spinor(prec, b, full.data(), 1, dbl);
  std::memcpy(out, full.data() + d * db, db);
}

void ddd_in_s_(scs_t* out, scs_t* in, int* domain) {
  drop_in(__func__, [&] { ddd_in_generic(qws::P32, out, in, *domain, "ddd_in_s_"); });
}
void ddd_in_d_(scd_t* out, scd_t* in, int* domain) {
  drop_in(__func__, [&] { ddd_in_generic(qws::P64, out, in, *domain, "ddd_in_d_"); });
}

// pre: record the faces of `in` (the whole field), pos: out (domain d's data) +=
// factor T^-1 H_face(recorded field) (ddd_out_s.cc:112-626, ddd_out_d.cc:84-235)
static void ddd_out_pre_generic(int prec, const void* in, const char* where) {
  require_domain_blocks(where);
  Lattice& L = lat_sync(where);
  const bool dbl = prec == qws::P64;
  L.upload_spinor(prec, dbuf(dbl ? 4 : 5, prec), in, 1, dbl);
}
static void ddd_out_pos_generic(int prec, void* out, int d, double factor, const char* where) {
  require_domain_blocks(where);
  Lattice& L = lat_sync(where);
  const bool dbl = prec == qws::P64;
  const size_t db = dbl ? dom_bytes_d() : dom_bytes_s();
  void* o = dbuf(dbl ? 0 : 2, prec);
  L.upload_spinor(prec, o, embed(0, out, db, d), 1, dbl);
  L.face(prec, o, dbuf(dbl ? 4 : 5, prec), factor, d);
  auto& full = g_qws.host_tmp[1];
  full.resize(2 * db);
  L.download_spinor(prec, o, full.data(), 1, dbl);
  std::memcpy(out, full.data() + d * db, db);
}
void ddd_out_pre_s_(scs_t* in, int* domain) {
  (void)domain;
  drop_in(__func__, [&] { ddd_out_pre_generic(qws::P32, in, "ddd_out_pre_s_"); });
}
void ddd_out_pos_s_(scs_t* out, scs_t* in, int* domain, float factor) {
  (void)in;
  drop_in(__func__, [&] { ddd_out_pos_generic(qws::P32, out, *domain, factor, "ddd_out_pos_s_"); });
}
void ddd_out_pre_d_(scd_t* in, int* domain) {
  (void)domain;
  drop_in(__func__, [&] { ddd_out_pre_generic(qws::P64, in, "ddd_out_pre_d_"); });
}
void ddd_out_pos_d_(scd_t* out, scd_t* in, int* domain) {
  (void)in;
  drop_in(__func__, [&] { ddd_out_pos_generic(qws::P64, out, *domain, mkappa, "ddd_out_pos_d_"); });
}

void jinv_ddd_in_s_(scs_t* x, scs_t* b, int* domain, int* maxiter) {
  drop_in(__func__, [&] {
    require_domain_blocks("jinv_ddd_in_s_");
    Lattice& L = lat_sync("jinv_ddd_in_s_");
    const int d = *domain;
    const size_t db = dom_bytes_s();
    void* bb = dbuf(2, qws::P32);
    void* xx = dbuf(3, qws::P32);
    L.upload_spinor(qws::P32, bb, embed(0, b, db, d), 1, 0);
    L.jinv(qws::P32, xx, bb, d, *maxiter);
    auto& full = g_qws.host_tmp[1];
    full.resize(2 * db);
    L.download_spinor(qws::P32, xx, full.data(), 1, 0);
    std::memcpy(x, full.data() + d * db, db);
  });
}

void prec_ddd_s_(scs_t* out, scs_t* in, int* nsap, int* nm) {
  if (*nsap > 10) { std::printf("prec_ddd_s_ nsap > 10\n"); std::exit(1); }
  if (*nm > 10) { std::printf("prec_ddd_s_ nn   > 10\n"); std::exit(1); }
  drop_in(__func__, [&] {
    Lattice& L = lat_sync("prec_ddd_s_");
    void* a = dbuf(2, qws::P32);
    void* b = dbuf(3, qws::P32);
    L.upload_spinor(qws::P32, a, in, 1, 0);
    L.prec_ddd(qws::P32, b, a, *nsap, *nm);
    L.download_spinor(qws::P32, b, out, 1, 0);
  });
}

void prec_s_(scs_t* out, scs_t* in, int* nsap, int* nm) {
  drop_in(__func__, [&] {
    Lattice& L = lat_sync("prec_s_");
    void* a = dbuf(2, qws::P32);
    void* b = dbuf(3, qws::P32);
    L.upload_spinor(qws::P32, a, in, 1, 0);
    L.prec_s(qws::P32, b, a, *nsap, *nm);
    L.download_spinor(qws::P32, b, out, 1, 0);
  });
}

void bicgstab_precdd_s_(scs_t* x, scs_t* b, double* tol, int* conviter, int* maxiter, int* nsap, int* nm) {
  drop_in(__func__, [&] {
    Lattice& L = lat_sync("bicgstab_precdd_s_");
    void* bb = dbuf(2, qws::P32);
    void* xx = dbuf(3, qws::P32);
    L.upload_spinor(qws::P32, bb, b, 1, 0);
    int it = L.bicgstab_precdd(qws::P32, xx, bb, *tol, *maxiter, *nsap, *nm);
    L.download_spinor(qws::P32, xx, x, 1, 0);
    if (conviter) *conviter = it;
  });
}

// The outer solve on the host arrays; returns |b - D x| / |b| (FP64, on the GPU) when
// true_res is set.  Throws on any failure; the callers decide whether that ends the process.
static bool dd_mix_body(const char* where, scd_t* x, const scd_t* b, const double* tol, int* conviter,
                        const int* maxiter, const double* tol_s, const int* maxiter_s, const int* nsap, const int* nm,
                        int sap_prec, double* true_res) {
    if (!g_qws.lat) throw std::runtime_error("qws_init_ has not been called");
    Lattice& L = lat_sync(where);
    double* xx = (double*)dbuf(0, qws::P64);
    double* bb = (double*)dbuf(1, qws::P64);
    L.upload_spinor(qws::P64, xx, x, 1, 1);
    L.upload_spinor(qws::P64, bb, b, 1, 1);
    qws::SolveParams p;
    p.tol = *tol;
    p.maxiter = *maxiter;
    p.tol_s = *tol_s;
    p.maxiter_s = *maxiter_s;
    p.nsap = *nsap;
    p.nm = *nm;
    p.sap_prec = sap_prec;
    p.escalate = 0;  // the reference's method throughout
    // bicgstab_dd_mix.cc:165 runs iter = 1..maxiter-1 and reports the last iter;
    // bicgstab_dd_mix2_hf.cc:169 runs iter = 0..maxiter-1 and reports iter only on convergence
    if (sap_prec == qws::P16) p.maxiter = *maxiter + 1;
    int inner = 0;
    int it = L.bicgstab_dd_mix(xx, bb, p, &inner);
    if (true_res) *true_res = L.true_residual(xx, bb);
    L.download_spinor(qws::P64, xx, x, 1, 1);
    if (sap_prec != qws::P16) {
      if (conviter) *conviter = it;
    } else if (it < p.maxiter && conviter) {
      *conviter = it - 1;
    }
    // the recursive residual met tol within the iterations run (a solve that ran out
    // reports the iteration limit)
    return it < p.maxiter;
}

static void dd_mix_generic(const char* where, scd_t* x, const scd_t* b, const double* tol, int* conviter,
                           const int* maxiter, const double* tol_s, const int* maxiter_s, const int* nsap, const int* nm,
                           int sap_prec) {
  // the reference's inner solve ends the process in prec_ddd_s_ (qws.cc:798-799)
  if (*nsap > 10) { std::printf("prec_ddd_s_ nsap > 10\n"); std::exit(1); }
  if (*nm > 10) { std::printf("prec_ddd_s_ nn   > 10\n"); std::exit(1); }
  drop_in(where, [&] { dd_mix_body(where, x, b, tol, conviter, maxiter, tol_s, maxiter_s, nsap, nm, sap_prec, nullptr); });
}

void bicgstab_dd_mix_(scd_t* x, scd_t* b, double* tol, int* conviter, int* maxiter, double* tol_s, int* maxiter_s,
                      int* nsap, int* nm) {
  dd_mix_generic("bicgstab_dd_mix_", x, b, tol, conviter, maxiter, tol_s, maxiter_s, nsap, nm, qws::P32);
}

// The status variants: the same solve, the outcome as a return code and the FP64 true
// residual, errors on stderr instead of ending the process.
static int dd_mix_status(const char* where, scd_t* x, const scd_t* b, const double* tol, int* conviter,
                         const int* maxiter, const double* tol_s, const int* maxiter_s, const int* nsap, const int* nm,
                         int sap_prec, double* true_residual) {
  if (true_residual) *true_residual = -1;
  if (!g_qws.lat) {
    std::fprintf(stderr, "ERROR at %s : qws_init_ has not been called\n", where);
    return -1;
  }
  if (!x || !b || !tol || !maxiter || !tol_s || !maxiter_s || !nsap || !nm) {
    std::fprintf(stderr, "ERROR at %s : null argument\n", where);
    return -2;
  }
  if (*nsap < 1 || *nsap > 10 || *nm < 1 || *nm > 10) {
    std::fprintf(stderr, "ERROR at %s : nsap and nm must be in 1..10\n", where);
    return -2;
  }
  double res = -1;
  bool converged = false;
  try {
    converged = dd_mix_body(where, x, b, tol, conviter, maxiter, tol_s, maxiter_s, nsap, nm, sap_prec, &res);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "ERROR at %s : %s\n", where, e.what());
    return -3;
  }
  if (true_residual) *true_residual = res;
  if (!std::isfinite(res)) return 2;
  return (*tol < 0 || converged) ? 0 : 1;
}

int bicgstab_dd_mix_status_(scd_t* x, scd_t* b, double* tol, int* conviter, int* maxiter, double* tol_s,
                            int* maxiter_s, int* nsap, int* nm, double* true_residual) {
  return dd_mix_status(__func__, x, b, tol, conviter, maxiter, tol_s, maxiter_s, nsap, nm, qws::P32, true_residual);
}
int bicgstab_dd_mix2_hf_status_(scd_t* x, const scd_t* b, const double* tol, int* conviter, const int* maxiter,
                                const double* tol_s, const int* maxiter_s, const int* nsap, const int* nm,
                                double* true_residual) {
  return dd_mix_status(__func__, x, b, tol, conviter, maxiter, tol_s, maxiter_s, nsap, nm, qws::P16, true_residual);
}

// ---- half-precision variant (qws_h.cc, prec_ddd_s_h.h, bicgstab_precdd_s_h.cc, bicgstab_dd_mix2_hf.cc)
void qws_h_init_(const glus_t* u, const clvs_t* c) {
  // The FP16 copies are made on the device from the FP32 fields the library holds;
  // the reference takes the FP32 arrays it copies here (qws_h.cc:150-166).
  if (u != glus || c != clvs) fatal("qws_h_init_", "expects the library's glus and clvs arrays");
  drop_in(__func__, [&] { lat_sync("qws_h_init_"); });
}

void assign_mult_mSAP_s_h_(scs_t* xs, const scs_t* bs, const int* nsap, const int* nm) {
  if (*nsap > 10) { std::printf("%s : %s nsap > 10\n", __FILE__, __func__); std::exit(1); }
  if (*nm > 10) { std::printf("%s : %s nn   > 10\n", __FILE__, __func__); std::exit(1); }
  drop_in(__func__, [&] {
    Lattice& L = lat_sync("assign_mult_mSAP_s_h_");
    float* a = (float*)dbuf(2, qws::P32);
    float* b = (float*)dbuf(3, qws::P32);
    L.upload_spinor(qws::P32, a, bs, 1, 0);
    L.msap_h(b, a, *nsap, *nm);
    L.download_spinor(qws::P32, b, xs, 1, 0);
  });
}

void assign_mult_wd_s_(scs_t* x, const scs_t* b) { ddd_s_(x, const_cast<scs_t*>(b)); }

void bicgstab_precdd_s_h_(scs_t* x, const scs_t* b, const double* tol, int* conviter, const int* maxiter,
                          const int* nsap, const int* nm) {
  drop_in(__func__, [&] {
    Lattice& L = lat_sync("bicgstab_precdd_s_h_");
    float* bb = (float*)dbuf(2, qws::P32);
    float* xx = (float*)dbuf(3, qws::P32);
    L.upload_spinor(qws::P32, bb, b, 1, 0);
    int it = L.bicgstab_precdd_h(xx, bb, *tol, *maxiter, *nsap, *nm);
    L.download_spinor(qws::P32, xx, x, 1, 0);
    if (conviter) *conviter = it;
  });
}

void bicgstab_dd_mix2_hf_(scd_t* x, const scd_t* b, const double* tol, int* conviter, const int* maxiter,
                          const double* tol_s, const int* maxiter_s, const int* nsap, const int* nm) {
  dd_mix_generic("bicgstab_dd_mix2_hf_", x, b, tol, conviter, maxiter, tol_s, maxiter_s, nsap, nm, qws::P16);
}

// clover inverse on the first vold elements with the clover of block `pe`
// (qws.cc:390-443; in the DD layout pe selects the x-domain)
static void clvd_generic(int pe, scd_t* out, const scd_t* in) {
  Lattice& L = lat_sync("clvd_vm_");
  const size_t db = dom_bytes_d();
  void* a = dbuf(0, qws::P64);
  void* b = dbuf(1, qws::P64);
  L.upload_spinor(qws::P64, a, embed(0, in, db, pe), 1, 1);
  L.clover(qws::P64, b, a);
  auto& full = g_qws.host_tmp[1];
  full.resize(2 * db);
  L.download_spinor(qws::P64, b, full.data(), 1, 1);
  std::memcpy(out, full.data() + pe * db, db);
}
// ---- even/odd family (deo_in_d.cc, deo_out_d.cc, deo_dag_in.cc, deo_dag_out.cc,
// deo_in_s.cc, deo_out_s.cc, qws.cc:518-605)
void deo_in_vm_(int* pe, int* po, scd_t* out, scd_t* in) {
  eo_call(__func__, qws::P64, pe, po, 1, 0, 0, 0, 0, out, in, nullptr);
}
void dee_deo_in_vm_(int* pe, int* po, scd_t* out, scd_t* in) {
  eo_call(__func__, qws::P64, pe, po, 1, 0, 1, 0, 0, out, in, nullptr);
}
void deo_dag_in_vm_(int* pe, int* po, scd_t* out, scd_t* in) {
  eo_call(__func__, qws::P64, pe, po, 1, 1, 0, 0, 0, out, in, nullptr);
}
void dee_deo_dag_in_vm_(int* pe, int* po, scd_t* out, scd_t* in) {
  eo_call(__func__, qws::P64, pe, po, 1, 1, 1, 0, 0, out, in, nullptr);
}
void one_minus_dee_deo_in_vm_(int* pe, int* po, scd_t* out, scd_t* in, scd_t* in0, double factor) {
  eo_call(__func__, qws::P64, pe, po, 1, 0, 1, 2, factor, out, in, in0);
}
void one_minus_deo_dag_in_vm_(int* pe, int* po, scd_t* out, scd_t* in, scd_t* in0, double factor) {
  eo_call(__func__, qws::P64, pe, po, 1, 1, 0, 2, factor, out, in, in0);
}
void deo_out_pre_vm_(int* pe, int* po, scd_t* out, scd_t* in) {
  (void)pe, (void)po, (void)out;
  eo_pre(__func__, qws::P64, in);
}
void deo_dag_out_pre_vm_(int* pe, int* po, scd_t* out, scd_t* in) {
  (void)pe, (void)po, (void)out;
  eo_pre(__func__, qws::P64, in);
}
void deo_out_pos_vm_(int* pe, int* po, scd_t* out, scd_t* in, double factor) {
  eo_call(__func__, qws::P64, pe, po, 2, 0, 0, 1, factor, out, in, nullptr);
}
void deo_dag_out_pos_vm_(int* pe, int* po, scd_t* out, scd_t* in, double factor) {
  eo_call(__func__, qws::P64, pe, po, 2, 1, 0, 1, factor, out, in, nullptr);
}
void dee_deo_out_pos_vm_(int* pe, int* po, scd_t* out, scd_t* in, double factor) {
  eo_call(__func__, qws::P64, pe, po, 2, 0, 1, 1, factor, out, in, nullptr);
}
void dee_deo_dag_out_pos_vm_(int* pe, int* po, scd_t* out, scd_t* in, double factor) {
  eo_call(__func__, qws::P64, pe, po, 2, 1, 1, 1, factor, out, in, nullptr);
}
void deo_vm_(int* pe, int* po, scd_t* out, scd_t* in) {
  eo_call(__func__, qws::P64, pe, po, 0, 0, 0, 0, 0, out, in, nullptr);
}
void dee_deo_vm_(int* pe, int* po, scd_t* out, scd_t* in) {
  eo_call(__func__, qws::P64, pe, po, 0, 0, 1, 0, 0, out, in, nullptr);
}
void one_minus_dee_deo_vm_(int* pe, int* po, scd_t* out, scd_t* in, scd_t* in0) {
  eo_call(__func__, qws::P64, pe, po, 0, 0, 1, 2, -kappa2, out, in, in0);
}
void deo_dag_vm_(int* pe, int* po, scd_t* out, scd_t* in) {
  eo_call(__func__, qws::P64, pe, po, 0, 1, 0, 0, 0, out, in, nullptr);
}
void dee_deo_dag_vm_(int* pe, int* po, scd_t* out, scd_t* in) {
  eo_call(__func__, qws::P64, pe, po, 0, 1, 1, 0, 0, out, in, nullptr);
}
void one_minus_deo_dag_vm_(int* pe, int* po, scd_t* out, scd_t* in, scd_t* in0) {
  eo_call(__func__, qws::P64, pe, po, 0, 1, 0, 2, -kappa2, out, in, in0);
}
void mtilde_vm_(scd_t* out, scd_t* in) { eo_mtilde(__func__, qws::P64, 0, -kappa2, out, in); }
// bicgstab.cc:65 -- BiCGStab on mtilde_vm_ from the initial guess in x, stopping at
// sqrt(|r|^2/|b|^2) < 1e-14
void bicgstab_vm_(scd_t* x, scd_t* b, int* conviter, int* maxiter) {
  if (!g_qws.pce || (*g_qws.pce != 0 && *g_qws.pce != 1)) fatal(__func__, "pce must point at 0 or 1");
  drop_in(__func__, [&] {
    Lattice& L = lat_sync_eo("bicgstab_vm_");
    double* xx = (double*)eobuf(qws::P64, 0);
    double* bb = (double*)eobuf(qws::P64, 2);
    L.eo_upload_spinor(qws::P64, xx, x, 1, QWS_VLEND, 1);
    L.eo_upload_spinor(qws::P64, bb, b, 1, QWS_VLEND, 1);
    const int it = L.bicgstab_eo(xx, bb, *maxiter, 1e-14, *g_qws.pce, -kappa2);
    L.eo_download_spinor(qws::P64, xx, x, 1, QWS_VLEND, 1);
    if (conviter) *conviter = it;
  });
}
// cg.cc:65 -- CG on mtdagmt_vm_ from the initial guess in x, stopping at
// sqrt(|r|^2/|b|^2) < 1e-14 or after maxiter-1 iterations
void cg_vm_(scd_t* x, scd_t* b, int* conviter, int* maxiter) {
  if (!g_qws.pce || (*g_qws.pce != 0 && *g_qws.pce != 1)) fatal(__func__, "pce must point at 0 or 1");
  drop_in(__func__, [&] {
    Lattice& L = lat_sync_eo("cg_vm_");
    double* xx = (double*)eobuf(qws::P64, 0);
    double* bb = (double*)eobuf(qws::P64, 2);
    L.eo_upload_spinor(qws::P64, xx, x, 1, QWS_VLEND, 1);
    L.eo_upload_spinor(qws::P64, bb, b, 1, QWS_VLEND, 1);
    const int it = L.cg_eo(xx, bb, *maxiter, 1e-14, *g_qws.pce, -kappa2);
    L.eo_download_spinor(qws::P64, xx, x, 1, QWS_VLEND, 1);
    if (conviter) *conviter = it;
  });
}
// mcg.cc:65 -- multi-shift CG: xx[k*vold .. (k+1)*vold) solves (mtdagmt + sigma[k]) x = b
void mcg_vm_(scd_t* xx, scd_t* b, int* n, double* sigma, int* maxiter) {
  if (!g_qws.pce || (*g_qws.pce != 0 && *g_qws.pce != 1)) fatal(__func__, "pce must point at 0 or 1");
  if (*n < 1 || *n > 64) fatal(__func__, "the number of shifts must be in 1..64");
  drop_in(__func__, [&] {
    Lattice& L = lat_sync_eo("mcg_vm_");
    const size_t eb = L.eo_spinor_bytes(qws::P64);
    double* bb = (double*)eobuf(qws::P64, 2);
    char* base = (char*)eobuf_n(qws::P64, *n);
    std::vector<double*> xs(*n);
    for (int k = 0; k < *n; ++k) xs[k] = (double*)(base + k * eb);
    L.eo_upload_spinor(qws::P64, bb, b, 1, QWS_VLEND, 1);
    L.mcg_eo(xs.data(), bb, *n, sigma, *maxiter, 1e-14, *g_qws.pce, -kappa2);
    for (int k = 0; k < *n; ++k) L.eo_download_spinor(qws::P64, xs[k], xx + (size_t)k * vold, 1, QWS_VLEND, 1);
  });
}
void mtilde_dag_vm_(scd_t* out, scd_t* in) { eo_mtilde(__func__, qws::P64, 1, -kappa2, out, in); }
void mtdagmt_vm_(scd_t* out, scd_t* in) { eo_mtilde(__func__, qws::P64, 2, -kappa2, out, in); }

void deo_in_s_(int* pe, int* po, scs_t* out, scs_t* in) {
  eo_call(__func__, qws::P32, pe, po, 1, 0, 0, 0, 0, out, in, nullptr);
}
void dee_deo_in_s_(int* pe, int* po, scs_t* out, scs_t* in) {
  eo_call(__func__, qws::P32, pe, po, 1, 0, 1, 0, 0, out, in, nullptr);
}
void one_minus_dee_deo_in_s_(int* pe, int* po, scs_t* out, scs_t* in, scs_t* in0, float factor) {
  eo_call(__func__, qws::P32, pe, po, 1, 0, 1, 2, factor, out, in, in0);
}
void deo_out_pre_s_(const int* pe, const int* po, const scs_t* out, const scs_t* in) {
  (void)pe, (void)po, (void)out;
  eo_pre(__func__, qws::P32, in);
}
void deo_out_pos_s_(const int* pe, const int* po, const scs_t* out, const scs_t* in, float factor) {
  eo_call(__func__, qws::P32, pe, po, 2, 0, 0, 1, factor, const_cast<scs_t*>(out), in, nullptr);
}
void dee_deo_out_pos_s_(int* pe, int* po, scs_t* out, scs_t* in, float factor) {
  eo_call(__func__, qws::P32, pe, po, 2, 0, 1, 1, factor, out, in, nullptr);
}
void clv_s_(int* pe, scs_t* inout) {
  const int po = 1 - *pe;
  eo_call(__func__, qws::P32, pe, &po, 3, 0, 1, 0, 0, inout, inout, nullptr);
}
void deo_s_(int* pe, int* po, scs_t* out, scs_t* in) {
  eo_call(__func__, qws::P32, pe, po, 0, 0, 0, 0, 0, out, in, nullptr);
}
void dee_deo_s_(int* pe, int* po, scs_t* out, scs_t* in) {
  eo_call(__func__, qws::P32, pe, po, 0, 0, 1, 0, 0, out, in, nullptr);
}
void one_minus_dee_deo_s_(int* pe, int* po, scs_t* out, scs_t* in, scs_t* in0) {
  eo_call(__func__, qws::P32, pe, po, 0, 0, 1, 2, -(float)kappa2, out, in, in0);
}
void mtilde_s_(scs_t* out, scs_t* in) { eo_mtilde(__func__, qws::P32, 0, -(float)kappa2, out, in); }

void clvd_vm_(int* pe, scd_t* inout) {
  drop_in(__func__, [&] { clvd_generic(*pe, inout, inout); });
}
void clvd_vm2_(int* pe, scd_t* out, scd_t* in) {
  drop_in(__func__, [&] { clvd_generic(*pe, out, in); });
}

// ---- FP64 <-> FP32 with normalisation (reordering.cc:108-204), host side
void fermi_reorder_d2s_dd_renorm_(scs_t* out, const scd_t* in, double* dnorm) {
  const size_t n = (size_t)vold * 2;
  double acc = 0;
  for (size_t i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c)
      for (int s = 0; s < 4; ++s)
        for (int ri = 0; ri < 2; ++ri)
          for (int j = 0; j < QWS_VLEND; ++j) acc += in[i].c[c][s][ri][j] * in[i].c[c][s][ri][j];
  acc = std::sqrt(acc);
  *dnorm = acc;
  const double inv = 1.0 / acc;
  for (int x = 0; x < nx; ++x)
    for (int t = 0; t < nt; ++t)
      for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y) {
          const int id = (x % nxh) / QWS_VLEND + nxd * (y + ny * (z + nz * t)) + vold * (x / nxh);
          const int jd = (x % nxh) % QWS_VLEND;
          const int is = (x % nxh) / QWS_VLENS + nxs * (y + ny * (z + nz * t)) + vols * (x / nxh);
          const int js = (x % nxh) % QWS_VLENS;
          for (int c = 0; c < 3; ++c)
            for (int s = 0; s < 4; ++s)
              for (int ri = 0; ri < 2; ++ri) out[is].c[c][s][ri][js] = (float)(in[id].c[c][s][ri][jd] * inv);
        }
}

void fermi_reorder_s2d_dd_renorm_(scd_t* out, const scs_t* in, const double* dnorm) {
  for (int x = 0; x < nx; ++x)
    for (int t = 0; t < nt; ++t)
      for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y) {
          const int id = (x % nxh) / QWS_VLEND + nxd * (y + ny * (z + nz * t)) + vold * (x / nxh);
          const int jd = (x % nxh) % QWS_VLEND;
          const int is = (x % nxh) / QWS_VLENS + nxs * (y + ny * (z + nz * t)) + vols * (x / nxh);
          const int js = (x % nxh) % QWS_VLENS;
          for (int c = 0; c < 3; ++c)
            for (int s = 0; s < 4; ++s)
              for (int ri = 0; ri < 2; ++ri)
                out[id].c[c][s][ri][jd] = (double)in[is].c[c][s][ri][js] * (*dnorm);
        }
}

// reordering.cc:59-105 -- the same site maps without the normalisation
void fermi_reorder_d2s_dd_(scs_t* out, const scd_t* in) {
  for (int x = 0; x < nx; ++x)
    for (int t = 0; t < nt; ++t)
      for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y) {
          const int id = (x % nxh) / QWS_VLEND + nxd * (y + ny * (z + nz * t)) + vold * (x / nxh);
          const int jd = (x % nxh) % QWS_VLEND;
          const int is = (x % nxh) / QWS_VLENS + nxs * (y + ny * (z + nz * t)) + vols * (x / nxh);
          const int js = (x % nxh) % QWS_VLENS;
          for (int c = 0; c < 3; ++c)
            for (int s = 0; s < 4; ++s)
              for (int ri = 0; ri < 2; ++ri) out[is].c[c][s][ri][js] = (float)in[id].c[c][s][ri][jd];
        }
}
void fermi_reorder_s2d_dd_(scd_t* out, const scs_t* in) {
  const double one = 1.0;
  fermi_reorder_s2d_dd_renorm_(out, in, &one);
}

// ddd_out_s.cc:96,100 -- the face terms without QWS's per-call timers (this library
// keeps none): the same functions
void ddd_out_pre_s_no_timer_(scs_t* in, int* domain) { ddd_out_pre_s_(in, domain); }
void ddd_out_pos_s_no_timer_(scs_t* out, scs_t* in, int* domain, float factor) {
  ddd_out_pos_s_(out, in, domain, factor);
}

// qws.cc:445-500 -- host helpers of QWS's drivers: clear a field, print |field|^2
// ("%30s : %24.14e"; the FP32 sum is formed in FP64 and rounded to FP32)
void zero_scd_field(scd_t* inout, int size) { std::memset(inout, 0, sizeof(scd_t) * (size_t)size); }
void zero_scs_field(scs_t* inout, int size) { std::memset(inout, 0, sizeof(scs_t) * (size_t)size); }
void print_scdnorm2(const char* a, scd_t* in, int size) {
  const double* v = reinterpret_cast<const double*>(in);
  double acc = 0;
  for (size_t i = 0; i < (size_t)size * 24 * QWS_VLEND; ++i) acc += v[i] * v[i];
  if (rank == 0) {
    std::printf("%30s : %24.14e\n", a, acc);
    std::fflush(stdout);
  }
}
void print_scsnorm2(const char* a, scs_t* in, int size) {
  const float* v = reinterpret_cast<const float*>(in);
  double acc = 0;
  for (size_t i = 0; i < (size_t)size * 24 * QWS_VLENS; ++i) acc += (double)v[i] * v[i];
  if (rank == 0) {
    std::printf("%30s : %24.14e\n", a, (double)(float)acc);
    std::fflush(stdout);
  }
}

// timing.c:171,185 -- QWS's driver clock and its per-function timing report; this
// library keeps no per-function timers, so the report says so
double get_clock(void) {
  timespec tp;
  clock_gettime(CLOCK_REALTIME, &tp);
  return (double)tp.tv_sec + (double)tp.tv_nsec * 1e-9;
}
void check_timing_(const char* name) {
  if (name == nullptr && rank == 0) std::printf("(libqws_mi355x keeps no per-function timers)\n");
}
void print_timing_(void) { check_timing_(nullptr); }

// qws.cc:634-639 -- D on one domain from the domain-local and face parts
void ddd_eo_s_(scs_t* out, scs_t* in, int* domain) {
  ddd_out_pre_s_(in, domain);
  ddd_in_s_(&out[(size_t)vols * *domain], &in[(size_t)vols * *domain], domain);
  ddd_out_pos_s_(&out[(size_t)vols * *domain], in, domain, (float)mkappa);
}

// bicgstab_dd_s.cc:67, bicgstab_dd_d.cc:64 -- plain BiCGStab on D
void bicgstab_dd_s_(scs_t* x, scs_t* b, int* conviter, int* maxiter) {
  drop_in(__func__, [&] {
    Lattice& L = lat_sync("bicgstab_dd_s_");
    void* xx = dbuf(2, qws::P32);
    void* bb = dbuf(3, qws::P32);
    L.upload_spinor(qws::P32, xx, x, 1, 0);
    L.upload_spinor(qws::P32, bb, b, 1, 0);
    const int it = L.bicgstab_dd(qws::P32, xx, bb, *maxiter, 1e-6);
    L.download_spinor(qws::P32, xx, x, 1, 0);
    if (conviter) *conviter = it;
  });
}
void bicgstab_dd_d_(scd_t* x, scd_t* b, int* conviter, int* maxiter) {
  drop_in(__func__, [&] {
    Lattice& L = lat_sync("bicgstab_dd_d_");
    void* xx = dbuf(0, qws::P64);
    void* bb = dbuf(1, qws::P64);
    L.upload_spinor(qws::P64, xx, x, 1, 1);
    L.upload_spinor(qws::P64, bb, b, 1, 1);
    const int it = L.bicgstab_dd(qws::P64, xx, bb, *maxiter, 1e-14);
    L.download_spinor(qws::P64, xx, x, 1, 1);
    if (conviter) *conviter = it;
  });
}

// bicgstab_precdd_s.cc:561-597 -- prints "@C iter |b - D M_SAP x| / sqrt(bnorm)"
void check_residual_s_(const scs_t* x, const scs_t* b, const int* iter, const float* bnorm, const int* nsap,
                       const int* nm) {
  const size_t n = (size_t)vols * 2;
  std::vector<scs_t> w(n);
  prec_ddd_s_(w.data(), const_cast<scs_t*>(x), const_cast<int*>(nsap), const_cast<int*>(nm));
  float rnorm = 0.0f;
  for (size_t i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c)
      for (int s = 0; s < 4; ++s)
        for (int ri = 0; ri < 2; ++ri)
          for (int v = 0; v < QWS_VLENS; ++v) {
            const float d = b[i].c[c][s][ri][v] - w[i].c[c][s][ri][v];
            rnorm += d * d;
          }
  const float err = (float)std::sqrt((double)(rnorm / *bnorm));
  if (rank == 0) std::printf("@C %5d %12.7e\n", *iter, err);
}

}  // extern "C"

// ---- BQCD loaders (qws_bqcd.cc:66-647).  BQCD numbers the sites of each parity
// itself: std_xyzt2i_ (1-based index within the parity) and e_o_ (the parity) come
// from the caller.  They are found as weak references at load time, or else in the
// process's global symbol scope at call time (a caller that loads them later), and
// tabulated once per geometry.  The permutations run on the GPU (launch_bqcd): a
// BQCD array goes up once, the kernel writes QWS's host AoSoA images in device
// memory, and these come back into the caller's arrays or QWS's field arrays.  The
// device fields are then built from the images already on the device.
namespace {

using bqcd_site_fn = int (*)(int*);

const int* bqcd_map(const char* where) {
  if (!g_qws.lat) fatal(where, "qws_init_ has not been called");
  if (g_qws.bqcd_map) return g_qws.bqcd_map;
  bqcd_site_fn idx = std_xyzt2i_ ? std_xyzt2i_ : (bqcd_site_fn)dlsym(RTLD_DEFAULT, "std_xyzt2i_");
  bqcd_site_fn par = e_o_ ? e_o_ : (bqcd_site_fn)dlsym(RTLD_DEFAULT, "e_o_");
  if (!idx || !par) fatal(where, "BQCD's std_xyzt2i_/e_o_ are not linked");
  const long V = (long)nx * ny * nz * nt;
  std::vector<int> map(V);
  for (int t = 0; t < nt; ++t)
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
          int j[4] = {x, y, z, t};
          const int i = idx(j), eo = par(j);
          if (i < 1 || i > V / 2 || (eo != 0 && eo != 1)) fatal(where, "std_xyzt2i_/e_o_ out of range");
          map[x + (long)nx * (y + (long)ny * (z + (long)nz * t))] = 2 * (i - 1) + eo;
        }
  drop_in(where, [&] {
    int* d = nullptr;
    QWS_HIP(hipMalloc((void**)&d, sizeof(int) * V));
    QWS_HIP(hipMemcpy(d, map.data(), sizeof(int) * V, hipMemcpyHostToDevice));
    g_qws.bqcd_map = d;
  });
  return g_qws.bqcd_map;
}

void set_kappa(double k) {
  kappa = k;
  mkappa = -k;
  kappa2 = k * k;
}

qws::BqcdGeo bqcd_geo(long volh, bool eo_layout) {
  qws::BqcdGeo b;
  b.L[0] = nx, b.L[1] = ny, b.L[2] = nz, b.L[3] = nt;
  b.volh = volh;
  b.eo_layout = eo_layout ? 1 : 0;
  return b;
}

// Links (kind 1) or clover (kind 2) from a BQCD array into QWS's FP64 and FP32 field
// arrays (glud/glus, clvd/clvs) and, from the same images, into the device fields of
// the reading the layout is for (DD or even/odd).  The other reading is rebuilt from
// the host arrays when it is next used.
void load_bqcd_field(const char* where, int kind, const double* src, long volh, bool eo_layout) {
  const int* map = bqcd_map(where);
  Lattice& L = *g_qws.lat;
  const long V = (long)nx * ny * nz * nt;
  if (2 * volh != V) fatal(where, "the BQCD array's parity volume does not match qws_init_'s lattice");
  const size_t n = (size_t)V * 72;  // reals: 4 x 18 links or 72 clover per site
  void* hd = kind == 1 ? (void*)glud : (void*)clvd;
  void* hs = kind == 1 ? (void*)glus : (void*)clvs;
  drop_in(where, [&] {
    double* bq = (double*)L.work("bqcd_in", n * sizeof(double));
    double* imd = (double*)L.work("bqcd_img_d", n * sizeof(double));
    float* ims = (float*)L.work("bqcd_img_s", n * sizeof(float));
    L.upload_bytes(bq, src, n * sizeof(double));
    const qws::BqcdGeo b = bqcd_geo(volh, eo_layout);
    qws::launch_bqcd<double>(b, kind, 1, map, bq, imd, QWS_VLEND, L.stream());
    qws::launch_bqcd<float>(b, kind, 1, map, bq, ims, QWS_VLENS, L.stream());
    L.download_bytes(hd, imd, n * sizeof(double));
    L.download_bytes(hs, ims, n * sizeof(float));
    bool& dd = kind == 1 ? g_qws.dirty_g : g_qws.dirty_c;
    bool& eo = kind == 1 ? g_qws.eo_dirty_g : g_qws.eo_dirty_c;
    dd = eo = true;
    if (!eo_layout) {
      if (kind == 1) L.load_gauge(imd, 1, 1, 1, true), L.load_gauge(ims, 1, 0, 2, true);
      else L.load_clover(imd, 1, 1, 1, true), L.load_clover(ims, 1, 0, 2, true);
      dd = false;
    } else {
      if (kind == 1) L.eo_load_gauge(imd, 1, QWS_VLEND, 1, 1, true), L.eo_load_gauge(ims, 1, QWS_VLENS, 0, 2, true);
      else L.eo_load_clover(imd, 1, QWS_VLEND, 1, 1, true), L.eo_load_clover(ims, 1, QWS_VLENS, 0, 2, true);
      eo = false;
    }
  });
}

// Spinors: BQCD's per-parity arrays he/ho ([i][24]) <-> a QWS host spinor array (scd_t
// or scs_t, lane width VL).  The DD form moves both parities into the full field
// (qws_bqcd.cc:500-642); the even/odd form moves the even sites only, into one parity
// field of x-compressed sites (qws_bqcd.cc:384-498; ho unused).
template <typename T, int VL>
void bqcd_spinor(const char* where, T* qws_arr, double* he, double* ho, bool load, bool eo_layout) {
  // (scd_t and scs_t have the same size: 8 FP64 lanes or 16 FP32 lanes)
  using E = typename std::conditional<std::is_same<T, scd_t>::value, double, float>::type;
  static_assert(std::is_same<T, scd_t>::value || std::is_same<T, scs_t>::value, "QWS spinor types only");
  const int* map = bqcd_map(where);
  Lattice& L = *g_qws.lat;
  const long V = (long)nx * ny * nz * nt, volh = V / 2;
  const size_t nbq = (size_t)(eo_layout ? volh : V) * 24;  // BQCD reals moved
  const size_t nimg = (size_t)(eo_layout ? volh : V) * 24;  // QWS image reals (one or both parities)
  drop_in(where, [&] {
    double* bq = (double*)L.work("bqcd_in", nbq * sizeof(double));
    E* img = (E*)L.work(sizeof(E) == 8 ? "bqcd_img_d" : "bqcd_img_s", nimg * sizeof(E));
    const qws::BqcdGeo b = bqcd_geo(volh, eo_layout);
    const size_t half = (size_t)volh * 24 * sizeof(double);
    if (load) {
      L.upload_bytes(bq, he, half);
      if (!eo_layout) L.upload_bytes(bq + (size_t)volh * 24, ho, half);
      qws::launch_bqcd<E>(b, 0, 1, map, bq, img, VL, L.stream());
      L.download_bytes(qws_arr, img, nimg * sizeof(E));
    } else {
      L.upload_bytes(img, qws_arr, nimg * sizeof(E));
      qws::launch_bqcd<E>(b, 0, 0, map, bq, img, VL, L.stream());
      L.download_bytes(he, bq, half);
      if (!eo_layout) L.download_bytes(ho, bq + (size_t)volh * 24, half);
    }
  });
}

}  // namespace

extern "C" {

void qws_loadg_dd_(double* u, int* volh_tot, double* in_kappa) {
  set_kappa(*in_kappa);
  load_bqcd_field(__func__, 1, u, *volh_tot, false);
}
void qws_loadc_dd_(double* c, int* volh) { load_bqcd_field(__func__, 2, c, *volh, false); }
void qws_loadfd_dd_(scd_t* out, double* ine, double* ino) {
  bqcd_spinor<scd_t, QWS_VLEND>(__func__, out, ine, ino, true, false);
}
void qws_storefd_dd_(double* oute, double* outo, scd_t* in) {
  bqcd_spinor<scd_t, QWS_VLEND>(__func__, in, oute, outo, false, false);
}
void qws_loadfs_dd_(scs_t* out, double* ine, double* ino) {
  bqcd_spinor<scs_t, QWS_VLENS>(__func__, out, ine, ino, true, false);
}
void qws_storefs_dd_(double* oute, double* outo, scs_t* in) {
  bqcd_spinor<scs_t, QWS_VLENS>(__func__, in, oute, outo, false, false);
}

void qws_loadg_bqcd_(double* u, int* volh_tot, double* in_kappa) {
  set_kappa(*in_kappa);
  load_bqcd_field(__func__, 1, u, *volh_tot, true);
}
void qws_loadc_bqcd_(double* c, int* volh) { load_bqcd_field(__func__, 2, c, *volh, true); }
void qws_loadfd_bqcd_(scd_t* out, double* in) {
  bqcd_spinor<scd_t, QWS_VLEND>(__func__, out, in, nullptr, true, true);
}
void qws_loadfs_bqcd_(scs_t* out, double* in) {
  bqcd_spinor<scs_t, QWS_VLENS>(__func__, out, in, nullptr, true, true);
}
void qws_storefd_bqcd_(double* out, scd_t* in) {
  bqcd_spinor<scd_t, QWS_VLEND>(__func__, in, out, nullptr, false, true);
}
void qws_storefs_bqcd_(double* out, scs_t* in) {
  bqcd_spinor<scs_t, QWS_VLENS>(__func__, in, out, nullptr, false, true);
}
void bqcd_qws_h_init_(void) { qws_h_init_(glus, clvs); }

}  // extern "C"
