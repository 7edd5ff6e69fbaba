// multi.cpp -- solves over several lattices at once, advanced in lockstep: every
// lattice runs its own FP64 outer BiCGStab (bicgstab_dd_mix.cc:105-388), FP32 inner
// BiCGStab (bicgstab_precdd_s.cc:85-559) and SAP preconditioner (qws.cc:776-975) with
// its own volume, fields, kappa and right-hand sides, but each SAP half-sweep, and the
// final face term of prec_ddd_s_, is one launch over the blocks of all lattices
// (launch_sap_multi / launch_sap_face_multi: a table maps each workgroup to a lattice
// and a block).  A small lattice thus shares the GPU with the others instead of
// launching a few workgroups of its own.  The vector updates and reductions stay
// per lattice, in the same kernels as a single solve, so every right-hand side gets
// the result it gets alone, bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>

#include "lattice.h"
#include "multi.h"
#include "sap.h"

namespace qws {

bool trace_on();

// ------------------------------------------------------------------ SapFusion
SapFusion::SapFusion(hipStream_t st, size_t bytes) : st_(st), cap_(bytes) {
  QWS_HIP(hipMalloc((void**)&dev_, cap_));
  QWS_HIP(hipHostMalloc((void**)&host_, cap_, hipHostMallocDefault));
  // the fused SAP sequences replay as HIP graphs unless QWS_MI355X_SAP_GRAPH=0
  const char* g = std::getenv("QWS_MI355X_SAP_GRAPH");
  use_graphs_ = !(g && g[0] == '0');
}

SapFusion::~SapFusion() {
  (void)hipStreamSynchronize(st_);
  drop_graphs();
  if (dev_) (void)hipFree(dev_);
  if (host_) (void)hipHostFree(host_);
}

void SapFusion::drop_graphs() {
  for (Graph& g : graphs_)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  graphs_.clear();
}

// The device copy of `bytes` of descriptor data: an earlier copy with the same
// contents, or a new one in the arena (uploaded on the stream, so it is in place
// before any later launch on it reads it).  A full arena starts over after the
// stream has drained (no launch still reads the old copies).
const void* SapFusion::upload(const void* data, size_t bytes) {
  const std::string key((const char*)data, bytes);
  for (const auto& e : entries_)
    if (e.first == key) return dev_ + e.second;
  const size_t need = (bytes + 255) & ~(size_t)255;
  if (need > cap_) throw std::invalid_argument("fused launch: descriptor table larger than its arena");
  if (off_ + need > cap_) {
    QWS_HIP(hipStreamSynchronize(st_));
    entries_.clear();
    drop_graphs();  // they read descriptor tables that are about to be overwritten
    off_ = 0;
  }
  std::memcpy(host_ + off_, data, bytes);
  QWS_HIP(hipMemcpyAsync(dev_ + off_, host_ + off_, bytes, hipMemcpyHostToDevice, st_));
  entries_.push_back({key, off_});
  const size_t at = off_;
  off_ += need;
  return dev_ + at;
}

// ------------------------------------------------------------------ helpers
namespace {

// the host copies of the scalars of every lattice in `lats` (one synchronisation per
// distinct stream)
void read_all(const std::vector<std::pair<Lattice*, std::pair<const Scalars*, int>>>& v) {
  std::vector<hipStream_t> streams;
  for (const auto& e : v) {
    Lattice& L = *e.first;
    QWS_HIP(hipMemcpyAsync(L.host_scalars(), e.second.first, sizeof(Scalars) * e.second.second,
                           hipMemcpyDeviceToHost, L.stream()));
    if (std::find(streams.begin(), streams.end(), L.stream()) == streams.end()) streams.push_back(L.stream());
  }
  for (hipStream_t s : streams) QWS_HIP(hipStreamSynchronize(s));
}

// the FP32-SAP settings that replace a failed FP16-stored SAP: the same sweeps
SolveParams escalated(const SolveParams& p) {
  SolveParams q = p;
  q.sap_prec = P32;
  q.nsap = p.nsap > 1 ? p.nsap - 1 : 1;
  return q;
}

bool any_of(const std::vector<int>& v) {
  for (int x : v)
    if (x) return true;
  return false;
}

}  // namespace

// ------------------------------------------------------------------ Fused::prec
// out_i = D M_SAP in_i (ddd) or M_SAP in_i for every job; fused over the jobs when
// they allow it (FP32 fused SAP with a face buffer, one block volume, one link storage,
// one stream, the same direction-update choice), else one lattice after another.
void Fused::prec(std::vector<PrecJob>& jobs, bool ddd, int nsap, int nm, SapFusion* fz) {
  if (jobs.empty()) return;
  const Lattice& L0 = *jobs[0].L;
  bool fuse = fz != nullptr && jobs.size() > 1;
  for (const PrecJob& j : jobs) {
    const Lattice& L = *j.L;
    fuse = fuse && L.fused_ok<float>() && L.use_fb_ && sap_face_buffer_floats(L.g_) > 0 && L.g_.Vb == L0.g_.Vb &&
           L.nl32_ == L0.nl32_ && L.st_ == L0.st_ && L.st_ == fz->stream() && (j.upd != nullptr) == (jobs[0].upd != nullptr);
  }
  if (!fuse) {
    for (PrecJob& j : jobs) {
      if (ddd) j.L->prec_ddd_t<float>(j.B, j.out, j.in, nsap, nm, j.upd);
      else j.L->prec_s_t<float>(j.B, j.out, j.in, nsap, nm);
    }
    return;
  }
  const size_t np = jobs.size();
  const bool fold = jobs[0].upd != nullptr && nsap > 0;
  std::vector<std::vector<SapStep>> plans(np);
  std::vector<float*> fb(np);
  std::vector<long long> fbs(np);
  for (size_t i = 0; i < np; ++i) {
    PrecJob& j = jobs[i];
    Lattice& L = *j.L;
    if (j.upd && !fold) launch_update_p<float>(j.B, L.g_.V * 12, j.upd->p, j.upd->r, j.upd->q, j.upd->sc, L.st_);
    plans[i] = L.sap_plan(j.B, ddd, j.out, j.in, nsap, fold, fb[i], fbs[i]);
    if (plans[i].size() != plans[0].size()) throw std::logic_error("fused SAP: plans of different length");
  }
  // workgroup -> (lattice, block) for the blocks of each colour, lattice after lattice
  const int2* table[2];
  int ntab[2];
  for (int c = 0; c < 2; ++c) {
    std::vector<int2> t;
    for (size_t i = 0; i < np; ++i)
      for (int b : jobs[i].L->h_blocks_[c]) t.push_back(make_int2((int)i, b));
    ntab[c] = (int)t.size();
    table[c] = ntab[c] ? (const int2*)fz->upload(t.data(), sizeof(int2) * t.size()) : nullptr;
  }
  const int vb = L0.g_.Vb, nl = L0.nl32_;
  std::vector<SapPart<float>> parts(np);
  struct Rec {
    SapMulti<float> m;
    int flags, face;
  };
  std::vector<Rec> recs;
  for (size_t k = 0; k < plans[0].size(); ++k) {
    const SapStep& s0 = plans[0][k];
    for (size_t i = 0; i < np; ++i) {
      const SapStep& st = plans[i][k];
      if (st.face != s0.face || st.colour != s0.colour || st.flags != s0.flags || st.pupd != s0.pupd)
        throw std::logic_error("fused SAP: launches differ between lattices");
      const Lattice& L = *jobs[i].L;
      const Lattice::FieldSet<float> f = L.fields<float>();
      SapPart<float>& P = parts[i];
      P = SapPart<float>();
      P.g = L.g_;
      P.active = jobs[i].B.active;
      P.stride = jobs[i].B.stride;
      P.nrhs = jobs[i].B.n;
      P.kappa = st.face ? -(float)L.kappa_ : (float)L.kappa_;
      P.s = st.s, P.in = st.in, P.qface = st.qface, P.qout = st.qout, P.out = st.out;
      P.U = f.U, P.C = f.C, P.fb = fb[i], P.fbs = fbs[i];
      if (st.pupd) P.upd = *jobs[i].upd;
    }
    SapMulti<float> m;
    m.parts = (const SapPart<float>*)fz->upload(parts.data(), sizeof(SapPart<float>) * np);
    m.table = table[s0.colour];
    m.n = ntab[s0.colour];
    m.xcd = L0.g_.xcd_remap;
    Rec r;
    std::memset(&r, 0, sizeof(r));
    r.m = m, r.flags = s0.flags | (s0.pupd ? SAP_SRC_PUPD : 0), r.face = s0.face ? 1 : 0;
    recs.push_back(r);
  }
  // the launches (with their descriptor tables already uploaded) as one HIP graph
  // from their second occurrence on
  std::string key((const char*)recs.data(), sizeof(Rec) * recs.size());
  const int shape[3] = {vb, nl, nm};
  key.append((const char*)shape, sizeof(shape));
  fz->run(key, [&] {
    for (const Rec& r : recs) {
      if (r.face) launch_sap_face_multi(r.m, vb, nl, fz->stream());
      else launch_sap_multi(r.m, vb, nl, nm, r.flags, fz->stream());
    }
  });
}

// ------------------------------------------------------------------ Fused::bicg
// The FP32 inner BiCGStab of Lattice::bicg_t (without hp) for several lattices in
// lockstep: per iteration one fused preconditioner application per half-step, the
// vector updates and reductions per lattice, one host synchronisation per half-step
// for the convergence test of all of them.
void Fused::bicg(std::vector<InnerJob>& J, double tol, int maxiter, int nsap, int nm, SapFusion* fz) {
  const size_t np = J.size();
  struct St {
    float *q, *r, *p, *t, *r0;
    Scalars* sc;
    double* part;
    std::vector<int> live;
    std::vector<double> bnorm;
    std::vector<int> fix;  // RHS leaving after the first half-step (host copy of the mask)
    long long n;
    Batch B;
    SapSrcUpdate pupd;
  };
  std::vector<St> S(np);
  auto upload = [&](size_t i) {
    Lattice& L = *J[i].L;
    QWS_HIP(hipMemcpyAsync(L.d_inner_mask_, S[i].live.data(), sizeof(int) * S[i].live.size(), hipMemcpyHostToDevice,
                           L.st_));
  };
  for (size_t i = 0; i < np; ++i) {
    Lattice& L = *J[i].L;
    St& s = S[i];
    const int nb = J[i].Bin.n;
    s.n = L.g_.V * 12;
    const size_t bytes = L.spinor_bytes(P32) * nb;
    s.q = (float*)L.work("bicg_q", bytes);
    s.r = (float*)L.work("bicg_r", bytes);
    s.p = (float*)L.work("bicg_p", bytes);
    s.t = (float*)L.work("bicg_t", bytes);
    s.r0 = (float*)L.work("bicg_r0", bytes);
    if (nb > L.sc_in_n_) {
      if (L.sc_in_) QWS_HIP(hipFree(L.sc_in_));
      QWS_HIP(hipMalloc((void**)&L.sc_in_, sizeof(Scalars) * nb));
      L.sc_in_n_ = nb;
    }
    s.sc = L.sc_in_;
    s.part = L.reduction_partials(nb);
    std::vector<int>& iters = *J[i].iters;
    s.live.assign(nb, 0);
    for (int k = 0; k < nb; ++k) s.live[k] = iters[k] >= 0 ? 1 : 0;  // iters[k] < 0: RHS outside the batch
    for (int k = 0; k < nb; ++k) iters[k] = 0;
    upload(i);
    s.B = Batch{nb, L.g_.V * 24, L.d_inner_mask_};
    launch_bicg_init_inner<float>(s.B, s.n, J[i].x, s.r, s.p, s.r0, J[i].b, s.part, L.st_);
    launch_finalize(s.B, s.part, 1, s.sc, FIN_INIT_INNER, L.st_);
    s.bnorm.assign(nb, 1.0);
    s.pupd.p = s.p, s.pupd.r = s.r, s.pupd.q = s.q, s.pupd.sc = s.sc;
  }
  auto read = [&](const std::vector<size_t>& which) {
    std::vector<std::pair<Lattice*, std::pair<const Scalars*, int>>> v;
    for (size_t i : which) v.push_back({J[i].L, {S[i].sc, J[i].Bin.n}});
    read_all(v);
  };
  auto live_parts = [&] {
    std::vector<size_t> v;
    for (size_t i = 0; i < np; ++i)
      if (any_of(S[i].live)) v.push_back(i);
    return v;
  };
  if (tol >= 0) {
    std::vector<size_t> all;
    for (size_t i = 0; i < np; ++i) all.push_back(i);
    read(all);
    for (size_t i = 0; i < np; ++i)
      for (int k = 0; k < J[i].Bin.n; ++k) S[i].bnorm[k] = J[i].L->host_scalars()[k].bnorm;
  }
  // drop converged RHS (and those whose residual is not finite); true if any remain
  auto converge = [&](const std::vector<size_t>& act) {
    if (tol < 0) return true;
    read(act);
    bool any = false;
    for (size_t i : act) {
      const Scalars* h = J[i].L->host_scalars();
      bool changed = false;
      for (int k = 0; k < J[i].Bin.n; ++k) {
        if (!S[i].live[k]) continue;
        if (std::sqrt(h[k].rnorm / S[i].bnorm[k]) < tol || !std::isfinite(h[k].rnorm)) {
          S[i].live[k] = 0;
          changed = true;
        } else {
          any = true;
        }
      }
      if (changed) upload(i);
    }
    return any;
  };
  std::vector<char> pupd_pending(np, 0);
  for (int iter = 0; iter < maxiter; ++iter) {
    std::vector<size_t> act = live_parts();
    if (act.empty()) break;
    std::vector<PrecJob> jobs;
    for (size_t i : act) {
      std::vector<int>& iters = *J[i].iters;
      for (int k = 0; k < J[i].Bin.n; ++k) iters[k] += S[i].live[k];
      jobs.push_back({J[i].L, S[i].B, S[i].q, S[i].p, pupd_pending[i] ? &S[i].pupd : nullptr});
      pupd_pending[i] = 0;
    }
    prec(jobs, true, nsap, nm, fz);
    for (size_t i : act) {
      St& s = S[i];
      hipStream_t st = J[i].L->st_;
      launch_dot<float>(s.B, s.n, s.r0, s.q, s.part, st);
      launch_finalize(s.B, s.part, 2, s.sc, FIN_ALPHA, st);
      // r -= alpha q; x += alpha p is deferred to the second half-step's update (two
      // spinor passes fewer), or done below for the RHS that leave here
      launch_update_xr<float>(s.B, s.n, J[i].x, s.p, s.r, s.q, s.sc, 0, nullptr, s.part, st, XR_R_ONLY);
      launch_finalize(s.B, s.part, 1, s.sc, FIN_RNORM, st);
    }
    std::vector<std::vector<int>> before(np);
    for (size_t i : act) before[i] = S[i].live;
    const bool more = converge(act);
    for (size_t i : act) {
      St& s = S[i];
      Lattice& L = *J[i].L;
      const int nb = J[i].Bin.n;
      s.fix.assign(nb, 0);
      bool any = false;
      for (int k = 0; k < nb; ++k) any |= (s.fix[k] = before[i][k] && !s.live[k]) != 0;
      if (!any) continue;
      int* dfix = (int*)L.work("bicg_fix", sizeof(int) * nb);
      QWS_HIP(hipMemcpyAsync(dfix, s.fix.data(), sizeof(int) * nb, hipMemcpyHostToDevice, L.st_));
      launch_update_x<float>(Batch{nb, s.B.stride, dfix}, s.n, J[i].x, s.p, s.sc, L.st_);
    }
    if (!more) break;
    act = live_parts();
    jobs.clear();
    for (size_t i : act) jobs.push_back({J[i].L, S[i].B, S[i].t, S[i].r, nullptr});
    prec(jobs, true, nsap, nm, fz);
    for (size_t i : act) {
      St& s = S[i];
      hipStream_t st = J[i].L->st_;
      launch_norm_dot<float>(s.B, s.n, s.t, s.r, s.part, st);
      launch_finalize(s.B, s.part, 3, s.sc, FIN_OMEGA, st);
      launch_update_xr<float>(s.B, s.n, J[i].x, s.r, s.r, s.t, s.sc, 1, s.r0, s.part, st, XR_DEFERRED, s.p);
      launch_finalize(s.B, s.part, 3, s.sc, FIN_RHO, st);
    }
    if (!converge(act)) break;
    // p = r + beta (p - omega q): done by the next q = D M p (bicgstab_precdd_s.cc:518-545)
    for (size_t i : live_parts()) {
      if (J[i].L->fold_pupd_) pupd_pending[i] = 1;
      else launch_update_p<float>(S[i].B, S[i].n, S[i].p, S[i].r, S[i].q, S[i].sc, J[i].L->st_);
    }
  }
  if (trace_on()) {
    std::vector<size_t> all;
    for (size_t i = 0; i < np; ++i) all.push_back(i);
    read(all);
    for (size_t i = 0; i < np; ++i)
      for (int k = 0; k < J[i].Bin.n; ++k)
        if ((*J[i].iters)[k] > 0)
          std::fprintf(stderr, "@C %zu.%d %5d %12.7e\n", i, k, (*J[i].iters)[k],
                       std::sqrt(J[i].L->host_scalars()[k].rnorm / S[i].bnorm[k]));
  }
}

// ------------------------------------------------------------------ Fused::approx_inv
// u = M p for every job (Lattice::approx_inv: normalise, FP32 inner solve, prec_s_,
// rescale), the inner solves and the final prec_s_ applications fused over the jobs.
// The FP16-storage method runs lattice after lattice.
void Fused::approx_inv(std::vector<AIJob>& jobs, const SolveParams& prm, SapFusion* fz) {
  if (prm.sap_prec == P16 || !fz || jobs.size() == 1) {
    for (AIJob& j : jobs) {
      if (!j.L->approx_inv(j.B, j.sc, *j.live, j.u, j.p, prm, *j.inner)) continue;
      // the FP16-stored SAP's inner solve failed: this batch goes on with the FP32 SAP,
      // starting with this preconditioner application, with the same number of sweeps
      // (the FP16 method runs nsap - 1 of them; near the critical kappa the Jacobi
      // block solves amplify, so an extra sweep would hurt)
      j.escalate = true;
      j.L->approx_inv(j.B, j.sc, *j.live, j.u, j.p, escalated(prm), *j.inner);
    }
    return;
  }
  const size_t np = jobs.size();
  std::vector<float*> bs(np), ts(np), xs(np);
  std::vector<std::vector<int>> it(np);
  std::vector<InnerJob> inner;
  for (size_t i = 0; i < np; ++i) {
    AIJob& j = jobs[i];
    Lattice& L = *j.L;
    const long long n = L.g_.V * 12;
    const int nb = j.B.n;
    const size_t bytes = L.spinor_bytes(P32) * nb;
    bs[i] = (float*)L.work("ai_b", bytes);
    ts[i] = (float*)L.work("ai_t", bytes);
    xs[i] = (float*)L.work("ai_x", bytes);
    double* part = L.reduction_partials(nb);
    launch_norm<double>(j.B, n, j.p, part, L.st_);
    launch_finalize(j.B, part, 1, j.sc, FIN_NRM, L.st_);
    launch_d2s_scaled<float>(j.B, L.g_.V, j.B.stride, j.B.stride, bs[i], j.p, j.sc, L.st_);
    it[i].resize(nb);
    for (int k = 0; k < nb; ++k) it[i][k] = (*j.live)[k] ? 0 : -1;
  }
  for (size_t i = 0; i < np; ++i)
    inner.push_back({jobs[i].L, Batch{jobs[i].B.n, jobs[i].L->g_.V * 24, jobs[i].L->d_inner_mask_}, ts[i], bs[i], &it[i]});
  bicg(inner, prm.tol_s, prm.maxiter_s, prm.nsap, prm.nm, fz);
  std::vector<PrecJob> pj;
  for (size_t i = 0; i < np; ++i) pj.push_back({jobs[i].L, jobs[i].B, xs[i], ts[i], nullptr});
  prec(pj, false, prm.nsap, prm.nm, fz);
  for (size_t i = 0; i < np; ++i) {
    AIJob& j = jobs[i];
    Lattice& L = *j.L;
    launch_s2d_scaled<float>(j.B, L.g_.V, j.B.stride, j.B.stride, j.u, xs[i], j.sc, L.st_);
    for (int k = 0; k < j.B.n; ++k)
      if ((*j.live)[k]) (*j.inner)[k] += it[i][k];
  }
}

// ------------------------------------------------------------------ Fused::outer_step
// One outer iteration (OuterBatch::step) of every running slot of every batch.
int Fused::outer_step(std::vector<OuterBatch*>& obs, SapFusion* fz) {
  const size_t nb = obs.size();
  std::vector<std::vector<int>> live(nb);
  std::vector<size_t> act;
  for (size_t i = 0; i < nb; ++i) {
    OuterBatch& o = *obs[i];
    live[i].assign(o.cap_, 0);
    int nlive = 0;
    for (int k = 0; k < o.cap_; ++k) nlive += (live[i][k] = o.state_[k] == OuterBatch::RUNNING);
    if (nlive) act.push_back(i);
  }
  if (act.empty()) return 0;
  auto batch = [&](size_t i) { return Batch{obs[i]->cap_, obs[i]->stride_, obs[i]->d_mask_}; };
  auto check_all = [&](const std::vector<size_t>& which, int half) {
    bool trace = trace_on();
    std::vector<std::pair<Lattice*, std::pair<const Scalars*, int>>> v;
    for (size_t i : which)
      if (obs[i]->prm_.tol >= 0 || trace) v.push_back({&obs[i]->L_, {obs[i]->sc_, obs[i]->cap_}});
    read_all(v);
    for (size_t i : which) obs[i]->check(live[i], half);
  };
  auto still = [&](const std::vector<size_t>& which) {
    std::vector<size_t> v;
    for (size_t i : which)
      if (any_of(live[i])) v.push_back(i);
    return v;
  };
  // u = M v for the batches in `which` (v = p or r), those with the same SAP precision
  // together; a batch whose FP16-stored SAP failed continues with the FP32 SAP
  auto apply_m = [&](const std::vector<size_t>& which, bool of_p) {
    // each batch's SAP precision as the half-step starts (a batch that switches during
    // the FP16 pass has its preconditioner application redone there, not again below)
    std::vector<char> f16(obs.size(), 0);
    for (size_t i : which) f16[i] = obs[i]->prm_.sap_prec == P16;
    for (int pass = 0; pass < 2; ++pass) {
      std::vector<AIJob> ai;
      std::vector<size_t> idx;
      for (size_t i : which) {
        OuterBatch& o = *obs[i];
        if ((f16[i] != 0) != (pass == 0)) continue;
        ai.push_back({&o.L_, batch(i), o.sc_, &live[i], o.u_, of_p ? o.p_ : o.r_, &o.inner_, false});
        idx.push_back(i);
      }
      if (ai.empty()) continue;
      approx_inv(ai, obs[idx[0]]->prm_, fz);
      for (size_t j = 0; j < ai.size(); ++j)
        if (ai[j].escalate) {
          obs[idx[j]]->prm_ = escalated(obs[idx[j]]->prm_);
          obs[idx[j]]->escalated_ = true;
        }
    }
  };
  for (size_t i : act) {
    OuterBatch& o = *obs[i];
    o.upload_mask(live[i]);
    for (int k = 0; k < o.cap_; ++k) o.iters_[k] += live[i][k];
  }
  // u = M p ; q = D u ; alpha ; x += alpha u ; r -= alpha q
  apply_m(act, true);
  for (size_t i : act) {
    OuterBatch& o = *obs[i];
    Lattice& L = o.L_;
    const Batch B = batch(i);
    const long long n = L.volume() * 12;
    double* part = L.reduction_partials(o.cap_);
    L.dslash_batch(B, o.q_, o.u_);
    launch_dot<double>(B, n, o.r0_, o.q_, part, L.st_);
    launch_finalize(B, part, 2, o.sc_, FIN_ALPHA, L.st_);
    launch_update_xr<double>(B, n, o.x_, o.u_, o.r_, o.q_, o.sc_, 0, nullptr, part, L.st_);
    launch_finalize(B, part, 1, o.sc_, FIN_RNORM, L.st_);
  }
  check_all(act, 0);
  std::vector<size_t> act2 = still(act);
  if (!act2.empty()) {
    // u = M r ; t = D u ; omega ; x += omega u ; r -= omega t ; rho, beta
    apply_m(act2, false);
    for (size_t i : act2) {
      OuterBatch& o = *obs[i];
      Lattice& L = o.L_;
      const Batch B = batch(i);
      const long long n = L.volume() * 12;
      double* part = L.reduction_partials(o.cap_);
      L.dslash_batch(B, o.t_, o.u_);
      launch_norm_dot<double>(B, n, o.t_, o.r_, part, L.st_);
      launch_finalize(B, part, 3, o.sc_, FIN_OMEGA, L.st_);
      launch_update_xr<double>(B, n, o.x_, o.u_, o.r_, o.t_, o.sc_, 1, o.r0_, part, L.st_);
      launch_finalize(B, part, 3, o.sc_, FIN_RHO, L.st_);
    }
    check_all(act2, 1);
    for (size_t i : still(act2)) {
      OuterBatch& o = *obs[i];
      launch_update_p<double>(batch(i), o.L_.volume() * 12, o.p_, o.r_, o.q_, o.sc_, o.L_.st_);
    }
  }
  // the reference loops iter = 1 .. maxiter-1 and reports maxiter when it ran out
  int running = 0;
  for (OuterBatch* ob : obs) {
    OuterBatch& o = *ob;
    for (int k = 0; k < o.cap_; ++k)
      if (o.state_[k] == OuterBatch::RUNNING && o.iters_[k] >= o.prm_.maxiter - 1) {
        o.state_[k] = OuterBatch::FINISHED;
        o.iters_[k] = o.prm_.maxiter;
      }
    running += o.active_count();
  }
  return running;
}

// ------------------------------------------------------------------ Fused::solve
std::vector<std::vector<SolveStats>> Fused::solve(const std::vector<Lattice*>& lats,
                                                  const std::vector<std::vector<double*>>& x,
                                                  const std::vector<std::vector<const double*>>& b,
                                                  const SolveParams& prm, SapFusion* fz) {
  const size_t nl = lats.size();
  std::vector<std::vector<SolveStats>> out(nl);
  auto t0 = std::chrono::steady_clock::now();
  std::vector<OuterBatch*> obs;
  std::vector<std::vector<int>> open(nl);
  for (size_t i = 0; i < nl; ++i) {
    Lattice& L = *lats[i];
    const int nrhs = (int)x[i].size();
    out[i].resize(nrhs);
    if (nrhs == 0) continue;
    if (!L.batch_ || L.batch_->capacity() != nrhs) {
      L.batch_.reset();
      L.batch_.reset(new OuterBatch(L, nrhs, prm));
    }
    OuterBatch& ob = *L.batch_;
    ob.set_params(prm);
    // a zero source has the exact solution x = 0 (the relative residual is undefined
    // there, and the normalised inner solve would divide by zero): it takes no slot
    open[i].assign(nrhs, 1);
    for (int k = 0; k < nrhs; ++k) {
      if (L.norm2(P64, b[i][k]) == 0.0) {
        open[i][k] = 0;
        out[i][k].true_residual = 0;
        QWS_HIP(hipMemsetAsync(x[i][k], 0, L.spinor_bytes(P64), L.st_));
        ob.release(k);
      } else {
        ob.admit(k, b[i][k], x[i][k]);
      }
    }
    obs.push_back(&ob);
  }
  for (;;) {
    while (outer_step(obs, fz) > 0) {
    }
    // FP64 true residual of every slot that just finished; restart those above true_tol
    bool again = false;
    for (size_t i = 0; i < nl; ++i) {
      if (x[i].empty()) continue;
      OuterBatch& ob = *lats[i]->batch_;
      for (size_t k = 0; k < x[i].size(); ++k) {
        if (!open[i][k]) continue;
        SolveStats& st = out[i][k];
        st.outer_iters += ob.iterations((int)k);
        st.inner_iters += ob.inner_iterations((int)k);
        st.true_residual = lats[i]->true_residual(ob.x((int)k), ob.b((int)k));
        if (prm.true_tol > 0 && st.true_residual > prm.true_tol && st.restarts < prm.max_restarts) {
          st.restarts++;
          ob.admit((int)k, ob.b((int)k), ob.x((int)k));
          again = true;
        } else {
          open[i][k] = 0;
        }
      }
    }
    if (!again) break;
  }
  for (size_t i = 0; i < nl; ++i) {
    if (x[i].empty()) continue;
    Lattice& L = *lats[i];
    OuterBatch& ob = *L.batch_;
    for (size_t k = 0; k < x[i].size(); ++k) out[i][k].sap_escalated = ob.escalated() ? 1 : 0;
    for (size_t k = 0; k < x[i].size(); ++k)
      if (ob.running((int)k) || ob.finished((int)k))
        QWS_HIP(hipMemcpyAsync(x[i][k], ob.x((int)k), L.spinor_bytes(P64), hipMemcpyDeviceToDevice, L.st_));
    L.sync();
  }
  const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  for (auto& v : out)
    for (auto& s : v) s.seconds = sec;
  return out;
}

}  // namespace qws
