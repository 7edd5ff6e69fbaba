// serve.cpp -- the solve server (serve.h).
#include "serve.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <numeric>
#include <stdexcept>
#include <thread>

#include "kernels.h"
#include "multi.h"

namespace qws {

// A pooled slot: a lattice on one resident field (sharing its links) with its own
// clover inverse and an outer batch of `cap` right-hand sides, on the server's stream.
struct Server::Slot {
  int field = -1, cap = 0;
  double bytes = 0;
  std::unique_ptr<Lattice> lat;
  std::unique_ptr<OuterBatch> ob;
  unsigned long long* dig = nullptr;  // device word for the solutions' digest
  bool busy = false;
  long long used = 0;  // when the slot last took a request (eviction order)
  // the request being served
  int req = -1, nrhs = 0;
  std::vector<SolveStats> st;
  std::vector<int> open;
  RequestResult res;

  ~Slot() {
    if (dig) (void)hipFree(dig);
  }
};

Server::Server(const std::vector<FieldSpec>& fields, const ServeConfig& cfg) : cfg_(cfg) {
  if (cfg_.slots < 1 || cfg_.slots > 16) throw std::invalid_argument("slots must be in 1..16");
  check_sap_params(cfg_.prm.nsap, cfg_.prm.nm);
  for (const FieldSpec& f : fields) {
    fields_.emplace_back(new Lattice(f.L, f.B, cfg_.tbc, cfg_.links32));
    fields_.back()->generate_gauge(f.eps, f.seed);
  }
  QWS_HIP(hipStreamCreate(&st_));
  fz_.reset(new SapFusion(st_));
  if (cfg_.mem_budget <= 0) {
    size_t free_b = 0, total_b = 0;
    QWS_HIP(hipMemGetInfo(&free_b, &total_b));
    cfg_.mem_budget = 0.9 * (double)free_b;
  }
}

Server::~Server() {
  if (st_) (void)hipStreamSynchronize(st_);
  pool_.clear();
  fz_.reset();
  fields_.clear();
  if (st_) (void)hipStreamDestroy(st_);
}

double Server::request_bytes(long long V, int nrhs) {
  return 1.1 * (double)V * (864.0 + 192.0 + 2592.0 * nrhs) + 64.0 * 1024 * 1024;
}

int Server::slot_capacity(int nrhs) {
  for (int c : {1, 2, 4, 8, 12, 16, 32, 64})
    if (nrhs <= c) return c;
  return MAX_BATCH;
}

std::vector<RequestResult> Server::run(const std::vector<Request>& trace) {
  const int n = (int)trace.size();
  std::vector<RequestResult> out(n);
  std::vector<double> need(n);
  for (int i = 0; i < n; ++i) {
    const Request& r = trace[i];
    if (r.field < 0 || r.field >= (int)fields_.size()) throw std::invalid_argument("request names no resident field");
    if (r.nrhs < 1 || r.nrhs > MAX_BATCH) throw std::invalid_argument("request RHS count must be in 1..64");
    need[i] = request_bytes(fields_[r.field]->volume(), slot_capacity(r.nrhs));
    if (need[i] > cfg_.mem_budget) throw std::invalid_argument("a request needs more memory than the budget");
  }
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return trace[a].arrival < trace[b].arrival; });

  using clock = std::chrono::steady_clock;
  const auto t0 = clock::now();
  auto since = [&] { return std::chrono::duration<double>(clock::now() - t0).count(); };

  // a free pooled slot on the field with room for the request (the smallest such), a
  // new one if it fits the budget (giving back idle slots, least recently used first),
  // or nullptr for now
  long long uses = 0;
  auto acquire = [&](int field, int cap, double bytes) -> Slot* {
    Slot* best = nullptr;
    for (auto& s : pool_)
      if (!s->busy && s->field == field && s->cap >= cap && (!best || s->cap < best->cap)) best = s.get();
    if (best) {
      best->used = ++uses;
      return best;
    }
    auto in_use = [&] {
      double u = 0;
      for (auto& s : pool_) u += s->bytes;
      return u;
    };
    while (in_use() + bytes > cfg_.mem_budget) {
      auto it = pool_.end();
      for (auto j = pool_.begin(); j != pool_.end(); ++j)
        if (!(*j)->busy && (it == pool_.end() || (*j)->used < (*it)->used)) it = j;
      if (it == pool_.end()) return nullptr;
      QWS_HIP(hipStreamSynchronize(st_));
      pool_.erase(it);
    }
    std::unique_ptr<Slot> s(new Slot);
    const Lattice& F = *fields_[field];
    const Geo& g = F.geo();
    s->field = field, s->cap = cap, s->bytes = bytes;
    s->lat.reset(new Lattice(g.L, g.B, cfg_.tbc, cfg_.links32));
    s->lat->share_gauge(F);
    s->lat->set_stream(st_);
    s->ob.reset(new OuterBatch(*s->lat, cap, cfg_.prm));
    QWS_HIP(hipMalloc((void**)&s->dig, sizeof(unsigned long long)));
    s->used = ++uses;
    pool_.push_back(std::move(s));
    return pool_.back().get();
  };

  // the request's clover inverse for its kappa, its sources, and its RHS admitted
  auto start = [&](Slot& s, int i) {
    const Request& r = trace[i];
    Lattice& L = *s.lat;
    OuterBatch& ob = *s.ob;
    s.busy = true, s.req = i, s.nrhs = r.nrhs;
    s.st.assign(r.nrhs, SolveStats());
    s.open.assign(r.nrhs, 1);
    s.res = RequestResult();
    s.res.arrival = r.arrival;
    s.res.admit = since();
    s.res.bytes = s.bytes;
    L.set_kappa(r.kappa);
    L.compute_clover(cfg_.csw);
    ob.set_params(cfg_.prm);
    for (int k = 0; k < s.cap; ++k) {
      if (k < r.nrhs) {
        L.gaussian_spinor(P64, ob.b(k), r.seed * 1000ULL + (unsigned long long)k);
        ob.admit(k, ob.b(k), nullptr);
      } else {
        ob.release(k);
      }
    }
  };

  // all RHS of the slot's outer batch have stopped: FP64 true residuals, restarts of
  // those above true_tol (bicgstab_dd_mix_ again from x); true when the request is done
  auto finish = [&](Slot& s) {
    Lattice& L = *s.lat;
    OuterBatch& ob = *s.ob;
    const SolveParams& prm = cfg_.prm;
    bool again = false;
    for (int k = 0; k < s.nrhs; ++k) {
      if (!s.open[k]) continue;
      SolveStats& st = s.st[k];
      st.outer_iters += ob.iterations(k);
      st.inner_iters += ob.inner_iterations(k);
      st.true_residual = L.true_residual(ob.x(k), ob.b(k));
      if (prm.true_tol > 0 && st.true_residual > prm.true_tol && st.restarts < prm.max_restarts) {
        st.restarts++;
        ob.admit(k, ob.b(k), ob.x(k));
        again = true;
      } else {
        s.open[k] = 0;
      }
    }
    if (again) return false;
    RequestResult& res = s.res;
    for (const SolveStats& st : s.st) {
      res.outer_iters += st.outer_iters;
      res.inner_iters += st.inner_iters;
      res.restarts += st.restarts;
      res.true_residual_max = std::max(res.true_residual_max, st.true_residual);
    }
    QWS_HIP(hipMemsetAsync(s.dig, 0, sizeof(unsigned long long), st_));
    for (int k = 0; k < s.nrhs; ++k) launch_digest(L.volume() * 24, ob.x(k), s.dig, st_);
    QWS_HIP(hipMemcpyAsync(&res.digest, s.dig, sizeof(unsigned long long), hipMemcpyDeviceToHost, st_));
    QWS_HIP(hipStreamSynchronize(st_));
    res.finish = since();
    out[s.req] = res;
    s.busy = false, s.req = -1;
    return true;
  };

  std::deque<int> pending;
  std::vector<Slot*> running;
  int next = 0, done = 0;
  while (done < n) {
    while (next < n && trace[order[next]].arrival <= since()) pending.push_back(order[next++]);
    // admission in arrival order: the head of the queue waits for a slot and memory
    while (!pending.empty() && (int)running.size() < cfg_.slots) {
      const int i = pending.front();
      Slot* s = acquire(trace[i].field, slot_capacity(trace[i].nrhs), need[i]);
      if (!s) break;
      pending.pop_front();
      start(*s, i);
      running.push_back(s);
    }
    if (running.empty()) {
      if (next >= n) throw std::logic_error("server: requests left that cannot be admitted");
      std::this_thread::sleep_until(t0 + std::chrono::duration_cast<clock::duration>(
                                             std::chrono::duration<double>(trace[order[next]].arrival)));
      continue;
    }
    // one outer iteration of every admitted request, fused over their lattices
    std::vector<OuterBatch*> obs;
    for (Slot* s : running) obs.push_back(s->ob.get());
    Fused::outer_step(obs, fz_.get());
    for (size_t j = 0; j < running.size();) {
      Slot* s = running[j];
      if (s->ob->active_count() == 0 && finish(*s)) {
        ++done;
        running.erase(running.begin() + (long)j);
      } else {
        ++j;
      }
    }
  }
  return out;
}

}  // namespace qws
