// serve.cpp -- the solve server (serve.h).
#include "serve.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <exception>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <thread>

#include "kernels.h"

namespace qws {

Server::Server(const std::vector<FieldSpec>& fields, const ServeConfig& cfg) : cfg_(cfg) {
  if (cfg_.slots < 1 || cfg_.slots > 16) throw std::invalid_argument("slots must be in 1..16");
  for (const FieldSpec& f : fields) {
    fields_.emplace_back(new Lattice(f.L, f.B, cfg_.tbc, cfg_.links32));
    fields_.back()->generate_gauge(f.eps, f.seed);
  }
  if (cfg_.mem_budget <= 0) {
    size_t free_b = 0, total_b = 0;
    QWS_HIP(hipMemGetInfo(&free_b, &total_b));
    cfg_.mem_budget = 0.9 * (double)free_b;
  }
}

Server::~Server() = default;

// clover FP64 + FP32 (864 B/site) and per RHS: x, b (FP64), 8 outer FP64 vectors,
// 5 inner and 3 approx_inv FP32 vectors, 3 SAP FP32 vectors (2976 B/site), +10 %
double Server::request_bytes(long long V, int nrhs) {
  return 1.1 * (double)V * (864.0 + 2976.0 * nrhs) + 64.0 * 1024 * 1024;
}

RequestResult Server::serve_one(const Request& r) {
  const Lattice& F = *fields_[r.field];
  const Geo& g = F.geo();
  Lattice lat(g.L, g.B, cfg_.tbc, cfg_.links32);
  lat.share_gauge(F);
  lat.set_kappa(r.kappa);
  lat.compute_clover(cfg_.csw);
  std::vector<double*> x(r.nrhs, nullptr), b(r.nrhs, nullptr);
  RequestResult res;
  unsigned long long* dig = nullptr;
  try {
    for (int k = 0; k < r.nrhs; ++k) {
      x[k] = (double*)lat.alloc_spinor(P64);
      b[k] = (double*)lat.alloc_spinor(P64);
      lat.gaussian_spinor(P64, b[k], r.seed * 1000ULL + (unsigned long long)k);
    }
    std::vector<SolveStats> st = lat.solve_batch(x.data(), b.data(), r.nrhs, cfg_.prm);
    for (const SolveStats& s : st) {
      res.outer_iters += s.outer_iters;
      res.inner_iters += s.inner_iters;
      res.restarts += s.restarts;
      res.true_residual_max = std::max(res.true_residual_max, s.true_residual);
    }
    QWS_HIP(hipMalloc(&dig, sizeof(unsigned long long)));
    QWS_HIP(hipMemsetAsync(dig, 0, sizeof(unsigned long long), lat.stream()));
    for (int k = 0; k < r.nrhs; ++k) launch_digest(lat.volume() * 24, x[k], dig, lat.stream());
    QWS_HIP(hipMemcpyAsync(&res.digest, dig, sizeof(unsigned long long), hipMemcpyDeviceToHost, lat.stream()));
    lat.sync();
  } catch (...) {
    for (double* p : x) (void)hipFree(p);
    for (double* p : b) (void)hipFree(p);
    if (dig) (void)hipFree(dig);
    throw;
  }
  for (double* p : x) QWS_HIP(hipFree(p));
  for (double* p : b) QWS_HIP(hipFree(p));
  QWS_HIP(hipFree(dig));
  return res;
}

std::vector<RequestResult> Server::run(const std::vector<Request>& trace) {
  const int n = (int)trace.size();
  std::vector<RequestResult> out(n);
  std::vector<double> need(n);
  for (int i = 0; i < n; ++i) {
    const Request& r = trace[i];
    if (r.field < 0 || r.field >= (int)fields_.size()) throw std::invalid_argument("request names no resident field");
    if (r.nrhs < 1 || r.nrhs > MAX_BATCH) throw std::invalid_argument("request RHS count must be in 1..64");
    need[i] = request_bytes(fields_[r.field]->volume(), r.nrhs);
    if (need[i] > cfg_.mem_budget) throw std::invalid_argument("a request needs more memory than the budget");
  }
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return trace[a].arrival < trace[b].arrival; });

  using clock = std::chrono::steady_clock;
  const auto t0 = clock::now();
  auto since = [&] { return std::chrono::duration<double>(clock::now() - t0).count(); };
  std::mutex m;
  std::condition_variable cv;
  std::deque<int> pending;
  int released = 0;
  double in_use = 0;
  std::exception_ptr err;

  auto worker = [&] {
    for (;;) {
      std::unique_lock<std::mutex> lk(m);
      cv.wait(lk, [&] {
        if (err) return true;
        if (!pending.empty()) return in_use == 0 || in_use + need[pending.front()] <= cfg_.mem_budget;
        return released == n;
      });
      if (err || pending.empty()) return;
      const int i = pending.front();
      pending.pop_front();
      in_use += need[i];
      lk.unlock();
      const double admit = since();
      try {
        RequestResult r = serve_one(trace[i]);
        r.arrival = trace[i].arrival;
        r.admit = admit;
        r.finish = since();
        r.bytes = need[i];
        out[i] = r;
      } catch (...) {
        std::lock_guard<std::mutex> g(m);
        if (!err) err = std::current_exception();
      }
      lk.lock();
      in_use -= need[i];
      lk.unlock();
      cv.notify_all();
    }
  };
  std::vector<std::thread> pool;
  for (int s = 0; s < cfg_.slots; ++s) pool.emplace_back(worker);
  for (int k = 0; k < n; ++k) {
    const int i = order[k];
    std::this_thread::sleep_until(t0 + std::chrono::duration_cast<clock::duration>(
                                           std::chrono::duration<double>(trace[i].arrival)));
    {
      std::lock_guard<std::mutex> g(m);
      if (err) break;
      pending.push_back(i);
      ++released;
    }
    cv.notify_all();
  }
  {
    std::lock_guard<std::mutex> g(m);
    released = n;
  }
  cv.notify_all();
  for (auto& t : pool) t.join();
  if (err) std::rethrow_exception(err);
  return out;
}

}  // namespace qws
