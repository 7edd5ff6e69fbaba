// serve.h -- a solve server on one GPU: gauge fields stay resident in HBM; requests of
// mixed volume, kappa and right-hand-side count are admitted into slots as slots and
// memory free up, and the admitted requests run at the same time, each on its own
// HIP stream.  Every request builds its clover inverse for its own kappa on the shared
// links and solves all its right-hand sides as one batch (qwsg_solve_batch); its
// results do not depend on what else runs (no shared state, deterministic reductions).
#pragma once

#include <memory>
#include <vector>

#include "lattice.h"

namespace qws {

struct FieldSpec {
  int L[4], B[4];
  double eps;               // links exp(i eps H), H Gaussian traceless Hermitian (Philox)
  unsigned long long seed;
};

struct Request {
  int field;                // index into the resident fields
  double kappa;
  int nrhs;                 // 1..MAX_BATCH Gaussian sources
  unsigned long long seed;  // source k uses seed * 1000 + k
  double arrival;           // seconds after the start of the trace
};

struct RequestResult {
  double arrival = 0, admit = 0, finish = 0;  // seconds after the start of the trace
  int outer_iters = 0, inner_iters = 0, restarts = 0;  // summed over the request's RHS
  double true_residual_max = 0;
  double bytes = 0;                       // memory reserved for the request
  unsigned long long digest = 0;          // digest of the FP64 solutions (launch_digest)
};

struct ServeConfig {
  int links32 = 12, tbc = -1;
  double csw = 1.0;
  SolveParams prm;
  int slots = 4;            // requests in flight at once
  double mem_budget = 0;    // bytes for in-flight requests (0: 90% of free HBM after the fields)
};

class Server {
 public:
  Server(const std::vector<FieldSpec>& fields, const ServeConfig& cfg);
  ~Server();
  // Replays the trace: each request is released at its arrival time, admitted in
  // arrival order when a slot is free and its memory fits the budget.
  std::vector<RequestResult> run(const std::vector<Request>& trace);
  // device memory one request needs beyond the resident links (clover, spinors, work)
  static double request_bytes(long long V, int nrhs);
  double mem_budget() const { return cfg_.mem_budget; }

 private:
  RequestResult serve_one(const Request& r);
  std::vector<std::unique_ptr<Lattice>> fields_;
  ServeConfig cfg_;
};

}  // namespace qws
