// serve.h -- a solve server on one GPU with continuous batching.  The gauge fields
// stay resident in HBM.  Requests of mixed volume, kappa and right-hand-side count are
// admitted, as slots and memory free up, into the one batch that runs on the GPU: all
// admitted requests advance together, one outer iteration at a time, each SAP
// half-sweep one launch over the blocks of all of them (Fused), and a request that
// arrives while others run joins at the next outer iteration.  Each admitted request
// takes a pooled slot (a lattice sharing its field's links, its clover inverse for its
// kappa, its solver vectors) that later requests reuse: once the pool has warmed up,
// serving allocates nothing.  A request's results do not depend on what else runs
// (its blocks compute as alone, its reductions are its own and deterministic).
#pragma once

#include <memory>
#include <vector>

#include "lattice.h"

namespace qws {

class SapFusion;

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
  double bytes = 0;                       // memory of the slot that served the request
  unsigned long long digest = 0;          // digest of the FP64 solutions (launch_digest)
};

struct ServeConfig {
  int links32 = 12, tbc = -1;
  double csw = 1.0;
  SolveParams prm;
  int slots = 4;            // requests in flight at once
  double mem_budget = 0;    // bytes for the slot pool (0: 90% of free HBM after the fields)
};

class Server {
 public:
  Server(const std::vector<FieldSpec>& fields, const ServeConfig& cfg);
  ~Server();
  // Replays the trace: each request is released at its arrival time and admitted in
  // arrival order when fewer than `slots` requests run and a pooled slot for it is free
  // (on its field, with room for its right-hand sides: the smallest such) or a new one
  // fits the budget (idle slots are given back to make room, least recently used first).
  std::vector<RequestResult> run(const std::vector<Request>& trace);
  // device memory of a slot for nrhs right-hand sides on V sites: clover inverse (FP64 +
  // FP32), the outer solver's 8 FP64 vectors, the inner solver's, preconditioner's and
  // SAP's 11 FP32 vectors (with the face buffer) per RHS, residual vector, +10 %
  static double request_bytes(long long V, int nrhs);
  // right-hand-side capacity of the slot that serves nrhs (1, 2, 4, 8, 12, 16, 32, 64)
  static int slot_capacity(int nrhs);
  double mem_budget() const { return cfg_.mem_budget; }

 private:
  struct Slot;
  std::vector<std::unique_ptr<Lattice>> fields_;
  std::vector<std::unique_ptr<Slot>> pool_;
  ServeConfig cfg_;
  hipStream_t st_ = nullptr;
  std::unique_ptr<SapFusion> fz_;
};

}  // namespace qws
