// multi.h -- solves over several lattices in lockstep with fused SAP launches
// (multi.cpp).
#pragma once

#include <string>
#include <utility>
#include <vector>

#include "lattice.h"

namespace qws {

// Device memory for the descriptor tables of fused launches (a fixed arena, filled
// through a pinned host copy on the solve's stream; tables with the same contents are
// uploaded once).  Nothing is allocated after construction.
class SapFusion {
 public:
  explicit SapFusion(hipStream_t st, size_t bytes = (size_t)4 << 20);
  ~SapFusion();
  SapFusion(const SapFusion&) = delete;
  SapFusion& operator=(const SapFusion&) = delete;
  hipStream_t stream() const { return st_; }
  const void* upload(const void* data, size_t bytes);
  // A sequence of launches identified by `key` (the bytes of its launch records, whose
  // descriptor tables live in the arena): the first time it runs as given by
  // `launch`, the second time it is captured into a HIP graph, and from then on the
  // graph replays it.  Graphs are dropped with the arena's contents.  The launches
  // must be kernels only (no memory set or copy nodes).
  template <typename F> void run(const std::string& key, F&& launch);

 private:
  hipStream_t st_;
  size_t cap_, off_ = 0;
  char* dev_ = nullptr;
  char* host_ = nullptr;
  std::vector<std::pair<std::string, size_t>> entries_;
  struct Graph {
    std::string key;
    int seen = 0;
    hipGraphExec_t exec = nullptr;
  };
  std::vector<Graph> graphs_;
  bool use_graphs_ = true;
  void drop_graphs();
};

template <typename F> void SapFusion::run(const std::string& key, F&& launch) {
  if (!use_graphs_) {
    launch();
    return;
  }
  Graph* g = nullptr;
  for (Graph& x : graphs_)
    if (x.key == key) g = &x;
  if (!g) {
    if (graphs_.size() >= 64) drop_graphs();
    graphs_.push_back(Graph{key, 0, nullptr});
    g = &graphs_.back();
  }
  if (g->exec) {
    QWS_HIP(hipGraphLaunch(g->exec, st_));
    return;
  }
  if (++g->seen < 2) {
    launch();
    return;
  }
  hipGraph_t graph = nullptr;
  QWS_HIP(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal));
  try {
    launch();
  } catch (...) {
    (void)hipStreamEndCapture(st_, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    throw;
  }
  QWS_HIP(hipStreamEndCapture(st_, &graph));
  hipGraphExec_t exec = nullptr;
  const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  QWS_HIP(e);
  g->exec = exec;
  QWS_HIP(hipGraphLaunch(g->exec, st_));
}

struct Fused {
  // out = D M_SAP in (prec_ddd_s_) or M_SAP in (prec_s_) on one lattice's batch
  struct PrecJob {
    Lattice* L;
    Batch B;
    float* out;
    const float* in;
    const SapSrcUpdate* upd;
  };
  static void prec(std::vector<PrecJob>& jobs, bool ddd, int nsap, int nm, SapFusion* fz);
  // the FP32 inner solve D M_SAP x = b, x0 = 0 (Lattice::bicg_t without hp)
  struct InnerJob {
    Lattice* L;
    Batch Bin;
    float* x;
    const float* b;
    std::vector<int>* iters;
  };
  static void bicg(std::vector<InnerJob>& jobs, double tol, int maxiter, int nsap, int nm, SapFusion* fz);
  // u = M p (Lattice::approx_inv)
  struct AIJob {
    Lattice* L;
    Batch B;
    Scalars* sc;
    const std::vector<int>* live;
    double* u;
    const double* p;
    std::vector<int>* inner;
    bool escalate;  // out: the FP16-stored SAP failed and the FP32 SAP redid the job
  };
  static void approx_inv(std::vector<AIJob>& jobs, const SolveParams& prm, SapFusion* fz);
  // one outer iteration of every running slot of every batch (each on its own
  // lattice; all on fz's stream when fz is given); returns the slots still running
  static int outer_step(std::vector<OuterBatch*>& obs, SapFusion* fz);
  // Lattice::solve_batch on several lattices at once: lattice i solves D x[i][k] =
  // b[i][k] (x holds the initial guess) for its own right-hand sides; the lattices
  // must share one stream (fz's, when fz is given).  Stats per lattice and RHS.
  static std::vector<std::vector<SolveStats>> solve(const std::vector<Lattice*>& lats,
                                                    const std::vector<std::vector<double*>>& x,
                                                    const std::vector<std::vector<const double*>>& b,
                                                    const SolveParams& prm, SapFusion* fz);
};

}  // namespace qws
