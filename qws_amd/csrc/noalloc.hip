// noalloc.hip -- qwsg_prec_ddd_noalloc: the SAP preconditioner application
// out = D M_SAP in (FP32) for callers that cannot allocate device memory (code under
// HIP graph capture, serving loops that must not call hipMalloc).  Its work buffers
// (the SAP iterates s and q and the face buffer, or the step-by-step path's third
// spinor) live in one fixed device array instead of allocated memory.
//
// Each GPU the library runs on holds its own copy of the array until the process
// ends.  Every call on a GPU uses that GPU's copy: calls that overlap on one GPU
// (other host threads, or other streams before the previous call has finished)
// race on it.  The array keeps the last call's contents; nothing in it has to start
// at zero, because prec_ddd writes each of its work buffers before reading it.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/qws_mi355x.h"
#include "handle.h"

namespace {

__device__ __align__(256) unsigned char scratch[QWSG_NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void* g_scratch[MAX_DEVICES];  // per device: the address of its copy of `scratch`

// The device address of the current device's copy of `scratch`, looked up on the
// first call on that device (with it current) and kept.  In host code the array's
// name is only a stand-in for the symbol; this address is what kernels receive.
void* scratch_on_current_device() {
  int dev = -1;
  QWS_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) throw std::runtime_error("device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    QWS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw std::runtime_error("hipGetSymbolAddress returned no address for the scratch array");
    g_scratch[dev] = p;
  }
  return g_scratch[dev];
}

}  // namespace

extern "C" {

int qwsg_prec_ddd_noalloc(qwsg_lattice* h, void* out, const void* in, int nsap, int nm) {
  return qws::guarded([&] {
    if (!h || !h->lat) throw std::invalid_argument("qwsg_prec_ddd_noalloc: no lattice");
    const size_t need = h->lat->prec_ddd_scratch_bytes();
    if (need > (size_t)QWSG_NOALLOC_SCRATCH_BYTES)
      throw std::invalid_argument("qwsg_prec_ddd_noalloc: this lattice needs " + std::to_string(need) +
                                  " bytes of scratch, more than the " +
                                  std::to_string((size_t)QWSG_NOALLOC_SCRATCH_BYTES) +
                                  " the library holds (with 4^4 blocks: 24^3x48 fits, 32^3x64 does not)");
    h->lat->prec_ddd_scratch(out, in, nsap, nm, scratch_on_current_device(), (size_t)QWSG_NOALLOC_SCRATCH_BYTES);
  });
}

}  // extern "C"
