// handle.h -- what the C ABI's source files share: the opaque lattice handle and
// the wrapper that turns exceptions into return codes and qwsg_last_error() text.
#pragma once

#include <stdexcept>
#include <string>

#include "lattice.h"

struct qwsg_lattice {
  qws::Lattice* lat;
};

namespace qws {

// the calling thread's last error message (qwsg_last_error)
std::string& last_error();

// invalid_argument when a pointer argument of entry point `what` is null
inline void need(const char*) {}
template <typename P, typename... R> void need(const char* what, P p, R... rest) {
  if (!p) throw std::invalid_argument(std::string(what) + ": null pointer argument");
  need(what, rest...);
}

// the lattice of a handle (invalid_argument for a null handle)
inline Lattice* lat_of(qwsg_lattice* h) {
  if (!h || !h->lat) throw std::invalid_argument("null lattice handle");
  return h->lat;
}

// 0 on success; 1 invalid argument, 2 HIP runtime error, 3 any other failure
template <typename F> int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const HipError& e) {
    last_error() = e.what();
    return 2;
  } catch (const std::invalid_argument& e) {
    last_error() = e.what();
    return 1;
  } catch (const std::exception& e) {
    last_error() = e.what();
    return 3;
  }
}

}  // namespace qws
