// hostcopy.cpp -- see hostcopy.h.  Plain C++ (no device code), so that the CPU tests
// can build it on its own.
#include "hostcopy.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <thread>

namespace qws {

void copy_part(size_t n, int T, int i, size_t& o, size_t& e) {
  const size_t part = ((n + T - 1) / T + 63) & ~(size_t)63;
  o = std::min(n, part * i);
  e = i == T - 1 ? n : std::min(n, part * (i + 1));
}

void par_memcpy(void* dst, const void* src, size_t n) {
  constexpr int T = 8;
  if (n < ((size_t)1 << 20)) {
    std::memcpy(dst, src, n);
    return;
  }
  std::thread th[T - 1];
  int started = 0;  // threads th[0 .. started-1] copy parts 1 .. started
  try {
    for (int i = 1; i < T; ++i, ++started) {
      size_t o, e;
      copy_part(n, T, i, o, e);
      th[i - 1] = std::thread([=] { std::memcpy((char*)dst + o, (const char*)src + o, e - o); });
    }
  } catch (...) {
    for (int i = started + 1; i < T; ++i) {
      size_t o, e;
      copy_part(n, T, i, o, e);
      std::memcpy((char*)dst + o, (const char*)src + o, e - o);
    }
  }
  size_t o, e;
  copy_part(n, T, 0, o, e);
  std::memcpy(dst, src, e);
  for (int i = 0; i < started; ++i) th[i].join();
}

size_t pin_chunk_from_env() {
  const char* v = std::getenv("QWS_MI355X_PIN_CHUNK");
  const long long n = v ? std::atoll(v) : 0;
  if (n <= 0) return (size_t)32 << 20;
  return ((size_t)n + 4095) & ~(size_t)4095;
}

}  // namespace qws
