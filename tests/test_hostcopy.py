"""The library's multi-threaded host copy (pageable <-> pinned staging of every
drop-in upload and download), built on its own with g++ and run on odd sizes:
every byte arrives, including the tail when n is not a multiple of the thread
count, and the parts tile [0, n) without overlap."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "qws_amd", "csrc", "hostcopy.cpp")

MAIN = r"""
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include "hostcopy.h"
int main() {
  const size_t sizes[] = {0, 1, 63, 4096, (1u << 20) - 1, 1u << 20, (1u << 20) + 5, (1u << 20) + 7,
                          (8u << 20) + 3, (32u << 20) + 4095, 96 * 100003};
  for (size_t n : sizes) {
    for (int T : {1, 3, 8}) {  // the parts tile [0, n) in order
      size_t prev = 0;
      for (int i = 0; i < T; ++i) {
        size_t o, e;
        qws::copy_part(n, T, i, o, e);
        if (o != prev || e < o || (i == T - 1 && e != n)) { std::printf("tiling n=%zu T=%d i=%d\n", n, T, i); return 1; }
        prev = e;
      }
    }
    std::vector<unsigned char> a(n + 16), b(n + 16, 0xAB);
    for (size_t i = 0; i < n + 16; ++i) a[i] = (unsigned char)(i * 131 + 7);
    qws::par_memcpy(b.data() + 3, a.data() + 5, n);
    if (n && std::memcmp(b.data() + 3, a.data() + 5, n) != 0) { std::printf("copy n=%zu\n", n); return 2; }
    for (size_t i = 0; i < 3; ++i) if (b[i] != 0xAB) { std::printf("underrun n=%zu\n", n); return 3; }
    for (size_t i = n + 3; i < n + 16; ++i) if (b[i] != 0xAB) { std::printf("overrun n=%zu\n", n); return 4; }
  }
  setenv("QWS_MI355X_PIN_CHUNK", "5000", 1);
  if (qws::pin_chunk_from_env() != 8192) return 5;
  unsetenv("QWS_MI355X_PIN_CHUNK");
  if (qws::pin_chunk_from_env() != (32u << 20)) return 6;
  std::printf("ok\n");
  return 0;
}
"""


def test_par_memcpy_odd_sizes(tmp_path):
    main = tmp_path / "main.cpp"
    main.write_text(MAIN)
    exe = tmp_path / "hostcopy_test"
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", "-I", os.path.dirname(SRC), str(main), SRC, "-o",
                    str(exe)], check=True)
    p = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0 and p.stdout.strip() == "ok", p.stdout + p.stderr
