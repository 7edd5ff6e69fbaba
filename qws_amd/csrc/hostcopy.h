// hostcopy.h -- host-side copies between caller memory and the pinned staging
// chunks of the library's host<->device transfers.
#pragma once

#include <cstddef>

namespace qws {

// Byte range [o, e) of part i (0 <= i < T) when n bytes are split into T parts of
// whole 64-byte lines; the last part takes whatever the others leave.
void copy_part(size_t n, int T, int i, size_t& o, size_t& e);

// memcpy of n bytes, split over 8 threads from 1 MiB on (one thread's memcpy is well
// below the host's memory bandwidth and the link's rate).  If a thread cannot be
// started, the calling thread copies the parts left over.
void par_memcpy(void* dst, const void* src, size_t n);

// Size of the two pinned staging chunks: 32 MiB, or QWS_MI355X_PIN_CHUNK bytes
// (rounded up to 4 KiB) when that is set to a positive number.
size_t pin_chunk_from_env();

}  // namespace qws
