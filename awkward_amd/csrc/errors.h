// errors.h -- turning the device error word into the CPU kernels' Error.
#pragma once

#include "common.h"
#include "scan.h"

namespace akhip {

// Host read of one device value (error path only).
template <typename T>
inline T read_device(const T* p, int64_t i) {
  T v{};
  (void)hipMemcpyAsync(&v, p + i, sizeof(T), hipMemcpyDeviceToHost, stream());
  (void)hipStreamSynchronize(stream());
  return v;
}

// After an error-capable launch: in synchronous mode wait, and rebuild the
// first error in CPU loop order.  `attempt(identity, code)` recomputes the
// kernel's `attempt` field from device memory (only called on failure).
template <class AttemptF>
inline Error finish_errors(const char* const* messages, const char* file, AttemptF attempt) {
  if (deferred_errors()) return success();
  const uint64_t key = take_error();
  if (key == kNoError) return success();
  const int code = error_code(key);
  const int64_t identity = error_identity(key);
  if (code == kScanTimeoutCode) return failure("offsets scan look-back timed out", identity, kSliceNone, file);
  return failure(messages[code], identity, attempt(identity, code), file);
}

inline int64_t no_attempt(int64_t, int) { return kSliceNone; }

}  // namespace akhip
