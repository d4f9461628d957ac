// common.h -- shared pieces of libawkward_hip.so (gfx950 / CDNA4).
//
// * struct Error and kSliceNone, as in awkward-cpp/include/awkward/common.h:39-78
// * the runtime state: current stream, error mode, and per (device, stream)
//   a device error word and grow-only scratch pools
// * small device helpers (wave reductions, launch geometry)
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstddef>

extern "C" {
struct Error {
  const char* str;
  const char* filename;
  int64_t identity;
  int64_t attempt;
};
}

namespace akhip {

constexpr int64_t kSliceNone = INT64_MAX;
constexpr int kWave = 64;
constexpr uint64_t kNoError = ~uint64_t(0);

inline Error success() { return Error{nullptr, nullptr, kSliceNone, kSliceNone}; }
inline Error failure(const char* msg, int64_t identity, int64_t attempt, const char* file) {
  return Error{msg, file, identity, attempt};
}

// ---- runtime (runtime.hip) ------------------------------------------------
// State is kept per (device, stream): the stream set by awkward_hip_set_stream
// on the calling thread, on the current device.
hipStream_t stream();
bool deferred_errors();
// Device error word of the current stream (uint64, kNoError when clear);
// nullptr if the stream's state could not be allocated.
uint64_t* error_word();
// Next invocation tag for an error-capable launch (fits 16 bits).
uint32_t next_invocation(const char* kernel, const char* const* messages);
// Synchronous mode: wait, read and clear the error word.  Returns kNoError
// or the packed key (invocation << 48 | identity << 8 | code).
uint64_t take_error();
// Scratch memory of at least `bytes` of the current stream, valid until the
// next call of scratch() on this stream (stream-ordered reuse).
void* scratch(size_t bytes);
// A second such pool, for work queues that live across the kernels of one
// entry point while its caller may hold scratch() memory.
void* queue_pool(size_t bytes);
// Error describing a failed HIP runtime call.
Error hip_failure(hipError_t err, const char* what);
// Number of compute units of the current device (cached).
int cu_count();

inline uint64_t pack_error(uint32_t invocation, int64_t identity, int code) {
  uint64_t id = identity < 0 ? 0 : (uint64_t)identity;
  if (id >= (uint64_t(1) << 40)) id = (uint64_t(1) << 40) - 1;
  return (uint64_t(invocation & 0xffff) << 48) | (id << 8) | uint64_t(code & 0xff);
}
inline int64_t error_identity(uint64_t key) { return (int64_t)((key >> 8) & ((uint64_t(1) << 40) - 1)); }
inline int error_code(uint64_t key) { return (int)(key & 0xff); }

__device__ inline void record_error(uint64_t* word, uint32_t invocation, int64_t identity, int code) {
  uint64_t id = identity < 0 ? 0 : (uint64_t)identity;
  if (id >= (uint64_t(1) << 40)) id = (uint64_t(1) << 40) - 1;
  const uint64_t key = (uint64_t(invocation & 0xffff) << 48) | (id << 8) | uint64_t(code & 0xff);
  atomicMin((unsigned long long*)word, (unsigned long long)key);
}

// Declares `var` = the current stream's error word, or fails the entry point
// when the stream's state cannot be allocated (no device memory, or a first
// use of the stream under graph capture).
#define AKHIP_ERROR_WORD(var)                                                                          \
  uint64_t* var = ::akhip::error_word();                                                               \
  if (var == nullptr)                                                                                  \
  return ::akhip::failure("no device state for this stream (out of device memory, or its first use "   \
                          "under graph capture)",                                                      \
                          ::akhip::kSliceNone, ::akhip::kSliceNone, "awkward_amd/csrc/runtime.hip")

#define AKHIP_LAUNCH_CHECK(what)                                  \
  do {                                                            \
    hipError_t _e = hipGetLastError();                            \
    if (_e != hipSuccess) return ::akhip::hip_failure(_e, what);  \
  } while (0)

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Grid for an elementwise launch over n items: enough blocks to fill the
// chip a few times over, each block striding.
inline unsigned grid_for(int64_t n, int block, int64_t per_thread = 4) {
  int64_t blocks = ceil_div(n, (int64_t)block * per_thread);
  const int64_t cap = (int64_t)cu_count() * 32;
  if (blocks > cap) blocks = cap;
  if (blocks < 1) blocks = 1;
  return (unsigned)blocks;
}

// ---- wave / block helpers ---------------------------------------------------
template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, kWave);
  return v;
}

// Inclusive scan across the 64 lanes of a wave (Hillis-Steele on shuffles).
template <typename T>
__device__ inline T wave_inclusive_scan(T v) {
  const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    T u = __shfl_up(v, d, kWave);
    if (lane >= d) v += u;
  }
  return v;
}

}  // namespace akhip
