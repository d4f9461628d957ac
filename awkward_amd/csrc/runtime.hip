// runtime.hip -- stream, error word and scratch pool of libawkward_hip.so.
//
// The reference CUDA backend keeps one error word per stream context and
// raises lazily (src/awkward/_connect/cuda/__init__.py:28-40, 257-294).  Here
// each device has one uint64 error word: error-capable kernels atomicMin a
// key (invocation << 48 | failing list << 8 | code), so the earliest failure
// in loop order of the earliest invocation wins, as in the CPU kernels.

#include <mutex>
#include <cstdio>
#include <cstring>
#include <vector>

#include "common.h"

namespace akhip {
namespace {

constexpr int kMaxDevices = 64;

struct DeviceState {
  uint64_t* error_word = nullptr;  // device
  uint64_t* host_word = nullptr;   // pinned host mirror
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  void* queue = nullptr;  // second pool: work queues of multi-kernel sequences
  size_t queue_bytes = 0;
  int cus = 0;
};

DeviceState g_dev[kMaxDevices];
std::mutex g_mutex;
thread_local hipStream_t t_stream = nullptr;
bool g_deferred = false;

struct Invocation {
  const char* kernel;
  const char* const* messages;
};
std::vector<Invocation> g_invocations;  // index = invocation tag (deferred mode)

int current_device() {
  int d = 0;
  (void)hipGetDevice(&d);
  return d;
}

__global__ void reset_word(uint64_t* w) { *w = kNoError; }

DeviceState& state() {
  const int d = current_device();
  DeviceState& s = g_dev[d];
  if (s.error_word == nullptr) {
    std::lock_guard<std::mutex> lock(g_mutex);
    if (s.error_word == nullptr) {
      uint64_t* w = nullptr;
      if (hipMalloc(&w, sizeof(uint64_t)) == hipSuccess) {
        (void)hipMemset(w, 0xff, sizeof(uint64_t));
        (void)hipHostMalloc((void**)&s.host_word, sizeof(uint64_t), hipHostMallocDefault);
        (void)hipDeviceSynchronize();
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, d) == hipSuccess) s.cus = prop.multiProcessorCount;
        s.error_word = w;
      }
    }
  }
  return s;
}

}  // namespace

hipStream_t stream() { return t_stream; }
bool deferred_errors() { return g_deferred; }
uint64_t* error_word() { return state().error_word; }

int cu_count() {
  int c = state().cus;
  return c > 0 ? c : 256;
}

uint32_t next_invocation(const char* kernel, const char* const* messages) {
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_deferred) {
    g_invocations.clear();
  }
  g_invocations.push_back(Invocation{kernel, messages});
  return (uint32_t)((g_invocations.size() - 1) & 0xffff);
}

uint64_t take_error() {
  DeviceState& s = state();
  hipStream_t st = t_stream;
  if (hipMemcpyAsync(s.host_word, s.error_word, sizeof(uint64_t), hipMemcpyDeviceToHost, st) != hipSuccess)
    return kNoError;
  if (hipStreamSynchronize(st) != hipSuccess) return kNoError;
  const uint64_t key = *s.host_word;
  if (key != kNoError) {
    reset_word<<<1, 1, 0, st>>>(s.error_word);
    (void)hipStreamSynchronize(st);
  }
  return key;
}

// Grow-only device buffer: at least `bytes`, contents not preserved on growth.
void* grow(void** buf, size_t* have, size_t bytes) {
  if (bytes < 256) bytes = 256;
  if (*have < bytes) {
    std::lock_guard<std::mutex> lock(g_mutex);
    if (*have < bytes) {
      size_t want = *have ? *have : (size_t)1 << 20;
      while (want < bytes) want *= 2;
      if (*buf) {
        (void)hipDeviceSynchronize();  // in-flight kernels may still use it
        (void)hipFree(*buf);
        *buf = nullptr;
        *have = 0;
      }
      void* p = nullptr;
      if (hipMalloc(&p, want) != hipSuccess) return nullptr;
      *buf = p;
      *have = want;
    }
  }
  return *buf;
}

void* scratch(size_t bytes) {
  DeviceState& s = state();
  return grow(&s.scratch, &s.scratch_bytes, bytes);
}

void* queue_pool(size_t bytes) {
  DeviceState& s = state();
  return grow(&s.queue, &s.queue_bytes, bytes);
}

Error hip_failure(hipError_t err, const char* what) {
  static thread_local char buf[256];
  snprintf(buf, sizeof(buf), "HIP error in %s: %s", what, hipGetErrorString(err));
  return Error{buf, "awkward_amd/csrc/runtime.hip", kSliceNone, kSliceNone};
}

}  // namespace akhip

using namespace akhip;

extern "C" {

void awkward_hip_set_stream(void* s) { t_stream = (hipStream_t)s; }
void* awkward_hip_get_stream(void) { return (void*)t_stream; }

void awkward_hip_set_deferred_errors(int deferred) {
  std::lock_guard<std::mutex> lock(g_mutex);
  g_deferred = deferred != 0;
  g_invocations.clear();
}

Error awkward_hip_error_poll(void) {
  const uint64_t key = take_error();
  Error out = success();
  if (key != kNoError) {
    std::lock_guard<std::mutex> lock(g_mutex);
    const uint32_t inv = (uint32_t)(key >> 48);
    const int code = error_code(key);
    if (inv < g_invocations.size() && g_invocations[inv].messages != nullptr) {
      out = Error{g_invocations[inv].messages[code], g_invocations[inv].kernel, error_identity(key), kSliceNone};
    } else {
      out = Error{"kernel error (unknown invocation)", "awkward_amd/csrc/runtime.hip", error_identity(key), kSliceNone};
    }
  }
  {
    std::lock_guard<std::mutex> lock(g_mutex);
    g_invocations.clear();
  }
  return out;
}

const char* awkward_hip_version(void) { return "awkward_hip 0.1 (gfx950)"; }

// Size the scratch pool ahead of time (e.g. before capturing a hipGraph, in
// which no allocation may happen).  Returns 0 on success.
int awkward_hip_reserve_scratch(int64_t bytes) {
  return scratch((size_t)bytes) != nullptr ? 0 : 1;
}

// The same for the work-queue pool (the batch reductions' long-list queue).
int awkward_hip_reserve_queue(int64_t bytes) {
  return queue_pool((size_t)bytes) != nullptr ? 0 : 1;
}

}  // extern "C"
