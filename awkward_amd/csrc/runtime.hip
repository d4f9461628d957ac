// runtime.hip -- streams, error words and scratch pools of libawkward_hip.so.
//
// The reference CUDA backend keeps its error state per stream
// (src/awkward/_connect/cuda/__init__.py:28 cuda_streamptr_to_contexts, and
// :257-294 for the lazy check).  Here every (device, stream) pair that calls
// the library owns a StreamState:
//
//   * a uint64 device error word: error-capable kernels atomicMin a key
//     (invocation << 48 | failing list << 8 | code), so the earliest failure
//     in loop order of the earliest invocation wins, as in the CPU kernels;
//     only kernels of this stream write it, and it is read and cleared on
//     this stream, so streams never see each other's errors;
//   * the invocation table of deferred mode (tag -> kernel and messages),
//     bounded: at 65,536 entries the word is read once (one stream
//     synchronisation), a recorded error is kept for the next poll, and the
//     table restarts, so tags never alias;
//   * two grow-only device pools (scratch, and work queues that must live
//     across the kernels of one entry point).  A pool that grows retires its
//     old buffer instead of freeing it: kernels in flight and captured
//     hipGraphs may still use it, so a retired buffer stays allocated until
//     the process ends, or until the caller, knowing no graph uses them any
//     more, releases the pools (awkward_hip_release_pools).  Pools double up to
//     1 GiB, so the retired buffers of a pool add up to less than its
//     current size; beyond, they grow to the request.  Nothing on the call
//     path synchronises the device or frees memory.
//
// Scratch is stream-ordered: the kernels of one stream use its pools one
// after another, so two streams never share scratch.  A hipGraph captured on
// a stream uses that stream's pools; replaying it concurrently with other
// work of the same stream would race (replay it on its own stream, or
// serialise).  The first call on a stream allocates its state (an 8-byte
// error word and its pinned mirror): make it before capturing on that
// stream (awkward_hip_reserve_scratch does).

#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "common.h"

namespace akhip {
namespace {

constexpr int kMaxDevices = 64;
constexpr size_t kMaxInvocations = 65536;  // the 16-bit tag of pack_error

struct Pool {
  void* buf = nullptr;
  size_t bytes = 0;
  std::vector<void*> retired;
};

struct Invocation {
  const char* kernel;
  const char* const* messages;
};

struct StreamState {
  uint64_t* error_word = nullptr;  // device
  uint64_t* host_word = nullptr;   // pinned host mirror
  Pool scratch;
  Pool queue;
  std::vector<Invocation> invocations;  // deferred mode: index = invocation tag
  Error pending{nullptr, nullptr, kSliceNone, kSliceNone};  // an error found when the table wrapped
};

int g_cus[kMaxDevices];
std::mutex g_mutex;
std::map<std::pair<int, hipStream_t>, StreamState*> g_states;
thread_local hipStream_t t_stream = nullptr;
thread_local StreamState* t_cached = nullptr;
thread_local hipStream_t t_cached_stream = nullptr;
thread_local int t_cached_device = -1;
bool g_deferred = false;

int current_device() {
  int d = 0;
  (void)hipGetDevice(&d);
  return d;
}

__global__ void reset_word(uint64_t* w) { *w = kNoError; }

// The state of (current device, current stream); nullptr if its error word
// cannot be allocated.
StreamState* state() {
  const int d = current_device();
  if (t_cached != nullptr && t_cached_device == d && t_cached_stream == t_stream) return t_cached;
  std::lock_guard<std::mutex> lock(g_mutex);
  StreamState*& s = g_states[std::make_pair(d, t_stream)];
  if (s == nullptr) {
    uint64_t* w = nullptr;
    uint64_t* h = nullptr;
    if (hipMalloc(&w, sizeof(uint64_t)) != hipSuccess) return nullptr;
    if (hipHostMalloc((void**)&h, sizeof(uint64_t), hipHostMallocDefault) != hipSuccess) {
      (void)hipFree(w);
      return nullptr;
    }
    reset_word<<<1, 1, 0, t_stream>>>(w);
    if (hipGetLastError() != hipSuccess) return nullptr;
    if (d >= 0 && d < kMaxDevices && g_cus[d] == 0) {
      hipDeviceProp_t prop;
      if (hipGetDeviceProperties(&prop, d) == hipSuccess) g_cus[d] = prop.multiProcessorCount;
    }
    s = new StreamState();
    s->error_word = w;
    s->host_word = h;
  }
  t_cached = s;
  t_cached_stream = t_stream;
  t_cached_device = d;
  return s;
}

// Reads and clears the stream's error word (synchronises the stream).
uint64_t read_and_clear(StreamState* s) {
  hipStream_t st = t_stream;
  if (hipMemcpyAsync(s->host_word, s->error_word, sizeof(uint64_t), hipMemcpyDeviceToHost, st) != hipSuccess)
    return kNoError;
  if (hipStreamSynchronize(st) != hipSuccess) return kNoError;
  const uint64_t key = *s->host_word;
  if (key != kNoError) reset_word<<<1, 1, 0, st>>>(s->error_word);
  return key;
}

Error describe(const StreamState* s, uint64_t key) {
  const uint32_t inv = (uint32_t)(key >> 48);
  const int code = error_code(key);
  if (inv < s->invocations.size() && s->invocations[inv].messages != nullptr)
    return Error{s->invocations[inv].messages[code], s->invocations[inv].kernel, error_identity(key), kSliceNone};
  return Error{"kernel error (unknown invocation)", "awkward_amd/csrc/runtime.hip", error_identity(key), kSliceNone};
}

// Grow-only device buffer: at least `bytes`, contents not preserved on
// growth; the old buffer is retired, not freed.  Up to 1 GiB the size
// doubles; beyond, it is the request rounded up to 64 MiB (doubling a
// 100 GB request would hold 128 GiB, plus the retired buffers).
void* grow(Pool* pool, size_t bytes) {
  if (bytes < 256) bytes = 256;
  if (pool->bytes < bytes) {
    size_t want = pool->bytes ? pool->bytes : (size_t)1 << 20;
    while (want < bytes && want < ((size_t)1 << 30)) want *= 2;
    if (want < bytes) want = (bytes + ((size_t)64 << 20) - 1) / ((size_t)64 << 20) * ((size_t)64 << 20);
    void* p = nullptr;
    if (hipMalloc(&p, want) != hipSuccess) return nullptr;
    if (pool->buf) pool->retired.push_back(pool->buf);
    pool->buf = p;
    pool->bytes = want;
  }
  return pool->buf;
}

}  // namespace

hipStream_t stream() { return t_stream; }
bool deferred_errors() { return g_deferred; }

uint64_t* error_word() {
  StreamState* s = state();
  return s ? s->error_word : nullptr;
}

int cu_count() {
  const int d = current_device();
  int c = (d >= 0 && d < kMaxDevices) ? g_cus[d] : 0;
  if (c == 0) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, d) == hipSuccess) c = prop.multiProcessorCount;
    if (d >= 0 && d < kMaxDevices) g_cus[d] = c;
  }
  return c > 0 ? c : 256;
}

uint32_t next_invocation(const char* kernel, const char* const* messages) {
  StreamState* s = state();
  if (s == nullptr) return 0;
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_deferred) {
    s->invocations.clear();
  } else if (s->invocations.size() >= kMaxInvocations) {
    // the tag space is used up: settle the errors recorded so far
    const uint64_t key = read_and_clear(s);
    if (key != kNoError && s->pending.str == nullptr) s->pending = describe(s, key);
    s->invocations.clear();
  }
  s->invocations.push_back(Invocation{kernel, messages});
  return (uint32_t)((s->invocations.size() - 1) & 0xffff);
}

uint64_t take_error() {
  StreamState* s = state();
  if (s == nullptr) return kNoError;
  return read_and_clear(s);
}

void* scratch(size_t bytes) {
  StreamState* s = state();
  if (s == nullptr) return nullptr;
  std::lock_guard<std::mutex> lock(g_mutex);
  return grow(&s->scratch, bytes);
}

void* queue_pool(size_t bytes) {
  StreamState* s = state();
  if (s == nullptr) return nullptr;
  std::lock_guard<std::mutex> lock(g_mutex);
  return grow(&s->queue, bytes);
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
  for (auto& kv : g_states) kv.second->invocations.clear();
}

Error awkward_hip_error_poll(void) {
  StreamState* s = state();
  if (s == nullptr) return Error{"no device state for this stream", "awkward_amd/csrc/runtime.hip", kSliceNone,
                                 kSliceNone};
  const uint64_t key = read_and_clear(s);
  std::lock_guard<std::mutex> lock(g_mutex);
  Error out = success();
  if (s->pending.str != nullptr) {
    out = s->pending;
    s->pending = success();
  } else if (key != kNoError) {
    out = describe(s, key);
  }
  s->invocations.clear();
  return out;
}

const char* awkward_hip_version(void) { return "awkward_hip 0.2 (gfx950)"; }

// Size the current stream's scratch pool ahead of time (e.g. before capturing
// a hipGraph on that stream, in which no allocation may happen); also creates
// the stream's state.  Returns 0 on success.
int awkward_hip_reserve_scratch(int64_t bytes) {
  return scratch((size_t)bytes) != nullptr ? 0 : 1;
}

// The same for the work-queue pool (the batch reductions' long-list queue).
int awkward_hip_reserve_queue(int64_t bytes) {
  return queue_pool((size_t)bytes) != nullptr ? 0 : 1;
}

// Frees the current stream's pool buffers, current and retired, after
// synchronising the stream; the next use allocates afresh.  Only for a
// caller that knows no captured hipGraph still uses them (e.g. between the
// waves of a batch larger than HBM, once the wave's graphs are gone).
// Returns 0 on success.
int awkward_hip_release_pools(void) {
  StreamState* s = state();
  if (s == nullptr) return 1;
  if (hipStreamSynchronize(t_stream) != hipSuccess) return 1;
  std::lock_guard<std::mutex> lock(g_mutex);
  int rc = 0;
  for (Pool* pool : {&s->scratch, &s->queue}) {
    for (void* b : pool->retired)
      if (hipFree(b) != hipSuccess) rc = 1;
    pool->retired.clear();
    if (pool->buf != nullptr && hipFree(pool->buf) != hipSuccess) rc = 1;
    pool->buf = nullptr;
    pool->bytes = 0;
  }
  return rc;
}

// Bytes held by the current stream's pools: current buffers and retired
// ones (diagnostics; out[0] scratch, out[1] queue, out[2] retired buffers).
void awkward_hip_pool_bytes(int64_t* out) {
  StreamState* s = state();
  out[0] = out[1] = out[2] = 0;
  if (s == nullptr) return;
  std::lock_guard<std::mutex> lock(g_mutex);
  out[0] = (int64_t)s->scratch.bytes;
  out[1] = (int64_t)s->queue.bytes;
  out[2] = (int64_t)(s->scratch.retired.size() + s->queue.retired.size());
}

}  // extern "C"
