// getitem.hip -- jagged and range slicing, local index and offsets kernels.
//
// Replaces awkward-cpp/src/cpu-kernels/:
//   awkward_localindex.cpp, awkward_ListArray_localindex.cpp,
//   awkward_RegularArray_getitem_jagged_expand.cpp,
//   awkward_ListArray_getitem_jagged_carrylen.cpp,
//   awkward_ListArray_getitem_jagged_apply.cpp,
//   awkward_ListArray_getitem_jagged_descend.cpp,
//   awkward_ListArray_getitem_next_range{,_carrylength,_counts}.cpp,
//   awkward_RegularArray_getitem_next_range.cpp,
//   awkward_ListArray_getitem_next_at.cpp,
//   awkward_ListArray_compact_offsets.cpp,
//   awkward_IndexedArray_getitem_nextcarry.cpp
//
// Everything the CPU kernels compute with a running counter `k` is an
// exclusive scan here (scan.h); per-list work runs one thread per list,
// per-element work is coalesced over elements with the owning list found by
// a binary search in the workgroup's LDS copy of the offsets.

#include <cstdio>
#include <mutex>

#include "errors.h"
#include "scan.h"

namespace akhip {

namespace {
const char* kFile = "awkward_amd/csrc/getitem.hip";

// ---- helpers ----------------------------------------------------------------
__global__ void fill_i64(int64_t* p, int64_t v) { *p = v; }

template <typename C>
__global__ void fill_c(C* p, C v) { *p = v; }

// Adds Body::value(i) over i < n into *out (which the caller zeroed).
template <class Body, int BLOCK>
__global__ __launch_bounds__(BLOCK) void sum_scalar_kernel(Body body, int64_t n, int64_t* out) {
  __shared__ int64_t s_w[BLOCK / kWave];
  int64_t s = 0;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK)
    s += body.value(i);
  s = wave_sum(s);
  if ((threadIdx.x & (kWave - 1)) == 0) s_w[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t t = 0;
    for (int w = 0; w < BLOCK / kWave; w++) t += s_w[w];
    if (t != 0) atomicAdd((unsigned long long*)out, (unsigned long long)t);
  }
}

template <class Body>
Error sum_to_scalar(const Body& body, int64_t n, int64_t* out) {
  hipStream_t st = stream();
  fill_i64<<<1, 1, 0, st>>>(out, 0);
  if (n > 0) sum_scalar_kernel<Body, 256><<<grid_for(n, 256, 8), 256, 0, st>>>(body, n, out);
  AKHIP_LAUNCH_CHECK("sum_scalar_kernel");
  return success();
}

// awkward_regularize_rangeslice (kernel-utils.cpp:7-39) and the number of
// positions the CPU loops visit, in integer arithmetic.
__device__ inline int64_t range_count(int64_t length, int64_t start, int64_t stop, int64_t step, int64_t* first) {
  const bool hasstart = start != kSliceNone, hasstop = stop != kSliceNone;
  int64_t a = start, b = stop;
  if (step > 0) {
    if (!hasstart) a = 0;
    else if (a < 0) a += length;
    if (a < 0) a = 0;
    if (a > length) a = length;
    if (!hasstop) b = length;
    else if (b < 0) b += length;
    if (b < 0) b = 0;
    if (b > length) b = length;
    if (b < a) b = a;
    *first = a;
    return (b > a) ? 1 + (b - a - 1) / step : 0;
  } else {
    if (!hasstart) a = length - 1;
    else if (a < 0) a += length;
    if (a < -1) a = -1;
    if (a > length - 1) a = length - 1;
    if (!hasstop) b = -1;
    else if (b < 0) b += length;
    if (b < -1) b = -1;
    if (b > length - 1) b = length - 1;
    if (b > a) b = a;
    *first = a;
    // step == 0 takes this branch too (step > 0 is false), where the CPU
    // loop visits nothing when a <= b and never ends otherwise: count 0.
    if (step == 0) return 0;
    const int64_t s = -step;
    return (a > b) ? 1 + (a - b - 1) / s : 0;
  }
}

// ---- local index ---------------------------------------------------------------
__global__ void iota_kernel(int64_t* __restrict__ out, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = i;
}

// Owning list of element k among s_off[0..nl]: s_off[lo] <= k < s_off[lo+1].
__device__ inline int find_list(const int64_t* s_off, int nl, int64_t k) {
  int lo = 0, hi = nl;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (s_off[mid] <= k) lo = mid;
    else hi = mid;
  }
  return lo;
}

template <typename C, int BLOCK>
__global__ __launch_bounds__(BLOCK) void list_localindex_kernel(int64_t* __restrict__ toindex,
                                                                const C* __restrict__ offsets, int64_t length) {
  __shared__ int64_t s_off[BLOCK + 1];
  const int64_t L0 = (int64_t)blockIdx.x * BLOCK;
  const int nl = (int)((length - L0) < BLOCK ? (length - L0) : BLOCK);
  for (int t = threadIdx.x; t <= nl; t += BLOCK) s_off[t] = (int64_t)offsets[L0 + t];
  __syncthreads();
  const int64_t E0 = s_off[0], E1 = s_off[nl];
  for (int64_t k = E0 + threadIdx.x; k < E1; k += BLOCK) {
    const int l = find_list(s_off, nl, k);
    toindex[k] = k - s_off[l];
  }
}

// ---- jagged getitem -----------------------------------------------------------------
__global__ void jagged_expand_kernel(int64_t* __restrict__ multistarts, int64_t* __restrict__ multistops,
                                     const int64_t* __restrict__ singleoffsets, int64_t size, int64_t total) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = k % size;
    multistarts[k] = singleoffsets[j];
    multistops[k] = singleoffsets[j + 1];
  }
}

struct CarrylenBody {
  const int64_t* starts;
  const int64_t* stops;
  __device__ int64_t value(int64_t i) const { return stops[i] - starts[i]; }
};

const char* const kJaggedApplyMsgs[] = {
    nullptr,
    "jagged slice's stops[i] < starts[i]",
    "jagged slice's offsets extend beyond its content",
    "stops[i] < starts[i]",
    "stops[i] > len(content)",
    "index out of range",
};

template <typename C>
struct JaggedApplyBody {
  int64_t* tooffsets;
  int64_t* tocarry;
  const int64_t* slicestarts;
  const int64_t* slicestops;
  const int64_t* sliceindex;
  int64_t sliceinnerlen;
  const C* fromstarts;
  const C* fromstops;
  int64_t contentlen;
  int64_t n;
  uint64_t* err;
  uint32_t inv;
  __device__ int64_t count(int64_t i) const {
    const int64_t s0 = slicestarts[i], s1 = slicestops[i];
    if (s0 == s1) return 0;
    int code = 0;
    if (s1 < s0) code = 1;
    else if (s1 > sliceinnerlen) code = 2;
    else {
      const int64_t start = (int64_t)fromstarts[i], stop = (int64_t)fromstops[i];
      if (stop < start) code = 3;
      else if (start != stop && stop > contentlen) code = 4;
    }
    if (code) {
      record_error(err, inv, i, code);
      return 0;
    }
    return s1 - s0;
  }
  __device__ void emit(int64_t i, int64_t excl, int64_t cnt) const {
    tooffsets[i] = excl;
    if (i == n - 1) tooffsets[n] = excl + cnt;
    if (cnt == 0) return;
    const int64_t s0 = slicestarts[i];
    const int64_t start = (int64_t)fromstarts[i];
    const int64_t count = (int64_t)fromstops[i] - start;
    for (int64_t t = 0; t < cnt; t++) {
      int64_t index = sliceindex[s0 + t];
      if (index < -count || index >= count) {
        record_error(err, inv, i, 5);
        return;
      }
      if (index < 0) index += count;
      tocarry[excl + t] = start + index;
    }
  }
};

template <typename C>
struct DescendBody {
  int64_t* tooffsets;
  const int64_t* slicestarts;
  const int64_t* slicestops;
  const C* fromstarts;
  const C* fromstops;
  uint64_t* err;
  uint32_t inv;
  __device__ int64_t count(int64_t i) const {
    const int64_t slicecount = slicestops[i] - slicestarts[i];
    const int64_t cnt = (int64_t)fromstops[i] - (int64_t)fromstarts[i];
    if (slicecount != cnt) {
      record_error(err, inv, i, 1);
      return 0;
    }
    return cnt < 0 ? 0 : cnt;
  }
  __device__ void emit(int64_t i, int64_t excl, int64_t cnt) const {
    const int64_t base = slicestarts[0];
    if (i == 0) tooffsets[0] = base;
    tooffsets[i + 1] = base + excl + cnt;
  }
};
const char* const kDescendMsgs[] = {nullptr, "jagged slice inner length differs from array inner length"};

// ---- range getitem --------------------------------------------------------------------
template <typename C>
struct RangeCountBody {
  const C* fromstarts;
  const C* fromstops;
  int64_t start, stop, step;
  __device__ int64_t value(int64_t i) const {
    int64_t first;
    return range_count((int64_t)fromstops[i] - (int64_t)fromstarts[i], start, stop, step, &first);
  }
};

template <typename C>
struct RangeBody {
  C* tooffsets;
  int64_t* tocarry;
  const C* fromstarts;
  const C* fromstops;
  int64_t start, stop, step;
  __device__ int64_t count(int64_t i) const {
    int64_t first;
    return range_count((int64_t)fromstops[i] - (int64_t)fromstarts[i], start, stop, step, &first);
  }
  __device__ void emit(int64_t i, int64_t excl, int64_t cnt) const {
    if (i == 0) tooffsets[0] = (C)0;
    tooffsets[i + 1] = (C)(excl + cnt);
    if (cnt == 0) return;
    int64_t first;
    const int64_t fs = (int64_t)fromstarts[i];
    range_count((int64_t)fromstops[i] - fs, start, stop, step, &first);
    for (int64_t t = 0; t < cnt; t++) tocarry[excl + t] = fs + first + t * step;
  }
};

template <typename C>
__global__ void range_counts_kernel(int64_t* total, const C* fromoffsets, int64_t lenstarts) {
  *total = (int64_t)fromoffsets[lenstarts] - (int64_t)fromoffsets[0];
}

__global__ void regular_range_kernel(int64_t* __restrict__ tocarry, int64_t regular_start, int64_t step,
                                     int64_t size, int64_t nextsize, int64_t total) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = k / nextsize, j = k - i * nextsize;
    tocarry[k] = i * size + regular_start + j * step;
  }
}

template <typename C>
__global__ void next_at_kernel(int64_t* __restrict__ tocarry, const C* __restrict__ fromstarts,
                               const C* __restrict__ fromstops, int64_t lenstarts, int64_t at, uint64_t* err,
                               uint32_t inv) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < lenstarts; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t length = (int64_t)fromstops[i] - (int64_t)fromstarts[i];
    int64_t r = at < 0 ? at + length : at;
    if (!(0 <= r && r < length)) {
      record_error(err, inv, i, 1);
      continue;
    }
    tocarry[i] = (int64_t)fromstarts[i] + r;
  }
}
const char* const kIndexMsgs[] = {nullptr, "index out of range"};

// ---- offsets --------------------------------------------------------------------------
template <typename C>
struct CompactBody {
  int64_t* tooffsets;
  const C* fromstarts;
  const C* fromstops;
  uint64_t* err;
  uint32_t inv;
  __device__ int64_t count(int64_t i) const {
    const C a = fromstarts[i], b = fromstops[i];
    if (b < a) {
      record_error(err, inv, i, 1);
      return 0;
    }
    return (int64_t)(C)(b - a);
  }
  __device__ void emit(int64_t i, int64_t excl, int64_t cnt) const {
    if (i == 0) tooffsets[0] = 0;
    tooffsets[i + 1] = excl + cnt;
  }
};
const char* const kCompactMsgs[] = {nullptr, "stops[i] < starts[i]"};

template <typename C>
__global__ void nextcarry_kernel(int64_t* __restrict__ tocarry, const C* __restrict__ fromindex, int64_t lenindex,
                                 int64_t lencontent, uint64_t* err, uint32_t inv) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < lenindex; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = (int64_t)fromindex[i];
    if (j < 0 || j >= lencontent) {
      record_error(err, inv, i, 1);
      continue;
    }
    tocarry[i] = j;
  }
}

template <class Body>
Error run_scan(const Body& body, int64_t n, uint32_t inv) {
  void* s = scratch(scan_scratch_bytes(scan_tiles(n)));
  if (!s) return failure("out of device memory for scan scratch", kSliceNone, kSliceNone, kFile);
  return launch_lookback_scan(body, n, inv, s);
}

}  // namespace

// ---- entry-point bodies (templated over the ListArray index type C) ------------------
template <typename C>
Error list_localindex(int64_t* toindex, const C* offsets, int64_t length) {
  if (length <= 0) return success();
  list_localindex_kernel<C, 256><<<(unsigned)ceil_div(length, 256), 256, 0, stream()>>>(toindex, offsets, length);
  AKHIP_LAUNCH_CHECK("list_localindex_kernel");
  return success();
}

template <typename C>
Error jagged_apply(int64_t* tooffsets, int64_t* tocarry, const int64_t* slicestarts, const int64_t* slicestops,
                   int64_t sliceouterlen, const int64_t* sliceindex, int64_t sliceinnerlen, const C* fromstarts,
                   const C* fromstops, int64_t contentlen) {
  if (sliceouterlen <= 0) {
    fill_i64<<<1, 1, 0, stream()>>>(tooffsets, 0);
    AKHIP_LAUNCH_CHECK("fill_i64");
    return success();
  }
  AKHIP_ERROR_WORD(ew);
  const uint32_t inv = next_invocation("awkward_ListArray_getitem_jagged_apply", kJaggedApplyMsgs);
  JaggedApplyBody<C> body{tooffsets, tocarry, slicestarts, slicestops, sliceindex, sliceinnerlen,
                          fromstarts, fromstops, contentlen, sliceouterlen, ew, inv};
  Error e = run_scan(body, sliceouterlen, inv);
  if (e.str) return e;
  return finish_errors(kJaggedApplyMsgs, kFile, [&](int64_t i, int code) -> int64_t {
    if (code == 2) return read_device(slicestops, i);
    if (code != 5) return kSliceNone;
    const int64_t s0 = read_device(slicestarts, i), s1 = read_device(slicestops, i);
    const int64_t count = (int64_t)read_device(fromstops, i) - (int64_t)read_device(fromstarts, i);
    for (int64_t j = s0; j < s1; j++) {
      const int64_t index = read_device(sliceindex, j);
      if (index < -count || index >= count) return index;
    }
    return kSliceNone;
  });
}

template <typename C>
Error jagged_descend(int64_t* tooffsets, const int64_t* slicestarts, const int64_t* slicestops, int64_t sliceouterlen,
                     const C* fromstarts, const C* fromstops) {
  if (sliceouterlen <= 0) {
    fill_i64<<<1, 1, 0, stream()>>>(tooffsets, 0);
    AKHIP_LAUNCH_CHECK("fill_i64");
    return success();
  }
  AKHIP_ERROR_WORD(ew);
  const uint32_t inv = next_invocation("awkward_ListArray_getitem_jagged_descend", kDescendMsgs);
  DescendBody<C> body{tooffsets, slicestarts, slicestops, fromstarts, fromstops, ew, inv};
  Error e = run_scan(body, sliceouterlen, inv);
  if (e.str) return e;
  return finish_errors(kDescendMsgs, kFile, no_attempt);
}

template <typename C>
Error next_range(C* tooffsets, int64_t* tocarry, const C* fromstarts, const C* fromstops, int64_t lenstarts,
                 int64_t start, int64_t stop, int64_t step) {
  if (lenstarts <= 0) {
    fill_c<C><<<1, 1, 0, stream()>>>(tooffsets, (C)0);
    AKHIP_LAUNCH_CHECK("fill_c");
    return success();
  }
  RangeBody<C> body{tooffsets, tocarry, fromstarts, fromstops, start, stop, step};
  return run_scan(body, lenstarts, 0);
}

template <typename C>
Error next_range_carrylength(int64_t* carrylength, const C* fromstarts, const C* fromstops, int64_t lenstarts,
                             int64_t start, int64_t stop, int64_t step) {
  return sum_to_scalar(RangeCountBody<C>{fromstarts, fromstops, start, stop, step}, lenstarts, carrylength);
}

template <typename C>
Error next_range_counts(int64_t* total, const C* fromoffsets, int64_t lenstarts) {
  range_counts_kernel<C><<<1, 1, 0, stream()>>>(total, fromoffsets, lenstarts < 0 ? 0 : lenstarts);
  AKHIP_LAUNCH_CHECK("range_counts_kernel");
  return success();
}

template <typename C>
Error next_at(int64_t* tocarry, const C* fromstarts, const C* fromstops, int64_t lenstarts, int64_t at) {
  if (lenstarts <= 0) return success();
  AKHIP_ERROR_WORD(ew);
  const uint32_t inv = next_invocation("awkward_ListArray_getitem_next_at", kIndexMsgs);
  next_at_kernel<C><<<grid_for(lenstarts, 256), 256, 0, stream()>>>(tocarry, fromstarts, fromstops, lenstarts, at,
                                                                    ew, inv);
  AKHIP_LAUNCH_CHECK("next_at_kernel");
  return finish_errors(kIndexMsgs, kFile, [&](int64_t, int) -> int64_t { return at; });
}

template <typename C>
Error compact_offsets(int64_t* tooffsets, const C* fromstarts, const C* fromstops, int64_t length) {
  if (length <= 0) {
    fill_i64<<<1, 1, 0, stream()>>>(tooffsets, 0);
    AKHIP_LAUNCH_CHECK("fill_i64");
    return success();
  }
  AKHIP_ERROR_WORD(ew);
  const uint32_t inv = next_invocation("awkward_ListArray_compact_offsets", kCompactMsgs);
  CompactBody<C> body{tooffsets, fromstarts, fromstops, ew, inv};
  Error e = run_scan(body, length, inv);
  if (e.str) return e;
  return finish_errors(kCompactMsgs, kFile, no_attempt);
}

template <typename C>
Error indexed_nextcarry(int64_t* tocarry, const C* fromindex, int64_t lenindex, int64_t lencontent) {
  if (lenindex <= 0) return success();
  AKHIP_ERROR_WORD(ew);
  const uint32_t inv = next_invocation("awkward_IndexedArray_getitem_nextcarry", kIndexMsgs);
  nextcarry_kernel<C><<<grid_for(lenindex, 256), 256, 0, stream()>>>(tocarry, fromindex, lenindex, lencontent,
                                                                     ew, inv);
  AKHIP_LAUNCH_CHECK("nextcarry_kernel");
  return finish_errors(kIndexMsgs, kFile,
                       [&](int64_t i, int) -> int64_t { return (int64_t)read_device(fromindex, i); });
}

}  // namespace akhip

// ---- next_range without allocation ---------------------------------------------------
// The look-back scan's tile counter and status words live in this file-scope
// device array instead of the stream's pool, for callers that cannot
// allocate (hipGraph capture, serving loops).  Each device holds its own copy
// (1 MiB) until the process ends.  Host code never passes the array's name to
// a kernel: its device address comes from hipGetSymbolAddress, once per
// device.  Layout: [0, 256) the scan's error word, [256, 512) the tile
// counter, then one 64-bit status word per tile.
namespace {
constexpr int64_t kNoAllocBytes = int64_t(1) << 20;
constexpr int64_t kNoAllocMaxTiles = (kNoAllocBytes - 512) / 8;
constexpr int64_t kNoAllocMaxLists = kNoAllocMaxTiles * akhip::kScanTile;
}  // namespace

static __device__ __align__(256) uint64_t scratch[kNoAllocBytes / 8];

namespace akhip {
namespace {
constexpr int kNoAllocDevices = 64;
std::mutex g_noalloc_mutex;
char* g_noalloc_ptr[kNoAllocDevices];

// The current device's address of `scratch` (nullptr on failure).
char* noalloc_scratch(hipError_t* err) {
  int d = 0;
  *err = hipGetDevice(&d);
  if (*err != hipSuccess || d < 0 || d >= kNoAllocDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_noalloc_mutex);
  if (g_noalloc_ptr[d] == nullptr) {
    void* p = nullptr;
    *err = hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch));
    if (*err != hipSuccess) return nullptr;
    g_noalloc_ptr[d] = (char*)p;
  }
  return g_noalloc_ptr[d];
}

__global__ void noalloc_prep_kernel(uint64_t* err, ScanScratch s, int64_t ntiles) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    *err = kNoError;
    *s.counter = 0u;
  }
  if (i < ntiles) s.status[i] = 0;
}
}  // namespace
}  // namespace akhip

using namespace akhip;

extern "C" {

// awkward_ListArray64_getitem_next_range_64 with the scan scratch in the
// file-scope array above: no hipMalloc, hipFree or host synchronisation.
// At most kNoAllocMaxLists (536,608,768) lists; calls that overlap on one
// device share the array and race.
Error awkward_hip_ListArray64_getitem_next_range_64_noalloc(int64_t* tooffsets, int64_t* tocarry,
                                                            const int64_t* fromstarts, const int64_t* fromstops,
                                                            int64_t lenstarts, int64_t start, int64_t stop,
                                                            int64_t step) {
  static thread_local char msg[160];
  if (lenstarts > kNoAllocMaxLists) {
    snprintf(msg, sizeof(msg), "lenstarts %lld exceeds the no-allocation scratch (at most %lld lists)",
             (long long)lenstarts, (long long)kNoAllocMaxLists);
    return failure(msg, kSliceNone, kSliceNone, kFile);
  }
  hipStream_t st = stream();
  if (lenstarts <= 0) {
    fill_c<int64_t><<<1, 1, 0, st>>>(tooffsets, (int64_t)0);
    AKHIP_LAUNCH_CHECK("fill_c");
    return success();
  }
  hipError_t herr = hipSuccess;
  char* base = noalloc_scratch(&herr);
  if (base == nullptr) return hip_failure(herr, "hipGetSymbolAddress of the no-allocation scratch");
  uint64_t* err = (uint64_t*)base;
  const ScanScratch s = scan_scratch_at(base + 256);
  const int64_t ntiles = scan_tiles(lenstarts);
  noalloc_prep_kernel<<<(unsigned)ceil_div(ntiles + 1, 256), 256, 0, st>>>(err, s, ntiles);
  RangeBody<int64_t> body{tooffsets, tocarry, fromstarts, fromstops, start, stop, step};
  lookback_scan_kernel<akhip::kScanBlock, akhip::kScanIpt, RangeBody<int64_t>><<<(unsigned)ntiles, akhip::kScanBlock, 0, st>>>(
      body, s, lenstarts, err, 0);
  AKHIP_LAUNCH_CHECK("lookback_scan_kernel (no allocation)");
  return success();
}

Error awkward_localindex_64(int64_t* toindex, int64_t length) {
  if (length <= 0) return success();
  iota_kernel<<<grid_for(length, 256), 256, 0, stream()>>>(toindex, length);
  AKHIP_LAUNCH_CHECK("iota_kernel");
  return success();
}

Error awkward_RegularArray_getitem_jagged_expand_64(int64_t* multistarts, int64_t* multistops,
                                                    const int64_t* singleoffsets, int64_t regularsize,
                                                    int64_t regularlength) {
  const int64_t total = regularsize * regularlength;
  if (total <= 0 || regularsize <= 0) return success();
  jagged_expand_kernel<<<grid_for(total, 256), 256, 0, stream()>>>(multistarts, multistops, singleoffsets,
                                                                   regularsize, total);
  AKHIP_LAUNCH_CHECK("jagged_expand_kernel");
  return success();
}

Error awkward_ListArray_getitem_jagged_carrylen_64(int64_t* carrylen, const int64_t* slicestarts,
                                                   const int64_t* slicestops, int64_t sliceouterlen) {
  return sum_to_scalar(CarrylenBody{slicestarts, slicestops}, sliceouterlen, carrylen);
}

Error awkward_RegularArray_getitem_next_range_64(int64_t* tocarry, int64_t regular_start, int64_t step, int64_t length,
                                                 int64_t size, int64_t nextsize) {
  const int64_t total = length * nextsize;
  if (total <= 0) return success();
  regular_range_kernel<<<grid_for(total, 256), 256, 0, stream()>>>(tocarry, regular_start, step, size, nextsize,
                                                                   total);
  AKHIP_LAUNCH_CHECK("regular_range_kernel");
  return success();
}

#define LISTARRAY(P, C)                                                                                       \
  Error awkward_ListArray##P##_localindex_64(int64_t* t, const C* o, int64_t n) {                             \
    return list_localindex<C>(t, o, n);                                                                       \
  }                                                                                                           \
  Error awkward_ListArray##P##_getitem_jagged_apply_64(int64_t* to, int64_t* tc, const int64_t* ss,           \
                                                       const int64_t* se, int64_t sol, const int64_t* si,     \
                                                       int64_t sil, const C* fs, const C* fe, int64_t cl) {   \
    return jagged_apply<C>(to, tc, ss, se, sol, si, sil, fs, fe, cl);                                         \
  }                                                                                                           \
  Error awkward_ListArray##P##_getitem_jagged_descend_64(int64_t* to, const int64_t* ss, const int64_t* se,   \
                                                         int64_t sol, const C* fs, const C* fe) {             \
    return jagged_descend<C>(to, ss, se, sol, fs, fe);                                                        \
  }                                                                                                           \
  Error awkward_ListArray##P##_getitem_next_range_64(C* to, int64_t* tc, const C* fs, const C* fe, int64_t n, \
                                                     int64_t a, int64_t b, int64_t s) {                       \
    return next_range<C>(to, tc, fs, fe, n, a, b, s);                                                         \
  }                                                                                                           \
  Error awkward_ListArray##P##_getitem_next_range_carrylength(int64_t* cl, const C* fs, const C* fe,         \
                                                              int64_t n, int64_t a, int64_t b, int64_t s) {   \
    return next_range_carrylength<C>(cl, fs, fe, n, a, b, s);                                                 \
  }                                                                                                           \
  Error awkward_ListArray##P##_getitem_next_range_counts_64(int64_t* t, const C* o, int64_t n) {              \
    return next_range_counts<C>(t, o, n);                                                                     \
  }                                                                                                           \
  Error awkward_ListArray##P##_getitem_next_at_64(int64_t* t, const C* fs, const C* fe, int64_t n,            \
                                                  int64_t at) {                                               \
    return next_at<C>(t, fs, fe, n, at);                                                                      \
  }                                                                                                           \
  Error awkward_ListArray##P##_compact_offsets_64(int64_t* t, const C* fs, const C* fe, int64_t n) {          \
    return compact_offsets<C>(t, fs, fe, n);                                                                  \
  }                                                                                                           \
  Error awkward_IndexedArray##P##_getitem_nextcarry_64(int64_t* t, const C* fi, int64_t li, int64_t lc) {     \
    return indexed_nextcarry<C>(t, fi, li, lc);                                                               \
  }
LISTARRAY(32, int32_t)
LISTARRAY(U32, uint32_t)
LISTARRAY(64, int64_t)
#undef LISTARRAY

}  // extern "C"
