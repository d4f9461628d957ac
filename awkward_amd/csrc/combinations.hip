// combinations.hip -- awkward_ListArray_combinations_length and
// awkward_ListArray_combinations on the device.
//
// Replaces awkward-cpp/src/cpu-kernels/awkward_ListArray_combinations_length.cpp
// and awkward_ListArray_combinations.cpp (+ the recursive
// awkward_ListArray_combinations_step of kernel-utils.cpp:41-79).
//
// The CPU enumerates each list's n-tuples recursively.  Here the tuples are
// decoded independently per output slot: a scan gives every list its first
// slot, each slot finds its list by binary search over those offsets and
// unranks its local rank in lexicographic order (combinatorial number system,
// one binary search per tuple position over closed-form counts).  Slots are
// spread over threads so that all n carries are written with coalesced
// stores, which is what bounds this kernel (8 * n bytes per tuple).

#include "errors.h"
#include "scan.h"

namespace akhip {
namespace {

const char* kFile = "awkward_amd/csrc/combinations.hip";

// C(m, k) by the CPU kernel's incremental multiply-then-divide
// (awkward_ListArray_combinations_length.cpp:16-43), for m >= 0.
__host__ __device__ inline int64_t choose(int64_t m, int64_t k) {
  if (k < 0 || k > m) return 0;
  if (k == m || k == 0) return 1;
  if (k * 2 > m) k = m - k;
  int64_t c = m;
  for (int64_t j = 2; j <= k; j++) {
    c *= (m - j + 1);
    c /= j;
  }
  return c;
}

// Number of n-combinations (or multisets with replacement) of `size` items.
__host__ __device__ inline int64_t comb_count(int64_t size, int64_t n, bool replacement) {
  if (replacement) size += n - 1;
  if (n > size) return 0;
  if (n == size) return 1;
  return choose(size, n);
}

// Tuples whose remaining kk positions are drawn from [x, s): C(s-x, kk), or
// the multiset count C(s-x+kk-1, kk) with replacement.
__device__ inline int64_t tail_count(int64_t s, int64_t x, int64_t kk, bool replacement) {
  const int64_t m = s - x;
  if (m <= 0) return 0;
  return replacement ? choose(m + kk - 1, kk) : choose(m, kk);
}

template <typename C>
struct LengthBody {
  int64_t* totallen;
  int64_t* tooffsets;
  const C* starts;
  const C* stops;
  int64_t n;
  bool replacement;
  int64_t length;
  __device__ int64_t count(int64_t i) const {
    return comb_count((int64_t)(stops[i] - starts[i]), n, replacement);
  }
  __device__ void emit(int64_t i, int64_t excl, int64_t cnt) const {
    if (i == 0) tooffsets[0] = 0;
    tooffsets[i + 1] = excl + cnt;
    if (i == length - 1) *totallen = excl + cnt;
  }
};

__global__ void zero_two(int64_t* a, int64_t* b) {
  *a = 0;
  *b = 0;
}

constexpr int kDecodeBlock = 256;
constexpr int kDecodeSlots = 4;  // slots per thread per tile
constexpr int kMaxN = 8;         // tuple sizes decoded on the device

// Persistent decode over P = offsets[length] slots (read on the device, so
// no host synchronisation is needed between the scan and the decode).
template <typename C>
__global__ __launch_bounds__(kDecodeBlock) void decode_kernel(int64_t* const* __restrict__ tocarry,
                                                              int64_t* __restrict__ toindex,
                                                              const int64_t* __restrict__ offsets,
                                                              const C* __restrict__ starts,
                                                              const C* __restrict__ stops, int64_t length,
                                                              int64_t n, bool replacement) {
  constexpr int64_t TILE = (int64_t)kDecodeBlock * kDecodeSlots;
  constexpr int kCache = 2048;  // offsets of the tile's lists kept in LDS
  __shared__ int64_t* s_out[kMaxN];
  __shared__ int64_t s_off[kCache + 1];
  __shared__ int64_t s_range[2];
  if (threadIdx.x < n) s_out[threadIdx.x] = tocarry[threadIdx.x];
  const int64_t P = offsets[length];
  if (blockIdx.x == 0 && threadIdx.x < n) toindex[threadIdx.x] = P;
  for (int64_t tile = blockIdx.x; tile * TILE < P; tile += gridDim.x) {
    const int64_t p0 = tile * TILE;
    const int64_t p1 = (p0 + TILE < P ? p0 + TILE : P) - 1;  // last slot of the tile
    if (threadIdx.x < 2) {
      const int64_t target = threadIdx.x == 0 ? p0 : p1;
      int64_t lo = 0, hi = length;  // offsets[lo] <= target < offsets[hi]
      while (hi - lo > 1) {
        const int64_t mid = (lo + hi) >> 1;
        if (offsets[mid] <= target) lo = mid;
        else hi = mid;
      }
      s_range[threadIdx.x] = lo;
    }
    __syncthreads();
    const int64_t l0 = s_range[0], l1 = s_range[1];
    const int64_t nl = l1 - l0 + 1;
    const bool cached = nl <= kCache;
    if (cached)
      for (int64_t t = threadIdx.x; t <= nl; t += kDecodeBlock) s_off[t] = offsets[l0 + t];
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kDecodeSlots; u++) {
      const int64_t p = p0 + (int64_t)u * kDecodeBlock + threadIdx.x;
      if (p > p1) break;
      // list with offsets[l] <= p < offsets[l+1] (lists without tuples repeat
      // an offset, so this is the last list starting at or before p)
      int64_t lo = 0, hi = nl, base;
      if (cached) {
        while (hi - lo > 1) {
          const int64_t mid = (lo + hi) >> 1;
          if (s_off[mid] <= p) lo = mid;
          else hi = mid;
        }
        base = s_off[lo];
      } else {
        while (hi - lo > 1) {
          const int64_t mid = (lo + hi) >> 1;
          if (offsets[l0 + mid] <= p) lo = mid;
          else hi = mid;
        }
        base = offsets[l0 + lo];
      }
      const int64_t l = l0 + lo;
      const int64_t start = (int64_t)starts[l];
      const int64_t s = (int64_t)stops[l] - start;
      int64_t r = p - base;
      int64_t prev = replacement ? 0 : -1;
      for (int64_t q = 0; q < n; q++) {
        const int64_t kk = n - q;
        const int64_t xlo = replacement ? prev : prev + 1;
        const int64_t xhi = replacement ? s - 1 : s - kk;
        const int64_t total = tail_count(s, xlo, kk, replacement);
        // largest x in [xlo, xhi] with total - tail_count(s, x, kk) <= r
        int64_t a = xlo, b = xhi;
        while (a < b) {
          const int64_t mid = (a + b + 1) >> 1;
          if (total - tail_count(s, mid, kk, replacement) <= r) a = mid;
          else b = mid - 1;
        }
        r -= total - tail_count(s, a, kk, replacement);
        s_out[q][p] = start + a;
        prev = a;
      }
    }
    __syncthreads();  // s_off and s_range are rewritten by the next tile
  }
}

template <typename C>
struct OffsetsBody {
  int64_t* offsets;
  const C* starts;
  const C* stops;
  int64_t n;
  bool replacement;
  __device__ int64_t count(int64_t i) const {
    return comb_count((int64_t)(stops[i] - starts[i]), n, replacement);
  }
  __device__ void emit(int64_t i, int64_t excl, int64_t cnt) const {
    if (i == 0) offsets[0] = 0;
    offsets[i + 1] = excl + cnt;
  }
};

template <typename C>
Error launch_decode(int64_t** tocarry, int64_t* toindex, const int64_t* offsets, const C* starts, const C* stops,
                    int64_t length, int64_t n, bool replacement, int64_t grid_hint) {
  int64_t grid = (int64_t)cu_count() * 16;
  if (grid_hint > 0 && grid_hint < grid) grid = grid_hint;
  decode_kernel<C><<<(unsigned)grid, kDecodeBlock, 0, stream()>>>(tocarry, toindex, offsets, starts, stops, length,
                                                                   n, replacement);
  AKHIP_LAUNCH_CHECK("combinations decode_kernel");
  return success();
}

}  // namespace

template <typename C>
Error combinations_length(int64_t* totallen, int64_t* tooffsets, int64_t n, bool replacement, const C* starts,
                          const C* stops, int64_t length) {
  if (length <= 0) {
    zero_two<<<1, 1, 0, stream()>>>(totallen, tooffsets);
    AKHIP_LAUNCH_CHECK("zero_two");
    return success();
  }
  LengthBody<C> body{totallen, tooffsets, starts, stops, n, replacement, length};
  const int64_t ntiles = ceil_div(length, (int64_t)256 * 8);
  void* s = scratch(scan_scratch_bytes(ntiles));
  if (!s) return failure("out of device memory for scan scratch", kSliceNone, kSliceNone, kFile);
  return launch_lookback_scan<256, 8>(body, length, 0, s);
}

template <typename C>
Error combinations(int64_t** tocarry, int64_t* toindex, int64_t*, int64_t n, bool replacement, const C* starts,
                   const C* stops, int64_t length) {
  if (n <= 0) return success();
  if (n > kMaxN) return failure("combinations with n > 8 are not supported on the hip backend", kSliceNone, kSliceNone, kFile);
  // scratch: per-list first slot (length + 1) followed by the scan state
  const int64_t ntiles = ceil_div(length > 0 ? length : 1, (int64_t)256 * 8);
  const size_t off_bytes = ((size_t)(length + 1) * sizeof(int64_t) + 255) / 256 * 256;
  char* s = (char*)scratch(off_bytes + scan_scratch_bytes(ntiles));
  if (!s) return failure("out of device memory for combinations scratch", kSliceNone, kSliceNone, kFile);
  int64_t* offsets = (int64_t*)s;
  if (length <= 0) {
    zero_two<<<1, 1, 0, stream()>>>(offsets, offsets);
  } else {
    OffsetsBody<C> body{offsets, starts, stops, n, replacement};
    Error e = launch_lookback_scan<256, 8>(body, length, 0, s + off_bytes);
    if (e.str) return e;
  }
  return launch_decode<C>(tocarry, toindex, offsets, starts, stops, length > 0 ? length : 0, n, replacement, 0);
}

}  // namespace akhip

using namespace akhip;

extern "C" {

#define COMB(P, C)                                                                                           \
  Error awkward_ListArray##P##_combinations_length_64(int64_t* totallen, int64_t* tooffsets, int64_t n,      \
                                                      bool replacement, const C* starts, const C* stops,     \
                                                      int64_t length) {                                      \
    return combinations_length<C>(totallen, tooffsets, n, replacement, starts, stops, length);                \
  }                                                                                                          \
  Error awkward_ListArray##P##_combinations_64(int64_t** tocarry, int64_t* toindex, int64_t* fromindex,     \
                                               int64_t n, bool replacement, const C* starts, const C* stops, \
                                               int64_t length) {                                             \
    return combinations<C>(tocarry, toindex, fromindex, n, replacement, starts, stops, length);               \
  }
COMB(32, int32_t)
COMB(U32, uint32_t)
COMB(64, int64_t)
#undef COMB

// Host-layer form (no counterpart in awkward-cpp): decode from the offsets the
// preceding combinations_length call already produced, skipping the recount.
// `totallen` is the host copy of the tuple count, used to size the grid.
Error awkward_hip_combinations_from_offsets(int64_t** tocarry, int64_t* toindex, const int64_t* tooffsets,
                                            const int64_t* starts, const int64_t* stops, int64_t length, int64_t n,
                                            bool replacement, int64_t totallen) {
  if (n <= 0 || length <= 0) return success();
  if (n > kMaxN) return failure("combinations with n > 8 are not supported on the hip backend", kSliceNone, kSliceNone, kFile);
  const int64_t tiles = ceil_div(totallen > 0 ? totallen : 1, (int64_t)kDecodeBlock * kDecodeSlots);
  return launch_decode<int64_t>(tocarry, toindex, tooffsets, starts, stops, length, n, replacement, tiles);
}

}  // extern "C"
