// combinations.hip -- awkward_ListArray_combinations_length and
// awkward_ListArray_combinations on the device.
//
// Replaces awkward-cpp/src/cpu-kernels/awkward_ListArray_combinations_length.cpp
// and awkward_ListArray_combinations.cpp (+ the recursive
// awkward_ListArray_combinations_step of kernel-utils.cpp:41-79).
//
// The CPU enumerates each list's n-tuples recursively.  Here the tuples are
// decoded independently per output slot: a scan gives every list its first
// slot, and each slot finds its list from those offsets (kept in LDS per
// segment) and decodes its local rank in lexicographic order: pairs and
// triples of short lists from a colex table, longer lists by unranking
// (closed forms for pairs and triples, else one binary search per tuple
// position over closed-form counts).  Slots are spread over threads so that
// all n carries are written with coalesced stores, which is what bounds
// this kernel (8 * n bytes per tuple).

#include <type_traits>

#include "combin.h"
#include "errors.h"
#include "scan.h"

namespace akhip {
namespace {

const char* kFile = "awkward_amd/csrc/combinations.hip";

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
constexpr int kWindow = 2048;  // lists whose offsets a tile may keep in LDS

// Tuples per thread: pairs and triples run 8 slots per thread, general n 4.
template <int N>
constexpr int decode_run() {
  return (N == 2 || N == 3) ? 8 : 4;
}

// The next tuple of a list in lexicographic order: combinations of [0, s)
// (x strictly increasing, x_q <= s - n + q) or, with replacement,
// multisets (x non-decreasing, x_q <= s - 1).  False when x was the last.
template <int NQ, typename X>
__device__ inline bool next_tuple(X* x, X s, int64_t n, bool replacement) {
  int q = -1;
#pragma unroll
  for (int k = 0; k < NQ; k++)
    if (k < n && x[k] < (replacement ? s - 1 : s - (X)n + k)) q = k;
  if (q < 0) return false;
  x[q]++;
#pragma unroll
  for (int k = 1; k < NQ; k++)
    if (k > q && k < n) x[k] = replacement ? x[q] : x[k - 1] + 1;
  return true;
}

// Persistent decode over P = offsets[length] slots (read on the device, so
// no host synchronisation is needed between the scan and the decode).  Each
// workgroup decodes a contiguous run of tiles: the offsets window starting
// at the previous tile's last list (sized from the mean tuples per list) is
// loaded in one coalesced pass and the lists are found by binary search in
// LDS (a tile whose lists run past the window searches global memory).  A
// thread unranks the first of its RUN consecutive slots only and steps to
// the others with next_tuple (moving on to the next list with tuples when
// one is exhausted); each carry then goes through LDS so that the stores are
// coalesced.
template <typename C, int N>
__global__ __launch_bounds__(kDecodeBlock) void decode_kernel(int64_t* const* __restrict__ tocarry,
                                                              int64_t* __restrict__ toindex,
                                                              const int64_t* __restrict__ offsets,
                                                              const C* __restrict__ starts,
                                                              const C* __restrict__ stops, int64_t length,
                                                              int64_t n, bool replacement) {
  constexpr int RUN = decode_run<N>();
  constexpr int NQ = N > 0 ? N : kMaxN;
  constexpr int64_t TILE = (int64_t)kDecodeBlock * RUN;
  __shared__ int64_t* s_out[kMaxN];
  __shared__ int64_t s_off[kWindow + 1];
  __shared__ int64_t s_hint;
  if (threadIdx.x < n) s_out[threadIdx.x] = tocarry[threadIdx.x];
  const int64_t P = offsets[length];
  if (blockIdx.x == 0 && threadIdx.x < n) toindex[threadIdx.x] = P;
  const int64_t ntiles = (P + TILE - 1) / TILE;
  const int64_t per = (ntiles + gridDim.x - 1) / gridDim.x;
  const int64_t t_begin = (int64_t)blockIdx.x * per;
  const int64_t t_end = t_begin + per < ntiles ? t_begin + per : ntiles;
  if (t_begin >= t_end) return;  // workgroup-uniform, before any barrier
  // window length: about twice the lists a tile spans on average
  int64_t want = (int64_t)(2.0 * (double)TILE * (double)length / (double)(P > 0 ? P : 1)) + 64;
  if (want > kWindow + 1) want = kWindow + 1;
  // largest l >= lo with offsets[l] <= slot: the list holding the slot (a
  // list without tuples repeats an offset)
  auto search = [&](int64_t lo, int64_t slot) {
    int64_t hi = length;  // offsets[length] = P > slot
    while (hi - lo > 1) {
      const int64_t mid = (lo + hi) >> 1;
      if (offsets[mid] <= slot) lo = mid;
      else hi = mid;
    }
    return lo;
  };
  if (threadIdx.x == 0) s_hint = search(0, t_begin * TILE);
  __syncthreads();
  int64_t hint = s_hint;
  for (int64_t tile = t_begin; tile < t_end; tile++) {
    const int64_t p0 = tile * TILE;
    const int64_t p1 = (p0 + TILE < P ? p0 + TILE : P) - 1;  // last slot of the tile
    const int64_t wn = want < length + 1 - hint ? want : length + 1 - hint;  // offsets[hint .. hint + wn)
    for (int64_t i = threadIdx.x; i < wn; i += kDecodeBlock) s_off[i] = offsets[hint + i];
    __syncthreads();
    const bool cached = s_off[wn - 1] > p1;  // the window holds every list of the tile
    auto off_at = [&](int64_t l) { return cached ? s_off[l - hint] : offsets[l]; };
    auto local = [&](int64_t p) {  // s_off[lo] <= p < s_off[lo + 1]
      int64_t lo = 0, hi = wn - 1;
      while (hi - lo > 1) {
        const int64_t mid = (lo + hi) >> 1;
        if (s_off[mid] <= p) lo = mid;
        else hi = mid;
      }
      return lo;
    };
    // a run: the list start of each slot and the tuple's positions in its
    // list (32-bit for pairs and triples: a list whose pairs fit in memory
    // has fewer than 2^31 elements)
    using X = typename std::conditional<(N == 2 || N == 3), int32_t, int64_t>::type;
    int64_t run_start[RUN];
    X out[RUN][NQ];
    const int64_t p = p0 + (int64_t)threadIdx.x * RUN;
    if (p <= p1) {
      int64_t l = cached ? hint + local(p) : search(hint, p);
      int64_t end = off_at(l + 1);  // first slot of the next list
      int64_t start = (int64_t)starts[l];
      X s = (X)((int64_t)stops[l] - start);
      X x[NQ];
      {
        int64_t xr[kMaxN];
        unrank<N>(p - off_at(l), (int64_t)s, n, replacement, xr);
#pragma unroll
        for (int q = 0; q < NQ; q++) x[q] = (X)xr[q];
      }
#pragma unroll
      for (int u = 0; u < RUN; u++) {
        const int64_t slot = p + u;
        if (slot > p1) break;
        if (slot == end) {  // the next list with tuples starts here
          do {
            l++;
            end = off_at(l + 1);
          } while (end <= slot);
          start = (int64_t)starts[l];
          s = (X)((int64_t)stops[l] - start);
#pragma unroll
          for (int q = 0; q < NQ; q++) x[q] = replacement ? 0 : (X)q;
        } else if (u > 0) {
          next_tuple<NQ, X>(x, s, n, replacement);
        }
        run_start[u] = start;
#pragma unroll
        for (int q = 0; q < NQ; q++) out[u][q] = x[q];
      }
    }
    // each carry: the run's RUN consecutive slots are RUN * 8 contiguous
    // bytes, written as 16-byte stores when aligned (whole lines fill in L2
    // from neighbouring threads)
    if (p <= p1) {
      const bool full = p + RUN - 1 <= p1;
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        if (q >= n) break;
        int64_t* dst = s_out[q] + p;
        if (full && (((uintptr_t)dst) & 15) == 0) {
#pragma unroll
          for (int u = 0; u < RUN; u += 2) {
            longlong2 v;
            v.x = run_start[u] + (int64_t)out[u][q];
            v.y = run_start[u + 1] + (int64_t)out[u + 1][q];
            *reinterpret_cast<longlong2*>(dst + u) = v;
          }
        } else {
#pragma unroll
          for (int u = 0; u < RUN; u++)
            if (p + u <= p1) dst[u] = run_start[u] + (int64_t)out[u][q];
        }
      }
    }
    if (threadIdx.x == 0) s_hint = cached ? hint + local(p1) : search(hint, p1);
    __syncthreads();  // s_off and s_hint are rewritten by the next tile
    hint = s_hint;
    __syncthreads();
  }
}

constexpr int kSegSlots = 4096;                    // slots decoded per segment
constexpr int kSegLists = 1024;                    // lists whose offsets a segment keeps in LDS
constexpr int kSegPer = kSegSlots / kDecodeBlock;  // slots per thread
__device__ inline int seg_pad(int i) { return i + (i >> 5); }

// Pairs and triples.  Each workgroup owns a contiguous range of slots and
// walks it in segments: the offsets, starts and sizes of the next kSegLists
// lists are loaded into LDS (coalesced); every list with a tuple in the
// segment marks its first slot there, and a max-scan hands each slot its
// list.  A thread then decodes the slots idx = k * 256 + t, so each carry is
// stored 64 consecutive slots per wave instruction.  A slot of a list with
// m = s (+ n - 1 with replacement) <= 64 (pairs) or 32 (triples) items is
// read from the colex table at j = offsets[l + 1] - 1 - slot; longer lists
// unrank.  Marks are list index + 1 (0: none) in 16 bits.  A segment ends
// where the window of lists ends, so any run of empty lists only costs
// window loads.  The carries are streamed out with nontemporal stores: they
// are not read again by this kernel, and keeping them out of the caches
// measured 1.9 -> 1.5 ms for 213M triples.
template <typename C, int N>
__global__ __launch_bounds__(kDecodeBlock) void decode_segments_kernel(int64_t* const* __restrict__ tocarry,
                                                                       int64_t* __restrict__ toindex,
                                                                       const int64_t* __restrict__ offsets,
                                                                       const C* __restrict__ starts,
                                                                       const C* __restrict__ stops, int64_t length,
                                                                       bool replacement) {
  static_assert(N == 2 || N == 3, "table decode is for pairs and triples");
  constexpr int TAB = N == 2 ? kColex2 : kColex3;
  constexpr int TABMAX = N == 2 ? kColex2Max : kColex3Max;
  constexpr int BITS = N == 2 ? 6 : 5;
  constexpr int MASK = (1 << BITS) - 1;
  __shared__ int64_t* s_out[N];
  __shared__ int64_t s_off[kSegLists + 1];
  __shared__ int64_t s_start[kSegLists];
  __shared__ int32_t s_size[kSegLists];
  __shared__ uint16_t s_list[kSegSlots + kSegSlots / 32];
  __shared__ uint16_t s_tab[TAB];
  __shared__ int32_t s_wmax[kDecodeBlock / kWave];
  __shared__ int64_t s_hint;
  const int t = threadIdx.x;
  const int lane = t & (kWave - 1);
  const int wave = t / kWave;
  if (t < N) s_out[t] = tocarry[t];
  const uint16_t* tab = N == 2 ? kColex.pairs : kColex.triples;
  for (int i = t; i < TAB; i += kDecodeBlock) s_tab[i] = tab[i];
  const int64_t P = offsets[length];
  if (blockIdx.x == 0 && t < N) toindex[t] = P;
  const int64_t per = (P + gridDim.x - 1) / gridDim.x;
  const int64_t wbeg = (int64_t)blockIdx.x * per;
  const int64_t wend = wbeg + per < P ? wbeg + per : P;
  if (wbeg >= wend) return;  // workgroup-uniform, before any barrier
  if (t == 0) {  // largest l with offsets[l] <= wbeg (offsets[length] = P > wbeg)
    int64_t lo = 0, hi = length;
    while (hi - lo > 1) {
      const int64_t mid = (lo + hi) >> 1;
      if (offsets[mid] <= wbeg) lo = mid;
      else hi = mid;
    }
    s_hint = lo;
  }
  __syncthreads();
  int64_t hint = s_hint;  // offsets[hint] <= q0 at every segment
  int64_t q0 = wbeg;
  const int64_t extra = replacement ? N - 1 : 0;
  while (q0 < wend) {
    const int64_t wn = kSegLists + 1 < length + 1 - hint ? kSegLists + 1 : length + 1 - hint;
    for (int64_t i = t; i < wn; i += kDecodeBlock) {
      s_off[i] = offsets[hint + i];
      if (i < wn - 1) {
        const int64_t st = (int64_t)starts[hint + i];
        s_start[i] = st;
        s_size[i] = (int32_t)((int64_t)stops[hint + i] - st);
      }
    }
    for (int i = t; i < kSegSlots; i += kDecodeBlock) s_list[seg_pad(i)] = 0;
    __syncthreads();
    const int64_t cover = s_off[wn - 1];  // slots before this belong to lists in the window
    int64_t q1 = q0 + kSegSlots < wend ? q0 + kSegSlots : wend;
    if (cover < q1) q1 = cover;
    if (q1 <= q0) {  // the window holds only lists that end at or before q0
      hint += wn - 1;
      __syncthreads();
      continue;
    }
    const int seg_n = (int)(q1 - q0);
    // each list with a tuple in [q0, q1) marks its first slot there
    for (int i = t; i < wn - 1; i += kDecodeBlock) {
      const int64_t lo = s_off[i], hi = s_off[i + 1];
      if (lo < hi && hi > q0 && lo < q1) s_list[seg_pad((int)((lo > q0 ? lo : q0) - q0))] = (uint16_t)(i + 1);
    }
    __syncthreads();
    // inclusive max-scan of the marks: thread t owns slots [t * PER, t * PER + PER)
    int v[kSegPer];
    int m = 0;
#pragma unroll
    for (int k = 0; k < kSegPer; k++) {
      const int x = (int)s_list[seg_pad(t * kSegPer + k)];
      m = x > m ? x : m;
      v[k] = m;
    }
    int incl = m;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      const int o = __shfl_up(incl, d, kWave);
      if (lane >= d) incl = o > incl ? o : incl;
    }
    int excl = __shfl_up(incl, 1, kWave);
    if (lane == 0) excl = 0;
    if (lane == kWave - 1) s_wmax[wave] = incl;
    __syncthreads();
    for (int w = 0; w < wave; w++) excl = s_wmax[w] > excl ? s_wmax[w] : excl;
#pragma unroll
    for (int k = 0; k < kSegPer; k++) s_list[seg_pad(t * kSegPer + k)] = (uint16_t)(v[k] > excl ? v[k] : excl);
    __syncthreads();
    // decode: slot q0 + idx, idx = k * 256 + t
#pragma unroll 2
    for (int k = 0; k < kSegPer; k++) {
      const int idx = k * kDecodeBlock + t;
      if (idx >= seg_n) break;
      const int64_t slot = q0 + idx;
      const int li = (int)s_list[seg_pad(idx)] - 1;
      const int64_t m_items = (int64_t)s_size[li] + extra;
      int64_t x[N];
      if (m_items <= TABMAX) {
        const uint32_t y = s_tab[(int)(s_off[li + 1] - 1 - slot)];
#pragma unroll
        for (int q = 0; q < N; q++) x[q] = m_items - 1 - (int64_t)((y >> (BITS * (N - 1 - q))) & MASK);
        if (replacement) {
#pragma unroll
          for (int q = 1; q < N; q++) x[q] -= q;
        }
      } else {
        unrank<N>(slot - s_off[li], (int64_t)s_size[li], N, replacement, x);
      }
      const int64_t st = s_start[li];
#pragma unroll
      for (int q = 0; q < N; q++) __builtin_nontemporal_store(st + x[q], s_out[q] + slot);
    }
    hint += (int)s_list[seg_pad(seg_n - 1)] - 1;  // the list holding q1 - 1
    q0 = q1;
    __syncthreads();  // the window and marks are rewritten by the next segment
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

// Workgroups of kernel K that fit on one CU at once (cached per kernel).
template <auto K>
int resident_per_cu() {
  static int cached = 0;
  if (cached == 0) {
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, K, kDecodeBlock, 0) != hipSuccess || b < 1) b = 4;
    cached = b;
  }
  return cached;
}

template <typename C, int N>
void decode_segments(int64_t** tocarry, int64_t* toindex, const int64_t* offsets, const C* starts, const C* stops,
                     int64_t length, bool replacement, int64_t grid_hint, hipStream_t st) {
  // workgroups per CU: 4 for pairs, 3 for triples (measured on 20M events:
  // a fourth workgroup of three concurrent carry streams slowed the triples)
  int per_cu = resident_per_cu<decode_segments_kernel<C, N>>();
  if (per_cu > (N == 3 ? 3 : 4)) per_cu = N == 3 ? 3 : 4;
  int64_t grid = (int64_t)cu_count() * per_cu;
  if (grid_hint > 0 && grid_hint < grid) grid = grid_hint;
  decode_segments_kernel<C, N><<<(unsigned)grid, kDecodeBlock, 0, st>>>(tocarry, toindex, offsets, starts, stops,
                                                                         length, replacement);
}

template <typename C, int N>
void decode(int64_t** tocarry, int64_t* toindex, const int64_t* offsets, const C* starts, const C* stops,
            int64_t length, int64_t n, bool replacement, int64_t grid_hint, hipStream_t st) {
  // persistent: exactly the workgroups the chip holds at once, so no
  // workgroup waits for another one's whole run of tiles
  int64_t grid = (int64_t)cu_count() * resident_per_cu<decode_kernel<C, N>>();
  if (grid_hint > 0 && grid_hint < grid) grid = grid_hint;
  decode_kernel<C, N><<<(unsigned)grid, kDecodeBlock, 0, st>>>(tocarry, toindex, offsets, starts, stops, length, n,
                                                               replacement);
}

template <typename C>
Error launch_decode(int64_t** tocarry, int64_t* toindex, const int64_t* offsets, const C* starts, const C* stops,
                    int64_t length, int64_t n, bool replacement, int64_t grid_hint) {
  hipStream_t st = stream();
  if (n == 2) decode_segments<C, 2>(tocarry, toindex, offsets, starts, stops, length, replacement, grid_hint, st);
  else if (n == 3) decode_segments<C, 3>(tocarry, toindex, offsets, starts, stops, length, replacement, grid_hint, st);
  else decode<C, 0>(tocarry, toindex, offsets, starts, stops, length, n, replacement, grid_hint, st);
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
  const int64_t ntiles = scan_tiles(length);
  void* s = scratch(scan_scratch_bytes(ntiles));
  if (!s) return failure("out of device memory for scan scratch", kSliceNone, kSliceNone, kFile);
  return launch_lookback_scan(body, length, 0, s);
}

template <typename C>
Error combinations(int64_t** tocarry, int64_t* toindex, int64_t*, int64_t n, bool replacement, const C* starts,
                   const C* stops, int64_t length) {
  if (n <= 0) return success();
  if (n > kMaxN) return failure("combinations with n > 8 are not supported on the hip backend", kSliceNone, kSliceNone, kFile);
  // scratch: per-list first slot (length + 1) followed by the scan state
  const int64_t ntiles = scan_tiles(length);
  const size_t off_bytes = ((size_t)(length + 1) * sizeof(int64_t) + 255) / 256 * 256;
  char* s = (char*)scratch(off_bytes + scan_scratch_bytes(ntiles));
  if (!s) return failure("out of device memory for combinations scratch", kSliceNone, kSliceNone, kFile);
  int64_t* offsets = (int64_t*)s;
  if (length <= 0) {
    zero_two<<<1, 1, 0, stream()>>>(offsets, offsets);
  } else {
    OffsetsBody<C> body{offsets, starts, stops, n, replacement};
    Error e = launch_lookback_scan(body, length, 0, s + off_bytes);
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
  const int64_t tiles = ceil_div(totallen > 0 ? totallen : 1, (int64_t)kDecodeBlock * decode_run<0>());
  return launch_decode<int64_t>(tocarry, toindex, tooffsets, starts, stops, length, n, replacement, tiles);
}

}  // extern "C"
