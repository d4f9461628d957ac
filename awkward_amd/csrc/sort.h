// sort.h -- segmented sort / argsort building blocks shared by the plain
// kernels (sort.hip) and the batch runner (batch.hip).
//
// Order: every value maps to an order-preserving uint64 key in which NaN is 0
// (first, in both directions, as awkward-cpp's comparators put it:
// awkward_sort.cpp:23-60, awkward_argsort.cpp:23-60) and -0.0 and +0.0 share
// one key (they compare equal there).  Descending order reverses the non-NaN
// keys.  Elements are sorted as (key, position) pairs, so every network below
// yields the unique stable order -- the std::stable_sort result; for
// stable=false the same order is a valid std::sort result.
//
// Segments are binned by length inside each tile of 256 segments:
//   len <= 16    one thread, a 16-wide bitonic network in registers
//   len <= 64    one wave, a 64-lane bitonic network over shuffles
//   longer       queued: the segment's 4096-element chunks go to a device
//                queue (one atomic per wave) and every workgroup of the chip
//                takes chunks: block_sort (16 per thread in registers, then
//                merge stages in LDS) sorts each chunk -- a segment of one
//                chunk is done there -- and merge-path passes
//                (long_merge_kernel), each workgroup producing one
//                4096-element block of a merged run, finish the longer ones;
//                the last pass of a segment writes the output.  Pass w exits
//                at once when no queued segment is longer than w (the maximum
//                length is kept on the device), so the host never
//                synchronises.
#pragma once

#include <type_traits>

#include "common.h"

namespace akhip {
namespace sortimpl {

constexpr int kBlock = 256;
constexpr int kCap = 4096;
constexpr uint64_t kPadKey = ~uint64_t(0);
constexpr uint32_t kPadIdx = ~uint32_t(0);

// Order keys of types of at most 4 bytes fit 32 bits: the LDS stages and
// merge buffers of the long-segment path carry them as uint32 (8 bytes per
// (key, index) pair instead of 12).
template <typename T>
using KeyOf = std::conditional_t<(sizeof(T) <= 4), uint32_t, uint64_t>;

template <typename T>
__device__ inline uint64_t order_key(T v, bool ascending) {
  if constexpr (std::is_same_v<T, bool>) {
    const uint64_t k = v ? 1 : 0;
    return ascending ? k : 1 - k;
  } else if constexpr (std::is_integral_v<T>) {
    using U = std::make_unsigned_t<T>;
    constexpr int bits = 8 * sizeof(T);
    U u = (U)v;
    if constexpr (std::is_signed_v<T>) u ^= (U)(U(1) << (bits - 1));
    const uint64_t k = (uint64_t)u;
    const uint64_t top = (bits == 64) ? ~uint64_t(0) : ((uint64_t(1) << bits) - 1);
    return ascending ? k : top - k;
  } else {
    if (v != v) return 0;  // NaN first
    using U = std::conditional_t<sizeof(T) == 4, uint32_t, uint64_t>;
    constexpr int bits = 8 * sizeof(T);
    if (v == T(0)) v = T(0);  // -0.0 and +0.0 tie
    U u;
    __builtin_memcpy(&u, &v, sizeof(T));
    const U sign = (U)(U(1) << (bits - 1));
    u = (u & sign) ? (U)~u : (U)(u | sign);  // >= 1 for every non-NaN value
    const uint64_t k = (uint64_t)u;
    const uint64_t top = (bits == 64) ? ~uint64_t(0) : ((uint64_t(1) << bits) - 1);
    return ascending ? k : top - k + 1;
  }
}

template <typename K>
__device__ inline bool pair_less(K ka, uint32_t ia, K kb, uint32_t ib) {
  return ka < kb || (ka == kb && ia < ib);
}

// Writes the sorted segment: argsort stores local positions, sort gathers.
template <typename T, bool ARG>
__device__ inline void store_sorted(void* out, const T* from, int64_t a, int64_t r, uint32_t idx) {
  if constexpr (ARG) ((int64_t*)out)[a + r] = (int64_t)idx;
  else ((T*)out)[a + r] = from[a + idx];
}

// ---- len <= 16: registers ------------------------------------------------------
// 16 (key, index) pairs in registers into ascending order (bitonic network).
template <typename K>
__device__ inline void bitonic16(K* k, uint32_t* ix) {
#pragma unroll
  for (int kk = 2; kk <= 16; kk <<= 1) {
#pragma unroll
    for (int j = kk >> 1; j > 0; j >>= 1) {
#pragma unroll
      for (int i = 0; i < 16; i++) {
        const int l = i ^ j;
        if (l > i) {
          const bool up = (i & kk) == 0;
          const bool swap = up ? pair_less(k[l], ix[l], k[i], ix[i]) : pair_less(k[i], ix[i], k[l], ix[l]);
          if (swap) {
            const K tk = k[i];
            k[i] = k[l];
            k[l] = tk;
            const uint32_t ti = ix[i];
            ix[i] = ix[l];
            ix[l] = ti;
          }
        }
      }
    }
  }
}

template <typename T, bool ARG>
__device__ void sort_thread(void* out, const T* from, int64_t a, int len, bool asc) {
  uint64_t k[16];
  uint32_t ix[16];
#pragma unroll
  for (int u = 0; u < 16; u++) {
    if (u < len) {
      k[u] = order_key<T>(from[a + u], asc);
      ix[u] = (uint32_t)u;
    } else {
      k[u] = kPadKey;
      ix[u] = kPadIdx;
    }
  }
  bitonic16(k, ix);
#pragma unroll
  for (int u = 0; u < 16; u++)
    if (u < len) store_sorted<T, ARG>(out, from, a, u, ix[u]);
}

// ---- len <= 64: one wave ---------------------------------------------------------
template <typename T, bool ARG>
__device__ void sort_wave(void* out, const T* from, int64_t a, int len, bool asc) {
  const int lane = threadIdx.x & (kWave - 1);
  uint64_t key = lane < len ? order_key<T>(from[a + lane], asc) : kPadKey;
  uint32_t idx = lane < len ? (uint32_t)lane : kPadIdx;
#pragma unroll
  for (int kk = 2; kk <= kWave; kk <<= 1) {
#pragma unroll
    for (int j = kk >> 1; j > 0; j >>= 1) {
      const uint64_t ok = (uint64_t)__shfl_xor((unsigned long long)key, j, kWave);
      const uint32_t oi = (uint32_t)__shfl_xor((int)idx, j, kWave);
      const bool lower = (lane & j) == 0;
      const bool up = (lane & kk) == 0;
      const bool other_less = pair_less(ok, oi, key, idx);
      // the lower lane keeps the smaller pair in an ascending run
      const bool take = (lower == up) ? other_less : !other_less;
      if (take) {
        key = ok;
        idx = oi;
      }
    }
  }
  if (lane < len) store_sorted<T, ARG>(out, from, a, lane, idx);
}

// ---- (key, index) pairs in LDS -------------------------------------------------------
// A 32-bit key and its index pack into one 64-bit word, key in the high half:
// a pair moves in one LDS access and word order is pair order.  64-bit keys
// keep two arrays.  A stage of `cap` pairs takes lds_pair_words<K>(cap)
// 64-bit words.
template <typename K>
struct LdsPairs;

template <>
struct LdsPairs<uint32_t> {
  uint64_t* w;
  __device__ void get(int i, uint32_t& k, uint32_t& x) const {
    const uint64_t v = w[i];
    k = (uint32_t)(v >> 32);
    x = (uint32_t)v;
  }
  __device__ void put(int i, uint32_t k, uint32_t x) const { w[i] = ((uint64_t)k << 32) | x; }
  __device__ uint32_t idx(int i) const { return (uint32_t)w[i]; }
  __device__ LdsPairs sub(int o) const { return LdsPairs{w + o}; }
};

template <>
struct LdsPairs<uint64_t> {
  uint64_t* k;
  uint32_t* x;
  __device__ void get(int i, uint64_t& kk, uint32_t& xx) const {
    kk = k[i];
    xx = x[i];
  }
  __device__ void put(int i, uint64_t kk, uint32_t xx) const {
    k[i] = kk;
    x[i] = xx;
  }
  __device__ uint32_t idx(int i) const { return x[i]; }
  __device__ LdsPairs sub(int o) const { return LdsPairs{k + o, x + o}; }
};

template <typename K>
constexpr int lds_pair_words(int cap) {
  return sizeof(K) == 4 ? cap : cap + (cap + 1) / 2;
}

template <typename K>
__device__ inline LdsPairs<K> lds_pairs(uint64_t* base, int cap) {
  if constexpr (sizeof(K) == 4) return LdsPairs<K>{base};
  else return LdsPairs<K>{base, (uint32_t*)(base + cap)};
}

// ---- len <= kCap: the workgroup, registers + LDS merges ------------------------------
// LDS position of element i: one pad slot per 16 elements, so the 16-element
// rows the lanes of a wave read and write fall in different banks.
__device__ inline int pad16(int i) { return i + (i >> 4); }
constexpr int kCapPad = kCap + kCap / 16;

// Sorts n <= THREADS * 16 pairs (order_key(from[a + i]), idx_base + i) with
// THREADS threads (t = 0 .. THREADS-1): each sorts 16 in registers, then
// log2(np / 16) merge stages in LDS, each thread finding its 16 outputs'
// merge-path split by binary search (np = n rounded up to a power of two, at
// least 16).  Every thread of the workgroup calls it with the same np_sync
// >= np (the merge stages and their barriers run up to np_sync), so several
// groups of threads may sort different segments side by side.  The sorted
// pairs end in `sp` at positions pad16(0 .. n-1).
template <typename T, int THREADS, typename K = KeyOf<T>>
__device__ void block_sort_part(const LdsPairs<K>& sp, const T* from, int64_t a, int n, bool asc,
                                uint32_t idx_base, int t, int np_sync) {
  constexpr int PER = 16;
  int np = PER;
  while (np < n) np <<= 1;
  for (int i = t; i < np; i += THREADS) {
    const int p = pad16(i);
    if (i < n) sp.put(p, (K)order_key<T>(from[a + i], asc), idx_base + (uint32_t)i);
    else sp.put(p, ~K(0), kPadIdx);
  }
  __syncthreads();
  const bool active = t * PER < np;
  K k[PER];
  uint32_t x[PER];
  const int o = t * PER;  // this thread's first position
  if (active) {
#pragma unroll
    for (int u = 0; u < PER; u++) sp.get(pad16(o + u), k[u], x[u]);
    bitonic16(k, x);
#pragma unroll
    for (int u = 0; u < PER; u++) sp.put(pad16(o + u), k[u], x[u]);
  }
  __syncthreads();
  for (int r = PER; r < np_sync; r <<= 1) {
    const bool merge = active && r < np;
    if (merge) {
      const int ps = o / (2 * r) * (2 * r);
      const int d = o - ps;
      const int A0 = ps, B0 = ps + r;
      int lo = d > r ? d - r : 0;
      int hi = d < r ? d : r;
      while (lo < hi) {
        const int i = (lo + hi) >> 1;
        const int j = d - i;
        K ka, kb;
        uint32_t ia, ib;
        sp.get(pad16(A0 + i), ka, ia);
        sp.get(pad16(B0 + j - 1), kb, ib);
        if (pair_less(ka, ia, kb, ib)) lo = i + 1;
        else hi = i;
      }
      int xa = lo, yb = d - lo;
      K ka, kb;
      uint32_t ia, ib;
      sp.get(pad16(A0 + (xa < r ? xa : r - 1)), ka, ia);
      sp.get(pad16(B0 + (yb < r ? yb : r - 1)), kb, ib);
#pragma unroll
      for (int u = 0; u < PER; u++) {
        const bool from_a = yb >= r || (xa < r && !pair_less(kb, ib, ka, ia));
        k[u] = from_a ? ka : kb;
        x[u] = from_a ? ia : ib;
        xa += from_a ? 1 : 0;
        yb += from_a ? 0 : 1;
        const int p = from_a ? pad16(A0 + (xa < r ? xa : r - 1)) : pad16(B0 + (yb < r ? yb : r - 1));
        K nk;
        uint32_t ni;
        sp.get(p, nk, ni);
        ka = from_a ? nk : ka;
        ia = from_a ? ni : ia;
        kb = from_a ? kb : nk;
        ib = from_a ? ib : ni;
      }
    }
    __syncthreads();
    if (merge) {
#pragma unroll
      for (int u = 0; u < PER; u++) sp.put(pad16(o + u), k[u], x[u]);
    }
    __syncthreads();
  }
}

// ---- longer: LDS runs + merge-path passes ------------------------------------------------
// Number of elements taken from run A among the first `diag` outputs of the
// stable merge of A (first) and B.
// A holds pairs [0, na) of `sp`, B pairs [na, na + nb).
template <typename K>
__device__ inline int merge_corank(const LdsPairs<K>& sp, int na, int nb, int diag) {
  int lo = diag > nb ? diag - nb : 0;
  int hi = diag < na ? diag : na;
  while (lo < hi) {
    const int i = (lo + hi) >> 1;  // take i from A, diag - i from B
    const int j = diag - i;
    // too few from A if A[i] < B[j-1]
    K ka, kb;
    uint32_t ia, ib;
    sp.get(i, ka, ia);
    sp.get(na + j - 1, kb, ib);
    if (pair_less(ka, ia, kb, ib)) lo = i + 1;
    else hi = i;
  }
  return lo;
}

// One 4096-element chunk of a queued long segment.
struct LongChunk {
  void* out;         // output array (argsort: int64 local positions; sort: values)
  const void* from;  // input array
  int64_t a;         // segment start in the arrays
  int64_t len;       // segment length
  int64_t buf;       // segment start in the merge buffers
  int64_t chunk;     // chunk index within the segment
};

// Segments of kWave < len <= kSmall elements go to a queue of their own and
// are sorted a wave each (four side by side per workgroup).
constexpr int kSmall = kWave * 16;

// Capacities: the small queue holds at most elements / (kWave + 1) segments;
// the chunk queue sum over longer segments of ceil(len / kCap), at most
// elements / kCap + elements / (kSmall + 1) + 1.
inline int64_t small_queue_capacity(int64_t elements) { return elements / (kWave + 1) + 2; }
inline int64_t chunk_queue_capacity(int64_t elements) {
  return (elements + kCap - 1) / kCap + elements / (kSmall + 1) + 2;
}

struct LongQueue {
  LongChunk* chunks;
  unsigned long long* count;   // chunks queued
  unsigned long long* maxlen;  // longest segment in the chunk queue
  int64_t capacity;
  LongChunk* small;
  unsigned long long* small_count;
  int64_t small_capacity;
  int64_t* coranks;            // 2 per chunk: a merge pass's split of its output range
};

// Bytes of the co-rank array for a chunk queue of `capacity` entries.
inline size_t corank_bytes(int64_t capacity) { return ((size_t)capacity * 16 + 255) / 256 * 256; }

// Called by every lane of a wave: the lanes with want set append their
// segment -- one entry in the small queue, or its chunks in the chunk queue;
// one atomic per wave and queue reserves the room.
__device__ inline void queue_segments(const LongQueue& q, bool want, void* out, const void* from, int64_t a,
                                      int64_t len, int64_t buf) {
  if (!__any(want)) return;
  const int lane = threadIdx.x & (kWave - 1);
  const bool small = want && len <= kSmall;
  if (__any(small)) {
    const unsigned long long one = small ? 1ull : 0ull;
    const unsigned long long sincl = wave_inclusive_scan(one);
    const unsigned long long stotal = __shfl(sincl, kWave - 1, kWave);
    unsigned long long sbase = 0;
    if (lane == 0) sbase = atomicAdd(q.small_count, stotal);
    sbase = __shfl(sbase, 0, kWave);
    const int64_t at = (int64_t)(sbase + sincl - one);
    if (small && at < q.small_capacity) q.small[at] = LongChunk{out, from, a, len, buf, 0};
  }
  want = want && !small;
  if (!__any(want)) return;
  const unsigned long long nch = want ? (unsigned long long)((len + kCap - 1) / kCap) : 0ull;
  const unsigned long long incl = wave_inclusive_scan(nch);
  const unsigned long long total = __shfl(incl, kWave - 1, kWave);
  unsigned long long longest = want ? (unsigned long long)len : 0ull;
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) {
    const unsigned long long o = __shfl_xor(longest, d, kWave);
    longest = o > longest ? o : longest;
  }
  unsigned long long base = 0;
  if (lane == 0) {
    base = atomicAdd(q.count, total);
    atomicMax(q.maxlen, longest);
  }
  base = __shfl(base, 0, kWave);
  const int64_t first = (int64_t)(base + incl - nch);
  for (int64_t c = 0; c < (int64_t)nch && first + c < q.capacity; c++)
    q.chunks[first + c] = LongChunk{out, from, a, len, buf, c};
}


// One tile: segments [S0, S0 + kBlock) of `offsets` (clipped to nseg), each
// segment of at most kWave elements sorted into `out` (argsort: local
// positions; sort: values); longer ones are queued (merge-buffer position
// buf_base + segment start).
template <typename T, typename U, bool ARG>
__device__ void sort_tile(void* out, const T* __restrict__ from, int64_t length, const U* __restrict__ offsets,
                          int64_t nseg, int64_t S0, bool asc, const LongQueue& q, int64_t buf_base) {
  const int tid = threadIdx.x;
  const int64_t seg = S0 + tid;
  int64_t a = 0, len = 0;
  if (seg < nseg) {
    a = (int64_t)offsets[seg];
    int64_t b = (int64_t)offsets[seg + 1];
    if (a > length) a = length;
    if (b > length) b = length;
    len = b > a ? b - a : 0;
  }
  queue_segments(q, len > kWave, out, from, a, len, buf_base + a);
  if (len > 0 && len <= 16) sort_thread<T, ARG>(out, from, a, (int)len, asc);
  uint64_t mid = __ballot(len > 16 && len <= kWave);
  while (mid) {
    const int b = __builtin_ctzll(mid);
    mid &= mid - 1;
    const int64_t sa = __shfl(a, b, kWave);
    const int sl = (int)__shfl(len, b, kWave);
    sort_wave<T, ARG>(out, from, sa, sl, asc);
  }
}

// ---- queued long segments -------------------------------------------------------------
// (a template so every translation unit may instantiate it)
template <int Dummy = 0>
__global__ void long_queue_reset(unsigned long long* count, unsigned long long* maxlen,
                                 unsigned long long* small_count) {
  *count = 0ull;
  *maxlen = 0ull;
  *small_count = 0ull;
}

// Small segments four at a time, a wave each, in quarters of the LDS tile;
// then every chunk of the longer ones with the whole workgroup: a segment of
// one chunk straight to its output, longer ones into merge buffer 0
// (segment-relative positions).
template <typename T, bool ARG>
__global__ __launch_bounds__(kBlock) void long_runs_kernel(LongQueue q, bool asc, KeyOf<T>* __restrict__ gk,
                                                           uint32_t* __restrict__ gi) {
  constexpr int NW = kBlock / kWave;
  constexpr int QUARTER = kCapPad / NW;  // pad16(kSmall) fits
  using K = KeyOf<T>;
  __shared__ uint64_t s_raw[lds_pair_words<K>(kCapPad)];
  __shared__ int s_np[NW];
  const LdsPairs<K> sp = lds_pairs<K>(s_raw, kCapPad);
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  int64_t ns = (int64_t)*q.small_count;
  if (ns > q.small_capacity) ns = q.small_capacity;
  for (int64_t g = blockIdx.x; g * NW < ns; g += gridDim.x) {
    const int64_t e = g * NW + wave;
    LongChunk ch{};
    int n = 0;
    if (e < ns) {
      ch = q.small[e];
      n = (int)ch.len;
    }
    int np = 16;
    while (np < n) np <<= 1;
    if (lane == 0) s_np[wave] = np;
    __syncthreads();
    int np_sync = s_np[0];
    for (int w = 1; w < NW; w++) np_sync = s_np[w] > np_sync ? s_np[w] : np_sync;
    block_sort_part<T, kWave>(sp.sub(wave * QUARTER), (const T*)ch.from, ch.a, n, asc, 0, lane, np_sync);
    for (int i = lane; i < n; i += kWave)
      store_sorted<T, ARG>(ch.out, (const T*)ch.from, ch.a, i, sp.idx(wave * QUARTER + pad16(i)));
    __syncthreads();
  }
  int64_t nc = (int64_t)*q.count;
  if (nc > q.capacity) nc = q.capacity;
  for (int64_t c = blockIdx.x; c < nc; c += gridDim.x) {
    const LongChunk ch = q.chunks[c];
    const int64_t c0 = ch.chunk * kCap;
    const int n = (int)((ch.len - c0) < kCap ? (ch.len - c0) : kCap);
    int np = 16;
    while (np < n) np <<= 1;
    block_sort_part<T, kBlock>(sp, (const T*)ch.from, ch.a + c0, n, asc, (uint32_t)c0, threadIdx.x, np);
    if (ch.len <= kCap) {
      for (int i = threadIdx.x; i < n; i += kBlock)
        store_sorted<T, ARG>(ch.out, (const T*)ch.from, ch.a, i, sp.idx(pad16(i)));
    } else {
      for (int i = threadIdx.x; i < n; i += kBlock) {
        K kk;
        uint32_t xx;
        sp.get(pad16(i), kk, xx);
        gk[ch.buf + c0 + i] = kk;
        gi[ch.buf + c0 + i] = xx;
      }
    }
    __syncthreads();
  }
}

// Co-ranks of two diagonals at once: threads [0, 128) search diag[0],
// threads [128, 256) diag[1]; each round tests 128 positions of the
// remaining range, so a range of n takes about log_128(n) rounds of
// parallel loads.  P(i) = "A[i] sorts before B[diag - i - 1]" is true below
// the co-rank and false from it on.
template <typename K>
__device__ inline void corank_pair(const K* ka, const uint32_t* ia, int64_t na, const K* kb,
                                   const uint32_t* ib, int64_t nb, const int64_t* diag, int64_t* result,
                                   int* s_cnt) {
  constexpr int G = kBlock / 2;
  const int g = threadIdx.x / G;
  const int t = threadIdx.x % G;
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int64_t d = diag[g];
  int64_t lo = d > nb ? d - nb : 0;
  int64_t hi = d < na ? d : na;
  for (;;) {
    const bool done = lo >= hi;
    if (__syncthreads_and(done ? 1 : 0)) break;
    const int64_t step = (hi - lo + G) / G;
    const int64_t i = lo + (int64_t)t * step;
    bool pred = false;
    if (!done && i < hi) {
      const int64_t j = d - i;
      pred = j > 0 && i < na && pair_less(ka[i], ia[i], kb[j - 1], ib[j - 1]);
    }
    const uint64_t bal = __ballot(pred);
    if (lane == 0) s_cnt[wave] = __popcll(bal);
    __syncthreads();
    const int k = s_cnt[2 * g] + s_cnt[2 * g + 1];
    if (!done) {
      const int64_t nlo = k > 0 ? lo + (int64_t)(k - 1) * step + 1 : lo;
      const int64_t cand = lo + (int64_t)k * step;
      hi = cand < hi ? cand : hi;
      lo = nlo;
    }
    __syncthreads();  // s_cnt is rewritten by the next round
  }
  if (t == 0) result[g] = lo;
  __syncthreads();
}

// Before each merge pass: the co-ranks of every chunk's two output
// boundaries (one thread per boundary, a binary search in global memory --
// all of them in flight at once, where a merging workgroup would wait on
// them one chunk after another).  P(i) = "A[i] sorts before B[d - i - 1]"
// is true below the co-rank and false from it on.
template <typename K>
__global__ __launch_bounds__(kBlock) void long_corank_kernel(const LongChunk* __restrict__ chunks,
                                                             const unsigned long long* __restrict__ count,
                                                             const unsigned long long* __restrict__ maxlen,
                                                             int64_t capacity, int64_t w,
                                                             const K* __restrict__ sk,
                                                             const uint32_t* __restrict__ si,
                                                             int64_t* __restrict__ coranks) {
  if ((int64_t)*maxlen <= w) return;
  int64_t nc = (int64_t)*count;
  if (nc > capacity) nc = capacity;
  for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < 2 * nc; t += (int64_t)gridDim.x * kBlock) {
    const LongChunk ch = chunks[t >> 1];
    if (ch.len <= w) continue;
    const int64_t d0 = ch.chunk * kCap;
    const int64_t d = (t & 1) ? ((d0 + kCap) < ch.len ? d0 + kCap : ch.len) : d0;
    const int64_t ps = d0 / (2 * w) * (2 * w);
    const int64_t na = (ch.len - ps) < w ? ch.len - ps : w;
    const int64_t rest = ch.len - ps - na;
    const int64_t nb = rest < w ? rest : w;
    const K* ka = sk + ch.buf + ps;
    const uint32_t* ia = si + ch.buf + ps;
    const K* kb = ka + na;
    const uint32_t* ib = ia + na;
    const int64_t dd = d - ps;
    int64_t lo = dd > nb ? dd - nb : 0;
    int64_t hi = dd < na ? dd : na;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      const int64_t j = dd - mid;
      if (j > 0 && pair_less(ka[mid], ia[mid], kb[j - 1], ib[j - 1])) lo = mid + 1;
      else hi = mid;
    }
    coranks[t] = lo;
  }
}

// Merge pass over runs of width w: the workgroup for chunk c produces
// elements [c * kCap, (c + 1) * kCap) of the merged pair of runs holding
// them, between the co-ranks long_corank_kernel found; the pass that
// completes a segment writes its output.
// Merge passes use workgroups of 512 threads, 8 outputs each: twice the
// waves per CU of 256-thread blocks at the same LDS (the merge is bound by
// LDS latency): 40.5 -> 32.8 ms on the 256-array batch.
constexpr int kMergeBlock = 512;

template <typename T, bool ARG>
__global__ __launch_bounds__(kMergeBlock) void long_merge_kernel(const LongChunk* __restrict__ chunks,
                                                            const unsigned long long* __restrict__ count,
                                                            const unsigned long long* __restrict__ maxlen,
                                                            int64_t capacity, int64_t w,
                                                            const KeyOf<T>* __restrict__ sk,
                                                            const uint32_t* __restrict__ si,
                                                            KeyOf<T>* __restrict__ dk, uint32_t* __restrict__ di,
                                                            const int64_t* __restrict__ coranks) {
  using K = KeyOf<T>;
  if ((int64_t)*maxlen <= w) return;  // every queued segment is already one run
  constexpr int PER = kCap / kMergeBlock;   // outputs per thread
  __shared__ uint64_t s_raw[lds_pair_words<K>(kCap)];
  const LdsPairs<K> sp = lds_pairs<K>(s_raw, kCap);
  int64_t nc = (int64_t)*count;
  if (nc > capacity) nc = capacity;
  const int tid = threadIdx.x;
  for (int64_t c = blockIdx.x; c < nc; c += gridDim.x) {
    const LongChunk ch = chunks[c];
    if (ch.len <= w) continue;  // workgroup-uniform
    const int64_t d0 = ch.chunk * kCap;
    const int64_t d1 = (d0 + kCap) < ch.len ? d0 + kCap : ch.len;
    const int64_t ps = d0 / (2 * w) * (2 * w);
    const int64_t na = (ch.len - ps) < w ? ch.len - ps : w;
    const K* ka = sk + ch.buf + ps;
    const uint32_t* ia = si + ch.buf + ps;
    const K* kb = ka + na;
    const uint32_t* ib = ia + na;
    const int64_t x0 = coranks[2 * c], x1 = coranks[2 * c + 1];
    const int64_t s_diag[2] = {d0 - ps, d1 - ps};
    const int64_t y0 = s_diag[0] - x0, y1 = s_diag[1] - x1;
    const int nal = (int)(x1 - x0), nbl = (int)(y1 - y0);
    const int n = nal + nbl;
    for (int i = tid; i < n; i += kMergeBlock) {
      if (i < nal) sp.put(i, ka[x0 + i], ia[x0 + i]);
      else sp.put(i, kb[y0 + i - nal], ib[y0 + i - nal]);
    }
    __syncthreads();
    // this thread's outputs [tid * PER, tid * PER + PER) of the block
    K rk[PER];
    uint32_t ri[PER];
    const int o0 = tid * PER < n ? tid * PER : n;
    const int o1 = o0 + PER < n ? o0 + PER : n;
    if (o0 < o1) {
      // the heads of both runs stay in registers; each output reloads only
      // the side it consumed (one key and one index read per output)
      int x = merge_corank(sp, nal, nbl, o0);
      int y = o0 - x;
      const int last = n - 1;
      K ka, kb;
      uint32_t ia, ib;
      sp.get(x < last ? x : last, ka, ia);
      sp.get(nal + y < last ? nal + y : last, kb, ib);
#pragma unroll
      for (int u = 0; u < PER; u++) {
        if (o0 + u < o1) {
          const bool from_a = y >= nbl || (x < nal && !pair_less(kb, ib, ka, ia));
          rk[u] = from_a ? ka : kb;
          ri[u] = from_a ? ia : ib;
          x += from_a ? 1 : 0;
          y += from_a ? 0 : 1;
          int p = from_a ? x : nal + y;
          p = p < last ? p : last;
          K nk;
          uint32_t ni;
          sp.get(p, nk, ni);
          ka = from_a ? nk : ka;
          ia = from_a ? ni : ia;
          kb = from_a ? kb : nk;
          ib = from_a ? ib : ni;
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PER; u++)
      if (o0 + u < o1) sp.put(o0 + u, rk[u], ri[u]);
    __syncthreads();
    if (2 * w >= ch.len) {
      for (int i = tid; i < n; i += kMergeBlock) store_sorted<T, ARG>(ch.out, (const T*)ch.from, ch.a, d0 + i, sp.idx(i));
    } else {
      for (int i = tid; i < n; i += kMergeBlock) {
        K kk;
        uint32_t xx;
        sp.get(i, kk, xx);
        dk[ch.buf + d0 + i] = kk;
        di[ch.buf + d0 + i] = xx;
      }
    }
    __syncthreads();
  }
}

// Sorts everything queued: runs, then merge passes up to runs of max_len
// (an upper bound of the longest segment known on the host).  Buffers
// (gk0, gi0) and (gk1, gi1) cover the merge-buffer positions.
template <typename T, bool ARG>
inline hipError_t launch_long_sort(const LongQueue& q, int64_t max_len, bool asc, void* gk0v, uint32_t* gi0,
                                   void* gk1v, uint32_t* gi1, hipStream_t st) {
  using K = KeyOf<T>;
  K* gk0 = (K*)gk0v;
  K* gk1 = (K*)gk1v;
  if (max_len <= kWave) return hipSuccess;
  int64_t grid = (int64_t)cu_count() * 8;
  if (grid > q.capacity + q.small_capacity) grid = q.capacity + q.small_capacity;
  if (grid < 1) grid = 1;
  long_runs_kernel<T, ARG><<<(unsigned)grid, kBlock, 0, st>>>(q, asc, gk0, gi0);
  K *sk = gk0, *dk = gk1;
  uint32_t *si = gi0, *di = gi1;
  int64_t cgrid = (int64_t)cu_count() * 16;
  if (cgrid > (2 * q.capacity + kBlock - 1) / kBlock) cgrid = (2 * q.capacity + kBlock - 1) / kBlock;
  if (cgrid < 1) cgrid = 1;
  for (int64_t w = kCap; w < max_len; w <<= 1) {
    long_corank_kernel<K><<<(unsigned)cgrid, kBlock, 0, st>>>(q.chunks, q.count, q.maxlen, q.capacity, w, sk, si,
                                                             q.coranks);
    long_merge_kernel<T, ARG><<<(unsigned)grid, kMergeBlock, 0, st>>>(q.chunks, q.count, q.maxlen, q.capacity, w, sk, si,
                                                                 dk, di, q.coranks);
    K* tk = sk;
    sk = dk;
    dk = tk;
    uint32_t* ti = si;
    si = di;
    di = ti;
  }
  return hipGetLastError();
}

}  // namespace sortimpl
}  // namespace akhip
