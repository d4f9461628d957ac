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
// Segments are binned by length inside each workgroup of 256 segments:
//   len <= 16    one thread, a 16-wide bitonic network in registers
//   len <= 64    one wave, a 64-lane bitonic network over shuffles
//   len <= 4096  the workgroup, a bitonic network in LDS
//   longer       the workgroup: 4096-element LDS runs, then merge-path
//                passes through a global scratch buffer
#pragma once

#include <type_traits>

#include "common.h"

namespace akhip {
namespace sortimpl {

constexpr int kBlock = 256;
constexpr int kCap = 4096;
constexpr uint64_t kPadKey = ~uint64_t(0);
constexpr uint32_t kPadIdx = ~uint32_t(0);

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

__device__ inline bool pair_less(uint64_t ka, uint32_t ia, uint64_t kb, uint32_t ib) {
  return ka < kb || (ka == kb && ia < ib);
}

// Writes the sorted segment: argsort stores local positions, sort gathers.
template <typename T, bool ARG>
__device__ inline void store_sorted(void* out, const T* from, int64_t a, int64_t r, uint32_t idx) {
  if constexpr (ARG) ((int64_t*)out)[a + r] = (int64_t)idx;
  else ((T*)out)[a + r] = from[a + idx];
}

// ---- len <= 16: registers ------------------------------------------------------
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
            const uint64_t tk = k[i];
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

// ---- len <= kCap: the workgroup, LDS bitonic -----------------------------------------
template <typename T>
__device__ void lds_bitonic(uint64_t* s_key, uint32_t* s_idx, const T* from, int64_t a, int len, bool asc,
                            uint32_t idx_base) {
  int np = 1;
  while (np < len) np <<= 1;
  for (int i = threadIdx.x; i < np; i += kBlock) {
    if (i < len) {
      s_key[i] = order_key<T>(from[a + i], asc);
      s_idx[i] = idx_base + (uint32_t)i;
    } else {
      s_key[i] = kPadKey;
      s_idx[i] = kPadIdx;
    }
  }
  __syncthreads();
  for (int kk = 2; kk <= np; kk <<= 1) {
    for (int j = kk >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < np; i += kBlock) {
        const int l = i ^ j;
        if (l > i) {
          const bool up = (i & kk) == 0;
          const uint64_t ki = s_key[i], kl = s_key[l];
          const uint32_t ii = s_idx[i], il = s_idx[l];
          const bool swap = up ? pair_less(kl, il, ki, ii) : pair_less(ki, ii, kl, il);
          if (swap) {
            s_key[i] = kl;
            s_key[l] = ki;
            s_idx[i] = il;
            s_idx[l] = ii;
          }
        }
      }
      __syncthreads();
    }
  }
}

// ---- longer: LDS runs + merge-path passes ------------------------------------------------
// Number of elements taken from run A among the first `diag` outputs of the
// stable merge of A (first) and B.
__device__ inline int64_t merge_corank(const uint64_t* ka, const uint32_t* ia, int64_t na, const uint64_t* kb,
                                       const uint32_t* ib, int64_t nb, int64_t diag) {
  int64_t lo = diag > nb ? diag - nb : 0;
  int64_t hi = diag < na ? diag : na;
  while (lo < hi) {
    const int64_t i = (lo + hi) >> 1;  // take i from A, diag - i from B
    const int64_t j = diag - i;
    // too few from A if A[i] < B[j-1]
    if (j > 0 && i < na && pair_less(ka[i], ia[i], kb[j - 1], ib[j - 1])) lo = i + 1;
    else hi = i;
  }
  return lo;
}

template <typename T, bool ARG>
__device__ void sort_long(void* out, const T* from, int64_t a, int64_t len, bool asc, uint64_t* s_key,
                          uint32_t* s_idx, uint64_t* gk0, uint32_t* gi0, uint64_t* gk1, uint32_t* gi1) {
  // runs of kCap sorted in LDS, written to buffer 0 (segment-relative)
  for (int64_t c = 0; c < len; c += kCap) {
    const int n = (int)((len - c) < kCap ? (len - c) : kCap);
    lds_bitonic<T>(s_key, s_idx, from, a + c, n, asc, (uint32_t)c);
    for (int i = threadIdx.x; i < n; i += kBlock) {
      gk0[a + c + i] = s_key[i];
      gi0[a + c + i] = s_idx[i];
    }
    __syncthreads();
  }
  uint64_t *sk = gk0, *dk = gk1;
  uint32_t *si = gi0, *di = gi1;
  for (int64_t w = kCap; w < len; w <<= 1) {
    for (int64_t s = 0; s < len; s += 2 * w) {
      const int64_t na = (len - s) < w ? (len - s) : w;
      const int64_t nb = (len - s - na) < w ? (len - s - na) : w;
      const uint64_t* ka = sk + a + s;
      const uint32_t* ia = si + a + s;
      const uint64_t* kb = ka + na;
      const uint32_t* ib = ia + na;
      const int64_t total = na + nb;
      const int64_t per = (total + kBlock - 1) / kBlock;
      const int64_t d0 = (int64_t)threadIdx.x * per < total ? (int64_t)threadIdx.x * per : total;
      const int64_t d1 = d0 + per < total ? d0 + per : total;
      int64_t x = merge_corank(ka, ia, na, kb, ib, nb, d0);
      int64_t y = d0 - x;
      for (int64_t d = d0; d < d1; d++) {
        const bool from_a = y >= nb || (x < na && !pair_less(kb[y], ib[y], ka[x], ia[x]));
        if (from_a) {
          dk[a + s + d] = ka[x];
          di[a + s + d] = ia[x];
          x++;
        } else {
          dk[a + s + d] = kb[y];
          di[a + s + d] = ib[y];
          y++;
        }
      }
    }
    __syncthreads();
    __threadfence_block();
    uint64_t* tk = sk;
    sk = dk;
    dk = tk;
    uint32_t* ti = si;
    si = di;
    di = ti;
  }
  for (int64_t r = threadIdx.x; r < len; r += kBlock) store_sorted<T, ARG>(out, from, a, r, si[a + r]);
  __syncthreads();
}

struct SortSmem {
  uint64_t key[kCap];
  uint32_t idx[kCap];
  uint64_t mask[kBlock / kWave][2];
};

// One tile: segments [S0, S0 + kBlock) of `offsets` (clipped to nseg), each
// segment sorted into `out` (argsort: local positions; sort: values).
// gk*/gi* are merge buffers over the element positions (long segments only).
// Ends with a barrier, so a persistent caller may reuse `sm` at once.
template <typename T, typename U, bool ARG>
__device__ void sort_tile(void* out, const T* __restrict__ from, int64_t length, const U* __restrict__ offsets,
                          int64_t nseg, int64_t S0, bool asc, uint64_t* gk0, uint32_t* gi0, uint64_t* gk1,
                          uint32_t* gi1, SortSmem& sm) {
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wave = tid / kWave;
  const int64_t seg = S0 + tid;
  int64_t a = 0, len = 0;
  if (seg < nseg) {
    a = (int64_t)offsets[seg];
    int64_t b = (int64_t)offsets[seg + 1];
    if (a > length) a = length;
    if (b > length) b = length;
    len = b > a ? b - a : 0;
  }
  if (len > 0 && len <= 16) sort_thread<T, ARG>(out, from, a, (int)len, asc);
  uint64_t mid = __ballot(len > 16 && len <= kWave);
  while (mid) {
    const int b = __builtin_ctzll(mid);
    mid &= mid - 1;
    const int64_t sa = __shfl(a, b, kWave);
    const int sl = (int)__shfl(len, b, kWave);
    sort_wave<T, ARG>(out, from, sa, sl, asc);
  }
  const uint64_t big = __ballot(len > kWave && len <= kCap);
  const uint64_t huge = __ballot(len > kCap);
  if (lane == 0) {
    sm.mask[wave][0] = big;
    sm.mask[wave][1] = huge;
  }
  __syncthreads();
  for (int w = 0; w < kBlock / kWave; w++) {
    for (int kind = 0; kind < 2; kind++) {
      uint64_t m = sm.mask[w][kind];
      while (m) {
        const int b = __builtin_ctzll(m);
        m &= m - 1;
        const int64_t sgi = S0 + w * kWave + b;
        int64_t sa = (int64_t)offsets[sgi], sb = (int64_t)offsets[sgi + 1];
        if (sa > length) sa = length;
        if (sb > length) sb = length;
        const int64_t sl = sb - sa;
        if (kind == 0) {
          lds_bitonic<T>(sm.key, sm.idx, from, sa, (int)sl, asc, 0);
          for (int i = tid; i < sl; i += kBlock) store_sorted<T, ARG>(out, from, sa, i, sm.idx[i]);
          __syncthreads();
        } else {
          sort_long<T, ARG>(out, from, sa, sl, asc, sm.key, sm.idx, gk0, gi0, gk1, gi1);
        }
      }
    }
  }
  __syncthreads();
}

}  // namespace sortimpl
}  // namespace akhip
