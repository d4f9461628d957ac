// tuples.h -- a workgroup's ak.combinations carries for n = 2 or 3 (without
// replacement), in the lexicographic order of
// awkward_ListArray_combinations (i < j (< k) by position within each list),
// written with coalesced stores: used by pairs.hip and by the pair kernels
// compiled at run time (pair_jit.h), so it needs nothing beyond the wave
// width and the integer types.
#pragma once

namespace akhip {

constexpr int kTupLists = 128;                  // lists (and threads) per workgroup
constexpr int kTupPer = 8;                      // consecutive slots a thread decodes per chunk
constexpr int kTupChunk = kTupLists * kTupPer;  // slots per chunk
constexpr int kTupPad = kTupChunk + kTupChunk / kTupPer;
// thread t's kTupPer slots at t * (kTupPer + 1): an odd stride, no bank conflicts
__device__ inline int tup_pad(int i) { return i + i / kTupPer; }

// The next tuple of the list in lexicographic order; false past its last.
template <int N>
__device__ inline bool next_tuple(int (&x)[N], int m) {
  if constexpr (N == 2) {
    if (++x[1] == m) {
      ++x[0];
      x[1] = x[0] + 1;
    }
    return x[0] < m - 1;
  } else {
    if (++x[2] == m) {
      if (++x[1] == m - 1) {
        ++x[0];
        x[1] = x[0] + 1;
      }
      x[2] = x[1] + 1;
    }
    return x[0] < m - 2;
  }
}

// The rank-r tuple of a list of m items by counting: the tuples starting at
// element i number C(m - 1 - i, N - 1).
template <int N>
__device__ inline void tuple_by_count(int r, int m, int (&x)[N]) {
  int i = 0;
  if constexpr (N == 3) {
    while (r >= (m - 1 - i) * (m - 2 - i) / 2) {
      r -= (m - 1 - i) * (m - 2 - i) / 2;
      i++;
    }
    x[0] = i++;
  }
  while (r >= m - 1 - i) {
    r -= m - 1 - i;
    i++;
  }
  x[N - 2] = i;
  x[N - 1] = i + 1 + r;
}

// Writes the N-tuples of this workgroup's lists, in chunks of kTupChunk
// slots: thread t decodes slots [t * kTupPer, (t + 1) * kTupPer) of the chunk -- one
// binary search over the group's tuple offsets and one decode (first(r, m,
// x): the rank-r tuple of a list of m items) for the first, then steps in
// lexicographic order -- into an LDS stage as element positions relative to
// the group's first element, and the chunk is then stored with each carry
// written 64 consecutive slots per wave instruction.  Group-relative values
// are 32-bit (the caller checks that the group's elements and tuples fit).
// s_tbl / s_offl: the group's tuple and element offsets relative to its
// first; stage: N * kTupPad ints.
template <int N, class First>
__device__ inline void group_tuples(int64_t* const* carry, int64_t g0, int64_t E0, const int* s_tbl,
                                    const int* s_offl, int nl, int* stage, First first) {
  const int tid = threadIdx.x;
  const int T = s_tbl[nl];
  for (int c0 = 0; c0 < T; c0 += kTupChunk) {
    const int cn = T - c0 < kTupChunk ? T - c0 : kTupChunk;
    int q = c0 + tid * kTupPer;
    const int qe = q + kTupPer < c0 + cn ? q + kTupPer : c0 + cn;
    if (q < qe) {
      int lo = 0, hi = nl;  // the last list e with s_tbl[e] <= q
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (s_tbl[mid] <= q) lo = mid;
        else hi = mid;
      }
      int e = lo;
      int base = s_offl[e], m = s_offl[e + 1] - base;
      int x[N];
      first(q - s_tbl[e], m, x);
      for (;;) {
        const int at = tup_pad(q - c0);
#pragma unroll
        for (int c = 0; c < N; c++) stage[c * kTupPad + at] = base + x[c];
        if (++q >= qe) break;
        if (!next_tuple<N>(x, m)) {  // on to the next list with tuples
          do {
            e++;
          } while (s_tbl[e + 1] == s_tbl[e]);
          base = s_offl[e];
          m = s_offl[e + 1] - base;
#pragma unroll
          for (int c = 0; c < N; c++) x[c] = c;
        }
      }
    }
    __syncthreads();
    // two consecutive slots per lane, one 16-byte store per carry: the
    // pairs start at an even absolute slot (an odd first slot goes alone)
    const int64_t gbase = g0 + c0;
    const int lead = (int)(gbase & 1);
    if (lead && tid == 0) {
#pragma unroll
      for (int c = 0; c < N; c++) carry[c][gbase] = E0 + stage[c * kTupPad + tup_pad(0)];
    }
    for (int i = lead + 2 * tid; i < cn; i += 2 * kTupLists) {
      const int a0 = tup_pad(i), a1 = tup_pad(i + 1);
#pragma unroll
      for (int c = 0; c < N; c++) {
        int64_t* dst = carry[c] + gbase;
        const int64_t v0 = E0 + stage[c * kTupPad + a0];
        if (i + 1 < cn) {
          typedef long long v2 __attribute__((ext_vector_type(2)));
          const v2 v = {(long long)v0, (long long)(E0 + stage[c * kTupPad + a1])};
          *(v2*)(dst + i) = v;
        } else {
          dst[i] = v0;
        }
      }
    }
    __syncthreads();  // the stage is rewritten by the next chunk
  }
}

}  // namespace akhip
