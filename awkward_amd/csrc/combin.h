// combin.h -- counting and unranking of n-combinations of a list's items
// (awkward_ListArray_combinations_length.cpp:16-43 counts; the tuples of
// awkward_ListArray_combinations.cpp in lexicographic order), shared by
// the combinations decode and the fused pairs kernel.
#pragma once

#include <cmath>

#include "common.h"

namespace akhip {
namespace {

constexpr int kMaxN = 8;  // tuple sizes decoded on the device

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

// C(m, K) for a compile-time K: the same incremental product, but every
// divisor is a constant, so no runtime 64-bit division is emitted.
template <int K>
__device__ inline int64_t choose_k(int64_t m) {
  if (m < K) return 0;
  int64_t c = 1;
#pragma unroll
  for (int j = 1; j <= K; j++) c = c * (m - j + 1) / j;  // exact at every step
  return c;
}

__device__ inline int64_t choose_small(int64_t m, int64_t k) {
  switch (k) {
    case 0: return 1;
    case 1: return m > 0 ? m : 0;
    case 2: return choose_k<2>(m);
    case 3: return choose_k<3>(m);
    case 4: return choose_k<4>(m);
    case 5: return choose_k<5>(m);
    case 6: return choose_k<6>(m);
    case 7: return choose_k<7>(m);
    default: return choose_k<8>(m);
  }
}

// Tuples whose remaining kk positions are drawn from [x, s): C(s-x, kk), or
// the multiset count C(s-x+kk-1, kk) with replacement (kk <= 8).
__device__ inline int64_t tail_count(int64_t s, int64_t x, int64_t kk, bool replacement) {
  const int64_t m = s - x;
  if (m <= 0) return 0;
  return replacement ? choose_small(m + kk - 1, kk) : choose_small(m, kk);
}

// Lexicographic unranking of n-combinations of [0, s).  Multisets (with
// replacement) map to combinations of s + n - 1 items by x_i' = x_i + i, an
// order-preserving bijection, so both use the same unrankers.
//
// pairs: the first element a is the largest with F(a) = a(2s - a - 1)/2 <= r
// (pairs starting before a); estimated from the quadratic, then corrected
// with exact integer counts.
__device__ inline void unrank2(int64_t r, int64_t s, int64_t* x) {
  const double t = 2.0 * (double)s - 1.0;
  double disc = t * t - 8.0 * (double)r;
  if (disc < 0.0) disc = 0.0;
  int64_t a = (int64_t)((t - sqrt(disc)) * 0.5);
  if (a < 0) a = 0;
  if (a > s - 2) a = s - 2;
  auto F = [s](int64_t q) { return q * (2 * s - q - 1) / 2; };
  while (a > 0 && F(a) > r) a--;
  while (a < s - 2 && F(a + 1) <= r) a++;
  x[0] = a;
  x[1] = a + 1 + (r - F(a));
}

// triples: with R = C(s,3) - r, the first element is a = s - m for the
// smallest m with C(m, 3) >= R (m ~ cbrt(6R)); the rest is a pair of the
// m - 1 items after a.
__device__ inline void unrank3(int64_t r, int64_t s, int64_t* x) {
  auto C3 = [](int64_t m) { return m < 3 ? (int64_t)0 : m * (m - 1) * (m - 2) / 6; };
  const int64_t R = C3(s) - r;
  int64_t m = (int64_t)cbrt(6.0 * (double)R) + 1;
  if (m < 3) m = 3;
  if (m > s) m = s;
  while (m > 3 && C3(m - 1) >= R) m--;
  while (m < s && C3(m) < R) m++;
  const int64_t a = s - m;
  const int64_t r2 = r - (C3(s) - C3(m));
  int64_t y[2];
  unrank2(r2, m - 1, y);
  x[0] = a;
  x[1] = a + 1 + y[0];
  x[2] = a + 1 + y[1];
}

// any n <= kMaxN: one binary search per position over closed-form counts
__device__ inline void unrank_general(int64_t r, int64_t s, int64_t n, int64_t* x) {
  int64_t prev = -1;
  for (int64_t q = 0; q < n; q++) {
    const int64_t kk = n - q;
    const int64_t xlo = prev + 1, xhi = s - kk;
    const int64_t total = tail_count(s, xlo, kk, false);
    int64_t a = xlo, b = xhi;
    while (a < b) {
      const int64_t mid = (a + b + 1) >> 1;
      if (total - tail_count(s, mid, kk, false) <= r) a = mid;
      else b = mid - 1;
    }
    r -= total - tail_count(s, a, kk, false);
    x[q] = a;
    prev = a;
  }
}

template <int N>
__device__ inline void unrank(int64_t r, int64_t s, int64_t n, bool replacement, int64_t* x) {
  const int64_t sz = replacement ? s + n - 1 : s;
  if constexpr (N == 2) unrank2(r, sz, x);
  else if constexpr (N == 3) unrank3(r, sz, x);
  else unrank_general(r, sz, n, x);
  if (replacement) {
#pragma unroll
    for (int q = 1; q < (N > 0 ? N : kMaxN); q++)
      if (q < n) x[q] -= q;
  }
}

// Colex tables of the pairs of [0, 64) and the triples of [0, 32): entry j
// is the j-th subset in colexicographic order (largest element first), its
// elements packed 6 (pairs) or 5 (triples) bits each, smallest lowest.
// Colex order of the subsets of [0, s) is a prefix of that of [0, s + 1),
// and lex rank r of a list of s items is colex rank C(s, n) - 1 - r of the
// complement y = s - 1 - x (reversed), so one table serves every short list.
constexpr int kColex2Max = 64, kColex2 = 2016;  // C(64, 2)
constexpr int kColex3Max = 32, kColex3 = 4960;  // C(32, 3)
struct ColexTables {
  uint16_t pairs[kColex2];
  uint16_t triples[kColex3];
  constexpr ColexTables() : pairs(), triples() {
    int k = 0;
    for (int c1 = 1; c1 < kColex2Max; c1++)
      for (int c0 = 0; c0 < c1; c0++) pairs[k++] = (uint16_t)(c0 | (c1 << 6));
    k = 0;
    for (int c2 = 2; c2 < kColex3Max; c2++)
      for (int c1 = 1; c1 < c2; c1++)
        for (int c0 = 0; c0 < c1; c0++) triples[k++] = (uint16_t)(c0 | (c1 << 5) | (c2 << 10));
  }
};
__constant__ ColexTables kColex = ColexTables();

// The lexicographic rank-r n-combination (n = 2 or 3, no replacement) of a
// list of m items: from the colex table when m fits it, else by unranking.
template <int N>
__device__ inline void decode_tuple(int64_t r, int64_t m, int64_t (&x)[N]) {
  static_assert(N == 2 || N == 3, "pairs and triples");
  constexpr int TABMAX = N == 2 ? kColex2Max : kColex3Max;
  constexpr int BITS = N == 2 ? 6 : 5;
  constexpr int MASK = (1 << BITS) - 1;
  if (m <= TABMAX) {
    const int64_t total = choose_k<N>(m);
    const uint32_t y = N == 2 ? kColex.pairs[total - 1 - r] : kColex.triples[total - 1 - r];
#pragma unroll
    for (int q = 0; q < N; q++) x[q] = m - 1 - (int64_t)((y >> (BITS * (N - 1 - q))) & MASK);
  } else if constexpr (N == 2) {
    unrank2(r, m, x);
  } else {
    unrank3(r, m, x);
  }
}

}  // namespace
}  // namespace akhip
