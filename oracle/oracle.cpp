// oracle.cpp -- CPU restatement of Awkward's jagged-array kernels.
//
// TEST INFRASTRUCTURE ONLY.  This library is the checker for the hip
// backend: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg load it.  The product path (awkward_amd) never links, loads or calls
// it, and has no CPU fallback.
//
// Every function keeps the name, argument order and error convention of the
// awkward-cpp kernel it restates (awkward-cpp/include/awkward/common.h:39-78)
// and takes HOST pointers.  Loops are serial and single-threaded, in the CPU
// kernels' order, so float sums accumulate left to right in the output type
// exactly as awkward-cpp does.  Each block cites the CPU kernel and the
// kernel-specification.yml entry whose semantics it follows.  The reference
// sources themselves cannot be compiled here (they include the generated
// awkward/kernels.h), so this file is written from the specification's
// definitions and the CPU sources read as text.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <type_traits>
#include <vector>

extern "C" {
struct Error {
  const char* str;
  const char* filename;
  int64_t identity;
  int64_t attempt;
};
}

namespace {

constexpr int64_t kSliceNone = INT64_MAX;  // common.h:50-51 (2**63 - 1)
const char* kFile = "oracle/oracle.cpp";

inline Error ok() { return Error{nullptr, nullptr, kSliceNone, kSliceNone}; }
inline Error fail(const char* msg, int64_t identity, int64_t attempt) {
  return Error{msg, kFile, identity, attempt};
}

// Integer sums wrap modulo 2**bits like the CPU kernels' native adds; doing
// the add in the unsigned type keeps that well defined.
template <typename OUT, typename IN>
inline OUT add_to(OUT acc, IN x) {
  if constexpr (std::is_floating_point_v<OUT>) {
    return acc + static_cast<OUT>(x);
  } else {
    using U = std::make_unsigned_t<OUT>;
    return static_cast<OUT>(static_cast<U>(acc) + static_cast<U>(static_cast<OUT>(x)));
  }
}

// ---- reductions -----------------------------------------------------------

// awkward_reduce_sum.cpp:11-25; kernel-specification.yml:5035-5195
template <typename OUT, typename IN>
Error reduce_sum(OUT* toptr, const IN* fromptr, const int64_t* offsets, int64_t outlength) {
  for (int64_t bin = 0; bin < outlength; bin++) {
    OUT acc = OUT(0);
    for (int64_t i = offsets[bin]; i < offsets[bin + 1]; i++) acc = add_to<OUT, IN>(acc, fromptr[i]);
    toptr[bin] = acc;
  }
  return ok();
}

// awkward_reduce_sum_int64_bool_64.cpp / awkward_reduce_sum_int32_bool_64.cpp
template <typename OUT>
Error reduce_sum_bool(OUT* toptr, const bool* fromptr, const int64_t* offsets, int64_t outlength) {
  for (int64_t bin = 0; bin < outlength; bin++) {
    OUT acc = 0;
    for (int64_t i = offsets[bin]; i < offsets[bin + 1]; i++)
      if (fromptr[i]) acc++;
    toptr[bin] = acc;
  }
  return ok();
}

// Integer products wrap modulo 2**bits too (multiplied in the unsigned type).
template <typename OUT, typename IN>
inline OUT mul_by(OUT acc, IN x) {
  if constexpr (std::is_floating_point_v<OUT>) {
    return acc * static_cast<OUT>(x);
  } else {
    using U = std::make_unsigned_t<OUT>;
    return static_cast<OUT>(static_cast<U>(acc) * static_cast<U>(static_cast<OUT>(x)));
  }
}

// awkward_reduce_prod.cpp:11-24; kernel-specification.yml:4794-4900: acc = 1,
// acc *= (OUT)x left to right
template <typename OUT, typename IN>
Error reduce_prod(OUT* toptr, const IN* fromptr, const int64_t* offsets, int64_t outlength) {
  for (int64_t bin = 0; bin < outlength; bin++) {
    OUT acc = OUT(1);
    for (int64_t i = offsets[bin]; i < offsets[bin + 1]; i++) acc = mul_by<OUT, IN>(acc, fromptr[i]);
    toptr[bin] = acc;
  }
  return ok();
}

// awkward_reduce_sum_bool.cpp / awkward_reduce_prod_bool.cpp;
// kernel-specification.yml:5477-5550 and :4931-5009: whether any (ANY) or
// every (!ANY) element of the bin is nonzero (x != 0: NaN counts as nonzero,
// -0.0 as zero); an empty bin gives false (any) / true (all)
template <typename IN, bool ANY>
Error reduce_anyall(bool* toptr, const IN* fromptr, const int64_t* offsets, int64_t outlength) {
  for (int64_t bin = 0; bin < outlength; bin++) {
    bool r = !ANY;
    for (int64_t i = offsets[bin]; i < offsets[bin + 1]; i++) {
      if ((fromptr[i] != 0) == ANY) {
        r = ANY;
        break;
      }
    }
    toptr[bin] = r;
  }
  return ok();
}

// awkward_reduce_max.cpp:8-26: best = identity; best = v > best ? v : best
template <typename T>
Error reduce_max(T* toptr, const T* fromptr, const int64_t* offsets, int64_t outlength, T identity) {
  for (int64_t bin = 0; bin < outlength; bin++) {
    T best = identity;
    for (int64_t i = offsets[bin]; i < offsets[bin + 1]; i++) {
      T v = fromptr[i];
      best = (v > best) ? v : best;
    }
    toptr[bin] = best;
  }
  return ok();
}

// awkward_reduce_min.cpp:25-43
template <typename T>
Error reduce_min(T* toptr, const T* fromptr, const int64_t* offsets, int64_t outlength, T identity) {
  for (int64_t bin = 0; bin < outlength; bin++) {
    T best = identity;
    for (int64_t i = offsets[bin]; i < offsets[bin + 1]; i++) {
      T v = fromptr[i];
      best = (v < best) ? v : best;
    }
    toptr[bin] = best;
  }
  return ok();
}

// awkward_reduce_argmax.cpp:8-33: the first element seeds (index, value);
// later elements replace it only when strictly greater.  Global indices;
// -1 for an empty bin.  `starts` is unused.
template <typename T, bool MAX>
Error reduce_arg(int64_t* toptr, const T* fromptr, const int64_t* offsets, const int64_t*,
                 int64_t outlength) {
  for (int64_t bin = 0; bin < outlength; bin++) {
    const int64_t start = offsets[bin], stop = offsets[bin + 1];
    int64_t best = -1;
    if (start < stop) {
      best = start;
      T best_val = fromptr[start];
      for (int64_t i = start + 1; i < stop; i++) {
        const T v = fromptr[i];
        if (MAX ? (v > best_val) : (v < best_val)) {
          best_val = v;
          best = i;
        }
      }
    }
    toptr[bin] = best;
  }
  return ok();
}

}  // namespace

extern "C" {

#define SUM(NAME, OUT, IN) \
  Error NAME(OUT* t, const IN* f, const int64_t* o, int64_t n) { return reduce_sum<OUT, IN>(t, f, o, n); }
SUM(awkward_reduce_sum_int64_int8_64, int64_t, int8_t)
SUM(awkward_reduce_sum_uint64_uint8_64, uint64_t, uint8_t)
SUM(awkward_reduce_sum_int64_int16_64, int64_t, int16_t)
SUM(awkward_reduce_sum_uint64_uint16_64, uint64_t, uint16_t)
SUM(awkward_reduce_sum_int64_int32_64, int64_t, int32_t)
SUM(awkward_reduce_sum_uint64_uint32_64, uint64_t, uint32_t)
SUM(awkward_reduce_sum_int64_int64_64, int64_t, int64_t)
SUM(awkward_reduce_sum_uint64_uint64_64, uint64_t, uint64_t)
SUM(awkward_reduce_sum_float32_float32_64, float, float)
SUM(awkward_reduce_sum_float64_float64_64, double, double)
SUM(awkward_reduce_sum_int32_int8_64, int32_t, int8_t)
SUM(awkward_reduce_sum_uint32_uint8_64, uint32_t, uint8_t)
SUM(awkward_reduce_sum_int32_int16_64, int32_t, int16_t)
SUM(awkward_reduce_sum_uint32_uint16_64, uint32_t, uint16_t)
SUM(awkward_reduce_sum_int32_int32_64, int32_t, int32_t)
SUM(awkward_reduce_sum_uint32_uint32_64, uint32_t, uint32_t)
SUM(awkward_reduce_sum_float64_int8_64, double, int8_t)
SUM(awkward_reduce_sum_float64_uint8_64, double, uint8_t)
SUM(awkward_reduce_sum_float64_int16_64, double, int16_t)
SUM(awkward_reduce_sum_float64_uint16_64, double, uint16_t)
SUM(awkward_reduce_sum_float64_int32_64, double, int32_t)
SUM(awkward_reduce_sum_float64_uint32_64, double, uint32_t)
SUM(awkward_reduce_sum_float64_int64_64, double, int64_t)
SUM(awkward_reduce_sum_float64_uint64_64, double, uint64_t)
SUM(awkward_reduce_sum_float64_float32_64, double, float)
#undef SUM

Error awkward_reduce_sum_int64_bool_64(int64_t* t, const bool* f, const int64_t* o, int64_t n) {
  return reduce_sum_bool<int64_t>(t, f, o, n);
}
Error awkward_reduce_sum_int32_bool_64(int32_t* t, const bool* f, const int64_t* o, int64_t n) {
  return reduce_sum_bool<int32_t>(t, f, o, n);
}

#define PROD(NAME, OUT, IN) \
  Error NAME(OUT* t, const IN* f, const int64_t* o, int64_t n) { return reduce_prod<OUT, IN>(t, f, o, n); }
PROD(awkward_reduce_prod_int64_int8_64, int64_t, int8_t)
PROD(awkward_reduce_prod_uint64_uint8_64, uint64_t, uint8_t)
PROD(awkward_reduce_prod_int64_int16_64, int64_t, int16_t)
PROD(awkward_reduce_prod_uint64_uint16_64, uint64_t, uint16_t)
PROD(awkward_reduce_prod_int64_int32_64, int64_t, int32_t)
PROD(awkward_reduce_prod_uint64_uint32_64, uint64_t, uint32_t)
PROD(awkward_reduce_prod_int64_int64_64, int64_t, int64_t)
PROD(awkward_reduce_prod_uint64_uint64_64, uint64_t, uint64_t)
PROD(awkward_reduce_prod_float32_float32_64, float, float)
PROD(awkward_reduce_prod_float64_float64_64, double, double)
PROD(awkward_reduce_prod_int32_int8_64, int32_t, int8_t)
PROD(awkward_reduce_prod_uint32_uint8_64, uint32_t, uint8_t)
PROD(awkward_reduce_prod_int32_int16_64, int32_t, int16_t)
PROD(awkward_reduce_prod_uint32_uint16_64, uint32_t, uint16_t)
PROD(awkward_reduce_prod_int32_int32_64, int32_t, int32_t)
PROD(awkward_reduce_prod_uint32_uint32_64, uint32_t, uint32_t)
#undef PROD

#define ANYALL(TNAME, T)                                                                             \
  Error awkward_reduce_sum_bool_##TNAME##_64(bool* t, const T* f, const int64_t* o, int64_t n) {     \
    return reduce_anyall<T, true>(t, f, o, n);                                                       \
  }                                                                                                  \
  Error awkward_reduce_prod_bool_##TNAME##_64(bool* t, const T* f, const int64_t* o, int64_t n) {    \
    return reduce_anyall<T, false>(t, f, o, n);                                                      \
  }
ANYALL(bool, bool)
ANYALL(int8, int8_t)
ANYALL(uint8, uint8_t)
ANYALL(int16, int16_t)
ANYALL(uint16, uint16_t)
ANYALL(int32, int32_t)
ANYALL(uint32, uint32_t)
ANYALL(int64, int64_t)
ANYALL(uint64, uint64_t)
ANYALL(float32, float)
ANYALL(float64, double)
#undef ANYALL

#define EXTREMA(TNAME, T)                                                                     \
  Error awkward_reduce_max_##TNAME##_##TNAME##_64(T* t, const T* f, const int64_t* o, int64_t n, \
                                                  T id) {                                    \
    return reduce_max<T>(t, f, o, n, id);                                                    \
  }                                                                                          \
  Error awkward_reduce_min_##TNAME##_##TNAME##_64(T* t, const T* f, const int64_t* o, int64_t n, \
                                                  T id) {                                    \
    return reduce_min<T>(t, f, o, n, id);                                                    \
  }                                                                                          \
  Error awkward_reduce_argmax_##TNAME##_64(int64_t* t, const T* f, const int64_t* o,         \
                                           const int64_t* s, int64_t n) {                    \
    return reduce_arg<T, true>(t, f, o, s, n);                                               \
  }                                                                                          \
  Error awkward_reduce_argmin_##TNAME##_64(int64_t* t, const T* f, const int64_t* o,         \
                                           const int64_t* s, int64_t n) {                    \
    return reduce_arg<T, false>(t, f, o, s, n);                                              \
  }
EXTREMA(int8, int8_t)
EXTREMA(uint8, uint8_t)
EXTREMA(int16, int16_t)
EXTREMA(uint16, uint16_t)
EXTREMA(int32, int32_t)
EXTREMA(uint32, uint32_t)
EXTREMA(int64, int64_t)
EXTREMA(uint64, uint64_t)
EXTREMA(float32, float)
EXTREMA(float64, double)
#undef EXTREMA

// awkward_reduce_count_64.cpp:7-15
Error awkward_reduce_count_64(int64_t* toptr, const int64_t* offsets, int64_t outlength) {
  for (int64_t b = 0; b < outlength; b++) toptr[b] = offsets[b + 1] - offsets[b];
  return ok();
}

// awkward_NumpyArray_reduce_adjust_starts_64.cpp:7-25
Error awkward_NumpyArray_reduce_adjust_starts_64(int64_t* toptr, int64_t outlength, const int64_t*,
                                                 const int64_t* starts) {
  for (int64_t k = 0; k < outlength; k++)
    if (toptr[k] >= 0) toptr[k] -= starts[k];
  return ok();
}

// awkward_NumpyArray_reduce_mask_ByteMaskedArray_64.cpp:7-22
Error awkward_NumpyArray_reduce_mask_ByteMaskedArray_64(int8_t* toptr, const int64_t* offsets,
                                                        int64_t outlength) {
  for (int64_t i = 0; i < outlength; i++) toptr[i] = (offsets[i + 1] - offsets[i] > 0) ? 0 : 1;
  return ok();
}

// ---- local index ----------------------------------------------------------

// awkward_localindex.cpp; kernel-specification.yml (awkward_localindex)
Error awkward_localindex_64(int64_t* toindex, int64_t length) {
  for (int64_t i = 0; i < length; i++) toindex[i] = i;
  return ok();
}

}  // extern "C"

namespace {
// awkward_ListArray_localindex.cpp; kernel-specification.yml:1919-1945
template <typename C>
Error list_localindex(int64_t* toindex, const C* offsets, int64_t length) {
  for (int64_t i = 0; i < length; i++) {
    const int64_t start = (int64_t)offsets[i], stop = (int64_t)offsets[i + 1];
    for (int64_t j = start; j < stop; j++) toindex[j] = j - start;
  }
  return ok();
}

// awkward_ListArray_getitem_jagged_apply.cpp (messages, identity and attempt
// as there; the first failing list in loop order wins)
template <typename C>
Error jagged_apply(int64_t* tooffsets, int64_t* tocarry, const int64_t* slicestarts,
                   const int64_t* slicestops, int64_t sliceouterlen, const int64_t* sliceindex,
                   int64_t sliceinnerlen, const C* fromstarts, const C* fromstops, int64_t contentlen) {
  int64_t k = 0;
  for (int64_t i = 0; i < sliceouterlen; i++) {
    const int64_t s0 = slicestarts[i], s1 = slicestops[i];
    tooffsets[i] = k;
    if (s0 == s1) continue;
    if (s1 < s0) return fail("jagged slice's stops[i] < starts[i]", i, kSliceNone);
    if (s1 > sliceinnerlen) return fail("jagged slice's offsets extend beyond its content", i, s1);
    const int64_t start = (int64_t)fromstarts[i], stop = (int64_t)fromstops[i];
    if (stop < start) return fail("stops[i] < starts[i]", i, kSliceNone);
    if (start != stop && stop > contentlen) return fail("stops[i] > len(content)", i, kSliceNone);
    const int64_t count = stop - start;
    for (int64_t j = s0; j < s1; j++) {
      int64_t index = sliceindex[j];
      if (index < -count || index >= count) return fail("index out of range", i, index);
      if (index < 0) index += count;
      tocarry[k++] = start + index;
    }
  }
  tooffsets[sliceouterlen] = k;
  return ok();
}

// awkward_ListArray_getitem_jagged_descend.cpp
template <typename C>
Error jagged_descend(int64_t* tooffsets, const int64_t* slicestarts, const int64_t* slicestops,
                     int64_t sliceouterlen, const C* fromstarts, const C* fromstops) {
  tooffsets[0] = (sliceouterlen == 0) ? 0 : slicestarts[0];
  for (int64_t i = 0; i < sliceouterlen; i++) {
    const int64_t slicecount = slicestops[i] - slicestarts[i];
    const int64_t count = (int64_t)fromstops[i] - (int64_t)fromstarts[i];
    if (slicecount != count)
      return fail("jagged slice inner length differs from array inner length", i, kSliceNone);
    tooffsets[i + 1] = tooffsets[i] + count;
  }
  return ok();
}

// kernel-utils.cpp:7-39 (awkward_regularize_rangeslice)
inline void regularize_rangeslice(int64_t* start, int64_t* stop, bool posstep, bool hasstart,
                                  bool hasstop, int64_t length) {
  if (posstep) {
    if (!hasstart) *start = 0;
    else if (*start < 0) *start += length;
    if (*start < 0) *start = 0;
    if (*start > length) *start = length;
    if (!hasstop) *stop = length;
    else if (*stop < 0) *stop += length;
    if (*stop < 0) *stop = 0;
    if (*stop > length) *stop = length;
    if (*stop < *start) *stop = *start;
  } else {
    if (!hasstart) *start = length - 1;
    else if (*start < 0) *start += length;
    if (*start < -1) *start = -1;
    if (*start > length - 1) *start = length - 1;
    if (!hasstop) *stop = -1;
    else if (*stop < 0) *stop += length;
    if (*stop < -1) *stop = -1;
    if (*stop > length - 1) *stop = length - 1;
    if (*stop > *start) *stop = *start;
  }
}

// awkward_ListArray_getitem_next_range_carrylength.cpp
template <typename C>
Error next_range_carrylength(int64_t* carrylength, const C* fromstarts, const C* fromstops,
                             int64_t lenstarts, int64_t start, int64_t stop, int64_t step) {
  *carrylength = 0;
  for (int64_t i = 0; i < lenstarts; i++) {
    const int64_t length = (int64_t)fromstops[i] - (int64_t)fromstarts[i];
    int64_t a = start, b = stop;
    regularize_rangeslice(&a, &b, step > 0, start != kSliceNone, stop != kSliceNone, length);
    if (step > 0) {
      for (int64_t j = a; j < b; j += step) (*carrylength)++;
    } else {
      for (int64_t j = a; j > b; j += step) (*carrylength)++;
    }
  }
  return ok();
}

// awkward_ListArray_getitem_next_range.cpp
template <typename C>
Error next_range(C* tooffsets, int64_t* tocarry, const C* fromstarts, const C* fromstops,
                 int64_t lenstarts, int64_t start, int64_t stop, int64_t step) {
  int64_t k = 0;
  tooffsets[0] = 0;
  for (int64_t i = 0; i < lenstarts; i++) {
    const int64_t length = (int64_t)fromstops[i] - (int64_t)fromstarts[i];
    int64_t a = start, b = stop;
    regularize_rangeslice(&a, &b, step > 0, start != kSliceNone, stop != kSliceNone, length);
    if (step > 0) {
      for (int64_t j = a; j < b; j += step) tocarry[k++] = (int64_t)fromstarts[i] + j;
    } else {
      for (int64_t j = a; j > b; j += step) tocarry[k++] = (int64_t)fromstarts[i] + j;
    }
    tooffsets[i + 1] = (C)k;
  }
  return ok();
}

// awkward_ListArray_getitem_next_range_counts.cpp
template <typename C>
Error next_range_counts(int64_t* total, const C* fromoffsets, int64_t lenstarts) {
  *total = 0;
  for (int64_t i = 0; i < lenstarts; i++) *total = *total + (int64_t)fromoffsets[i + 1] - (int64_t)fromoffsets[i];
  return ok();
}

// awkward_ListArray_getitem_next_at.cpp
template <typename C>
Error next_at(int64_t* tocarry, const C* fromstarts, const C* fromstops, int64_t lenstarts, int64_t at) {
  for (int64_t i = 0; i < lenstarts; i++) {
    const int64_t length = (int64_t)fromstops[i] - (int64_t)fromstarts[i];
    int64_t r = at;
    if (r < 0) r += length;
    if (!(0 <= r && r < length)) return fail("index out of range", i, at);
    tocarry[i] = (int64_t)fromstarts[i] + r;
  }
  return ok();
}

// awkward_ListArray_compact_offsets.cpp
template <typename C>
Error compact_offsets(int64_t* tooffsets, const C* fromstarts, const C* fromstops, int64_t length) {
  tooffsets[0] = 0;
  for (int64_t i = 0; i < length; i++) {
    const C start = fromstarts[i], stop = fromstops[i];
    if (stop < start) return fail("stops[i] < starts[i]", i, kSliceNone);
    tooffsets[i + 1] = tooffsets[i] + (int64_t)(stop - start);
  }
  return ok();
}

// awkward_ListArray_combinations_length.cpp: C(size, n) by the incremental
// multiply-then-divide of the CPU kernel (exact while the products fit).
inline int64_t combinations_count(int64_t size, int64_t n, bool replacement) {
  if (replacement) size += n - 1;
  int64_t thisn = n;
  if (thisn > size) return 0;
  if (thisn == size) return 1;
  if (thisn * 2 > size) thisn = size - thisn;
  int64_t c = size;
  for (int64_t j = 2; j <= thisn; j++) {
    c *= (size - j + 1);
    c /= j;
  }
  return c;
}

template <typename C>
Error combinations_length(int64_t* totallen, int64_t* tooffsets, int64_t n, bool replacement,
                          const C* starts, const C* stops, int64_t length) {
  *totallen = 0;
  tooffsets[0] = 0;
  for (int64_t i = 0; i < length; i++) {
    const int64_t c = combinations_count((int64_t)(stops[i] - starts[i]), n, replacement);
    *totallen += c;
    tooffsets[i + 1] = tooffsets[i] + c;
  }
  return ok();
}

// awkward_ListArray_combinations.cpp + kernel-utils.cpp:41-79: every n-tuple
// of positions in [start, stop) in lexicographic order, strictly increasing
// (or non-decreasing with replacement).  Written as an iterative odometer.
template <typename C>
Error combinations(int64_t** tocarry, int64_t* toindex, int64_t* fromindex, int64_t n,
                   bool replacement, const C* starts, const C* stops, int64_t length) {
  for (int64_t j = 0; j < n; j++) toindex[j] = 0;
  if (n <= 0) return ok();
  for (int64_t i = 0; i < length; i++) {
    const int64_t start = (int64_t)starts[i], stop = (int64_t)stops[i];
    // first tuple
    for (int64_t k = 0; k < n; k++) fromindex[k] = replacement ? start : start + k;
    if (fromindex[n - 1] >= stop) continue;
    while (true) {
      for (int64_t k = 0; k < n; k++) tocarry[k][toindex[k]++] = fromindex[k];
      // advance: rightmost position that can still move
      int64_t p = n - 1;
      while (p >= 0) {
        const int64_t limit = replacement ? stop - 1 : stop - (n - p);
        if (fromindex[p] < limit) break;
        p--;
      }
      if (p < 0) break;
      fromindex[p]++;
      for (int64_t k = p + 1; k < n; k++) fromindex[k] = replacement ? fromindex[p] : fromindex[k - 1] + 1;
    }
  }
  return ok();
}

// awkward_IndexedArray_getitem_nextcarry.cpp
template <typename C>
Error indexed_nextcarry(int64_t* tocarry, const C* fromindex, int64_t lenindex, int64_t lencontent) {
  int64_t k = 0;
  for (int64_t i = 0; i < lenindex; i++) {
    const C j = fromindex[i];
    if ((int64_t)j < 0 || (int64_t)j >= lencontent) return fail("index out of range", i, (int64_t)j);
    tocarry[k++] = (int64_t)j;
  }
  return ok();
}

// ---- sorting: awkward_sort.cpp:23-117, awkward_argsort.cpp:23-110 ----------
// Integers and bool compare directly.  Floats: NaN orders before every
// number in both directions, and -0.0 == +0.0.
template <typename T>
inline bool before(T l, T r, bool ascending) {
  if constexpr (std::is_floating_point_v<T>) {
    if (std::isnan(r)) return false;
    if (std::isnan(l)) return true;
  }
  return ascending ? (l < r) : (l > r);
}

template <typename T, typename U>
Error argsort(int64_t* toptr, const T* fromptr, int64_t length, const U* offsets, int64_t offsetslength,
              bool ascending, bool stable) {
  std::iota(toptr, toptr + length, int64_t{0});
  for (int64_t i = 0; i + 1 < offsetslength; i++) {
    int64_t a = (int64_t)offsets[i], b = (int64_t)offsets[i + 1];
    a = std::min(a, length);
    b = std::min(b, length);
    if (a >= b) continue;
    auto cmp = [fromptr, ascending](int64_t x, int64_t y) { return before<T>(fromptr[x], fromptr[y], ascending); };
    if (stable) std::stable_sort(toptr + a, toptr + b, cmp);
    else std::sort(toptr + a, toptr + b, cmp);
    for (int64_t j = a; j < b; j++) toptr[j] -= a;
  }
  return ok();
}

template <typename T>
Error sort(T* toptr, const T* fromptr, int64_t length, const int64_t* offsets, int64_t offsetslength,
           bool ascending, bool stable) {
  std::vector<int64_t> index(length);
  std::iota(index.begin(), index.end(), int64_t{0});
  for (int64_t i = 0; i + 1 < offsetslength; i++) {
    int64_t a = std::min(offsets[i], length), b = std::min(offsets[i + 1], length);
    if (a >= b) continue;
    auto cmp = [fromptr, ascending](int64_t x, int64_t y) { return before<T>(fromptr[x], fromptr[y], ascending); };
    if (stable) std::stable_sort(index.begin() + a, index.begin() + b, cmp);
    else std::sort(index.begin() + a, index.begin() + b, cmp);
  }
  const int64_t total = offsetslength > 0 ? offsets[offsetslength - 1] : 0;
  const int64_t copy = std::min(total, length);
  for (int64_t i = 0; i < copy; i++) toptr[i] = fromptr[index[i]];
  return ok();
}

}  // namespace

extern "C" {

#define LISTARRAY(P, C)                                                                          \
  Error awkward_ListArray##P##_localindex_64(int64_t* t, const C* o, int64_t n) {                \
    return list_localindex<C>(t, o, n);                                                          \
  }                                                                                              \
  Error awkward_ListArray##P##_getitem_jagged_apply_64(                                          \
      int64_t* to, int64_t* tc, const int64_t* ss, const int64_t* se, int64_t sol,               \
      const int64_t* si, int64_t sil, const C* fs, const C* fe, int64_t cl) {                    \
    return jagged_apply<C>(to, tc, ss, se, sol, si, sil, fs, fe, cl);                            \
  }                                                                                              \
  Error awkward_ListArray##P##_getitem_jagged_descend_64(int64_t* to, const int64_t* ss,         \
                                                         const int64_t* se, int64_t sol,         \
                                                         const C* fs, const C* fe) {             \
    return jagged_descend<C>(to, ss, se, sol, fs, fe);                                           \
  }                                                                                              \
  Error awkward_ListArray##P##_getitem_next_range_64(C* to, int64_t* tc, const C* fs, const C* fe, \
                                                     int64_t n, int64_t a, int64_t b, int64_t s) { \
    return next_range<C>(to, tc, fs, fe, n, a, b, s);                                            \
  }                                                                                              \
  Error awkward_ListArray##P##_getitem_next_range_carrylength(                                   \
      int64_t* cl, const C* fs, const C* fe, int64_t n, int64_t a, int64_t b, int64_t s) {       \
    return next_range_carrylength<C>(cl, fs, fe, n, a, b, s);                                    \
  }                                                                                              \
  Error awkward_ListArray##P##_getitem_next_range_counts_64(int64_t* t, const C* o, int64_t n) { \
    return next_range_counts<C>(t, o, n);                                                        \
  }                                                                                              \
  Error awkward_ListArray##P##_getitem_next_at_64(int64_t* t, const C* fs, const C* fe,          \
                                                  int64_t n, int64_t at) {                       \
    return next_at<C>(t, fs, fe, n, at);                                                         \
  }                                                                                              \
  Error awkward_ListArray##P##_compact_offsets_64(int64_t* t, const C* fs, const C* fe,          \
                                                  int64_t n) {                                   \
    return compact_offsets<C>(t, fs, fe, n);                                                     \
  }                                                                                              \
  Error awkward_ListArray##P##_combinations_length_64(int64_t* tl, int64_t* to, int64_t n,       \
                                                      bool r, const C* s, const C* e,            \
                                                      int64_t len) {                             \
    return combinations_length<C>(tl, to, n, r, s, e, len);                                      \
  }                                                                                              \
  Error awkward_ListArray##P##_combinations_64(int64_t** tc, int64_t* ti, int64_t* fi,           \
                                               int64_t n, bool r, const C* s, const C* e,        \
                                               int64_t len) {                                    \
    return combinations<C>(tc, ti, fi, n, r, s, e, len);                                         \
  }                                                                                              \
  Error awkward_IndexedArray##P##_getitem_nextcarry_64(int64_t* t, const C* fi, int64_t li,      \
                                                       int64_t lc) {                             \
    return indexed_nextcarry<C>(t, fi, li, lc);                                                  \
  }
LISTARRAY(32, int32_t)
LISTARRAY(U32, uint32_t)
LISTARRAY(64, int64_t)
#undef LISTARRAY

// awkward_RegularArray_getitem_jagged_expand.cpp
Error awkward_RegularArray_getitem_jagged_expand_64(int64_t* multistarts, int64_t* multistops,
                                                    const int64_t* singleoffsets, int64_t regularsize,
                                                    int64_t regularlength) {
  for (int64_t i = 0; i < regularlength; i++)
    for (int64_t j = 0; j < regularsize; j++) {
      multistarts[i * regularsize + j] = singleoffsets[j];
      multistops[i * regularsize + j] = singleoffsets[j + 1];
    }
  return ok();
}

// awkward_ListArray_getitem_jagged_carrylen.cpp
Error awkward_ListArray_getitem_jagged_carrylen_64(int64_t* carrylen, const int64_t* slicestarts,
                                                   const int64_t* slicestops, int64_t sliceouterlen) {
  *carrylen = 0;
  for (int64_t i = 0; i < sliceouterlen; i++) *carrylen += slicestops[i] - slicestarts[i];
  return ok();
}

// awkward_RegularArray_getitem_next_range.cpp
Error awkward_RegularArray_getitem_next_range_64(int64_t* tocarry, int64_t regular_start, int64_t step,
                                                 int64_t length, int64_t size, int64_t nextsize) {
  for (int64_t i = 0; i < length; i++)
    for (int64_t j = 0; j < nextsize; j++) tocarry[i * nextsize + j] = i * size + regular_start + j * step;
  return ok();
}

// awkward_ListOffsetArray_local_preparenext_64.cpp: an argsort of fromindex.
// The CPU kernel uses std::sort, whose order among equal keys is not
// specified; the oracle uses the stable order, which is one valid answer.
Error awkward_ListOffsetArray_local_preparenext_64(int64_t* tocarry, const int64_t* fromindex,
                                                   int64_t length) {
  std::iota(tocarry, tocarry + length, int64_t{0});
  std::stable_sort(tocarry, tocarry + length,
                   [fromindex](int64_t a, int64_t b) { return fromindex[a] < fromindex[b]; });
  return ok();
}

// ---- option-type contents --------------------------------------------------
// awkward_ByteMaskedArray_numnull.cpp
Error awkward_ByteMaskedArray_numnull(int64_t* numnull, const int8_t* mask, int64_t length, bool validwhen) {
  *numnull = 0;
  for (int64_t i = 0; i < length; i++)
    if ((mask[i] != 0) != validwhen) *numnull = *numnull + 1;
  return ok();
}

// awkward_ByteMaskedArray_reduce_next_64.cpp: the valid elements of each bin,
// in bin order, become the next carry; outindex maps each element to its
// rank among them (-1 if masked); nextoffsets counts them per bin.
Error awkward_ByteMaskedArray_reduce_next_64(int64_t* nextcarry, int64_t* nextoffsets, int64_t* outindex,
                                             const int8_t* mask, const int64_t* offsets, int64_t outlength,
                                             bool validwhen) {
  int64_t k = 0;
  nextoffsets[0] = 0;
  for (int64_t bin = 0; bin < outlength; bin++) {
    for (int64_t i = offsets[bin]; i < offsets[bin + 1]; i++) {
      if ((mask[i] != 0) == validwhen) {
        nextcarry[k] = i;
        outindex[i] = k++;
      } else {
        outindex[i] = -1;
      }
    }
    nextoffsets[bin + 1] = k;
  }
  return ok();
}

// awkward_ByteMaskedArray_reduce_next_nonlocal_nextshifts_64.cpp and
// ..._fromshifts_64.cpp: for each valid element, the masked ones before it
// (plus its incoming shift).
Error awkward_ByteMaskedArray_reduce_next_nonlocal_nextshifts_64(int64_t* nextshifts, const int8_t* mask,
                                                                 int64_t length, bool valid_when) {
  int64_t nullsum = 0, k = 0;
  for (int64_t i = 0; i < length; i++) {
    if ((mask[i] != 0) == valid_when) nextshifts[k++] = nullsum;
    else nullsum++;
  }
  return ok();
}
Error awkward_ByteMaskedArray_reduce_next_nonlocal_nextshifts_fromshifts_64(int64_t* nextshifts, const int8_t* mask,
                                                                           int64_t length, bool valid_when,
                                                                           const int64_t* shifts) {
  int64_t nullsum = 0, k = 0;
  for (int64_t i = 0; i < length; i++) {
    if ((mask[i] != 0) == valid_when) nextshifts[k++] = shifts[i] + nullsum;
    else nullsum++;
  }
  return ok();
}

// awkward_ByteMaskedArray_getitem_nextcarry.cpp and ..._outindex.cpp
Error awkward_ByteMaskedArray_getitem_nextcarry_64(int64_t* tocarry, const int8_t* mask, int64_t length,
                                                   bool validwhen) {
  int64_t k = 0;
  for (int64_t i = 0; i < length; i++)
    if ((mask[i] != 0) == validwhen) tocarry[k++] = i;
  return ok();
}
Error awkward_ByteMaskedArray_getitem_nextcarry_outindex_64(int64_t* tocarry, int64_t* outindex, const int8_t* mask,
                                                            int64_t length, bool validwhen) {
  int64_t k = 0;
  for (int64_t i = 0; i < length; i++) {
    if ((mask[i] != 0) == validwhen) {
      tocarry[k] = i;
      outindex[i] = k++;
    } else {
      outindex[i] = -1;
    }
  }
  return ok();
}

// awkward_NumpyArray_reduce_adjust_starts_shifts_64.cpp: a global arg index
// into the compacted content becomes local to the bin, counting the nulls
// that were removed before it.
Error awkward_NumpyArray_reduce_adjust_starts_shifts_64(int64_t* toptr, int64_t outlength, const int64_t*,
                                                        const int64_t* starts, const int64_t* shifts) {
  for (int64_t k = 0; k < outlength; k++) {
    const int64_t i = toptr[k];
    if (i >= 0) toptr[k] += shifts[i] - starts[k];
  }
  return ok();
}

// ---- non-local (axis != -1) reductions ------------------------------------------
// awkward_ListOffsetArray_reduce_nonlocal_maxcount_offsetscopy_64.cpp
Error awkward_ListOffsetArray_reduce_nonlocal_maxcount_offsetscopy_64(int64_t* maxcount, int64_t* offsetscopy,
                                                                      const int64_t* offsets, int64_t length) {
  *maxcount = 0;
  offsetscopy[0] = offsets[0];
  for (int64_t i = 0; i < length; i++) {
    const int64_t count = offsets[i + 1] - offsets[i];
    if (*maxcount < count) *maxcount = count;
    offsetscopy[i + 1] = offsets[i + 1];
  }
  return ok();
}

// awkward_ListOffsetArray_reduce_nonlocal_preparenext_64.cpp: the elements
// of each outer bin regrouped by column (position in their row), rows in
// order within a column; a group per (bin, column), nextbin = bin *
// maxcount + column.
Error awkward_ListOffsetArray_reduce_nonlocal_preparenext_64(int64_t* nextcarry, int64_t* nextoffsets, int64_t,
                                                             int64_t* maxnextparents, int64_t* distincts,
                                                             int64_t distinctslen, int64_t*, const int64_t* offsets,
                                                             int64_t, const int64_t* outer_offsets, int64_t outlength,
                                                             int64_t maxcount) {
  *maxnextparents = -1;
  for (int64_t i = 0; i < distinctslen; i++) distincts[i] = -1;
  int64_t k = 0;
  nextoffsets[0] = 0;
  for (int64_t bin = 0; bin < outlength; bin++) {
    for (int64_t col = 0; col < maxcount; col++) {
      const int64_t nextbin = bin * maxcount + col;
      bool any = false;
      for (int64_t row = outer_offsets[bin]; row < outer_offsets[bin + 1]; row++) {
        if (col < offsets[row + 1] - offsets[row]) {
          nextcarry[k++] = offsets[row] + col;
          any = true;
        }
      }
      nextoffsets[nextbin + 1] = k;
      if (any) {
        if (nextbin < distinctslen) distincts[nextbin] = nextbin;
        if (nextbin > *maxnextparents) *maxnextparents = nextbin;
      }
    }
  }
  return ok();
}

// awkward_ListOffsetArray_reduce_nonlocal_nextshifts_64.cpp: missing[e] for
// the element in column j of a row is the number of earlier rows of its bin
// that are too short to have a column j; nextshifts gathers it through
// nextcarry; nummissing is left as the last bin's per-column counts.
Error awkward_ListOffsetArray_reduce_nonlocal_nextshifts_64(int64_t* nummissing, int64_t* missing,
                                                            int64_t* nextshifts, const int64_t* offsets, int64_t,
                                                            const int64_t*, const int64_t* outer_offsets,
                                                            int64_t outlength, int64_t maxcount, int64_t nextlen,
                                                            const int64_t* nextcarry) {
  for (int64_t bin = 0; bin < outlength; bin++) {
    for (int64_t k = 0; k < maxcount; k++) nummissing[k] = 0;
    for (int64_t i = outer_offsets[bin]; i < outer_offsets[bin + 1]; i++) {
      const int64_t start = offsets[i], count = offsets[i + 1] - offsets[i];
      for (int64_t k = count; k < maxcount; k++) nummissing[k]++;
      for (int64_t j = 0; j < count; j++) missing[start + j] = nummissing[j];
    }
  }
  for (int64_t j = 0; j < nextlen; j++) nextshifts[j] = missing[nextcarry[j]];
  return ok();
}

// awkward_ListOffsetArray_reduce_nonlocal_outstartsstops_64.cpp: each outer
// bin's output list spans its columns up to the last one that had data.
Error awkward_ListOffsetArray_reduce_nonlocal_outstartsstops_64(int64_t* outstarts, int64_t* outstops,
                                                                const int64_t* distincts, int64_t lendistincts,
                                                                int64_t outlength) {
  if (outlength > 0 && lendistincts > 0) {
    const int64_t maxcount = lendistincts / outlength;
    int64_t k = 0, i_next = 0;
    for (int64_t i = 0; i < lendistincts; i++) {
      if (i == i_next) {
        i_next += maxcount;
        outstarts[k] = i;
        outstops[k] = i;
        k++;
      }
      if (distincts[i] != -1) outstops[k - 1] = i + 1;
    }
  } else {
    for (int64_t k = 0; k < outlength; k++) outstarts[k] = outstops[k] = 0;
  }
  return ok();
}

}  // extern "C"

namespace {

// awkward_IndexedArray_numnull.cpp (negative index = null; an unsigned
// index is never null)
template <typename C>
Error indexed_numnull(int64_t* numnull, const C* fromindex, int64_t lenindex) {
  *numnull = 0;
  for (int64_t i = 0; i < lenindex; i++)
    if (fromindex[i] < 0) *numnull = *numnull + 1;
  return ok();
}

// awkward_IndexedArray_reduce_next_64.cpp
template <typename C>
Error indexed_reduce_next(int64_t* nextcarry, int64_t* nextoffsets, int64_t* outindex, const C* index,
                          const int64_t* offsets, int64_t outlength) {
  int64_t k = 0;
  nextoffsets[0] = 0;
  for (int64_t bin = 0; bin < outlength; bin++) {
    for (int64_t i = offsets[bin]; i < offsets[bin + 1]; i++) {
      if (index[i] >= 0) {
        nextcarry[k] = (int64_t)index[i];
        outindex[i] = k++;
      } else {
        outindex[i] = -1;
      }
    }
    nextoffsets[bin + 1] = k;
  }
  return ok();
}

// awkward_IndexedArray_reduce_next_nonlocal_nextshifts_64.cpp and ..._fromshifts_64.cpp
template <typename C>
Error indexed_nextshifts(int64_t* nextshifts, const C* index, int64_t length, const int64_t* shifts) {
  int64_t nullsum = 0, k = 0;
  for (int64_t i = 0; i < length; i++) {
    if (index[i] >= 0) nextshifts[k++] = (shifts ? shifts[i] : 0) + nullsum;
    else nullsum++;
  }
  return ok();
}

// awkward_IndexedArray_getitem_nextcarry_outindex.cpp:13-28: an index at or
// past lencontent is an error (identity i, attempt j); a negative one is null.
template <typename C>
Error indexed_nextcarry_outindex(int64_t* tocarry, C* toindex, const C* fromindex, int64_t lenindex,
                                 int64_t lencontent) {
  int64_t k = 0;
  for (int64_t i = 0; i < lenindex; i++) {
    const C j = fromindex[i];
    if ((int64_t)j >= lencontent) {
      return fail("index out of range", i, (int64_t)j);
    } else if (j < 0) {
      toindex[i] = (C)-1;
    } else {
      tocarry[k] = (int64_t)j;
      toindex[i] = (C)k;
      k++;
    }
  }
  return ok();
}

}  // namespace

extern "C" {

#define OPTION_INDEXED(P, C)                                                                                     \
  Error awkward_IndexedArray##P##_numnull(int64_t* n, const C* fi, int64_t li) { return indexed_numnull<C>(n, fi, li); } \
  Error awkward_IndexedArray##P##_reduce_next_64(int64_t* nc, int64_t* no, int64_t* oi, const C* ix,             \
                                                 const int64_t* off, int64_t ol) {                              \
    return indexed_reduce_next<C>(nc, no, oi, ix, off, ol);                                                     \
  }                                                                                                            \
  Error awkward_IndexedArray##P##_reduce_next_nonlocal_nextshifts_64(int64_t* ns, const C* ix, int64_t len) {    \
    return indexed_nextshifts<C>(ns, ix, len, nullptr);                                                          \
  }                                                                                                            \
  Error awkward_IndexedArray##P##_reduce_next_nonlocal_nextshifts_fromshifts_64(int64_t* ns, const C* ix,        \
                                                                               int64_t len, const int64_t* sh) { \
    return indexed_nextshifts<C>(ns, ix, len, sh);                                                               \
  }                                                                                                            \
  Error awkward_IndexedArray##P##_getitem_nextcarry_outindex_64(int64_t* tc, C* ti, const C* fi, int64_t li,     \
                                                                int64_t lc) {                                   \
    return indexed_nextcarry_outindex<C>(tc, ti, fi, li, lc);                                                   \
  }
OPTION_INDEXED(32, int32_t)
OPTION_INDEXED(U32, uint32_t)
OPTION_INDEXED(64, int64_t)
#undef OPTION_INDEXED

#define SORT(TNAME, T)                                                                          \
  Error awkward_sort_##TNAME(T* t, const T* f, int64_t len, const int64_t* o, int64_t ol, bool a, \
                             bool s) {                                                          \
    return sort<T>(t, f, len, o, ol, a, s);                                                     \
  }                                                                                             \
  Error awkward_argsort_##TNAME##_int32(int64_t* t, const T* f, int64_t len, const int32_t* o,  \
                                        int64_t ol, bool a, bool s) {                           \
    return argsort<T, int32_t>(t, f, len, o, ol, a, s);                                         \
  }                                                                                             \
  Error awkward_argsort_##TNAME##_uint32(int64_t* t, const T* f, int64_t len, const uint32_t* o, \
                                         int64_t ol, bool a, bool s) {                          \
    return argsort<T, uint32_t>(t, f, len, o, ol, a, s);                                        \
  }                                                                                             \
  Error awkward_argsort_##TNAME##_uint64(int64_t* t, const T* f, int64_t len, const uint64_t* o, \
                                         int64_t ol, bool a, bool s) {                          \
    return argsort<T, uint64_t>(t, f, len, o, ol, a, s);                                        \
  }                                                                                             \
  Error awkward_argsort_##TNAME##_int64(int64_t* t, const T* f, int64_t len, const int64_t* o,  \
                                        int64_t ol, bool a, bool s) {                           \
    return argsort<T, int64_t>(t, f, len, o, ol, a, s);                                         \
  }
SORT(bool, bool)
SORT(int8, int8_t)
SORT(uint8, uint8_t)
SORT(int16, int16_t)
SORT(uint16, uint16_t)
SORT(int32, int32_t)
SORT(uint32, uint32_t)
SORT(int64, int64_t)
SORT(uint64, uint64_t)
SORT(float32, float)
SORT(float64, double)
#undef SORT

}  // extern "C"
