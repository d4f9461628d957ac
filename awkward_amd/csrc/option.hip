// option.hip -- option-type contents (ByteMaskedArray, IndexedOptionArray)
// under the reducers and getitem.
//
// Replaces awkward-cpp/src/cpu-kernels/
//   awkward_ByteMaskedArray_numnull.cpp
//   awkward_ByteMaskedArray_reduce_next_64.cpp
//   awkward_ByteMaskedArray_reduce_next_nonlocal_nextshifts_64.cpp (+ _fromshifts_64)
//   awkward_ByteMaskedArray_getitem_nextcarry.cpp (+ _outindex)
//   awkward_IndexedArray_numnull.cpp
//   awkward_IndexedArray_reduce_next_64.cpp
//   awkward_IndexedArray_reduce_next_nonlocal_nextshifts_64.cpp (+ _fromshifts_64)
//   awkward_IndexedArray_getitem_nextcarry_outindex.cpp
//   awkward_NumpyArray_reduce_adjust_starts_shifts_64.cpp
//
// Every CPU loop here walks the elements once with a running count k of the
// valid ones.  On the device that is a stream compaction: one pass of the
// decoupled look-back scan over the elements (count 1 for a valid element)
// hands each valid element its rank k, and the emit writes whatever the CPU
// loop writes at that k -- the carry, outindex, or the shift (the nulls
// before element i are i - k).  The reduce_next kernels also need each
// bin's running count: the valid positions are ascending, so
// nextoffsets[b + 1] is the number of them below offsets[b + 1], found by
// binary search per bin.

#include "errors.h"
#include "scan.h"

namespace akhip {
namespace {

const char* kFile = "awkward_amd/csrc/option.hip";
const char* const kIndexMsgs[] = {nullptr, "index out of range"};
constexpr int kBlock = 256;
constexpr int kIpt = 8;

// ---- validity of one element -------------------------------------------------
struct MaskValid {
  const int8_t* mask;
  bool validwhen;
  __device__ bool operator()(int64_t i) const { return (mask[i] != 0) == validwhen; }
};

template <typename C>
struct IndexValid {  // a negative index is null; an unsigned index never is
  const C* index;
  __device__ bool operator()(int64_t i) const { return index[i] >= 0; }
};

// ---- scan bodies (items are elements) ------------------------------------------
// getitem_nextcarry (+ _outindex): tocarry[k] = i; outindex[i] = k or -1
template <class V>
struct NextcarryBody {
  int64_t* tocarry;
  int64_t* outindex;  // may be null
  V valid;
  __device__ int64_t count(int64_t i) const { return valid(i) ? 1 : 0; }
  __device__ void emit(int64_t i, int64_t k, int64_t c) const {
    if (c) tocarry[k] = i;
    if (outindex) outindex[i] = c ? k : -1;
  }
};

// nonlocal_nextshifts (+ _fromshifts): nextshifts[k] = shifts[i] + nulls before i
template <class V>
struct ShiftsBody {
  int64_t* nextshifts;
  const int64_t* shifts;  // may be null
  V valid;
  __device__ int64_t count(int64_t i) const { return valid(i) ? 1 : 0; }
  __device__ void emit(int64_t i, int64_t k, int64_t c) const {
    if (c) nextshifts[k] = (shifts ? shifts[i] : 0) + (i - k);
  }
};

// reduce_next over the elements [base, base + n): the carry of a valid
// element (its position, or its index), outindex, the valid positions in
// order (for the bins' counts) and, from the last item, the total.
template <class V, typename C>
struct ReduceNextBody {
  int64_t* nextcarry;
  int64_t* outindex;
  int64_t* positions;  // == nextcarry when the carry is the position
  int64_t* total;
  const C* index;      // null: the carry is the position
  int64_t base;
  int64_t n;
  V valid;
  __device__ int64_t count(int64_t t) const { return valid(base + t) ? 1 : 0; }
  __device__ void emit(int64_t t, int64_t k, int64_t c) const {
    const int64_t p = base + t;
    if (c) {
      nextcarry[k] = index ? (int64_t)index[p] : p;
      if (positions != nextcarry) positions[k] = p;
      outindex[p] = k;
    } else {
      outindex[p] = -1;
    }
    if (t == n - 1) *total = k + c;
  }
};

// IndexedArray_getitem_nextcarry_outindex: an index at or past lencontent is
// an error (the first in loop order is reported); a negative one is null.
template <typename C>
struct IndexedOutindexBody {
  int64_t* tocarry;
  C* toindex;
  const C* fromindex;
  int64_t lencontent;
  uint64_t* err;
  uint32_t inv;
  __device__ int64_t count(int64_t i) const {
    const int64_t j = (int64_t)fromindex[i];
    return (j >= 0 && j < lencontent) ? 1 : 0;
  }
  __device__ void emit(int64_t i, int64_t k, int64_t c) const {
    const int64_t j = (int64_t)fromindex[i];
    if (j >= lencontent) {
      record_error(err, inv, i, 1);
    } else if (c == 0) {
      toindex[i] = (C)-1;
    } else {
      tocarry[k] = j;
      toindex[i] = (C)k;
    }
  }
};

// ---- small kernels ---------------------------------------------------------------
__global__ void set_i64(int64_t* p, int64_t v) { *p = v; }

// nulls among [0, n) into *out (zeroed first by the caller)
template <class V>
__global__ __launch_bounds__(kBlock) void count_nulls_kernel(V valid, int64_t n, int64_t* out) {
  __shared__ int64_t s_w[kBlock / kWave];
  int64_t s = 0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    s += valid(i) ? 0 : 1;
  s = wave_sum(s);
  if ((threadIdx.x & (kWave - 1)) == 0) s_w[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t t = 0;
    for (int w = 0; w < kBlock / kWave; w++) t += s_w[w];
    if (t != 0) atomicAdd((unsigned long long*)out, (unsigned long long)t);
  }
}

// nextoffsets[0] = 0; nextoffsets[b + 1] = valid positions below offsets[b + 1]
__global__ __launch_bounds__(kBlock) void bin_counts_kernel(int64_t* __restrict__ nextoffsets,
                                                            const int64_t* __restrict__ offsets,
                                                            const int64_t* __restrict__ positions,
                                                            const int64_t* __restrict__ total, int64_t outlength) {
  const int64_t b = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (b == 0) nextoffsets[0] = 0;
  if (b >= outlength) return;
  const int64_t key = offsets[b + 1];
  int64_t lo = 0, hi = *total;  // first position >= key
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (positions[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  nextoffsets[b + 1] = lo;
}

__global__ __launch_bounds__(kBlock) void adjust_shifts_kernel(int64_t* __restrict__ toptr, int64_t outlength,
                                                               const int64_t* __restrict__ starts,
                                                               const int64_t* __restrict__ shifts) {
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k >= outlength) return;
  const int64_t i = toptr[k];
  if (i >= 0) toptr[k] = i + shifts[i] - starts[k];
}

// ---- drivers ---------------------------------------------------------------------
template <class Body>
Error scan_elements(const Body& body, int64_t n, uint32_t inv, void* scan_scratch) {
  return launch_lookback_scan(body, n, inv, scan_scratch);
}

inline size_t round256(size_t b) { return (b + 255) / 256 * 256; }

template <class Body>
Error compact(const Body& body, int64_t n, uint32_t inv = 0) {
  if (n <= 0) return success();
  void* s = scratch(scan_scratch_bytes(scan_tiles(n)));
  if (!s) return failure("out of device memory for scan scratch", kSliceNone, kSliceNone, kFile);
  return scan_elements(body, n, inv, s);
}

template <class V>
Error numnull(int64_t* out, V valid, int64_t n) {
  hipStream_t st = stream();
  set_i64<<<1, 1, 0, st>>>(out, 0);
  if (n > 0) count_nulls_kernel<V><<<grid_for(n, kBlock, 8), kBlock, 0, st>>>(valid, n, out);
  AKHIP_LAUNCH_CHECK("count_nulls_kernel");
  return success();
}

// The element range [offsets[0], offsets[outlength]) is read from the device
// (it sizes the scan), as the CPU loop's bounds are.
template <class V, typename C>
Error reduce_next(int64_t* nextcarry, int64_t* nextoffsets, int64_t* outindex, const C* index, V valid,
                  const int64_t* offsets, int64_t outlength) {
  hipStream_t st = stream();
  if (outlength <= 0) {
    set_i64<<<1, 1, 0, st>>>(nextoffsets, 0);
    AKHIP_LAUNCH_CHECK("set_i64");
    return success();
  }
  const int64_t base = read_device(offsets, 0);
  const int64_t n = read_device(offsets, outlength) - base;
  const int64_t items = n > 0 ? n : 0;
  // scratch: [total | positions (only for an index carry) | scan state]
  const size_t pos_bytes = index ? round256((size_t)items * sizeof(int64_t)) : 0;
  const size_t scan_bytes = scan_scratch_bytes(scan_tiles(items));
  char* s = (char*)scratch(256 + pos_bytes + scan_bytes);
  if (!s) return failure("out of device memory for option-type scratch", kSliceNone, kSliceNone, kFile);
  int64_t* total = (int64_t*)s;
  int64_t* positions = index ? (int64_t*)(s + 256) : nextcarry;
  set_i64<<<1, 1, 0, st>>>(total, 0);
  if (items > 0) {
    ReduceNextBody<V, C> body{nextcarry, outindex, positions, total, index, base, items, valid};
    Error e = scan_elements(body, items, 0, s + 256 + pos_bytes);
    if (e.str) return e;
  }
  bin_counts_kernel<<<(unsigned)ceil_div(outlength, (int64_t)kBlock), kBlock, 0, st>>>(nextoffsets, offsets, positions,
                                                                                      total, outlength);
  AKHIP_LAUNCH_CHECK("bin_counts_kernel");
  return success();
}

template <typename C>
Error indexed_nextcarry_outindex(int64_t* tocarry, C* toindex, const C* fromindex, int64_t lenindex,
                                 int64_t lencontent) {
  if (lenindex <= 0) return success();
  AKHIP_ERROR_WORD(ew);
  const uint32_t inv = next_invocation("awkward_IndexedArray_getitem_nextcarry_outindex", kIndexMsgs);
  IndexedOutindexBody<C> body{tocarry, toindex, fromindex, lencontent, ew, inv};
  Error e = compact(body, lenindex, inv);
  if (e.str) return e;
  return finish_errors(kIndexMsgs, kFile,
                       [&](int64_t i, int) -> int64_t { return (int64_t)read_device(fromindex, i); });
}

}  // namespace
}  // namespace akhip

using namespace akhip;

extern "C" {

Error awkward_ByteMaskedArray_numnull(int64_t* numnull, const int8_t* mask, int64_t length, bool validwhen) {
  return akhip::numnull(numnull, MaskValid{mask, validwhen}, length);
}

Error awkward_ByteMaskedArray_reduce_next_64(int64_t* nextcarry, int64_t* nextoffsets, int64_t* outindex,
                                             const int8_t* mask, const int64_t* offsets, int64_t outlength,
                                             bool validwhen) {
  return reduce_next<MaskValid, int64_t>(nextcarry, nextoffsets, outindex, nullptr, MaskValid{mask, validwhen},
                                         offsets, outlength);
}

Error awkward_ByteMaskedArray_reduce_next_nonlocal_nextshifts_64(int64_t* nextshifts, const int8_t* mask,
                                                                 int64_t length, bool valid_when) {
  return compact(ShiftsBody<MaskValid>{nextshifts, nullptr, MaskValid{mask, valid_when}}, length);
}

Error awkward_ByteMaskedArray_reduce_next_nonlocal_nextshifts_fromshifts_64(int64_t* nextshifts, const int8_t* mask,
                                                                           int64_t length, bool valid_when,
                                                                           const int64_t* shifts) {
  return compact(ShiftsBody<MaskValid>{nextshifts, shifts, MaskValid{mask, valid_when}}, length);
}

Error awkward_ByteMaskedArray_getitem_nextcarry_64(int64_t* tocarry, const int8_t* mask, int64_t length,
                                                   bool validwhen) {
  return compact(NextcarryBody<MaskValid>{tocarry, nullptr, MaskValid{mask, validwhen}}, length);
}

Error awkward_ByteMaskedArray_getitem_nextcarry_outindex_64(int64_t* tocarry, int64_t* outindex, const int8_t* mask,
                                                            int64_t length, bool validwhen) {
  return compact(NextcarryBody<MaskValid>{tocarry, outindex, MaskValid{mask, validwhen}}, length);
}

Error awkward_NumpyArray_reduce_adjust_starts_shifts_64(int64_t* toptr, int64_t outlength, const int64_t*,
                                                        const int64_t* starts, const int64_t* shifts) {
  if (outlength <= 0) return success();
  adjust_shifts_kernel<<<(unsigned)ceil_div(outlength, (int64_t)kBlock), kBlock, 0, stream()>>>(toptr, outlength,
                                                                                               starts, shifts);
  AKHIP_LAUNCH_CHECK("adjust_shifts_kernel");
  return success();
}

#define INDEXED(P, C)                                                                                            \
  Error awkward_IndexedArray##P##_numnull(int64_t* n, const C* fi, int64_t li) {                                 \
    return akhip::numnull(n, IndexValid<C>{fi}, li);                                                            \
  }                                                                                                            \
  Error awkward_IndexedArray##P##_reduce_next_64(int64_t* nc, int64_t* no, int64_t* oi, const C* ix,             \
                                                 const int64_t* off, int64_t ol) {                              \
    return reduce_next<IndexValid<C>, C>(nc, no, oi, ix, IndexValid<C>{ix}, off, ol);                           \
  }                                                                                                            \
  Error awkward_IndexedArray##P##_reduce_next_nonlocal_nextshifts_64(int64_t* ns, const C* ix, int64_t len) {    \
    return compact(ShiftsBody<IndexValid<C>>{ns, nullptr, IndexValid<C>{ix}}, len);                              \
  }                                                                                                            \
  Error awkward_IndexedArray##P##_reduce_next_nonlocal_nextshifts_fromshifts_64(int64_t* ns, const C* ix,        \
                                                                               int64_t len, const int64_t* sh) { \
    return compact(ShiftsBody<IndexValid<C>>{ns, sh, IndexValid<C>{ix}}, len);                                   \
  }                                                                                                            \
  Error awkward_IndexedArray##P##_getitem_nextcarry_outindex_64(int64_t* tc, C* ti, const C* fi, int64_t li,     \
                                                                int64_t lc) {                                   \
    return indexed_nextcarry_outindex<C>(tc, ti, fi, li, lc);                                                   \
  }
INDEXED(32, int32_t)
INDEXED(U32, uint32_t)
INDEXED(64, int64_t)
#undef INDEXED

}  // extern "C"
