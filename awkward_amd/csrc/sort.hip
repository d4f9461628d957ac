// sort.hip -- segmented sort and argsort (awkward_sort_*, awkward_argsort_*),
// and awkward_ListOffsetArray_local_preparenext_64.
//
// Replaces awkward-cpp/src/cpu-kernels/awkward_sort.cpp:23-139,
// awkward_argsort.cpp:23-174 and awkward_ListOffsetArray_local_preparenext_64.cpp.
// The algorithms are in sort.h.

#include "sort.h"

namespace akhip {
namespace {

using namespace sortimpl;
const char* kFile = "awkward_amd/csrc/sort.hip";

template <typename T, typename U, bool ARG>
__global__ __launch_bounds__(kBlock) void segmented_sort_kernel(void* out, const T* __restrict__ from, int64_t length,
                                                                const U* __restrict__ offsets, int64_t nseg,
                                                                bool asc, LongQueue q) {
  sort_tile<T, U, ARG>(out, from, length, offsets, nseg, (int64_t)blockIdx.x * kBlock, asc, q, 0);
}

// Positions outside every segment: argsort leaves the identity index there,
// sort copies the input (awkward_sort.cpp:106-115, awkward_argsort.cpp:70-72).
template <typename T, typename U, bool ARG>
__global__ void uncovered_kernel(void* out, const T* __restrict__ from, int64_t length, const U* __restrict__ offsets,
                                 int64_t offsetslength) {
  int64_t first = offsetslength > 0 ? (int64_t)offsets[0] : length;
  int64_t last = offsetslength > 0 ? (int64_t)offsets[offsetslength - 1] : length;
  if (first > length) first = length;
  if (last > length) last = length;
  if (last < first) last = first;
  // sort copies only positions below offsets[last]
  const int64_t tail_end = ARG ? length : last;
  const int64_t head = first;
  const int64_t tail = tail_end > last ? tail_end - last : 0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < head + tail; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t < head ? t : last + (t - head);
    if constexpr (ARG) ((int64_t*)out)[i] = i;
    else ((T*)out)[i] = from[i];
  }
}

// Scratch layout: 256 bytes of head room (local_preparenext keeps its
// single-segment offsets there), 256 bytes of queue counters, the long-chunk
// queue, then two (key, index) merge buffers over the element positions.
struct SortScratch {
  char* head;
  LongQueue q;
  uint64_t *gk0, *gk1;
  uint32_t *gi0, *gi1;
};

inline bool sort_scratch(int64_t length, SortScratch* out) {
  const size_t kbytes = ((size_t)length * 8 + 255) / 256 * 256;
  const size_t ibytes = ((size_t)length * 4 + 255) / 256 * 256;
  const int64_t capacity = chunk_queue_capacity(length);
  const int64_t small_capacity = small_queue_capacity(length);
  const size_t qbytes = ((size_t)capacity * sizeof(LongChunk) + 255) / 256 * 256;
  const size_t sbytes = ((size_t)small_capacity * sizeof(LongChunk) + 255) / 256 * 256;
  const size_t cbytes = corank_bytes(capacity);
  char* s = (char*)scratch(512 + qbytes + sbytes + cbytes + 2 * (kbytes + ibytes));
  if (!s) return false;
  out->head = s;
  out->q.count = (unsigned long long*)(s + 256);
  out->q.maxlen = (unsigned long long*)(s + 256 + 8);
  out->q.small_count = (unsigned long long*)(s + 256 + 16);
  out->q.chunks = (LongChunk*)(s + 512);
  out->q.capacity = capacity;
  out->q.small = (LongChunk*)(s + 512 + qbytes);
  out->q.small_capacity = small_capacity;
  out->q.coranks = (int64_t*)(s + 512 + qbytes + sbytes);
  char* b = s + 512 + qbytes + sbytes + cbytes;
  out->gk0 = (uint64_t*)b;
  out->gk1 = (uint64_t*)(b + kbytes);
  out->gi0 = (uint32_t*)(b + 2 * kbytes);
  out->gi1 = (uint32_t*)(b + 2 * kbytes + ibytes);
  return true;
}

template <typename T, typename U, bool ARG>
Error segmented_sort(void* out, const T* from, int64_t length, const U* offsets, int64_t offsetslength, bool asc,
                     const SortScratch& sc) {
  if (length <= 0) return success();
  hipStream_t st = stream();
  uncovered_kernel<T, U, ARG><<<grid_for(length, 256, 16), 256, 0, st>>>(out, from, length, offsets, offsetslength);
  const int64_t nseg = offsetslength - 1;
  if (nseg <= 0) {
    AKHIP_LAUNCH_CHECK("uncovered_kernel");
    return success();
  }
  long_queue_reset<><<<1, 1, 0, st>>>(sc.q.count, sc.q.maxlen, sc.q.small_count);
  segmented_sort_kernel<T, U, ARG><<<(unsigned)ceil_div(nseg, kBlock), kBlock, 0, st>>>(out, from, length, offsets,
                                                                                       nseg, asc, sc.q);
  AKHIP_LAUNCH_CHECK("segmented_sort_kernel");
  const hipError_t e = launch_long_sort<T, ARG>(sc.q, length, asc, sc.gk0, sc.gi0, sc.gk1, sc.gi1, st);
  if (e != hipSuccess) return hip_failure(e, "long segment sort");
  return success();
}

template <typename T, typename U, bool ARG>
Error segmented_sort(void* out, const T* from, int64_t length, const U* offsets, int64_t offsetslength, bool asc) {
  if (length <= 0) return success();
  SortScratch sc;
  if (!sort_scratch(length, &sc)) return failure("out of device memory for sort scratch", kSliceNone, kSliceNone, kFile);
  return segmented_sort<T, U, ARG>(out, from, length, offsets, offsetslength, asc, sc);
}

// awkward_ListOffsetArray_local_preparenext_64: the stable argsort of one
// int64 array (a single segment).
__global__ void single_segment_offsets(int64_t* off, int64_t length) {
  off[0] = 0;
  off[1] = length;
}

}  // namespace
}  // namespace akhip

using namespace akhip;

extern "C" {

#define SORT(TN, T)                                                                                      \
  Error awkward_sort_##TN(T* toptr, const T* fromptr, int64_t length, const int64_t* offsets,            \
                          int64_t offsetslength, bool ascending, bool) {                                 \
    return segmented_sort<T, int64_t, false>(toptr, fromptr, length, offsets, offsetslength, ascending); \
  }                                                                                                      \
  Error awkward_argsort_##TN##_int32(int64_t* toptr, const T* fromptr, int64_t length,                   \
                                     const int32_t* offsets, int64_t offsetslength, bool ascending, bool) { \
    return segmented_sort<T, int32_t, true>(toptr, fromptr, length, offsets, offsetslength, ascending);   \
  }                                                                                                      \
  Error awkward_argsort_##TN##_uint32(int64_t* toptr, const T* fromptr, int64_t length,                  \
                                      const uint32_t* offsets, int64_t offsetslength, bool ascending, bool) { \
    return segmented_sort<T, uint32_t, true>(toptr, fromptr, length, offsets, offsetslength, ascending);  \
  }                                                                                                      \
  Error awkward_argsort_##TN##_uint64(int64_t* toptr, const T* fromptr, int64_t length,                  \
                                      const uint64_t* offsets, int64_t offsetslength, bool ascending, bool) { \
    return segmented_sort<T, uint64_t, true>(toptr, fromptr, length, offsets, offsetslength, ascending);  \
  }                                                                                                      \
  Error awkward_argsort_##TN##_int64(int64_t* toptr, const T* fromptr, int64_t length,                   \
                                     const int64_t* offsets, int64_t offsetslength, bool ascending, bool) { \
    return segmented_sort<T, int64_t, true>(toptr, fromptr, length, offsets, offsetslength, ascending);   \
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

Error awkward_ListOffsetArray_local_preparenext_64(int64_t* tocarry, const int64_t* fromindex, int64_t length) {
  if (length <= 0) return success();
  SortScratch sc;
  if (!sort_scratch(length, &sc)) return failure("out of device memory", kSliceNone, kSliceNone, kFile);
  int64_t* off = (int64_t*)sc.head;  // head room holds the offsets pair
  single_segment_offsets<<<1, 1, 0, stream()>>>(off, length);
  return segmented_sort<int64_t, int64_t, true>(tocarry, fromindex, length, off, 2, true, sc);
}

}  // extern "C"
