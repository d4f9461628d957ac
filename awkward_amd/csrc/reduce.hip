// reduce.hip -- C-ABI entry points of the segmented reductions and their
// small companions (count, adjust_starts, reduce_mask), plus the tuned
// variants the batch runner and the dispatcher select among.
//
// Replaces awkward-cpp/src/cpu-kernels/awkward_reduce_{sum,sum_int64_bool_64,
// sum_int32_bool_64,max,min,argmax,argmin,count_64}.cpp,
// awkward_NumpyArray_reduce_adjust_starts_64.cpp and
// awkward_NumpyArray_reduce_mask_ByteMaskedArray_64.cpp.

#include "reduce.h"

namespace akhip {

template <class Op, int BLOCK, int CAP, class Epi>
static Error launch_reduce_v(const Op& op, const Epi& epi, typename Op::Out* toptr, const typename Op::In* fromptr,
                             const int64_t* offsets, int64_t outlength) {
  if (outlength <= 0) return success();
  const unsigned grid = (unsigned)ceil_div(outlength, BLOCK);
  segmented_reduce_kernel<Op, BLOCK, CAP, Epi><<<grid, BLOCK, 0, stream()>>>(op, epi, toptr, fromptr, offsets, outlength);
  AKHIP_LAUNCH_CHECK("segmented_reduce_kernel");
  return success();
}

// Variant table (tile shapes).  Variant 0 is the fixed default of the plain
// entry points; the others are the autotuner's search space.
constexpr int kDefaultBlock = 256;
constexpr int kDefaultCap = 32768;

template <class Op>
static Error launch_reduce(const Op& op, typename Op::Out* toptr, const typename Op::In* fromptr,
                           const int64_t* offsets, int64_t outlength) {
  return launch_reduce_v<Op, kDefaultBlock, kDefaultCap>(op, PlainEpilogue{}, toptr, fromptr, offsets, outlength);
}

template <class Op, class Epi>
static Error launch_reduce(const Op& op, typename Op::Out* toptr, const typename Op::In* fromptr,
                           const int64_t* offsets, int64_t outlength, int variant, Epi epi) {
  switch (variant) {
    case 1: return launch_reduce_v<Op, 64, 8192>(op, epi, toptr, fromptr, offsets, outlength);
    case 2: return launch_reduce_v<Op, 128, 16384>(op, epi, toptr, fromptr, offsets, outlength);
    case 3: return launch_reduce_v<Op, 256, 16384>(op, epi, toptr, fromptr, offsets, outlength);
    case 4: return launch_reduce_v<Op, 256, 65536>(op, epi, toptr, fromptr, offsets, outlength);
    case 5: return launch_reduce_v<Op, 128, 4096>(op, epi, toptr, fromptr, offsets, outlength);
    default: return launch_reduce_v<Op, kDefaultBlock, kDefaultCap>(op, epi, toptr, fromptr, offsets, outlength);
  }
}

// ---- elementwise companions -------------------------------------------------
__global__ void count_kernel(int64_t* __restrict__ toptr, const int64_t* __restrict__ offsets, int64_t n) {
  for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < n; b += (int64_t)gridDim.x * blockDim.x)
    toptr[b] = offsets[b + 1] - offsets[b];
}

__global__ void adjust_starts_kernel(int64_t* __restrict__ toptr, const int64_t* __restrict__ starts, int64_t n) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = toptr[k];
    if (i >= 0) toptr[k] = i - starts[k];
  }
}

__global__ void reduce_mask_kernel(int8_t* __restrict__ toptr, const int64_t* __restrict__ offsets, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    toptr[i] = (offsets[i + 1] - offsets[i] > 0) ? 0 : 1;
}

// Fused epilogue of the host layer's local (axis=-1) arg-reductions: the
// global index becomes local (awkward_NumpyArray_reduce_adjust_starts_64 with
// starts = offsets[:-1]) and the empty-list mask is written in the same pass
// (awkward_NumpyArray_reduce_mask_ByteMaskedArray_64).
struct LocalArgEpilogue {
  int8_t* mask;  // may be null
  __device__ int64_t apply(int64_t v, int64_t bin, int64_t gb, int64_t ge) const {
    if (mask) mask[bin] = (ge - gb > 0) ? 0 : 1;
    return v >= 0 ? v - gb : v;
  }
};

// Fused mask for value reductions (max/min with mask_identity=True).
struct MaskEpilogue {
  int8_t* mask;
  template <typename Out>
  __device__ Out apply(Out v, int64_t bin, int64_t gb, int64_t ge) const {
    if (mask) mask[bin] = (ge - gb > 0) ? 0 : 1;
    return v;
  }
};

}  // namespace akhip

using namespace akhip;

extern "C" {

#define SUM(NAME, OUT, IN)                                                                    \
  Error NAME(OUT* toptr, const IN* fromptr, const int64_t* offsets, int64_t outlength) {      \
    return launch_reduce(SumOp<OUT, IN>{}, toptr, fromptr, offsets, outlength);               \
  }
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

Error awkward_reduce_sum_int64_bool_64(int64_t* toptr, const bool* fromptr, const int64_t* offsets, int64_t outlength) {
  return launch_reduce(SumBoolOp<int64_t>{}, toptr, fromptr, offsets, outlength);
}
Error awkward_reduce_sum_int32_bool_64(int32_t* toptr, const bool* fromptr, const int64_t* offsets, int64_t outlength) {
  return launch_reduce(SumBoolOp<int32_t>{}, toptr, fromptr, offsets, outlength);
}

#define EXTREMA(TN, T)                                                                                         \
  Error awkward_reduce_max_##TN##_##TN##_64(T* toptr, const T* fromptr, const int64_t* offsets,                \
                                            int64_t outlength, T identity) {                                   \
    return launch_reduce(ExtremumOp<T, true>{identity}, toptr, fromptr, offsets, outlength);                   \
  }                                                                                                            \
  Error awkward_reduce_min_##TN##_##TN##_64(T* toptr, const T* fromptr, const int64_t* offsets,                \
                                            int64_t outlength, T identity) {                                   \
    return launch_reduce(ExtremumOp<T, false>{identity}, toptr, fromptr, offsets, outlength);                  \
  }                                                                                                            \
  Error awkward_reduce_argmax_##TN##_64(int64_t* toptr, const T* fromptr, const int64_t* offsets,              \
                                        const int64_t*, int64_t outlength) {                                   \
    return launch_reduce(ArgOp<T, true>{}, toptr, fromptr, offsets, outlength);                                \
  }                                                                                                            \
  Error awkward_reduce_argmin_##TN##_64(int64_t* toptr, const T* fromptr, const int64_t* offsets,              \
                                        const int64_t*, int64_t outlength) {                                   \
    return launch_reduce(ArgOp<T, false>{}, toptr, fromptr, offsets, outlength);                               \
  }                                                                                                            \
  /* host-layer fused forms (no counterpart in awkward-cpp): local index + mask */                            \
  Error awkward_hip_reduce_argmax_local_##TN(int64_t* toptr, int8_t* mask, const T* fromptr,                   \
                                             const int64_t* offsets, int64_t outlength, int variant) {        \
    return launch_reduce(ArgOp<T, true>{}, toptr, fromptr, offsets, outlength, variant, LocalArgEpilogue{mask}); \
  }                                                                                                            \
  Error awkward_hip_reduce_argmin_local_##TN(int64_t* toptr, int8_t* mask, const T* fromptr,                   \
                                             const int64_t* offsets, int64_t outlength, int variant) {        \
    return launch_reduce(ArgOp<T, false>{}, toptr, fromptr, offsets, outlength, variant, LocalArgEpilogue{mask}); \
  }                                                                                                            \
  Error awkward_hip_reduce_max_masked_##TN(T* toptr, int8_t* mask, const T* fromptr, const int64_t* offsets,   \
                                           int64_t outlength, T identity, int variant) {                      \
    return launch_reduce(ExtremumOp<T, true>{identity}, toptr, fromptr, offsets, outlength, variant,           \
                         MaskEpilogue{mask});                                                                  \
  }                                                                                                            \
  Error awkward_hip_reduce_min_masked_##TN(T* toptr, int8_t* mask, const T* fromptr, const int64_t* offsets,   \
                                           int64_t outlength, T identity, int variant) {                      \
    return launch_reduce(ExtremumOp<T, false>{identity}, toptr, fromptr, offsets, outlength, variant,          \
                         MaskEpilogue{mask});                                                                  \
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

// Tuned sums: the same kernel with an explicit tile variant.
#define SUMV(TN, OUT, IN)                                                                                   \
  Error awkward_hip_reduce_sum_##TN(OUT* toptr, const IN* fromptr, const int64_t* offsets, int64_t outlength, \
                                    int variant) {                                                           \
    return launch_reduce(SumOp<OUT, IN>{}, toptr, fromptr, offsets, outlength, variant, PlainEpilogue{});                     \
  }
SUMV(int64_int32, int64_t, int32_t)
SUMV(int64_int64, int64_t, int64_t)
SUMV(float32_float32, float, float)
SUMV(float64_float64, double, double)
#undef SUMV

Error awkward_reduce_count_64(int64_t* toptr, const int64_t* offsets, int64_t outlength) {
  if (outlength <= 0) return success();
  count_kernel<<<grid_for(outlength, 256), 256, 0, stream()>>>(toptr, offsets, outlength);
  AKHIP_LAUNCH_CHECK("count_kernel");
  return success();
}

Error awkward_NumpyArray_reduce_adjust_starts_64(int64_t* toptr, int64_t outlength, const int64_t*, const int64_t* starts) {
  if (outlength <= 0) return success();
  adjust_starts_kernel<<<grid_for(outlength, 256), 256, 0, stream()>>>(toptr, starts, outlength);
  AKHIP_LAUNCH_CHECK("adjust_starts_kernel");
  return success();
}

Error awkward_NumpyArray_reduce_mask_ByteMaskedArray_64(int8_t* toptr, const int64_t* offsets, int64_t outlength) {
  if (outlength <= 0) return success();
  reduce_mask_kernel<<<grid_for(outlength, 256), 256, 0, stream()>>>(toptr, offsets, outlength);
  AKHIP_LAUNCH_CHECK("reduce_mask_kernel");
  return success();
}

int awkward_hip_reduce_variant_count(void) { return 6; }

}  // extern "C"
