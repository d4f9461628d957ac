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

}  // namespace akhip

using namespace akhip;

extern "C" {

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

int awkward_hip_reduce_variant_count(void) { return kVariants; }

}  // extern "C"
