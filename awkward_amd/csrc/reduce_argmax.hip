// reduce_argmax.hip -- awkward_reduce_argmax_<T>_64 (awkward-cpp/src/cpu-kernels/awkward_reduce_argmax.cpp),
// plus awkward_hip_reduce_argmax_local_<T>: the same reduction with an explicit tile variant whose
// epilogue also makes the index local (awkward_NumpyArray_reduce_adjust_starts_64 with
// starts = offsets[:-1]) and writes the empty-list mask, and awkward_hip_reduce_argmax_shifts_<T>:
// the same below an option type (awkward_NumpyArray_reduce_adjust_starts_shifts_64 fused); no
// counterpart in awkward-cpp.

#include "reduce.h"

using namespace akhip;

extern "C" {

#define ARG(TN, T)                                                                                          \
  Error awkward_reduce_argmax_##TN##_64(int64_t* toptr, const T* fromptr, const int64_t* offsets,            \
                                        const int64_t*, int64_t outlength) {                               \
    return launch_reduce(ArgOp<T, true>{}, toptr, fromptr, offsets, outlength);                 \
  }                                                                                                        \
  Error awkward_hip_reduce_argmax_local_##TN(int64_t* toptr, int8_t* mask, const T* fromptr,                 \
                                          const int64_t* offsets, int64_t outlength, int variant) {        \
    return launch_reduce(ArgOp<T, true>{}, fromptr, offsets, outlength, variant, \
                         LocalArgEpilogue{toptr, mask});                                                          \
  }                                                                                                        \
  Error awkward_hip_reduce_argmax_shifts_##TN(int64_t* toptr, int8_t* mask, const T* fromptr,                \
                                           const int64_t* offsets, const int64_t* starts,                    \
                                           const int64_t* shifts, int64_t outlength, int variant) {          \
    return launch_reduce(ArgOp<T, true>{}, fromptr, offsets, outlength, variant,                              \
                         LocalArgEpilogue{toptr, mask, shifts, starts});                                             \
  }
ARG(int8, int8_t)
ARG(uint8, uint8_t)
ARG(int16, int16_t)
ARG(uint16, uint16_t)
ARG(int32, int32_t)
ARG(uint32, uint32_t)
ARG(int64, int64_t)
ARG(uint64, uint64_t)
ARG(float32, float)
ARG(float64, double)
#undef ARG

}  // extern "C"
