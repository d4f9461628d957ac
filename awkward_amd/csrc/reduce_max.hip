// reduce_max.hip -- awkward_reduce_max_<T>_<T>_64 (awkward-cpp/src/cpu-kernels/awkward_reduce_max.cpp),
// plus awkward_hip_reduce_max_masked_<T>: the same reduction with an explicit tile variant
// that also writes the empty-list mask of mask_identity=True
// (awkward_NumpyArray_reduce_mask_ByteMaskedArray_64) in the same pass; no
// counterpart in awkward-cpp.

#include "reduce.h"

using namespace akhip;

extern "C" {

#define EXT(TN, T)                                                                                          \
  Error awkward_reduce_max_##TN##_##TN##_64(T* toptr, const T* fromptr, const int64_t* offsets,            \
                                            int64_t outlength, T identity) {                               \
    return launch_reduce(ExtremumOp<T, true>{identity}, toptr, fromptr, offsets, outlength);        \
  }                                                                                                        \
  Error awkward_hip_reduce_max_masked_##TN(T* toptr, int8_t* mask, const T* fromptr, const int64_t* offsets, \
                                           int64_t outlength, T identity, int variant) {                   \
    return launch_reduce(ExtremumOp<T, true>{identity}, fromptr, offsets, outlength, variant, \
                         MaskEpilogue<T>{toptr, mask});                                                              \
  }
EXT(int8, int8_t)
EXT(uint8, uint8_t)
EXT(int16, int16_t)
EXT(uint16, uint16_t)
EXT(int32, int32_t)
EXT(uint32, uint32_t)
EXT(int64, int64_t)
EXT(uint64, uint64_t)
EXT(float32, float)
EXT(float64, double)
#undef EXT

}  // extern "C"
