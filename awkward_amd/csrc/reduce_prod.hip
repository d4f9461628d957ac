// reduce_prod.hip -- awkward_reduce_prod_<OUT>_<IN>_64, awkward_reduce_sum_bool_<T>_64 and
// awkward_reduce_prod_bool_<T>_64 (awkward-cpp/src/cpu-kernels/awkward_reduce_prod.cpp,
// awkward_reduce_sum_bool.cpp, awkward_reduce_prod_bool.cpp): ak.prod, ak.any / ak.all, and
// ak.max / ak.min of booleans.  Plus the tuned forms with an explicit tile variant
// (awkward_hip_reduce_prod_<OUT>_<IN>) and, for any/all, the empty-list mask of
// mask_identity=True written in the same pass (awkward_hip_reduce_{any,all}_masked_<T>);
// those have no counterpart in awkward-cpp.

#include "reduce.h"

using namespace akhip;

extern "C" {

#define PROD(OUTN, INN, OUT, IN)                                                                          \
  Error awkward_reduce_prod_##OUTN##_##INN##_64(OUT* toptr, const IN* fromptr, const int64_t* offsets,    \
                                                int64_t outlength) {                                      \
    return launch_reduce(ProdOp<OUT, IN>{}, toptr, fromptr, offsets, outlength);                          \
  }                                                                                                       \
  Error awkward_hip_reduce_prod_##OUTN##_##INN(OUT* toptr, const IN* fromptr, const int64_t* offsets,     \
                                               int64_t outlength, int variant) {                          \
    return launch_reduce(ProdOp<OUT, IN>{}, fromptr, offsets, outlength, variant, PlainEpilogue<OUT>{toptr}); \
  }
PROD(int64, int8, int64_t, int8_t)
PROD(uint64, uint8, uint64_t, uint8_t)
PROD(int64, int16, int64_t, int16_t)
PROD(uint64, uint16, uint64_t, uint16_t)
PROD(int64, int32, int64_t, int32_t)
PROD(uint64, uint32, uint64_t, uint32_t)
PROD(int64, int64, int64_t, int64_t)
PROD(uint64, uint64, uint64_t, uint64_t)
PROD(float32, float32, float, float)
PROD(float64, float64, double, double)
PROD(int32, int8, int32_t, int8_t)
PROD(uint32, uint8, uint32_t, uint8_t)
PROD(int32, int16, int32_t, int16_t)
PROD(uint32, uint16, uint32_t, uint16_t)
PROD(int32, int32, int32_t, int32_t)
PROD(uint32, uint32, uint32_t, uint32_t)
#undef PROD

#define ANYALL(TN, T)                                                                                     \
  Error awkward_reduce_sum_bool_##TN##_64(bool* toptr, const T* fromptr, const int64_t* offsets,          \
                                          int64_t outlength) {                                            \
    return launch_reduce(AnyAllOp<T, true>{}, toptr, fromptr, offsets, outlength);                        \
  }                                                                                                       \
  Error awkward_reduce_prod_bool_##TN##_64(bool* toptr, const T* fromptr, const int64_t* offsets,         \
                                           int64_t outlength) {                                           \
    return launch_reduce(AnyAllOp<T, false>{}, toptr, fromptr, offsets, outlength);                       \
  }                                                                                                       \
  Error awkward_hip_reduce_any_masked_##TN(bool* toptr, int8_t* mask, const T* fromptr,                   \
                                           const int64_t* offsets, int64_t outlength, int variant) {      \
    return launch_reduce(AnyAllOp<T, true>{}, fromptr, offsets, outlength, variant,                       \
                         MaskEpilogue<bool>{toptr, mask});                                                \
  }                                                                                                       \
  Error awkward_hip_reduce_all_masked_##TN(bool* toptr, int8_t* mask, const T* fromptr,                   \
                                           const int64_t* offsets, int64_t outlength, int variant) {      \
    return launch_reduce(AnyAllOp<T, false>{}, fromptr, offsets, outlength, variant,                      \
                         MaskEpilogue<bool>{toptr, mask});                                                \
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

}  // extern "C"
