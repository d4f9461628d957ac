// reduce_sum.hip -- awkward_reduce_sum_* and awkward_reduce_sum_{int64,int32}_bool_64
// (awkward-cpp/src/cpu-kernels/awkward_reduce_sum.cpp, awkward_reduce_sum_int64_bool_64.cpp,
// awkward_reduce_sum_int32_bool_64.cpp), plus awkward_hip_reduce_sum_<OUT>_<IN>: the same
// reduction with an explicit tile variant (no counterpart in awkward-cpp).

#include "reduce.h"

using namespace akhip;

extern "C" {

#define SUM(OUTN, INN, OUT, IN)                                                                          \
  Error awkward_reduce_sum_##OUTN##_##INN##_64(OUT* toptr, const IN* fromptr, const int64_t* offsets,    \
                                               int64_t outlength) {                                      \
    return launch_reduce(SumOp<OUT, IN>{}, toptr, fromptr, offsets, outlength);                          \
  }                                                                                                      \
  Error awkward_hip_reduce_sum_##OUTN##_##INN(OUT* toptr, const IN* fromptr, const int64_t* offsets,     \
                                              int64_t outlength, int variant) {                          \
    return launch_reduce(SumOp<OUT, IN>{}, fromptr, offsets, outlength, variant, PlainEpilogue<OUT>{toptr}); \
  }
SUM(int64, int8, int64_t, int8_t)
SUM(uint64, uint8, uint64_t, uint8_t)
SUM(int64, int16, int64_t, int16_t)
SUM(uint64, uint16, uint64_t, uint16_t)
SUM(int64, int32, int64_t, int32_t)
SUM(uint64, uint32, uint64_t, uint32_t)
SUM(int64, int64, int64_t, int64_t)
SUM(uint64, uint64, uint64_t, uint64_t)
SUM(float32, float32, float, float)
SUM(float64, float64, double, double)
SUM(int32, int8, int32_t, int8_t)
SUM(uint32, uint8, uint32_t, uint8_t)
SUM(int32, int16, int32_t, int16_t)
SUM(uint32, uint16, uint32_t, uint16_t)
SUM(int32, int32, int32_t, int32_t)
SUM(uint32, uint32, uint32_t, uint32_t)
SUM(float64, int8, double, int8_t)
SUM(float64, uint8, double, uint8_t)
SUM(float64, int16, double, int16_t)
SUM(float64, uint16, double, uint16_t)
SUM(float64, int32, double, int32_t)
SUM(float64, uint32, double, uint32_t)
SUM(float64, int64, double, int64_t)
SUM(float64, uint64, double, uint64_t)
SUM(float64, float32, double, float)
#undef SUM
#define SUMBOOL(OUTN, OUT)                                                                               \
  Error awkward_reduce_sum_##OUTN##_bool_64(OUT* toptr, const bool* fromptr, const int64_t* offsets,     \
                                            int64_t outlength) {                                         \
    return launch_reduce(SumBoolOp<OUT>{}, toptr, fromptr, offsets, outlength);                          \
  }                                                                                                      \
  Error awkward_hip_reduce_sum_##OUTN##_bool(OUT* toptr, const bool* fromptr, const int64_t* offsets,    \
                                             int64_t outlength, int variant) {                           \
    return launch_reduce(SumBoolOp<OUT>{}, fromptr, offsets, outlength, variant, PlainEpilogue<OUT>{toptr}); \
  }
SUMBOOL(int64, int64_t)
SUMBOOL(int32, int32_t)
#undef SUMBOOL

}  // extern "C"
