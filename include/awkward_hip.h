/* awkward_hip.h -- C ABI of libawkward_hip.so (generated by
 * awkward_amd/_gen_signatures.py from kernel-specification.yml; the runtime
 * section at the end is written by hand in the generator).
 *
 * Every kernel entry point keeps the name, argument order and meaning of the
 * awkward-cpp CPU kernel it replaces, and returns the same struct Error by
 * value (awkward-cpp/include/awkward/common.h:39-78: str == NULL means
 * success; identity / attempt carry the failing position, kSliceNone when
 * absent).  The differences from the CPU ABI:
 *
 *   - every pointer argument is a DEVICE pointer on the current HIP device,
 *     including the array of output pointers of the combinations kernel
 *     (tocarry is a device array of n device pointers);
 *   - kernels are enqueued on the stream set by awkward_hip_set_stream()
 *     (default: the null stream) and return before they complete, except
 *     where a result is needed on the host to size an allocation (none of
 *     the entry points here do that: the *_carrylen / *_carrylength /
 *     totallen outputs stay in device memory, as in the reference);
 *   - data-dependent errors (bad indices, stops < starts) are detected on
 *     the device.  In the default synchronous-error mode the entry point
 *     waits for its kernels and returns the first error in the CPU kernel's
 *     loop order, with the same message, identity and attempt.  In deferred
 *     mode (awkward_hip_set_deferred_errors(1)) it returns success at once
 *     and awkward_hip_error_poll() reports the first recorded error later,
 *     like the reference CUDA backend's error word
 *     (src/awkward/_connect/cuda/__init__.py:257-294).
 */
#ifndef AWKWARD_HIP_H_
#define AWKWARD_HIP_H_

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

struct Error {
  const char* str;
  const char* filename;
  int64_t identity;
  int64_t attempt;
};

/* ---- awkward_reduce_sum ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_sum.cpp:34; contract kernel-specification.yml:5037 */
struct Error awkward_reduce_sum_int64_int8_64(int64_t* toptr, const int8_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_sum.cpp:35; contract kernel-specification.yml:5043 */
struct Error awkward_reduce_sum_uint64_uint8_64(uint64_t* toptr, const uint8_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_sum.cpp:36; contract kernel-specification.yml:5049 */
struct Error awkward_reduce_sum_int64_int16_64(int64_t* toptr, const int16_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_sum.cpp:37; contract kernel-specification.yml:5055 */
struct Error awkward_reduce_sum_uint64_uint16_64(uint64_t* toptr, const uint16_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_sum.cpp:38; contract kernel-specification.yml:5061 */
struct Error awkward_reduce_sum_int64_int32_64(int64_t* toptr, const int32_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_sum.cpp:39; contract kernel-specification.yml:5067 */
struct Error awkward_reduce_sum_uint64_uint32_64(uint64_t* toptr, const uint32_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_sum.cpp:40; contract kernel-specification.yml:5073 */
struct Error awkward_reduce_sum_int64_int64_64(int64_t* toptr, const int64_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_sum.cpp:41; contract kernel-specification.yml:5079 */
struct Error awkward_reduce_sum_uint64_uint64_64(uint64_t* toptr, const uint64_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_sum.cpp:42; contract kernel-specification.yml:5085 */
struct Error awkward_reduce_sum_float32_float32_64(float* toptr, const float* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_sum.cpp:43; contract kernel-specification.yml:5091 */
struct Error awkward_reduce_sum_float64_float64_64(double* toptr, const double* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_sum.cpp:44; contract kernel-specification.yml:5097 */
struct Error awkward_reduce_sum_int32_int8_64(int32_t* toptr, const int8_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_sum.cpp:45; contract kernel-specification.yml:5103 */
struct Error awkward_reduce_sum_uint32_uint8_64(uint32_t* toptr, const uint8_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_sum.cpp:46; contract kernel-specification.yml:5109 */
struct Error awkward_reduce_sum_int32_int16_64(int32_t* toptr, const int16_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_sum.cpp:47; contract kernel-specification.yml:5115 */
struct Error awkward_reduce_sum_uint32_uint16_64(uint32_t* toptr, const uint16_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_sum.cpp:48; contract kernel-specification.yml:5121 */
struct Error awkward_reduce_sum_int32_int32_64(int32_t* toptr, const int32_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_sum.cpp:49; contract kernel-specification.yml:5127 */
struct Error awkward_reduce_sum_uint32_uint32_64(uint32_t* toptr, const uint32_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_sum.cpp:50; contract kernel-specification.yml:5133 */
struct Error awkward_reduce_sum_float64_int8_64(double* toptr, const int8_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_sum.cpp:51; contract kernel-specification.yml:5139 */
struct Error awkward_reduce_sum_float64_uint8_64(double* toptr, const uint8_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_sum.cpp:52; contract kernel-specification.yml:5145 */
struct Error awkward_reduce_sum_float64_int16_64(double* toptr, const int16_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_sum.cpp:53; contract kernel-specification.yml:5151 */
struct Error awkward_reduce_sum_float64_uint16_64(double* toptr, const uint16_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_sum.cpp:54; contract kernel-specification.yml:5157 */
struct Error awkward_reduce_sum_float64_int32_64(double* toptr, const int32_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_sum.cpp:55; contract kernel-specification.yml:5163 */
struct Error awkward_reduce_sum_float64_uint32_64(double* toptr, const uint32_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_sum.cpp:56; contract kernel-specification.yml:5169 */
struct Error awkward_reduce_sum_float64_int64_64(double* toptr, const int64_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_sum.cpp:57; contract kernel-specification.yml:5175 */
struct Error awkward_reduce_sum_float64_uint64_64(double* toptr, const uint64_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_sum.cpp:58; contract kernel-specification.yml:5181 */
struct Error awkward_reduce_sum_float64_float32_64(double* toptr, const float* fromptr, const int64_t* offsets, int64_t outlength);

/* ---- awkward_reduce_sum_int64_bool_64 ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_sum_int64_bool_64.cpp:3; contract kernel-specification.yml:5599 */
struct Error awkward_reduce_sum_int64_bool_64(int64_t* toptr, const bool* fromptr, const int64_t* offsets, int64_t outlength);

/* ---- awkward_reduce_sum_int32_bool_64 ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_sum_int32_bool_64.cpp:3; contract kernel-specification.yml:5581 */
struct Error awkward_reduce_sum_int32_bool_64(int32_t* toptr, const bool* fromptr, const int64_t* offsets, int64_t outlength);

/* ---- awkward_reduce_max ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_max.cpp:35; contract kernel-specification.yml:4564 */
struct Error awkward_reduce_max_int8_int8_64(int8_t* toptr, const int8_t* fromptr, const int64_t* offsets, int64_t outlength, int8_t identity);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_max.cpp:36; contract kernel-specification.yml:4571 */
struct Error awkward_reduce_max_uint8_uint8_64(uint8_t* toptr, const uint8_t* fromptr, const int64_t* offsets, int64_t outlength, uint8_t identity);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_max.cpp:37; contract kernel-specification.yml:4578 */
struct Error awkward_reduce_max_int16_int16_64(int16_t* toptr, const int16_t* fromptr, const int64_t* offsets, int64_t outlength, int16_t identity);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_max.cpp:38; contract kernel-specification.yml:4585 */
struct Error awkward_reduce_max_uint16_uint16_64(uint16_t* toptr, const uint16_t* fromptr, const int64_t* offsets, int64_t outlength, uint16_t identity);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_max.cpp:39; contract kernel-specification.yml:4592 */
struct Error awkward_reduce_max_int32_int32_64(int32_t* toptr, const int32_t* fromptr, const int64_t* offsets, int64_t outlength, int32_t identity);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_max.cpp:40; contract kernel-specification.yml:4599 */
struct Error awkward_reduce_max_uint32_uint32_64(uint32_t* toptr, const uint32_t* fromptr, const int64_t* offsets, int64_t outlength, uint32_t identity);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_max.cpp:41; contract kernel-specification.yml:4606 */
struct Error awkward_reduce_max_int64_int64_64(int64_t* toptr, const int64_t* fromptr, const int64_t* offsets, int64_t outlength, int64_t identity);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_max.cpp:42; contract kernel-specification.yml:4613 */
struct Error awkward_reduce_max_uint64_uint64_64(uint64_t* toptr, const uint64_t* fromptr, const int64_t* offsets, int64_t outlength, uint64_t identity);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_max.cpp:43; contract kernel-specification.yml:4620 */
struct Error awkward_reduce_max_float32_float32_64(float* toptr, const float* fromptr, const int64_t* offsets, int64_t outlength, float identity);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_max.cpp:44; contract kernel-specification.yml:4627 */
struct Error awkward_reduce_max_float64_float64_64(double* toptr, const double* fromptr, const int64_t* offsets, int64_t outlength, double identity);

/* ---- awkward_reduce_min ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_min.cpp:52; contract kernel-specification.yml:4680 */
struct Error awkward_reduce_min_int8_int8_64(int8_t* toptr, const int8_t* fromptr, const int64_t* offsets, int64_t outlength, int8_t identity);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_min.cpp:53; contract kernel-specification.yml:4687 */
struct Error awkward_reduce_min_uint8_uint8_64(uint8_t* toptr, const uint8_t* fromptr, const int64_t* offsets, int64_t outlength, uint8_t identity);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_min.cpp:54; contract kernel-specification.yml:4694 */
struct Error awkward_reduce_min_int16_int16_64(int16_t* toptr, const int16_t* fromptr, const int64_t* offsets, int64_t outlength, int16_t identity);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_min.cpp:55; contract kernel-specification.yml:4701 */
struct Error awkward_reduce_min_uint16_uint16_64(uint16_t* toptr, const uint16_t* fromptr, const int64_t* offsets, int64_t outlength, uint16_t identity);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_min.cpp:56; contract kernel-specification.yml:4708 */
struct Error awkward_reduce_min_int32_int32_64(int32_t* toptr, const int32_t* fromptr, const int64_t* offsets, int64_t outlength, int32_t identity);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_min.cpp:57; contract kernel-specification.yml:4715 */
struct Error awkward_reduce_min_uint32_uint32_64(uint32_t* toptr, const uint32_t* fromptr, const int64_t* offsets, int64_t outlength, uint32_t identity);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_min.cpp:58; contract kernel-specification.yml:4722 */
struct Error awkward_reduce_min_int64_int64_64(int64_t* toptr, const int64_t* fromptr, const int64_t* offsets, int64_t outlength, int64_t identity);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_min.cpp:59; contract kernel-specification.yml:4729 */
struct Error awkward_reduce_min_uint64_uint64_64(uint64_t* toptr, const uint64_t* fromptr, const int64_t* offsets, int64_t outlength, uint64_t identity);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_min.cpp:60; contract kernel-specification.yml:4736 */
struct Error awkward_reduce_min_float32_float32_64(float* toptr, const float* fromptr, const int64_t* offsets, int64_t outlength, float identity);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_min.cpp:61; contract kernel-specification.yml:4743 */
struct Error awkward_reduce_min_float64_float64_64(double* toptr, const double* fromptr, const int64_t* offsets, int64_t outlength, double identity);

/* ---- awkward_reduce_prod ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_prod.cpp:34; contract kernel-specification.yml:4796 */
struct Error awkward_reduce_prod_int64_int8_64(int64_t* toptr, const int8_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_prod.cpp:35; contract kernel-specification.yml:4802 */
struct Error awkward_reduce_prod_uint64_uint8_64(uint64_t* toptr, const uint8_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_prod.cpp:36; contract kernel-specification.yml:4808 */
struct Error awkward_reduce_prod_int64_int16_64(int64_t* toptr, const int16_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_prod.cpp:37; contract kernel-specification.yml:4814 */
struct Error awkward_reduce_prod_uint64_uint16_64(uint64_t* toptr, const uint16_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_prod.cpp:38; contract kernel-specification.yml:4820 */
struct Error awkward_reduce_prod_int64_int32_64(int64_t* toptr, const int32_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_prod.cpp:39; contract kernel-specification.yml:4826 */
struct Error awkward_reduce_prod_uint64_uint32_64(uint64_t* toptr, const uint32_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_prod.cpp:40; contract kernel-specification.yml:4832 */
struct Error awkward_reduce_prod_int64_int64_64(int64_t* toptr, const int64_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_prod.cpp:41; contract kernel-specification.yml:4838 */
struct Error awkward_reduce_prod_uint64_uint64_64(uint64_t* toptr, const uint64_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_prod.cpp:42; contract kernel-specification.yml:4844 */
struct Error awkward_reduce_prod_float32_float32_64(float* toptr, const float* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_prod.cpp:43; contract kernel-specification.yml:4850 */
struct Error awkward_reduce_prod_float64_float64_64(double* toptr, const double* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_prod.cpp:44; contract kernel-specification.yml:4856 */
struct Error awkward_reduce_prod_int32_int8_64(int32_t* toptr, const int8_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_prod.cpp:45; contract kernel-specification.yml:4862 */
struct Error awkward_reduce_prod_uint32_uint8_64(uint32_t* toptr, const uint8_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_prod.cpp:46; contract kernel-specification.yml:4868 */
struct Error awkward_reduce_prod_int32_int16_64(int32_t* toptr, const int16_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_prod.cpp:47; contract kernel-specification.yml:4874 */
struct Error awkward_reduce_prod_uint32_uint16_64(uint32_t* toptr, const uint16_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_prod.cpp:48; contract kernel-specification.yml:4880 */
struct Error awkward_reduce_prod_int32_int32_64(int32_t* toptr, const int32_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_prod.cpp:49; contract kernel-specification.yml:4886 */
struct Error awkward_reduce_prod_uint32_uint32_64(uint32_t* toptr, const uint32_t* fromptr, const int64_t* offsets, int64_t outlength);

/* ---- awkward_reduce_sum_bool ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_sum_bool.cpp:42; contract kernel-specification.yml:5479 */
struct Error awkward_reduce_sum_bool_bool_64(bool* toptr, const bool* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_sum_bool.cpp:43; contract kernel-specification.yml:5485 */
struct Error awkward_reduce_sum_bool_int8_64(bool* toptr, const int8_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_sum_bool.cpp:44; contract kernel-specification.yml:5491 */
struct Error awkward_reduce_sum_bool_uint8_64(bool* toptr, const uint8_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_sum_bool.cpp:45; contract kernel-specification.yml:5497 */
struct Error awkward_reduce_sum_bool_int16_64(bool* toptr, const int16_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_sum_bool.cpp:46; contract kernel-specification.yml:5503 */
struct Error awkward_reduce_sum_bool_uint16_64(bool* toptr, const uint16_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_sum_bool.cpp:47; contract kernel-specification.yml:5509 */
struct Error awkward_reduce_sum_bool_int32_64(bool* toptr, const int32_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_sum_bool.cpp:48; contract kernel-specification.yml:5515 */
struct Error awkward_reduce_sum_bool_uint32_64(bool* toptr, const uint32_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_sum_bool.cpp:49; contract kernel-specification.yml:5521 */
struct Error awkward_reduce_sum_bool_int64_64(bool* toptr, const int64_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_sum_bool.cpp:50; contract kernel-specification.yml:5527 */
struct Error awkward_reduce_sum_bool_uint64_64(bool* toptr, const uint64_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_sum_bool.cpp:51; contract kernel-specification.yml:5533 */
struct Error awkward_reduce_sum_bool_float32_64(bool* toptr, const float* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_sum_bool.cpp:52; contract kernel-specification.yml:5539 */
struct Error awkward_reduce_sum_bool_float64_64(bool* toptr, const double* fromptr, const int64_t* offsets, int64_t outlength);

/* ---- awkward_reduce_prod_bool ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_prod_bool.cpp:34; contract kernel-specification.yml:4933 */
struct Error awkward_reduce_prod_bool_bool_64(bool* toptr, const bool* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_prod_bool.cpp:35; contract kernel-specification.yml:4939 */
struct Error awkward_reduce_prod_bool_int8_64(bool* toptr, const int8_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_prod_bool.cpp:36; contract kernel-specification.yml:4945 */
struct Error awkward_reduce_prod_bool_uint8_64(bool* toptr, const uint8_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_prod_bool.cpp:37; contract kernel-specification.yml:4951 */
struct Error awkward_reduce_prod_bool_int16_64(bool* toptr, const int16_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_prod_bool.cpp:38; contract kernel-specification.yml:4957 */
struct Error awkward_reduce_prod_bool_uint16_64(bool* toptr, const uint16_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_prod_bool.cpp:39; contract kernel-specification.yml:4963 */
struct Error awkward_reduce_prod_bool_int32_64(bool* toptr, const int32_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_prod_bool.cpp:40; contract kernel-specification.yml:4969 */
struct Error awkward_reduce_prod_bool_uint32_64(bool* toptr, const uint32_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_prod_bool.cpp:41; contract kernel-specification.yml:4975 */
struct Error awkward_reduce_prod_bool_int64_64(bool* toptr, const int64_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_prod_bool.cpp:42; contract kernel-specification.yml:4981 */
struct Error awkward_reduce_prod_bool_uint64_64(bool* toptr, const uint64_t* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_prod_bool.cpp:43; contract kernel-specification.yml:4987 */
struct Error awkward_reduce_prod_bool_float32_64(bool* toptr, const float* fromptr, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_prod_bool.cpp:44; contract kernel-specification.yml:4993 */
struct Error awkward_reduce_prod_bool_float64_64(bool* toptr, const double* fromptr, const int64_t* offsets, int64_t outlength);

/* ---- awkward_reduce_argmax ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_argmax.cpp:43; contract kernel-specification.yml:4227 */
struct Error awkward_reduce_argmax_int8_64(int64_t* toptr, const int8_t* fromptr, const int64_t* offsets, const int64_t* starts, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_argmax.cpp:44; contract kernel-specification.yml:4234 */
struct Error awkward_reduce_argmax_uint8_64(int64_t* toptr, const uint8_t* fromptr, const int64_t* offsets, const int64_t* starts, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_argmax.cpp:45; contract kernel-specification.yml:4241 */
struct Error awkward_reduce_argmax_int16_64(int64_t* toptr, const int16_t* fromptr, const int64_t* offsets, const int64_t* starts, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_argmax.cpp:46; contract kernel-specification.yml:4248 */
struct Error awkward_reduce_argmax_uint16_64(int64_t* toptr, const uint16_t* fromptr, const int64_t* offsets, const int64_t* starts, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_argmax.cpp:47; contract kernel-specification.yml:4255 */
struct Error awkward_reduce_argmax_int32_64(int64_t* toptr, const int32_t* fromptr, const int64_t* offsets, const int64_t* starts, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_argmax.cpp:48; contract kernel-specification.yml:4262 */
struct Error awkward_reduce_argmax_uint32_64(int64_t* toptr, const uint32_t* fromptr, const int64_t* offsets, const int64_t* starts, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_argmax.cpp:49; contract kernel-specification.yml:4269 */
struct Error awkward_reduce_argmax_int64_64(int64_t* toptr, const int64_t* fromptr, const int64_t* offsets, const int64_t* starts, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_argmax.cpp:50; contract kernel-specification.yml:4276 */
struct Error awkward_reduce_argmax_uint64_64(int64_t* toptr, const uint64_t* fromptr, const int64_t* offsets, const int64_t* starts, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_argmax.cpp:51; contract kernel-specification.yml:4283 */
struct Error awkward_reduce_argmax_float32_64(int64_t* toptr, const float* fromptr, const int64_t* offsets, const int64_t* starts, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_argmax.cpp:52; contract kernel-specification.yml:4290 */
struct Error awkward_reduce_argmax_float64_64(int64_t* toptr, const double* fromptr, const int64_t* offsets, const int64_t* starts, int64_t outlength);

/* ---- awkward_reduce_argmin ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_argmin.cpp:59; contract kernel-specification.yml:4338 */
struct Error awkward_reduce_argmin_int8_64(int64_t* toptr, const int8_t* fromptr, const int64_t* offsets, const int64_t* starts, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_argmin.cpp:60; contract kernel-specification.yml:4345 */
struct Error awkward_reduce_argmin_uint8_64(int64_t* toptr, const uint8_t* fromptr, const int64_t* offsets, const int64_t* starts, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_argmin.cpp:61; contract kernel-specification.yml:4352 */
struct Error awkward_reduce_argmin_int16_64(int64_t* toptr, const int16_t* fromptr, const int64_t* offsets, const int64_t* starts, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_argmin.cpp:62; contract kernel-specification.yml:4359 */
struct Error awkward_reduce_argmin_uint16_64(int64_t* toptr, const uint16_t* fromptr, const int64_t* offsets, const int64_t* starts, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_argmin.cpp:63; contract kernel-specification.yml:4366 */
struct Error awkward_reduce_argmin_int32_64(int64_t* toptr, const int32_t* fromptr, const int64_t* offsets, const int64_t* starts, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_argmin.cpp:64; contract kernel-specification.yml:4373 */
struct Error awkward_reduce_argmin_uint32_64(int64_t* toptr, const uint32_t* fromptr, const int64_t* offsets, const int64_t* starts, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_argmin.cpp:65; contract kernel-specification.yml:4380 */
struct Error awkward_reduce_argmin_int64_64(int64_t* toptr, const int64_t* fromptr, const int64_t* offsets, const int64_t* starts, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_argmin.cpp:66; contract kernel-specification.yml:4387 */
struct Error awkward_reduce_argmin_uint64_64(int64_t* toptr, const uint64_t* fromptr, const int64_t* offsets, const int64_t* starts, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_argmin.cpp:67; contract kernel-specification.yml:4394 */
struct Error awkward_reduce_argmin_float32_64(int64_t* toptr, const float* fromptr, const int64_t* offsets, const int64_t* starts, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_argmin.cpp:68; contract kernel-specification.yml:4401 */
struct Error awkward_reduce_argmin_float64_64(int64_t* toptr, const double* fromptr, const int64_t* offsets, const int64_t* starts, int64_t outlength);

/* ---- awkward_reduce_count_64 ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_reduce_count_64.cpp:3; contract kernel-specification.yml:4447 */
struct Error awkward_reduce_count_64(int64_t* toptr, const int64_t* offsets, int64_t outlength);

/* ---- awkward_NumpyArray_reduce_adjust_starts_64 ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_NumpyArray_reduce_adjust_starts_64.cpp:3; contract kernel-specification.yml:2550 */
struct Error awkward_NumpyArray_reduce_adjust_starts_64(int64_t* toptr, int64_t outlength, const int64_t* offsets, const int64_t* starts);

/* ---- awkward_NumpyArray_reduce_mask_ByteMaskedArray_64 ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_NumpyArray_reduce_mask_ByteMaskedArray_64.cpp:3; contract kernel-specification.yml:2585 */
struct Error awkward_NumpyArray_reduce_mask_ByteMaskedArray_64(int8_t* toptr, const int64_t* offsets, int64_t outlength);

/* ---- awkward_localindex ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_localindex.cpp:15; contract kernel-specification.yml:4196 */
struct Error awkward_localindex_64(int64_t* toindex, int64_t length);

/* ---- awkward_ListArray_localindex ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_ListArray_localindex.cpp:27; contract kernel-specification.yml:1921 */
struct Error awkward_ListArray32_localindex_64(int64_t* toindex, const int32_t* offsets, int64_t length);
/* replaces awkward-cpp/src/cpu-kernels/awkward_ListArray_localindex.cpp:29; contract kernel-specification.yml:1926 */
struct Error awkward_ListArray64_localindex_64(int64_t* toindex, const int64_t* offsets, int64_t length);
/* replaces awkward-cpp/src/cpu-kernels/awkward_ListArray_localindex.cpp:28; contract kernel-specification.yml:1931 */
struct Error awkward_ListArrayU32_localindex_64(int64_t* toindex, const uint32_t* offsets, int64_t length);

/* ---- awkward_RegularArray_getitem_jagged_expand ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_RegularArray_getitem_jagged_expand.cpp:22; contract kernel-specification.yml:2912 */
struct Error awkward_RegularArray_getitem_jagged_expand_64(int64_t* multistarts, int64_t* multistops, const int64_t* singleoffsets, int64_t regularsize, int64_t regularlength);

/* ---- awkward_ListArray_getitem_jagged_carrylen ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_ListArray_getitem_jagged_carrylen.cpp:19; contract kernel-specification.yml:1398 */
struct Error awkward_ListArray_getitem_jagged_carrylen_64(int64_t* carrylen, const int64_t* slicestarts, const int64_t* slicestops, int64_t sliceouterlen);

/* ---- awkward_ListArray_getitem_jagged_apply ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_ListArray_getitem_jagged_apply.cpp:62; contract kernel-specification.yml:1318 */
struct Error awkward_ListArray32_getitem_jagged_apply_64(int64_t* tooffsets, int64_t* tocarry, const int64_t* slicestarts, const int64_t* slicestops, int64_t sliceouterlen, const int64_t* sliceindex, int64_t sliceinnerlen, const int32_t* fromstarts, const int32_t* fromstops, int64_t contentlen);
/* replaces awkward-cpp/src/cpu-kernels/awkward_ListArray_getitem_jagged_apply.cpp:64; contract kernel-specification.yml:1330 */
struct Error awkward_ListArray64_getitem_jagged_apply_64(int64_t* tooffsets, int64_t* tocarry, const int64_t* slicestarts, const int64_t* slicestops, int64_t sliceouterlen, const int64_t* sliceindex, int64_t sliceinnerlen, const int64_t* fromstarts, const int64_t* fromstops, int64_t contentlen);
/* replaces awkward-cpp/src/cpu-kernels/awkward_ListArray_getitem_jagged_apply.cpp:63; contract kernel-specification.yml:1342 */
struct Error awkward_ListArrayU32_getitem_jagged_apply_64(int64_t* tooffsets, int64_t* tocarry, const int64_t* slicestarts, const int64_t* slicestops, int64_t sliceouterlen, const int64_t* sliceindex, int64_t sliceinnerlen, const uint32_t* fromstarts, const uint32_t* fromstops, int64_t contentlen);

/* ---- awkward_ListArray_getitem_jagged_descend ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_ListArray_getitem_jagged_descend.cpp:39; contract kernel-specification.yml:1416 */
struct Error awkward_ListArray32_getitem_jagged_descend_64(int64_t* tooffsets, const int64_t* slicestarts, const int64_t* slicestops, int64_t sliceouterlen, const int32_t* fromstarts, const int32_t* fromstops);
/* replaces awkward-cpp/src/cpu-kernels/awkward_ListArray_getitem_jagged_descend.cpp:41; contract kernel-specification.yml:1424 */
struct Error awkward_ListArray64_getitem_jagged_descend_64(int64_t* tooffsets, const int64_t* slicestarts, const int64_t* slicestops, int64_t sliceouterlen, const int64_t* fromstarts, const int64_t* fromstops);
/* replaces awkward-cpp/src/cpu-kernels/awkward_ListArray_getitem_jagged_descend.cpp:40; contract kernel-specification.yml:1432 */
struct Error awkward_ListArrayU32_getitem_jagged_descend_64(int64_t* tooffsets, const int64_t* slicestarts, const int64_t* slicestops, int64_t sliceouterlen, const uint32_t* fromstarts, const uint32_t* fromstops);

/* ---- awkward_ListArray_getitem_next_range ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_ListArray_getitem_next_range.cpp:49; contract kernel-specification.yml:1741 */
struct Error awkward_ListArray32_getitem_next_range_64(int32_t* tooffsets, int64_t* tocarry, const int32_t* fromstarts, const int32_t* fromstops, int64_t lenstarts, int64_t start, int64_t stop, int64_t step);
/* replaces awkward-cpp/src/cpu-kernels/awkward_ListArray_getitem_next_range.cpp:51; contract kernel-specification.yml:1751 */
struct Error awkward_ListArray64_getitem_next_range_64(int64_t* tooffsets, int64_t* tocarry, const int64_t* fromstarts, const int64_t* fromstops, int64_t lenstarts, int64_t start, int64_t stop, int64_t step);
/* replaces awkward-cpp/src/cpu-kernels/awkward_ListArray_getitem_next_range.cpp:50; contract kernel-specification.yml:1761 */
struct Error awkward_ListArrayU32_getitem_next_range_64(uint32_t* tooffsets, int64_t* tocarry, const uint32_t* fromstarts, const uint32_t* fromstops, int64_t lenstarts, int64_t start, int64_t stop, int64_t step);

/* ---- awkward_ListArray_getitem_next_range_carrylength ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_ListArray_getitem_next_range_carrylength.cpp:44; contract kernel-specification.yml:1806 */
struct Error awkward_ListArray32_getitem_next_range_carrylength(int64_t* carrylength, const int32_t* fromstarts, const int32_t* fromstops, int64_t lenstarts, int64_t start, int64_t stop, int64_t step);
/* replaces awkward-cpp/src/cpu-kernels/awkward_ListArray_getitem_next_range_carrylength.cpp:46; contract kernel-specification.yml:1815 */
struct Error awkward_ListArray64_getitem_next_range_carrylength(int64_t* carrylength, const int64_t* fromstarts, const int64_t* fromstops, int64_t lenstarts, int64_t start, int64_t stop, int64_t step);
/* replaces awkward-cpp/src/cpu-kernels/awkward_ListArray_getitem_next_range_carrylength.cpp:45; contract kernel-specification.yml:1824 */
struct Error awkward_ListArrayU32_getitem_next_range_carrylength(int64_t* carrylength, const uint32_t* fromstarts, const uint32_t* fromstops, int64_t lenstarts, int64_t start, int64_t stop, int64_t step);

/* ---- awkward_ListArray_getitem_next_range_counts ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_ListArray_getitem_next_range_counts.cpp:24; contract kernel-specification.yml:1865 */
struct Error awkward_ListArray32_getitem_next_range_counts_64(int64_t* total, const int32_t* fromoffsets, int64_t lenstarts);
/* replaces awkward-cpp/src/cpu-kernels/awkward_ListArray_getitem_next_range_counts.cpp:26; contract kernel-specification.yml:1870 */
struct Error awkward_ListArray64_getitem_next_range_counts_64(int64_t* total, const int64_t* fromoffsets, int64_t lenstarts);
/* replaces awkward-cpp/src/cpu-kernels/awkward_ListArray_getitem_next_range_counts.cpp:25; contract kernel-specification.yml:1875 */
struct Error awkward_ListArrayU32_getitem_next_range_counts_64(int64_t* total, const uint32_t* fromoffsets, int64_t lenstarts);

/* ---- awkward_RegularArray_getitem_next_range ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_RegularArray_getitem_next_range.cpp:22; contract kernel-specification.yml:3014 */
struct Error awkward_RegularArray_getitem_next_range_64(int64_t* tocarry, int64_t regular_start, int64_t step, int64_t length, int64_t size, int64_t nextsize);

/* ---- awkward_ListArray_getitem_next_at ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_ListArray_getitem_next_at.cpp:33; contract kernel-specification.yml:1705 */
struct Error awkward_ListArray32_getitem_next_at_64(int64_t* tocarry, const int32_t* fromstarts, const int32_t* fromstops, int64_t lenstarts, int64_t at);
/* replaces awkward-cpp/src/cpu-kernels/awkward_ListArray_getitem_next_at.cpp:35; contract kernel-specification.yml:1712 */
struct Error awkward_ListArray64_getitem_next_at_64(int64_t* tocarry, const int64_t* fromstarts, const int64_t* fromstops, int64_t lenstarts, int64_t at);
/* replaces awkward-cpp/src/cpu-kernels/awkward_ListArray_getitem_next_at.cpp:34; contract kernel-specification.yml:1719 */
struct Error awkward_ListArrayU32_getitem_next_at_64(int64_t* tocarry, const uint32_t* fromstarts, const uint32_t* fromstops, int64_t lenstarts, int64_t at);

/* ---- awkward_ListArray_compact_offsets ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_ListArray_compact_offsets.cpp:30; contract kernel-specification.yml:1237 */
struct Error awkward_ListArray32_compact_offsets_64(int64_t* tooffsets, const int32_t* fromstarts, const int32_t* fromstops, int64_t length);
/* replaces awkward-cpp/src/cpu-kernels/awkward_ListArray_compact_offsets.cpp:32; contract kernel-specification.yml:1243 */
struct Error awkward_ListArray64_compact_offsets_64(int64_t* tooffsets, const int64_t* fromstarts, const int64_t* fromstops, int64_t length);
/* replaces awkward-cpp/src/cpu-kernels/awkward_ListArray_compact_offsets.cpp:31; contract kernel-specification.yml:1249 */
struct Error awkward_ListArrayU32_compact_offsets_64(int64_t* tooffsets, const uint32_t* fromstarts, const uint32_t* fromstops, int64_t length);

/* ---- awkward_ListArray_combinations_length ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_ListArray_combinations_length.cpp:52; contract kernel-specification.yml:1177 */
struct Error awkward_ListArray32_combinations_length_64(int64_t* totallen, int64_t* tooffsets, int64_t n, bool replacement, const int32_t* starts, const int32_t* stops, int64_t length);
/* replaces awkward-cpp/src/cpu-kernels/awkward_ListArray_combinations_length.cpp:54; contract kernel-specification.yml:1186 */
struct Error awkward_ListArray64_combinations_length_64(int64_t* totallen, int64_t* tooffsets, int64_t n, bool replacement, const int64_t* starts, const int64_t* stops, int64_t length);
/* replaces awkward-cpp/src/cpu-kernels/awkward_ListArray_combinations_length.cpp:53; contract kernel-specification.yml:1195 */
struct Error awkward_ListArrayU32_combinations_length_64(int64_t* totallen, int64_t* tooffsets, int64_t n, bool replacement, const uint32_t* starts, const uint32_t* stops, int64_t length);

/* ---- awkward_ListArray_combinations ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_ListArray_combinations.cpp:42; contract kernel-specification.yml:1127 */
struct Error awkward_ListArray32_combinations_64(int64_t** tocarry, int64_t* toindex, int64_t* fromindex, int64_t n, bool replacement, const int32_t* starts, const int32_t* stops, int64_t length);
/* replaces awkward-cpp/src/cpu-kernels/awkward_ListArray_combinations.cpp:44; contract kernel-specification.yml:1137 */
struct Error awkward_ListArray64_combinations_64(int64_t** tocarry, int64_t* toindex, int64_t* fromindex, int64_t n, bool replacement, const int64_t* starts, const int64_t* stops, int64_t length);
/* replaces awkward-cpp/src/cpu-kernels/awkward_ListArray_combinations.cpp:43; contract kernel-specification.yml:1147 */
struct Error awkward_ListArrayU32_combinations_64(int64_t** tocarry, int64_t* toindex, int64_t* fromindex, int64_t n, bool replacement, const uint32_t* starts, const uint32_t* stops, int64_t length);

/* ---- awkward_IndexedArray_getitem_nextcarry ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_IndexedArray_getitem_nextcarry.cpp:32; contract kernel-specification.yml:474 */
struct Error awkward_IndexedArray32_getitem_nextcarry_64(int64_t* tocarry, const int32_t* fromindex, int64_t lenindex, int64_t lencontent);
/* replaces awkward-cpp/src/cpu-kernels/awkward_IndexedArray_getitem_nextcarry.cpp:34; contract kernel-specification.yml:480 */
struct Error awkward_IndexedArray64_getitem_nextcarry_64(int64_t* tocarry, const int64_t* fromindex, int64_t lenindex, int64_t lencontent);
/* replaces awkward-cpp/src/cpu-kernels/awkward_IndexedArray_getitem_nextcarry.cpp:33; contract kernel-specification.yml:486 */
struct Error awkward_IndexedArrayU32_getitem_nextcarry_64(int64_t* tocarry, const uint32_t* fromindex, int64_t lenindex, int64_t lencontent);

/* ---- awkward_ListOffsetArray_local_preparenext_64 ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_ListOffsetArray_local_preparenext_64.cpp:3; contract kernel-specification.yml:2151 */
struct Error awkward_ListOffsetArray_local_preparenext_64(int64_t* tocarry, const int64_t* fromindex, int64_t length);

/* ---- awkward_sort ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_sort.cpp:129; contract kernel-specification.yml:5619 */
struct Error awkward_sort_bool(bool* toptr, const bool* fromptr, int64_t length, const int64_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_sort.cpp:130; contract kernel-specification.yml:5628 */
struct Error awkward_sort_int8(int8_t* toptr, const int8_t* fromptr, int64_t length, const int64_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_sort.cpp:132; contract kernel-specification.yml:5637 */
struct Error awkward_sort_int16(int16_t* toptr, const int16_t* fromptr, int64_t length, const int64_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_sort.cpp:134; contract kernel-specification.yml:5646 */
struct Error awkward_sort_int32(int32_t* toptr, const int32_t* fromptr, int64_t length, const int64_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_sort.cpp:136; contract kernel-specification.yml:5655 */
struct Error awkward_sort_int64(int64_t* toptr, const int64_t* fromptr, int64_t length, const int64_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_sort.cpp:131; contract kernel-specification.yml:5664 */
struct Error awkward_sort_uint8(uint8_t* toptr, const uint8_t* fromptr, int64_t length, const int64_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_sort.cpp:133; contract kernel-specification.yml:5673 */
struct Error awkward_sort_uint16(uint16_t* toptr, const uint16_t* fromptr, int64_t length, const int64_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_sort.cpp:135; contract kernel-specification.yml:5682 */
struct Error awkward_sort_uint32(uint32_t* toptr, const uint32_t* fromptr, int64_t length, const int64_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_sort.cpp:137; contract kernel-specification.yml:5691 */
struct Error awkward_sort_uint64(uint64_t* toptr, const uint64_t* fromptr, int64_t length, const int64_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_sort.cpp:138; contract kernel-specification.yml:5700 */
struct Error awkward_sort_float32(float* toptr, const float* fromptr, int64_t length, const int64_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_sort.cpp:139; contract kernel-specification.yml:5709 */
struct Error awkward_sort_float64(double* toptr, const double* fromptr, int64_t length, const int64_t* offsets, int64_t offsetslength, bool ascending, bool stable);

/* ---- awkward_argsort ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:121; contract kernel-specification.yml:3724 */
struct Error awkward_argsort_bool_int32(int64_t* toptr, const bool* fromptr, int64_t length, const int32_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:122; contract kernel-specification.yml:3733 */
struct Error awkward_argsort_bool_uint32(int64_t* toptr, const bool* fromptr, int64_t length, const uint32_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:123; contract kernel-specification.yml:3742 */
struct Error awkward_argsort_bool_uint64(int64_t* toptr, const bool* fromptr, int64_t length, const uint64_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:124; contract kernel-specification.yml:3751 */
struct Error awkward_argsort_bool_int64(int64_t* toptr, const bool* fromptr, int64_t length, const int64_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:126; contract kernel-specification.yml:3760 */
struct Error awkward_argsort_int8_int32(int64_t* toptr, const int8_t* fromptr, int64_t length, const int32_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:127; contract kernel-specification.yml:3769 */
struct Error awkward_argsort_int8_uint32(int64_t* toptr, const int8_t* fromptr, int64_t length, const uint32_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:128; contract kernel-specification.yml:3778 */
struct Error awkward_argsort_int8_uint64(int64_t* toptr, const int8_t* fromptr, int64_t length, const uint64_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:129; contract kernel-specification.yml:3787 */
struct Error awkward_argsort_int8_int64(int64_t* toptr, const int8_t* fromptr, int64_t length, const int64_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:136; contract kernel-specification.yml:3796 */
struct Error awkward_argsort_int16_int32(int64_t* toptr, const int16_t* fromptr, int64_t length, const int32_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:137; contract kernel-specification.yml:3805 */
struct Error awkward_argsort_int16_uint32(int64_t* toptr, const int16_t* fromptr, int64_t length, const uint32_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:138; contract kernel-specification.yml:3814 */
struct Error awkward_argsort_int16_uint64(int64_t* toptr, const int16_t* fromptr, int64_t length, const uint64_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:139; contract kernel-specification.yml:3823 */
struct Error awkward_argsort_int16_int64(int64_t* toptr, const int16_t* fromptr, int64_t length, const int64_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:146; contract kernel-specification.yml:3832 */
struct Error awkward_argsort_int32_int32(int64_t* toptr, const int32_t* fromptr, int64_t length, const int32_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:147; contract kernel-specification.yml:3841 */
struct Error awkward_argsort_int32_uint32(int64_t* toptr, const int32_t* fromptr, int64_t length, const uint32_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:148; contract kernel-specification.yml:3850 */
struct Error awkward_argsort_int32_uint64(int64_t* toptr, const int32_t* fromptr, int64_t length, const uint64_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:149; contract kernel-specification.yml:3859 */
struct Error awkward_argsort_int32_int64(int64_t* toptr, const int32_t* fromptr, int64_t length, const int64_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:156; contract kernel-specification.yml:3868 */
struct Error awkward_argsort_int64_int32(int64_t* toptr, const int64_t* fromptr, int64_t length, const int32_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:157; contract kernel-specification.yml:3877 */
struct Error awkward_argsort_int64_uint32(int64_t* toptr, const int64_t* fromptr, int64_t length, const uint32_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:158; contract kernel-specification.yml:3886 */
struct Error awkward_argsort_int64_uint64(int64_t* toptr, const int64_t* fromptr, int64_t length, const uint64_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:159; contract kernel-specification.yml:3895 */
struct Error awkward_argsort_int64_int64(int64_t* toptr, const int64_t* fromptr, int64_t length, const int64_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:131; contract kernel-specification.yml:3904 */
struct Error awkward_argsort_uint8_int32(int64_t* toptr, const uint8_t* fromptr, int64_t length, const int32_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:132; contract kernel-specification.yml:3913 */
struct Error awkward_argsort_uint8_uint32(int64_t* toptr, const uint8_t* fromptr, int64_t length, const uint32_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:133; contract kernel-specification.yml:3922 */
struct Error awkward_argsort_uint8_uint64(int64_t* toptr, const uint8_t* fromptr, int64_t length, const uint64_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:134; contract kernel-specification.yml:3931 */
struct Error awkward_argsort_uint8_int64(int64_t* toptr, const uint8_t* fromptr, int64_t length, const int64_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:141; contract kernel-specification.yml:3940 */
struct Error awkward_argsort_uint16_int32(int64_t* toptr, const uint16_t* fromptr, int64_t length, const int32_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:142; contract kernel-specification.yml:3949 */
struct Error awkward_argsort_uint16_uint32(int64_t* toptr, const uint16_t* fromptr, int64_t length, const uint32_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:143; contract kernel-specification.yml:3958 */
struct Error awkward_argsort_uint16_uint64(int64_t* toptr, const uint16_t* fromptr, int64_t length, const uint64_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:144; contract kernel-specification.yml:3967 */
struct Error awkward_argsort_uint16_int64(int64_t* toptr, const uint16_t* fromptr, int64_t length, const int64_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:151; contract kernel-specification.yml:3976 */
struct Error awkward_argsort_uint32_int32(int64_t* toptr, const uint32_t* fromptr, int64_t length, const int32_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:152; contract kernel-specification.yml:3985 */
struct Error awkward_argsort_uint32_uint32(int64_t* toptr, const uint32_t* fromptr, int64_t length, const uint32_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:153; contract kernel-specification.yml:3994 */
struct Error awkward_argsort_uint32_uint64(int64_t* toptr, const uint32_t* fromptr, int64_t length, const uint64_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:154; contract kernel-specification.yml:4003 */
struct Error awkward_argsort_uint32_int64(int64_t* toptr, const uint32_t* fromptr, int64_t length, const int64_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:161; contract kernel-specification.yml:4012 */
struct Error awkward_argsort_uint64_int32(int64_t* toptr, const uint64_t* fromptr, int64_t length, const int32_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:162; contract kernel-specification.yml:4021 */
struct Error awkward_argsort_uint64_uint32(int64_t* toptr, const uint64_t* fromptr, int64_t length, const uint32_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:163; contract kernel-specification.yml:4030 */
struct Error awkward_argsort_uint64_uint64(int64_t* toptr, const uint64_t* fromptr, int64_t length, const uint64_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:164; contract kernel-specification.yml:4039 */
struct Error awkward_argsort_uint64_int64(int64_t* toptr, const uint64_t* fromptr, int64_t length, const int64_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:166; contract kernel-specification.yml:4048 */
struct Error awkward_argsort_float32_int32(int64_t* toptr, const float* fromptr, int64_t length, const int32_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:167; contract kernel-specification.yml:4057 */
struct Error awkward_argsort_float32_uint32(int64_t* toptr, const float* fromptr, int64_t length, const uint32_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:168; contract kernel-specification.yml:4066 */
struct Error awkward_argsort_float32_uint64(int64_t* toptr, const float* fromptr, int64_t length, const uint64_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:169; contract kernel-specification.yml:4075 */
struct Error awkward_argsort_float32_int64(int64_t* toptr, const float* fromptr, int64_t length, const int64_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:171; contract kernel-specification.yml:4084 */
struct Error awkward_argsort_float64_int32(int64_t* toptr, const double* fromptr, int64_t length, const int32_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:172; contract kernel-specification.yml:4093 */
struct Error awkward_argsort_float64_uint32(int64_t* toptr, const double* fromptr, int64_t length, const uint32_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:173; contract kernel-specification.yml:4102 */
struct Error awkward_argsort_float64_uint64(int64_t* toptr, const double* fromptr, int64_t length, const uint64_t* offsets, int64_t offsetslength, bool ascending, bool stable);
/* replaces awkward-cpp/src/cpu-kernels/awkward_argsort.cpp:174; contract kernel-specification.yml:4111 */
struct Error awkward_argsort_float64_int64(int64_t* toptr, const double* fromptr, int64_t length, const int64_t* offsets, int64_t offsetslength, bool ascending, bool stable);

/* ---- awkward_ByteMaskedArray_numnull ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_ByteMaskedArray_numnull.cpp:3; contract kernel-specification.yml:196 */
struct Error awkward_ByteMaskedArray_numnull(int64_t* numnull, const int8_t* mask, int64_t length, bool validwhen);

/* ---- awkward_ByteMaskedArray_reduce_next_64 ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_ByteMaskedArray_reduce_next_64.cpp:3; contract kernel-specification.yml:231 */
struct Error awkward_ByteMaskedArray_reduce_next_64(int64_t* nextcarry, int64_t* nextoffsets, int64_t* outindex, const int8_t* mask, const int64_t* offsets, int64_t outlength, bool validwhen);

/* ---- awkward_ByteMaskedArray_reduce_next_nonlocal_nextshifts_64 ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_ByteMaskedArray_reduce_next_nonlocal_nextshifts_64.cpp:3; contract kernel-specification.yml:259 */
struct Error awkward_ByteMaskedArray_reduce_next_nonlocal_nextshifts_64(int64_t* nextshifts, const int8_t* mask, int64_t length, bool valid_when);

/* ---- awkward_ByteMaskedArray_reduce_next_nonlocal_nextshifts_fromshifts_64 ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_ByteMaskedArray_reduce_next_nonlocal_nextshifts_fromshifts_64.cpp:3; contract kernel-specification.yml:282 */
struct Error awkward_ByteMaskedArray_reduce_next_nonlocal_nextshifts_fromshifts_64(int64_t* nextshifts, const int8_t* mask, int64_t length, bool valid_when, const int64_t* shifts);

/* ---- awkward_ByteMaskedArray_getitem_nextcarry ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_ByteMaskedArray_getitem_nextcarry.cpp:22; contract kernel-specification.yml:156 */
struct Error awkward_ByteMaskedArray_getitem_nextcarry_64(int64_t* tocarry, const int8_t* mask, int64_t length, bool validwhen);

/* ---- awkward_ByteMaskedArray_getitem_nextcarry_outindex ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_ByteMaskedArray_getitem_nextcarry_outindex.cpp:27; contract kernel-specification.yml:174 */
struct Error awkward_ByteMaskedArray_getitem_nextcarry_outindex_64(int64_t* tocarry, int64_t* outindex, const int8_t* mask, int64_t length, bool validwhen);

/* ---- awkward_IndexedArray_numnull ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_IndexedArray_numnull.cpp:26; contract kernel-specification.yml:573 */
struct Error awkward_IndexedArray32_numnull(int64_t* numnull, const int32_t* fromindex, int64_t lenindex);
/* replaces awkward-cpp/src/cpu-kernels/awkward_IndexedArray_numnull.cpp:28; contract kernel-specification.yml:578 */
struct Error awkward_IndexedArray64_numnull(int64_t* numnull, const int64_t* fromindex, int64_t lenindex);
/* replaces awkward-cpp/src/cpu-kernels/awkward_IndexedArray_numnull.cpp:27; contract kernel-specification.yml:583 */
struct Error awkward_IndexedArrayU32_numnull(int64_t* numnull, const uint32_t* fromindex, int64_t lenindex);

/* ---- awkward_IndexedArray_reduce_next_64 ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_IndexedArray_reduce_next_64.cpp:38; contract kernel-specification.yml:707 */
struct Error awkward_IndexedArray32_reduce_next_64(int64_t* nextcarry, int64_t* nextoffsets, int64_t* outindex, const int32_t* index, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_IndexedArray_reduce_next_64.cpp:40; contract kernel-specification.yml:715 */
struct Error awkward_IndexedArray64_reduce_next_64(int64_t* nextcarry, int64_t* nextoffsets, int64_t* outindex, const int64_t* index, const int64_t* offsets, int64_t outlength);
/* replaces awkward-cpp/src/cpu-kernels/awkward_IndexedArray_reduce_next_64.cpp:39; contract kernel-specification.yml:723 */
struct Error awkward_IndexedArrayU32_reduce_next_64(int64_t* nextcarry, int64_t* nextoffsets, int64_t* outindex, const uint32_t* index, const int64_t* offsets, int64_t outlength);

/* ---- awkward_IndexedArray_reduce_next_nonlocal_nextshifts_64 ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_IndexedArray_reduce_next_nonlocal_nextshifts_64.cpp:31; contract kernel-specification.yml:799 */
struct Error awkward_IndexedArray32_reduce_next_nonlocal_nextshifts_64(int64_t* nextshifts, const int32_t* index, int64_t length);
/* replaces awkward-cpp/src/cpu-kernels/awkward_IndexedArray_reduce_next_nonlocal_nextshifts_64.cpp:33; contract kernel-specification.yml:804 */
struct Error awkward_IndexedArray64_reduce_next_nonlocal_nextshifts_64(int64_t* nextshifts, const int64_t* index, int64_t length);
/* replaces awkward-cpp/src/cpu-kernels/awkward_IndexedArray_reduce_next_nonlocal_nextshifts_64.cpp:32; contract kernel-specification.yml:809 */
struct Error awkward_IndexedArrayU32_reduce_next_nonlocal_nextshifts_64(int64_t* nextshifts, const uint32_t* index, int64_t length);

/* ---- awkward_IndexedArray_reduce_next_nonlocal_nextshifts_fromshifts_64 ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_IndexedArray_reduce_next_nonlocal_nextshifts_fromshifts_64.cpp:32; contract kernel-specification.yml:829 */
struct Error awkward_IndexedArray32_reduce_next_nonlocal_nextshifts_fromshifts_64(int64_t* nextshifts, const int32_t* index, int64_t length, const int64_t* shifts);
/* replaces awkward-cpp/src/cpu-kernels/awkward_IndexedArray_reduce_next_nonlocal_nextshifts_fromshifts_64.cpp:34; contract kernel-specification.yml:835 */
struct Error awkward_IndexedArray64_reduce_next_nonlocal_nextshifts_fromshifts_64(int64_t* nextshifts, const int64_t* index, int64_t length, const int64_t* shifts);
/* replaces awkward-cpp/src/cpu-kernels/awkward_IndexedArray_reduce_next_nonlocal_nextshifts_fromshifts_64.cpp:33; contract kernel-specification.yml:841 */
struct Error awkward_IndexedArrayU32_reduce_next_nonlocal_nextshifts_fromshifts_64(int64_t* nextshifts, const uint32_t* index, int64_t length, const int64_t* shifts);

/* ---- awkward_IndexedArray_getitem_nextcarry_outindex ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_IndexedArray_getitem_nextcarry_outindex.cpp:37; contract kernel-specification.yml:507 */
struct Error awkward_IndexedArray32_getitem_nextcarry_outindex_64(int64_t* tocarry, int32_t* toindex, const int32_t* fromindex, int64_t lenindex, int64_t lencontent);
/* replaces awkward-cpp/src/cpu-kernels/awkward_IndexedArray_getitem_nextcarry_outindex.cpp:39; contract kernel-specification.yml:514 */
struct Error awkward_IndexedArray64_getitem_nextcarry_outindex_64(int64_t* tocarry, int64_t* toindex, const int64_t* fromindex, int64_t lenindex, int64_t lencontent);
/* replaces awkward-cpp/src/cpu-kernels/awkward_IndexedArray_getitem_nextcarry_outindex.cpp:38; contract kernel-specification.yml:521 */
struct Error awkward_IndexedArrayU32_getitem_nextcarry_outindex_64(int64_t* tocarry, uint32_t* toindex, const uint32_t* fromindex, int64_t lenindex, int64_t lencontent);

/* ---- awkward_NumpyArray_reduce_adjust_starts_shifts_64 ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_NumpyArray_reduce_adjust_starts_shifts_64.cpp:3; contract kernel-specification.yml:2566 */
struct Error awkward_NumpyArray_reduce_adjust_starts_shifts_64(int64_t* toptr, int64_t outlength, const int64_t* offsets, const int64_t* starts, const int64_t* shifts);

/* ---- awkward_ListOffsetArray_reduce_nonlocal_maxcount_offsetscopy_64 ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_ListOffsetArray_reduce_nonlocal_maxcount_offsetscopy_64.cpp:3; contract kernel-specification.yml:2197 */
struct Error awkward_ListOffsetArray_reduce_nonlocal_maxcount_offsetscopy_64(int64_t* maxcount, int64_t* offsetscopy, const int64_t* offsets, int64_t length);

/* ---- awkward_ListOffsetArray_reduce_nonlocal_preparenext_64 ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_ListOffsetArray_reduce_nonlocal_preparenext_64.cpp:3; contract kernel-specification.yml:2296 */
struct Error awkward_ListOffsetArray_reduce_nonlocal_preparenext_64(int64_t* nextcarry, int64_t* nextoffsets, int64_t nextlen, int64_t* maxnextparents, int64_t* distincts, int64_t distinctslen, int64_t* offsetscopy, const int64_t* offsets, int64_t length, const int64_t* outer_offsets, int64_t outlength, int64_t maxcount);

/* ---- awkward_ListOffsetArray_reduce_nonlocal_nextshifts_64 ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_ListOffsetArray_reduce_nonlocal_nextshifts_64.cpp:3; contract kernel-specification.yml:2219 */
struct Error awkward_ListOffsetArray_reduce_nonlocal_nextshifts_64(int64_t* nummissing, int64_t* missing, int64_t* nextshifts, const int64_t* offsets, int64_t length, const int64_t* starts, const int64_t* outer_offsets, int64_t outlength, int64_t maxcount, int64_t nextlen, const int64_t* nextcarry);

/* ---- awkward_ListOffsetArray_reduce_nonlocal_outstartsstops_64 ---- */
/* replaces awkward-cpp/src/cpu-kernels/awkward_ListOffsetArray_reduce_nonlocal_outstartsstops_64.cpp:6; contract kernel-specification.yml:2256 */
struct Error awkward_ListOffsetArray_reduce_nonlocal_outstartsstops_64(int64_t* outstarts, int64_t* outstops, const int64_t* distincts, int64_t lendistincts, int64_t outlength);

/* ---- runtime (no counterpart in awkward-cpp, which is synchronous CPU code) ---- */

/* Stream that subsequent entry points enqueue on (a hipStream_t; NULL = the
   null stream).  Mirrors the reference CUDA backend's use of the current
   CuPy stream (src/awkward/_kernels.py:188-232). */
void awkward_hip_set_stream(void* stream);
void* awkward_hip_get_stream(void);

/* 0 (default): entry points with device-detected errors synchronise and
   return the error by value.  1: they return at once; errors accumulate in
   a device error word read by awkward_hip_error_poll. */
void awkward_hip_set_deferred_errors(int deferred);

/* Synchronises the current stream and returns the first error recorded in
   deferred mode since the last poll (then clears it). */
struct Error awkward_hip_error_poll(void);

/* Library version string and the gfx target it was built for. */
const char* awkward_hip_version(void);

#ifdef __cplusplus
}
#endif

#endif /* AWKWARD_HIP_H_ */
