/* awkward_hip_ext.h -- entry points of libawkward_hip.so that have no
 * counterpart in awkward-cpp.  They serve the same results as sequences of
 * the registry kernels declared in awkward_hip.h (tuned tile variants with
 * fused epilogues, fused op sequences, the batch runner, the learned tile
 * dispatcher) or are runtime plumbing.  Same conventions: device pointers
 * unless a comment says host, kernels enqueued on the stream set by
 * awkward_hip_set_stream(), errors returned as struct Error (str == NULL is
 * success).  Every symbol declared here is exported by the library.
 */
#ifndef AWKWARD_HIP_EXT_H
#define AWKWARD_HIP_EXT_H

#include <stdbool.h>
#include <stdint.h>

#include "awkward_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- runtime ------------------------------------------------------------ */
/* State is kept per (device, stream): each stream that calls the library has
   its own device error word, deferred-mode invocation table and scratch
   pools, so work on different streams never shares scratch or errors.  A
   pool that grows retires its old buffer instead of freeing it (in-flight
   kernels and captured hipGraphs may still use it); retired buffers stay
   allocated until the process ends or awkward_hip_release_pools frees
   them.  Pools double up to 1 GiB and then grow to the request (rounded to
   64 MiB).  No entry point synchronises the device or frees memory to grow
   a pool. */
/* Grow the current stream's scratch pool to at least `bytes` now (and create
   the stream's state), e.g. before capturing a hipGraph on that stream.
   0 on success. */
int awkward_hip_reserve_scratch(int64_t bytes);
/* The same for the work-queue pool (the batch reductions' long-list queue). */
int awkward_hip_reserve_queue(int64_t bytes);
/* Diagnostics: out[0] = bytes of the current stream's scratch pool, out[1] =
   of its queue pool, out[2] = number of retired pool buffers. */
void awkward_hip_pool_bytes(int64_t* out);
/* Synchronise the current stream and free its pools' buffers, current and
   retired (the next call allocates afresh); only when no captured hipGraph
   still uses them (no counterpart in the reference: for callers that stream
   batches larger than HBM).  0 on success. */
int awkward_hip_release_pools(void);
/* Number of reduce tile variants (threads x lists per thread x LDS stage). */
int awkward_hip_reduce_variant_count(void);

/* ---- tuned reductions: one tile variant (0 .. 9, or 10: lists longer than
   4096 elements queued in 8192-element chunks across the chip: a fixed-capacity
   queue, nothing read back, capturable), fused epilogues ---------
   sum:  awkward_reduce_sum_<OUT>_<IN>_64 with the variant's tile.
   max/min_masked: awkward_reduce_max/min + the empty-list mask
     (awkward_NumpyArray_reduce_mask_ByteMaskedArray_64) in one pass; mask may
     be NULL.
   argmax/argmin_local: awkward_reduce_argmax/argmin + adjust_starts with
     starts = offsets[:-1] (local indices) + the empty-list mask. */
struct Error awkward_hip_reduce_sum_int64_int8(int64_t* toptr, const int8_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_sum_uint64_uint8(uint64_t* toptr, const uint8_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_sum_int64_int16(int64_t* toptr, const int16_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_sum_uint64_uint16(uint64_t* toptr, const uint16_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_sum_int64_int32(int64_t* toptr, const int32_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_sum_uint64_uint32(uint64_t* toptr, const uint32_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_sum_int64_int64(int64_t* toptr, const int64_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_sum_uint64_uint64(uint64_t* toptr, const uint64_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_sum_float32_float32(float* toptr, const float* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_sum_float64_float64(double* toptr, const double* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_sum_int32_int8(int32_t* toptr, const int8_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_sum_uint32_uint8(uint32_t* toptr, const uint8_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_sum_int32_int16(int32_t* toptr, const int16_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_sum_uint32_uint16(uint32_t* toptr, const uint16_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_sum_int32_int32(int32_t* toptr, const int32_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_sum_uint32_uint32(uint32_t* toptr, const uint32_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_sum_float64_int8(double* toptr, const int8_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_sum_float64_uint8(double* toptr, const uint8_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_sum_float64_int16(double* toptr, const int16_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_sum_float64_uint16(double* toptr, const uint16_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_sum_float64_int32(double* toptr, const int32_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_sum_float64_uint32(double* toptr, const uint32_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_sum_float64_int64(double* toptr, const int64_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_sum_float64_uint64(double* toptr, const uint64_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_sum_float64_float32(double* toptr, const float* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_sum_int64_bool(int64_t* toptr, const bool* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_sum_int32_bool(int32_t* toptr, const bool* fromptr, const int64_t* offsets, int64_t outlength, int variant);
/* prod:  awkward_reduce_prod_<OUT>_<IN>_64 with the variant's tile. */
struct Error awkward_hip_reduce_prod_int64_int8(int64_t* toptr, const int8_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_prod_uint64_uint8(uint64_t* toptr, const uint8_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_prod_int64_int16(int64_t* toptr, const int16_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_prod_uint64_uint16(uint64_t* toptr, const uint16_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_prod_int64_int32(int64_t* toptr, const int32_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_prod_uint64_uint32(uint64_t* toptr, const uint32_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_prod_int64_int64(int64_t* toptr, const int64_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_prod_uint64_uint64(uint64_t* toptr, const uint64_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_prod_float32_float32(float* toptr, const float* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_prod_float64_float64(double* toptr, const double* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_prod_int32_int8(int32_t* toptr, const int8_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_prod_uint32_uint8(uint32_t* toptr, const uint8_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_prod_int32_int16(int32_t* toptr, const int16_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_prod_uint32_uint16(uint32_t* toptr, const uint16_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_prod_int32_int32(int32_t* toptr, const int32_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_prod_uint32_uint32(uint32_t* toptr, const uint32_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
/* any / all (ak.any, ak.all; ak.max / ak.min of booleans): awkward_reduce_sum_bool_<T>_64 /
   awkward_reduce_prod_bool_<T>_64 with the variant's tile, plus the empty-list mask of
   mask_identity=True (mask[k] = 1 for an empty list; NULL: not written). */
struct Error awkward_hip_reduce_any_masked_bool(bool* toptr, int8_t* mask, const bool* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_any_masked_int8(bool* toptr, int8_t* mask, const int8_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_any_masked_uint8(bool* toptr, int8_t* mask, const uint8_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_any_masked_int16(bool* toptr, int8_t* mask, const int16_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_any_masked_uint16(bool* toptr, int8_t* mask, const uint16_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_any_masked_int32(bool* toptr, int8_t* mask, const int32_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_any_masked_uint32(bool* toptr, int8_t* mask, const uint32_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_any_masked_int64(bool* toptr, int8_t* mask, const int64_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_any_masked_uint64(bool* toptr, int8_t* mask, const uint64_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_any_masked_float32(bool* toptr, int8_t* mask, const float* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_any_masked_float64(bool* toptr, int8_t* mask, const double* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_all_masked_bool(bool* toptr, int8_t* mask, const bool* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_all_masked_int8(bool* toptr, int8_t* mask, const int8_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_all_masked_uint8(bool* toptr, int8_t* mask, const uint8_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_all_masked_int16(bool* toptr, int8_t* mask, const int16_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_all_masked_uint16(bool* toptr, int8_t* mask, const uint16_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_all_masked_int32(bool* toptr, int8_t* mask, const int32_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_all_masked_uint32(bool* toptr, int8_t* mask, const uint32_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_all_masked_int64(bool* toptr, int8_t* mask, const int64_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_all_masked_uint64(bool* toptr, int8_t* mask, const uint64_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_all_masked_float32(bool* toptr, int8_t* mask, const float* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_all_masked_float64(bool* toptr, int8_t* mask, const double* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_max_masked_int8(int8_t* toptr, int8_t* mask, const int8_t* fromptr, const int64_t* offsets, int64_t outlength, int8_t identity, int variant);
struct Error awkward_hip_reduce_max_masked_uint8(uint8_t* toptr, int8_t* mask, const uint8_t* fromptr, const int64_t* offsets, int64_t outlength, uint8_t identity, int variant);
struct Error awkward_hip_reduce_max_masked_int16(int16_t* toptr, int8_t* mask, const int16_t* fromptr, const int64_t* offsets, int64_t outlength, int16_t identity, int variant);
struct Error awkward_hip_reduce_max_masked_uint16(uint16_t* toptr, int8_t* mask, const uint16_t* fromptr, const int64_t* offsets, int64_t outlength, uint16_t identity, int variant);
struct Error awkward_hip_reduce_max_masked_int32(int32_t* toptr, int8_t* mask, const int32_t* fromptr, const int64_t* offsets, int64_t outlength, int32_t identity, int variant);
struct Error awkward_hip_reduce_max_masked_uint32(uint32_t* toptr, int8_t* mask, const uint32_t* fromptr, const int64_t* offsets, int64_t outlength, uint32_t identity, int variant);
struct Error awkward_hip_reduce_max_masked_int64(int64_t* toptr, int8_t* mask, const int64_t* fromptr, const int64_t* offsets, int64_t outlength, int64_t identity, int variant);
struct Error awkward_hip_reduce_max_masked_uint64(uint64_t* toptr, int8_t* mask, const uint64_t* fromptr, const int64_t* offsets, int64_t outlength, uint64_t identity, int variant);
struct Error awkward_hip_reduce_max_masked_float32(float* toptr, int8_t* mask, const float* fromptr, const int64_t* offsets, int64_t outlength, float identity, int variant);
struct Error awkward_hip_reduce_max_masked_float64(double* toptr, int8_t* mask, const double* fromptr, const int64_t* offsets, int64_t outlength, double identity, int variant);
struct Error awkward_hip_reduce_min_masked_int8(int8_t* toptr, int8_t* mask, const int8_t* fromptr, const int64_t* offsets, int64_t outlength, int8_t identity, int variant);
struct Error awkward_hip_reduce_min_masked_uint8(uint8_t* toptr, int8_t* mask, const uint8_t* fromptr, const int64_t* offsets, int64_t outlength, uint8_t identity, int variant);
struct Error awkward_hip_reduce_min_masked_int16(int16_t* toptr, int8_t* mask, const int16_t* fromptr, const int64_t* offsets, int64_t outlength, int16_t identity, int variant);
struct Error awkward_hip_reduce_min_masked_uint16(uint16_t* toptr, int8_t* mask, const uint16_t* fromptr, const int64_t* offsets, int64_t outlength, uint16_t identity, int variant);
struct Error awkward_hip_reduce_min_masked_int32(int32_t* toptr, int8_t* mask, const int32_t* fromptr, const int64_t* offsets, int64_t outlength, int32_t identity, int variant);
struct Error awkward_hip_reduce_min_masked_uint32(uint32_t* toptr, int8_t* mask, const uint32_t* fromptr, const int64_t* offsets, int64_t outlength, uint32_t identity, int variant);
struct Error awkward_hip_reduce_min_masked_int64(int64_t* toptr, int8_t* mask, const int64_t* fromptr, const int64_t* offsets, int64_t outlength, int64_t identity, int variant);
struct Error awkward_hip_reduce_min_masked_uint64(uint64_t* toptr, int8_t* mask, const uint64_t* fromptr, const int64_t* offsets, int64_t outlength, uint64_t identity, int variant);
struct Error awkward_hip_reduce_min_masked_float32(float* toptr, int8_t* mask, const float* fromptr, const int64_t* offsets, int64_t outlength, float identity, int variant);
struct Error awkward_hip_reduce_min_masked_float64(double* toptr, int8_t* mask, const double* fromptr, const int64_t* offsets, int64_t outlength, double identity, int variant);
struct Error awkward_hip_reduce_argmax_local_int8(int64_t* toptr, int8_t* mask, const int8_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_argmax_local_uint8(int64_t* toptr, int8_t* mask, const uint8_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_argmax_local_int16(int64_t* toptr, int8_t* mask, const int16_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_argmax_local_uint16(int64_t* toptr, int8_t* mask, const uint16_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_argmax_local_int32(int64_t* toptr, int8_t* mask, const int32_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_argmax_local_uint32(int64_t* toptr, int8_t* mask, const uint32_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_argmax_local_int64(int64_t* toptr, int8_t* mask, const int64_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_argmax_local_uint64(int64_t* toptr, int8_t* mask, const uint64_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_argmax_local_float32(int64_t* toptr, int8_t* mask, const float* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_argmax_local_float64(int64_t* toptr, int8_t* mask, const double* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_argmin_local_int8(int64_t* toptr, int8_t* mask, const int8_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_argmin_local_uint8(int64_t* toptr, int8_t* mask, const uint8_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_argmin_local_int16(int64_t* toptr, int8_t* mask, const int16_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_argmin_local_uint16(int64_t* toptr, int8_t* mask, const uint16_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_argmin_local_int32(int64_t* toptr, int8_t* mask, const int32_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_argmin_local_uint32(int64_t* toptr, int8_t* mask, const uint32_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_argmin_local_int64(int64_t* toptr, int8_t* mask, const int64_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_argmin_local_uint64(int64_t* toptr, int8_t* mask, const uint64_t* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_argmin_local_float32(int64_t* toptr, int8_t* mask, const float* fromptr, const int64_t* offsets, int64_t outlength, int variant);
struct Error awkward_hip_reduce_argmin_local_float64(int64_t* toptr, int8_t* mask, const double* fromptr, const int64_t* offsets, int64_t outlength, int variant);
/* arg-reductions below an option type: the content was compacted, shifts[i]
   counts the nulls removed before element i and starts[k] is list k's start
   before compaction; the local index is v + shifts[v] - starts[k]
   (awkward_NumpyArray_reduce_adjust_starts_shifts_64 fused, replacing the
   sequence at awkward/_reducers.py:85-129 with shifts; reference
   awkward-cpp/src/cpu-kernels/awkward_NumpyArray_reduce_adjust_starts_shifts_64.cpp). */
struct Error awkward_hip_reduce_argmax_shifts_int8(int64_t* toptr, int8_t* mask, const int8_t* fromptr, const int64_t* offsets, const int64_t* starts, const int64_t* shifts, int64_t outlength, int variant);
struct Error awkward_hip_reduce_argmax_shifts_uint8(int64_t* toptr, int8_t* mask, const uint8_t* fromptr, const int64_t* offsets, const int64_t* starts, const int64_t* shifts, int64_t outlength, int variant);
struct Error awkward_hip_reduce_argmax_shifts_int16(int64_t* toptr, int8_t* mask, const int16_t* fromptr, const int64_t* offsets, const int64_t* starts, const int64_t* shifts, int64_t outlength, int variant);
struct Error awkward_hip_reduce_argmax_shifts_uint16(int64_t* toptr, int8_t* mask, const uint16_t* fromptr, const int64_t* offsets, const int64_t* starts, const int64_t* shifts, int64_t outlength, int variant);
struct Error awkward_hip_reduce_argmax_shifts_int32(int64_t* toptr, int8_t* mask, const int32_t* fromptr, const int64_t* offsets, const int64_t* starts, const int64_t* shifts, int64_t outlength, int variant);
struct Error awkward_hip_reduce_argmax_shifts_uint32(int64_t* toptr, int8_t* mask, const uint32_t* fromptr, const int64_t* offsets, const int64_t* starts, const int64_t* shifts, int64_t outlength, int variant);
struct Error awkward_hip_reduce_argmax_shifts_int64(int64_t* toptr, int8_t* mask, const int64_t* fromptr, const int64_t* offsets, const int64_t* starts, const int64_t* shifts, int64_t outlength, int variant);
struct Error awkward_hip_reduce_argmax_shifts_uint64(int64_t* toptr, int8_t* mask, const uint64_t* fromptr, const int64_t* offsets, const int64_t* starts, const int64_t* shifts, int64_t outlength, int variant);
struct Error awkward_hip_reduce_argmax_shifts_float32(int64_t* toptr, int8_t* mask, const float* fromptr, const int64_t* offsets, const int64_t* starts, const int64_t* shifts, int64_t outlength, int variant);
struct Error awkward_hip_reduce_argmax_shifts_float64(int64_t* toptr, int8_t* mask, const double* fromptr, const int64_t* offsets, const int64_t* starts, const int64_t* shifts, int64_t outlength, int variant);
struct Error awkward_hip_reduce_argmin_shifts_int8(int64_t* toptr, int8_t* mask, const int8_t* fromptr, const int64_t* offsets, const int64_t* starts, const int64_t* shifts, int64_t outlength, int variant);
struct Error awkward_hip_reduce_argmin_shifts_uint8(int64_t* toptr, int8_t* mask, const uint8_t* fromptr, const int64_t* offsets, const int64_t* starts, const int64_t* shifts, int64_t outlength, int variant);
struct Error awkward_hip_reduce_argmin_shifts_int16(int64_t* toptr, int8_t* mask, const int16_t* fromptr, const int64_t* offsets, const int64_t* starts, const int64_t* shifts, int64_t outlength, int variant);
struct Error awkward_hip_reduce_argmin_shifts_uint16(int64_t* toptr, int8_t* mask, const uint16_t* fromptr, const int64_t* offsets, const int64_t* starts, const int64_t* shifts, int64_t outlength, int variant);
struct Error awkward_hip_reduce_argmin_shifts_int32(int64_t* toptr, int8_t* mask, const int32_t* fromptr, const int64_t* offsets, const int64_t* starts, const int64_t* shifts, int64_t outlength, int variant);
struct Error awkward_hip_reduce_argmin_shifts_uint32(int64_t* toptr, int8_t* mask, const uint32_t* fromptr, const int64_t* offsets, const int64_t* starts, const int64_t* shifts, int64_t outlength, int variant);
struct Error awkward_hip_reduce_argmin_shifts_int64(int64_t* toptr, int8_t* mask, const int64_t* fromptr, const int64_t* offsets, const int64_t* starts, const int64_t* shifts, int64_t outlength, int variant);
struct Error awkward_hip_reduce_argmin_shifts_uint64(int64_t* toptr, int8_t* mask, const uint64_t* fromptr, const int64_t* offsets, const int64_t* starts, const int64_t* shifts, int64_t outlength, int variant);
struct Error awkward_hip_reduce_argmin_shifts_float32(int64_t* toptr, int8_t* mask, const float* fromptr, const int64_t* offsets, const int64_t* starts, const int64_t* shifts, int64_t outlength, int variant);
struct Error awkward_hip_reduce_argmin_shifts_float64(int64_t* toptr, int8_t* mask, const double* fromptr, const int64_t* offsets, const int64_t* starts, const int64_t* shifts, int64_t outlength, int variant);

/* ---- without allocation ------------------------------------------------------
   No counterpart in the reference: for callers that cannot allocate device
   memory (code under hipGraph capture, tight serving loops, harnesses that
   forbid hipMalloc).  awkward_ListArray64_getitem_next_range_64 with the
   look-back scan's scratch (tile counter and status words) in a file-scope
   1 MiB __device__ array instead of the stream's pool; no hipMalloc,
   hipFree or host synchronisation.  At most 536,608,768 lists (a clear
   error beyond).  Every call on one device shares that device's array, and
   each call clears what it uses first: calls that overlap on one device
   (concurrent streams, threads) race. */
struct Error awkward_hip_ListArray64_getitem_next_range_64_noalloc(int64_t* tooffsets, int64_t* tocarry, const int64_t* fromstarts, const int64_t* fromstops, int64_t lenstarts, int64_t start, int64_t stop, int64_t step);

/* ---- elementwise expressions -------------------------------------------------
   A program (struct ExprProgram of awkward_amd/csrc/expr.h: up to 8 leaf
   columns, each optionally gathered through an int64 index, and up to 40
   typed ops over 10 registers) is a chain of numpy ufuncs with the operand
   types already resolved by numpy's rules (array_module.py:254-332 applies
   them one at a time).  The program pointer is a HOST pointer; the leaves
   are device pointers.
   expr_eval: out[i] = program(start + i), 0 <= i < n, in the program's
     out_type.
   mask_getitem: array[mask] for a jagged boolean mask program of masklen
     elements over int64 offsets moffsets (n lists): the program's bits
     (one word per 64 elements, in the stream's scratch), then the result's
     offsets (n + 1, from 0) and
     its carry into the content of the lists [fromstarts[i], fromstops[i]),
     with awkward_ListArray_getitem_jagged_apply's checks and messages --
     the reference's _normalise_item_bool_to_int + jagged_expand + carrylen
     + jagged_apply sequence in one scan.  tocarry must hold the mask's
     true count (an upper bound such as its element count will do); the
     count is tooffsets[n].  A program that is a conjunction of float32
     column comparisons (optionally of abs) with constants -- a selection
     cut -- takes a dedicated kernel for its bits (the same comparisons).
   expr_cut_terms: host only, no device work: the number of comparison
     terms when the program is such a cut, else 0. */
struct Error awkward_hip_expr_eval(void* out, const void* program, int64_t start, int64_t n);
int awkward_hip_expr_cut_terms(const void* program);
struct Error awkward_hip_mask_getitem(int64_t* tooffsets, int64_t* tocarry, const int64_t* fromstarts, const int64_t* fromstops, int64_t contentlen, const int64_t* moffsets, int64_t n, int64_t masklen, const void* program);

/* ---- gathers and small host-layer steps ------------------------------------
   take_<bytes>: out[i] = data[index[i]] (NumpyArray._carry, numpyarray.py:367-373).
   bool_to_index: the jagged boolean mask -> jagged int index of
     _slicing._normalise_item_bool_to_int (_slicing.py:476-530): toindex gets
     the local positions of the True entries (room for len(mask)), tooffsets
     (n + 1) their per-list counts' prefix.
   subtract_i64: out = a - b (ak.num's stops - starts, ak_num.py:133-142). */
struct Error awkward_hip_take_1(void* out, const void* data, const int64_t* index, int64_t n);
struct Error awkward_hip_take_2(void* out, const void* data, const int64_t* index, int64_t n);
struct Error awkward_hip_take_4(void* out, const void* data, const int64_t* index, int64_t n);
struct Error awkward_hip_take_8(void* out, const void* data, const int64_t* index, int64_t n);
struct Error awkward_hip_bool_to_index(int64_t* toindex, int64_t* tooffsets, const bool* mask, const int64_t* offsets, int64_t n);
struct Error awkward_hip_subtract_i64(int64_t* out, const int64_t* a, const int64_t* b, int64_t n);
/* Measurement only: a plain streaming read of nbytes (rounded down to 16)
   of 16-byte-aligned `data`, the read-rate ceiling a reduction over the same
   buffer is compared with; `sink` (one uint64) is written only if a 16-byte
   word folds (x ^ z, y ^ w) to one fixed 64-bit pattern. */
struct Error awkward_hip_stream_read(const void* data, int64_t nbytes, uint64_t* sink);

/* ---- combinations without a host round trip ---------------------------------
   awkward_ListArray_combinations for lists [starts, stops) whose output
   offsets tooffsets (length + 1, from awkward_ListArray_combinations_length)
   are already on the device; tocarry: device array of n device pointers;
   totallen: the number of tuples (tooffsets[length]). */
struct Error awkward_hip_combinations_from_offsets(int64_t** tocarry, int64_t* toindex, const int64_t* tooffsets, const int64_t* starts, const int64_t* stops, int64_t length, int64_t n, bool replacement, int64_t totallen);

/* ---- fused op sequences (capturable in a hipGraph) ----------------------------
   event_pipeline: per event of records, sel = ev[AND of terms][:, start:stop:step]
     then num(sel), sum/argmax/max of the float32 field `red` -- the results of
     the mask ufuncs + bool_to_index + jagged_expand/carrylen/apply +
     next_range + gathers + reduce kernels.  field_ptrs, field_types,
     term_* are HOST arrays; types 0 f32, 1 f64, 2 i8, 3 i32, 4 i64; ops
     0 >, 1 >=, 2 <, 3 <=, 4 ==, 5 !=; any output may be NULL.
   pair_max: per event, max and first argmax over its particle pairs (the
     order of awkward_ListArray_combinations, n = 2) of
     2 pt_i pt_j (cosh(eta_i - eta_j) - cos(phi_i - phi_j)) in float32.
   pair_expr: that expression over already gathered legs. */
struct Error awkward_hip_event_pipeline(const int64_t* offsets, int64_t n, int64_t n_particles, const void* const* field_ptrs, const int32_t* field_types, int32_t nfields, const int32_t* term_field, const int32_t* term_op, const int32_t* term_abs, const double* term_value, int32_t nterms, int64_t start, int64_t stop, int64_t step, int32_t red, int64_t* out_num, float* out_sum, int64_t* out_argmax, int8_t* out_argmask, float* out_max, int8_t* out_maxmask);
struct Error awkward_hip_pair_max(const int64_t* offsets, int64_t n, const float* pt, const float* eta, const float* phi, float* out_max, int64_t* out_argmax, int8_t* out_mask, float identity);
struct Error awkward_hip_pair_expr(float* out, const float* pt1, const float* eta1, const float* phi1, const float* pt2, const float* eta2, const float* phi2, int64_t n);

/* ---- reductions over the pairs of each list, from an expression program ------
   pair_reduce: ak.max / ak.min / ak.argmax / ak.argmin / ak.sum at axis=-1
     of an elementwise expression of ak.combinations(array, 2) -- the
     reference's awkward_ListArray_combinations_length + _combinations,
     per-leg carries, one apply_ufunc per op and awkward_reduce_{max,min,
     argmax,argmin,sum} (listoffsetarray.py:1152-1261, array_module.py:
     254-332) -- without writing the pairs or the expression's values.
     program (a HOST pointer) is an expr.h ExprProgram of float32 or
     float64 whose leaves have leg 0 (read at the pair's first element) or
     1 (its second); offsets are the n lists' int64 offsets, absolute
     positions in the leaves' columns.  Pairs are taken i < j in the
     lexicographic order of awkward_ListArray_combinations; max / min start
     from -inf / +inf, argmax / argmin are local pair indices (first of
     equals, -1 without pairs), sums accumulate left to right in the
     program's type, mask is 1 for a list without pairs; any output may be
     NULL.  The reductions run in a kernel compiled at run time (hipRTC) for
     the program's shape on first use -- the program's ops as straight-line
     code, the same values as the interpreter's -- and cached per device;
     AWKWARD_HIP_JIT=0 in the environment keeps the interpreter. */
struct Error awkward_hip_pair_reduce(const int64_t* offsets, int64_t n, const void* program, void* out_max, void* out_min, void* out_sum, int64_t* out_argmax, int64_t* out_argmin, int8_t* out_mask);
/* combinations_pairs: ak.combinations(array, k) for up to two k in {2, 3}
     (without replacement) -- the reference's awkward_ListArray_combinations
     per k -- and, when program is not NULL, pair_reduce's reductions, in one
     launch.  req_n (nreq sizes), req_tooffsets (for each request the n + 1
     tuple offsets that awkward_ListArray64_combinations_length_64 wrote for
     these offsets) and req_carries (req_n[0] carry pointers, then req_n[1])
     are HOST arrays of device pointers; every carry holds the request's
     tuple count; carries are absolute element positions.  With a program,
     the run-time compiled kernel of pair_reduce also writes the carries
     (one launch). */
struct Error awkward_hip_combinations_pairs(const int64_t* offsets, int64_t n, int32_t nreq, const int32_t* req_n, const int64_t* const* req_tooffsets, int64_t* const* req_carries, const void* program, void* out_max, void* out_min, void* out_sum, int64_t* out_argmax, int64_t* out_argmin, int8_t* out_mask);
/* jit_kernels: host only -- the number of kernels compiled at run time so
     far in this process (one per program shape and device). */
int awkward_hip_jit_kernels(void);
/* jit_selftest: host only, no device needed -- compiles the run-time pair
     kernel of two representative programs (float32 with every reduction and
     the n = 2 / 3 carries; float64 with an int8 leaf's cast) for target (an
     --offload-arch value such as "gfx950"): 0 when both build, else the
     number that failed (the compiler's log on stderr).  The build runs it. */
int awkward_hip_jit_selftest(const char* target);

/* ---- batch runner: many arrays of one dtype, persistent work queues ------------
   items: device table of 64-byte rows {data, offsets, out_sum, out_min,
   out_argmin, out_mask, n, tile_begin} (reduce) or 56-byte rows {data,
   offsets, out, n, length, tile_begin, elem_begin} (argsort), tile_begin
   ascending in units of the variant's lists per tile; counter: a device
   uint64 cleared by awkward_hip_batch_reset before each launch.
   reduce: per list sum (awkward_reduce_sum's output type), min (with the
   type's identity) and local argmin + empty mask in one pass.
   reduce_<T>_chosen: all ten tiles enqueued, each predicated on the device
   int choice[0] (awkward_hip_dispatch_infer); tables / total_tiles are HOST
   arrays of ten.
   total_elements: the batch's element count; lists longer than 4096
   elements are reduced in 8192-element chunks by every workgroup (queue
   from the work-queue pool).
   argsort: stable ascending/descending argsort per list, local indices. */
struct Error awkward_hip_batch_reset(unsigned long long* counter);
int64_t awkward_hip_batch_reduce_tile_lists(int variant);
int64_t awkward_hip_batch_sort_tile_lists(void);
struct Error awkward_hip_batch_reduce_int32(const void* items, int32_t nitems, int64_t total_tiles, unsigned long long* counter, int32_t identity, int32_t variant, int64_t total_elements);
struct Error awkward_hip_batch_reduce_int32_chosen(const void* const* tables, const int64_t* total_tiles, int32_t nitems, unsigned long long* counter, int32_t identity, const int32_t* choice, int64_t total_elements);
struct Error awkward_hip_batch_argsort_int32(const void* items, int32_t nitems, int64_t total_tiles, unsigned long long* counter, bool ascending, int64_t total_elements);
struct Error awkward_hip_batch_reduce_int64(const void* items, int32_t nitems, int64_t total_tiles, unsigned long long* counter, int64_t identity, int32_t variant, int64_t total_elements);
struct Error awkward_hip_batch_reduce_int64_chosen(const void* const* tables, const int64_t* total_tiles, int32_t nitems, unsigned long long* counter, int64_t identity, const int32_t* choice, int64_t total_elements);
struct Error awkward_hip_batch_argsort_int64(const void* items, int32_t nitems, int64_t total_tiles, unsigned long long* counter, bool ascending, int64_t total_elements);
struct Error awkward_hip_batch_reduce_float32(const void* items, int32_t nitems, int64_t total_tiles, unsigned long long* counter, float identity, int32_t variant, int64_t total_elements);
struct Error awkward_hip_batch_reduce_float32_chosen(const void* const* tables, const int64_t* total_tiles, int32_t nitems, unsigned long long* counter, float identity, const int32_t* choice, int64_t total_elements);
struct Error awkward_hip_batch_argsort_float32(const void* items, int32_t nitems, int64_t total_tiles, unsigned long long* counter, bool ascending, int64_t total_elements);
struct Error awkward_hip_batch_reduce_float64(const void* items, int32_t nitems, int64_t total_tiles, unsigned long long* counter, double identity, int32_t variant, int64_t total_elements);
struct Error awkward_hip_batch_reduce_float64_chosen(const void* const* tables, const int64_t* total_tiles, int32_t nitems, unsigned long long* counter, double identity, const int32_t* choice, int64_t total_elements);
struct Error awkward_hip_batch_argsort_float64(const void* items, int32_t nitems, int64_t total_tiles, unsigned long long* counter, bool ascending, int64_t total_elements);

/* ---- learned tile dispatcher ---------------------------------------------------
   layout(what): 0 features (16), 1 hidden (64), 2 outputs (11), 3 trained
     parameters, 4 parameter array length (parameters + normalisation), 5
     accumulator words per slot, 6 lists per statistics tile, 7 rows per
     training workgroup.
   reset / stats: integer statistics of list lengths per slot; items: device
     rows {offsets, n, tile_begin, slot} (32 bytes).
   infer: meta (device int32[4 * nslots]) {itemsize, op, fallback, 0} with op
     0 sum, 1 min/max, 2 argmin/argmax, 3 batched sum+min+argmin; writes
     features (16 per slot), costs (11 per slot) -- either may be NULL -- and
     choice (2 per slot: tile, 1 if the fallback was taken).
   choose: stats then infer in one launch (the workgroup that finishes the
     statistics last runs the network); acc must hold zeros on entry and
     holds zeros again when the launch ends; launches sharing one acc must not
     overlap.
   mlp_forward: rows of raw features -> 11 outputs (+ inside-range flags).
   mlp_train_step / mlp_adam: one minibatch's gradients (one row per
     workgroup of 64) and the Adam step over their fixed-order sum.
   dispatch_time_variants: the exhaustive sweep; ms_out is a HOST float[10];
     dtype 0 int32, 1 int64, 2 float32, 3 float64; synchronises. */
int32_t awkward_hip_dispatch_layout(int32_t what);
struct Error awkward_hip_dispatch_reset(unsigned long long* acc, int64_t nslots);
struct Error awkward_hip_dispatch_stats(const void* items, int32_t nitems, int64_t total_tiles, unsigned long long* acc);
struct Error awkward_hip_dispatch_infer(const unsigned long long* acc, const int32_t* meta, int32_t nslots, const float* params, float* features, float* costs, int32_t* choice);
struct Error awkward_hip_dispatch_choose(const void* items, int32_t nitems, int64_t total_tiles, unsigned long long* acc, const int32_t* meta, int32_t nslots, const float* params, float* features, float* costs, int32_t* choice);
struct Error awkward_hip_mlp_forward(const float* X, int64_t rows, const float* params, float* Y, int32_t* inside);
struct Error awkward_hip_mlp_train_step(const float* X, const float* T, const float* ow, int64_t R, int64_t pa, int64_t pb, int64_t row0, int64_t bsz, const float* params, float* grads, float* losses);
struct Error awkward_hip_mlp_adam(float* params, float* m, float* v, const float* grads, const float* losses, int32_t nparts, float lr, float beta1, float beta2, float eps, int64_t step, float* loss_acc);
struct Error awkward_hip_dispatch_time_variants(int32_t op, int32_t dtype, const void* data, const int64_t* offsets, int64_t n, int64_t length, int32_t reps, float* ms_out);

#ifdef __cplusplus
}
#endif

#endif /* AWKWARD_HIP_EXT_H */
