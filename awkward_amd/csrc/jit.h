// jit.h -- kernels compiled at run time for one expression program (hipRTC).
#pragma once

#include "common.h"
#include "expr.h"

namespace akhip {

// The reductions of a program over the pairs of each list (pair_jit.h),
// and optionally the lists' ak.combinations carries for n = 2 / 3:
// pair_jit_kernel compiles the kernel for this program's shape on first use
// (cached per device) and returns it, or null when run-time compilation is
// switched off (AWKWARD_HIP_JIT=0) or the kernel cannot be built -- the
// caller then interprets the program; pair_jit_run launches it on the
// current stream.  slot / first: the program's staged columns (leaf k reads
// column slot[k]; column s is loaded through leaf first[s]).  mx / mn / sum:
// max or argmax / min or argmin / sum requested; tuples: carries requested
// (nreq > 0 at the launch; req_* as awkward_hip_combinations_pairs takes
// them).
const void* pair_jit_kernel(const ExprProgram& p, const int8_t* slot, const int8_t* first, int nslots, bool mx,
                            bool mn, bool sum, bool tuples);
void pair_jit_run(const void* kernel, const ExprProgram& p, int nslots, const int64_t* offsets, int64_t n,
                  void* out_max, void* out_min, void* out_sum, int64_t* out_argmax, int64_t* out_argmin,
                  int8_t* out_mask, int nreq, const int32_t* req_n, const int64_t* const* req_tooffsets,
                  int64_t* const* req_carries);

}  // namespace akhip
