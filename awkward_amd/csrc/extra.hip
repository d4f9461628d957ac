// extra.hip -- array operations the layout code performs through its nplike
// rather than through a registry kernel, written as device kernels for the
// hip nplike:
//
//   * awkward_hip_take_{1,2,4,8}: gather out[i] = data[index[i]] by element
//     size (NumpyArray._carry, src/awkward/contents/numpyarray.py:367-373);
//   * awkward_hip_bool_to_index: a jagged boolean mask to the jagged integer
//     index of its true positions, in one scan (the localindex + flatten +
//     boolean compaction + cumsum + gather sequence of
//     src/awkward/_slicing.py:476-530);
//   * awkward_hip_subtract_i64: stops - starts (ak.num, operations/ak_num.py:133-142).

#include "common.h"
#include "scan.h"

namespace akhip {
namespace {

const char* kFile = "awkward_amd/csrc/extra.hip";

template <typename W>
__global__ void take_kernel(W* __restrict__ out, const W* __restrict__ data, const int64_t* __restrict__ index,
                            int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = data[index[i]];
}

template <typename W>
Error take(void* out, const void* data, const int64_t* index, int64_t n) {
  if (n <= 0) return success();
  take_kernel<W><<<grid_for(n, 256, 8), 256, 0, stream()>>>((W*)out, (const W*)data, index, n);
  AKHIP_LAUNCH_CHECK("take_kernel");
  return success();
}

struct BoolToIndexBody {
  int64_t* toindex;
  int64_t* tooffsets;
  const bool* mask;
  const int64_t* offsets;
  int64_t base;  // offsets[0] is read on the device: positions are relative to it
  __device__ int64_t count(int64_t i) const {
    int64_t c = 0;
    for (int64_t j = offsets[i]; j < offsets[i + 1]; j++) c += mask[j] ? 1 : 0;
    return c;
  }
  __device__ void emit(int64_t i, int64_t excl, int64_t cnt) const {
    if (i == 0) tooffsets[0] = 0;
    tooffsets[i + 1] = excl + cnt;
    if (cnt == 0) return;
    const int64_t a = offsets[i], b = offsets[i + 1];
    int64_t k = excl;
    for (int64_t j = a; j < b; j++)
      if (mask[j]) toindex[k++] = j - a;
  }
};

__global__ void sub_kernel(int64_t* __restrict__ out, const int64_t* __restrict__ a, const int64_t* __restrict__ b,
                           int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = a[i] - b[i];
}

__global__ void zero1(int64_t* p) { *p = 0; }

}  // namespace
}  // namespace akhip

using namespace akhip;

extern "C" {

Error awkward_hip_take_1(void* out, const void* data, const int64_t* index, int64_t n) {
  return take<uint8_t>(out, data, index, n);
}
Error awkward_hip_take_2(void* out, const void* data, const int64_t* index, int64_t n) {
  return take<uint16_t>(out, data, index, n);
}
Error awkward_hip_take_4(void* out, const void* data, const int64_t* index, int64_t n) {
  return take<uint32_t>(out, data, index, n);
}
Error awkward_hip_take_8(void* out, const void* data, const int64_t* index, int64_t n) {
  return take<uint64_t>(out, data, index, n);
}

// toindex must hold at least as many entries as the mask has true values in
// [offsets[0], offsets[n]); tooffsets (n + 1) starts at 0.
Error awkward_hip_bool_to_index(int64_t* toindex, int64_t* tooffsets, const bool* mask, const int64_t* offsets,
                                int64_t n) {
  if (n <= 0) {
    zero1<<<1, 1, 0, stream()>>>(tooffsets);
    AKHIP_LAUNCH_CHECK("zero1");
    return success();
  }
  const int64_t ntiles = ceil_div(n, (int64_t)256 * 8);
  void* s = scratch(scan_scratch_bytes(ntiles));
  if (!s) return failure("out of device memory for scan scratch", kSliceNone, kSliceNone, kFile);
  BoolToIndexBody body{toindex, tooffsets, mask, offsets, 0};
  return launch_lookback_scan<256, 8>(body, n, 0, s);
}

Error awkward_hip_subtract_i64(int64_t* out, const int64_t* a, const int64_t* b, int64_t n) {
  if (n <= 0) return success();
  sub_kernel<<<grid_for(n, 256, 8), 256, 0, stream()>>>(out, a, b, n);
  AKHIP_LAUNCH_CHECK("sub_kernel");
  return success();
}

}  // extern "C"
