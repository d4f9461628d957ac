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

// A plain streaming read (16-byte loads, one per thread): the device's own
// read rate for a buffer, the ceiling a reduction over it is measured
// against.  Each loaded word is compared with a key known only at run time,
// so the compiler cannot drop the loads; the host passes a key with bits
// that no folded word has, so `sink` is never written.
__global__ __launch_bounds__(256) void stream_read_kernel(const uint4* __restrict__ p, int64_t n16, uint64_t key,
                                                          uint64_t* __restrict__ sink) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n16) return;
  const uint4 v = p[i];
  const uint64_t x = ((uint64_t)(v.x ^ v.z) << 32) | (uint64_t)(v.y ^ v.w);
  if (x == key) sink[0] = x;
}

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

// Measurement: reads the first nbytes / 16 * 16 bytes of `data` (16-byte
// aligned); `sink` (8 bytes) is written only if a 16-byte word folds to one
// fixed 64-bit pattern.
Error awkward_hip_stream_read(const void* data, int64_t nbytes, uint64_t* sink) {
  const int64_t n16 = nbytes / 16;
  if (n16 <= 0) return success();
  if (((uintptr_t)data & 15) != 0) return failure("stream_read needs 16-byte aligned data", kSliceNone, kSliceNone, kFile);
  stream_read_kernel<<<(unsigned)ceil_div(n16, (int64_t)256), 256, 0, stream()>>>((const uint4*)data, n16,
                                                                                     0x0123456789abcdefull, sink);
  AKHIP_LAUNCH_CHECK("stream_read_kernel");
  return success();
}

Error awkward_hip_subtract_i64(int64_t* out, const int64_t* a, const int64_t* b, int64_t n) {
  if (n <= 0) return success();
  sub_kernel<<<grid_for(n, 256, 8), 256, 0, stream()>>>(out, a, b, n);
  AKHIP_LAUNCH_CHECK("sub_kernel");
  return success();
}

}  // extern "C"
