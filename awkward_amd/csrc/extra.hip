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

// A look-back scan over lists (scan.h's structure) whose tile first stages
// its mask bytes in LDS (4-byte coalesced loads of the tile's element span),
// so the per-list count and emit loops read LDS instead of walking global
// memory byte by byte; a tile whose span exceeds the stage reads global
// memory.  count(list) = selected elements; emit writes tooffsets[i + 1] and
// the selected elements' local positions.  3072-list tiles of 768 threads
// (61 KiB of LDS, two workgroups per CU): the look-back costs per tile, and
// on 50M lists of about 3 elements this shape measured 0.74 ms against 0.94
// for 1024-list tiles of 256 threads (0.80 at 512 x 4 and 512 x 3, 0.90 at
// 640 x 4, 0.92 at 1024 x 4, 1.00 at 896 x 4, 1.03 at 256 x 8).
constexpr int kB2IBlock = 768;
constexpr int kB2IIPT = 4;
constexpr int kB2ITile = kB2IBlock * kB2IIPT;  // lists per tile
constexpr int kB2IWords = 3072;                // 12 KiB of mask bytes
__global__ __launch_bounds__(kB2IBlock) void bool_to_index_kernel(int64_t* __restrict__ toindex,
                                                                  int64_t* __restrict__ tooffsets,
                                                                  const bool* __restrict__ mask,
                                                                  const int64_t* __restrict__ offsets, int64_t n,
                                                                  ScanScratch s, uint64_t* errword, uint32_t inv) {
  constexpr int NW = kB2IBlock / kWave;
  __shared__ int64_t s_off[kB2ITile + 1];
  __shared__ uint32_t s_mask[kB2IWords];
  __shared__ int64_t s_v[kB2ITile + kB2ITile / 32];
  __shared__ int64_t s_wave[NW];
  __shared__ int64_t s_prefix;
  __shared__ unsigned s_tile;
  auto pad = [](int p) { return p + (p >> 5); };
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wave = tid / kWave;

  if (tid == 0) s_tile = atomicAdd(s.counter, 1u);
  __syncthreads();
  const int64_t tile = s_tile;
  const int64_t L0 = tile * (int64_t)kB2ITile;
  const int nl = (int)((n - L0) < kB2ITile ? (n - L0) : kB2ITile);
  for (int t = tid; t <= nl; t += kB2IBlock) s_off[t] = offsets[L0 + t];
  __syncthreads();
  // the span's bytes, from the 4-byte word holding its first byte
  const uintptr_t first = (uintptr_t)(mask + s_off[0]);
  const uintptr_t abase = first & ~(uintptr_t)3;
  const int64_t words = (int64_t)(((uintptr_t)(mask + s_off[nl]) + 3 - abase) >> 2);
  const bool staged = s_off[nl] >= s_off[0] && words <= kB2IWords;
  if (staged)
    for (int w = tid; w < (int)words; w += kB2IBlock) s_mask[w] = ((const uint32_t*)abase)[w];
  __syncthreads();
  const uint8_t* lmask = (const uint8_t*)s_mask;
  auto selected = [&](int64_t j) -> bool {
    return staged ? lmask[(uintptr_t)(mask + j) - abase] != 0 : mask[j];
  };

  int64_t c[kB2IIPT];
#pragma unroll
  for (int u = 0; u < kB2IIPT; u++) {
    const int t = u * kB2IBlock + tid;
    int64_t cnt = 0;
    if (t < nl)
      for (int64_t j = s_off[t]; j < s_off[t + 1]; j++) cnt += selected(j) ? 1 : 0;
    c[u] = cnt;
    s_v[pad(t)] = cnt;
  }
  __syncthreads();
  int64_t b[kB2IIPT];
  int64_t total = 0;
#pragma unroll
  for (int v = 0; v < kB2IIPT; v++) {
    b[v] = s_v[pad(tid * kB2IIPT + v)];
    total += b[v];
  }
  const int64_t incl = wave_inclusive_scan(total);
  if (lane == kWave - 1) s_wave[wave] = incl;
  __syncthreads();
  int64_t wave_off = 0, block_total = 0;
#pragma unroll
  for (int w = 0; w < NW; w++) {
    const int64_t v = s_wave[w];
    wave_off += (w < wave) ? v : 0;
    block_total += v;
  }
  if (wave == 0) {
    const int64_t prefix = lookback_prefix(s, tile, block_total, errword, inv);
    if (lane == 0) s_prefix = prefix;
  }
  __syncthreads();
  int64_t run = s_prefix + wave_off + incl - total;
#pragma unroll
  for (int v = 0; v < kB2IIPT; v++) {
    s_v[pad(tid * kB2IIPT + v)] = run;
    run += b[v];
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kB2IIPT; u++) {
    const int t = u * kB2IBlock + tid;
    if (t < nl) {
      const int64_t i = L0 + t;
      const int64_t excl = s_v[pad(t)];
      if (i == 0) tooffsets[0] = 0;
      tooffsets[i + 1] = excl + c[u];
      int64_t k = excl;
      const int64_t a = s_off[t];
      if (c[u] > 0)
        for (int64_t j = a; j < s_off[t + 1]; j++)
          if (selected(j)) toindex[k++] = j - a;
    }
  }
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
  const int64_t ntiles = ceil_div(n, (int64_t)kB2ITile);
  void* base = scratch(scan_scratch_bytes(ntiles));
  if (!base) return failure("out of device memory for scan scratch", kSliceNone, kSliceNone, kFile);
  const ScanScratch s = scan_scratch_at(base);
  AKHIP_ERROR_WORD(ew);
  hipStream_t st = stream();
  scan_prep_kernel<0><<<(unsigned)ceil_div(ntiles + 1, 256), 256, 0, st>>>(s, ntiles);
  bool_to_index_kernel<<<(unsigned)ntiles, kB2IBlock, 0, st>>>(toindex, tooffsets, mask, offsets, n, s, ew, 0);
  AKHIP_LAUNCH_CHECK("bool_to_index_kernel");
  return success();
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
