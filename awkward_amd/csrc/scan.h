// scan.h -- single-pass exclusive scan with decoupled look-back (gfx950).
//
// Used wherever the CPU kernels build offsets with a running `k`
// (compact_offsets, combinations_length, getitem_jagged_apply,
// getitem_next_range, boolean-mask compaction).  One launch reads each input
// once and writes each output once:
//
//   * each workgroup claims a tile id from an atomic counter, so every tile
//     it waits on has already started (no dependence on dispatch order);
//   * the tile's items are counted (Body::count) striped over the threads
//     (item u * BLOCK + t), so the loads coalesce, and transposed through LDS
//     so that each thread then scans IPT consecutive items in registers;
//     across the wave with shuffles, then across waves through LDS;
//   * the tile publishes its aggregate, and wave 0 looks back over up to 64
//     predecessors at a time until it meets a published inclusive prefix;
//   * the exclusive prefixes go back through LDS and Body::emit(i,
//     exclusive_prefix, count) runs striped again, so neighbouring lanes
//     write neighbouring outputs.
//
// Status words pack a 2-bit flag with a 62-bit non-negative value, so one
// 64-bit atomic carries both and no other memory needs ordering: the loads
// and stores are relaxed atomics at agent scope.  (Acquire/release at agent
// scope would write back / invalidate the XCD's L2 on every poll, since L2
// is not coherent across the 8 XCDs -- that costs more than the whole scan.)
// The look-back spin is bounded: a tile that waited ~2^26 polls records an
// error instead of hanging the device.
#pragma once

#include "common.h"

namespace akhip {

constexpr uint64_t kScanAgg = uint64_t(1) << 62;
constexpr uint64_t kScanPrefix = uint64_t(2) << 62;
constexpr uint64_t kScanFlags = uint64_t(3) << 62;
constexpr uint64_t kScanValue = ~kScanFlags;
constexpr int kScanTimeoutCode = 250;

struct ScanScratch {
  unsigned* counter;
  uint64_t* status;
};

inline size_t scan_scratch_bytes(int64_t ntiles) { return 256 + (size_t)ntiles * sizeof(uint64_t); }

inline ScanScratch scan_scratch_at(void* base) {
  return ScanScratch{(unsigned*)base, (uint64_t*)((char*)base + 256)};
}

// Clears the tile counter and the status words before a scan (a kernel, not
// hipMemsetAsync, so the sequence can be captured into a hipGraph).
template <int Dummy = 0>
__global__ void scan_prep_kernel(ScanScratch s, int64_t ntiles) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) *s.counter = 0u;
  if (i < ntiles) s.status[i] = 0;
}

// Called by every lane of one wave of tile `tile`'s workgroup: publishes the
// tile's aggregate and returns (to every lane) the sum of all earlier tiles,
// publishing this tile's inclusive prefix in turn.
__device__ inline int64_t lookback_prefix(ScanScratch s, int64_t tile, int64_t block_total, uint64_t* errword,
                                          uint32_t inv) {
  const int lane = threadIdx.x & (kWave - 1);
  int64_t prefix = 0;
  if (tile == 0) {
    if (lane == 0)
      __hip_atomic_store(&s.status[0], kScanPrefix | (uint64_t)block_total, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    return 0;
  }
  if (lane == 0)
    __hip_atomic_store(&s.status[tile], kScanAgg | (uint64_t)block_total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  int64_t look = tile - 1;
  bool done = false;
  while (!done) {
    const int64_t idx = look - lane;
    uint64_t w = kScanPrefix;  // lanes before tile 0 read as an empty prefix
    if (idx >= 0) {
      uint32_t spins = 0;
      do {
        w = __hip_atomic_load(&s.status[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((w & kScanFlags) != 0) break;
        __builtin_amdgcn_s_sleep(1);
      } while (++spins < (1u << 26));
      if ((w & kScanFlags) == 0) {
        record_error(errword, inv, tile, kScanTimeoutCode);
        w = kScanPrefix;
      }
    }
    const uint64_t prefix_lanes = __ballot((w & kScanFlags) == kScanPrefix);
    int64_t v = (int64_t)(w & kScanValue);
    if (prefix_lanes != 0) {
      const int first = __builtin_ctzll(prefix_lanes);
      v = (lane <= first) ? v : 0;
      done = true;
    } else {
      look -= kWave;
    }
    prefix += wave_sum(v);
  }
  if (lane == 0)
    __hip_atomic_store(&s.status[tile], kScanPrefix | (uint64_t)(prefix + block_total), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  return prefix;
}

// The scan's tiles: 512 threads x 8 items.  Every tile claims its id with one
// atomic on a single counter, and those atomics serialise at about 11 ns each
// on MI355X: at 2048-item tiles the claims alone took 0.28 ms of a 50M-item
// scan.  At 4096 items the tile's LDS (33 KiB) lets four workgroups share a
// CU where 8192-item tiles (67 KiB) allowed two: measured on the 50M-list
// event workload, jagged_apply 0.99 -> 0.79 ms, next_range 0.46 -> 0.39 ms,
// combinations_length 0.14 -> 0.13 ms (2048-item tiles: 0.89, 0.49, 0.17 ms).
constexpr int kScanBlock = 512;
constexpr int kScanIpt = 8;
constexpr int kScanTile = kScanBlock * kScanIpt;

inline int64_t scan_tiles(int64_t n) { return ceil_div(n > 0 ? n : 1, (int64_t)kScanTile); }

// Per-item counts and then exclusive prefixes live in LDS (the prefix of the
// next item minus an item's prefix is its count, so the emits need no
// per-thread arrays): counts are taken striped (item u * BLOCK + t: the
// loads coalesce), each thread sums IPT consecutive items, the tile is
// scanned across waves, and the prefixes are written back in place.
template <int BLOCK, int IPT, class Body>
__global__ __launch_bounds__(BLOCK) void lookback_scan_kernel(Body body, ScanScratch s, int64_t n,
                                                              uint64_t* errword, uint32_t inv) {
  constexpr int NW = BLOCK / kWave;
  constexpr int TILE = BLOCK * IPT;
  __shared__ int64_t s_wave[NW];
  __shared__ int64_t s_prefix;
  __shared__ unsigned s_tile;
  __shared__ int64_t s_v[TILE + TILE / 32 + 1];  // one pad slot per 32 values, then the tile's end prefix
  auto pad = [](int p) { return p + (p >> 5); };
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wave = tid / kWave;

  if (tid == 0) s_tile = atomicAdd(s.counter, 1u);
  __syncthreads();
  const int64_t tile = s_tile;
  const int64_t tbase = tile * (int64_t)TILE;

#pragma unroll 8
  for (int u = 0; u < IPT; u++) {
    const int64_t i = tbase + (int64_t)u * BLOCK + tid;
    s_v[pad(u * BLOCK + tid)] = (i < n) ? body.count(i) : 0;
  }
  __syncthreads();
  int64_t total = 0;
#pragma unroll
  for (int v = 0; v < IPT; v++) total += s_v[pad(tid * IPT + v)];
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
  for (int v = 0; v < IPT; v++) {
    const int q = pad(tid * IPT + v);
    const int64_t c = s_v[q];
    s_v[q] = run;
    run += c;
  }
  if (tid == BLOCK - 1) s_v[pad(TILE)] = run;
  __syncthreads();
#pragma unroll 4
  for (int u = 0; u < IPT; u++) {
    const int t = u * BLOCK + tid;
    const int64_t i = tbase + t;
    if (i < n) {
      const int64_t e = s_v[pad(t)];
      body.emit(i, e, s_v[pad(t + 1)] - e);
    }
  }
}

// Enqueue the scan of n items (n >= 1) on the current stream; the scratch
// holds scan_scratch_bytes(scan_tiles(n)) bytes.
template <class Body>
Error launch_lookback_scan(const Body& body, int64_t n, uint32_t inv, void* scratch_base) {
  const int64_t ntiles = scan_tiles(n);
  ScanScratch s = scan_scratch_at(scratch_base);
  AKHIP_ERROR_WORD(ew);
  hipStream_t st = stream();
  scan_prep_kernel<0><<<(unsigned)ceil_div(ntiles + 1, 256), 256, 0, st>>>(s, ntiles);
  lookback_scan_kernel<kScanBlock, kScanIpt, Body><<<(unsigned)ntiles, kScanBlock, 0, st>>>(body, s, n, ew, inv);
  AKHIP_LAUNCH_CHECK("lookback_scan_kernel");
  return success();
}

}  // namespace akhip
