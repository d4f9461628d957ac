// nonlocal.hip -- reductions at an axis above the innermost lists
// (axis != -1): the column regrouping of ListOffsetArray._reduce_next's
// non-local branch.
//
// Replaces awkward-cpp/src/cpu-kernels/
//   awkward_ListOffsetArray_reduce_nonlocal_maxcount_offsetscopy_64.cpp
//   awkward_ListOffsetArray_reduce_nonlocal_preparenext_64.cpp
//   awkward_ListOffsetArray_reduce_nonlocal_nextshifts_64.cpp
//   awkward_ListOffsetArray_reduce_nonlocal_outstartsstops_64.cpp
//
// The CPU loops walk, per outer bin, every column j < maxcount and every row
// of the bin, emitting the element in column j of each row long enough:
// the bin's elements regrouped column-major, rows in order within a column.
// On the device that order is a stable segmented sort of the elements by
// column, one segment per outer bin (the library's own argsort), and a
// histogram of (bin, column) groups scanned into their offsets.  An element
// at sorted position t of bin b then lands at nextcarry index
// t - segment start + (offset of group (b, 0)); its rank r within its
// column group gives the rows of the bin before it that lack that column,
// (row - first row of the bin) - r, which is nextshifts' `missing`.
//
// For maxcount <= 256 (the usual multiplicities) no sort is needed: the rank
// of the element in column c of row r within its (bin, column) group is
// S(r, c) - S(first row of the bin, c), where S(r, c) counts the rows before
// r longer than c.  Rows go 256 to a workgroup, a row per thread; a wave
// ballot per column gives the in-wave counts, a scan over the workgroups per
// column the rest (colranks below).

#include "errors.h"
#include "scan.h"

extern "C" Error awkward_argsort_int64_int64(int64_t* toptr, const int64_t* fromptr, int64_t length,
                                             const int64_t* offsets, int64_t offsetslength, bool ascending,
                                             bool stable);
extern "C" Error awkward_argsort_int32_int64(int64_t* toptr, const int32_t* fromptr, int64_t length,
                                             const int64_t* offsets, int64_t offsetslength, bool ascending,
                                             bool stable);

namespace akhip {
namespace {

const char* kFile = "awkward_amd/csrc/nonlocal.hip";
constexpr int kBlock = 256;

inline size_t round256(size_t b) { return (b + 255) / 256 * 256; }

// first index in [lo, hi) with a[idx] > v (hi if none)
__device__ inline int64_t upper_bound(const int64_t* a, int64_t lo, int64_t hi, int64_t v) {
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (a[mid] <= v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__global__ void set_i64(int64_t* p, int64_t v) { *p = v; }

__global__ __launch_bounds__(kBlock) void fill_kernel(int64_t* p, int64_t n, int64_t v) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) p[i] = v;
}

// ---- maxcount_offsetscopy ---------------------------------------------------------
__global__ __launch_bounds__(kBlock) void maxcount_kernel(unsigned long long* maxcount, int64_t* offsetscopy,
                                                          const int64_t* offsets, int64_t length) {
  __shared__ unsigned long long s_max[kBlock / kWave];
  unsigned long long m = 0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < length; i += (int64_t)gridDim.x * kBlock) {
    const int64_t c = offsets[i + 1] - offsets[i];
    if (c > 0 && (unsigned long long)c > m) m = (unsigned long long)c;
    offsetscopy[i + 1] = offsets[i + 1];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) offsetscopy[0] = offsets[0];
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) {
    const unsigned long long o = __shfl_xor(m, d, kWave);
    m = o > m ? o : m;
  }
  if ((threadIdx.x & (kWave - 1)) == 0) s_max[threadIdx.x / kWave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kBlock / kWave; w++) m = s_max[w] > m ? s_max[w] : m;
    if (m > 0) atomicMax(maxcount, m);
  }
}

// ---- outstartsstops -----------------------------------------------------------------
// group k covers distincts[k * maxcount, (k + 1) * maxcount); its stop is one
// past the last entry that is not -1 (or its start)
__global__ __launch_bounds__(kBlock) void outstartsstops_kernel(int64_t* outstarts, int64_t* outstops,
                                                                const int64_t* distincts, int64_t lendistincts,
                                                                int64_t outlength, int64_t maxcount) {
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k >= outlength) return;
  const int64_t a = k * maxcount;
  if (a >= lendistincts) return;
  const int64_t b = a + maxcount < lendistincts ? a + maxcount : lendistincts;
  int64_t stop = a;
  for (int64_t i = b - 1; i >= a; i--)
    if (distincts[i] != -1) {
      stop = i + 1;
      break;
    }
  outstarts[k] = a;
  outstops[k] = stop;
}

// the CPU loop itself, for the degenerate split lendistincts < outlength
__global__ void outstartsstops_serial(int64_t* outstarts, int64_t* outstops, const int64_t* distincts,
                                      int64_t lendistincts, int64_t outlength) {
  const int64_t maxcount = lendistincts / outlength;
  int64_t k = 0, i_next = 0;
  for (int64_t i = 0; i < lendistincts; i++) {
    if (i == i_next) {
      i_next += maxcount;
      if (k < outlength) {
        outstarts[k] = i;
        outstops[k] = i;
      }
      k++;
    }
    if (distincts[i] != -1 && k - 1 < outlength) outstops[k - 1] = i + 1;
  }
}

// ---- the column-major regrouping ------------------------------------------------------
struct ColMajor {
  int64_t row0, row1;  // rows of the outer bins: [outer_offsets[0], outer_offsets[outlength])
  int64_t E0, n;       // their elements: [offsets[row0], offsets[row1])
  int64_t groups;      // outlength * maxcount
  // n: column of each element (maxcount: past the last column); 32-bit keys
  // (the sort's narrow path) whenever maxcount fits
  int32_t* keys32;
  int64_t* keys64;
  int64_t* segoff;     // outlength + 1: each bin's first element, relative to E0
  int64_t* order;      // n: argsort of keys within each bin (local positions)
  int64_t* cnt;        // groups: elements per (bin, column)
  int64_t* prefix;     // groups + 1: prefix[g] = elements of the groups before g
  unsigned long long* maxgroup;  // 1 + the last non-empty group (0: none)
  int64_t last_bin_row;          // outer_offsets[outlength - 1] (host)
  __device__ int64_t key(int64_t t) const { return keys32 ? (int64_t)keys32[t] : keys64[t]; }
};

// Element keys, 2048 consecutive elements per workgroup iteration: the rows
// they fall in are found once per iteration, each element's row by a
// search within them.
constexpr int kKeysTile = kBlock * 8;
__global__ __launch_bounds__(kBlock) void keys_kernel(ColMajor cm, const int64_t* offsets, int64_t maxcount) {
  __shared__ int64_t s_rows[2];
  for (int64_t t0 = (int64_t)blockIdx.x * kKeysTile; t0 < cm.n; t0 += (int64_t)gridDim.x * kKeysTile) {
    const int64_t t1 = t0 + kKeysTile < cm.n ? t0 + kKeysTile : cm.n;
    if (threadIdx.x < 2) {
      const int64_t e = cm.E0 + (threadIdx.x == 0 ? t0 : t1 - 1);
      s_rows[threadIdx.x] = upper_bound(offsets, cm.row0, cm.row1 + 1, e) - 1;
    }
    __syncthreads();
    const int64_t ra = s_rows[0], rb = s_rows[1];
    __syncthreads();
    for (int64_t t = t0 + threadIdx.x; t < t1; t += kBlock) {
      const int64_t e = cm.E0 + t;
      const int64_t row = upper_bound(offsets, ra, rb + 1, e) - 1;
      const int64_t col = e - offsets[row];
      const int64_t k = col < maxcount ? col : maxcount;
      if (cm.keys32) cm.keys32[t] = (int32_t)k;
      else cm.keys64[t] = k;
    }
  }
}

// cnt[b * maxcount + c] = rows of bin b longer than c.  A workgroup takes
// 2048 consecutive rows; their bins [b0, b0 + span) share an LDS histogram of
// row lengths clipped to CL = min(maxcount, kHist / span - 1), turned into
// "longer than c" counts by a suffix sum and added to cnt once per
// workgroup -- so rows of one bin do not all hit the same counters.  The
// columns [CL, length) of rows longer than CL (few, unless the bins are many,
// when the counters they hit are many too) go to cnt directly, the lanes of
// a wave splitting each such row.
constexpr int kHist = 8192;
constexpr int kCountRows = 2048;
__global__ __launch_bounds__(kBlock) void colcount_kernel(ColMajor cm, const int64_t* offsets,
                                                          const int64_t* outer_offsets, int64_t outlength,
                                                          int64_t maxcount) {
  __shared__ unsigned int hist[kHist];
  __shared__ int64_t s_bins[2];
  __shared__ int s_ml;  // the longest clipped length of the rows (bounds the suffix sums)
  const int lane = threadIdx.x & (kWave - 1);
  for (int64_t r0 = cm.row0 + (int64_t)blockIdx.x * kCountRows; r0 < cm.row1;
       r0 += (int64_t)gridDim.x * kCountRows) {
    const int64_t r1 = r0 + kCountRows < cm.row1 ? r0 + kCountRows : cm.row1;
    if (threadIdx.x < 2) s_bins[threadIdx.x] = upper_bound(outer_offsets, 0, outlength + 1,
                                                           threadIdx.x == 0 ? r0 : r1 - 1) - 1;
    __syncthreads();
    const int64_t b0 = s_bins[0], span = s_bins[1] - s_bins[0] + 1;
    __syncthreads();
    int64_t cl = kHist / span - 1;
    if (cl > maxcount) cl = maxcount;
    if (cl < 0) cl = 0;
    const int64_t w = cl + 1;  // histogram entries per bin: clipped lengths 0 .. cl
    if (cl > 0)
      for (int64_t i = threadIdx.x; i < span * w; i += kBlock) hist[i] = 0u;
    if (threadIdx.x == 0) s_ml = 0;
    __syncthreads();
    for (int64_t rbase = r0; rbase < r1; rbase += kBlock) {
      const int64_t r = rbase + threadIdx.x;
      const bool valid = r < r1;
      int64_t b = 0, len = 0;
      if (valid) {
        b = upper_bound(outer_offsets, b0, b0 + span + 1 <= outlength + 1 ? b0 + span + 1 : outlength + 1, r) - 1;
        len = offsets[r + 1] - offsets[r];
        len = len < 0 ? 0 : len > maxcount ? maxcount : len;
        if (cl > 0) {
          const int64_t k = len < cl ? len : cl;
          atomicAdd(&hist[(b - b0) * w + k], 1u);
          atomicMax(&s_ml, (int)k);
        }
      }
      uint64_t over = __ballot(valid && len > cl);
      while (over) {
        const int src = __builtin_ctzll(over);
        over &= over - 1;
        const int64_t ob = __shfl(b, src, kWave), olen = __shfl(len, src, kWave);
        for (int64_t c = cl + lane; c < olen; c += kWave)
          atomicAdd((unsigned long long*)&cm.cnt[ob * maxcount + c], 1ull);
      }
    }
    __syncthreads();
    const int64_t ml = s_ml;
    if (ml > 0) {
      // per bin: entry c + 1 becomes the rows of length > c (lengths clipped at
      // cl; entries above ml are zero)
      for (int64_t bb = threadIdx.x; bb < span; bb += kBlock) {
        unsigned int run = 0;
        for (int64_t c = ml - 1; c >= 0; c--) {
          run += hist[bb * w + c + 1];
          hist[bb * w + c + 1] = run;
        }
      }
      __syncthreads();
      for (int64_t i = threadIdx.x; i < span * ml; i += kBlock) {
        const int64_t bb = i / ml, c = i - bb * ml;
        const unsigned int v = hist[bb * w + c + 1];
        if (v) atomicAdd((unsigned long long*)&cm.cnt[(b0 + bb) * maxcount + c], (unsigned long long)v);
      }
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(kBlock) void segoff_kernel(ColMajor cm, const int64_t* offsets,
                                                        const int64_t* outer_offsets, int64_t outlength) {
  const int64_t b = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (b <= outlength) cm.segoff[b] = offsets[outer_offsets[b]] - cm.E0;
}

__global__ void bounds_kernel(int64_t* out, const int64_t* offsets, const int64_t* outer_offsets, int64_t outlength) {
  const int64_t r0 = outer_offsets[0], r1 = outer_offsets[outlength];
  out[0] = r0;
  out[1] = r1;
  out[2] = offsets[r0];
  out[3] = offsets[r1];
  out[4] = outer_offsets[outlength > 0 ? outlength - 1 : 0];
}

// The bins' row range, its element range and the last bin's first row, in
// one host round trip.
Error read_bounds(ColMajor& cm, const int64_t* offsets, const int64_t* outer_offsets, int64_t outlength) {
  hipStream_t st = stream();
  int64_t* d = (int64_t*)scratch(256);
  if (!d) return failure("out of device memory for scratch", kSliceNone, kSliceNone, kFile);
  int64_t h[5] = {0, 0, 0, 0, 0};
  bounds_kernel<<<1, 1, 0, st>>>(d, offsets, outer_offsets, outlength);
  hipError_t err = hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, st);
  if (err == hipSuccess) err = hipStreamSynchronize(st);
  if (err != hipSuccess) return hip_failure(err, "reading the column regrouping's bounds");
  cm.row0 = h[0];
  cm.row1 = h[1];
  cm.E0 = h[2];
  cm.n = h[3] - h[2];
  cm.last_bin_row = h[4];
  return success();
}

struct PrefixBody {
  const int64_t* cnt;
  int64_t* prefix;
  unsigned long long* maxgroup;
  __device__ int64_t count(int64_t g) const { return cnt[g]; }
  __device__ void emit(int64_t g, int64_t excl, int64_t c) const {
    if (g == 0) prefix[0] = 0;
    prefix[g + 1] = excl + c;
    if (c > 0) atomicMax(maxgroup, (unsigned long long)(g + 1));
  }
};

// Fills every array of `cm` (host values included).  Arrays live in the
// work-queue pool: the argsort and the scan use the scratch pool.
Error colmajor(ColMajor& cm, const int64_t* offsets, const int64_t* outer_offsets, int64_t outlength,
               int64_t maxcount) {
  hipStream_t st = stream();
  {
    Error e = read_bounds(cm, offsets, outer_offsets, outlength);
    if (e.str) return e;
  }
  if (cm.n < 0) cm.n = 0;
  cm.groups = outlength * maxcount;
  const bool narrow = maxcount < INT32_MAX;
  const size_t nb = round256((size_t)cm.n * 8), kb = round256((size_t)cm.n * (narrow ? 4 : 8)),
               sb = round256((size_t)(outlength + 1) * 8), gb = round256((size_t)cm.groups * 8),
               pb = round256((size_t)(cm.groups + 1) * 8);
  char* q = (char*)queue_pool(256 + nb + kb + sb + gb + pb);
  if (!q) return failure("out of device memory for the column regrouping", kSliceNone, kSliceNone, kFile);
  cm.maxgroup = (unsigned long long*)q;
  cm.keys32 = narrow ? (int32_t*)(q + 256 + nb + sb + gb + pb) : nullptr;
  cm.keys64 = narrow ? nullptr : (int64_t*)(q + 256 + nb + sb + gb + pb);
  cm.order = (int64_t*)(q + 256);
  cm.segoff = (int64_t*)(q + 256 + nb);
  cm.cnt = (int64_t*)(q + 256 + nb + sb);
  cm.prefix = (int64_t*)(q + 256 + nb + sb + gb);
  set_i64<<<1, 1, 0, st>>>((int64_t*)cm.maxgroup, 0);
  if (cm.groups > 0) fill_kernel<<<grid_for(cm.groups, kBlock, 8), kBlock, 0, st>>>(cm.cnt, cm.groups, 0);
  segoff_kernel<<<(unsigned)ceil_div(outlength + 1, (int64_t)kBlock), kBlock, 0, st>>>(cm, offsets, outer_offsets,
                                                                                       outlength);
  if (cm.row1 > cm.row0 && cm.groups > 0) {
    int64_t grid = ceil_div(cm.row1 - cm.row0, (int64_t)kCountRows);
    if (grid > (int64_t)cu_count() * 4) grid = (int64_t)cu_count() * 4;
    colcount_kernel<<<(unsigned)grid, kBlock, 0, st>>>(cm, offsets, outer_offsets, outlength, maxcount);
    AKHIP_LAUNCH_CHECK("colcount_kernel");
  }
  if (cm.n > 0) {
    int64_t grid = ceil_div(cm.n, (int64_t)kKeysTile);
    if (grid > (int64_t)cu_count() * 16) grid = (int64_t)cu_count() * 16;
    keys_kernel<<<(unsigned)grid, kBlock, 0, st>>>(cm, offsets, maxcount);
    AKHIP_LAUNCH_CHECK("keys_kernel");
    Error e = cm.keys32 ? awkward_argsort_int32_int64(cm.order, cm.keys32, cm.n, cm.segoff, outlength + 1, true, true)
                        : awkward_argsort_int64_int64(cm.order, cm.keys64, cm.n, cm.segoff, outlength + 1, true, true);
    if (e.str) return e;
  }
  set_i64<<<1, 1, 0, st>>>(cm.prefix, 0);
  if (cm.groups > 0) {
    const int64_t ntiles = scan_tiles(cm.groups);
    void* s = scratch(scan_scratch_bytes(ntiles));
    if (!s) return failure("out of device memory for scan scratch", kSliceNone, kSliceNone, kFile);
    Error e = launch_lookback_scan(PrefixBody{cm.cnt, cm.prefix, cm.maxgroup}, cm.groups, 0, s);
    if (e.str) return e;
  }
  AKHIP_LAUNCH_CHECK("colmajor");
  return success();
}

// sorted position t -> (bin, element, column)
struct Sorted {
  int64_t bin, e, col, local;
};
__device__ inline Sorted sorted_at(const ColMajor& cm, int64_t t, int64_t outlength) {
  const int64_t b = upper_bound(cm.segoff, 0, outlength + 1, t) - 1;
  const int64_t local = cm.order[t];
  const int64_t rel = cm.segoff[b] + local;
  return Sorted{b, cm.E0 + rel, cm.key(rel), t - cm.segoff[b]};
}

__global__ __launch_bounds__(kBlock) void place_kernel(ColMajor cm, int64_t* nextcarry, int64_t outlength,
                                                       int64_t maxcount) {
  for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < cm.n; t += (int64_t)gridDim.x * kBlock) {
    const Sorted s = sorted_at(cm, t, outlength);
    if (s.col < maxcount) nextcarry[cm.prefix[s.bin * maxcount] + s.local] = s.e;
  }
}

__global__ __launch_bounds__(kBlock) void groups_out_kernel(ColMajor cm, int64_t* nextoffsets, int64_t* distincts,
                                                            int64_t distinctslen) {
  const int64_t hi = cm.groups + 1 > distinctslen ? cm.groups + 1 : distinctslen;
  for (int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x; g < hi; g += (int64_t)gridDim.x * kBlock) {
    if (g <= cm.groups) nextoffsets[g] = cm.prefix[g];
    if (g < distinctslen) distincts[g] = (g < cm.groups && cm.cnt[g] > 0) ? g : -1;
  }
}

__global__ void maxgroup_out(const ColMajor cm, int64_t* maxnextparents) {
  *maxnextparents = (int64_t)*cm.maxgroup - 1;
}

__global__ __launch_bounds__(kBlock) void missing_kernel(ColMajor cm, int64_t* missing, const int64_t* offsets,
                                                         const int64_t* outer_offsets, int64_t outlength,
                                                         int64_t maxcount) {
  for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < cm.n; t += (int64_t)gridDim.x * kBlock) {
    const Sorted s = sorted_at(cm, t, outlength);
    if (s.col >= maxcount) continue;
    const int64_t g0 = s.bin * maxcount;
    const int64_t rank = s.local - (cm.prefix[g0 + s.col] - cm.prefix[g0]);
    const int64_t row = upper_bound(offsets, cm.row0, cm.row1 + 1, s.e) - 1;
    missing[s.e] = (row - outer_offsets[s.bin]) - rank;
  }
}

__global__ __launch_bounds__(kBlock) void gather_kernel(int64_t* out, const int64_t* src, const int64_t* index,
                                                        int64_t n) {
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < n; k += (int64_t)gridDim.x * kBlock)
    out[k] = src[index[k]];
}

// nummissing[k] after the CPU loop: rows of the last bin shorter than k + 1.
// A histogram of the short rows' lengths, per workgroup in LDS for lengths
// below kHist (the common ones), added to `hist` once per workgroup.
__global__ __launch_bounds__(kBlock) void short_rows_kernel(int64_t* hist, const int64_t* offsets, int64_t r0,
                                                            int64_t r1, int64_t maxcount) {
  __shared__ unsigned int h[kHist];
  const int64_t lim = maxcount < kHist ? maxcount : kHist;
  for (int64_t i = threadIdx.x; i < lim; i += kBlock) h[i] = 0u;
  __syncthreads();
  for (int64_t r = r0 + (int64_t)blockIdx.x * kBlock + threadIdx.x; r < r1; r += (int64_t)gridDim.x * kBlock) {
    int64_t c = offsets[r + 1] - offsets[r];
    c = c > 0 ? c : 0;
    if (c < lim) atomicAdd(&h[c], 1u);
    else if (c < maxcount) atomicAdd((unsigned long long*)&hist[c], 1ull);
  }
  __syncthreads();
  for (int64_t i = threadIdx.x; i < lim; i += kBlock)
    if (h[i]) atomicAdd((unsigned long long*)&hist[i], (unsigned long long)h[i]);
}
__global__ void nummissing_kernel(int64_t* nummissing, const int64_t* hist, int64_t maxcount) {
  int64_t run = 0;
  for (int64_t k = 0; k < maxcount; k++) {
    run += hist[k];
    nummissing[k] = run;
  }
}

// ---- sort-free regrouping (maxcount <= kFastMax) ------------------------------------
constexpr int kFastMax = 256;
constexpr int kRowsPerGroup = kBlock;  // one row per thread

// Shared by the passes over row groups: the group's rows, a row per thread,
// with lengths clipped to [0, maxcount]; the per-wave counts of rows longer
// than c (c below the group's longest row) in s_w, and that bound in s_ml.
struct RowGroup {
  int64_t r;    // this thread's row (valid when r < row1)
  int64_t len;  // clipped length (0 past row1)
  int ml;       // longest clipped length in the workgroup
};

__device__ inline RowGroup row_group(const int64_t* offsets, int64_t row0, int64_t row1, int64_t maxcount, int64_t j,
                                     unsigned int (*s_w)[kFastMax], int* s_ml) {
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  RowGroup g;
  g.r = row0 + j * kRowsPerGroup + threadIdx.x;
  g.len = 0;
  if (g.r < row1) {
    const int64_t l = offsets[g.r + 1] - offsets[g.r];
    g.len = l < 0 ? 0 : l > maxcount ? maxcount : l;
  }
  if (threadIdx.x == 0) *s_ml = 0;
  __syncthreads();
  if (g.len > 0) atomicMax(s_ml, (int)g.len);
  __syncthreads();
  g.ml = *s_ml;
  for (int c = 0; c < g.ml; c++) {
    const uint64_t bal = __ballot(g.len > c);
    if (lane == 0) s_w[wave][c] = (unsigned int)__popcll(bal);
  }
  __syncthreads();
  return g;
}

// chunkpre[c * (ngroups + 1) + j]: rows of group j longer than c (scanned in
// place over j by colscan_kernel)
__global__ __launch_bounds__(kBlock) void colgroup_kernel(int64_t* chunkpre, const int64_t* offsets, int64_t row0,
                                                          int64_t row1, int64_t maxcount, int64_t ngroups) {
  __shared__ unsigned int s_w[kBlock / kWave][kFastMax];
  __shared__ int s_ml;
  const int64_t j = blockIdx.x;
  const RowGroup g = row_group(offsets, row0, row1, maxcount, j, s_w, &s_ml);
  for (int c = threadIdx.x; c < maxcount; c += kBlock) {
    int64_t t = 0;
    if (c < g.ml)
      for (int w = 0; w < kBlock / kWave; w++) t += s_w[w][c];
    chunkpre[(int64_t)c * (ngroups + 1) + j] = t;
  }
}

// per column c (one workgroup each): exclusive scan over the groups, the
// total at position ngroups
__global__ __launch_bounds__(kBlock) void colscan_kernel(int64_t* chunkpre, int64_t ngroups) {
  __shared__ int64_t s_sum[kBlock];
  int64_t* col = chunkpre + (int64_t)blockIdx.x * (ngroups + 1);
  const int64_t per = (ngroups + kBlock - 1) / kBlock;
  const int64_t a = (int64_t)threadIdx.x * per;
  const int64_t b = a + per < ngroups ? a + per : ngroups;
  int64_t local = 0;
  for (int64_t i = a; i < b; i++) local += col[i];
  s_sum[threadIdx.x] = local;
  __syncthreads();
  for (int d = 1; d < kBlock; d <<= 1) {
    const int64_t v = threadIdx.x >= d ? s_sum[threadIdx.x - d] : 0;
    __syncthreads();
    s_sum[threadIdx.x] += v;
    __syncthreads();
  }
  int64_t run = s_sum[threadIdx.x] - local;
  for (int64_t i = a; i < b; i++) {
    const int64_t v = col[i];
    col[i] = run;
    run += v;
  }
  if (threadIdx.x == kBlock - 1) col[ngroups] = s_sum[kBlock - 1];
}

// S(r, c) of this thread's row for column c < g.ml; calls ballot, so every
// lane of the wave calls it with the same c
__device__ inline int64_t rows_before(const RowGroup& g, int c, const int64_t* chunkpre, int64_t ngroups, int64_t j,
                                      unsigned int (*s_w)[kFastMax]) {
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const uint64_t bal = __ballot(g.len > c);
  int64_t s = chunkpre[(int64_t)c * (ngroups + 1) + j] + __popcll(bal & ((1ull << lane) - 1ull));
  for (int w = 0; w < wave; w++) s += s_w[w][c];
  return s;
}

// binS[b * maxcount + c] = S(outer_offsets[b], c) for the bins starting at a
// row of the group (several when bins are empty)
__global__ __launch_bounds__(kBlock) void binstart_kernel(int64_t* binS, const int64_t* chunkpre,
                                                          const int64_t* offsets, const int64_t* outer_offsets,
                                                          int64_t outlength, int64_t row0, int64_t row1,
                                                          int64_t maxcount, int64_t ngroups) {
  __shared__ unsigned int s_w[kBlock / kWave][kFastMax];
  __shared__ int s_ml;
  const int64_t j = blockIdx.x;
  const RowGroup g = row_group(offsets, row0, row1, maxcount, j, s_w, &s_ml);
  int64_t b0 = outlength;  // first bin starting at this row (outlength: none)
  if (g.r < row1) {
    int64_t lo = 0, hi = outlength;  // first b with outer_offsets[b] >= r
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (outer_offsets[mid] < g.r) lo = mid + 1;
      else hi = mid;
    }
    if (lo < outlength && outer_offsets[lo] == g.r) b0 = lo;
  }
  if (!__any(b0 < outlength)) return;  // no bin starts at this wave's rows (wave-uniform)
  // columns at or past the group's longest row: S(r, c) is the count before the group
  for (int c = 0; c < (int)maxcount; c++) {
    const int64_t s = c < g.ml ? rows_before(g, c, chunkpre, ngroups, j, s_w)
                               : chunkpre[(int64_t)c * (ngroups + 1) + j];
    for (int64_t b = b0; b < outlength && outer_offsets[b] == g.r; b++) binS[b * maxcount + c] = s;
  }
}

// the bins starting at row1 (empty trailing bins, and b = outlength): every
// row is before them
__global__ __launch_bounds__(kBlock) void binstart_tail_kernel(int64_t* binS, const int64_t* chunkpre,
                                                               const int64_t* outer_offsets, int64_t outlength,
                                                               int64_t row1, int64_t maxcount, int64_t ngroups) {
  const int64_t total = (outlength + 1) * maxcount;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < total; i += (int64_t)gridDim.x * kBlock) {
    const int64_t b = i / maxcount, c = i - b * maxcount;
    if (outer_offsets[b] == row1) binS[i] = chunkpre[c * (ngroups + 1) + ngroups];
  }
}

__global__ __launch_bounds__(kBlock) void cnt_from_bins_kernel(int64_t* cnt, const int64_t* binS, int64_t groups,
                                                               int64_t maxcount) {
  for (int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x; g < groups; g += (int64_t)gridDim.x * kBlock)
    cnt[g] = binS[g + maxcount] - binS[g];
}

// Every element (row r, column c < maxcount): its rank in group (bin, c) is
// S(r, c) - binS[bin, c]; PLACE writes nextcarry, otherwise `missing`.
template <bool PLACE>
__global__ __launch_bounds__(kBlock) void colrank_scatter_kernel(int64_t* out, const int64_t* chunkpre,
                                                                 const int64_t* binS, const int64_t* prefix,
                                                                 const int64_t* offsets,
                                                                 const int64_t* outer_offsets, int64_t outlength,
                                                                 int64_t row0, int64_t row1, int64_t maxcount,
                                                                 int64_t ngroups) {
  // `missing` is written in row-major order: each wave stages its rows'
  // values (when they fit and every element of the rows is written) and
  // stores them coalesced
  constexpr int kStage = PLACE ? 1 : 1024;
  __shared__ unsigned int s_w[kBlock / kWave][kFastMax];
  __shared__ int s_ml;
  __shared__ int64_t s_out[kBlock / kWave][kStage];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int64_t j = blockIdx.x;
  const RowGroup g = row_group(offsets, row0, row1, maxcount, j, s_w, &s_ml);
  if (g.ml == 0) return;  // workgroup-uniform
  int64_t bin = 0, e0 = 0, bstart = 0;
  if (g.len > 0) {
    bin = upper_bound(outer_offsets, 0, outlength + 1, g.r) - 1;
    e0 = offsets[g.r];
    bstart = outer_offsets[bin];
  }
  bool stage = false;
  int64_t we0 = 0, we1 = 0;
  if constexpr (!PLACE) {
    const bool valid = g.r < row1;
    int64_t rs = 0, re = 0;
    if (valid) {
      rs = offsets[g.r];
      re = offsets[g.r + 1];
    }
    const uint64_t vb = __ballot(valid);
    if (vb != 0) {
      const int last = 63 - __builtin_clzll(vb);
      we0 = __shfl(rs, 0, kWave);
      we1 = __shfl(re, last, kWave);
      stage = !__any(valid && (re - rs > maxcount || re < rs)) && we1 >= we0 && we1 - we0 <= kStage;
    }
  }
  for (int c = 0; c < g.ml; c++) {
    const int64_t s = rows_before(g, c, chunkpre, ngroups, j, s_w);
    if (c < g.len) {
      const int64_t rank = s - binS[bin * maxcount + c];
      if constexpr (PLACE) {
        out[prefix[bin * maxcount + c] + rank] = e0 + c;
      } else {
        const int64_t v = (g.r - bstart) - rank;
        if (stage) s_out[wave][e0 + c - we0] = v;
        else out[e0 + c] = v;
      }
    }
  }
  if constexpr (!PLACE) {
    __syncthreads();
    if (stage)
      for (int64_t k = lane; k < we1 - we0; k += kWave) out[we0 + k] = s_out[wave][k];
  }
}

// The sort-free regrouping's tables, from the work-queue pool after the
// ColMajor arrays: chunkpre (maxcount x (ngroups + 1)) and binS
// ((outlength + 1) x maxcount).
struct ColRanks {
  int64_t ngroups;
  int64_t* chunkpre;
  int64_t* binS;
};

Error colranks(ColRanks& cr, const ColMajor& cm, const int64_t* offsets, const int64_t* outer_offsets,
               int64_t outlength, int64_t maxcount, char* mem) {
  hipStream_t st = stream();
  cr.ngroups = ceil_div(cm.row1 - cm.row0 > 0 ? cm.row1 - cm.row0 : 0, (int64_t)kRowsPerGroup);
  cr.chunkpre = (int64_t*)mem;
  cr.binS = (int64_t*)(mem + round256((size_t)maxcount * (cr.ngroups + 1) * 8));
  if (cr.ngroups > 0) {
    colgroup_kernel<<<(unsigned)cr.ngroups, kBlock, 0, st>>>(cr.chunkpre, offsets, cm.row0, cm.row1, maxcount,
                                                             cr.ngroups);
  } else {
    fill_kernel<<<grid_for(maxcount, kBlock, 8), kBlock, 0, st>>>(cr.chunkpre, maxcount, 0);  // ngroups + 1 == 1
  }
  colscan_kernel<<<(unsigned)maxcount, kBlock, 0, st>>>(cr.chunkpre, cr.ngroups);
  if (cr.ngroups > 0)
    binstart_kernel<<<(unsigned)cr.ngroups, kBlock, 0, st>>>(cr.binS, cr.chunkpre, offsets, outer_offsets, outlength,
                                                             cm.row0, cm.row1, maxcount, cr.ngroups);
  binstart_tail_kernel<<<grid_for((outlength + 1) * maxcount, kBlock, 8), kBlock, 0, st>>>(
      cr.binS, cr.chunkpre, outer_offsets, outlength, cm.row1, maxcount, cr.ngroups);
  AKHIP_LAUNCH_CHECK("colranks");
  return success();
}

inline size_t colranks_bytes(int64_t rows, int64_t outlength, int64_t maxcount) {
  const int64_t ngroups = ceil_div(rows > 0 ? rows : 0, (int64_t)kRowsPerGroup);
  return round256((size_t)maxcount * (ngroups + 1) * 8) + round256((size_t)(outlength + 1) * maxcount * 8);
}

// Host values and the (bin, column) counts, prefix and maxgroup of `cm`
// without the sort (keys and order stay unset).
Error colmajor_fast(ColMajor& cm, ColRanks& cr, const int64_t* offsets, const int64_t* outer_offsets,
                    int64_t outlength, int64_t maxcount, bool with_prefix) {
  hipStream_t st = stream();
  {
    Error e = read_bounds(cm, offsets, outer_offsets, outlength);
    if (e.str) return e;
  }
  if (cm.n < 0) cm.n = 0;
  cm.groups = outlength * maxcount;
  const size_t gb = round256((size_t)cm.groups * 8), pb = round256((size_t)(cm.groups + 1) * 8);
  char* q = (char*)queue_pool(256 + gb + pb + colranks_bytes(cm.row1 - cm.row0, outlength, maxcount));
  if (!q) return failure("out of device memory for the column regrouping", kSliceNone, kSliceNone, kFile);
  cm.maxgroup = (unsigned long long*)q;
  cm.keys32 = nullptr;
  cm.keys64 = nullptr;
  cm.order = nullptr;
  cm.segoff = nullptr;
  cm.cnt = (int64_t*)(q + 256);
  cm.prefix = (int64_t*)(q + 256 + gb);
  Error e = colranks(cr, cm, offsets, outer_offsets, outlength, maxcount, q + 256 + gb + pb);
  if (e.str) return e;
  set_i64<<<1, 1, 0, st>>>((int64_t*)cm.maxgroup, 0);
  set_i64<<<1, 1, 0, st>>>(cm.prefix, 0);
  if (cm.groups > 0) {
    cnt_from_bins_kernel<<<grid_for(cm.groups, kBlock, 8), kBlock, 0, st>>>(cm.cnt, cr.binS, cm.groups, maxcount);
    if (with_prefix) {
      const int64_t ntiles = scan_tiles(cm.groups);
      void* s = scratch(scan_scratch_bytes(ntiles));
      if (!s) return failure("out of device memory for scan scratch", kSliceNone, kSliceNone, kFile);
      e = launch_lookback_scan(PrefixBody{cm.cnt, cm.prefix, cm.maxgroup}, cm.groups, 0, s);
      if (e.str) return e;
    }
  }
  AKHIP_LAUNCH_CHECK("colmajor_fast");
  return success();
}

}  // namespace
}  // namespace akhip

using namespace akhip;

extern "C" {

Error awkward_ListOffsetArray_reduce_nonlocal_maxcount_offsetscopy_64(int64_t* maxcount, int64_t* offsetscopy,
                                                                      const int64_t* offsets, int64_t length) {
  hipStream_t st = stream();
  set_i64<<<1, 1, 0, st>>>(maxcount, 0);
  if (length <= 0) {
    hipMemcpyAsync(offsetscopy, offsets, sizeof(int64_t), hipMemcpyDeviceToDevice, st);
    return success();
  }
  maxcount_kernel<<<grid_for(length, kBlock, 8), kBlock, 0, st>>>((unsigned long long*)maxcount, offsetscopy, offsets,
                                                                  length);
  AKHIP_LAUNCH_CHECK("maxcount_kernel");
  return success();
}

Error awkward_ListOffsetArray_reduce_nonlocal_outstartsstops_64(int64_t* outstarts, int64_t* outstops,
                                                                const int64_t* distincts, int64_t lendistincts,
                                                                int64_t outlength) {
  if (outlength <= 0) return success();
  hipStream_t st = stream();
  const unsigned grid = (unsigned)ceil_div(outlength, (int64_t)kBlock);
  if (lendistincts <= 0) {
    fill_kernel<<<grid, kBlock, 0, st>>>(outstarts, outlength, 0);
    fill_kernel<<<grid, kBlock, 0, st>>>(outstops, outlength, 0);
  } else if (lendistincts < outlength) {
    outstartsstops_serial<<<1, 1, 0, st>>>(outstarts, outstops, distincts, lendistincts, outlength);
  } else {
    outstartsstops_kernel<<<grid, kBlock, 0, st>>>(outstarts, outstops, distincts, lendistincts, outlength,
                                                   lendistincts / outlength);
  }
  AKHIP_LAUNCH_CHECK("outstartsstops_kernel");
  return success();
}

Error awkward_ListOffsetArray_reduce_nonlocal_preparenext_64(int64_t* nextcarry, int64_t* nextoffsets, int64_t,
                                                             int64_t* maxnextparents, int64_t* distincts,
                                                             int64_t distinctslen, int64_t*, const int64_t* offsets,
                                                             int64_t, const int64_t* outer_offsets, int64_t outlength,
                                                             int64_t maxcount) {
  hipStream_t st = stream();
  if (outlength <= 0 || maxcount <= 0) {
    set_i64<<<1, 1, 0, st>>>(maxnextparents, -1);
    const int64_t groups = outlength > 0 && maxcount > 0 ? outlength * maxcount : 0;
    fill_kernel<<<grid_for(groups + 1, kBlock, 8), kBlock, 0, st>>>(nextoffsets, groups + 1, 0);
    if (distinctslen > 0) fill_kernel<<<grid_for(distinctslen, kBlock, 8), kBlock, 0, st>>>(distincts, distinctslen, -1);
    AKHIP_LAUNCH_CHECK("preparenext (empty)");
    return success();
  }
  ColMajor cm;
  if (maxcount <= kFastMax) {
    ColRanks cr;
    Error e = colmajor_fast(cm, cr, offsets, outer_offsets, outlength, maxcount, true);
    if (e.str) return e;
    if (cr.ngroups > 0)
      colrank_scatter_kernel<true><<<(unsigned)cr.ngroups, kBlock, 0, st>>>(
          nextcarry, cr.chunkpre, cr.binS, cm.prefix, offsets, outer_offsets, outlength, cm.row0, cm.row1, maxcount,
          cr.ngroups);
  } else {
    Error e = colmajor(cm, offsets, outer_offsets, outlength, maxcount);
    if (e.str) return e;
    if (cm.n > 0) place_kernel<<<grid_for(cm.n, kBlock, 8), kBlock, 0, st>>>(cm, nextcarry, outlength, maxcount);
  }
  const int64_t hi = cm.groups + 1 > distinctslen ? cm.groups + 1 : distinctslen;
  groups_out_kernel<<<grid_for(hi, kBlock, 8), kBlock, 0, st>>>(cm, nextoffsets, distincts, distinctslen);
  maxgroup_out<<<1, 1, 0, st>>>(cm, maxnextparents);
  AKHIP_LAUNCH_CHECK("preparenext");
  return success();
}

Error awkward_ListOffsetArray_reduce_nonlocal_nextshifts_64(int64_t* nummissing, int64_t* missing,
                                                            int64_t* nextshifts, const int64_t* offsets, int64_t,
                                                            const int64_t*, const int64_t* outer_offsets,
                                                            int64_t outlength, int64_t maxcount, int64_t nextlen,
                                                            const int64_t* nextcarry) {
  if (outlength <= 0) return success();  // the CPU loop writes nothing
  hipStream_t st = stream();
  if (maxcount > 0) {
    ColMajor cm;
    if (maxcount <= kFastMax) {
      ColRanks cr;
      Error e = colmajor_fast(cm, cr, offsets, outer_offsets, outlength, maxcount, false);
      if (e.str) return e;
      if (cr.ngroups > 0)
        colrank_scatter_kernel<false><<<(unsigned)cr.ngroups, kBlock, 0, st>>>(
            missing, cr.chunkpre, cr.binS, nullptr, offsets, outer_offsets, outlength, cm.row0, cm.row1, maxcount,
            cr.ngroups);
    } else {
      Error e = colmajor(cm, offsets, outer_offsets, outlength, maxcount);
      if (e.str) return e;
      if (cm.n > 0)
        missing_kernel<<<grid_for(cm.n, kBlock, 8), kBlock, 0, st>>>(cm, missing, offsets, outer_offsets, outlength,
                                                                     maxcount);
    }
    // nummissing: the last bin's short rows (histogram by length, then a running sum)
    int64_t* hist = cm.cnt;  // groups >= maxcount entries, free again after the scan
    fill_kernel<<<grid_for(maxcount, kBlock, 8), kBlock, 0, st>>>(hist, maxcount, 0);
    const int64_t r0 = cm.last_bin_row, r1 = cm.row1;
    if (r1 > r0) short_rows_kernel<<<grid_for(r1 - r0, kBlock, 8), kBlock, 0, st>>>(hist, offsets, r0, r1, maxcount);
    nummissing_kernel<<<1, 1, 0, st>>>(nummissing, hist, maxcount);
  }
  if (nextlen > 0) gather_kernel<<<grid_for(nextlen, kBlock, 8), kBlock, 0, st>>>(nextshifts, missing, nextcarry, nextlen);
  AKHIP_LAUNCH_CHECK("nextshifts");
  return success();
}

}  // extern "C"
