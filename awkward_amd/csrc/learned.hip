// learned.hip -- the learned tile dispatcher (no counterpart in awkward-cpp).
//
// Which reduce tile (threads x lists per thread x LDS stage: the ten
// variants of launch_reduce and of the batch kernels) is fastest depends on
// the list-length distribution, the element size and the op.  The
// dispatcher predicts every variant's cost from 16 features of the array
// with a small FP32 MLP (16 -> 64 -> 64 -> 11: ten log2 costs relative to
// their mean, plus that mean) and takes the cheapest -- on the device, so a
// captured sequence can choose without a host round trip.
//
//   awkward_hip_dispatch_reset / _stats   one pass over the offsets of one or
//       many arrays; integer accumulators per slot (an array, or a group of
//       arrays that feed one batched launch), so features are bit-identical
//       from run to run whatever the workgroup order;
//   awkward_hip_dispatch_infer            features -> MLP -> choice, one
//       workgroup per slot; a feature outside the training range selects the
//       caller's fallback (the autotuner table's entry) instead;
//   awkward_hip_dispatch_choose           both in one launch: the workgroup
//       that finishes the statistics last runs the inference (what the
//       dispatcher and the batch runner call);
//   awkward_hip_mlp_forward              the same network over feature rows;
//   awkward_hip_mlp_train_step            forward + backward of a minibatch,
//       one gradient row per workgroup (64 rows each);
//   awkward_hip_mlp_adam                  sums the gradient rows in a fixed
//       order and applies Adam -- training is deterministic;
//   awkward_hip_dispatch_time_variants    the exhaustive sweep that yields
//       the training targets and the regret baseline.
//
// The network is ~6k parameters: every layer is a handful of FMAs per
// thread with weights in LDS; nothing here is large enough to be worth the
// matrix cores.

#include "common.h"

#include <cmath>

namespace akhip {
namespace learn {

constexpr int kFeatures = 16;
constexpr int kHidden = 64;
constexpr int kVariantsN = 10;
constexpr int kOut = kVariantsN + 1;
// parameter layout (floats): W1[16][64] b1[64] W2[64][64] b2[64] W3[64][11] b3[11]
// then the normalisation: mu[16] inv_sigma[16] lo[16] hi[16] (lo/hi normalised)
constexpr int kW1 = 0;
constexpr int kB1 = kW1 + kFeatures * kHidden;
constexpr int kW2 = kB1 + kHidden;
constexpr int kB2 = kW2 + kHidden * kHidden;
constexpr int kW3 = kB2 + kHidden;
constexpr int kB3 = kW3 + kHidden * kOut;
constexpr int kParams = kB3 + kOut;  // trained parameters
constexpr int kMu = kParams;
constexpr int kInvSigma = kMu + kFeatures;
constexpr int kLo = kInvSigma + kFeatures;
constexpr int kHi = kLo + kFeatures;
constexpr int kTotal = kHi + kFeatures;

// accumulator words per slot
constexpr int kAcc = 16;
enum { A_LISTS = 0, A_EMPTY, A_ELEMS, A_MAXLEN, A_HEAVY, A_SUMSQ, A_BIN0 };
constexpr int kBins = 7;
constexpr int kStatsBlock = 256;
constexpr int kStatsLPT = 16;
constexpr int64_t kStatsLists = (int64_t)kStatsBlock * kStatsLPT;
constexpr int64_t kHeavyLen = 32;

struct StatsItem {
  const int64_t* offsets;
  int64_t n;
  int64_t tile_begin;  // in tiles of kStatsLists lists
  int64_t slot;
};

__device__ inline int find_stats_item(const StatsItem* items, int nitems, int64_t tile) {
  int lo = 0, hi = nitems;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (items[mid].tile_begin <= tile) lo = mid;
    else hi = mid;
  }
  return lo;
}

__global__ void acc_reset_kernel(unsigned long long* acc, int64_t words) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (int64_t)gridDim.x * blockDim.x)
    acc[i] = 0ull;
}

// element-weighted length bins: [1,8] (8,32] (32,128] (128,512] (512,2048] (2048,8192] >8192
__device__ inline int length_bin(int64_t len) {
  if (len <= 8) return 0;
  const int clog = 64 - __clzll((unsigned long long)(len - 1));  // ceil(log2 len)
  const int b = (clog - 2) >> 1;
  return b < kBins - 1 ? b : kBins - 1;
}

__device__ inline unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_xor(v, d);
  return v;
}
__device__ inline unsigned long long wave_max_u64(unsigned long long v) {
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) {
    const unsigned long long o = __shfl_xor(v, d);
    v = o > v ? o : v;
  }
  return v;
}

// The statistics of one tile of kStatsLists lists of one item, added to its
// slot's accumulators (one workgroup of kStatsBlock threads).
__device__ inline void stats_tile(const StatsItem* __restrict__ items, int nitems, unsigned long long* __restrict__ acc,
                                  int64_t tile) {
  constexpr int NW = kStatsBlock / kWave;
  __shared__ int s_item;
  __shared__ unsigned long long s_red[NW][13];
  if (threadIdx.x == 0) s_item = find_stats_item(items, nitems, tile);
  __syncthreads();
  const StatsItem it = items[s_item];
  const int64_t l0 = (tile - it.tile_begin) * kStatsLists;
  unsigned long long v[13];
#pragma unroll
  for (int k = 0; k < 13; k++) v[k] = 0ull;
#pragma unroll 4
  for (int j = 0; j < kStatsLPT; j++) {
    const int64_t l = l0 + (int64_t)j * kStatsBlock + threadIdx.x;
    if (l < it.n) {
      int64_t len = it.offsets[l + 1] - it.offsets[l];
      if (len < 0) len = 0;
      const unsigned long long ul = (unsigned long long)len;
      v[A_LISTS] += 1;
      v[A_EMPTY] += len == 0;
      v[A_ELEMS] += ul;
      v[A_MAXLEN] = ul > v[A_MAXLEN] ? ul : v[A_MAXLEN];
      v[A_HEAVY] += len > kHeavyLen;
      const unsigned long long c = ul < (1ull << 21) ? ul : (1ull << 21);
      v[A_SUMSQ] += c * c;
      if (len > 0) {
        const int b = length_bin(len);
#pragma unroll
        for (int k = 0; k < kBins; k++) v[A_BIN0 + k] += (b == k) ? ul : 0ull;
      }
    }
  }
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < 13; k++) {
    const unsigned long long r = (k == A_MAXLEN) ? wave_max_u64(v[k]) : wave_sum_u64(v[k]);
    if (lane == 0) s_red[wave][k] = r;
  }
  __syncthreads();
  if (threadIdx.x < 13) {
    const int k = threadIdx.x;
    unsigned long long r = s_red[0][k];
    for (int w = 1; w < NW; w++) {
      const unsigned long long o = s_red[w][k];
      r = (k == A_MAXLEN) ? (o > r ? o : r) : r + o;
    }
    unsigned long long* dst = acc + it.slot * kAcc + k;
    if (r != 0ull) {
      if (k == A_MAXLEN) atomicMax(dst, r);
      else atomicAdd(dst, r);
    }
  }
}

// One workgroup per tile.
__global__ __launch_bounds__(kStatsBlock) void stats_kernel(const StatsItem* __restrict__ items, int nitems,
                                                            unsigned long long* __restrict__ acc) {
  stats_tile(items, nitems, acc, (int64_t)blockIdx.x);
}

// The 16 features of a slot from its accumulators (double arithmetic, then
// rounded to float).
__device__ inline void slot_features(const unsigned long long* a, int itemsize, int op, float* x) {
  const double n = (double)a[A_LISTS];
  const double N = (double)a[A_ELEMS];
  const double nn = n > 0 ? n : 1.0, NN = N > 0 ? N : 1.0;
  const double mean = N / nn;
  double var = (double)a[A_SUMSQ] / nn - mean * mean;
  if (var < 0) var = 0;
  x[0] = (float)log2(1.0 + n);
  x[1] = (float)log2(1.0 + N);
  x[2] = (float)log2(1.0 + mean);
  x[3] = (float)log2(1.0 + sqrt(var));
  x[4] = (float)log2(1.0 + (double)a[A_MAXLEN]);
  x[5] = (float)((double)a[A_EMPTY] / nn);
  for (int k = 0; k < kBins; k++) x[6 + k] = (float)((double)a[A_BIN0 + k] / NN);
  x[13] = (float)(itemsize / 8.0);
  x[14] = (float)(op / 3.0);
  x[15] = (float)((double)a[A_HEAVY] / nn);
}

// The network for one row: xraw (LDS, 16 raw features) -> y (LDS, kOut
// outputs), computed by the first kHidden threads of the workgroup (every
// thread reaches the barriers).  Returns (to every thread) whether every
// normalised feature lies inside the training range.
__device__ inline bool mlp_eval(const float* __restrict__ P, const float* xraw, float* xn, float* h1, float* h2,
                                float* y, int* s_in) {
  const int j = threadIdx.x;
  const bool unit = j < kHidden;
  if (j == 0) *s_in = 1;
  __syncthreads();
  if (j < kFeatures) {
    const float z = (xraw[j] - P[kMu + j]) * P[kInvSigma + j];
    xn[j] = z;
    if (!(z >= P[kLo + j] && z <= P[kHi + j])) *s_in = 0;
  }
  __syncthreads();
  if (unit) {
    float a = P[kB1 + j];
    for (int k = 0; k < kFeatures; k++) a = fmaf(P[kW1 + k * kHidden + j], xn[k], a);
    h1[j] = fmaxf(a, 0.0f);
  }
  __syncthreads();
  if (unit) {
    float a = P[kB2 + j];
    for (int k = 0; k < kHidden; k++) a = fmaf(P[kW2 + k * kHidden + j], h1[k], a);
    h2[j] = fmaxf(a, 0.0f);
  }
  __syncthreads();
  if (j < kOut) {
    float a = P[kB3 + j];
    for (int k = 0; k < kHidden; k++) a = fmaf(P[kW3 + k * kOut + j], h2[k], a);
    y[j] = a;
  }
  __syncthreads();
  return *s_in != 0;
}

struct InferSmem {
  float x[kFeatures], xn[kFeatures], h1[kHidden], h2[kHidden], y[kOut];
  int in;
};

// Features -> network -> choice of one slot from its accumulators `a`;
// meta m: {itemsize, op, fallback variant, unused}.  Every thread of the
// workgroup calls it (at least kHidden of them).
__device__ inline void infer_slot(const unsigned long long* a, const int32_t* m, const float* __restrict__ P,
                                  int slot, float* __restrict__ features, float* __restrict__ costs,
                                  int32_t* __restrict__ choice, InferSmem& s) {
  if (threadIdx.x == 0) slot_features(a, m[0], m[1], s.x);
  __syncthreads();
  const bool inside = mlp_eval(P, s.x, s.xn, s.h1, s.h2, s.y, &s.in);
  if (threadIdx.x < kFeatures && features) features[slot * kFeatures + threadIdx.x] = s.x[threadIdx.x];
  if (threadIdx.x < kOut && costs) costs[slot * kOut + threadIdx.x] = s.y[threadIdx.x];
  if (threadIdx.x == 0) {
    int best = 0;
    for (int v = 1; v < kVariantsN; v++)
      if (s.y[v] < s.y[best]) best = v;
    const bool empty = a[A_LISTS] == 0ull;
    const bool use_fallback = !inside || empty;
    choice[2 * slot] = use_fallback ? m[2] : best;
    choice[2 * slot + 1] = use_fallback ? 1 : 0;
  }
}

// One workgroup of 64 per slot, after stats_kernel.
__global__ __launch_bounds__(64) void infer_kernel(const unsigned long long* __restrict__ acc,
                                                   const int32_t* __restrict__ meta, const float* __restrict__ P,
                                                   float* __restrict__ features, float* __restrict__ costs,
                                                   int32_t* __restrict__ choice) {
  __shared__ InferSmem s;
  const int slot = blockIdx.x;
  infer_slot(acc + (int64_t)slot * kAcc, meta + 4 * slot, P, slot, features, costs, choice, s);
}

// Statistics and inference in one launch: max(total_tiles, 1) workgroups
// each add one tile's statistics, then count themselves done in the spare
// accumulator word A_DONE of slot 0; the workgroup that counts last (every
// other tile's atomics are then visible to it) copies each slot's
// accumulators, clears them -- the accumulators start and end every launch
// at zero, so no clearing launch or memset is needed, under a captured graph
// too -- and runs the slot's inference.
constexpr int A_DONE = kAcc - 1;
static_assert(A_DONE >= A_BIN0 + kBins, "the completion count needs a spare accumulator word");
static_assert(kStatsBlock >= kHidden, "the last workgroup runs the network");

__global__ __launch_bounds__(kStatsBlock) void choose_kernel(const StatsItem* __restrict__ items, int nitems,
                                                             int64_t total_tiles, unsigned long long* acc,
                                                             const int32_t* __restrict__ meta, int nslots,
                                                             const float* __restrict__ P, float* __restrict__ features,
                                                             float* __restrict__ costs, int32_t* __restrict__ choice) {
  __shared__ int s_last;
  __shared__ unsigned long long s_a[kAcc];
  __shared__ InferSmem s;
  if (nitems > 0 && (int64_t)blockIdx.x < total_tiles) stats_tile(items, nitems, acc, (int64_t)blockIdx.x);
  __threadfence();  // this workgroup's accumulator atomics before its count
  __syncthreads();
  if (threadIdx.x == 0)
    s_last = __hip_atomic_fetch_add(acc + A_DONE, 1ull, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) ==
             (unsigned long long)gridDim.x - 1ull;
  __syncthreads();
  if (!s_last) return;
  __threadfence();  // every other workgroup's atomics before the reads below
  for (int slot = 0; slot < nslots; slot++) {
    unsigned long long* a = acc + (int64_t)slot * kAcc;
    if (threadIdx.x < kAcc) s_a[threadIdx.x] = __hip_atomic_load(a + threadIdx.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    if (threadIdx.x < kAcc) __hip_atomic_store(a + threadIdx.x, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    infer_slot(s_a, meta + 4 * slot, P, slot, features, costs, choice, s);
    __syncthreads();  // s_a and s are rewritten for the next slot
  }
}

// rows of raw features -> outputs
__global__ __launch_bounds__(64) void forward_kernel(const float* __restrict__ X, const float* __restrict__ P,
                                                     float* __restrict__ Y, int32_t* __restrict__ inside) {
  __shared__ float x[kFeatures], xn[kFeatures], h1[kHidden], h2[kHidden], y[kOut];
  __shared__ int s_in;
  const int64_t row = blockIdx.x;
  if (threadIdx.x < kFeatures) x[threadIdx.x] = X[row * kFeatures + threadIdx.x];
  __syncthreads();
  const bool in = mlp_eval(P, x, xn, h1, h2, y, &s_in);
  if (threadIdx.x < kOut) Y[row * kOut + threadIdx.x] = y[threadIdx.x];
  if (threadIdx.x == 0 && inside) inside[row] = in ? 1 : 0;
}

// ---- training ---------------------------------------------------------------
constexpr int kRows = 64;      // rows per workgroup
constexpr int kTrainBlock = 256;
constexpr int kDyPitch = 12;

struct TrainSmem {
  float w[kParams];
  float x[kRows * kFeatures];
  float t[kRows * kOut];
  float h1[kRows * kHidden];  // H1, then dL/dH1
  float h2[kRows * kHidden];  // H2, then dL/dH2
  float dy[kRows * kDyPitch];
  float valid[kRows];
  float red[kTrainBlock / kWave];
};

// X: normalised features (R x 16), T: targets (R x kOut), ow: per-output loss
// weights.  Minibatch rows row0 .. row0+bsz-1 of the permutation
// r -> (pa * r + pb) mod R.  Loss = sum over rows and outputs of
// ow[o] (y - t)^2 / bsz; this workgroup's gradient goes to grads[blockIdx.x].
__global__ __launch_bounds__(kTrainBlock) void train_kernel(const float* __restrict__ X, const float* __restrict__ T,
                                                            const float* __restrict__ ow, int64_t R, int64_t pa,
                                                            int64_t pb, int64_t row0, int64_t bsz,
                                                            const float* __restrict__ P, float* __restrict__ grads,
                                                            float* __restrict__ losses) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  TrainSmem& s = *reinterpret_cast<TrainSmem*>(smem_raw);
  const int tid = threadIdx.x;
  float* g = grads + (int64_t)blockIdx.x * kParams;
  int64_t end = row0 + bsz;
  if (end > R) end = R;
  const float inv_b = 1.0f / (float)(end - row0);

  for (int i = tid; i < kParams; i += kTrainBlock) s.w[i] = P[i];
  for (int r = tid; r < kRows; r += kTrainBlock) {
    const int64_t gi = row0 + (int64_t)blockIdx.x * kRows + r;
    s.valid[r] = gi < end ? 1.0f : 0.0f;
  }
  for (int e = tid; e < kRows * kFeatures; e += kTrainBlock) {
    const int r = e / kFeatures, k = e % kFeatures;
    const int64_t gi = row0 + (int64_t)blockIdx.x * kRows + r;
    float v = 0.0f;
    if (gi < end) v = X[((pa * gi + pb) % R) * kFeatures + k];
    s.x[e] = v;
  }
  for (int e = tid; e < kRows * kOut; e += kTrainBlock) {
    const int r = e / kOut, o = e % kOut;
    const int64_t gi = row0 + (int64_t)blockIdx.x * kRows + r;
    float v = 0.0f;
    if (gi < end) v = T[((pa * gi + pb) % R) * kOut + o];
    s.t[e] = v;
  }
  __syncthreads();

  const int j = tid & (kHidden - 1);
  const int grp = tid / kHidden;  // 0..3
  constexpr int RG = kRows / (kTrainBlock / kHidden);  // 16 rows per group
  // H1 = relu(X W1 + b1)
  {
    float a[RG];
#pragma unroll
    for (int i = 0; i < RG; i++) a[i] = s.w[kB1 + j];
    for (int k = 0; k < kFeatures; k++) {
      const float w = s.w[kW1 + k * kHidden + j];
#pragma unroll
      for (int i = 0; i < RG; i++) a[i] = fmaf(w, s.x[(grp * RG + i) * kFeatures + k], a[i]);
    }
#pragma unroll
    for (int i = 0; i < RG; i++) s.h1[(grp * RG + i) * kHidden + j] = fmaxf(a[i], 0.0f);
  }
  __syncthreads();
  // H2 = relu(H1 W2 + b2)
  {
    float a[RG];
#pragma unroll
    for (int i = 0; i < RG; i++) a[i] = s.w[kB2 + j];
    for (int k = 0; k < kHidden; k++) {
      const float w = s.w[kW2 + k * kHidden + j];
#pragma unroll
      for (int i = 0; i < RG; i++) a[i] = fmaf(w, s.h1[(grp * RG + i) * kHidden + k], a[i]);
    }
#pragma unroll
    for (int i = 0; i < RG; i++) s.h2[(grp * RG + i) * kHidden + j] = fmaxf(a[i], 0.0f);
  }
  __syncthreads();
  // Y = H2 W3 + b3, dY, loss
  float lsum = 0.0f;
  for (int e = tid; e < kRows * kOut; e += kTrainBlock) {
    const int r = e / kOut, o = e % kOut;
    float a = s.w[kB3 + o];
    for (int k = 0; k < kHidden; k++) a = fmaf(s.w[kW3 + k * kOut + o], s.h2[r * kHidden + k], a);
    const float d = s.valid[r] * (a - s.t[e]);
    lsum += ow[o] * d * d;
    s.dy[r * kDyPitch + o] = 2.0f * ow[o] * d * inv_b;
  }
  {
    float w = lsum;
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) w += __shfl_xor(w, d);
    if ((tid & (kWave - 1)) == 0) s.red[tid / kWave] = w;
  }
  __syncthreads();
  if (tid == 0) {
    float w = 0.0f;
    for (int k = 0; k < kTrainBlock / kWave; k++) w += s.red[k];
    losses[blockIdx.x] = w * inv_b;
  }
  // dW3, db3
  for (int e = tid; e < kHidden * kOut; e += kTrainBlock) {
    const int k = e / kOut, o = e % kOut;
    float a = 0.0f;
    for (int r = 0; r < kRows; r++) a = fmaf(s.h2[r * kHidden + k], s.dy[r * kDyPitch + o], a);
    g[kW3 + e] = a;
  }
  if (tid < kOut) {
    float a = 0.0f;
    for (int r = 0; r < kRows; r++) a += s.dy[r * kDyPitch + tid];
    g[kB3 + tid] = a;
  }
  __syncthreads();
  // dH2 = (dY W3^T) * [H2 > 0], in place
  for (int e = tid; e < kRows * kHidden; e += kTrainBlock) {
    const int r = e / kHidden, k = e % kHidden;
    float a = 0.0f;
    if (s.h2[e] > 0.0f)
      for (int o = 0; o < kOut; o++) a = fmaf(s.dy[r * kDyPitch + o], s.w[kW3 + k * kOut + o], a);
    s.h2[e] = a;
  }
  __syncthreads();
  // dW2[k][j] = sum_r H1[r][k] dH2[r][j]; db2
  {
    float a[RG];
#pragma unroll
    for (int i = 0; i < RG; i++) a[i] = 0.0f;
    float b = 0.0f;
    for (int r = 0; r < kRows; r++) {
      const float d = s.h2[r * kHidden + j];
      b += d;
#pragma unroll
      for (int i = 0; i < RG; i++) a[i] = fmaf(s.h1[r * kHidden + grp * RG + i], d, a[i]);
    }
#pragma unroll
    for (int i = 0; i < RG; i++) g[kW2 + (grp * RG + i) * kHidden + j] = a[i];
    if (grp == 0) g[kB2 + j] = b;
  }
  __syncthreads();
  // dH1 = (dH2 W2^T) * [H1 > 0], in place; the rotation spreads the W2 reads over banks
  for (int e = tid; e < kRows * kHidden; e += kTrainBlock) {
    const int r = e / kHidden, k = e % kHidden;
    float a = 0.0f;
    if (s.h1[e] > 0.0f)
      for (int q = 0; q < kHidden; q++) {
        const int jj = (q + k) & (kHidden - 1);
        a = fmaf(s.h2[r * kHidden + jj], s.w[kW2 + k * kHidden + jj], a);
      }
    s.h1[e] = a;
  }
  __syncthreads();
  // dW1[k][j] = sum_r X[r][k] dH1[r][j]; db1
  {
    constexpr int KG = kFeatures / (kTrainBlock / kHidden);  // 4 inputs per group
    float a[KG];
#pragma unroll
    for (int i = 0; i < KG; i++) a[i] = 0.0f;
    float b = 0.0f;
    for (int r = 0; r < kRows; r++) {
      const float d = s.h1[r * kHidden + j];
      b += d;
#pragma unroll
      for (int i = 0; i < KG; i++) a[i] = fmaf(s.x[r * kFeatures + grp * KG + i], d, a[i]);
    }
#pragma unroll
    for (int i = 0; i < KG; i++) g[kW1 + (grp * KG + i) * kHidden + j] = a[i];
    if (grp == 0) g[kB1 + j] = b;
  }
}

// Adam over the summed gradient rows (fixed order); loss_acc[0] += sum of the
// workgroup losses (thread 0).
__global__ void adam_kernel(float* __restrict__ P, float* __restrict__ m, float* __restrict__ v,
                            const float* __restrict__ grads, const float* __restrict__ losses, int nparts, float lr,
                            float beta1, float beta2, float eps, float bc1, float bc2, float* __restrict__ loss_acc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    float l = 0.0f;
    for (int b = 0; b < nparts; b++) l += losses[b];
    loss_acc[0] += l;
  }
  if (i >= kParams) return;
  float gs = 0.0f;
  for (int b = 0; b < nparts; b++) gs += grads[(int64_t)b * kParams + i];
  const float mi = beta1 * m[i] + (1.0f - beta1) * gs;
  const float vi = beta2 * v[i] + (1.0f - beta2) * gs * gs;
  m[i] = mi;
  v[i] = vi;
  const float denom = sqrtf(vi) / sqrtf(bc2) + eps;
  P[i] -= (lr / bc1) * (mi / denom);
}

}  // namespace learn
}  // namespace akhip

using namespace akhip;
using namespace akhip::learn;

extern "C" {

// entry points of the ops the sweep times (reduce_*.hip, batch.hip)
Error awkward_hip_reduce_sum_int64_int32(int64_t*, const int32_t*, const int64_t*, int64_t, int);
Error awkward_hip_reduce_sum_int64_int64(int64_t*, const int64_t*, const int64_t*, int64_t, int);
Error awkward_hip_reduce_sum_float32_float32(float*, const float*, const int64_t*, int64_t, int);
Error awkward_hip_reduce_sum_float64_float64(double*, const double*, const int64_t*, int64_t, int);
Error awkward_hip_reduce_min_masked_int32(int32_t*, int8_t*, const int32_t*, const int64_t*, int64_t, int32_t, int);
Error awkward_hip_reduce_min_masked_int64(int64_t*, int8_t*, const int64_t*, const int64_t*, int64_t, int64_t, int);
Error awkward_hip_reduce_min_masked_float32(float*, int8_t*, const float*, const int64_t*, int64_t, float, int);
Error awkward_hip_reduce_min_masked_float64(double*, int8_t*, const double*, const int64_t*, int64_t, double, int);
Error awkward_hip_reduce_argmin_local_int32(int64_t*, int8_t*, const int32_t*, const int64_t*, int64_t, int);
Error awkward_hip_reduce_argmin_local_int64(int64_t*, int8_t*, const int64_t*, const int64_t*, int64_t, int);
Error awkward_hip_reduce_argmin_local_float32(int64_t*, int8_t*, const float*, const int64_t*, int64_t, int);
Error awkward_hip_reduce_argmin_local_float64(int64_t*, int8_t*, const double*, const int64_t*, int64_t, int);
Error awkward_hip_batch_reset(unsigned long long* counter);
int64_t awkward_hip_batch_reduce_tile_lists(int variant);
Error awkward_hip_batch_reduce_int32(const void*, int32_t, int64_t, unsigned long long*, int32_t, int32_t, int64_t);
Error awkward_hip_batch_reduce_int64(const void*, int32_t, int64_t, unsigned long long*, int64_t, int32_t, int64_t);
Error awkward_hip_batch_reduce_float32(const void*, int32_t, int64_t, unsigned long long*, float, int32_t, int64_t);
Error awkward_hip_batch_reduce_float64(const void*, int32_t, int64_t, unsigned long long*, double, int32_t, int64_t);

int32_t awkward_hip_dispatch_layout(int32_t what) {
  switch (what) {
    case 0: return kFeatures;
    case 1: return kHidden;
    case 2: return kOut;
    case 3: return kParams;
    case 4: return kTotal;
    case 5: return kAcc;
    case 6: return (int32_t)kStatsLists;
    case 7: return kRows;
    default: return -1;
  }
}

// Zero `nslots` accumulator slots (a kernel, so it can be captured).
Error awkward_hip_dispatch_reset(unsigned long long* acc, int64_t nslots) {
  if (nslots <= 0) return success();
  const int64_t words = nslots * kAcc;
  acc_reset_kernel<<<(unsigned)ceil_div(words, 256) < 1024u ? (unsigned)ceil_div(words, 256) : 1024u, 256, 0,
                     stream()>>>(acc, words);
  AKHIP_LAUNCH_CHECK("acc_reset_kernel");
  return success();
}

// items: device table of StatsItem {offsets, n, tile_begin, slot} (32 bytes
// each, tile_begin in tiles of awkward_hip_dispatch_layout(6) lists, ascending).
Error awkward_hip_dispatch_stats(const void* items, int32_t nitems, int64_t total_tiles, unsigned long long* acc) {
  if (nitems <= 0 || total_tiles <= 0) return success();
  if (total_tiles > 0x7fffffffLL)
    return failure("too many lists for one statistics launch", kSliceNone, kSliceNone, "awkward_amd/csrc/learned.hip");
  stats_kernel<<<(unsigned)total_tiles, kStatsBlock, 0, stream()>>>((const StatsItem*)items, nitems, acc);
  AKHIP_LAUNCH_CHECK("stats_kernel");
  return success();
}

// meta: device int32[4 * nslots] {itemsize, op (0 sum, 1 min/max, 2 argmin/argmax,
// 3 batched sum+min+argmin), fallback variant, 0}; params: device float[kTotal].
// Writes features (nslots x 16, may be null), costs (nslots x 11, may be
// null) and choice (nslots x 2: variant, 1 if the fallback was taken).
Error awkward_hip_dispatch_infer(const unsigned long long* acc, const int32_t* meta, int32_t nslots,
                                 const float* params, float* features, float* costs, int32_t* choice) {
  if (nslots <= 0) return success();
  infer_kernel<<<(unsigned)nslots, 64, 0, stream()>>>(acc, meta, params, features, costs, choice);
  AKHIP_LAUNCH_CHECK("infer_kernel");
  return success();
}

// Statistics and inference in one launch (the outputs of dispatch_stats
// followed by dispatch_infer).  acc (nslots accumulator slots) must hold
// zeros on entry and holds zeros again when the launch ends; launches that
// share one acc must not overlap.
Error awkward_hip_dispatch_choose(const void* items, int32_t nitems, int64_t total_tiles, unsigned long long* acc,
                                  const int32_t* meta, int32_t nslots, const float* params, float* features,
                                  float* costs, int32_t* choice) {
  if (nslots <= 0) return success();
  if (nitems < 0 || total_tiles < 0 || total_tiles > 0x7fffffffLL)
    return failure("invalid statistics table for one launch", kSliceNone, kSliceNone, "awkward_amd/csrc/learned.hip");
  const unsigned grid = total_tiles > 0 ? (unsigned)total_tiles : 1u;
  choose_kernel<<<grid, kStatsBlock, 0, stream()>>>((const StatsItem*)items, nitems, total_tiles, acc, meta, nslots,
                                                     params, features, costs, choice);
  AKHIP_LAUNCH_CHECK("choose_kernel");
  return success();
}

// X: rows x 16 raw features -> Y: rows x 11; inside (may be null): 1 where
// every feature is inside the training range.
Error awkward_hip_mlp_forward(const float* X, int64_t rows, const float* params, float* Y, int32_t* inside) {
  if (rows <= 0) return success();
  if (rows > 0x7fffffffLL)
    return failure("too many rows", kSliceNone, kSliceNone, "awkward_amd/csrc/learned.hip");
  forward_kernel<<<(unsigned)rows, 64, 0, stream()>>>(X, params, Y, inside);
  AKHIP_LAUNCH_CHECK("forward_kernel");
  return success();
}

// One minibatch: ceil(bsz / 64) workgroups; grads must hold that many rows of
// kParams floats, losses that many floats.
Error awkward_hip_mlp_train_step(const float* X, const float* T, const float* ow, int64_t R, int64_t pa, int64_t pb,
                                 int64_t row0, int64_t bsz, const float* params, float* grads, float* losses) {
  if (R <= 0 || bsz <= 0 || row0 >= R) return success();
  if (R > 0x7fffffffLL || pa < 0 || pa >= R || pb < 0 || pb >= R)
    return failure("invalid training permutation", kSliceNone, kSliceNone, "awkward_amd/csrc/learned.hip");
  int64_t rows = bsz;
  if (row0 + rows > R) rows = R - row0;
  const unsigned grid = (unsigned)ceil_div(rows, kRows);
  const size_t lds = sizeof(TrainSmem);  // ~65 KiB of the 160 KiB per CU
  train_kernel<<<grid, kTrainBlock, lds, stream()>>>(X, T, ow, R, pa, pb, row0, bsz, params, grads, losses);
  AKHIP_LAUNCH_CHECK("train_kernel");
  return success();
}

Error awkward_hip_mlp_adam(float* params, float* m, float* v, const float* grads, const float* losses, int32_t nparts,
                           float lr, float beta1, float beta2, float eps, int64_t step, float* loss_acc) {
  const float bc1 = 1.0f - powf(beta1, (float)step);
  const float bc2 = 1.0f - powf(beta2, (float)step);
  adam_kernel<<<(unsigned)ceil_div(kParams, 256), 256, 0, stream()>>>(params, m, v, grads, losses, nparts, lr, beta1,
                                                                      beta2, eps, bc1, bc2, loss_acc);
  AKHIP_LAUNCH_CHECK("adam_kernel");
  return success();
}

// The exhaustive sweep: time every variant of `op` (0 sum, 1 min with mask,
// 2 local argmin with mask, 3 the batch kernel's sum+min+argmin over this one
// array) on one array of `length` elements, dtype (0 int32, 1 int64, 2
// float32, 3 float64):
// one untimed launch, then `reps` launches between two events on the current
// stream.  ms_out (host) receives kVariantsN milliseconds per launch.  Uses
// the scratch pool for the outputs; synchronises.
Error awkward_hip_dispatch_time_variants(int32_t op, int32_t dtype, const void* data, const int64_t* offsets,
                                         int64_t n, int64_t length, int32_t reps, float* ms_out) {
  const char* file = "awkward_amd/csrc/learned.hip";
  if (op < 0 || op > 3 || dtype < 0 || dtype > 3 || reps < 1 || n < 0)
    return failure("invalid timing request", kSliceNone, kSliceNone, file);
  const int64_t nn = n > 0 ? n : 1;
  const size_t out8 = ((size_t)nn * 8 + 255) / 256 * 256;
  const size_t mask = ((size_t)nn + 255) / 256 * 256;
  char* s = (char*)scratch(512 + 3 * out8 + mask);
  if (!s) return failure("out of device memory for the timing sweep", kSliceNone, kSliceNone, file);
  void* o1 = s + 512;
  void* o2 = s + 512 + out8;
  int64_t* o3 = (int64_t*)(s + 512 + 2 * out8);
  int8_t* mk = (int8_t*)(s + 512 + 3 * out8);
  unsigned long long* counter = (unsigned long long*)(s + 256);
  hipStream_t st = stream();

  auto launch = [&](int v) -> Error {
    switch (op) {
      case 0:
        switch (dtype) {
          case 0: return awkward_hip_reduce_sum_int64_int32((int64_t*)o1, (const int32_t*)data, offsets, n, v);
          case 1: return awkward_hip_reduce_sum_int64_int64((int64_t*)o1, (const int64_t*)data, offsets, n, v);
          case 2: return awkward_hip_reduce_sum_float32_float32((float*)o1, (const float*)data, offsets, n, v);
          default: return awkward_hip_reduce_sum_float64_float64((double*)o1, (const double*)data, offsets, n, v);
        }
      case 1:
        switch (dtype) {
          case 0: return awkward_hip_reduce_min_masked_int32((int32_t*)o1, mk, (const int32_t*)data, offsets, n,
                                                             INT32_MAX, v);
          case 1: return awkward_hip_reduce_min_masked_int64((int64_t*)o1, mk, (const int64_t*)data, offsets, n,
                                                             INT64_MAX, v);
          case 2: return awkward_hip_reduce_min_masked_float32((float*)o1, mk, (const float*)data, offsets, n,
                                                               INFINITY, v);
          default: return awkward_hip_reduce_min_masked_float64((double*)o1, mk, (const double*)data, offsets, n,
                                                                (double)INFINITY, v);
        }
      case 2:
        switch (dtype) {
          case 0: return awkward_hip_reduce_argmin_local_int32(o3, mk, (const int32_t*)data, offsets, n, v);
          case 1: return awkward_hip_reduce_argmin_local_int64(o3, mk, (const int64_t*)data, offsets, n, v);
          case 2: return awkward_hip_reduce_argmin_local_float32(o3, mk, (const float*)data, offsets, n, v);
          default: return awkward_hip_reduce_argmin_local_float64(o3, mk, (const double*)data, offsets, n, v);
        }
      default: {
        const int64_t L = awkward_hip_batch_reduce_tile_lists(v);
        const int64_t tiles = ceil_div(n, L);
        Error e = awkward_hip_batch_reset(counter);
        if (e.str) return e;
        switch (dtype) {
          case 0: return awkward_hip_batch_reduce_int32(s, 1, tiles, counter, INT32_MAX, v, length);
          case 1: return awkward_hip_batch_reduce_int64(s, 1, tiles, counter, INT64_MAX, v, length);
          case 2: return awkward_hip_batch_reduce_float32(s, 1, tiles, counter, INFINITY, v, length);
          default: return awkward_hip_batch_reduce_float64(s, 1, tiles, counter, (double)INFINITY, v, length);
        }
      }
    }
  };

  if (op == 3) {
    // one BatchReduceItem {data, offsets, sum, min, argmin, mask, n, tile_begin}
    const uint64_t item[8] = {(uint64_t)(uintptr_t)data, (uint64_t)(uintptr_t)offsets, (uint64_t)(uintptr_t)o1,
                              (uint64_t)(uintptr_t)o2, (uint64_t)(uintptr_t)o3, (uint64_t)(uintptr_t)mk,
                              (uint64_t)n, 0ull};
    hipError_t e = hipMemcpyAsync(s, item, sizeof(item), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_failure(e, "hipMemcpyAsync(timing item)");
  }
  hipEvent_t e0, e1;
  hipError_t he = hipEventCreate(&e0);
  if (he != hipSuccess) return hip_failure(he, "hipEventCreate");
  he = hipEventCreate(&e1);
  if (he != hipSuccess) {
    (void)hipEventDestroy(e0);
    return hip_failure(he, "hipEventCreate");
  }
  Error result = success();
  for (int v = 0; v < kVariantsN && !result.str; v++) {
    result = launch(v);
    if (result.str) break;
    he = hipEventRecord(e0, st);
    for (int r = 0; r < reps && !result.str && he == hipSuccess; r++) result = launch(v);
    if (he == hipSuccess) he = hipEventRecord(e1, st);
    if (he == hipSuccess) he = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (he == hipSuccess) he = hipEventElapsedTime(&ms, e0, e1);
    if (he != hipSuccess) {
      result = hip_failure(he, "timing events");
      break;
    }
    ms_out[v] = ms / (float)reps;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return result;
}

}  // extern "C"
