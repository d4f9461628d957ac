// reduce.h -- segmented reductions over int64 offsets (gfx950).
//
// One kernel template serves awkward_reduce_{sum,max,min,argmax,argmin} and
// the bool sums (awkward-cpp/src/cpu-kernels/awkward_reduce_*.cpp).  A
// workgroup of BLOCK threads owns BLOCK * LPT consecutive lists (thread t the
// lists t + j * BLOCK, so loads of offsets and stores of results coalesce):
//
//   * light tile (the lists' elements fit in CAP_BYTES of LDS): the span is
//     staged with coalesced loads, then each thread reduces its own list
//     sequentially from LDS with the CPU kernel's own loop -- the result is
//     bit-identical to awkward-cpp, float sums included;
//   * heavy tile: lists of at most kShortLen elements are still reduced
//     sequentially by their thread (from global memory); medium lists
//     (<= kWaveLen) by one wave each, long lists by the whole workgroup.
//     Cooperative partials carry (value, index) for extrema so that ties
//     resolve to the first occurrence exactly as the serial loop does; only
//     float sums of lists longer than kShortLen in heavy tiles change their
//     summation order (within the stated tolerance).
//   * a caller may pass a ReduceQueue: lists longer than kWaveLen then go to
//     it in kReduceChunk-element chunks instead of to the workgroup, and
//     reduce_chunks_kernel (every workgroup of the chip, one chunk each) +
//     reduce_finalize_kernel (chunk partials combined in order) finish them
//     -- so a few very long lists no longer serialise on a few workgroups.
//
// Ops (struct with the members used below):
//   In, Out, Part
//   Out  sequential(const In* p, int64_t b, int64_t e, int64_t gb) -- the CPU loop
//   Part init(), elem(In v, int64_t gi), combine(Part a, Part b)
//   Out  finalize(Part p, const In* global, int64_t gb, int64_t ge)
//   Part shfl_xor(Part p, int mask)
#pragma once

#include <cmath>
#include <type_traits>

#include "common.h"
#include "errors.h"

namespace akhip {

constexpr int64_t kShortLen = 32;
constexpr int64_t kWaveLen = 4096;

// ---- shuffles of arbitrary small trivially-copyable values ----------------
template <typename T>
__device__ inline T shfl_xor_any(T v, int mask) {
  if constexpr (sizeof(T) < 4) {
    int x = (int)v;
    x = __shfl_xor(x, mask, kWave);
    return (T)x;
  } else if constexpr (sizeof(T) == 4) {
    int x;
    __builtin_memcpy(&x, &v, 4);
    x = __shfl_xor(x, mask, kWave);
    T out;
    __builtin_memcpy(&out, &x, 4);
    return out;
  } else {
    static_assert(sizeof(T) % 4 == 0, "shuffle of odd-sized type");
    int w[sizeof(T) / 4];
    __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
    for (unsigned k = 0; k < sizeof(T) / 4; k++) w[k] = __shfl_xor(w[k], mask, kWave);
    T out;
    __builtin_memcpy(&out, w, sizeof(T));
    return out;
  }
}

template <typename T>
__device__ inline bool is_nan(T v) {
  if constexpr (std::is_floating_point_v<T>) return v != v;
  else return false;
}

// Integer adds wrap (two's complement) like the CPU kernels' native adds.
template <typename OUT>
__device__ inline OUT wrap_add(OUT a, OUT b) {
  if constexpr (std::is_floating_point_v<OUT>) {
    return a + b;
  } else {
    using U = std::make_unsigned_t<OUT>;
    return (OUT)((U)a + (U)b);
  }
}

// ---- ops --------------------------------------------------------------------
template <typename OUT, typename IN>
struct SumOp {
  using In = IN;
  using Out = OUT;
  using Part = OUT;
  __device__ Out sequential(const In* p, int64_t b, int64_t e, int64_t) const {
    // the CPU loop's order; four loads issued ahead of their adds
    OUT acc = OUT(0);
    int64_t i = b;
    for (; i + 4 <= e; i += 4) {
      const In v0 = p[i], v1 = p[i + 1], v2 = p[i + 2], v3 = p[i + 3];
      acc = wrap_add<OUT>(acc, (OUT)v0);
      acc = wrap_add<OUT>(acc, (OUT)v1);
      acc = wrap_add<OUT>(acc, (OUT)v2);
      acc = wrap_add<OUT>(acc, (OUT)v3);
    }
    for (; i < e; i++) acc = wrap_add<OUT>(acc, (OUT)p[i]);
    return acc;
  }
  __device__ Part init() const { return OUT(0); }
  __device__ Part elem(In v, int64_t) const { return (OUT)v; }
  __device__ Part combine(Part a, Part b) const { return wrap_add<OUT>(a, b); }
  __device__ Out finalize(Part p, const In*, int64_t, int64_t) const { return p; }
  __device__ Part shfl_xor(Part p, int m) const { return shfl_xor_any(p, m); }
};

template <typename OUT>
struct SumBoolOp {  // awkward_reduce_sum_{int64,int32}_bool_64: count of true
  using In = bool;
  using Out = OUT;
  using Part = OUT;
  __device__ Out sequential(const In* p, int64_t b, int64_t e, int64_t) const {
    OUT acc = 0;
    for (int64_t i = b; i < e; i++)
      if (p[i]) acc++;
    return acc;
  }
  __device__ Part init() const { return 0; }
  __device__ Part elem(In v, int64_t) const { return v ? 1 : 0; }
  __device__ Part combine(Part a, Part b) const { return a + b; }
  __device__ Out finalize(Part p, const In*, int64_t, int64_t) const { return p; }
  __device__ Part shfl_xor(Part p, int m) const { return shfl_xor_any(p, m); }
};

// Integer products wrap (two's complement) like the CPU kernels' native multiplies.
template <typename OUT>
__device__ inline OUT wrap_mul(OUT a, OUT b) {
  if constexpr (std::is_floating_point_v<OUT>) {
    return a * b;
  } else {
    using U = std::make_unsigned_t<OUT>;
    return (OUT)((U)a * (U)b);
  }
}

// awkward_reduce_prod: acc = 1; acc *= (OUT)x in the CPU loop's order.
template <typename OUT, typename IN>
struct ProdOp {
  using In = IN;
  using Out = OUT;
  using Part = OUT;
  __device__ Out sequential(const In* p, int64_t b, int64_t e, int64_t) const {
    OUT acc = OUT(1);
    int64_t i = b;
    for (; i + 4 <= e; i += 4) {
      const In v0 = p[i], v1 = p[i + 1], v2 = p[i + 2], v3 = p[i + 3];
      acc = wrap_mul<OUT>(acc, (OUT)v0);
      acc = wrap_mul<OUT>(acc, (OUT)v1);
      acc = wrap_mul<OUT>(acc, (OUT)v2);
      acc = wrap_mul<OUT>(acc, (OUT)v3);
    }
    for (; i < e; i++) acc = wrap_mul<OUT>(acc, (OUT)p[i]);
    return acc;
  }
  __device__ Part init() const { return OUT(1); }
  __device__ Part elem(In v, int64_t) const { return (OUT)v; }
  __device__ Part combine(Part a, Part b) const { return wrap_mul<OUT>(a, b); }
  __device__ Out finalize(Part p, const In*, int64_t, int64_t) const { return p; }
  __device__ Part shfl_xor(Part p, int m) const { return shfl_xor_any(p, m); }
};

// awkward_reduce_sum_bool (ANY) / awkward_reduce_prod_bool (!ANY): whether
// any / every element is nonzero (x != 0, so NaN is nonzero and -0.0 zero).
// Order-free, so every path gives the CPU kernel's bits.
template <typename IN, bool ANY>
struct AnyAllOp {
  using In = IN;
  using Out = bool;
  using Part = int;
  __device__ Out sequential(const In* p, int64_t b, int64_t e, int64_t) const {
    int r = ANY ? 0 : 1;
    for (int64_t i = b; i < e; i++) r = ANY ? (r | (p[i] != In(0))) : (r & (p[i] != In(0)));
    return r != 0;
  }
  __device__ Part init() const { return ANY ? 0 : 1; }
  __device__ Part elem(In v, int64_t) const { return v != In(0) ? 1 : 0; }
  __device__ Part combine(Part a, Part b) const { return ANY ? (a | b) : (a & b); }
  __device__ Out finalize(Part p, const In*, int64_t, int64_t) const { return p != 0; }
  __device__ Part shfl_xor(Part p, int m) const { return __shfl_xor(p, m, kWave); }
};

template <typename T>
struct ValIdx {
  T val;
  int64_t idx;  // -1: no candidate
};

// Index-aware extremum combine: ignores missing candidates, prefers the
// strictly better value, and on equal values (including -0.0 == +0.0) the
// smaller index -- the element the serial loop would have kept.
template <typename T, bool MAX>
__device__ inline ValIdx<T> better(ValIdx<T> a, ValIdx<T> b) {
  if (a.idx < 0) return b;
  if (b.idx < 0) return a;
  const bool b_wins = MAX ? (b.val > a.val) : (b.val < a.val);
  const bool a_wins = MAX ? (a.val > b.val) : (a.val < b.val);
  if (b_wins) return b;
  if (a_wins) return a;
  return (b.idx < a.idx) ? b : a;
}

template <typename T>
__device__ inline ValIdx<T> shfl_xor_vi(ValIdx<T> p, int m) {
  return ValIdx<T>{shfl_xor_any(p.val, m), shfl_xor_any(p.idx, m)};
}

// awkward_reduce_max / awkward_reduce_min: best = identity; best = v OP best ? v : best
template <typename T, bool MAX>
struct ExtremumOp {
  using In = T;
  using Out = T;
  using Part = ValIdx<T>;
  T identity;
  __device__ Out sequential(const In* p, int64_t b, int64_t e, int64_t) const {
    T best = identity;
    int64_t i = b;
    for (; i + 4 <= e; i += 4) {
      const T v[4] = {p[i], p[i + 1], p[i + 2], p[i + 3]};
#pragma unroll
      for (int u = 0; u < 4; u++) best = (MAX ? (v[u] > best) : (v[u] < best)) ? v[u] : best;
    }
    for (; i < e; i++) {
      const T v = p[i];
      best = (MAX ? (v > best) : (v < best)) ? v : best;
    }
    return best;
  }
  __device__ Part init() const { return Part{identity, -1}; }
  __device__ Part elem(In v, int64_t gi) const { return Part{v, is_nan(v) ? -1 : gi}; }
  __device__ Part combine(Part a, Part b) const { return better<T, MAX>(a, b); }
  __device__ Out finalize(Part p, const In*, int64_t, int64_t) const {
    if (p.idx < 0) return identity;
    return (MAX ? (p.val > identity) : (p.val < identity)) ? p.val : identity;
  }
  __device__ Part shfl_xor(Part p, int m) const { return shfl_xor_vi(p, m); }
};

// awkward_reduce_argmax / argmin: first element seeds; strict improvement
// only.  A NaN first element therefore wins; later NaNs never do.
template <typename T, bool MAX>
struct ArgOp {
  using In = T;
  using Out = int64_t;
  using Part = ValIdx<T>;
  __device__ Out sequential(const In* p, int64_t b, int64_t e, int64_t gb) const {
    if (b >= e) return -1;
    // positions within the list in 32 bits (lists here hold < 2^31 elements)
    const T* q = p + b;
    const int n = (int)(e - b);
    int best = 0;
    T best_val = q[0];
    int i = 1;
    for (; i + 4 <= n; i += 4) {
      const T v[4] = {q[i], q[i + 1], q[i + 2], q[i + 3]};
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const bool take = MAX ? (v[u] > best_val) : (v[u] < best_val);
        best_val = take ? v[u] : best_val;
        best = take ? i + u : best;
      }
    }
    for (; i < n; i++) {
      const T v = q[i];
      const bool take = MAX ? (v > best_val) : (v < best_val);
      best_val = take ? v : best_val;
      best = take ? i : best;
    }
    return gb + best;
  }
  __device__ Part init() const { return Part{T(0), -1}; }
  __device__ Part elem(In v, int64_t gi) const { return Part{v, is_nan(v) ? -1 : gi}; }
  __device__ Part combine(Part a, Part b) const { return better<T, MAX>(a, b); }
  __device__ Out finalize(Part p, const In* g, int64_t gb, int64_t ge) const {
    if (gb >= ge) return -1;
    if (is_nan(g[gb])) return gb;
    return p.idx;
  }
  __device__ Part shfl_xor(Part p, int m) const { return shfl_xor_vi(p, m); }
};

// sum, min and argmin of each list from one pass over its elements, with
// the semantics of awkward_reduce_sum, awkward_reduce_min and
// awkward_reduce_argmin (the batch runner's fused op).
template <typename T, typename S>
struct MultiOut {
  S sum;
  T min;
  int64_t arg;  // global index, -1 for an empty list
};

template <typename T, typename S>
struct SumMinArgminOp {
  using In = T;
  using Out = MultiOut<T, S>;
  struct Part {
    S sum;
    ValIdx<T> m;
  };
  T identity;  // the min identity
  __device__ Out sequential(const In* p, int64_t b, int64_t e, int64_t gb) const {
    S acc = S(0);
    T best = identity;
    int64_t arg = -1;
    T arg_val = T(0);
    for (int64_t i = b; i < e; i++) {
      const T v = p[i];
      acc = wrap_add<S>(acc, (S)v);
      best = (v < best) ? v : best;
      if (i == b) {
        arg = i;
        arg_val = v;
      } else if (v < arg_val) {
        arg_val = v;
        arg = i;
      }
    }
    return Out{acc, best, arg < 0 ? -1 : gb + (arg - b)};
  }
  __device__ Part init() const { return Part{S(0), ValIdx<T>{identity, -1}}; }
  __device__ Part elem(In v, int64_t gi) const { return Part{(S)v, ValIdx<T>{v, is_nan(v) ? -1 : gi}}; }
  __device__ Part combine(Part a, Part b) const { return Part{wrap_add<S>(a.sum, b.sum), better<T, false>(a.m, b.m)}; }
  __device__ Out finalize(Part p, const In* g, int64_t gb, int64_t ge) const {
    T mn = identity;
    if (p.m.idx >= 0 && p.m.val < identity) mn = p.m.val;
    int64_t arg = -1;
    if (gb < ge) arg = is_nan(g[gb]) ? gb : p.m.idx;
    return Out{p.sum, mn, arg};
  }
  __device__ Part shfl_xor(Part p, int m) const { return Part{shfl_xor_any(p.sum, m), shfl_xor_vi(p.m, m)}; }
};

// ---- the long-list queue ----------------------------------------------------------
constexpr int64_t kReduceChunk = 8192;

struct ReduceChunk {
  int64_t bin;     // output position; -1 marks an unused slot
  int64_t gb, ge;  // the list's element range
  int64_t first;   // queue position of the list's first chunk (its partials start there)
  int32_t chunk, nchunks;
  int32_t item, pad;  // item: which array of a batch
};

struct ReduceQueue {
  ReduceChunk* chunks;
  void* partials;  // one Part per queue position
  unsigned long long* count;
  int64_t capacity;  // 0: no queue
};

// Queue capacity for `elements` elements: every queued list exceeds kWaveLen.
inline int64_t reduce_queue_capacity(int64_t elements) {
  return (elements + kReduceChunk - 1) / kReduceChunk + elements / (kWaveLen + 1) + 2;
}

// Called by every lane of a wave; lanes with `want` try to queue their list
// (one atomic per wave).  If the queue is full the wave's lists stay with
// the workgroup (false) and its reserved slots are marked unused.
__device__ inline bool queue_reduce(const ReduceQueue& q, bool want, int64_t bin, int64_t gb, int64_t ge,
                                    int item) {
  if (!__any(want)) return false;
  const int lane = threadIdx.x & (kWave - 1);
  const unsigned long long nch = want ? (unsigned long long)((ge - gb + kReduceChunk - 1) / kReduceChunk) : 0ull;
  const unsigned long long incl = wave_inclusive_scan(nch);
  const unsigned long long total = __shfl(incl, kWave - 1, kWave);
  unsigned long long base = 0;
  if (lane == 0) base = atomicAdd(q.count, total);
  base = __shfl(base, 0, kWave);
  const int64_t first = (int64_t)(base + incl - nch);
  if ((int64_t)(base + total) > q.capacity) {
    for (int64_t c = 0; c < (int64_t)nch; c++)
      if (first + c < q.capacity) q.chunks[first + c].bin = -1;
    return false;
  }
  // the wave writes each list's chunk entries together, a lane per entry
  uint64_t lists = __ballot(want);
  while (lists) {
    const int src = __builtin_ctzll(lists);
    lists &= lists - 1;
    const int64_t f = __shfl(first, src, kWave);
    const int64_t n = (int64_t)__shfl((long long)nch, src, kWave);
    const int64_t lb = __shfl(bin, src, kWave), lgb = __shfl(gb, src, kWave), lge = __shfl(ge, src, kWave);
    for (int64_t c = lane; c < n; c += kWave)
      q.chunks[f + c] = ReduceChunk{lb, lgb, lge, f, (int32_t)c, (int32_t)n, item, 0};
  }
  return want;
}

template <int Dummy = 0>
__global__ void reduce_queue_reset(unsigned long long* count) { *count = 0ull; }

// One queued chunk per workgroup iteration: a Part over its elements
// (four interleaved accumulators per thread, then the waves in order).
template <class Op, class Src>
__global__ __launch_bounds__(256) void reduce_chunks_kernel(Op op, Src src, ReduceQueue q) {
  using Part = typename Op::Part;
  using In = typename Op::In;
  constexpr int NW = 256 / kWave;
  __shared__ Part s_part[NW];
  int64_t nc = (int64_t)*q.count;
  if (nc > q.capacity) nc = q.capacity;
  Part* parts = (Part*)q.partials;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  for (int64_t c = blockIdx.x; c < nc; c += gridDim.x) {
    const ReduceChunk ch = q.chunks[c];
    if (ch.bin < 0) continue;  // workgroup-uniform
    const In* __restrict__ from = src.data(ch.item);
    const int64_t i0 = ch.gb + (int64_t)ch.chunk * kReduceChunk;
    const int64_t i1 = (i0 + kReduceChunk) < ch.ge ? i0 + kReduceChunk : ch.ge;
    Part p[4] = {op.init(), op.init(), op.init(), op.init()};
    for (int64_t i = i0 + tid; i < i1; i += 4 * 256) {
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int64_t k = i + u * 256;
        if (k < i1) p[u] = op.combine(p[u], op.elem(from[k], k));
      }
    }
    Part r = op.combine(op.combine(p[0], p[1]), op.combine(p[2], p[3]));
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) r = op.combine(r, op.shfl_xor(r, d));
    if (lane == 0) s_part[wave] = r;
    __syncthreads();
    if (tid == 0) {
      Part t = s_part[0];
      for (int w = 1; w < NW; w++) t = op.combine(t, s_part[w]);
      parts[c] = t;
    }
    __syncthreads();
  }
}

// Each queued list: its chunks' partials combined in chunk order, stored by
// the item's epilogue.
template <class Op, class Src>
__global__ void reduce_finalize_kernel(Op op, Src src, ReduceQueue q) {
  using Part = typename Op::Part;
  int64_t nc = (int64_t)*q.count;
  if (nc > q.capacity) nc = q.capacity;
  const Part* parts = (const Part*)q.partials;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x) {
    const ReduceChunk ch = q.chunks[c];
    if (ch.bin < 0 || ch.chunk != 0) continue;
    Part p = parts[ch.first];
    for (int k = 1; k < ch.nchunks; k++) p = op.combine(p, parts[ch.first + k]);
    src.epi(ch.item).store(ch.bin, op.finalize(p, src.data(ch.item), ch.gb, ch.ge), ch.gb, ch.ge);
  }
}

// ---- the tile -----------------------------------------------------------------
// Epilogues own the stores: store(bin, value, gb, ge) with the list's global
// element range [gb, ge).  PlainEpilogue writes the value; the fused
// host-layer variants also make argmin/argmax local and write the empty-list
// mask; the batch runner's epilogue scatters several results.
template <typename Out>
struct PlainEpilogue {
  Out* toptr;
  __device__ void store(int64_t bin, Out v, int64_t, int64_t) const { toptr[bin] = v; }
};

template <class Op, int BLOCK, int LPT, int CAP_BYTES>
struct ReduceSmem {
  using In = typename Op::In;
  using Part = typename Op::Part;
  static constexpr int CAP = CAP_BYTES / (int)sizeof(In);
  static constexpr int NW = BLOCK / kWave;
  static constexpr int L = BLOCK * LPT;
  In data[CAP];
  int64_t off[L + 1];
  uint64_t lng[LPT][NW];
  Part part[NW];
};

// One tile: the lists [L0, L0 + BLOCK * LPT) of `offsets` (clipped to
// outlength); thread t owns lists t + j * BLOCK.  With a queue (capacity >
// 0), lists longer than kWaveLen are queued (item: the batch item they
// belong to).  Ends with a barrier, so a persistent caller may reuse `sm`
// for its next tile at once.
template <class Op, int BLOCK, int LPT, int CAP_BYTES, class Epi>
__device__ inline void reduce_tile(const Op& op, const Epi& epi, const typename Op::In* __restrict__ fromptr,
                                   const int64_t* __restrict__ offsets, int64_t outlength, int64_t L0,
                                   ReduceSmem<Op, BLOCK, LPT, CAP_BYTES>& sm, const ReduceQueue& q = ReduceQueue{},
                                   int item = 0) {
  using In = typename Op::In;
  using Out = typename Op::Out;
  using Part = typename Op::Part;
  constexpr int CAP = CAP_BYTES / (int)sizeof(In);
  constexpr int NW = BLOCK / kWave;
  constexpr int L = BLOCK * LPT;

  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wave = tid / kWave;
  const int nl = (int)((outlength - L0) < L ? (outlength - L0) : L);

  // one wave, one list per lane: each lane holds its list's bounds in
  // registers and the tile's come by shuffles, so a light tile never passes
  // its offsets through LDS (the heavy path below still stages them)
  constexpr bool kRegs = BLOCK == kWave && LPT == 1;
  int64_t own_a = 0, own_b = 0, E0, span;
  if constexpr (kRegs) {
    if (tid < nl) {
      own_a = offsets[L0 + tid];
      own_b = offsets[L0 + tid + 1];
    }
    E0 = __shfl(own_a, 0, kWave);
    span = __shfl(own_b, nl - 1, kWave) - E0;
  } else {
    int64_t r[LPT + 1];
#pragma unroll
    for (int j = 0; j <= LPT; j++) {
      const int t = tid + j * BLOCK;
      if (t <= nl) r[j] = offsets[L0 + t];
    }
#pragma unroll
    for (int j = 0; j <= LPT; j++) {
      const int t = tid + j * BLOCK;
      if (t <= nl) sm.off[t] = r[j];
    }
    __syncthreads();
    E0 = sm.off[0];
    span = sm.off[nl] - E0;
  }

  if (span <= CAP) {
    // light tile: stage the span with coalesced loads, then serial loops.
    // U loads per thread cover the whole stage (up to 16), so a span that
    // fits is staged in one round trip to memory.
    constexpr int U = (CAP + BLOCK - 1) / BLOCK < 16 ? (CAP + BLOCK - 1) / BLOCK : 16;
    for (int64_t b = 0; b < span; b += (int64_t)BLOCK * U) {
      In r[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int64_t k = b + (int64_t)u * BLOCK + tid;
        if (k < span) r[u] = fromptr[E0 + k];
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int64_t k = b + (int64_t)u * BLOCK + tid;
        if (k < span) sm.data[k] = r[u];
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < LPT; j++) {
      const int t = tid + j * BLOCK;
      if (t < nl) {
        const int64_t gb = kRegs ? own_a : sm.off[t], ge = kRegs ? own_b : sm.off[t + 1];
        epi.store(L0 + t, op.sequential(sm.data, gb - E0, ge - E0, gb), gb, ge);
      }
    }
    __syncthreads();
    return;
  }

  // heavy tile
  if constexpr (kRegs) {
    if (tid < nl) sm.off[tid] = own_a;
    if (tid == nl - 1) sm.off[nl] = own_b;
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < LPT; j++) {
    const int t = tid + j * BLOCK;
    int64_t gb = 0, ge = 0;
    if (t < nl) {
      gb = sm.off[t];
      ge = sm.off[t + 1];
      if (ge - gb <= kShortLen) epi.store(L0 + t, op.sequential(fromptr, gb, ge, gb), gb, ge);
    }
    const int64_t len = ge - gb;
    // medium lists: the owning wave reduces each one cooperatively
    uint64_t medium = __ballot(t < nl && len > kShortLen && len <= kWaveLen);
    while (medium) {
      const int b = __builtin_ctzll(medium);
      medium &= medium - 1;
      const int owner = j * BLOCK + wave * kWave + b;
      const int64_t lb = sm.off[owner], le = sm.off[owner + 1];
      Part p = op.init();
      for (int64_t i = lb + lane; i < le; i += kWave) p = op.combine(p, op.elem(fromptr[i], i));
#pragma unroll
      for (int d = kWave / 2; d > 0; d >>= 1) p = op.combine(p, op.shfl_xor(p, d));
      if (lane == 0) epi.store(L0 + owner, op.finalize(p, fromptr, lb, le), lb, le);
    }
    const bool want = t < nl && len > kWaveLen;
    const bool queued = q.capacity > 0 && queue_reduce(q, want, L0 + t, gb, ge, item);
    const uint64_t lng = __ballot(want && !queued);
    if (lane == 0) sm.lng[j][wave] = lng;
  }
  __syncthreads();
  // long lists: the whole workgroup
  for (int j = 0; j < LPT; j++) {
    for (int w = 0; w < NW; w++) {
      uint64_t m = sm.lng[j][w];
      while (m) {
        const int b = __builtin_ctzll(m);
        m &= m - 1;
        const int owner = j * BLOCK + w * kWave + b;
        const int64_t lb = sm.off[owner], le = sm.off[owner + 1];
        Part p = op.init();
#pragma unroll 4
        for (int64_t i = lb + tid; i < le; i += BLOCK) p = op.combine(p, op.elem(fromptr[i], i));
#pragma unroll
        for (int d = kWave / 2; d > 0; d >>= 1) p = op.combine(p, op.shfl_xor(p, d));
        if (lane == 0) sm.part[wave] = p;
        __syncthreads();
        if (tid == 0) {
          Part q = sm.part[0];
          for (int k = 1; k < NW; k++) q = op.combine(q, sm.part[k]);
          epi.store(L0 + owner, op.finalize(q, fromptr, lb, le), lb, le);
        }
        __syncthreads();
      }
    }
  }
  __syncthreads();
  (void)sizeof(Out);
}

template <class Op, int BLOCK, int LPT, int CAP_BYTES, class Epi>
__global__ __launch_bounds__(BLOCK) void segmented_reduce_kernel(Op op, Epi epi,
                                                                 const typename Op::In* __restrict__ fromptr,
                                                                 const int64_t* __restrict__ offsets,
                                                                 int64_t outlength) {
  __shared__ ReduceSmem<Op, BLOCK, LPT, CAP_BYTES> sm;
  reduce_tile<Op, BLOCK, LPT, CAP_BYTES, Epi>(op, epi, fromptr, offsets, outlength, (int64_t)blockIdx.x * BLOCK * LPT,
                                             sm);
}

}  // namespace akhip

namespace akhip {

// Fused epilogue of the host layer's local (axis=-1) arg-reductions: the
// global index becomes local (awkward_NumpyArray_reduce_adjust_starts_64 with
// starts = offsets[:-1]) and the empty-list mask is written in the same pass
// (awkward_NumpyArray_reduce_mask_ByteMaskedArray_64).
//
// Below an option type the reduced content was compacted: `shifts[i]` counts
// the nulls removed before element i and `starts[bin]` is the list's start
// before compaction, so the local index is v + shifts[v] - starts[bin]
// (awkward_NumpyArray_reduce_adjust_starts_shifts_64).
struct LocalArgEpilogue {
  int64_t* toptr;
  int8_t* mask;                     // may be null
  const int64_t* shifts = nullptr;  // null: starts are the offsets
  const int64_t* starts = nullptr;
  __device__ void store(int64_t bin, int64_t v, int64_t gb, int64_t ge) const {
    if (mask) mask[bin] = (ge - gb > 0) ? 0 : 1;
    toptr[bin] = v < 0 ? v : shifts ? v + shifts[v] - starts[bin] : v - gb;
  }
};

// Fused mask for value reductions (max/min with mask_identity=True).
template <typename Out>
struct MaskEpilogue {
  Out* toptr;
  int8_t* mask;
  __device__ void store(int64_t bin, Out v, int64_t gb, int64_t ge) const {
    if (mask) mask[bin] = (ge - gb > 0) ? 0 : 1;
    toptr[bin] = v;
  }
};

template <class Op, int BLOCK, int LPT, int CAP, class Epi>
inline Error launch_reduce_v(const Op& op, const Epi& epi, const typename Op::In* fromptr, const int64_t* offsets,
                             int64_t outlength) {
  if (outlength <= 0) return success();
  const unsigned grid = (unsigned)ceil_div(outlength, (int64_t)BLOCK * LPT);
  segmented_reduce_kernel<Op, BLOCK, LPT, CAP, Epi><<<grid, BLOCK, 0, stream()>>>(op, epi, fromptr, offsets, outlength);
  AKHIP_LAUNCH_CHECK("segmented_reduce_kernel");
  return success();
}

// ---- long lists across the chip ---------------------------------------------------------
// Few lists of many elements (a reduction across lists regroups the data
// into maxcount columns): a tile pass that queues every list longer than
// kWaveLen in kReduceChunk-element chunks, one chunk per workgroup
// iteration over the whole chip, and the chunk partials combined in order
// per list.  The queue has a fixed capacity (kQueuedChunks chunks: 2^31
// queued elements), so nothing is read back to size it and the sequence can
// be captured; a wave whose lists do not fit keeps them with its workgroup
// (queue_reduce), which is slower but gives the same result.
template <typename In, class Epi>
struct SingleSrc {
  const In* p;
  Epi e;
  __device__ const In* data(int) const { return p; }
  __device__ Epi epi(int) const { return e; }
};

template <class Op, class Epi>
__global__ __launch_bounds__(256) void segmented_reduce_queued_kernel(Op op, Epi epi,
                                                                      const typename Op::In* __restrict__ fromptr,
                                                                      const int64_t* __restrict__ offsets,
                                                                      int64_t outlength, ReduceQueue q) {
  __shared__ ReduceSmem<Op, 256, 1, 8192> sm;
  reduce_tile<Op, 256, 1, 8192, Epi>(op, epi, fromptr, offsets, outlength, (int64_t)blockIdx.x * 256, sm, q, 0);
}

constexpr int64_t kQueuedChunks = int64_t(1) << 18;

template <class Op, class Epi>
inline Error launch_reduce_queued(const Op& op, const Epi& epi, const typename Op::In* fromptr,
                                  const int64_t* offsets, int64_t outlength) {
  using Part = typename Op::Part;
  using In = typename Op::In;
  if (outlength <= 0) return success();
  const int64_t cap = kQueuedChunks;
  const size_t cb = ((size_t)cap * sizeof(ReduceChunk) + 255) / 256 * 256;
  const size_t pb = ((size_t)cap * sizeof(Part) + 255) / 256 * 256;
  char* m = (char*)scratch(256 + cb + pb);
  if (!m) return failure("out of device memory for the long-list queue", kSliceNone, kSliceNone, "awkward_amd/csrc/reduce.h");
  const ReduceQueue q{(ReduceChunk*)(m + 256), (void*)(m + 256 + cb), (unsigned long long*)m, cap};
  hipStream_t st = stream();
  reduce_queue_reset<><<<1, 1, 0, st>>>(q.count);
  segmented_reduce_queued_kernel<Op, Epi><<<(unsigned)ceil_div(outlength, (int64_t)256), 256, 0, st>>>(
      op, epi, fromptr, offsets, outlength, q);
  const SingleSrc<In, Epi> src{fromptr, epi};
  reduce_chunks_kernel<Op, SingleSrc<In, Epi>><<<(unsigned)cu_count() * 8, 256, 0, st>>>(op, src, q);
  reduce_finalize_kernel<Op, SingleSrc<In, Epi>><<<(unsigned)cu_count(), 256, 0, st>>>(op, src, q);
  AKHIP_LAUNCH_CHECK("segmented_reduce_queued_kernel");
  return success();
}

// Tile variants (threads, lists per thread, LDS staging bytes).  The host
// layer picks one from the average list length
// (elements / lists, known from the array shapes without a device sync), and
// the autotuner and the learned dispatcher search all of them.  Small static
// LDS tiles let more workgroups -- and more loads -- be in flight per CU.
//   0: 256 x 1, 32 KiB    1: 256 x 4, 8 KiB     2: 256 x 2, 16 KiB
//   3: 128 x 1, 16 KiB    4: 256 x 1, 64 KiB    5: 256 x 8, 8 KiB
//   6:  64 x 1,  6 KiB    7: 128 x 2, 16 KiB    8:  64 x 4,  8 KiB
//   9: 128 x 1,  8 KiB
//  10: the long-list queue above (chosen for mean lengths of thousands; not
//      one of the searched tiles: it synchronises once)
constexpr int kVariants = 10;
constexpr int kLongLists = 10;

// The plain entry points use variant 3 (128 threads, one list each, 16 KiB):
// the fastest tile for lists of ~4-30 elements at every size measured.
template <class Op>
inline Error launch_reduce(const Op& op, typename Op::Out* toptr, const typename Op::In* fromptr,
                           const int64_t* offsets, int64_t outlength) {
  using E = PlainEpilogue<typename Op::Out>;
  return launch_reduce_v<Op, 128, 1, 16384, E>(op, E{toptr}, fromptr, offsets, outlength);
}

template <class Op, class Epi>
inline Error launch_reduce(const Op& op, const typename Op::In* fromptr, const int64_t* offsets, int64_t outlength,
                           int variant, Epi epi) {
  switch (variant) {
    case kLongLists: return launch_reduce_queued<Op, Epi>(op, epi, fromptr, offsets, outlength);
    case 1: return launch_reduce_v<Op, 256, 4, 8192, Epi>(op, epi, fromptr, offsets, outlength);
    case 2: return launch_reduce_v<Op, 256, 2, 16384, Epi>(op, epi, fromptr, offsets, outlength);
    case 3: return launch_reduce_v<Op, 128, 1, 16384, Epi>(op, epi, fromptr, offsets, outlength);
    case 4: return launch_reduce_v<Op, 256, 1, 65536, Epi>(op, epi, fromptr, offsets, outlength);
    case 5: return launch_reduce_v<Op, 256, 8, 8192, Epi>(op, epi, fromptr, offsets, outlength);
    case 6: return launch_reduce_v<Op, 64, 1, 6144, Epi>(op, epi, fromptr, offsets, outlength);
    case 7: return launch_reduce_v<Op, 128, 2, 16384, Epi>(op, epi, fromptr, offsets, outlength);
    case 8: return launch_reduce_v<Op, 64, 4, 8192, Epi>(op, epi, fromptr, offsets, outlength);
    case 9: return launch_reduce_v<Op, 128, 1, 8192, Epi>(op, epi, fromptr, offsets, outlength);
    default: return launch_reduce_v<Op, 256, 1, 32768, Epi>(op, epi, fromptr, offsets, outlength);
  }
}

}  // namespace akhip
