// pair_jit.h -- the pair-reduction kernel for one program, compiled at run
// time (hipRTC) by jit.hip: ak.max / ak.min / ak.argmax / ak.argmin /
// ak.sum at axis=-1 of an elementwise expression of ak.combinations(array,
// 2) (listoffsetarray.py:1152-1261, array_module.py:254-332,
// awkward_reduce_{max,min,argmax,argmin,sum}.cpp).
//
// The including source defines, in akhip::pjit, the program's shape as
// constants -- W (register word), F (output float), kMX / kMN / kSUM (which
// reductions), kTUP (carries written), kStage (elements per staged column),
// kCap (pair values kept per workgroup), kU (pairs a thread evaluates
// together), kNLeaves, kNSlots, kSlot[] / kLeg[] / kType[] / kIndexed[] per
// leaf, kFirst[] per staged column -- and
// `eval(l)`, the program as straight-line code over the leaves' canonical
// values, each op computed as expr.h's interpreter computes it, so every
// value is the interpreter's, bit for bit (the source is compiled with
// contraction off).
//
// A workgroup takes 128 lists and stages the distinct columns of their
// elements in LDS.  When the lists' pairs fit kCap, they are spread evenly
// over the threads in combinations order (i < j by position, lexicographic):
// each thread finds its first pair's list by binary search over the pair
// prefix, steps through the rest and writes the values to LDS.  Then every
// thread reduces its own list's values in order with the CPU kernels' loops
// -- the first strict improvement for arg-reductions, left-to-right
// accumulation for sums -- so the results are the elementwise form's bits.
// Lists whose pairs do not fit are enumerated by their own thread, from the
// stage or, past it, from memory.  With kTUP, the workgroup then writes its
// lists' ak.combinations carries (n = 2 and / or 3, tuples.h): a store
// stream that overlaps the arithmetic of the other workgroups on the CU.
#pragma once

#include "tuples.h"

namespace akhip {
namespace pjit {

constexpr int kLists = kTupLists;  // lists (and threads) per workgroup

struct Leaves {
  const void* data[kExprMaxLeaves];
  const int64_t* index[kExprMaxLeaves];
};

// up to two tuple sizes n (2 or 3), each with the lists' tuple offsets
// (the combinations_length scan's tooffsets) and its n carries
struct Tuples {
  int32_t nreq;
  int32_t n[2];
  const int64_t* tooffsets[2];
  int64_t* carry[2][3];
};

// One list's N-tuples written by its own thread (a group too large for
// 32-bit relative positions): positions a .. a + m - 1, slots from g.
template <int N>
__device__ inline void list_tuples(int64_t* const* carry, int64_t g, int64_t a, int64_t m) {
  if constexpr (N == 2) {
    for (int64_t i = 0; i < m; i++)
      for (int64_t j = i + 1; j < m; j++, g++) {
        carry[0][g] = a + i;
        carry[1][g] = a + j;
      }
  } else {
    for (int64_t i = 0; i < m; i++)
      for (int64_t j = i + 1; j < m; j++)
        for (int64_t k = j + 1; k < m; k++, g++) {
          carry[0][g] = a + i;
          carry[1][g] = a + j;
          carry[2][g] = a + k;
        }
  }
}

struct Out {
  void* max;  // F
  void* min;
  void* sum;
  int64_t* argmax;  // local pair index, -1 for a list without pairs
  int64_t* argmin;
  int8_t* mask;  // 1 for a list without pairs
};

// leaf k as the interpreter's loader takes it; whether it reads through an
// index is part of the program's shape (kIndexed), so for the plain columns
// the loader's index test folds away
__device__ inline ExprLeaf leaf_of(const Leaves& lv, int k) {
  return ExprLeaf{lv.data[k], kIndexed[k] ? lv.index[k] : nullptr, kType[k], 0};
}

__device__ inline F as_f(W w) {
  if constexpr (sizeof(F) == 4) return __uint_as_float((uint32_t)w);
  else return __longlong_as_double((long long)w);
}

__device__ inline int wave_scan(int v) {
  const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const int u = __shfl_up(v, d, kWave);
    if (lane >= d) v += u;
  }
  return v;
}

}  // namespace pjit
}  // namespace akhip

extern "C" __global__ __launch_bounds__(128) void akhip_pair_jit(akhip::pjit::Leaves lv,
                                                                 const int64_t* __restrict__ offsets, int64_t n,
                                                                 akhip::pjit::Out out, akhip::pjit::Tuples tp) {
  using namespace akhip;
  using namespace akhip::pjit;
  extern __shared__ uint64_t s_dyn[];
  W* s_col = (W*)s_dyn;  // kNSlots columns of kStage
  F* s_val = (F*)(s_col + (kNSlots > 0 ? kNSlots : 1) * kStage);
  __shared__ int64_t s_off[kLists + 1];
  __shared__ int s_pbeg[kLists + 1];
  __shared__ int s_wsum[kLists / kWave];
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1), wave = tid / kWave;
  const int64_t L0 = (int64_t)blockIdx.x * kLists;
  const int nl = (int)((n - L0) < kLists ? (n - L0) : kLists);
  for (int t = tid; t <= nl; t += kLists) s_off[t] = offsets[L0 + t];
  // the carries' tuple offsets, read before any store: a load issued after
  // stores waits for them to complete (loads and stores share one counter)
  int64_t t_own[2] = {0, 0}, t_first[2] = {0, 0}, t_last[2] = {0, 0};
  if constexpr (kTUP) {
#pragma unroll
    for (int k = 0; k < 2; k++) {
      if (k < tp.nreq) {
        if (tid < nl) t_own[k] = tp.tooffsets[k][L0 + tid];
        t_first[k] = tp.tooffsets[k][L0];
        t_last[k] = tp.tooffsets[k][L0 + nl];
      }
    }
  }
  __syncthreads();
  const int64_t E0 = s_off[0];
  const int64_t span = s_off[nl] - E0;
  const bool staged = span >= 0 && span <= kStage;
  if (staged) {
    for (int q = tid; q < (int)span; q += kLists) {
#pragma unroll
      for (int s = 0; s < kNSlots; s++) s_col[s * kStage + q] = x_load<W>(leaf_of(lv, kFirst[s]), E0 + q);
    }
  }
  // the pairs of this thread's list and their prefix over the workgroup
  // (staged: at most kStage elements, so the total stays below 2^31)
  const int m_own = (staged && tid < nl) ? (int)(s_off[tid + 1] - s_off[tid]) : 0;
  const int c_own = m_own > 1 ? m_own * (m_own - 1) / 2 : 0;
  const int incl = wave_scan(c_own);
  if (lane == kWave - 1) s_wsum[wave] = incl;
  __syncthreads();
  int wbase = 0, total = 0;
#pragma unroll
  for (int w = 0; w < kLists / kWave; w++) {
    wbase += (w < wave) ? s_wsum[w] : 0;
    total += s_wsum[w];
  }
  s_pbeg[tid] = wbase + incl - c_own;
  if (tid == 0) s_pbeg[kLists] = total;
  __syncthreads();
  const bool pairwise = staged && total <= kCap;
  if (pairwise && total > 0) {
    const int per = (total + kLists - 1) / kLists;
    int q = tid * per;
    const int qend = q + per < total ? q + per : total;
    if (q < qend) {
      int lo = 0, hi = nl;  // the list holding pair q: the last e with s_pbeg[e] <= q
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (s_pbeg[mid] <= q) lo = mid;
        else hi = mid;
      }
      int e = lo;
      int m = (int)(s_off[e + 1] - s_off[e]);
      int r = q - s_pbeg[e];
      int i = 0;
      while (r >= m - 1 - i) {
        r -= m - 1 - i;
        i++;
      }
      int j = i + 1 + r;
      int a = (int)(s_off[e] - E0);
      while (q < qend) {
        int li[kU], lj[kU];
        const int qs = q;
#pragma unroll
        for (int u = 0; u < kU; u++) {
          li[u] = a + i;
          lj[u] = a + j;
          if (q < qend) {
            q++;
            if (++j == m) {
              ++i;
              j = i + 1;
            }
            if (i >= m - 1 && q < qend) {  // on to the next list with pairs
              do {
                e++;
              } while (s_pbeg[e + 1] == s_pbeg[e]);
              m = (int)(s_off[e + 1] - s_off[e]);
              a = (int)(s_off[e] - E0);
              i = 0;
              j = 1;
            }
          }
        }
        W v[kU];
#pragma unroll
        for (int u = 0; u < kU; u++) {
          W l[kNLeaves > 0 ? kNLeaves : 1];
#pragma unroll
          for (int k = 0; k < kNLeaves; k++) l[k] = s_col[kSlot[k] * kStage + (kLeg[k] ? lj[u] : li[u])];
          v[u] = eval(l);
        }
#pragma unroll
        for (int u = 0; u < kU; u++)
          if (qs + u < qend) s_val[qs + u] = as_f(v[u]);
      }
    }
  }
  __syncthreads();
  if (tid < nl) {
    const int64_t b = L0 + tid;
    F best_max = (F)-__builtin_inf(), best_min = (F)__builtin_inf(), sum = F(0), vmax = F(0), vmin = F(0);
    int64_t amax = -1, amin = -1, rk = 0;
    auto take = [&](F v) {
      if constexpr (kMX) best_max = (v > best_max) ? v : best_max;  // awkward_reduce_max.cpp
      if constexpr (kMN) best_min = (v < best_min) ? v : best_min;  // awkward_reduce_min.cpp
      if constexpr (kSUM) sum += v;                                 // awkward_reduce_sum.cpp: left to right
      if (rk == 0) {                                                // awkward_reduce_argmax.cpp: the first seeds
        amax = amin = 0;
        vmax = vmin = v;
      } else {
        if constexpr (kMX) {
          if (v > vmax) {
            vmax = v;
            amax = rk;
          }
        }
        if constexpr (kMN) {
          if (v < vmin) {
            vmin = v;
            amin = rk;
          }
        }
      }
      rk++;
    };
    if (pairwise) {
      for (int q = s_pbeg[tid]; q < s_pbeg[tid + 1]; q++) take(s_val[q]);
    } else if (staged) {  // this list's pairs one at a time from the stage
      const int a = (int)(s_off[tid] - E0), e = (int)(s_off[tid + 1] - E0);
      for (int i = a; i < e; i++)
        for (int j = i + 1; j < e; j++) {
          W l[kNLeaves > 0 ? kNLeaves : 1];
#pragma unroll
          for (int k = 0; k < kNLeaves; k++) l[k] = s_col[kSlot[k] * kStage + (kLeg[k] ? j : i)];
          take(as_f(eval(l)));
        }
    } else {  // a span too long to stage: the leaves from memory
      const int64_t a = s_off[tid], e = s_off[tid + 1];
      for (int64_t i = a; i < e; i++)
        for (int64_t j = i + 1; j < e; j++) {
          W l[kNLeaves > 0 ? kNLeaves : 1];
#pragma unroll
          for (int k = 0; k < kNLeaves; k++) l[k] = x_load<W>(leaf_of(lv, k), kLeg[k] ? j : i);
          take(as_f(eval(l)));
        }
    }
    if constexpr (kMX) {
      if (out.max) ((F*)out.max)[b] = best_max;
      if (out.argmax) out.argmax[b] = amax;
    }
    if constexpr (kMN) {
      if (out.min) ((F*)out.min)[b] = best_min;
      if (out.argmin) out.argmin[b] = amin;
    }
    if constexpr (kSUM) {
      if (out.sum) ((F*)out.sum)[b] = sum;
    }
    if (out.mask) out.mask[b] = rk > 0 ? 0 : 1;
  }
  if constexpr (kTUP) {
    // the carries: the stage, the values and s_pbeg are free now
    int* stage = (int*)s_dyn;
    int* s_tbl = s_pbeg;
    __shared__ int s_offl[kLists + 1];
#pragma unroll
    for (int k = 0; k < 2; k++) {
      if (k >= tp.nreq) break;
      const int64_t g0 = t_first[k];
      const bool narrow = span >= 0 && span < (int64_t)0x7fffffff && t_last[k] - g0 < (int64_t)0x7fffffff;
      __syncthreads();  // the previous request's stage and s_tbl
      if (narrow) {  // uniform
        if (tid < nl) {
          s_offl[tid] = (int)(s_off[tid] - E0);
          s_tbl[tid] = (int)(t_own[k] - g0);
        }
        if (tid == 0) {
          s_offl[nl] = (int)(s_off[nl] - E0);
          s_tbl[nl] = (int)(t_last[k] - g0);
        }
        __syncthreads();
        if (tp.n[k] == 2) group_tuples<2>(tp.carry[k], g0, E0, s_tbl, s_offl, nl, stage, tuple_by_count<2>);
        else group_tuples<3>(tp.carry[k], g0, E0, s_tbl, s_offl, nl, stage, tuple_by_count<3>);
      } else if (tid < nl) {
        const int64_t a = s_off[tid], m = s_off[tid + 1] - a;
        if (tp.n[k] == 2) list_tuples<2>(tp.carry[k], t_own[k], a, m);
        else list_tuples<3>(tp.carry[k], t_own[k], a, m);
      }
    }
  }
}
