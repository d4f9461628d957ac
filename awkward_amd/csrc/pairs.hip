// pairs.hip -- reductions over the pairs of each list of a program (expr.h):
// ak.max / ak.min / ak.argmax / ak.argmin / ak.sum at axis=-1 of an
// elementwise expression of ak.combinations(array, 2) (the reference runs
// awkward_ListArray_combinations_length + awkward_ListArray_combinations,
// carries every field of both legs, applies the ufuncs one at a time and
// reduces: listoffsetarray.py:1152-1261, array_module.py:254-332,
// awkward_reduce_{max,min,argmax,argmin,sum}.cpp).  Here the pairs are never
// materialised: each workgroup takes 128 lists, stages the columns its
// program reads (each distinct column once, in canonical register form) for
// the lists' element span in LDS, spreads the lists' pairs evenly over its
// threads in combinations order (i < j by position, lexicographic),
// evaluates the program with its leg-0 leaves at element i and its leg-1
// leaves at element j, keeps the values in LDS, and each thread then reduces
// its list's values in order with the CPU kernels' loops -- the first strict
// improvement for arg-reductions, left-to-right accumulation for sums -- so
// the results are the elementwise form's bits.

#include "combin.h"
#include "common.h"
#include "expr.h"
#include "jit.h"
#include "tuples.h"

namespace akhip {
namespace {

const char* kFile = "awkward_amd/csrc/pairs.hip";

constexpr int kPRLists = kTupLists;  // lists (and threads) per workgroup
constexpr int kPRCap = 2048;   // pair values kept per workgroup
constexpr int kPRStage = 768;  // elements per staged column

struct PairOut {
  void* max;        // out_type
  void* min;        // out_type
  void* sum;        // out_type
  int64_t* argmax;  // local pair index, -1 for a list without pairs
  int64_t* argmin;
  int8_t* mask;  // mask_identity: 1 for a list without pairs
};

// The program's distinct columns: leaf k reads staged column slot[k]; staged
// column s is loaded through leaf first[s].
struct PairSlots {
  int8_t slot[kExprMaxLeaves];
  int8_t first[kExprMaxLeaves];
  int32_t nslots;
};

// Combinations decoded by the same launch: up to two tuple sizes n (2 or 3,
// without replacement), each with the lists' tuple offsets (the
// combinations_length scan's tooffsets) and its n carries.
struct CombReq {
  int32_t nreq;
  int32_t n[2];
  const int64_t* tooffsets[2];
  int64_t* carry[2][3];
};

// The rank-r N-tuple (lexicographic) of a list of m items, in 32 bits.
template <int N>
__device__ inline void decode_tuple32(int r, int m, int (&x)[N]) {
  constexpr int TABMAX = N == 2 ? kColex2Max : kColex3Max;
  constexpr int BITS = N == 2 ? 6 : 5;
  constexpr int MASK = (1 << BITS) - 1;
  if (m <= TABMAX) {
    const int total = N == 2 ? m * (m - 1) / 2 : m * (m - 1) * (m - 2) / 6;
    const uint32_t y = N == 2 ? kColex.pairs[total - 1 - r] : kColex.triples[total - 1 - r];
#pragma unroll
    for (int q = 0; q < N; q++) x[q] = m - 1 - (int)((y >> (BITS * (N - 1 - q))) & MASK);
  } else {
    int64_t x64[N];
    if constexpr (N == 2) unrank2(r, m, x64);
    else unrank3(r, m, x64);
#pragma unroll
    for (int q = 0; q < N; q++) x[q] = (int)x64[q];
  }
}

// The same for a group whose elements or tuples overflow 32 bits: slot by
// slot (q = t, t + 128, ...), each by binary search and decode.
template <int N>
__device__ inline void group_tuples_wide(int64_t* const* carry, const int64_t* s_tb, const int64_t* s_off, int nl) {
  const int64_t g0 = s_tb[0];
  const int64_t tot = s_tb[nl] - g0;
  for (int64_t q = threadIdx.x; q < tot; q += kPRLists) {
    const int64_t slot = g0 + q;
    int lo = 0, hi = nl;  // the last list e with s_tb[e] <= slot
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (s_tb[mid] <= slot) lo = mid;
      else hi = mid;
    }
    int64_t x[N];
    decode_tuple<N>(slot - s_tb[lo], s_off[lo + 1] - s_off[lo], x);
#pragma unroll
    for (int c = 0; c < N; c++) carry[c][slot] = s_off[lo] + x[c];
  }
}

template <typename F>
__device__ inline F as_value(uint64_t w) {
  if constexpr (sizeof(F) == 4) return __uint_as_float((uint32_t)w);
  else return __longlong_as_double((long long)w);
}

// The program at U pairs whose elements sit at staged positions li[u] (leg 0)
// and lj[u] (leg 1); li[u] < 0 marks an unused slot.  Staged columns are
// STRIDE elements apart.
template <int U, typename W, int STRIDE = kPRStage>
__device__ inline void eval_staged(const ExprProgram& p, const PairSlots& ps, const W* s_col, const int (&li)[U],
                                   const int (&lj)[U], W (&out)[U]) {
  XFileV<U, W> f = 0;
#pragma unroll
  for (int k = 0; k < kExprMaxLeaves && k < XFile<U, W>::regs; k++) {
    if (k < p.nleaves) {
      const W* col = s_col + ps.slot[k] * STRIDE;
      const bool second = p.leaf[k].leg != 0;
#pragma unroll
      for (int u = 0; u < U; u++)
        if (li[u] >= 0) f[k * U + u] = col[second ? lj[u] : li[u]];
    }
  }
  expr_run<U, W>(p, f, out);
}

template <typename W, int U, typename F, bool PROG>
__global__ __launch_bounds__(kPRLists) void pairs_kernel(ExprProgram p, PairSlots ps, int prog,
                                                         const int64_t* __restrict__ offsets, int64_t n, PairOut out,
                                                         CombReq cr) {
  // dynamic LDS: the tuple stage (phase A), then the staged columns and the
  // pair values (phase B)
  extern __shared__ uint64_t s_dyn[];
  int* stage = (int*)s_dyn;
  W* s_col = (W*)s_dyn;
  F* s_val = (F*)(s_col + (ps.nslots > 0 ? ps.nslots : 1) * kPRStage);
  __shared__ int64_t s_off[kPRLists + 1];
  __shared__ int64_t s_tb[kPRLists + 1];
  __shared__ int s_offl[kPRLists + 1];
  __shared__ int s_tbl[kPRLists + 1];
  __shared__ int s_pbeg[kPRLists + 1];
  __shared__ int s_wsum[kPRLists / kWave];
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1), wave = tid / kWave;
  const int64_t L0 = (int64_t)blockIdx.x * kPRLists;
  const int nl = (int)((n - L0) < kPRLists ? (n - L0) : kPRLists);
  for (int t = tid; t <= nl; t += kPRLists) s_off[t] = offsets[L0 + t];
  // the requests' tuple offsets (entries tid and tid + 128), read before any
  // store: a load issued after stores waits for them to complete (loads and
  // stores share one counter)
  int64_t t_pre[2][2] = {{0, 0}, {0, 0}};
#pragma unroll
  for (int k = 0; k < 2; k++)
#pragma unroll
    for (int j = 0; j < 2; j++)
      if (k < cr.nreq && tid + j * kPRLists <= nl) t_pre[k][j] = cr.tooffsets[k][L0 + tid + j * kPRLists];
  // phase A: the carries of the requested combinations (write-bound; the
  // program's evaluation below is ALU-bound, and the workgroups on a CU
  // overlap the two)
#pragma unroll
  for (int k = 0; k < 2; k++) {
    if (k >= cr.nreq) break;
#pragma unroll
    for (int j = 0; j < 2; j++)
      if (tid + j * kPRLists <= nl) s_tb[tid + j * kPRLists] = t_pre[k][j];
    __syncthreads();
    const int64_t E0 = s_off[0], g0 = s_tb[0];
    const bool narrow = s_off[nl] - E0 < (int64_t)INT32_MAX && s_tb[nl] - g0 < (int64_t)INT32_MAX;  // uniform
    if (narrow) {
      for (int t = tid; t <= nl; t += kPRLists) {
        s_offl[t] = (int)(s_off[t] - E0);
        s_tbl[t] = (int)(s_tb[t] - g0);
      }
      __syncthreads();
      if (cr.n[k] == 2) group_tuples<2>(cr.carry[k], g0, E0, s_tbl, s_offl, nl, stage, tuple_by_count<2>);
      else group_tuples<3>(cr.carry[k], g0, E0, s_tbl, s_offl, nl, stage, tuple_by_count<3>);
    } else {
      if (cr.n[k] == 2) group_tuples_wide<2>(cr.carry[k], s_tb, s_off, nl);
      else group_tuples_wide<3>(cr.carry[k], s_tb, s_off, nl);
    }
    __syncthreads();  // s_tb is rewritten by the next request
  }
  if constexpr (!PROG) return;
  if (!prog) return;  // uniform
  __syncthreads();    // s_off (with no requests, nothing above waited for it)
  const int64_t E0 = s_off[0];
  const int64_t span = s_off[nl] - E0;
  const bool staged = span <= kPRStage;
  if (staged) {
    for (int q = tid; q < (int)span; q += kPRLists) {
#pragma unroll
      for (int c = 0; c < kExprMaxLeaves; c++)
        if (c < ps.nslots) s_col[c * kPRStage + q] = x_load<W>(p.leaf[ps.first[c]], E0 + q);
    }
  }
  // the pairs of this thread's list and their prefix over the workgroup
  // (staged: at most kPRStage elements, so the total stays below 2^31)
  const int m_own = (staged && tid < nl) ? (int)(s_off[tid + 1] - s_off[tid]) : 0;
  const int c_own = m_own * (m_own - 1) / 2;
  const int incl = wave_inclusive_scan(c_own);
  if (lane == kWave - 1) s_wsum[wave] = incl;
  __syncthreads();
  int wbase = 0, total = 0;
  for (int w = 0; w < kPRLists / kWave; w++) {
    wbase += (w < wave) ? s_wsum[w] : 0;
    total += s_wsum[w];
  }
  s_pbeg[tid] = wbase + incl - c_own;
  if (tid == 0) s_pbeg[kPRLists] = total;
  __syncthreads();
  const bool pairwise = staged && total <= kPRCap;
  if (pairwise && total > 0) {
    const int per = (total + kPRLists - 1) / kPRLists;
    int q = tid * per;
    const int qend = q + per < total ? q + per : total;
    if (q < qend) {
      // the list holding pair q: the last e with s_pbeg[e] <= q
      int lo = 0, hi = nl;
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
        int li[U], lj[U];
        const int qs = q;
#pragma unroll
        for (int u = 0; u < U; u++) {
          if (q < qend) {
            li[u] = a + i;
            lj[u] = a + j;
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
          } else {
            li[u] = lj[u] = -1;
          }
        }
        W v[U];
        eval_staged<U, W>(p, ps, s_col, li, lj, v);
#pragma unroll
        for (int u = 0; u < U; u++)
          if (li[u] >= 0) s_val[qs + u] = as_value<F>((uint64_t)v[u]);
      }
    }
  }
  __syncthreads();
  if (tid >= nl) return;
  const int64_t b = L0 + tid;
  F best_max = -INFINITY, best_min = INFINITY, sum = F(0), vmax = F(0), vmin = F(0);
  int64_t amax = -1, amin = -1, r = 0;
  auto take = [&](F v) {
    best_max = (v > best_max) ? v : best_max;  // awkward_reduce_max.cpp
    best_min = (v < best_min) ? v : best_min;  // awkward_reduce_min.cpp
    sum += v;                                  // awkward_reduce_sum.cpp: left to right
    if (r == 0) {                              // awkward_reduce_argmax.cpp: the first seeds
      amax = amin = 0;
      vmax = vmin = v;
    } else {
      if (v > vmax) {
        vmax = v;
        amax = r;
      }
      if (v < vmin) {
        vmin = v;
        amin = r;
      }
    }
    r++;
  };
  if (pairwise) {
    for (int q = s_pbeg[tid]; q < s_pbeg[tid + 1]; q++) take(s_val[q]);
  } else if (staged) {  // this list's pairs one at a time from the staged columns
    const int a = (int)(s_off[tid] - E0), e = (int)(s_off[tid + 1] - E0);
    for (int i = a; i < e; i++)
      for (int j = i + 1; j < e; j++) {
        const int li[1] = {i}, lj[1] = {j};
        W v[1];
        eval_staged<1, W>(p, ps, s_col, li, lj, v);
        take(as_value<F>((uint64_t)v[0]));
      }
  } else {  // a span too long to stage: the leaves from memory
    const int64_t a = s_off[tid], e = s_off[tid + 1];
    for (int64_t i = a; i < e; i++)
      for (int64_t j = i + 1; j < e; j++) {
        const int64_t pos0[1] = {i}, pos1[1] = {j};
        W v[1];
        expr_eval_pairs<1, W>(p, pos0, pos1, v);
        take(as_value<F>((uint64_t)v[0]));
      }
  }
  if (out.max) ((F*)out.max)[b] = best_max;
  if (out.min) ((F*)out.min)[b] = best_min;
  if (out.sum) ((F*)out.sum)[b] = sum;
  if (out.argmax) out.argmax[b] = amax;
  if (out.argmin) out.argmin[b] = amin;
  if (out.mask) out.mask[b] = r > 0 ? 0 : 1;
}

template <typename F>
inline void launch_pairs(const ExprProgram& p, const PairSlots& ps, int prog, const int64_t* offsets, int64_t n,
                         const PairOut& out, const CombReq& cr) {
  const unsigned grid = (unsigned)ceil_div(n, (int64_t)kPRLists);
  int maxn = 0;
  for (int k = 0; k < cr.nreq; k++) maxn = cr.n[k] > maxn ? cr.n[k] : maxn;
  const size_t lds_a = (size_t)maxn * kTupPad * sizeof(int);
  if (!prog) {  // carries only: no staged columns
    pairs_kernel<uint32_t, 4, F, false><<<grid, kPRLists, lds_a, stream()>>>(p, ps, 0, offsets, n, out, cr);
    return;
  }
  const size_t w = p.wide ? 8 : 4;
  size_t lds = (size_t)(ps.nslots > 0 ? ps.nslots : 1) * kPRStage * w + kPRCap * sizeof(F);
  if (lds < lds_a) lds = lds_a;
  expr_dispatch(p, [&](auto kind) {
    using W = typename decltype(kind)::Word;
    constexpr int U = decltype(kind)::slots;
    pairs_kernel<W, U, F, true><<<grid, kPRLists, lds, stream()>>>(p, ps, 1, offsets, n, out, cr);
  });
}

// Leaves that read the same column (as the two legs of one field do) share
// a staged column.
inline PairSlots pair_slots(const ExprProgram& p) {
  PairSlots ps{};
  for (int k = 0; k < p.nleaves; k++) {
    int s = -1;
    for (int c = 0; c < ps.nslots && s < 0; c++) {
      const ExprLeaf& f = p.leaf[ps.first[c]];
      if (f.data == p.leaf[k].data && f.index == p.leaf[k].index && f.type == p.leaf[k].type) s = c;
    }
    if (s < 0) {
      s = ps.nslots++;
      ps.first[s] = (int8_t)k;
    }
    ps.slot[k] = (int8_t)s;
  }
  return ps;
}

}  // namespace
}  // namespace akhip

using namespace akhip;

extern "C" {

Error awkward_hip_combinations_pairs(const int64_t* offsets, int64_t n, int32_t nreq, const int32_t* req_n,
                                     const int64_t* const* req_tooffsets, int64_t* const* req_carries,
                                     const ExprProgram* program, void* out_max, void* out_min, void* out_sum,
                                     int64_t* out_argmax, int64_t* out_argmin, int8_t* out_mask);

// Reductions over the pairs (i < j, lexicographic) of the n lists of int64
// offsets (absolute positions in the program's columns): every output may be
// null; max / min / sum have the program's out_type (float32 or float64),
// argmax / argmin are local pair indices (-1 without pairs), mask is 1 for a
// list without pairs.  max / min start from -inf / +inf (the reducers'
// identities).  The program pointer is a HOST pointer.
Error awkward_hip_pair_reduce(const int64_t* offsets, int64_t n, const ExprProgram* program, void* out_max,
                              void* out_min, void* out_sum, int64_t* out_argmax, int64_t* out_argmin,
                              int8_t* out_mask) {
  return awkward_hip_combinations_pairs(offsets, n, 0, nullptr, nullptr, nullptr, program, out_max, out_min,
                                        out_sum, out_argmax, out_argmin, out_mask);
}

// ak.combinations(array, k) for up to two k in {2, 3} (no replacement) and,
// optionally, the pair reductions of awkward_hip_pair_reduce, in one launch.
// req_n, req_tooffsets (the k-th request's tuple offsets: the output of
// combinations_length for it) and req_carries (req_n[0] carry pointers,
// then req_n[1]) are HOST arrays of device pointers; program may be NULL
// (no reductions: the outputs are then ignored).
Error awkward_hip_combinations_pairs(const int64_t* offsets, int64_t n, int32_t nreq, const int32_t* req_n,
                                     const int64_t* const* req_tooffsets, int64_t* const* req_carries,
                                     const ExprProgram* program, void* out_max, void* out_min, void* out_sum,
                                     int64_t* out_argmax, int64_t* out_argmin, int8_t* out_mask) {
  if (nreq < 0 || nreq > 2) return failure("at most two combinations per launch", kSliceNone, kSliceNone, kFile);
  CombReq cr{};
  cr.nreq = nreq;
  int used = 0;
  for (int k = 0; k < nreq; k++) {
    if (req_n[k] != 2 && req_n[k] != 3)
      return failure("fused combinations take n = 2 or 3", kSliceNone, kSliceNone, kFile);
    cr.n[k] = req_n[k];
    cr.tooffsets[k] = req_tooffsets[k];
    for (int c = 0; c < req_n[k]; c++) cr.carry[k][c] = req_carries[used++];
  }
  ExprProgram p{};
  const int prog = program != nullptr;
  if (prog) {
    p = *program;
    if (!expr_program_ok(p)) return failure("malformed pair program", kSliceNone, kSliceNone, kFile);
    if (p.out_type != kTFloat32 && p.out_type != kTFloat64)
      return failure("pair reductions take float32 or float64 programs", kSliceNone, kSliceNone, kFile);
  }
  if (n <= 0 || (!prog && nreq == 0)) return success();
  const PairOut out{out_max, out_min, out_sum, out_argmax, out_argmin, out_mask};
  const PairSlots ps = prog ? pair_slots(p) : PairSlots{};
  // the program's own kernel compiled at run time (jit.h), with the
  // carries, when it builds
  const void* jit = prog ? pair_jit_kernel(p, ps.slot, ps.first, ps.nslots, out_max || out_argmax,
                                           out_min || out_argmin, out_sum != nullptr, nreq > 0)
                         : nullptr;
  if (jit != nullptr) {
    pair_jit_run(jit, p, ps.nslots, offsets, n, out_max, out_min, out_sum, out_argmax, out_argmin, out_mask, nreq,
                 req_n, req_tooffsets, req_carries);
    AKHIP_LAUNCH_CHECK("akhip_pair_jit");
    return success();
  }
  if (prog && p.out_type == kTFloat64) launch_pairs<double>(p, ps, prog, offsets, n, out, cr);
  else launch_pairs<float>(p, ps, prog, offsets, n, out, cr);
  AKHIP_LAUNCH_CHECK("pairs_kernel");
  return success();
}

}  // extern "C"
