// fused.hip -- fused forms of multi-kernel op sequences (no counterpart in
// awkward-cpp; each gives the same bits as the sequence of registry kernels
// it replaces, and none synchronises with the host, so a whole pipeline can
// be captured in one hipGraph).
//
// awkward_hip_event_pipeline: for every event (list) of records
//     sel = ev[(predicate over fields)][:, start:stop:step]
//     num(sel), sum(sel.f), argmax(sel.f) (local index + empty mask),
//         max(sel.f) (+ mask)
// which the per-kernel path computes with the mask ufuncs,
// awkward_hip_bool_to_index, RegularArray_getitem_jagged_expand,
// ListArray_getitem_jagged_carrylen/_apply, ListArray_getitem_next_range
// _carrylength/_next_range, two gathers, awkward_reduce_sum and
// awkward_reduce_argmax + adjust_starts + reduce_mask.  Here a workgroup of
// events loads the float32 fields of its particles in one coalesced pass,
// evaluates the predicate on the registers they land in and keeps only the
// reduced value and its flag per particle in LDS (one 8-byte word); then each thread walks its
// event: once to count the selected particles (the slice is regularised
// against that count, as awkward_regularize_rangeslice does) and once to
// reduce the ones the slice keeps, in particle order -- the same serial
// order as the per-kernel path, so sums are bit-identical.  A [start:] slice
// with step 1 needs no count (the kept particles are the selected ones of
// rank >= start): one walk.
//
// awkward_hip_pair_max: for every event, the maximum and first argmax over
// its particle pairs (i < j, lexicographic, as awkward_ListArray_combinations
// orders them) of 2*pt_i*pt_j*(cosh(eta_i - eta_j) - cos(phi_i - phi_j)),
// evaluated in float32 with the same operation order as the elementwise
// form; the pairs are never materialised.

#include <cmath>

#include "common.h"

namespace akhip {
namespace {

constexpr int kMaxFields = 4;
constexpr int kMaxTerms = 8;

enum FieldType : int { F32 = 0, F64 = 1, I8 = 2, I32 = 3, I64 = 4 };
enum CmpOp : int { GT = 0, GE = 1, LT = 2, LE = 3, EQ = 4, NE = 5 };

struct Term {
  int field;
  int op;
  int absolute;  // compare |x|
  int pad;
  double value;  // compared in the field's type for floats, as double for ints
};

struct Fields {
  const void* ptr[kMaxFields];
  int type[kMaxFields];
  int nfields;
};

struct Predicate {
  Term term[kMaxTerms];
  int nterms;
};

__device__ inline double load_as_double(const void* p, int type, int64_t i) {
  switch (type) {
    case F32: return (double)((const float*)p)[i];
    case F64: return ((const double*)p)[i];
    case I8: return (double)((const int8_t*)p)[i];
    case I32: return (double)((const int32_t*)p)[i];
    default: return (double)((const int64_t*)p)[i];
  }
}

template <typename T>
__device__ inline bool compare(T x, int op, T c) {
  switch (op) {
    case GT: return x > c;
    case GE: return x >= c;
    case LT: return x < c;
    case LE: return x <= c;
    case EQ: return x == c;
    default: return x != c;
  }
}

// awkward_regularize_rangeslice (kernel-utils.cpp:7-39) -> (first, count)
__device__ inline int64_t slice_count(int64_t length, int64_t start, int64_t stop, int64_t step, int64_t* first) {
  const bool hasstart = start != kSliceNone, hasstop = stop != kSliceNone;
  int64_t a = start, b = stop;
  if (step > 0) {
    if (!hasstart) a = 0;
    else if (a < 0) a += length;
    if (a < 0) a = 0;
    if (a > length) a = length;
    if (!hasstop) b = length;
    else if (b < 0) b += length;
    if (b < 0) b = 0;
    if (b > length) b = length;
    if (b < a) b = a;
    *first = a;
    if (step == 1) return b - a;
    return (b > a) ? 1 + (b - a - 1) / step : 0;
  }
  if (!hasstart) a = length - 1;
  else if (a < 0) a += length;
  if (a < -1) a = -1;
  if (a > length - 1) a = length - 1;
  if (!hasstop) b = -1;
  else if (b < 0) b += length;
  if (b < -1) b = -1;
  if (b > length - 1) b = length - 1;
  if (b > a) b = a;
  *first = a;
  if (step == 0) return 0;
  return (a > b) ? 1 + (a - b - 1) / (-step) : 0;
}

struct PipelineOut {
  int64_t* num;
  float* sum;       // float32 reduce field only
  int64_t* argmax;  // local index in the sliced list, -1 if empty
  int8_t* argmask;  // 1 where the sliced list is empty
  float* max;
  int8_t* maxmask;
};

// Staging plan: slot[k] >= 0 for the float32 fields loaded by the staging
// pass (-1: read from global memory where a term needs it); the LDS stage
// holds cap (reduced value, predicate flag) pairs of 8 bytes.
struct Stage {
  int slot[kMaxFields];
  int nslots;
  int cap;
};

// A term resolved once per thread before the particle loops: its field
// (loaded by the staging pass, or a global column), type, operator and
// constant live in registers, so the loops never re-read kernel arguments.
struct TermRT {
  const void* gptr;  // global column
  int type;
  int op;
  int absolute;
  float fval;
  double dval;
  int field;
  bool staged;  // the field is loaded into registers by the staging loop
};

__device__ inline bool eval_rt(const TermRT& t, int64_t i, int64_t local) {
  // float32 compares in float32 against the constant rounded to float32, as
  // numpy's weak-scalar rule and torch's scalar comparisons do
  if (t.type == F32) {
    float x = ((const float*)t.gptr)[i];
    if (t.absolute) x = fabsf(x);
    return compare<float>(x, t.op, t.fval);
  }
  double x = load_as_double(t.gptr, t.type, i);
  if (t.absolute) x = fabs(x);
  return compare<double>(x, t.op, t.dval);
}

// NT (the number of terms) is a template parameter so the unrolled loops
// index the resolved terms statically (registers, no local memory).
template <int NT>
__device__ inline bool keep_particle(const TermRT* rt, int64_t i, int64_t local) {
#pragma unroll
  for (int t = 0; t < NT; t++)
    if (!eval_rt(rt[t], i, local)) return false;
  return true;
}

// BLOCK threads, LPT events each (thread t owns events t + j * BLOCK, so
// output stores coalesce); the workgroup's particles are loaded with
// coalesced loads -- every float32 field's loads issued together -- and
// staged (reduced value and predicate flag, one 8-byte word per particle,
// so each particle costs one LDS read in the per-event loops) in an LDS stage sized at launch
// (dynamic LDS), to keep many loads in flight per CU.  Eight waves per SIMD
// are asked for: left to itself the compiler keeps ~106 SGPRs (the resolved
// terms and field pointers), which holds a SIMD to seven of these one-wave,
// latency-bound workgroups; the extra scalars spill to VGPR lanes instead.
template <int BLOCK, int LPT, int NT>
__global__ __launch_bounds__(BLOCK, 8) void event_pipeline_kernel(const int64_t* __restrict__ offsets, int64_t n,
                                                               Fields fields, Predicate pred, Stage st,
                                                               int64_t start, int64_t stop, int64_t step, int red,
                                                               PipelineOut out) {
  constexpr int L = BLOCK * LPT;
  constexpr int U = 4;
  extern __shared__ float s_stage[];
  // one wave, one event per lane: each lane holds its event's bounds in
  // registers (no LDS round trip for the offsets)
  constexpr bool kRegs = BLOCK == kWave && LPT == 1;
  __shared__ int64_t s_off[kRegs ? 1 : L + 1];
  const int tid = threadIdx.x;
  const int64_t L0 = (int64_t)blockIdx.x * L;
  const int nl = (int)((n - L0) < L ? (n - L0) : L);
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
      if (t <= nl) s_off[t] = r[j];
    }
    __syncthreads();
    E0 = s_off[0];
    span = s_off[nl] - E0;
  }
  const bool staged = span <= st.cap;
  TermRT rt[NT > 0 ? NT : 1];
#pragma unroll
  for (int t = 0; t < NT; t++) {
    const Term& term = pred.term[t];
    rt[t].gptr = fields.ptr[term.field];
    rt[t].type = fields.type[term.field];
    rt[t].op = term.op;
    rt[t].absolute = term.absolute;
    rt[t].fval = (float)term.value;
    rt[t].dval = term.value;
    rt[t].field = term.field;
    rt[t].staged = st.slot[term.field] >= 0;
  }
  // staged: the float32 fields' particles are loaded with coalesced loads;
  // the predicate is evaluated on the registers they land in (terms on other
  // fields read global memory), and only the reduced value and its flag
  // go to LDS, one word per particle
  uint2* s_pv = (uint2*)s_stage;  // (reduced value bits, predicate flag) per particle
  if (staged) {
    const float* src[kMaxFields];
#pragma unroll
    for (int k = 0; k < kMaxFields; k++) {
      const int slot = k < fields.nfields ? st.slot[k] : -1;
      src[k] = slot >= 0 ? (const float*)fields.ptr[k] + E0 : nullptr;
    }
    const int sp = (int)span;
    for (int b = 0; b < sp; b += BLOCK * U) {
      float r[kMaxFields][U];
#pragma unroll
      for (int k = 0; k < kMaxFields; k++)
#pragma unroll
        for (int u = 0; u < U; u++) {
          const int q = b + u * BLOCK + tid;
          r[k][u] = (src[k] && q < sp) ? src[k][q] : 0.0f;
        }
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int q = b + u * BLOCK + tid;
        if (q < sp) {
          bool keep = true;
#pragma unroll
          for (int t = 0; t < NT; t++) {
            if (!keep) break;
            if (rt[t].staged) {
              float x = 0.0f;
#pragma unroll
              for (int k = 0; k < kMaxFields; k++)
                if (k == rt[t].field) x = r[k][u];
              if (rt[t].absolute) x = fabsf(x);
              keep = compare<float>(x, rt[t].op, rt[t].fval);
            } else {
              keep = eval_rt(rt[t], E0 + q, q);
            }
          }
          const bool kept = keep;
          float v = 0.0f;
#pragma unroll
          for (int k = 0; k < kMaxFields; k++)
            if (k == red) v = r[k][u];
          s_pv[q] = make_uint2(__float_as_uint(v), kept ? 1u : 0u);
        }
      }
    }
  }
  __syncthreads();

  const float* red_g = (const float*)fields.ptr[red];
  // [start:] with step 1 (start >= 0, or omitted): the kept particles are the
  // selected ones of rank >= start, known without the count -- one pass
  const int64_t start0 = start == kSliceNone ? 0 : start;
  const bool one_pass = staged && step == 1 && stop == kSliceNone && start0 >= 0;
#pragma unroll
  for (int j = 0; j < LPT; j++) {
    const int e = tid + j * BLOCK;
    if (e >= nl) break;
    const int64_t ev = L0 + e;
    int64_t a, b;
    if constexpr (kRegs) {
      a = own_a;
      b = own_b;
    } else {
      a = s_off[e];
      b = s_off[e + 1];
    }
    if (one_pass) {
      float sum = 0.0f, best_val = 0.0f, maxv = -INFINITY;
      int64_t best = -1, t = 0, sel = 0;
      for (int q = (int)(a - E0); q < (int)(b - E0); q++) {
        const uint2 w = s_pv[q];
        if (!w.y) continue;
        if (sel >= start0) {
          const float v = __uint_as_float(w.x);
          sum += v;
          if (t == 0) {
            best = 0;
            best_val = v;
          } else if (v > best_val) {
            best_val = v;
            best = t;
          }
          maxv = (v > maxv) ? v : maxv;
          t++;
        }
        sel++;
      }
      if (out.num) out.num[ev] = t;
      if (out.sum) out.sum[ev] = sum;
      if (out.argmax) out.argmax[ev] = best;
      if (out.argmask) out.argmask[ev] = t > 0 ? 0 : 1;
      if (out.max) out.max[ev] = maxv;
      if (out.maxmask) out.maxmask[ev] = t > 0 ? 0 : 1;
      continue;
    }
    // pass 1: number of selected particles
    int64_t m = 0;
    if (staged) {
      for (int q = (int)(a - E0); q < (int)(b - E0); q++) m += s_pv[q].y;
    } else {
      for (int64_t i = a; i < b; i++) m += keep_particle<NT>(rt, i, i - E0) ? 1 : 0;
    }
    int64_t first;
    const int64_t cnt = slice_count(m, start, stop, step, &first);
    // pass 2: reduce the kept particles (selected positions first + t*step)
    float sum = 0.0f, best_val = 0.0f, maxv = -INFINITY;
    int64_t best = -1, t = 0, sel = 0;
    int64_t next = first;
    for (int64_t i = a; i < b && t < cnt; i++) {
      const bool keep = staged ? s_pv[i - E0].y != 0 : keep_particle<NT>(rt, i, i - E0);
      if (!keep) continue;
      if (sel == next) {
        const float v = staged ? __uint_as_float(s_pv[i - E0].x) : red_g[i];
        sum += v;
        if (t == 0) {
          best = 0;
          best_val = v;
        } else if (v > best_val) {
          best_val = v;
          best = t;
        }
        maxv = (v > maxv) ? v : maxv;
        t++;
        next += step;
      }
      sel++;
    }
    if (out.num) out.num[ev] = cnt;
    if (out.sum) out.sum[ev] = sum;
    if (out.argmax) out.argmax[ev] = best;
    if (out.argmask) out.argmask[ev] = cnt > 0 ? 0 : 1;
    if (out.max) out.max[ev] = maxv;
    if (out.maxmask) out.maxmask[ev] = cnt > 0 ? 0 : 1;
  }
}

// ---- the common case, streamlined ------------------------------------------------
// Every term and the reduced field on float32 columns that the staging pass
// loads (NS of them), one event per lane of a one-wave workgroup: the
// generic kernel above carries code for terms on other columns and types and
// dispatches each term's operator by branches, which leaves it issue-bound
// (about 370 VALU instructions per 64-event wave).  Here a term is
// evaluated straight-line from uniform values -- y = x, or |x| through a
// sign mask, then (y > c & gt) | (y < c & lt) | (y == c & eq), inverted for
// != -- which is each IEEE comparison, NaN included (every one of y > c,
// y < c, y == c is false for a NaN, so >, >=, <, <=, == are false and != is
// true); the staging loads are unguarded (indices clamped into the span), and
// the per-event walk counts in 32 bits.  The same bits as
// event_pipeline_kernel.
struct FastTerm {
  int slot;          // staged column the term reads
  uint32_t absmask;  // 0x7fffffff: |x|; 0xffffffff: x
  float c;
  int gt, lt, eq, inv;  // 0 or 1
};

struct FastCut {
  const float* col[kMaxFields];  // staged columns by slot
  FastTerm term[kMaxTerms];
  int red;  // slot of the reduced column
};

template <int NS>
__device__ inline float fast_pick(const float (&x)[NS], int slot) {
  float v = x[0];
#pragma unroll
  for (int s = 1; s < NS; s++) v = slot == s ? x[s] : v;
  return v;
}

template <int NS, int NT>
__device__ inline bool fast_keep(const FastCut& cut, const float (&x)[NS]) {
  bool k = true;
#pragma unroll
  for (int t = 0; t < NT; t++) {
    const FastTerm& ft = cut.term[t];
    const float y = __uint_as_float(__float_as_uint(fast_pick<NS>(x, ft.slot)) & ft.absmask);
    const bool r = ((y > ft.c) & (ft.gt != 0)) | ((y < ft.c) & (ft.lt != 0)) | ((y == ft.c) & (ft.eq != 0));
    k = k & (r != (ft.inv != 0));
  }
  return k;
}

template <int NS, int NT>
__global__ __launch_bounds__(kWave, 8) void event_pipeline_fast_kernel(const int64_t* __restrict__ offsets,
                                                                       int64_t n, FastCut cut, int cap,
                                                                       int64_t start, int64_t stop, int64_t step,
                                                                       PipelineOut out) {
  constexpr int U = 4;
  extern __shared__ float s_stage[];
  uint2* s_pv = (uint2*)s_stage;  // (reduced value bits, predicate flag) per particle
  const int lane = threadIdx.x;
  const int64_t L0 = (int64_t)blockIdx.x * kWave;
  const int nl = (int)((n - L0) < kWave ? (n - L0) : kWave);
  int64_t own_a = 0, own_b = 0;
  if (lane < nl) {
    own_a = offsets[L0 + lane];
    own_b = offsets[L0 + lane + 1];
  }
  // the tile's first particle and span are wave-uniform: scalar registers
  auto uniform = [](int64_t v) -> int64_t {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
  };
  const int64_t E0 = uniform(__shfl(own_a, 0, kWave));
  const int64_t span = uniform(__shfl(own_b, nl - 1, kWave)) - E0;
  const bool staged = span <= cap;
  if (staged && span > 0) {
    const int sp = (int)span;
    const float* base[NS];  // uniform: scalar base + 32-bit lane offset loads
#pragma unroll
    for (int s = 0; s < NS; s++) base[s] = cut.col[s] + E0;
    for (int b = 0; b < sp; b += kWave * U) {
      float r[NS][U];
#pragma unroll
      for (int s = 0; s < NS; s++)
#pragma unroll
        for (int u = 0; u < U; u++) {
          const int q = b + u * kWave + lane;
          // a 32-bit byte offset from the uniform base: one scalar-base load
          const uint32_t off4 = (uint32_t)(q < sp ? q : sp - 1) * 4u;
          r[s][u] = *(const float*)((const char*)base[s] + off4);
        }
      // every load issued before any is used (the compiler would otherwise
      // sink each slot's loads into the branch that stores it)
#pragma unroll
      for (int s = 0; s < NS; s++)
#pragma unroll
        for (int u = 0; u < U; u++) asm volatile("" ::"v"(r[s][u]));
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int q = b + u * kWave + lane;
        float x[NS];
#pragma unroll
        for (int s = 0; s < NS; s++) x[s] = r[s][u];
        const bool k = fast_keep<NS, NT>(cut, x);
        if (q < sp) s_pv[q] = make_uint2(__float_as_uint(fast_pick<NS>(x, cut.red)), k ? 1u : 0u);
      }
    }
  }
  __syncthreads();
  if (lane >= nl) return;
  const int64_t start0 = start == kSliceNone ? 0 : start;
  float sum = 0.0f, best_val = 0.0f, maxv = -INFINITY;
  int best = -1, t = 0;
  int64_t cnt = 0;
  // The lane's event over positions [lo, hi): get(q, v, k) gives particle
  // q's value and flag (staged: q is the 32-bit position in the stage).
  auto walk = [&](auto lo, auto hi, auto get) {
    using Q = decltype(lo);
    if (step == 1 && stop == kSliceNone && start0 >= 0) {
      // [start:]: the kept particles are the selected ones of rank >= start
      // branch-free: each step selects (a conditional add is a select of
      // sum + v, so the sums' bits are the serial loop's)
      const int first = start0 < INT32_MAX ? (int)start0 : INT32_MAX;
      int sel = 0;
      for (Q q = lo; q < hi; q++) {
        float v;
        bool k;
        get(q, v, k);
        const bool take = k & (sel >= first);
        const bool better = take & ((t == 0) | (v > best_val));
        sum = take ? sum + v : sum;
        best = better ? t : best;
        best_val = better ? v : best_val;
        maxv = (take & (v > maxv)) ? v : maxv;
        t += take ? 1 : 0;
        sel += k ? 1 : 0;
      }
      cnt = t;
      return;
    }
    // pass 1: the number of selected particles; the slice regularised against it
    int m = 0;
    for (Q q = lo; q < hi; q++) {
      float v;
      bool k;
      get(q, v, k);
      m += k ? 1 : 0;
    }
    int64_t first;
    cnt = slice_count(m, start, stop, step, &first);
    // pass 2: the kept particles (selected positions first + t * step)
    int64_t next = first, sel = 0;
    for (Q q = lo; q < hi && t < cnt; q++) {
      float v;
      bool k;
      get(q, v, k);
      const bool take = k & (sel == next);
      const bool better = take & ((t == 0) | (v > best_val));
      sum = take ? sum + v : sum;
      best = better ? t : best;
      best_val = better ? v : best_val;
      maxv = (take & (v > maxv)) ? v : maxv;
      t += take ? 1 : 0;
      next += take ? step : 0;
      sel += k ? 1 : 0;
    }
  };
  if (staged) {
    walk((int)(own_a - E0), (int)(own_b - E0), [&](int q, float& v, bool& k) {
      const uint2 w = s_pv[q];
      v = __uint_as_float(w.x);
      k = w.y != 0;
    });
  } else {  // a tile past the stage: the columns from memory
    walk(own_a, own_b, [&](int64_t i, float& v, bool& k) {
      float x[NS];
#pragma unroll
      for (int s = 0; s < NS; s++) x[s] = cut.col[s][i];
      k = fast_keep<NS, NT>(cut, x);
      v = fast_pick<NS>(x, cut.red);
    });
  }
  // stores from uniform bases at the tile's first event, 32-bit lane offsets
  const uint32_t l = (uint32_t)lane;
  if (out.num) *(int64_t*)((char*)(out.num + L0) + l * 8u) = cnt;
  if (out.sum) *(float*)((char*)(out.sum + L0) + l * 4u) = sum;
  if (out.argmax) *(int64_t*)((char*)(out.argmax + L0) + l * 8u) = (int64_t)best;
  if (out.argmask) *((int8_t*)(out.argmask + L0) + l) = cnt > 0 ? 0 : 1;
  if (out.max) *(float*)((char*)(out.max + L0) + l * 4u) = maxv;
  if (out.maxmask) *((int8_t*)(out.maxmask + L0) + l) = cnt > 0 ? 0 : 1;
}

// ---- pairs ---------------------------------------------------------------------
__device__ inline float pair_expr(float pt1, float eta1, float phi1, float pt2, float eta2, float phi2) {
  // 2 * pt1 * pt2 * (cosh(eta1 - eta2) - cos(phi1 - phi2)), left to right
  const float a = 2.0f * pt1;
  const float b = a * pt2;
  const float c = coshf(eta1 - eta2);
  const float d = cosf(phi1 - phi2);
  return b * (c - d);
}

constexpr int kPairEvents = 128;  // events (and threads) per workgroup
constexpr int kPairStage = 768;   // particles staged per workgroup (~1.5x a workgroup at Poisson(4))
constexpr int kPairCap = 1536;    // pair values kept per workgroup (small tiles: more workgroups per CU)

// Workgroup of 128 events.  When the events' particles and pair values fit
// the LDS tiles, the pairs are spread evenly over the threads -- each takes
// a contiguous run, finds the first pair's event by binary search over the
// pair-count prefix and steps through the rest in combinations order -- and
// the values go to LDS; then every thread reduces its own event's values in
// order.  Otherwise each thread enumerates its event's pairs itself.
__global__ __launch_bounds__(kPairEvents) void pair_max_kernel(const int64_t* __restrict__ offsets, int64_t n,
                                                               const float* __restrict__ pt,
                                                               const float* __restrict__ eta,
                                                               const float* __restrict__ phi,
                                                               float* __restrict__ outmax,
                                                               int64_t* __restrict__ outarg,
                                                               int8_t* __restrict__ outmask, float identity) {
  __shared__ float s_pt[kPairStage], s_eta[kPairStage], s_phi[kPairStage];
  __shared__ float s_val[kPairCap];
  __shared__ int64_t s_off[kPairEvents + 1];
  __shared__ int s_pbeg[kPairEvents + 1];  // pair-count prefix (valid when staged)
  __shared__ int s_wsum[kPairEvents / kWave];
  const int tid = threadIdx.x;
  const int64_t L0 = (int64_t)blockIdx.x * kPairEvents;
  const int nl = (int)((n - L0) < kPairEvents ? (n - L0) : kPairEvents);
  for (int t = tid; t <= nl; t += kPairEvents) s_off[t] = offsets[L0 + t];
  __syncthreads();
  const int64_t E0 = s_off[0];
  const int64_t span = s_off[nl] - E0;
  const bool staged = span <= kPairStage;
  if (staged) {
    for (int64_t q = tid; q < span; q += kPairEvents) {
      s_pt[q] = pt[E0 + q];
      s_eta[q] = eta[E0 + q];
      s_phi[q] = phi[E0 + q];
    }
  }
  // pairs of this thread's event, and their exclusive prefix over the
  // workgroup (staged: every event has at most kPairStage particles, so the
  // total stays below 128 * C(768, 2) < 2^31)
  const int m_own = tid < nl ? (int)(s_off[tid + 1] - s_off[tid]) : 0;
  const int c_own = staged ? m_own * (m_own - 1) / 2 : 0;
  const int incl = wave_inclusive_scan(c_own);
  const int lane = tid & (kWave - 1), wave = tid / kWave;
  if (lane == kWave - 1) s_wsum[wave] = incl;
  __syncthreads();
  int wbase = 0;
  for (int w = 0; w < wave; w++) wbase += s_wsum[w];
  int total = 0;
  for (int w = 0; w < kPairEvents / kWave; w++) total += s_wsum[w];
  s_pbeg[tid] = wbase + incl - c_own;
  if (tid == 0) s_pbeg[kPairEvents] = total;
  __syncthreads();
  const bool pairwise = staged && total <= kPairCap;
  if (pairwise && total > 0) {
    const int per = (total + kPairEvents - 1) / kPairEvents;
    int q = tid * per;
    const int qend = q + per < total ? q + per : total;
    if (q < qend) {
      // the event holding pair q: last e with s_pbeg[e] <= q
      int lo = 0, hi = nl;
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (s_pbeg[mid] <= q) lo = mid;
        else hi = mid;
      }
      int e = lo;
      while (s_pbeg[e + 1] <= q) e++;  // events without pairs share a prefix value
      int m = (int)(s_off[e + 1] - s_off[e]);
      int r = q - s_pbeg[e];
      int i = 0;
      while (r >= m - 1 - i) {
        r -= m - 1 - i;
        i++;
      }
      int j = i + 1 + r;
      int a = (int)(s_off[e] - E0);
      for (; q < qend; q++) {
        s_val[q] = pair_expr(s_pt[a + i], s_eta[a + i], s_phi[a + i], s_pt[a + j], s_eta[a + j], s_phi[a + j]);
        if (++j == m) {
          ++i;
          j = i + 1;
        }
        if (i >= m - 1 && q + 1 < qend) {  // on to the next event with pairs
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
  }
  __syncthreads();
  if (tid >= nl) return;
  const int64_t ev = L0 + tid;
  // ak.max: best = identity; best = v > best ? v : best   (awkward_reduce_max.cpp)
  // ak.argmax: first pair seeds, strict improvement only  (awkward_reduce_argmax.cpp)
  float best = identity, argval = 0.0f;
  int64_t arg = -1, r = 0;
  auto take = [&](float v) {
    best = (v > best) ? v : best;
    if (r == 0) {
      arg = 0;
      argval = v;
    } else if (v > argval) {
      argval = v;
      arg = r;
    }
    r++;
  };
  if (pairwise) {
    for (int q = s_pbeg[tid]; q < s_pbeg[tid + 1]; q++) take(s_val[q]);
  } else {
    const int64_t a = s_off[tid], b = s_off[tid + 1];
    // LDS is indexed with in-range local positions only (a pointer formed
    // below an LDS array wraps in the 32-bit LDS address space)
    const int64_t base = staged ? E0 : 0;
    const float* P = staged ? s_pt : pt;
    const float* H = staged ? s_eta : eta;
    const float* F = staged ? s_phi : phi;
    for (int64_t i = a - base; i < b - base; i++) {
      const float pi = P[i], ei = H[i], fi = F[i];
      for (int64_t j = i + 1; j < b - base; j++) take(pair_expr(pi, ei, fi, P[j], H[j], F[j]));
    }
  }
  outmax[ev] = best;
  outarg[ev] = arg;
  outmask[ev] = r > 0 ? 0 : 1;
}

__global__ void pair_expr_kernel(float* __restrict__ out, const float* __restrict__ pt1,
                                 const float* __restrict__ eta1, const float* __restrict__ phi1,
                                 const float* __restrict__ pt2, const float* __restrict__ eta2,
                                 const float* __restrict__ phi2, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = pair_expr(pt1[i], eta1[i], phi1[i], pt2[i], eta2[i], phi2[i]);
}

}  // namespace
}  // namespace akhip

using namespace akhip;

extern "C" {

// field_ptrs / field_types / term_* are HOST arrays (copied into the kernel
// arguments).  fields/types: nfields (<= 4) device pointers and type codes (0 f32, 1 f64,
// 2 i8, 3 i32, 4 i64); terms: nterms (<= 8) of (field, op, absolute, value),
// ops 0 >, 1 >=, 2 <, 3 <=, 4 ==, 5 !=; n_particles: the length of the fields
// (sizes the LDS stage; any value is correct); red: index of the float32 field that
// is reduced.  Any output pointer may be null.
Error awkward_hip_event_pipeline(const int64_t* offsets, int64_t n, int64_t n_particles, const void* const* field_ptrs,
                                 const int32_t* field_types, int32_t nfields, const int32_t* term_field,
                                 const int32_t* term_op, const int32_t* term_abs, const double* term_value,
                                 int32_t nterms, int64_t start, int64_t stop, int64_t step, int32_t red,
                                 int64_t* out_num, float* out_sum, int64_t* out_argmax, int8_t* out_argmask,
                                 float* out_max, int8_t* out_maxmask) {
  if (nfields < 1 || nfields > kMaxFields || nterms < 0 || nterms > kMaxTerms || red < 0 || red >= nfields)
    return failure("event pipeline: bad field or term count", kSliceNone, kSliceNone, "awkward_amd/csrc/fused.hip");
  if (field_types[red] != F32)
    return failure("event pipeline: the reduced field must be float32", kSliceNone, kSliceNone,
                   "awkward_amd/csrc/fused.hip");
  if (step <= 0)
    return failure("event pipeline: slice step must be positive", kSliceNone, kSliceNone,
                   "awkward_amd/csrc/fused.hip");
  if (n <= 0) return success();
  Fields f{};
  f.nfields = nfields;
  for (int k = 0; k < nfields; k++) {
    f.ptr[k] = field_ptrs[k];
    f.type[k] = field_types[k];
  }
  Predicate p{};
  p.nterms = nterms;
  for (int t = 0; t < nterms; t++) {
    if (term_field[t] < 0 || term_field[t] >= nfields)
      return failure("event pipeline: term on a missing field", kSliceNone, kSliceNone, "awkward_amd/csrc/fused.hip");
    p.term[t] = Term{term_field[t], term_op[t], term_abs[t], 0, term_value[t]};
  }
  PipelineOut out{out_num, out_sum, out_argmax, out_argmask, out_max, out_maxmask};
  // stage every float32 field the pipeline reads, sized for the workgroup's
  // expected particle span (mean multiplicity from the array shapes) with
  // room for its fluctuation; larger spans read global memory directly.
  Stage st{};
  st.nslots = 0;
  for (int k = 0; k < kMaxFields; k++) st.slot[k] = -1;
  for (int k = 0; k < nfields; k++)
    if (field_types[k] == F32) st.slot[k] = st.nslots++;
  // events per thread: at least ~4 KiB of staged particles per workgroup
  // (one event per thread at the configs' multiplicities, which measured best);
  // 64 threads: 0.83 -> 0.82 ms against 128 (256: 0.83)
  constexpr int B = 64;
  const double mean = n_particles > 0 ? (double)n_particles / (double)n : 0.0;
  const double per_event = mean * (st.nslots > 0 ? st.nslots : 1) * sizeof(float) * 1.5;
  int lpt = 1;
  while (lpt < 8 && (double)B * (2 * lpt) * per_event <= 4096.0) lpt *= 2;
  int64_t cap = (int64_t)(B * lpt * mean * 1.5) + 256;
  cap = (cap + 63) / 64 * 64;
  if (cap > 8192) cap = 8192;  // 8-byte (value, flag) pairs: at most 64 KiB of LDS
  st.cap = (int)cap;
  const size_t lds = (size_t)st.cap * 8;  // (reduced value, predicate flag) pairs
  const unsigned grid = (unsigned)ceil_div(n, (int64_t)B * lpt);
  hipStream_t s = stream();
  // the streamlined kernel: one event per lane, every term and the reduced
  // field on staged float32 columns, at most four terms
  bool fast = lpt == 1 && nterms <= 4 && st.slot[red] >= 0;
  for (int t = 0; t < nterms; t++) fast = fast && st.slot[term_field[t]] >= 0;
  if (fast) {
    FastCut cut{};
    for (int k = 0; k < nfields; k++)
      if (st.slot[k] >= 0) cut.col[st.slot[k]] = (const float*)field_ptrs[k];
    cut.red = st.slot[red];
    for (int t = 0; t < nterms; t++) {
      FastTerm& ft = cut.term[t];
      ft.slot = st.slot[term_field[t]];
      ft.absmask = term_abs[t] ? 0x7fffffffu : 0xffffffffu;
      ft.c = (float)term_value[t];
      const int op = term_op[t];
      ft.gt = op == GT || op == GE;
      ft.lt = op == LT || op == LE;
      ft.eq = op == GE || op == LE || op == EQ || op == NE;
      ft.inv = op == NE;
    }
    switch (st.nslots * 16 + nterms) {
#define FCASE(NS, K) \
  case NS * 16 + K: event_pipeline_fast_kernel<NS, K><<<grid, B, lds, s>>>(offsets, n, cut, st.cap, start, stop, step, out); break;
#define FCASES(NS) FCASE(NS, 0) FCASE(NS, 1) FCASE(NS, 2) FCASE(NS, 3) FCASE(NS, 4)
      FCASES(1)
      FCASES(2)
      FCASES(3)
      FCASES(4)
#undef FCASES
#undef FCASE
    }
    AKHIP_LAUNCH_CHECK("event_pipeline_fast_kernel");
    return success();
  }
  switch (lpt * 16 + nterms) {
#define CASE(LPT, K)                                                                                              \
  case LPT * 16 + K:                                                                                              \
    event_pipeline_kernel<B, LPT, K><<<grid, B, lds, s>>>(offsets, n, f, p, st, start, stop, step, red, out); \
    break;
#define CASES(LPT) CASE(LPT, 0) CASE(LPT, 1) CASE(LPT, 2) CASE(LPT, 3) CASE(LPT, 4) CASE(LPT, 5) CASE(LPT, 6) \
    CASE(LPT, 7) CASE(LPT, 8)
    CASES(1)
    CASES(2)
    CASES(4)
    CASES(8)
#undef CASES
#undef CASE
  }
  AKHIP_LAUNCH_CHECK("event_pipeline_kernel");
  return success();
}

Error awkward_hip_pair_max(const int64_t* offsets, int64_t n, const float* pt, const float* eta, const float* phi,
                           float* out_max, int64_t* out_argmax, int8_t* out_mask, float identity) {
  if (n <= 0) return success();
  pair_max_kernel<<<(unsigned)ceil_div(n, kPairEvents), kPairEvents, 0, stream()>>>(offsets, n, pt, eta, phi, out_max,
                                                                                    out_argmax, out_mask, identity);
  AKHIP_LAUNCH_CHECK("pair_max_kernel");
  return success();
}

// The pair expression over already-gathered legs (the elementwise form).
Error awkward_hip_pair_expr(float* out, const float* pt1, const float* eta1, const float* phi1, const float* pt2,
                            const float* eta2, const float* phi2, int64_t n) {
  if (n <= 0) return success();
  pair_expr_kernel<<<grid_for(n, 256, 8), 256, 0, stream()>>>(out, pt1, eta1, phi1, pt2, eta2, phi2, n);
  AKHIP_LAUNCH_CHECK("pair_expr_kernel");
  return success();
}

}  // extern "C"
