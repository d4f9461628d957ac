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
// 256 events stages the fields of its particles in LDS (one coalesced pass),
// then each thread walks its event twice: once to count the selected
// particles (the slice is regularised against that count, as
// awkward_regularize_rangeslice does) and once to reduce the ones the slice
// keeps, in particle order -- the same serial order as the per-kernel path,
// so sums are bit-identical.
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

// One term against element i of its field (float32 fields compare in
// float32 against the constant rounded to float32, as numpy's weak-scalar
// rule and torch's scalar comparisons do).
__device__ inline bool eval_term(const Term& t, const Fields& f, const float* s_f32, int64_t i, int64_t local,
                                 bool staged) {
  const int type = f.type[t.field];
  if (type == F32) {
    float x = staged ? s_f32[local] : ((const float*)f.ptr[t.field])[i];
    if (t.absolute) x = fabsf(x);
    return compare<float>(x, t.op, (float)t.value);
  }
  double x = load_as_double(f.ptr[t.field], type, i);
  if (t.absolute) x = fabs(x);
  return compare<double>(x, t.op, t.value);
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

// Staging plan: slot[k] is field k's row in the LDS stage (-1: read from
// global memory); the stage holds nslots rows of cap floats.
struct Stage {
  int slot[kMaxFields];
  int nslots;
  int cap;
};

__device__ inline bool keep_particle(const Predicate& pred, const Fields& fields, const Stage& st,
                                     const float* s_stage, bool staged, int64_t i, int64_t local) {
  for (int t = 0; t < pred.nterms; t++) {
    const Term& term = pred.term[t];
    const int slot = st.slot[term.field];
    bool ok;
    if (staged && slot >= 0) {
      float x = s_stage[slot * st.cap + local];
      if (term.absolute) x = fabsf(x);
      ok = compare<float>(x, term.op, (float)term.value);
    } else {
      ok = eval_term(term, fields, nullptr, i, local, false);
    }
    if (!ok) return false;
  }
  return true;
}

// BLOCK threads, one event each; the workgroup's particles are staged with
// coalesced loads into a stage sized at launch (dynamic LDS) to keep many
// workgroups -- and many loads -- in flight per CU.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void event_pipeline_kernel(const int64_t* __restrict__ offsets, int64_t n,
                                                               Fields fields, Predicate pred, Stage st,
                                                               int64_t start, int64_t stop, int64_t step, int red,
                                                               PipelineOut out) {
  extern __shared__ float s_stage[];
  __shared__ int64_t s_off[BLOCK + 1];
  const int tid = threadIdx.x;
  const int64_t L0 = (int64_t)blockIdx.x * BLOCK;
  const int nl = (int)((n - L0) < BLOCK ? (n - L0) : BLOCK);
  for (int t = tid; t <= nl; t += BLOCK) s_off[t] = offsets[L0 + t];
  __syncthreads();
  const int64_t E0 = s_off[0];
  const int64_t span = s_off[nl] - E0;
  const bool staged = span <= st.cap;
  if (staged) {
    for (int k = 0; k < fields.nfields; k++) {
      const int slot = st.slot[k];
      if (slot < 0) continue;
      const float* src = (const float*)fields.ptr[k];
      float* dst = s_stage + slot * st.cap;
      for (int64_t b = 0; b < span; b += (int64_t)BLOCK * 8) {
        float r[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int64_t q = b + (int64_t)u * BLOCK + tid;
          if (q < span) r[u] = src[E0 + q];
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int64_t q = b + (int64_t)u * BLOCK + tid;
          if (q < span) dst[q] = r[u];
        }
      }
    }
  }
  __syncthreads();
  if (tid >= nl) return;
  const int64_t ev = L0 + tid;
  const int64_t a = s_off[tid], b = s_off[tid + 1];
  const int red_slot = st.slot[red];
  // pass 1: number of selected particles
  int64_t m = 0;
  for (int64_t i = a; i < b; i++) m += keep_particle(pred, fields, st, s_stage, staged, i, i - E0) ? 1 : 0;
  int64_t first;
  const int64_t cnt = slice_count(m, start, stop, step, &first);
  // pass 2: reduce the kept particles (selected positions first + t*step)
  float sum = 0.0f, best_val = 0.0f, maxv = -INFINITY;
  int64_t best = -1, t = 0, sel = 0;
  int64_t next = first;
  for (int64_t i = a; i < b && t < cnt; i++) {
    if (!keep_particle(pred, fields, st, s_stage, staged, i, i - E0)) continue;
    if (sel == next) {
      const float v = (staged && red_slot >= 0) ? s_stage[red_slot * st.cap + (i - E0)]
                                                : ((const float*)fields.ptr[red])[i];
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

// ---- pairs ---------------------------------------------------------------------
__device__ inline float pair_expr(float pt1, float eta1, float phi1, float pt2, float eta2, float phi2) {
  // 2 * pt1 * pt2 * (cosh(eta1 - eta2) - cos(phi1 - phi2)), left to right
  const float a = 2.0f * pt1;
  const float b = a * pt2;
  const float c = coshf(eta1 - eta2);
  const float d = cosf(phi1 - phi2);
  return b * (c - d);
}

constexpr int kBlock = 256;
constexpr int kPairStage = 2048;

__global__ __launch_bounds__(kBlock) void pair_max_kernel(const int64_t* __restrict__ offsets, int64_t n,
                                                          const float* __restrict__ pt, const float* __restrict__ eta,
                                                          const float* __restrict__ phi, float* __restrict__ outmax,
                                                          int64_t* __restrict__ outarg, int8_t* __restrict__ outmask,
                                                          float identity) {
  __shared__ float s_pt[kPairStage], s_eta[kPairStage], s_phi[kPairStage];
  __shared__ int64_t s_off[kBlock + 1];
  const int tid = threadIdx.x;
  const int64_t L0 = (int64_t)blockIdx.x * kBlock;
  const int nl = (int)((n - L0) < kBlock ? (n - L0) : kBlock);
  for (int t = tid; t <= nl; t += kBlock) s_off[t] = offsets[L0 + t];
  __syncthreads();
  const int64_t E0 = s_off[0];
  const int64_t span = s_off[nl] - E0;
  const bool staged = span <= kPairStage;
  if (staged) {
    for (int64_t q = tid; q < span; q += kBlock) {
      s_pt[q] = pt[E0 + q];
      s_eta[q] = eta[E0 + q];
      s_phi[q] = phi[E0 + q];
    }
  }
  __syncthreads();
  if (tid >= nl) return;
  const int64_t ev = L0 + tid;
  const int64_t a = s_off[tid], b = s_off[tid + 1];
  // LDS is indexed with in-range local positions only (a pointer formed
  // below an LDS array wraps in the 32-bit LDS address space)
  const int64_t base = staged ? E0 : 0;
  const float* P = staged ? s_pt : pt;
  const float* H = staged ? s_eta : eta;
  const float* F = staged ? s_phi : phi;
  // ak.max: best = identity; best = v > best ? v : best   (awkward_reduce_max.cpp)
  // ak.argmax: first pair seeds, strict improvement only  (awkward_reduce_argmax.cpp)
  float best = identity, argval = 0.0f;
  int64_t arg = -1, r = 0;
  for (int64_t i = a - base; i < b - base; i++) {
    const float pi = P[i], ei = H[i], fi = F[i];
    for (int64_t j = i + 1; j < b - base; j++, r++) {
      const float v = pair_expr(pi, ei, fi, P[j], H[j], F[j]);
      best = (v > best) ? v : best;
      if (r == 0) {
        arg = 0;
        argval = v;
      } else if (v > argval) {
        argval = v;
        arg = r;
      }
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
  constexpr int B = 128;
  const double mean = n_particles > 0 ? (double)n_particles / (double)n : 0.0;
  int64_t cap = (int64_t)(B * mean * 1.5) + 256;
  cap = (cap + 63) / 64 * 64;
  if (cap > 8192) cap = 8192;
  st.cap = (int)cap;
  const size_t lds = (size_t)st.nslots * st.cap * sizeof(float);
  event_pipeline_kernel<B><<<(unsigned)ceil_div(n, B), B, lds, stream()>>>(offsets, n, f, p, st, start, stop, step,
                                                                          red, out);
  AKHIP_LAUNCH_CHECK("event_pipeline_kernel");
  return success();
}

Error awkward_hip_pair_max(const int64_t* offsets, int64_t n, const float* pt, const float* eta, const float* phi,
                           float* out_max, int64_t* out_argmax, int8_t* out_mask, float identity) {
  if (n <= 0) return success();
  pair_max_kernel<<<(unsigned)ceil_div(n, kBlock), kBlock, 0, stream()>>>(offsets, n, pt, eta, phi, out_max,
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
