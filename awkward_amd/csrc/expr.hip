// expr.hip -- elementwise expressions (expr.h) as kernels of the hip nplike:
//
//   * awkward_hip_expr_eval: out[i] = program(start + i) for n elements --
//     a chain of numpy ufuncs on lists of numbers in one pass
//     (src/awkward/_nplikes/array_module.py:254-332 applies them one by one,
//     each with a temporary);
//   * awkward_hip_mask_getitem: array[mask] for a jagged boolean mask given
//     as a program (a recorded predicate, or a load of a materialised mask):
//     the program's bits (one word per 64 elements, 1/8 of a byte mask),
//     then one look-back scan over the lists.  It writes what the reference's
//     sequence writes -- _slicing._normalise_item_bool_to_int
//     (src/awkward/_slicing.py:476-530: local index, compaction, cumsum) and
//     then RegularArray_getitem_jagged_expand + ListArray_getitem_jagged_carrylen
//     + ListArray_getitem_jagged_apply (contents/listarray.py:549-618) -- the
//     output offsets and the carry into the array's content, with
//     jagged_apply's checks, messages, identity and attempt
//     (awkward-cpp/src/cpu-kernels/awkward_ListArray_getitem_jagged_apply.cpp).
//     No byte mask is written, and neither are the integer index, the
//     expanded starts/stops or the carry length.

#include <vector>

#include "errors.h"
#include "expr.h"
#include "scan.h"

namespace akhip {
namespace {

const char* kFile = "awkward_amd/csrc/expr.hip";

__global__ void fill_first_i64(int64_t* p) { *p = 0; }

constexpr int kEvalBlock = 256;

// U elements per thread per program pass (expr_slots: the register file
// holds 32 VGPRs)
template <typename W, int U>
__global__ __launch_bounds__(kEvalBlock) void expr_eval_kernel(ExprProgram p, void* __restrict__ out, int64_t start,
                                                               int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * kEvalBlock * U;
  for (int64_t base = (int64_t)blockIdx.x * kEvalBlock * U; base < n; base += stride) {
    int64_t pos[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t i = base + (int64_t)u * kEvalBlock + threadIdx.x;
      pos[u] = i < n ? start + i : -1;
    }
    W v[U];
    expr_eval<U, W>(p, pos, v);
#pragma unroll
    for (int u = 0; u < U; u++)
      if (pos[u] >= 0) x_store<W>(out, p.out_type, pos[u] - start, v[u]);
  }
}

inline void launch_eval(const ExprProgram& p, void* out, int64_t start, int64_t n) {
  expr_dispatch(p, [&](auto kind) {
    using W = typename decltype(kind)::Word;
    constexpr int U = decltype(kind)::slots;
    expr_eval_kernel<W, U><<<grid_for(n, kEvalBlock, U), kEvalBlock, 0, stream()>>>(p, out, start, n);
  });
}

// jagged_apply's messages (the same codes and order: per list, the checks
// of the lists' bounds first, then the indices)
const char* const kMaskMsgs[] = {
    nullptr,
    "jagged slice's stops[i] < starts[i]",
    "jagged slice's offsets extend beyond its content",
    "stops[i] < starts[i]",
    "stops[i] > len(content)",
    "index out of range",
};

// The predicate as bits: bit (i & 63) of word i >> 6 is program(i) for
// 0 <= i < n (the words past n are not written).  Each wave evaluates 64
// consecutive elements per slot and stores their ballot as one word.
template <typename W, int U>
__global__ __launch_bounds__(kEvalBlock) void expr_bits_kernel(ExprProgram p, uint64_t* __restrict__ bits,
                                                               int64_t n) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int64_t stride = (int64_t)gridDim.x * kEvalBlock * U;
  for (int64_t base = (int64_t)blockIdx.x * kEvalBlock * U; base < n; base += stride) {
    int64_t pos[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t i = base + (int64_t)u * kEvalBlock + threadIdx.x;
      pos[u] = i < n ? i : -1;
    }
    W v[U];
    expr_eval<U, W>(p, pos, v);
#pragma unroll
    for (int u = 0; u < U; u++) {
      const uint64_t word = __ballot(pos[u] >= 0 && v[u] != 0);
      const int64_t first = base + (int64_t)u * kEvalBlock + (int64_t)wave * kWave;
      if (lane == 0 && first < n) bits[first >> 6] = word;
    }
  }
}

// ---- selection cuts -------------------------------------------------------------
// A predicate that is a conjunction of comparisons of float32 columns (or of
// their absolute values) with constants -- a selection cut such as
// (pt > 25) & (abs(eta) < 2.4) -- needs no interpreter: cut_bits_kernel
// evaluates the same float32 comparisons (so the same bits) with each lane's
// loads of every term in flight together.
constexpr int kMaxCutTerms = 8;

struct CutTerm {
  const float* col;
  float c;
  int op;  // kXGt .. kXNe
  int absolute;
};

struct Cut {
  CutTerm t[kMaxCutTerms];
  int n;
};

// A symbolic run of the program: a register holds a float32 column (a leaf
// read without an index), its absolute value, or a conjunction of terms (a
// set of cut.t); any other op, type or operand means the program is not a cut.
inline bool match_cut(const ExprProgram& p, Cut* cut) {
  if (p.out_type != kTBool || p.nops < 1 || p.wide) return false;
  enum Kind { kNone = 0, kCol, kAbs, kTerms };
  int kind[kExprRegs] = {}, col[kExprRegs] = {};
  uint32_t set[kExprRegs] = {};
  for (int k = 0; k < p.nleaves; k++)
    if (p.leaf[k].type == kTFloat32 && p.leaf[k].index == nullptr) {
      kind[k] = kCol;
      col[k] = k;
    }
  Cut all{};
  for (int i = 0; i < p.nops; i++) {
    const ExprOp& o = p.op[i];
    int nk = kNone, nc = 0;
    uint32_t ns = 0;
    switch (o.code) {
      case kXAbs:
        if (o.type == kTFloat32 && kind[o.a] == kCol) {
          nk = kAbs;
          nc = col[o.a];
        }
        break;
      case kXGt:
      case kXGe:
      case kXLt:
      case kXLe:
      case kXEq:
      case kXNe:
        if (o.type == kTFloat32 && o.bimm && o.to == 0 && (kind[o.a] == kCol || kind[o.a] == kAbs) &&
            all.n < kMaxCutTerms) {
          CutTerm& t = all.t[all.n];
          t.col = (const float*)p.leaf[col[o.a]].data;
          const uint32_t b = (uint32_t)o.imm;
          __builtin_memcpy(&t.c, &b, sizeof(float));
          t.op = o.code;
          t.absolute = kind[o.a] == kAbs;
          nk = kTerms;
          ns = 1u << all.n;
          all.n++;
        }
        break;
      case kXAnd:
      case kXLogicalAnd:
      case kXMul:  // bool * bool is &
        if (o.type == kTBool && !o.bimm && kind[o.a] == kTerms && kind[o.b] == kTerms) {
          nk = kTerms;
          ns = set[o.a] | set[o.b];
        }
        break;
      default:
        break;
    }
    if (nk == kNone) return false;
    kind[o.dst] = nk;
    col[o.dst] = nc;
    set[o.dst] = ns;
  }
  if (kind[p.out] != kTerms) return false;
  cut->n = 0;
  for (int k = 0; k < all.n; k++)
    if (set[p.out] & (1u << k)) cut->t[cut->n++] = all.t[k];
  return cut->n > 0;
}

__device__ inline bool cut_compare(float x, int op, float c) {
  switch (op) {
    case kXGt: return x > c;
    case kXGe: return x >= c;
    case kXLt: return x < c;
    case kXLe: return x <= c;
    case kXEq: return x == c;
    default: return x != c;
  }
}

// Bits of a cut with NT terms: each wave takes groups of 64 * U consecutive
// elements, lane l elements base + 64 u + l (every load coalesced, all NT * U
// of a lane in flight at once), and stores one ballot word per 64 elements.
constexpr int kCutBlock = 256;
constexpr int kCutU = 8;

template <int NT>
__global__ __launch_bounds__(kCutBlock) void cut_bits_kernel(Cut cut, uint64_t* __restrict__ bits, int64_t n) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t nwaves = (int64_t)gridDim.x * (kCutBlock / kWave);
  const float* col[NT];
  float c[NT];
  int op[NT], ab[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) {
    col[t] = cut.t[t].col;
    c[t] = cut.t[t].c;
    op[t] = cut.t[t].op;
    ab[t] = cut.t[t].absolute;
  }
  for (int64_t g = (int64_t)blockIdx.x * (kCutBlock / kWave) + threadIdx.x / kWave; g * (kWave * kCutU) < n;
       g += nwaves) {
    const int64_t base = g * (kWave * kCutU);
    float x[NT][kCutU];
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
      for (int u = 0; u < kCutU; u++) {
        const int64_t i = base + u * kWave + lane;
        x[t][u] = i < n ? col[t][i] : 0.0f;
      }
#pragma unroll
    for (int u = 0; u < kCutU; u++) {
      const int64_t first = base + u * kWave;
      bool keep = first + lane < n;
#pragma unroll
      for (int t = 0; t < NT; t++) keep = keep && cut_compare(ab[t] ? fabsf(x[t][u]) : x[t][u], op[t], c[t]);
      const uint64_t word = __ballot(keep);
      if (lane == 0 && first < n) bits[first >> 6] = word;
    }
  }
}

inline void launch_cut_bits(const Cut& cut, uint64_t* bits, int64_t n, hipStream_t st) {
  int64_t grid = ceil_div(n, (int64_t)kCutBlock * kCutU);
  const int64_t cap = (int64_t)cu_count() * 32;
  if (grid > cap) grid = cap;
  switch (cut.n) {
#define CUT_CASE(K) \
  case K: cut_bits_kernel<K><<<(unsigned)grid, kCutBlock, 0, st>>>(cut, bits, n); break;
    CUT_CASE(1)
    CUT_CASE(2)
    CUT_CASE(3)
    CUT_CASE(4)
    CUT_CASE(5)
    CUT_CASE(6)
    CUT_CASE(7)
    default: cut_bits_kernel<8><<<(unsigned)grid, kCutBlock, 0, st>>>(cut, bits, n); break;
#undef CUT_CASE
  }
}

__device__ inline bool bit_at(const uint64_t* bits, int64_t j) { return (bits[j >> 6] >> (j & 63)) & 1; }

// true bits in [a, b)
__device__ inline int64_t count_bits(const uint64_t* bits, int64_t a, int64_t b) {
  if (a >= b) return 0;
  int64_t c = 0;
  const int64_t w0 = a >> 6, w1 = (b - 1) >> 6;
  for (int64_t w = w0; w <= w1; w++) {
    uint64_t x = bits[w];
    if (w == w0) x &= ~uint64_t(0) << (a & 63);
    if (w == w1 && ((b & 63) != 0)) x &= ~uint64_t(0) >> (64 - (b & 63));
    c += __popcll(x);
  }
  return c;
}

// array[mask] from the mask's bits: a look-back scan over the lists (tiles
// of 4096; the per-list counts, then their exclusive prefixes, in LDS).
// count(list) = true bits of the mask list; the emit writes tooffsets and the
// carry start + local position of each true bit, with jagged_apply's checks.
constexpr int kMGBlock = 512;
constexpr int kMGIpt = 8;
constexpr int kMGTile = kMGBlock * kMGIpt;

__global__ __launch_bounds__(kMGBlock) void mask_bits_getitem_kernel(
    const uint64_t* __restrict__ bits, int64_t* __restrict__ tooffsets, int64_t* __restrict__ tocarry,
    const int64_t* __restrict__ fromstarts, const int64_t* __restrict__ fromstops, int64_t contentlen,
    const int64_t* __restrict__ moffsets, int64_t n, int64_t elements, ScanScratch s, uint64_t* errword,
    uint32_t inv) {
  constexpr int NW = kMGBlock / kWave;
  __shared__ int64_t s_v[kMGTile + kMGTile / 32 + 1];  // counts, then exclusive prefixes (+ the tile's end)
  __shared__ int64_t s_wave[NW];
  __shared__ int64_t s_prefix;
  __shared__ unsigned s_tile;
  auto pad = [](int q) { return q + (q >> 5); };
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wave = tid / kWave;

  if (tid == 0) s_tile = atomicAdd(s.counter, 1u);
  __syncthreads();
  const int64_t tile = s_tile;
  const int64_t L0 = tile * (int64_t)kMGTile;
  // a list's element range, clamped to the bits (offsets that are not
  // ascending select nothing beyond them)
  auto range = [&](int64_t i, int64_t& a, int64_t& b) {
    a = moffsets[i];
    b = moffsets[i + 1];
    a = a < 0 ? 0 : (a > elements ? elements : a);
    b = b < a ? a : (b > elements ? elements : b);
  };
#pragma unroll 4
  for (int u = 0; u < kMGIpt; u++) {
    const int t = u * kMGBlock + tid;
    const int64_t i = L0 + t;
    int64_t a = 0, b = 0;
    if (i < n) range(i, a, b);
    s_v[pad(t)] = count_bits(bits, a, b);
  }
  __syncthreads();
  int64_t total = 0;
#pragma unroll
  for (int v = 0; v < kMGIpt; v++) total += s_v[pad(tid * kMGIpt + v)];
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
  for (int v = 0; v < kMGIpt; v++) {
    const int q = pad(tid * kMGIpt + v);
    const int64_t c = s_v[q];
    s_v[q] = run;
    run += c;
  }
  if (tid == kMGBlock - 1) s_v[pad(kMGTile)] = run;
  __syncthreads();
  for (int u = 0; u < kMGIpt; u++) {
    const int t = u * kMGBlock + tid;
    const int64_t i = L0 + t;
    if (i >= n) continue;
    const int64_t excl = s_v[pad(t)];
    const int64_t cnt = s_v[pad(t + 1)] - excl;
    if (i == 0) tooffsets[0] = 0;
    tooffsets[i + 1] = excl + cnt;
    if (cnt == 0) continue;  // an empty slice list checks nothing (jagged_apply)
    const int64_t start = fromstarts[i], stop = fromstops[i];
    if (stop < start) {
      record_error(errword, inv, i, 3);
      continue;
    }
    if (start != stop && stop > contentlen) {
      record_error(errword, inv, i, 4);
      continue;
    }
    const int64_t count = stop - start;
    int64_t a, b;
    range(i, a, b);
    int64_t k = excl;
    for (int64_t w = a >> 6; w <= (b - 1) >> 6; w++) {
      uint64_t x = bits[w];
      if (w == (a >> 6)) x &= ~uint64_t(0) << (a & 63);
      if (w == ((b - 1) >> 6) && ((b & 63) != 0)) x &= ~uint64_t(0) >> (64 - (b & 63));
      bool bad = false;
      while (x) {
        const int64_t j = (w << 6) + __builtin_ctzll(x);
        x &= x - 1;
        if (j - a >= count) {
          record_error(errword, inv, i, 5);
          bad = true;
          break;
        }
        tocarry[k++] = start + (j - a);
      }
      if (bad) break;
    }
  }
}

}  // namespace
}  // namespace akhip

using namespace akhip;

extern "C" {

// out[i] = program(start + i) for 0 <= i < n; out has the program's out_type.
Error awkward_hip_expr_eval(void* out, const ExprProgram* program, int64_t start, int64_t n) {
  if (n <= 0) return success();
  const ExprProgram p = *program;
  if (!expr_program_ok(p)) return failure("malformed expression program", kSliceNone, kSliceNone, kFile);
  launch_eval(p, out, start, n);
  AKHIP_LAUNCH_CHECK("expr_eval_kernel");
  return success();
}

// The number of comparison terms of a program that is a selection cut, else 0
// (host only).
int awkward_hip_expr_cut_terms(const ExprProgram* program) {
  Cut cut;
  return (expr_program_ok(*program) && match_cut(*program, &cut)) ? cut.n : 0;
}

// array[mask] for a jagged boolean mask with int64 offsets `moffsets` (n
// lists, absolute positions into the program's masklen elements): tooffsets (n + 1,
// from 0) and tocarry (as many entries as the mask has true values; the
// caller allocates an upper bound, e.g. the mask's element count) are the
// output offsets and the carry into the indexed array's content, whose
// lists are [fromstarts[i], fromstops[i]) of a content of contentlen
// elements.  The number of carried elements is tooffsets[n].
Error awkward_hip_mask_getitem(int64_t* tooffsets, int64_t* tocarry, const int64_t* fromstarts,
                               const int64_t* fromstops, int64_t contentlen, const int64_t* moffsets, int64_t n,
                               int64_t masklen, const ExprProgram* program) {
  hipStream_t st = stream();
  if (n <= 0) {
    fill_first_i64<<<1, 1, 0, st>>>(tooffsets);
    AKHIP_LAUNCH_CHECK("fill_first_i64");
    return success();
  }
  const ExprProgram p = *program;
  if (!expr_program_ok(p) || p.out_type != kTBool)
    return failure("malformed mask program", kSliceNone, kSliceNone, kFile);
  // the bits cover the mask's elements [0, masklen) of the leaves
  const int64_t elements = masklen > 0 ? masklen : 0;
  const int64_t ntiles = ceil_div(n, (int64_t)kMGTile);
  const size_t bits_bytes = ((size_t)ceil_div(elements > 0 ? elements : 1, 64) * 8 + 255) / 256 * 256;
  char* base = (char*)scratch(bits_bytes + scan_scratch_bytes(ntiles));
  if (!base) return failure("out of device memory for the mask's bits", kSliceNone, kSliceNone, kFile);
  uint64_t* bits = (uint64_t*)base;
  AKHIP_ERROR_WORD(ew);
  const uint32_t inv = next_invocation("awkward_ListArray_getitem_jagged_apply", kMaskMsgs);
  const ScanScratch s = scan_scratch_at(base + bits_bytes);
  Cut cut;
  if (elements > 0 && match_cut(p, &cut)) {
    launch_cut_bits(cut, bits, elements, st);
  } else if (elements > 0) {
    expr_dispatch(p, [&](auto kind) {
      using W = typename decltype(kind)::Word;
      constexpr int U = decltype(kind)::slots;
      expr_bits_kernel<W, U><<<grid_for(elements, kEvalBlock, U), kEvalBlock, 0, st>>>(p, bits, elements);
    });
  }
  scan_prep_kernel<0><<<(unsigned)ceil_div(ntiles + 1, 256), 256, 0, st>>>(s, ntiles);
  mask_bits_getitem_kernel<<<(unsigned)ntiles, kMGBlock, 0, st>>>(bits, tooffsets, tocarry, fromstarts, fromstops,
                                                                  contentlen, moffsets, n, elements, s, ew, inv);
  AKHIP_LAUNCH_CHECK("mask_bits_getitem_kernel");
  return finish_errors(kMaskMsgs, kFile, [&](int64_t i, int code) -> int64_t {
    if (code != 5) return kSliceNone;
    // the first true position of list i past the indexed list's length
    const int64_t a = read_device(moffsets, i), e = read_device(moffsets, i + 1);
    const int64_t count = read_device(fromstops, i) - read_device(fromstarts, i);
    if (e <= a) return kSliceNone;
    void* flags = scratch((size_t)(e - a));
    if (!flags) return kSliceNone;
    launch_eval(p, flags, a, e - a);
    std::vector<uint8_t> h((size_t)(e - a));
    (void)hipMemcpyAsync(h.data(), flags, h.size(), hipMemcpyDeviceToHost, stream());
    (void)hipStreamSynchronize(stream());
    for (int64_t j = 0; j < e - a; j++)
      if (h[(size_t)j] && j >= count) return j;
    return kSliceNone;
  });
}

}  // extern "C"
