// expr.h -- elementwise expressions evaluated on the device from a small
// op list (the hip nplike's ufuncs, src/awkward/_nplikes/array_module.py:254-332).
//
// The host layer records a chain of numpy ufuncs on lists of numbers
// (comparisons, arithmetic, logic, abs, cosh, cos ...) as one ExprProgram:
// leaves (columns, optionally gathered through an int64 index: a lazy carry)
// and typed ops over registers.  The host resolves every operand's type with
// numpy's own promotion rules (NEP 50 for Python scalars) and inserts the
// casts, so each op here computes in exactly one type and the result is what
// numpy's loop for that type gives:
//
//   * integers wrap to their width (two's complement), like numpy's loops;
//   * float32 ops stay in float32 (no widening), float64 in float64;
//   * comparisons of floats are IEEE (NaN compares false, except !=);
//   * cosh/cos/sinh/sin/exp/sqrt call the device math library.
//
// A register holds a value's canonical bits: bool 0/1, signed integers
// sign-extended, unsigned zero-extended, float32 in the low 32 bits, float64
// as its bits.  Registers are 32 bits wide when every type of the program
// fits in 32 bits (the host sets `wide` = 0: half the VGPRs, no float64
// code paths, more waves in flight) and 64 bits otherwise.
//
// The program is uniform across the grid (a kernel argument), so the op
// dispatch is scalar control flow; each op is dispatched once for all U
// element slots of a thread; every leaf is loaded before the first op, so
// all of a thread's loads are in flight together.  Operands 0-7 name the
// leaves, 8-15 the temporaries that op results go to.
#pragma once

#include <cmath>
#include <type_traits>

#include "common.h"

namespace akhip {

enum ExprType : uint8_t {
  kTBool = 0,
  kTInt8,
  kTUInt8,
  kTInt16,
  kTUInt16,
  kTInt32,
  kTUInt32,
  kTInt64,
  kTUInt64,
  kTFloat32,
  kTFloat64,
};

enum ExprCode : uint8_t {
  kXConst = 0,  // dst = imm (canonical bits of type `type`)
  kXCast,       // dst = (to) a, a of type `type`
  kXAdd,
  kXSub,
  kXMul,
  kXDiv,  // true division; floats only (the host casts integers to float64 first)
  kXGt,
  kXGe,
  kXLt,
  kXLe,
  kXEq,
  kXNe,
  kXAnd,  // bitwise (logical for bool)
  kXOr,
  kXXor,
  kXLogicalAnd,
  kXLogicalOr,
  kXLogicalNot,
  kXInvert,  // bitwise not (logical for bool)
  kXNeg,
  kXAbs,
  kXCosh,
  kXCos,
  kXSinh,
  kXSin,
  kXExp,
  kXSqrt,
  kXMax,  // numpy maximum / minimum: NaN propagates
  kXMin,
};

constexpr int kExprMaxLeaves = 8;
constexpr int kExprMaxOps = 40;
constexpr int kExprRegs = 16;  // registers 0 .. nleaves - 1 start with the leaves

struct ExprLeaf {
  const void* data;
  const int64_t* index;  // nullptr: data[i]; else data[index[i]]
  int32_t type;
  int32_t leg;  // the pair kernel's leaves: which element of the pair (0 or 1) the leaf reads
};

struct ExprOp {
  uint8_t code, type, dst, a, b;
  uint8_t to;    // kXCast: the target type; comparisons: 1 / 2 for numpy's mixed uint64 / int64 loops
  uint8_t bimm;  // 1: the second operand is `imm` (a constant), not register b
  uint8_t pad1;
  uint64_t imm;  // kXConst's value, or the immediate second operand
};

struct ExprProgram {
  int32_t nleaves, nops, out, out_type;
  int32_t wide;   // 64-bit registers (some type of the program needs 64 bits)
  int32_t nregs;  // registers the program uses (every register number is below it)
  ExprLeaf leaf[kExprMaxLeaves];
  ExprOp op[kExprMaxOps];
};

inline bool expr_program_ok(const ExprProgram& p) {
  if (!(p.nleaves >= 0 && p.nleaves <= kExprMaxLeaves && p.nops >= 0 && p.nops <= kExprMaxOps &&
        p.nregs >= 1 && p.nregs <= kExprRegs && p.nleaves <= p.nregs && p.out >= 0 && p.out < p.nregs &&
        p.out_type >= kTBool && p.out_type <= kTFloat64))
    return false;
  for (int k = 0; k < p.nops; k++) {
    const ExprOp& o = p.op[k];
    if (o.dst >= p.nregs || o.a >= p.nregs || o.b >= p.nregs || o.code > kXMin || o.type > kTFloat64) return false;
  }
  return true;
}

// Element slots per thread: the register file is 32 VGPRs (the widest
// register tuple), so U slots leave 32 / (U * words) registers -- 4 slots
// of 32-bit registers when the program needs at most 8, else 2 (64-bit
// registers: 2, else 1).
inline int expr_slots(const ExprProgram& p) {
  const bool small = p.nregs <= 8;
  return p.wide ? (small ? 2 : 1) : (small ? 4 : 2);
}

// Calls f(XKind<W, U>{}) for the program's register width and slots.
template <typename W, int U>
struct XKind {
  using Word = W;
  static constexpr int slots = U;
};
template <class F>
inline void expr_dispatch(const ExprProgram& p, F f) {
  switch (expr_slots(p) + (p.wide ? 8 : 0)) {
    case 4: f(XKind<uint32_t, 4>{}); break;
    case 2: f(XKind<uint32_t, 2>{}); break;
    case 10: f(XKind<uint64_t, 2>{}); break;
    default: f(XKind<uint64_t, 1>{}); break;
  }
}

__device__ inline bool x_is_signed(int t) { return t == kTInt8 || t == kTInt16 || t == kTInt32 || t == kTInt64; }

template <typename W>
__device__ inline float x_f(W r) {
  return __uint_as_float((uint32_t)r);
}
template <typename W>
__device__ inline W x_from_f(float f) {
  return (W)__float_as_uint(f);
}
__device__ inline double x_d(uint64_t r) { return __longlong_as_double((long long)r); }
__device__ inline uint64_t x_from_d(double d) { return (uint64_t)__double_as_longlong(d); }

// a signed value of width W (sign-extended canonical bits)
template <typename W>
using XS = std::make_signed_t<W>;

template <typename W>
__device__ inline W x_load(const ExprLeaf& l, int64_t i) {
  const int64_t j = l.index ? l.index[i] : i;
  switch (l.type) {
    case kTBool: return ((const uint8_t*)l.data)[j] != 0;
    case kTInt8: return (W)(XS<W>)((const int8_t*)l.data)[j];
    case kTUInt8: return ((const uint8_t*)l.data)[j];
    case kTInt16: return (W)(XS<W>)((const int16_t*)l.data)[j];
    case kTUInt16: return ((const uint16_t*)l.data)[j];
    case kTInt32: return (W)(XS<W>)((const int32_t*)l.data)[j];
    case kTUInt32:
    case kTFloat32: return ((const uint32_t*)l.data)[j];
    default:
      if constexpr (sizeof(W) == 8) return ((const uint64_t*)l.data)[j];
      else return 0;  // 64-bit leaves never reach a narrow program
  }
}

// An integer result wrapped to its type's width, in canonical form.
template <typename W>
__device__ inline W x_wrap(W v, int t) {
  switch (t) {
    case kTBool: return v != 0;
    case kTInt8: return (W)(XS<W>)(int8_t)v;
    case kTUInt8: return v & 0xffu;
    case kTInt16: return (W)(XS<W>)(int16_t)v;
    case kTUInt16: return v & 0xffffu;
    case kTInt32: return (W)(XS<W>)(int32_t)v;
    case kTUInt32: return v & 0xffffffffu;
    default: return v;
  }
}

// Truth value of a canonical register of type t (x != 0).
template <typename W>
__device__ inline bool x_truth(W v, int t) {
  if (t == kTFloat32) return x_f(v) != 0.0f;
  if constexpr (sizeof(W) == 8) {
    if (t == kTFloat64) return x_d(v) != 0.0;
  }
  return v != 0;
}

template <typename W>
__device__ inline W x_cast(W v, int from, int to) {
  if (from == to) return v;
  if (to == kTBool) return x_truth(v, from);
  if (to == kTFloat32) {
    if constexpr (sizeof(W) == 8) {
      if (from == kTFloat64) return x_from_f<W>((float)x_d(v));
    }
    if (x_is_signed(from)) return x_from_f<W>((float)(XS<W>)v);
    return x_from_f<W>((float)v);
  }
  if constexpr (sizeof(W) == 8) {
    if (to == kTFloat64) {
      if (from == kTFloat32) return x_from_d((double)x_f(v));
      if (x_is_signed(from)) return x_from_d((double)(int64_t)v);
      return x_from_d((double)v);
    }
    if (from == kTFloat64) return x_wrap<W>(x_is_signed(to) ? (W)(int64_t)x_d(v) : (W)x_d(v), to);
  }
  // to an integer type: floats truncate toward zero (numpy's astype)
  if (from == kTFloat32) return x_wrap<W>(x_is_signed(to) ? (W)(XS<W>)x_f(v) : (W)x_f(v), to);
  return x_wrap<W>(v, to);
}

// ---- vectorised interpretation ---------------------------------------------------
// U slots as one vector value: a first-class value, never addressed through
// memory, so selecting a register by a run-time index stays in VGPRs.
template <int U, typename W>
struct XV {
  typedef W vec __attribute__((ext_vector_type(U)));
  vec v;
};

template <int U, typename W, class F>
__device__ inline XV<U, W> x_map1(const XV<U, W>& a, F f) {
  XV<U, W> r;
#pragma unroll
  for (int u = 0; u < U; u++) r.v[u] = f(a.v[u]);
  return r;
}

template <int U, typename W, class F>
__device__ inline XV<U, W> x_map2(const XV<U, W>& a, const XV<U, W>& b, F f) {
  XV<U, W> r;
#pragma unroll
  for (int u = 0; u < U; u++) r.v[u] = f(a.v[u], b.v[u]);
  return r;
}

// integer results wrapped to the width of type t (one dispatch)
template <int U, typename W>
__device__ inline XV<U, W> x_wrap_v(const XV<U, W>& a, int t) {
  switch (t) {
    case kTBool: return x_map1<U, W>(a, [](W v) -> W { return v != 0; });
    case kTInt8: return x_map1<U, W>(a, [](W v) -> W { return (W)(XS<W>)(int8_t)v; });
    case kTUInt8: return x_map1<U, W>(a, [](W v) -> W { return v & 0xffu; });
    case kTInt16: return x_map1<U, W>(a, [](W v) -> W { return (W)(XS<W>)(int16_t)v; });
    case kTUInt16: return x_map1<U, W>(a, [](W v) -> W { return v & 0xffffu; });
    case kTInt32: return x_map1<U, W>(a, [](W v) -> W { return (W)(XS<W>)(int32_t)v; });
    case kTUInt32: return x_map1<U, W>(a, [](W v) -> W { return v & 0xffffffffu; });
    default: return a;
  }
}

template <int U, typename W, class CMP>
__device__ inline XV<U, W> x_compare(const ExprOp& o, const XV<U, W>& a, const XV<U, W>& b, CMP cmp) {
  const int t = o.type;
  if (t == kTFloat32) return x_map2<U, W>(a, b, [&](W x, W y) -> W { return cmp(x_f(x), x_f(y)); });
  if constexpr (sizeof(W) == 8) {
    if (o.to != 0) {
      // numpy's mixed uint64 / int64 loops compare exactly: to = 1 when a
      // is the uint64 operand, 2 when b is
      const int side = o.to;
      return x_map2<U, W>(a, b, [&](W x, W y) -> W {
        const uint64_t uu = side == 1 ? x : y;
        const int64_t vv = (int64_t)(side == 1 ? y : x);
        const int c = vv < 0 ? 1 : (uu > (uint64_t)vv ? 1 : (uu < (uint64_t)vv ? -1 : 0));  // sign of uu - vv
        return cmp(side == 1 ? c : -c, 0);
      });
    }
    if (t == kTFloat64) return x_map2<U, W>(a, b, [&](W x, W y) -> W { return cmp(x_d(x), x_d(y)); });
  }
  if (x_is_signed(t)) return x_map2<U, W>(a, b, [&](W x, W y) -> W { return cmp((XS<W>)x, (XS<W>)y); });
  return x_map2<U, W>(a, b, [&](W x, W y) -> W { return cmp(x, y); });
}

template <int U, typename W, class FF, class FD, class FI>
__device__ inline XV<U, W> x_arith_v(int t, const XV<U, W>& a, const XV<U, W>& b, FF ff, FD fd, FI fi) {
  if (t == kTFloat32) return x_map2<U, W>(a, b, [&](W x, W y) -> W { return x_from_f<W>(ff(x_f(x), x_f(y))); });
  if constexpr (sizeof(W) == 8) {
    if (t == kTFloat64) return x_map2<U, W>(a, b, [&](W x, W y) -> W { return x_from_d(fd(x_d(x), x_d(y))); });
  }
  return x_wrap_v<U, W>(x_map2<U, W>(a, b, fi), t);
}

template <int U, typename W, class FF, class FD>
__device__ inline XV<U, W> x_math_v(int t, const XV<U, W>& a, FF ff, FD fd) {
  if constexpr (sizeof(W) == 8) {
    if (t == kTFloat64) return x_map1<U, W>(a, [&](W x) -> W { return x_from_d(fd(x_d(x))); });
  }
  return x_map1<U, W>(a, [&](W x) -> W { return x_from_f<W>(ff(x_f(x))); });
}

template <typename F>
__device__ inline F x_fmax(F a, F b) {
  return (a != a || b != b) ? (a != a ? a : b) : (a > b ? a : b);
}
template <typename F>
__device__ inline F x_fmin(F a, F b) {
  return (a != a || b != b) ? (a != a ? a : b) : (a < b ? a : b);
}

template <int U, typename W>
__device__ inline XV<U, W> x_op_v(const ExprOp& o, const XV<U, W>& a, const XV<U, W>& b) {
  const int t = o.type;
  switch (o.code) {
    case kXConst: {
      XV<U, W> r;
      r.v = (W)o.imm;
      return r;
    }
    case kXCast: {
      const int to = o.to;
      return x_map1<U, W>(a, [&](W v) { return x_cast<W>(v, t, to); });
    }
    case kXAdd:
      if (t == kTBool) return x_map2<U, W>(a, b, [](W x, W y) -> W { return x | y; });
      return x_arith_v<U, W>(t, a, b, [](float x, float y) { return x + y; }, [](double x, double y) { return x + y; },
                             [](W x, W y) -> W { return x + y; });
    case kXSub:
      return x_arith_v<U, W>(t, a, b, [](float x, float y) { return x - y; }, [](double x, double y) { return x - y; },
                             [](W x, W y) -> W { return x - y; });
    case kXMul:
      if (t == kTBool) return x_map2<U, W>(a, b, [](W x, W y) -> W { return x & y; });
      return x_arith_v<U, W>(t, a, b, [](float x, float y) { return x * y; }, [](double x, double y) { return x * y; },
                             [](W x, W y) -> W { return x * y; });
    case kXDiv:
      return x_arith_v<U, W>(t, a, b, [](float x, float y) { return x / y; }, [](double x, double y) { return x / y; },
                             [](W x, W) -> W { return x; });
    case kXMax:
    case kXMin: {
      const bool mx = o.code == kXMax;
      if (t == kTFloat32)
        return x_map2<U, W>(a, b, [&](W x, W y) -> W {
          return x_from_f<W>(mx ? x_fmax(x_f(x), x_f(y)) : x_fmin(x_f(x), x_f(y)));
        });
      if constexpr (sizeof(W) == 8) {
        if (t == kTFloat64)
          return x_map2<U, W>(a, b, [&](W x, W y) -> W {
            return x_from_d(mx ? x_fmax(x_d(x), x_d(y)) : x_fmin(x_d(x), x_d(y)));
          });
      }
      if (x_is_signed(t))
        return x_map2<U, W>(a, b, [&](W x, W y) -> W { return (mx == ((XS<W>)x > (XS<W>)y)) ? x : y; });
      return x_map2<U, W>(a, b, [&](W x, W y) -> W { return (mx == (x > y)) ? x : y; });
    }
    case kXGt: return x_compare<U, W>(o, a, b, [](auto x, auto y) -> W { return x > y; });
    case kXGe: return x_compare<U, W>(o, a, b, [](auto x, auto y) -> W { return x >= y; });
    case kXLt: return x_compare<U, W>(o, a, b, [](auto x, auto y) -> W { return x < y; });
    case kXLe: return x_compare<U, W>(o, a, b, [](auto x, auto y) -> W { return x <= y; });
    case kXEq: return x_compare<U, W>(o, a, b, [](auto x, auto y) -> W { return x == y; });
    case kXNe: return x_compare<U, W>(o, a, b, [](auto x, auto y) -> W { return x != y; });
    case kXAnd: return x_map2<U, W>(a, b, [](W x, W y) -> W { return x & y; });
    case kXOr: return x_map2<U, W>(a, b, [](W x, W y) -> W { return x | y; });
    case kXXor: return x_map2<U, W>(a, b, [](W x, W y) -> W { return x ^ y; });
    case kXLogicalAnd:
      return x_map2<U, W>(a, b, [&](W x, W y) -> W { return x_truth<W>(x, t) && x_truth<W>(y, t); });
    case kXLogicalOr:
      return x_map2<U, W>(a, b, [&](W x, W y) -> W { return x_truth<W>(x, t) || x_truth<W>(y, t); });
    case kXLogicalNot: return x_map1<U, W>(a, [&](W x) -> W { return !x_truth<W>(x, t); });
    case kXInvert:
      if (t == kTBool) return x_map1<U, W>(a, [](W x) -> W { return x == 0; });
      return x_wrap_v<U, W>(x_map1<U, W>(a, [](W x) -> W { return ~x; }), t);
    case kXNeg:
      if (t == kTFloat32) return x_map1<U, W>(a, [](W x) -> W { return x_from_f<W>(-x_f(x)); });
      if constexpr (sizeof(W) == 8) {
        if (t == kTFloat64) return x_map1<U, W>(a, [](W x) -> W { return x_from_d(-x_d(x)); });
      }
      return x_wrap_v<U, W>(x_map1<U, W>(a, [](W x) -> W { return (W)0 - x; }), t);
    case kXAbs:
      if (t == kTFloat32) return x_map1<U, W>(a, [](W x) -> W { return x_from_f<W>(fabsf(x_f(x))); });
      if constexpr (sizeof(W) == 8) {
        if (t == kTFloat64) return x_map1<U, W>(a, [](W x) -> W { return x_from_d(fabs(x_d(x))); });
      }
      if (x_is_signed(t))
        return x_wrap_v<U, W>(x_map1<U, W>(a, [](W x) -> W { return (XS<W>)x < 0 ? (W)0 - x : x; }), t);
      return a;
    case kXCosh: return x_math_v<U, W>(t, a, [](float x) { return coshf(x); }, [](double x) { return cosh(x); });
    case kXCos: return x_math_v<U, W>(t, a, [](float x) { return cosf(x); }, [](double x) { return cos(x); });
    case kXSinh: return x_math_v<U, W>(t, a, [](float x) { return sinhf(x); }, [](double x) { return sinh(x); });
    case kXSin: return x_math_v<U, W>(t, a, [](float x) { return sinf(x); }, [](double x) { return sin(x); });
    case kXExp: return x_math_v<U, W>(t, a, [](float x) { return expf(x); }, [](double x) { return exp(x); });
    default: return x_math_v<U, W>(t, a, [](float x) { return sqrtf(x); }, [](double x) { return sqrt(x); });
  }
}

// The register file: one vector of 32 VGPRs -- the widest register tuple --
// holding 32 / (U * words) registers of U slots, indexed at run time by the
// op's uniform register number, which the compiler turns into VGPR-indexed
// moves (s_set_gpr_idx) rather than a select over every register or a spill
// to scratch.  (Two files, or one wider than 32 VGPRs, lose that: the
// compiler then copies or selects whole files.)  The file is passed and
// returned as a value -- whole-vector extracts and inserts -- so its address
// is never taken.
template <int U, typename W>
struct XFile {
  static constexpr int elems = 32 / (sizeof(W) / 4);
  static constexpr int regs = elems / U;
  typedef W vec __attribute__((ext_vector_type(elems)));
};
template <int U, typename W>
using XFileV = typename XFile<U, W>::vec;

template <int U, typename W>
__device__ inline XV<U, W> x_get(XFileV<U, W> f, int r) {
  const int base = (r & (XFile<U, W>::regs - 1)) * U;
  XV<U, W> out;
#pragma unroll
  for (int u = 0; u < U; u++) out.v[u] = f[base + u];
  return out;
}

template <int U, typename W>
__device__ inline XFileV<U, W> x_set(XFileV<U, W> f, int r, XV<U, W> val) {
  const int base = (r & (XFile<U, W>::regs - 1)) * U;
#pragma unroll
  for (int u = 0; u < U; u++) f[base + u] = val.v[u];
  return f;
}

// Runs the program on a file whose first registers hold the leaves (the host
// picks U so that every register number of the program is in the file).
// Op k of the program read as four dwords (one scalar load) and unpacked: a
// field-by-field read of the bytes at offsets 5 and 6 is not dword aligned
// and would become a vector memory load per op.
__device__ inline ExprOp x_fetch(const ExprProgram& p, int k) {
  const uint32_t* w = reinterpret_cast<const uint32_t*>(&p.op[k]);
  const uint32_t w0 = w[0], w1 = w[1];
  ExprOp o;
  o.code = (uint8_t)w0;
  o.type = (uint8_t)(w0 >> 8);
  o.dst = (uint8_t)(w0 >> 16);
  o.a = (uint8_t)(w0 >> 24);
  o.b = (uint8_t)w1;
  o.to = (uint8_t)(w1 >> 8);
  o.bimm = (uint8_t)(w1 >> 16);
  o.pad1 = 0;
  o.imm = (uint64_t)w[2] | ((uint64_t)w[3] << 32);
  return o;
}
static_assert(sizeof(ExprOp) == 16, "ExprOp is four dwords");

template <int U, typename W>
__device__ inline void expr_run(const ExprProgram& p, XFileV<U, W> f, W (&out)[U]) {
  for (int k = 0; k < p.nops; k++) {
    const ExprOp o = x_fetch(p, k);
    const XV<U, W> a = x_get<U, W>(f, o.a);
    XV<U, W> b;
    if (o.bimm)
      b.v = (W)o.imm;
    else
      b = x_get<U, W>(f, o.b);
    f = x_set<U, W>(f, o.dst, x_op_v<U, W>(o, a, b));
  }
  const XV<U, W> res = x_get<U, W>(f, p.out);
#pragma unroll
  for (int u = 0; u < U; u++) out[u] = res.v[u];
}

// Evaluates the program at U element positions (pos[u] < 0: slot unused),
// with registers of width W (uint32_t: the program's `wide` is 0).
template <int U, typename W>
__device__ inline void expr_eval(const ExprProgram& p, const int64_t (&pos)[U], W (&out)[U]) {
  XFileV<U, W> l = 0;
#pragma unroll
  for (int k = 0; k < kExprMaxLeaves; k++)
#pragma unroll
    for (int u = 0; u < U; u++)
      if (k < p.nleaves && pos[u] >= 0) l[k * U + u] = x_load<W>(p.leaf[k], pos[u]);
  expr_run<U, W>(p, l, out);
}

// The same over pairs: leaf k reads element pos0[u] (leg 0) or pos1[u]
// (leg 1) of its column.
template <int U, typename W>
__device__ inline void expr_eval_pairs(const ExprProgram& p, const int64_t (&pos0)[U], const int64_t (&pos1)[U],
                                       W (&out)[U]) {
  XFileV<U, W> l = 0;
#pragma unroll
  for (int k = 0; k < kExprMaxLeaves; k++) {
    if (k < p.nleaves) {
      const bool second = p.leaf[k].leg != 0;
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int64_t q = second ? pos1[u] : pos0[u];
        if (q >= 0) l[k * U + u] = x_load<W>(p.leaf[k], q);
      }
    }
  }
  expr_run<U, W>(p, l, out);
}

// Stores a canonical register as element i of an array of type t.
template <typename W>
__device__ inline void x_store(void* outp, int t, int64_t i, W v) {
  switch (t) {
    case kTBool:
    case kTInt8:
    case kTUInt8: ((uint8_t*)outp)[i] = (uint8_t)v; break;
    case kTInt16:
    case kTUInt16: ((uint16_t*)outp)[i] = (uint16_t)v; break;
    case kTInt32:
    case kTUInt32:
    case kTFloat32: ((uint32_t*)outp)[i] = (uint32_t)v; break;
    default:
      if constexpr (sizeof(W) == 8) ((uint64_t*)outp)[i] = v;
      break;
  }
}

}  // namespace akhip
