// jit.hip -- kernels compiled at run time for one expression program.
//
// The interpreter (expr.h) runs any program, but each op costs a dispatch
// and indexed register moves around its arithmetic: over the pairs of
// ak.combinations that is most of the time.  Here the pair-reduction kernel
// (pair_jit.h) is compiled with hipRTC for the program's shape -- its ops
// as straight-line code, each computed as the interpreter computes it (the
// floating-point ones written out, the rest through expr.h's own x_op_v),
// so the values are the interpreter's bit for bit (contraction off: no fused
// multiply-adds across ops) -- once per shape and device, and cached for the
// life of the process (a handful of shapes per program; the modules stay
// loaded).  The build embeds expr.h, pair_jit.h and tuples.h as text
// (jit_embed.h); the run-time source needs nothing else from this library.

#include <hip/hiprtc.h>

#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "jit.h"
#include "jit_embed.h"
#include "tuples.h"

namespace akhip {
namespace {

// common.h for the run-time source: the integer types and the wave width
const char kJitCommonH[] =
    "#pragma once\n"
    "using __hip_internal::int8_t; using __hip_internal::uint8_t; using __hip_internal::int16_t;\n"
    "using __hip_internal::uint16_t; using __hip_internal::int32_t; using __hip_internal::uint32_t;\n"
    "using __hip_internal::int64_t; using __hip_internal::uint64_t;\n"
    "namespace akhip { constexpr int kWave = 64; }\n";

constexpr int kJitStage = 768;  // elements per staged column (pjit::kStage)
constexpr int kJitCap = 1536;   // pair values kept per workgroup (pjit::kCap)
constexpr int kJitLists = kTupLists;  // lists (and threads) per workgroup (pjit::kLists)

// pjit::Leaves, pjit::Tuples and pjit::Out, member for member
struct JitLeaves {
  const void* data[kExprMaxLeaves];
  const int64_t* index[kExprMaxLeaves];
};
struct JitTuples {
  int32_t nreq;
  int32_t n[2];
  const int64_t* tooffsets[2];
  int64_t* carry[2][3];
};
struct JitOut {
  void* max;
  void* min;
  void* sum;
  int64_t* argmax;
  int64_t* argmin;
  int8_t* mask;
};

std::mutex g_mu;
std::unordered_map<std::string, hipFunction_t> g_kernels;  // device + source -> kernel (null: not buildable)
int g_built = 0;

bool jit_enabled() {
  const char* e = std::getenv("AWKWARD_HIP_JIT");
  return !(e != nullptr && e[0] == '0');
}

std::string int_list(const std::vector<int>& v) {
  std::string s = "{";
  for (size_t k = 0; k < v.size(); k++) s += (k ? ", " : "") + std::to_string(v[k]);
  return s + "}";
}

// One op of the run-time source, over registers R0 .. R15 of canonical bits.
// Floating-point arithmetic, comparisons, max / min and the math functions
// are written out as the interpreter computes them (x_arith_v, x_compare,
// x_math_v: the op in the float type, the result's bits); every other op
// calls the interpreter's own x_op_v with the op as a constant.  (Written
// out, the compiler keeps the math functions' rare paths -- cosf's reduction
// of large arguments -- behind branches; through x_op_v's one-slot vectors
// it computes them for every pair and selects.)
std::string op_source(const ExprOp& o) {
  const std::string d = "R" + std::to_string(o.dst), a = "R" + std::to_string(o.a);
  const std::string b = o.bimm ? "(W)" + std::to_string(o.imm) + "ull" : "R" + std::to_string(o.b);
  const bool f32 = o.type == kTFloat32, f64 = o.type == kTFloat64;
  if (f32 || f64) {
    const std::string X = f32 ? "x_f<W>(" : "x_d(", TO = f32 ? "x_from_f<W>(" : "x_from_d(";
    const std::string A = X + a + ")", B = X + b + ")";
    const char* bin = nullptr;
    switch (o.code) {
      case kXAdd: bin = "+"; break;
      case kXSub: bin = "-"; break;
      case kXMul: bin = "*"; break;
      case kXDiv: bin = "/"; break;
      default: break;
    }
    if (bin != nullptr) return "  " + d + " = " + TO + A + " " + bin + " " + B + ");\n";
    const char* cmp = nullptr;
    switch (o.code) {
      case kXGt: cmp = ">"; break;
      case kXGe: cmp = ">="; break;
      case kXLt: cmp = "<"; break;
      case kXLe: cmp = "<="; break;
      case kXEq: cmp = "=="; break;
      case kXNe: cmp = "!="; break;
      default: break;
    }
    if (cmp != nullptr && o.to == 0) return "  " + d + " = (W)(" + A + " " + cmp + " " + B + ");\n";
    const char* fn = nullptr;
    switch (o.code) {
      case kXCosh: fn = f32 ? "coshf" : "cosh"; break;
      case kXCos: fn = f32 ? "cosf" : "cos"; break;
      case kXSinh: fn = f32 ? "sinhf" : "sinh"; break;
      case kXSin: fn = f32 ? "sinf" : "sin"; break;
      case kXExp: fn = f32 ? "expf" : "exp"; break;
      case kXSqrt: fn = f32 ? "sqrtf" : "sqrt"; break;
      case kXAbs: fn = f32 ? "fabsf" : "fabs"; break;
      default: break;
    }
    if (fn != nullptr) return "  " + d + " = " + TO + fn + "(" + A + "));\n";
    if (o.code == kXNeg) return "  " + d + " = " + TO + "-" + A + ");\n";
    if (o.code == kXMax || o.code == kXMin)
      return "  " + d + " = " + TO + (o.code == kXMax ? "x_fmax(" : "x_fmin(") + A + ", " + B + "));\n";
  }
  return "  {\n    constexpr ExprOp o = {" + std::to_string(o.code) + ", " + std::to_string(o.type) + ", " +
         std::to_string(o.dst) + ", " + std::to_string(o.a) + ", " + std::to_string(o.b) + ", " +
         std::to_string(o.to) + ", " + std::to_string(o.bimm) + ", 0, " + std::to_string(o.imm) + "ull};\n" +
         "    XV<1, W> A, B;\n    A.v = " + a + ";\n    B.v = " + b + ";\n    " + d +
         " = x_op_v<1, W>(o, A, B).v[0];\n  }\n";
}

// The run-time source for program p's reductions: its shape as constants and
// its ops as straight-line code (registers R0 .. R(nleaves - 1) start with
// the leaves, as in the interpreter).
std::string pair_source(const ExprProgram& p, const int8_t* slot, const int8_t* first, int nslots, bool mx, bool mn,
                        bool sum, bool tup) {
  const int nl = p.nleaves;
  std::vector<int> vslot(nl > 0 ? nl : 1, 0), vleg(nl > 0 ? nl : 1, 0), vtype(nl > 0 ? nl : 1, 0);
  std::vector<int> vindexed(nl > 0 ? nl : 1, 0);
  std::vector<int> vfirst(nslots > 0 ? nslots : 1, 0);
  for (int k = 0; k < nl; k++) {
    vslot[k] = slot[k];
    vleg[k] = p.leaf[k].leg != 0;
    vtype[k] = p.leaf[k].type;
    vindexed[k] = p.leaf[k].index != nullptr;
  }
  for (int s = 0; s < nslots; s++) vfirst[s] = first[s];
  std::string s = "#include \"expr.h\"\nnamespace akhip {\nnamespace pjit {\n";
  s += p.wide ? "using W = uint64_t;\n" : "using W = uint32_t;\n";
  s += p.out_type == kTFloat64 ? "using F = double;\n" : "using F = float;\n";
  s += std::string("constexpr bool kMX = ") + (mx ? "true" : "false") + ", kMN = " + (mn ? "true" : "false") +
       ", kSUM = " + (sum ? "true" : "false") + ", kTUP = " + (tup ? "true" : "false") + ";\n";
  s += "constexpr int kStage = " + std::to_string(kJitStage) + ", kCap = " + std::to_string(kJitCap) +
       ", kU = 1;\n";
  s += "constexpr int kNLeaves = " + std::to_string(nl) + ", kNSlots = " + std::to_string(nslots) + ";\n";
  s += "constexpr int kSlot[] = " + int_list(vslot) + ";\n";
  s += "constexpr int kLeg[] = " + int_list(vleg) + ";\n";
  s += "constexpr int kType[] = " + int_list(vtype) + ";\n";
  s += "constexpr bool kIndexed[] = " + int_list(vindexed) + ";\n";
  s += "constexpr int kFirst[] = " + int_list(vfirst) + ";\n";
  s += "__device__ inline W eval(const W (&l)[" + std::to_string(nl > 0 ? nl : 1) + "]) {\n";
  for (int r = 0; r < kExprRegs; r++)
    s += "  W R" + std::to_string(r) + " = " + (r < nl ? "l[" + std::to_string(r) + "]" : std::string("0")) + ";\n";
  for (int k = 0; k < p.nops; k++) s += op_source(p.op[k]);
  s += "  return R" + std::to_string(p.out) + ";\n}\n}  // namespace pjit\n}  // namespace akhip\n";
  s += "#include \"pair_jit.h\"\n";
  return s;
}

// Compiles the source for a target (`--offload-arch` value) into a code
// object; false, with one message and the compiler's log on stderr, when
// hipRTC cannot build it.
bool compile_code(const std::string& src, const std::string& target, std::vector<char>& code) {
  const std::string arch = "--offload-arch=" + target;
  const char* headers[] = {kJitExprH, kJitCommonH, kJitPairH, kJitTuplesH};
  const char* names[] = {"expr.h", "common.h", "pair_jit.h", "tuples.h"};
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "pair_jit.hip", 4, headers, names) != HIPRTC_SUCCESS) return false;
  const char* opts[] = {arch.c_str(), "-O3", "-std=c++17", "-ffp-contract=off", "-Xclang", "-target-feature",
                        "-Xclang", "-packed-fp32-ops"};
  bool ok = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts) == HIPRTC_SUCCESS;
  if (ok) {
    size_t size = 0;
    ok = hiprtcGetCodeSize(prog, &size) == HIPRTC_SUCCESS && size > 0;
    if (ok) {
      code.resize(size);
      ok = hiprtcGetCode(prog, code.data()) == HIPRTC_SUCCESS;
    }
  } else {
    size_t size = 0;
    hiprtcGetProgramLogSize(prog, &size);
    std::string log(size, '\0');
    if (size > 0) hiprtcGetProgramLog(prog, &log[0]);
    std::fprintf(stderr, "awkward_hip: run-time compilation of a pair kernel failed; interpreting it\n%s\n",
                 log.c_str());
  }
  hiprtcDestroyProgram(&prog);
  return ok;
}

// Compiles and loads the source on device dev; null when it cannot be built
// or loaded.
hipFunction_t build_kernel(const std::string& src, int dev) {
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return nullptr;
  std::vector<char> code;
  if (!compile_code(src, prop.gcnArchName, code)) return nullptr;
  hipModule_t mod;
  hipFunction_t fn = nullptr;
  if (hipModuleLoadData(&mod, code.data()) != hipSuccess) return nullptr;
  if (hipModuleGetFunction(&fn, mod, "akhip_pair_jit") != hipSuccess) return nullptr;
  return fn;
}

}  // namespace

const void* pair_jit_kernel(const ExprProgram& p, const int8_t* slot, const int8_t* first, int nslots, bool mx,
                            bool mn, bool sum, bool tuples) {
  if (!jit_enabled()) return nullptr;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  const std::string src = pair_source(p, slot, first, nslots, mx, mn, sum, tuples);
  const std::string key = std::to_string(dev) + "\n" + src;
  std::lock_guard<std::mutex> lock(g_mu);
  auto it = g_kernels.find(key);
  if (it != g_kernels.end()) return (const void*)it->second;
  hipFunction_t fn = build_kernel(src, dev);
  g_kernels.emplace(key, fn);
  if (fn != nullptr) g_built++;
  return (const void*)fn;
}

void pair_jit_run(const void* kernel, const ExprProgram& p, int nslots, const int64_t* offsets, int64_t n,
                  void* out_max, void* out_min, void* out_sum, int64_t* out_argmax, int64_t* out_argmin,
                  int8_t* out_mask, int nreq, const int32_t* req_n, const int64_t* const* req_tooffsets,
                  int64_t* const* req_carries) {
  JitLeaves lv{};
  for (int k = 0; k < p.nleaves; k++) {
    lv.data[k] = p.leaf[k].data;
    lv.index[k] = p.leaf[k].index;
  }
  JitOut out{out_max, out_min, out_sum, out_argmax, out_argmin, out_mask};
  JitTuples tp{};
  tp.nreq = nreq;
  int maxn = 0, used = 0;
  for (int k = 0; k < nreq; k++) {
    tp.n[k] = req_n[k];
    tp.tooffsets[k] = req_tooffsets[k];
    for (int c = 0; c < req_n[k]; c++) tp.carry[k][c] = req_carries[used++];
    maxn = req_n[k] > maxn ? req_n[k] : maxn;
  }
  const unsigned grid = (unsigned)ceil_div(n, (int64_t)kJitLists);
  size_t lds = (size_t)(nslots > 0 ? nslots : 1) * kJitStage * (p.wide ? 8 : 4) +
               (size_t)kJitCap * (p.out_type == kTFloat64 ? 8 : 4);
  const size_t lds_tuples = (size_t)maxn * kTupPad * sizeof(int);  // the carries' stage reuses the same LDS
  if (lds < lds_tuples) lds = lds_tuples;
  void* args[] = {&lv, (void*)&offsets, &n, &out, &tp};
  // a failed launch is reported by the caller's launch check
  (void)hipModuleLaunchKernel((hipFunction_t)kernel, grid, 1, 1, kJitLists, 1, 1, (unsigned)lds, stream(), args,
                              nullptr);
}

}  // namespace akhip

extern "C" {

// The number of kernels compiled at run time so far in this process (one
// per program shape and device).
int awkward_hip_jit_kernels(void) {
  std::lock_guard<std::mutex> lock(akhip::g_mu);
  return akhip::g_built;
}

// Compiles, without a device, the pair kernels of two representative
// programs for `target` (e.g. "gfx950"): 2 * pt0 * pt1 * (cosh(eta0 - eta1) -
// cos(phi0 - phi1)) in float32 with every reduction and the n = 2 / 3
// carries, and the same in float64 after an int8 leaf's cast (an op the
// generated source takes through x_op_v).  0 when both build, else the
// number that failed (the compiler's log on stderr).
int awkward_hip_jit_selftest(const char* target) {
  using namespace akhip;
  int failed = 0;
  for (int wide = 0; wide < 2; wide++) {
    ExprProgram p{};
    const int ft = wide ? kTFloat64 : kTFloat32;
    p.nleaves = 6 + wide;
    p.out_type = ft;
    p.wide = wide;
    for (int k = 0; k < 6; k++) {
      p.leaf[k].type = kTFloat32;
      p.leaf[k].leg = k & 1;
    }
    int8_t slot[7] = {0, 0, 1, 1, 2, 2, 3}, first[4] = {0, 2, 4, 6};
    int nops = 0;
    auto op = [&](int code, int type, int dst, int a, int b, int to = 0, int bimm = 0, uint64_t imm = 0) {
      p.op[nops++] = ExprOp{(uint8_t)code, (uint8_t)type, (uint8_t)dst, (uint8_t)a, (uint8_t)b, (uint8_t)to,
                            (uint8_t)bimm, 0, imm};
    };
    op(kXMul, kTFloat32, 8, 0, 0, 0, 1, 0x40000000u);
    op(kXMul, kTFloat32, 8, 8, 1);
    op(kXSub, kTFloat32, 9, 2, 3);
    op(kXCosh, kTFloat32, 9, 9, 9);
    op(kXSub, kTFloat32, 10, 4, 5);
    op(kXCos, kTFloat32, 10, 10, 10);
    op(kXSub, kTFloat32, 9, 9, 10);
    op(kXMul, kTFloat32, 8, 8, 9);
    if (wide) {
      p.leaf[6].type = kTInt8;
      op(kXCast, kTFloat32, 8, 8, 8, kTFloat64);
      op(kXCast, kTInt8, 11, 6, 6, kTFloat64);
      op(kXAdd, kTFloat64, 8, 8, 11);
    }
    p.nops = nops;
    p.nregs = 12;
    p.out = 8;
    const std::string src = pair_source(p, slot, first, 3 + wide, true, true, true, true);
    std::vector<char> code;
    if (!compile_code(src, target, code)) failed++;
  }
  return failed;
}

}  // extern "C"
