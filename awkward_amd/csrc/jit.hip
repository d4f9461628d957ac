// jit.hip -- kernels compiled at run time for one expression program.
//
// The interpreter (expr.h) runs any program, but each op costs a dispatch
// and indexed register moves around its arithmetic: over the pairs of
// ak.combinations that is most of the time.  Here the pair-extrema kernel
// (pair_jit.h) is compiled with hipRTC for the program's shape -- its ops
// as straight-line code, each op being expr.h's own x_op_v with the op a
// compile-time constant, so the values are the interpreter's bit for bit
// (contraction off: no fused multiply-adds across ops) -- once per shape and
// device, and cached for the life of the process.  The build embeds expr.h
// and pair_jit.h as text (jit_embed.h); the run-time source needs nothing
// else from this library.

#include <hip/hiprtc.h>

#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "jit.h"
#include "jit_embed.h"

namespace akhip {
namespace {

// common.h for the run-time source: the integer types and the wave width
const char kJitCommonH[] =
    "#pragma once\n"
    "using __hip_internal::int8_t; using __hip_internal::uint8_t; using __hip_internal::int16_t;\n"
    "using __hip_internal::uint16_t; using __hip_internal::int32_t; using __hip_internal::uint32_t;\n"
    "using __hip_internal::int64_t; using __hip_internal::uint64_t;\n"
    "namespace akhip { constexpr int kWave = 64; }\n";

constexpr int kJitStage = 384;  // elements per staged column and wave (pjit::kStage)

// pjit::Leaves and pjit::Out, member for member
struct JitLeaves {
  const void* data[kExprMaxLeaves];
  const int64_t* index[kExprMaxLeaves];
};
struct JitOut {
  void* max;
  void* min;
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

// The run-time source for program p's extrema: its shape as constants and
// its ops as straight-line code (registers R[0 .. nleaves) start with the
// leaves, as in the interpreter).
std::string pair_source(const ExprProgram& p, const int8_t* slot, const int8_t* first, int nslots, bool mx, bool mn) {
  const int nl = p.nleaves;
  std::vector<int> vslot(nl > 0 ? nl : 1, 0), vleg(nl > 0 ? nl : 1, 0), vtype(nl > 0 ? nl : 1, 0);
  std::vector<int> vfirst(nslots > 0 ? nslots : 1, 0);
  for (int k = 0; k < nl; k++) {
    vslot[k] = slot[k];
    vleg[k] = p.leaf[k].leg != 0;
    vtype[k] = p.leaf[k].type;
  }
  for (int s = 0; s < nslots; s++) vfirst[s] = first[s];
  std::string s = "#include \"expr.h\"\nnamespace akhip {\nnamespace pjit {\n";
  s += p.wide ? "using W = uint64_t;\n" : "using W = uint32_t;\n";
  s += p.out_type == kTFloat64 ? "using F = double;\n" : "using F = float;\n";
  s += std::string("constexpr bool kMX = ") + (mx ? "true" : "false") + ", kMN = " + (mn ? "true" : "false") + ";\n";
  s += "constexpr int kStage = " + std::to_string(kJitStage) + ", kU = " + std::to_string(p.wide ? 2 : 4) + ";\n";
  s += "constexpr int kNLeaves = " + std::to_string(nl) + ", kNSlots = " + std::to_string(nslots) + ";\n";
  s += "constexpr int kSlot[] = " + int_list(vslot) + ";\n";
  s += "constexpr int kLeg[] = " + int_list(vleg) + ";\n";
  s += "constexpr int kType[] = " + int_list(vtype) + ";\n";
  s += "constexpr int kFirst[] = " + int_list(vfirst) + ";\n";
  s += "__device__ inline W eval(const W (&l)[" + std::to_string(nl > 0 ? nl : 1) + "]) {\n";
  s += "  XV<1, W> R[" + std::to_string(kExprRegs) + "] = {};\n";
  for (int k = 0; k < nl; k++) s += "  R[" + std::to_string(k) + "].v = l[" + std::to_string(k) + "];\n";
  for (int k = 0; k < p.nops; k++) {
    const ExprOp& o = p.op[k];
    s += "  {\n    constexpr ExprOp o = {" + std::to_string(o.code) + ", " + std::to_string(o.type) + ", " +
         std::to_string(o.dst) + ", " + std::to_string(o.a) + ", " + std::to_string(o.b) + ", " +
         std::to_string(o.to) + ", " + std::to_string(o.bimm) + ", 0, " + std::to_string(o.imm) + "ull};\n";
    s += "    XV<1, W> B;\n";
    s += o.bimm ? "    B.v = (W)o.imm;\n" : "    B = R[" + std::to_string(o.b) + "];\n";
    s += "    R[" + std::to_string(o.dst) + "] = x_op_v<1, W>(o, R[" + std::to_string(o.a) + "], B);\n  }\n";
  }
  s += "  return R[" + std::to_string(p.out) + "].v[0];\n}\n}  // namespace pjit\n}  // namespace akhip\n";
  s += "#include \"pair_jit.h\"\n";
  return s;
}

// Compiles and loads the source on device dev; null (after one message on
// stderr) when hipRTC cannot build it.
hipFunction_t build_kernel(const std::string& src, int dev) {
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return nullptr;
  const std::string arch = std::string("--offload-arch=") + prop.gcnArchName;
  const char* headers[] = {kJitExprH, kJitCommonH, kJitPairH};
  const char* names[] = {"expr.h", "common.h", "pair_jit.h"};
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "pair_jit.hip", 3, headers, names) != HIPRTC_SUCCESS) return nullptr;
  const char* opts[] = {arch.c_str(), "-O3", "-std=c++17", "-ffp-contract=off", "-Xclang", "-target-feature",
                        "-Xclang", "-packed-fp32-ops"};
  hipFunction_t fn = nullptr;
  if (hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts) == HIPRTC_SUCCESS) {
    size_t size = 0;
    hiprtcGetCodeSize(prog, &size);
    std::vector<char> code(size);
    hipModule_t mod;
    if (size > 0 && hiprtcGetCode(prog, code.data()) == HIPRTC_SUCCESS &&
        hipModuleLoadData(&mod, code.data()) == hipSuccess)
      if (hipModuleGetFunction(&fn, mod, "akhip_pair_jit") != hipSuccess) fn = nullptr;
  } else {
    size_t size = 0;
    hiprtcGetProgramLogSize(prog, &size);
    std::string log(size, '\0');
    if (size > 0) hiprtcGetProgramLog(prog, &log[0]);
    std::fprintf(stderr, "awkward_hip: run-time compilation of a pair kernel failed; interpreting it\n%s\n",
                 log.c_str());
  }
  hiprtcDestroyProgram(&prog);
  return fn;
}

}  // namespace

const void* pair_jit_kernel(const ExprProgram& p, const int8_t* slot, const int8_t* first, int nslots, bool mx,
                            bool mn) {
  if (!jit_enabled()) return nullptr;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  const std::string src = pair_source(p, slot, first, nslots, mx || !mn, mn);
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
                  void* out_max, void* out_min, int64_t* out_argmax, int64_t* out_argmin, int8_t* out_mask) {
  JitLeaves lv{};
  for (int k = 0; k < p.nleaves; k++) {
    lv.data[k] = p.leaf[k].data;
    lv.index[k] = p.leaf[k].index;
  }
  JitOut out{out_max, out_min, out_argmax, out_argmin, out_mask};
  const unsigned grid = (unsigned)ceil_div(n, (int64_t)kWave);
  const size_t lds = (size_t)(nslots > 0 ? nslots : 1) * kJitStage * (p.wide ? 8 : 4);
  void* args[] = {&lv, (void*)&offsets, &n, &out};
  // a failed launch is reported by the caller's launch check
  (void)hipModuleLaunchKernel((hipFunction_t)kernel, grid, 1, 1, kWave, 1, 1, (unsigned)lds, stream(), args,
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

}  // extern "C"
