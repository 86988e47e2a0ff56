// jit.hip — straight-line register tile passes, compiled at run time.
//
// The interpreting kernel k_tile_reg dispatches every op of a pass's register
// program through one loop over a switch; the amplitudes it carries from op to
// op then pass through a merge point after every op, which costs about 64
// register moves per op on gfx950, as much as the op's own arithmetic.  Here
// each pass becomes its own kernel whose body calls the same device functions
// (regtile.h) one after another with the slot numbers as template arguments,
// so values stay in place.  Matrices, controls and remap tables are still
// read from the uploaded program at run time: a module depends only on the
// program's structure, and circuits that differ only in their angles reuse it.
// Modules are compiled with hipRTC for the current device and cached by their
// source text.  The cache holds at most kJitCacheModules modules: past that,
// the least recently used one leaves the cache, and it is unloaded once no
// run that launched its kernels still holds it (a run holds its modules
// until its stream is synchronised).
#include <hip/hiprtc.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <thread>
#include <type_traits>

#include "runtime.h"
#include "jit_src.inc"  // kJitPrelude: integer typedefs + devtypes.h + regtile.h (generated by build.py)

namespace nqpu {

namespace {

struct JitModule {
  hipModule_t mod = nullptr;
  std::vector<hipFunction_t> fns;
  uint64_t last_use = 0;
  ~JitModule() {
    if (mod) (void)hipModuleUnload(mod);
  }
};

constexpr size_t kJitCacheModules = 512;  // 8 passes per module
std::mutex g_mu;
std::map<std::string, std::shared_ptr<JitModule>> g_cache;
uint64_t g_compiled_passes = 0;
uint64_t g_use_clock = 0;

// Drop least recently used modules beyond the cap (caller holds g_mu).
void evict_locked() {
  while (g_cache.size() > kJitCacheModules) {
    auto victim = g_cache.begin();
    for (auto it = g_cache.begin(); it != g_cache.end(); ++it)
      if (it->second->last_use < victim->second->last_use) victim = it;
    g_cache.erase(victim);  // unloaded when the last run holding it lets go
  }
}

// C++ names of the amplitude types in generated source
template <typename R> struct JitNames;
template <> struct JitNames<float> {
  static constexpr const char* r = "float";
  static constexpr const char* t = "float2";
};
template <> struct JitNames<double> {
  static constexpr const char* r = "double";
  static constexpr const char* t = "double2";
};

template <typename R>
void emit_op(std::ostringstream& o, const RegOp<R>& r, int k, int eb) {
  const char* RT = JitNames<R>::r;
  const char* TT = JitNames<R>::t;
  switch (r.kind) {
    case R_REMAP:
      o << "    nqpu::reg_remap<" << RT << ", " << TT << ", " << eb << ">(v, sm, opv[" << k << "], tid, true);\n";
      return;
    case R_MFMA:
      o << "    { int var = 0;";
      if (r.pad > 0) o << " var |= (int)((base >> " << (r.s1 & 255) << ") & 1);";
      if (r.pad > 1) o << " var |= (int)((base >> " << (r.s1 >> 8) << ") & 1) << 1;";
      o << " nqpu::reg_mfma<float2, " << eb << ">(v, blob, " << r.s0 << " + var, lane); }\n";
      return;
    case R_XCHG:
      if (r.s1 < 16)
        o << "    nqpu::reg_xchg_lane<" << TT << ", " << eb << ", " << r.s0 << ", " << r.s1 << ">(v, lane);\n";
      else
        o << "    nqpu::reg_xchg_reg<" << TT << ", " << eb << ", " << r.s0 << ", " << (r.s1 - 16) << ">(v);\n";
      return;
    case R_QPHASE: {  // chunk masks are structure: compile-time constants
      o << "    { const uint64_t qm[4] = {";
      for (int c = 0; c < 4; ++c) o << (c ? ", " : "") << (c < r.pad ? qphase_chunk(r, c) : 0) << "ull";
      o << "}; nqpu::reg_qphase_n<float, float2, " << eb << ", " << r.pad << ">(v, opv[" << k
        << "], blob, tid, base, qm); }\n";
      return;
    }
    default:
      break;
  }
  o << "    if ((base & " << r.gc << "ull) == " << r.gc << "ull) { const bool on = (tid & " << r.ct << "u) == "
    << r.ct << "u; (void)on; ";
  switch (r.kind) {
    case R_U1:
      o << "nqpu::reg_u1<" << RT << ", " << TT << ", " << eb << ", " << r.s0 << ", "
        << ((r.ct == 0 && r.cr == 0) ? 1 : 0) << ">(v, opv[" << k << "], on);";
      break;
    case R_FLIP:
      o << "nqpu::reg_flip<" << RT << ", " << TT << ", " << eb << ", " << r.s0 << ", " << (r.ct == 0 ? 1 : 0)
        << ">(v, opv[" << k << "], on);";
      break;
    case R_D1:
    case R_D2: o << "nqpu::reg_diag<" << RT << ", " << TT << ", " << eb << ">(v, opv[" << k << "], tid, on, base);"; break;
    case R_U2:
      o << "nqpu::reg_u2<" << RT << ", " << TT << ", " << eb << ", " << r.s0 << ", " << r.s1 << ">(v, opv[" << k << "], on);";
      break;
    case R_SWAP:
      o << "nqpu::reg_swap<" << RT << ", " << TT << ", " << eb << ", " << r.s0 << ", " << r.s1 << ">(v, opv[" << k << "], on);";
      break;
    default: break;
  }
  o << " }\n";
}

template <typename R>
std::string kernel_source(int idx, const std::vector<RegOp<R>>& prog, const PassArgs& pa, int eb) {
  const int op0 = pa.op0, nops = pa.nops;
  const char* RT = JitNames<R>::r;
  const char* TT = JitNames<R>::t;
  std::ostringstream o;
  o << "extern \"C\" __global__ __launch_bounds__(256, 1) void nqpu_pass_" << idx << "(" << TT
    << "* __restrict__ st, nqpu::PassArgs pa, uint64_t ntiles,\n"
       "    const nqpu::RegOp<" << RT << ">* __restrict__ prog, const unsigned char* __restrict__ blob) {\n"
       "  const nqpu::RegOp<" << RT << ">* ops = prog + pa.op0;\n";
  o << "  nqpu::reg_tile_pass<" << RT << ", " << TT << ", " << eb << ">(st, pa, ntiles, [&](" << TT << " (&v)["
    << (1 << eb) << "], " << TT << "* sm, uint64_t base, int tid) __attribute__((always_inline)) {\n";
  o << "    const int lane = tid & 63; (void)lane; (void)sm; (void)ops; (void)blob;\n";
  if (std::is_same<R, double>::value)
    // the ops' operands (FP64 matrices) are read where they are used: an op
    // pointer that depends on the tile (base >> 63 is 0) keeps the compiler
    // from hoisting every op's scalar loads out of the tile loop, which
    // overflows the scalar registers into vector-register lanes
    o << "    const nqpu::RegOp<double>* opv = ops + (base >> 63);\n";
  else
    o << "    const nqpu::RegOp<float>* opv = ops;\n";
  for (int k = 0; k < nops; ++k) emit_op<R>(o, prog[op0 + k], k, eb);
  o << "  });\n}\n";
  return o.str();
}

NqpuError compile_code(const std::string& src, std::vector<char>* code) {
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "nqpu_passes.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
    set_error("hipRTC: could not create the pass program");
    return NQPU_DEVICE_ERROR;
  }
  // the same code generation settings as the precompiled kernels
  const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-mllvm", "-amdgpu-mfma-vgpr-form",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  const hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
  if (rc != HIPRTC_SUCCESS) {
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    std::string log(ls, '\0');
    if (ls) hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    set_error(std::string("hipRTC compilation of register tile passes failed: ") + hiprtcGetErrorString(rc) +
              "\n" + log.substr(0, 4000));
    return NQPU_DEVICE_ERROR;
  }
  size_t cs = 0;
  hiprtcGetCodeSize(prog, &cs);
  code->resize(cs);
  hiprtcGetCode(prog, code->data());
  hiprtcDestroyProgram(&prog);
  return NQPU_OK;
}

NqpuError compile_module(const std::string& src, int nfn, JitModule* out) {
  std::vector<char> code;
  const NqpuError e = compile_code(src, &code);
  if (e != NQPU_OK) return e;
  if (hipModuleLoadData(&out->mod, code.data()) != hipSuccess) {
    set_error("hipModuleLoadData failed for the compiled register tile passes");
    return NQPU_DEVICE_ERROR;
  }
  out->fns.resize(nfn);
  for (int i = 0; i < nfn; ++i) {
    const std::string name = "nqpu_pass_" + std::to_string(i);
    if (hipModuleGetFunction(&out->fns[i], out->mod, name.c_str()) != hipSuccess) {
      set_error("hipModuleGetFunction failed for " + name);
      return NQPU_DEVICE_ERROR;
    }
  }
  return NQPU_OK;
}

constexpr int kPassesPerModule = 8;

template <typename R>
std::vector<std::string> module_sources(const std::vector<RegOp<R>>& prog, const std::vector<PassArgs>& args,
                                        int eb) {
  const int np = (int)args.size(), per = kPassesPerModule;
  std::vector<std::string> srcs((np + per - 1) / per);
  for (int m = 0; m < (int)srcs.size(); ++m) {
    std::ostringstream o;
    o << kJitPrelude;
    for (int i = m * per; i < std::min(np, (m + 1) * per); ++i)
      o << kernel_source<R>(i - m * per, prog, args[i], eb);
    srcs[m] = o.str();
  }
  return srcs;
}

// With NQPU_JIT_DUMP=<directory>, every generated module source is also
// written there (nqpu_passes_<hash>.hip), for inspection and profiling.
void dump_sources(const std::vector<std::string>& srcs) {
  const char* dir = std::getenv("NQPU_JIT_DUMP");
  if (!dir || !*dir) return;
  for (const std::string& src : srcs) {
    char name[64];
    std::snprintf(name, sizeof name, "/nqpu_passes_%016llx.hip", (unsigned long long)std::hash<std::string>()(src));
    if (FILE* f = std::fopen((std::string(dir) + name).c_str(), "w")) {
      std::fwrite(src.data(), 1, src.size(), f);
      std::fclose(f);
    }
  }
}

}  // namespace

template <typename R>
NqpuError jit_check_compile(const std::vector<RegOp<R>>& prog, const std::vector<PassArgs>& args, int eb) {
  const std::vector<std::string> srcs = module_sources<R>(prog, args, eb);
  dump_sources(srcs);
  for (const std::string& src : srcs) {
    std::vector<char> code;
    const NqpuError e = compile_code(src, &code);
    if (e != NQPU_OK) return e;
  }
  return NQPU_OK;
}

template <typename R>
NqpuError jit_pass_functions(const std::vector<RegOp<R>>& prog, const std::vector<PassArgs>& args, int eb,
                             std::vector<hipFunction_t>* fns, std::vector<std::shared_ptr<void>>* hold) {
  // passes are split over a few modules compiled on parallel threads
  const int np = (int)args.size();
  const int per = kPassesPerModule;
  const std::vector<std::string> srcs = module_sources<R>(prog, args, eb);
  dump_sources(srcs);
  const int nmod = (int)srcs.size();
  std::vector<std::shared_ptr<JitModule>> mods(nmod);
  std::vector<int> todo;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    for (int m = 0; m < nmod; ++m) {
      auto it = g_cache.find(srcs[m]);
      if (it != g_cache.end()) {
        mods[m] = it->second;
        mods[m]->last_use = ++g_use_clock;
      } else {
        todo.push_back(m);
      }
    }
  }
  if (!todo.empty()) {
    std::vector<std::shared_ptr<JitModule>> built(todo.size());
    std::vector<NqpuError> errs(todo.size(), NQPU_OK);
    std::vector<std::string> msgs(todo.size());
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return NQPU_DEVICE_ERROR;
    std::vector<std::thread> th;
    for (size_t t = 0; t < todo.size(); ++t) {
      th.emplace_back([&, t] {
        (void)hipSetDevice(dev);
        built[t] = std::make_shared<JitModule>();
        const int m = todo[t];
        const int nfn = std::min(np, (m + 1) * per) - m * per;
        errs[t] = compile_module(srcs[m], nfn, built[t].get());
        if (errs[t] != NQPU_OK) msgs[t] = last_error();
      });
    }
    for (auto& x : th) x.join();
    for (size_t t = 0; t < todo.size(); ++t)
      if (errs[t] != NQPU_OK) {
        set_error(msgs[t]);
        return errs[t];
      }
    std::lock_guard<std::mutex> lk(g_mu);
    for (size_t t = 0; t < todo.size(); ++t) {
      const int m = todo[t];
      auto& slot = g_cache[srcs[m]];
      if (!slot) {
        slot = built[t];
        g_compiled_passes += slot->fns.size();
      }
      slot->last_use = ++g_use_clock;
      mods[m] = slot;
    }
    evict_locked();
  }
  fns->resize(np);
  for (int i = 0; i < np; ++i) (*fns)[i] = mods[i / per]->fns[i % per];
  for (auto& m : mods) hold->push_back(m);
  return NQPU_OK;
}

template NqpuError jit_check_compile<float>(const std::vector<RegOp<float>>&, const std::vector<PassArgs>&, int);
template NqpuError jit_check_compile<double>(const std::vector<RegOp<double>>&, const std::vector<PassArgs>&, int);
template NqpuError jit_pass_functions<float>(const std::vector<RegOp<float>>&, const std::vector<PassArgs>&, int,
                                             std::vector<hipFunction_t>*, std::vector<std::shared_ptr<void>>*);
template NqpuError jit_pass_functions<double>(const std::vector<RegOp<double>>&, const std::vector<PassArgs>&, int,
                                              std::vector<hipFunction_t>*, std::vector<std::shared_ptr<void>>*);

uint64_t jit_compiled_passes() {
  std::lock_guard<std::mutex> lk(g_mu);
  return g_compiled_passes;
}

}  // namespace nqpu
