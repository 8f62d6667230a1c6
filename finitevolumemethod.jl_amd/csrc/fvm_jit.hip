// Kernels for closures outside the traced forms, compiled at run time.
//
// The reference accepts any Julia closure for the diffusion and source functions
// (problem.jl:118-163, 425-440); Julia compiles each one into fvm_eqs! on first
// call.  The host package turns such a closure into HIP source (its expression
// DAG: D(x, y, t, u), S(x, y, t, u) and their u-derivatives) and this file does what
// Julia's compiler does there: it compiles the tile, fix-up and VJP kernels of
// fvm_rhs_kernels.hpp with that source as JitPhysics<FVM_FLUX_JIT> through hipRTC
// for gfx950, loads the code object and hands back the kernel handles.  Modules are
// cached per (source, tile shape) for the life of the process.
#include <hip/hiprtc.h>

#include <cstdio>
#include <map>
#include <memory>
#include <mutex>
#include <string>

#include "fvm_internal.hpp"
#include "fvm_jit.hpp"

extern const char fvm_src_device[];
extern const char fvm_src_rhs_kernels[];

namespace fvm {
namespace {

std::mutex g_jit_mu;
std::map<std::string, std::unique_ptr<JitKernels>> g_jit_cache;

const char* kPrelude =
    "typedef signed char int8_t;\n"
    "typedef unsigned char uint8_t;\n"
    "typedef short int16_t;\n"
    "typedef unsigned short uint16_t;\n"
    "typedef int int32_t;\n"
    "typedef unsigned int uint32_t;\n"
    "typedef long long int64_t;\n"
    "typedef unsigned long long uint64_t;\n";

std::string variant(int block, int tpt) {
  return std::to_string(block) + ", " + std::to_string(tpt);
}

}  // namespace

// The closure kernels of user_src for the tile shape (block, tpt) compiled to a
// gfx950 code object; lowered[] receives the kernels' symbol names in the order
// tile RHS, tile JVP, fix-up RHS, fix-up JVP, VJP tile, VJP fix-up.
int jit_compile(const char* user_src, int block, int tpt, std::string* code, std::string lowered[6]) {
  char defs[512];
  std::snprintf(defs, sizeof defs,
                "#define FVM_GEO_LEVEL %d\n"
                "enum { FVM_FLUX_D_POLY_U = %d, FVM_FLUX_D_NODAL = %d, FVM_FLUX_D_EDGE = %d,\n"
                "       FVM_FLUX_LINEAR = %d, FVM_FLUX_D_POWER = %d, FVM_FLUX_JIT = %d };\n",
                FVM_GEO_LEVEL, FVM_FLUX_D_POLY_U, FVM_FLUX_D_NODAL, FVM_FLUX_D_EDGE, FVM_FLUX_LINEAR,
                FVM_FLUX_D_POWER, FVM_FLUX_JIT);
  const std::string v = variant(block, tpt);
  const std::string fk = std::to_string(FVM_FLUX_JIT);
  const std::string names[6] = {"fvm_k::tile_kernel<" + fk + ", 0, " + v + ">",
                                "fvm_k::tile_kernel<" + fk + ", 1, " + v + ">",
                                "fvm_k::fix_kernel<" + fk + ", 0>",
                                "fvm_k::fix_kernel<" + fk + ", 1>",
                                "fvm_k::vjp_tile_kernel<" + fk + ", " + v + ">",
                                "fvm_k::vjp_fix_kernel<" + fk + ">"};
  std::string src = std::string(kPrelude) + defs + "#include \"fvm_rhs_kernels.hpp\"\n" + user_src +
                    "\nnamespace fvm_k {\n"
                    "template <> struct JitPhysics<FVM_FLUX_JIT> {\n"
                    "  __device__ static double D(double x, double y, double t, double u) { return fvm_jit_D(x, y, t, u); }\n"
                    "  __device__ static double dD(double x, double y, double t, double u) { return fvm_jit_dD(x, y, t, u); }\n"
                    "  __device__ static double S(double x, double y, double t, double u) { return fvm_jit_S(x, y, t, u); }\n"
                    "  __device__ static double dS(double x, double y, double t, double u) { return fvm_jit_dS(x, y, t, u); }\n"
                    "};\n";
  for (const std::string& n : names) src += "template __global__ void " + n + "(KArgs);\n";
  src += "}  // namespace fvm_k\n";

  const char* hdrs[2] = {fvm_src_device, fvm_src_rhs_kernels};
  const char* hnames[2] = {"fvm_device.hpp", "fvm_rhs_kernels.hpp"};
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "fvm_jit.hip", 2, hdrs, hnames) != HIPRTC_SUCCESS) {
    set_error("jit: hiprtcCreateProgram failed");
    return FVM_ERR_HIP;
  }
  for (const std::string& n : names) hiprtcAddNameExpression(prog, n.c_str());
  // the library's own flags: no FMA contraction (the reference's rounding), no
  // packed FP32 instructions
  const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  hiprtcResult r = hiprtcCompileProgram(prog, 8, opts);
  if (r != HIPRTC_SUCCESS) {
    size_t n = 0;
    hiprtcGetProgramLogSize(prog, &n);
    std::string log(n, '\0');
    if (n) hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    set_error("jit: compiling the closure kernels failed: " + log.substr(0, 4000));
    return FVM_ERR_ARG;
  }
  for (int k = 0; k < 6; ++k) {
    const char* ln = nullptr;
    if (hiprtcGetLoweredName(prog, names[k].c_str(), &ln) != HIPRTC_SUCCESS || !ln) {
      hiprtcDestroyProgram(&prog);
      set_error("jit: no lowered name for " + names[k]);
      return FVM_ERR_HIP;
    }
    lowered[k] = ln;
  }
  size_t code_size = 0;
  hiprtcGetCodeSize(prog, &code_size);
  code->assign(code_size, '\0');
  hiprtcGetCode(prog, &(*code)[0]);
  hiprtcDestroyProgram(&prog);
  return FVM_OK;
}

int jit_kernels(const char* user_src, int tile_tris, const JitKernels** out) {
  FVM_CHECK(user_src && out, "jit: null argument");
  const int block = tile_tris <= 512 ? 256 : 512;
  const int tpt = tile_tris <= 256 ? 1 : 2;
  const std::string key = std::to_string(block) + "/" + std::to_string(tpt) + "\n" + user_src;
  std::lock_guard<std::mutex> lk(g_jit_mu);
  auto it = g_jit_cache.find(key);
  if (it != g_jit_cache.end()) {
    *out = it->second.get();
    return FVM_OK;
  }
  std::string code, lowered[6];
  int rc = jit_compile(user_src, block, tpt, &code, lowered);
  if (rc != FVM_OK) return rc;
  auto jk = std::make_unique<JitKernels>();
  jk->block = block;
  jk->tpt = tpt;
  FVM_HIP(hipModuleLoadData(&jk->module, code.data()));
  hipFunction_t* fs[6] = {&jk->tile_rhs, &jk->tile_jvp, &jk->fix_rhs, &jk->fix_jvp, &jk->vjp_tile, &jk->vjp_fix};
  for (int k = 0; k < 6; ++k) FVM_HIP(hipModuleGetFunction(fs[k], jk->module, lowered[k].c_str()));
  *out = jk.get();
  g_jit_cache.emplace(key, std::move(jk));
  return FVM_OK;
}

int jit_launch(hipFunction_t f, unsigned grid, unsigned block, const void* args, hipStream_t s) {
  void* params[] = {const_cast<void*>(args)};
  FVM_HIP(hipModuleLaunchKernel(f, grid, 1, 1, block, 1, 1, 0, s, params, nullptr));
  return FVM_OK;
}

}  // namespace fvm

extern "C" int fvm_jit_check(const char* jit_source, int32_t tile_triangles, size_t* code_bytes) {
  FVM_CHECK(jit_source, "fvm_jit_check: null source");
  const int tt = tile_triangles > 0 ? tile_triangles : 512;
  std::string code, lowered[6];
  int rc = fvm::jit_compile(jit_source, tt <= 512 ? 256 : 512, tt <= 256 ? 1 : 2, &code, lowered);
  if (rc == FVM_OK && code_bytes) *code_bytes = code.size();
  return rc;
}
