// Run-time compiled kernels of the FVM_FLUX_JIT closure form (fvm_jit.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace fvm {

// Handles of one compiled module: the tile kernel (RHS and JVP modes), the fix-up
// (RHS and JVP) and the VJP pair, for tiles of `block` threads x `tpt` triangles.
struct JitKernels {
  hipModule_t module = nullptr;
  hipFunction_t tile_rhs = nullptr, tile_jvp = nullptr, fix_rhs = nullptr, fix_jvp = nullptr;
  hipFunction_t vjp_tile = nullptr, vjp_fix = nullptr;
  int block = 256, tpt = 2;
};

// Compiles (or finds in the process cache) the kernels for the closures in
// user_src (definitions of fvm_jit_D, fvm_jit_dD, fvm_jit_S, fvm_jit_dS) and the
// tile shape that holds tile_tris triangles.  The module lives until the process ends.
int jit_kernels(const char* user_src, int tile_tris, const JitKernels** out);
// Launches f on a 1-D grid with one by-value argument struct.
int jit_launch(hipFunction_t f, unsigned grid, unsigned block, const void* args, hipStream_t s);

}  // namespace fvm
