// Device code shared by the precompiled kernels and the kernels compiled at run time
// (hipRTC) for closures outside the traced catalogue: this header includes nothing
// when compiled by hipRTC, whose prelude supplies the fixed-width integer types.
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>

#include <cstdint>
#endif

namespace fvm {

// Number of XCDs of an MI355X; consecutive workgroups are dispatched round-robin
// over them, each XCD with its own L2.
constexpr int kXCD = 8;

// Element i of an array the whole launch shares, addressed by a 32-bit byte offset:
// the access takes its base from scalar registers and the offset from one VGPR, with
// no 64-bit address arithmetic per lane.  fvm_mesh_create (and fvm_system_create for
// a system's state) keep the arrays the tile kernels address this way below 4 GiB.
template <class T>
__device__ __forceinline__ T ld32(const T* base, uint32_t i) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + i * (uint32_t)sizeof(T));
}
template <class T>
__device__ __forceinline__ void st32(T* base, uint32_t i, T v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + i * (uint32_t)sizeof(T)) = v;
}
__device__ __forceinline__ uint32_t umin(uint32_t x, uint32_t y) { return x < y ? x : y; }

// Maps workgroup id to a work index so that each XCD processes one contiguous
// range of work items (neighbouring tiles share nodes, hence L2 lines).  The
// grid must hold kXCD * ceil(n / kXCD) workgroups; ids >= n are idle.
__device__ __forceinline__ int xcd_remap(int bid, int nblocks) {
  int per = nblocks / kXCD;
  return (bid % kXCD) * per + bid / kXCD;
}
// The fix-up kernels' workgroup -> range of fix nodes: XCD-contiguous when the grid is
// a multiple of kXCD (the fix nodes go in the order of their first tile, so each XCD
// then reads the partial sums its own tiles wrote), the identity otherwise.
__device__ __forceinline__ int fix_wid() {
  return gridDim.x % kXCD == 0 ? xcd_remap((int)blockIdx.x, (int)gridDim.x) : (int)blockIdx.x;
}
// Per-triangle record in the device SoA: field f of sorted triangle s lives at
// props[f * T + s].  Fields follow TriangleProperties (geometry.jl:21-26).
enum : int {
  P_S1 = 0,   // s1..s9 at 0..8
  P_MX = 9,   // CV-edge midpoints m1x m1y m2x m2y m3x m3y at 9..14
  P_NX = 15,  // normals n1x n1y n2x n2y n3x n3y at 15..20
  P_L = 21,   // lengths l1 l2 l3 at 21..23
  P_COUNT = 24
};

// a / b correctly rounded, given y = RN(1 / b): Markstein's correction step
// (P. Markstein, IBM J. Res. Dev. 34(1), 1990; Muller et al., Handbook of
// Floating-Point Arithmetic, 2nd ed., 2018, sec. 4.9): q0 = RN(a y), the remainder
// r = a - b q0 exactly by one FMA, then RN(q0 + r y).  Equal to a / b for finite
// operands away from overflow and underflow (checked against IEEE division on 2e8
// random pairs and on every quotient of the geometry of 4M-node meshes).
__device__ __forceinline__ double div_rn(double a, double b, double y) {
  const double q0 = a * y;
  const double r = __builtin_fma(-b, q0, a);
  return __builtin_fma(r, y, q0);
}

// The triangle record (TriangleProperties, geometry.jl:99-169) of vertices p, q, r,
// with the geometry kernel's operations in the reference's order, so every value
// equals the stored one bit for bit up to LEVEL 3.  LEVEL >= 2: s1..s6; always
// s7..s9 and the CV-edge midpoints; LEVEL 3: normals and lengths; LEVEL 4: the
// quotients by Delta and by 3 from one reciprocal (div_rn, still bit-identical),
// and the CV-edge normals unnormalised with lengths 1 ((q . n) l = q . (e_y, -e_x)).
template <int LEVEL>
__device__ __forceinline__ void rebuild_record(double px, double py, double qx, double qy,
                                               double rx, double ry, double* P) {
  constexpr double kThird = 1.0 / 3;
  const double sx = px + qx + rx, sy = py + qy + ry;
  const double cx = LEVEL >= 4 ? div_rn(sx, 3.0, kThird) : sx / 3;
  const double cy = LEVEL >= 4 ? div_rn(sy, 3.0, kThird) : sy / 3;
  const double m1x = (px + qx) / 2, m1y = (py + qy) / 2;
  const double m2x = (qx + rx) / 2, m2y = (qy + ry) / 2;
  const double m3x = (rx + px) / 2, m3y = (ry + py) / 2;
  const double D = qx * ry - qy * rx - px * ry + rx * py + px * qy - qx * py;
  const double yD = LEVEL >= 4 ? 1.0 / D : 0.0;
  auto quot = [&](double a) { return LEVEL >= 4 ? div_rn(a, D, yD) : a / D; };
  if (LEVEL >= 2) {
    P[0] = quot(qy - ry);
    P[1] = quot(ry - py);
    P[2] = quot(py - qy);
    P[3] = quot(rx - qx);
    P[4] = quot(px - rx);
    P[5] = quot(qx - px);
  }
  P[6] = quot(qx * ry - rx * qy);
  P[7] = quot(rx * py - px * ry);
  P[8] = quot(px * qy - qx * py);
  P[9] = (m1x + cx) / 2;
  P[10] = (m1y + cy) / 2;
  P[11] = (m2x + cx) / 2;
  P[12] = (m2y + cy) / 2;
  P[13] = (m3x + cx) / 2;
  P[14] = (m3y + cy) / 2;
  if (LEVEL >= 4) {
    // (q . n) l with n = (e_y, -e_x) / l is q . (e_y, -e_x): keep the normal
    // unnormalised and the length 1, saving six divisions and three square roots
    const double e1x = cx - m1x, e1y = cy - m1y;
    const double e2x = cx - m2x, e2y = cy - m2y;
    const double e3x = cx - m3x, e3y = cy - m3y;
    P[15] = e1y;
    P[16] = -e1x;
    P[17] = e2y;
    P[18] = -e2x;
    P[19] = e3y;
    P[20] = -e3x;
    P[21] = 1.0;
    P[22] = 1.0;
    P[23] = 1.0;
  } else if (LEVEL >= 3) {
    const double e1x = cx - m1x, e1y = cy - m1y;
    const double e2x = cx - m2x, e2y = cy - m2y;
    const double e3x = cx - m3x, e3y = cy - m3y;
    const double l1 = sqrt(e1x * e1x + e1y * e1y);
    const double l2 = sqrt(e2x * e2x + e2y * e2y);
    const double l3 = sqrt(e3x * e3x + e3y * e3y);
    P[15] = e1y / l1;
    P[16] = -e1x / l1;
    P[17] = e2y / l2;
    P[18] = -e2x / l2;
    P[19] = e3y / l3;
    P[20] = -e3x / l3;
    P[21] = l1;
    P[22] = l2;
    P[23] = l3;
  }
}

// Condition functions a(x,y,t,u) = c0 + cu u + ct t + cx x + cy y + g, with g
// multiplied by exp(lam t) where lam (may be null) is nonzero.
struct CondTable {
  int32_t n;
  const double* c0;
  const double* cu;
  const double* ct;
  const double* cx;
  const double* cy;
  const double* lam;      // may be null
  const double* node_g;   // may be null
  const double* bedge_g;  // may be null
};

__device__ __forceinline__ double eval_cond(const CondTable& C, int f, double x, double y,
                                            double t, double u, double g) {
  double v = C.c0[f] + C.cu[f] * u;
  v = v + C.ct[f] * t;
  v = v + C.cx[f] * x;
  v = v + C.cy[f] * y;
  if (C.lam && C.lam[f] != 0.0) return v + g * exp(C.lam[f] * t);
  return v + g;
}

}  // namespace fvm
