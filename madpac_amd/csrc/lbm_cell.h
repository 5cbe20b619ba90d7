// Per-cell D3Q19 arithmetic shared by the dense, patch and adjoint kernels (device code).
//
// Two collision paths:
//  * collide_f64: FP64, the reference's exact operation order so results are bit-identical
//    to the reference CPU build (computeDensity/computeVelocity/computeFeq/BlockBGK::collide,
//    BLB/BlockComputeCellValues.cpp:9-68, BLB/collisionmodels/BlockBGK.cpp:43).  Every file that
//    includes this header is compiled with -ffp-contract=off so no multiply-add is fused.
//  * collide_f32s: FP32 on shifted populations g = f - w_q (rest density removed), which keeps
//    the small deviations from equilibrium at full single precision.
// Citations: LB/ = madpac/peano/applications/latticeboltzmann/, BLB/ = LB/blocklatticeboltzmann/.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbm {

constexpr int Q = 19;
constexpr int REST = 9;
constexpr uint32_t FLUID = 0x80000000u;
constexpr int CLASS_SHIFT = 24;
constexpr uint32_t CLASS_MASK = 0x7Fu;
constexpr int MAX_CLASSES = 8;

// LB/LatticeBoltzmannDefinitions.cpp:113-131 (z-major; q' = 18 - q)
__host__ __device__ constexpr int CX(int q) {
  return (q == 3 || q == 7 || q == 10 || q == 13 || q == 17) ? 1
       : (q == 1 || q == 5 || q == 8 || q == 11 || q == 15) ? -1 : 0;
}
__host__ __device__ constexpr int CY(int q) {
  return (q == 4 || q == 11 || q == 12 || q == 13 || q == 18) ? 1
       : (q == 0 || q == 5 || q == 6 || q == 7 || q == 14) ? -1 : 0;
}
__host__ __device__ constexpr int CZ(int q) { return q <= 4 ? -1 : (q >= 14 ? 1 : 0); }
// LB/LatticeBoltzmannDefinitions.cpp:149-151
__host__ __device__ constexpr double W64(int q) {
  return q == REST ? 1.0 / 3.0
       : (q == 2 || q == 6 || q == 8 || q == 10 || q == 12 || q == 16) ? 1.0 / 18.0 : 1.0 / 36.0;
}
__host__ __device__ constexpr float W32(int q) {
  return q == REST ? 1.0f / 3.0f
       : (q == 2 || q == 6 || q == 8 || q == 10 || q == 12 || q == 16) ? 1.0f / 18.0f : 1.0f / 36.0f;
}

// Moving-wall classes and body force, uploaded per context into constant memory slots that
// the kernels receive by value.
struct WallTable {
  double uw[MAX_CLASSES][3];
  uint32_t mask[MAX_CLASSES];
};

struct CollideParams {
  double tau_eff;   // LB/MultiLevelSimData.h:160
  double omega;     // 1.0 / tau_eff, as BLB/collisionmodels/BlockBGK.cpp:43 evaluates it
  double g[3];      // body acceleration (Guo forcing)
};

// c_q . v as tarch::la::dot evaluates it (tarch/la/VectorVectorOperations.cpph:194-209: the
// products in d order, summed left to right).  With c in {-1,0,1} every product is exact, and
// a zero product never changes a non-zero partial sum, so the zero terms are skipped; only the
// sign of an all-zero result can differ, and no caller is sensitive to the sign of zero.
__device__ __forceinline__ double cdot64(int q, double v0, double v1, double v2) {
  double r = 0.0;
  if (CX(q) == 1) r += v0;
  if (CX(q) == -1) r -= v0;
  if (CY(q) == 1) r += v1;
  if (CY(q) == -1) r -= v1;
  if (CZ(q) == 1) r += v2;
  if (CZ(q) == -1) r -= v2;
  return r;
}

// ---------------------------------------------------------------------------------------------
// FP64, reference operation order.  f[] in, post-collision f*[] out (in place).  Returns rho.
// FORCE: Guo forcing with equilibrium velocity u + g/2.  ALPHA: equilibrium velocity scaled by
// (1 - alpha).  Both go beyond the reference and match the oracle's collide_cell.
template <bool FORCE, bool ALPHA>
__device__ __forceinline__ double collide_f64(double (&f)[Q], const CollideParams& p, double alpha) {
  // computeDensity (BLB/BlockComputeCellValues.cpp:36): sequential sum from 0.0
  double rho = 0.0;
#pragma unroll
  for (int q = 0; q < Q; q++) rho += f[q];
  // computeVelocity (:57-61): buffer += f_q c_q in q order, then buffer * (1.0/rho)
  // (the products with c = 0 add +-0 to a partial sum that is never -0, so they are skipped)
  double b0 = 0.0, b1 = 0.0, b2 = 0.0;
#pragma unroll
  for (int q = 0; q < Q; q++) {
    if (CX(q) == 1) b0 += f[q];
    if (CX(q) == -1) b0 -= f[q];
    if (CY(q) == 1) b1 += f[q];
    if (CY(q) == -1) b1 -= f[q];
    if (CZ(q) == 1) b2 += f[q];
    if (CZ(q) == -1) b2 -= f[q];
  }
  const double inv = 1.0 / rho;
  double u0 = b0 * inv, u1 = b1 * inv, u2 = b2 * inv;
  if (FORCE) {
    u0 = u0 + 0.5 * p.g[0];
    u1 = u1 + 0.5 * p.g[1];
    u2 = u2 + 0.5 * p.g[2];
  }
  if (ALPHA) {
    u0 = (1.0 - alpha) * u0;
    u1 = (1.0 - alpha) * u1;
    u2 = (1.0 - alpha) * u2;
  }
  // computeFeq (:14-20) and BGK (BlockBGK.cpp:43)
  double uu = u0 * u0;
  uu += u1 * u1;
  uu += u2 * u2;
  const double omega = p.omega;
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const double cu = cdot64(q, u0, u1, u2);
    const double feq = W64(q) * rho *
        (1.0 + cu / (1.0 / 3.0) + (cu * cu) / (2.0 * (1.0 / 3.0) * (1.0 / 3.0)) - (uu / (2.0 * (1.0 / 3.0))));
    f[q] -= omega * (f[q] - feq);
  }
  if (FORCE) {
    const double pref = (1.0 - 0.5 / p.tau_eff) * rho;
#pragma unroll
    for (int q = 0; q < Q; q++) {
      const double m0 = (double)CX(q) - u0, m1 = (double)CY(q) - u1, m2 = (double)CZ(q) - u2;
      double cmg = m0 * p.g[0];
      cmg += m1 * p.g[1];
      cmg += m2 * p.g[2];
      const double cu = cdot64(q, u0, u1, u2);
      const double cg = cdot64(q, p.g[0], p.g[1], p.g[2]);
      f[q] += pref * W64(q) * (cmg / (1.0 / 3.0) + cu * cg / ((1.0 / 3.0) * (1.0 / 3.0)));
    }
  }
  return rho;
}

// BLB/boundaryconditions/BlockMovingWall.cpp:23-28 for outgoing direction q:
// ciVel = u_w . c_q; ciVel = ciVel*2.0*w_q*rho/cs2.
__device__ __forceinline__ double moving_wall_f64(int q, const double* uw, double rho) {
  double civel = cdot64(q, uw[0], uw[1], uw[2]);
  civel = civel * 2.0 * W64(q) * rho / (1.0 / 3.0);
  return civel;
}

// ---------------------------------------------------------------------------------------------
// FP32 on shifted populations g_q = f_q - w_q.  Returns rho (as float).
struct CollideParams32 {
  float omega;
  float fpref;    // 1 - omega/2 (Guo)
  float g[3];
};

template <bool FORCE, bool ALPHA>
__device__ __forceinline__ float collide_f32s(float (&g)[Q], const CollideParams32& p, float alpha) {
  float drho = 0.f, j0 = 0.f, j1 = 0.f, j2 = 0.f;
#pragma unroll
  for (int q = 0; q < Q; q++) {
    drho += g[q];
    if (CX(q) == 1) j0 += g[q];
    if (CX(q) == -1) j0 -= g[q];
    if (CY(q) == 1) j1 += g[q];
    if (CY(q) == -1) j1 -= g[q];
    if (CZ(q) == 1) j2 += g[q];
    if (CZ(q) == -1) j2 -= g[q];
  }
  const float rho = 1.0f + drho;
  const float inv = 1.0f / rho;
  float u0 = j0 * inv, u1 = j1 * inv, u2 = j2 * inv;
  if (FORCE) {
    u0 += 0.5f * p.g[0];
    u1 += 0.5f * p.g[1];
    u2 += 0.5f * p.g[2];
  }
  if (ALPHA) {
    const float s = 1.0f - alpha;
    u0 *= s; u1 *= s; u2 *= s;
  }
  const float uu15 = 1.5f * (u0 * u0 + u1 * u1 + u2 * u2);
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const float cu = (float)CX(q) * u0 + (float)CY(q) * u1 + (float)CZ(q) * u2;
    // feq - w = w (drho + rho (3 cu + 4.5 cu^2 - 1.5 u^2))
    const float geq = W32(q) * (drho + rho * (3.0f * cu + 4.5f * cu * cu - uu15));
    float v = g[q] - p.omega * (g[q] - geq);
    if (FORCE) {
      const float cg = (float)CX(q) * p.g[0] + (float)CY(q) * p.g[1] + (float)CZ(q) * p.g[2];
      const float ug = u0 * p.g[0] + u1 * p.g[1] + u2 * p.g[2];
      v += p.fpref * W32(q) * rho * (3.0f * (cg - ug) + 9.0f * cu * cg);
    }
    g[q] = v;
  }
  return rho;
}

__device__ __forceinline__ float moving_wall_f32(int q, const double* uw, float rho) {
  const float cu = (float)CX(q) * (float)uw[0] + (float)CY(q) * (float)uw[1] + (float)CZ(q) * (float)uw[2];
  return 6.0f * W32(q) * rho * cu;
}

}  // namespace lbm
