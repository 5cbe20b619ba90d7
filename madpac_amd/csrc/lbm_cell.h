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
constexpr int CLASS_SHIFT = 24;       // bits 24..28 of the cell word: boundary class
constexpr int MAX_CLASSES = 32;       // class 0 = plain halfway bounce-back
constexpr uint32_t CLASS_BITS = MAX_CLASSES - 1;

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

// Boundary classes.  A fluid cell's class selects one condition and its target state; the
// class mask says which of the cell's bounce-back links take it (the others stay halfway
// bounce-back).  Types follow LB/BoundaryData.h:44-61 and the handlers registered in
// BLB/boundaryconditions/BlockBoundaryConditionManager.cpp:15-20.  BC_CELL_VELOCITY or-ed into
// the type takes the target velocity per cell from a separate [3*N] array.  The table lives in
// device memory (MAX_CLASSES entries) and is read only by cells that carry a class.
enum : uint32_t { BC_MOVING_WALL = 0, BC_PRESSURE_OUTLET = 1, BC_PRESSURE_INLET = 2, BC_VELOCITY_INLET = 3 };
constexpr uint32_t BC_TYPE_MASK = 15u;
constexpr uint32_t BC_CELL_VELOCITY = 16u;

struct BoundaryClass {
  double u[3];     // wall velocity u_w, or target velocity u_L
  double rho;      // target density rho_L
  uint32_t mask;   // bit q: outgoing link q (a bounce-back link of the cell) takes this condition
  uint32_t type;   // BC_* | BC_CELL_VELOCITY
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
// uo receives the cell velocity as the boundary conditions read it: (sum f c) / rho, plus g/2
// with a body force, before any porosity scaling.
template <bool FORCE, bool ALPHA>
__device__ __forceinline__ double collide_f64(double (&f)[Q], const CollideParams& p, double alpha,
                                              double (&uo)[3]) {
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
  uo[0] = u0;
  uo[1] = u1;
  uo[2] = u2;
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
    // Guo source (1 - 1/(2 tau)) w rho [3 (c - u).g + 9 (c.u)(c.g)] with cs2 = 1/3 folded into the
    // factors (not a reference routine: the multiplications replace two divisions per population)
    const double pref = (1.0 - 0.5 / p.tau_eff) * rho;
    double ug = u0 * p.g[0];
    ug += u1 * p.g[1];
    ug += u2 * p.g[2];
#pragma unroll
    for (int q = 0; q < Q; q++) {
      const double cu = cdot64(q, u0, u1, u2);
      const double cg = cdot64(q, p.g[0], p.g[1], p.g[2]);
      f[q] += pref * W64(q) * (3.0 * (cg - ug) + 9.0 * cu * cg);
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

// uo receives the velocity the boundary conditions read (as in collide_f64) and, in uo[3],
// the density deviation rho - 1 as summed from the shifted populations.
template <bool FORCE, bool ALPHA>
__device__ __forceinline__ float collide_f32s(float (&g)[Q], const CollideParams32& p, float alpha,
                                              float (&uo)[4]) {
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
  uo[0] = u0;
  uo[1] = u1;
  uo[2] = u2;
  uo[3] = drho;
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

// FP32-shifted storage, hybrid arithmetic: the moments, rho, u and each equilibrium value in FP64
// (the parts whose repeated FP32 rounding biases the momentum), the difference g - geq and the
// force term in FP32, the product omega (g - geq) in FP64: an FP32 omega would change the
// viscosity by up to ~1e-6 relative near tau = 0.5, where nu = (1/omega - 1/2) / 3 is small.
// Against an all-FP64 evaluation of the same update (widen, collide, round once at the store) it
// holds fewer FP64 values live: 60 instead of 76 VGPRs in the dense step, 8 instead of 6 waves per
// SIMD, +6 % throughput.  Outputs as collide_f32s.
template <bool FORCE, bool ALPHA>
__device__ __forceinline__ float collide_f32s_h(float (&gs)[Q], const CollideParams& p, const CollideParams32& p32,
                                                float alpha, float (&uo)[4]) {
  double drho = 0.0, j0 = 0.0, j1 = 0.0, j2 = 0.0;
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const double g = gs[q];
    drho += g;
    if (CX(q) == 1) j0 += g;
    if (CX(q) == -1) j0 -= g;
    if (CY(q) == 1) j1 += g;
    if (CY(q) == -1) j1 -= g;
    if (CZ(q) == 1) j2 += g;
    if (CZ(q) == -1) j2 -= g;
  }
  const double rho = 1.0 + drho;
  const double inv = 1.0 / rho;
  double u0 = j0 * inv, u1 = j1 * inv, u2 = j2 * inv;
  if (FORCE) {
    u0 += 0.5 * p.g[0];
    u1 += 0.5 * p.g[1];
    u2 += 0.5 * p.g[2];
  }
  uo[0] = (float)u0;
  uo[1] = (float)u1;
  uo[2] = (float)u2;
  uo[3] = (float)drho;
  if (ALPHA) {
    const double s = 1.0 - (double)alpha;
    u0 *= s; u1 *= s; u2 *= s;
  }
  const double uu15 = 1.5 * (u0 * u0 + u1 * u1 + u2 * u2);
  const float rho32 = (float)rho, v0 = (float)u0, v1 = (float)u1, v2 = (float)u2;
  const float ug = FORCE ? v0 * p32.g[0] + v1 * p32.g[1] + v2 * p32.g[2] : 0.f;
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const double cu = cdot64(q, u0, u1, u2);
    const float geq = (float)(W64(q) * (drho + rho * (3.0 * cu + 4.5 * cu * cu - uu15)));
    float v = gs[q] - (float)(p.omega * (double)(gs[q] - geq));
    if (FORCE) {
      const float cg = (float)CX(q) * p32.g[0] + (float)CY(q) * p32.g[1] + (float)CZ(q) * p32.g[2];
      const float cuf = (float)cu;
      v += p32.fpref * W32(q) * rho32 * (3.0f * (cg - ug) + 9.0f * cuf * cg);
    }
    gs[q] = v;
  }
  return rho32;
}

__device__ __forceinline__ float moving_wall_f32(int q, const double* uw, float rho) {
  const float cu = (float)CX(q) * (float)uw[0] + (float)CY(q) * (float)uw[1] + (float)CZ(q) * (float)uw[2];
  return 6.0f * W32(q) * rho * cu;
}

// ---------------------------------------------------------------------------------------------
// Boundary links of a classed cell, applied to the post-collision populations of the outgoing
// slots q in m (the swap streaming later moves each into the incoming slot 18 - q):
//  moving wall      f_q - 2 w_q rho (u_w . c_q) / cs2           BLB/boundaryconditions/BlockMovingWall.cpp:23-28
//  pressure in/out  -f_q + feq_{18-q} + feq_q of (rho_t, u_t)    BlockAbstractPressureCondition.cpp:24-37;
//                   outlet: rho_t = rho_L, u_t = cell velocity    BlockPressureOutletCondition.h:32-50
//                   inlet:  rho_t = cell density, u_t = u_L       BlockPressureInletCondition.h:32-50
//  velocity inlet   feq_{18-q}(rho_L, u_L)                        BlockVelocityInletEquilibrium.cpp:19-23
// computeFeq (BLB/BlockComputeCellValues.cpp:14-20) for one direction, reference order:
__device__ __forceinline__ double feq_f64(int q, double rho, double u0, double u1, double u2, double uu) {
  const double cu = cdot64(q, u0, u1, u2);
  return W64(q) * rho *
         (1.0 + cu / (1.0 / 3.0) + (cu * cu) / (2.0 * (1.0 / 3.0) * (1.0 / 3.0)) - (uu / (2.0 * (1.0 / 3.0))));
}

// (the cell velocity comes by value: a pointer select between it and the table would force it
// into scratch memory)
__device__ __forceinline__ void apply_boundary_f64(double (&f)[Q], uint32_t m, const BoundaryClass& b,
                                                   const double* __restrict__ bc_u, int64_t i, double rho,
                                                   double u0, double u1, double u2) {
  const uint32_t ty = b.type & BC_TYPE_MASK;
  if (ty == BC_MOVING_WALL) {
#pragma unroll
    for (int q = 0; q < Q; q++)
      if ((m >> q) & 1u) f[q] -= moving_wall_f64(q, b.u, rho);
    return;
  }
  const double* ub = (b.type & BC_CELL_VELOCITY) ? bc_u + 3 * i : b.u;
  double rt, t0, t1, t2;
  if (ty == BC_PRESSURE_OUTLET) {
    rt = b.rho; t0 = u0; t1 = u1; t2 = u2;
  } else {
    rt = ty == BC_PRESSURE_INLET ? rho : b.rho;
    t0 = ub[0]; t1 = ub[1]; t2 = ub[2];
  }
  double uu = t0 * t0;
  uu += t1 * t1;
  uu += t2 * t2;
  const bool vin = ty == BC_VELOCITY_INLET;
#pragma unroll
  for (int q = 0; q < Q; q++) {
    if ((m >> q) & 1u) {
      const double fin = feq_f64(Q - 1 - q, rt, t0, t1, t2, uu);
      f[q] = vin ? fin : -f[q] + fin + feq_f64(q, rt, t0, t1, t2, uu);
    }
  }
}

// FP32 shifted form: with geq_q = feq_q - w_q = w_q (drho_t + rho_t (3 cu + 4.5 cu^2 - 1.5 u^2))
// and w_q = w_{18-q}, the pressure rule is g_q <- -g_q + geq_{18-q} + geq_q and the velocity
// inlet g_q <- geq_{18-q}.
__device__ __forceinline__ void apply_boundary_f32(float (&g)[Q], uint32_t m, const BoundaryClass& b,
                                                   const double* __restrict__ bc_u, int64_t i, float rho,
                                                   float drho, float u0, float u1, float u2) {
  const uint32_t ty = b.type & BC_TYPE_MASK;
  if (ty == BC_MOVING_WALL) {
#pragma unroll
    for (int q = 0; q < Q; q++)
      if ((m >> q) & 1u) g[q] -= moving_wall_f32(q, b.u, rho);
    return;
  }
  const double* ub = (b.type & BC_CELL_VELOCITY) ? bc_u + 3 * i : b.u;
  float rt, drt, t0, t1, t2;
  if (ty == BC_PRESSURE_OUTLET) {
    rt = (float)b.rho; drt = (float)(b.rho - 1.0); t0 = u0; t1 = u1; t2 = u2;
  } else {
    if (ty == BC_PRESSURE_INLET) {
      rt = rho; drt = drho;
    } else {
      rt = (float)b.rho; drt = (float)(b.rho - 1.0);
    }
    t0 = (float)ub[0]; t1 = (float)ub[1]; t2 = (float)ub[2];
  }
  const float uu15 = 1.5f * (t0 * t0 + t1 * t1 + t2 * t2);
  const bool vin = ty == BC_VELOCITY_INLET;
#pragma unroll
  for (int q = 0; q < Q; q++) {
    if ((m >> q) & 1u) {
      const float cu = (float)CX(q) * t0 + (float)CY(q) * t1 + (float)CZ(q) * t2;
      // W_q = W_{18-q}; c_{18-q} . u = -cu
      const float gin = W32(q) * (drt + rt * (-3.0f * cu + 4.5f * cu * cu - uu15));
      const float gout = W32(q) * (drt + rt * (3.0f * cu + 4.5f * cu * cu - uu15));
      g[q] = vin ? gin : -g[q] + gin + gout;
    }
  }
}

}  // namespace lbm
