// MD <-> LB coupling hooks at a list of cells (the MaDPaC use of the block-LB solver).
//
// Reference: madpac/peanocoupling/PeanoDataExchangeService.cpp.  getAndSetV (:69-158) runs per
// block between computeVelocity and the collision: at coupled cells it either overwrites the LB
// velocity (and, with density coupling, the density) with the MD values, or, in the Dirichlet
// mode, computes g = (u_MD - u)/dt and shifts u by dt/2 g; adjustF (:189-218) then adds dt g_i
// (computeGi, :161-186) after the BGK relaxation; and it exports the LB density and velocity of
// the same cells to the MD side.  Here the fused step kernel collides every cell as usual, and a
// second small kernel re-collides only the coupled cells from the same source state and overwrites
// their output slots (own slots in the pull form; in the in-place layout the slots the step wrote,
// from populations saved before the step); the hot kernel is unchanged.  A gather kernel returns rho
// and u of the coupled cells as getAndSetV exports them (after its GET block).  Coupled cells use the reference's block-solver
// collision: no body force and no porosity (the host rejects those combinations).
#include <cmath>
#include <string>
#include <vector>

#include "lbm_ctx.h"
#include "lbm_geom.h"

namespace lbm {

struct CoupleArgs {
  const int64_t* cells;   // coupled cell indices
  int64_t n;
  const double* u;        // [3n] MD velocity (lattice units)
  const double* rho;      // [n] MD density or null
  int mode;               // LBM_COUPLE_*
  double dt;              // LB time step of the Dirichlet forcing
  double tau;             // configured tau (computeGi's _tau)
  double omega;           // 1 / tau_eff (BlockBGK)
};


// The GET block of getAndSetV (:100-129) for coupled cell k, on the cell's rho and u: overwrite sets
// u = u_MD (:121-122); Dirichlet computes g = (u_MD - u)/dt and shifts u += dt/2 g (:110-117); with
// density coupling rho = rho_MD (:124-129).  g stays 0 in the overwrite mode.
__device__ __forceinline__ void couple_get(const CoupleArgs& a, int64_t k, double& rho, double (&u)[3],
                                           double (&g)[3]) {
  const double* um = a.u + 3 * k;
  g[0] = g[1] = g[2] = 0.0;
  if (a.mode == LBM_COUPLE_DIRICHLET) {
    for (int d = 0; d < 3; d++) {
      g[d] = (um[d] - u[d]) / a.dt;
      u[d] += a.dt / 2.0 * g[d];
    }
  } else {
    for (int d = 0; d < 3; d++) u[d] = um[d];
  }
  if (a.rho) rho = a.rho[k];
}

// FP64 collision of one coupled cell (reference operation order, as the oracle restates it).
__device__ __forceinline__ double collide_coupled_f64(double (&f)[Q], const CoupleArgs& a, int64_t k,
                                                      double (&uo)[3]) {
  double rho = 0.0;
#pragma unroll
  for (int q = 0; q < Q; q++) rho += f[q];
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
  double u[3] = {b0 * inv, b1 * inv, b2 * inv};
  double g[3];
  couple_get(a, k, rho, u, g);
  const bool dir = a.mode == LBM_COUPLE_DIRICHLET;
  uo[0] = u[0];
  uo[1] = u[1];
  uo[2] = u[2];
  double uu = u[0] * u[0];
  uu += u[1] * u[1];
  uu += u[2] * u[2];
#pragma unroll
  for (int q = 0; q < Q; q++) f[q] -= a.omega * (f[q] - feq_f64(q, rho, u[0], u[1], u[2], uu));
  if (dir) {
    // adjustF / computeGi, with (c.u)^2 as the reference writes it
    const double pref = 1.0 - 0.5 / a.tau;
#pragma unroll
    for (int q = 0; q < Q; q++) {
      const double m0 = (double)CX(q) - u[0], m1 = (double)CY(q) - u[1], m2 = (double)CZ(q) - u[2];
      double cmg = m0 * g[0];
      cmg += m1 * g[1];
      cmg += m2 * g[2];
      const double cu = cdot64(q, u[0], u[1], u[2]);
      const double cg = cdot64(q, g[0], g[1], g[2]);
      const double gi = pref * W64(q) * rho *
                        (1.0 / (1.0 / 3.0) * cmg + (cu * cu) / ((1.0 / 3.0) * (1.0 / 3.0)) * cg);
      f[q] += a.dt * gi;
    }
  }
  return rho;
}

// Coupled collision of cell k (stored index i, flag word w), then the boundary class of its links.
template <bool WALLS>
__device__ __forceinline__ void couple_collide(double (&f)[Q], const CoupleArgs& a, int64_t k, int64_t i, uint32_t w,
                                               const BoundaryArgs& bc) {
  double u[3];
  const double rho = collide_coupled_f64(f, a, k, u);
  if (WALLS) {
    const uint32_t cls = (w >> CLASS_SHIFT) & CLASS_BITS;
    if (cls) {
      const BoundaryClass& b = bc.table[cls];
      const uint32_t m = w & b.mask;
      if (m) apply_boundary_f64(f, m, b, bc.cell_u, i, rho, u[0], u[1], u[2]);
    }
  }
}

// Re-collide the coupled cells from src and overwrite their own output slots in dst (run after the
// step kernel that wrote dst from src).  FP32-shifted contexts are lifted to FP64 for these cells.
template <typename T, bool WALLS>
__global__ __launch_bounds__(256) void k_couple(DenseGeom g, const T* __restrict__ src, T* __restrict__ dst,
                                                const uint32_t* __restrict__ words, CoupleArgs a, BoundaryArgs bc) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= a.n) return;
  const int64_t i = a.cells[k];
  const uint32_t w = words[i];
  if (!(w & FLUID)) return;
  const NbrOff o = nbr_offsets(g, pos_of(g, i));
  double f[Q];
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const T v = src[pull_index(g, q, w, i, o)];
    f[q] = sizeof(T) == 8 ? (double)v : (double)v + W64(q);
  }
  couple_collide<WALLS>(f, a, k, i, w, bc);
#pragma unroll
  for (int q = 0; q < Q; q++) dst[pop_at(g, q, i)] = sizeof(T) == 8 ? (T)f[q] : (T)(f[q] - W64(q));
}

// In-place (AA) layout: the step overwrites the populations the re-collision starts from, so they
// are saved before it: save[q * n + k] = population q of coupled cell k in the reference state, read
// through the layout the step starts from (g.layout).
template <typename T>
__global__ __launch_bounds__(256) void k_couple_save(DenseGeom g, const T* __restrict__ s,
                                                     const uint32_t* __restrict__ words, const int64_t* __restrict__ cells,
                                                     int64_t n, double* __restrict__ save) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  const int64_t i = cells[k];
  const uint32_t w = words[i];
  if (!(w & FLUID)) return;
  const NbrOff o = nbr_offsets(g, pos_of(g, i));
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const T v = s[state_index(g, q, w, i, o)];
    save[(int64_t)q * n + k] = sizeof(T) == 8 ? (double)v : (double)v + W64(q);
  }
}

// After the in-place step: re-collide from the saved populations and overwrite the slots the step
// wrote for this cell (k_aa_step): an even step the own slot 18 - q, an odd step the slot q of
// x + c_q, or the own slot 18 - q across a bounce-back link.
template <typename T, bool WALLS, bool ODD>
__global__ __launch_bounds__(256) void k_couple_aa(DenseGeom g, const double* __restrict__ save, T* __restrict__ s,
                                                   const uint32_t* __restrict__ words, CoupleArgs a, BoundaryArgs bc) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= a.n) return;
  const int64_t i = a.cells[k];
  const uint32_t w = words[i];
  if (!(w & FLUID)) return;
  const NbrOff o = nbr_offsets(g, pos_of(g, i));
  double f[Q];
#pragma unroll
  for (int q = 0; q < Q; q++) f[q] = save[(int64_t)q * a.n + k];
  couple_collide<WALLS>(f, a, k, i, w, bc);
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const bool bounce = (w >> q) & 1u;
    const int64_t di = (!ODD || bounce) ? pop_at(g, Q - 1 - q, i) : pop_at(g, q, i + dst_off(q, o));
    s[di] = sizeof(T) == 8 ? (T)f[q] : (T)(f[q] - W64(q));
  }
}

// Export (getAndSetV :136-151): rho and u = (sum f c) / rho of the reference state at the coupled
// cells, passed through the GET block when the coupling is on (`on`), as the reference sends the
// block's velocity and density after GET changed them; 0 at solid cells.
template <typename T>
__global__ __launch_bounds__(256) void k_couple_gather(DenseGeom g, const T* __restrict__ s,
                                                       const uint32_t* __restrict__ words, CoupleArgs a, int on,
                                                       double* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= a.n) return;
  const int64_t i = a.cells[k];
  const uint32_t w = words[i];
  double rho = 0.0, u0 = 0.0, u1 = 0.0, u2 = 0.0;
  if (w & FLUID) {
    const NbrOff o = nbr_offsets(g, pos_of(g, i));
    double f[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
      const T v = s[state_index(g, q, w, i, o)];
      f[q] = sizeof(T) == 8 ? (double)v : (double)v + W64(q);
    }
    double b0 = 0.0, b1 = 0.0, b2 = 0.0;
#pragma unroll
    for (int q = 0; q < Q; q++) {
      rho += f[q];
      if (CX(q) == 1) b0 += f[q];
      if (CX(q) == -1) b0 -= f[q];
      if (CY(q) == 1) b1 += f[q];
      if (CY(q) == -1) b1 -= f[q];
      if (CZ(q) == 1) b2 += f[q];
      if (CZ(q) == -1) b2 -= f[q];
    }
    const double inv = 1.0 / rho;
    double u[3] = {b0 * inv, b1 * inv, b2 * inv}, gd[3];
    if (on) couple_get(a, k, rho, u, gd);
    u0 = u[0];
    u1 = u[1];
    u2 = u[2];
  }
  out[4 * k] = rho;
  out[4 * k + 1] = u0;
  out[4 * k + 2] = u1;
  out[4 * k + 3] = u2;
}

}  // namespace lbm

using lbm::Q;

namespace {
int cfail(lbm_ctx* c, int code, const std::string& msg) {
  c->err = msg;
  return code;
}
}  // namespace

// Launch the re-collision of the coupled cells for the step src -> dst (called by the step loop).
int lbm_enqueue_coupling(lbm_ctx* c, const void* src, void* dst) {
  if (!c->couple_on || c->ncouple == 0) return LBM_OK;
  lbm::CoupleArgs a{c->d_ccells, c->ncouple, c->d_cu, c->couple_rho ? c->d_crho : nullptr, c->couple_mode,
                    c->couple_dt, c->cfg.tau, c->p64.omega};
  const lbm::BoundaryArgs bc{c->d_bct, c->d_bcu};
  const dim3 grid((unsigned)((c->ncouple + 255) / 256)), block(256);
  if (c->fp32) {
    if (c->walls)
      hipLaunchKernelGGL((lbm::k_couple<float, true>), grid, block, 0, c->stream, c->g, (const float*)src, (float*)dst,
                         c->words, a, bc);
    else
      hipLaunchKernelGGL((lbm::k_couple<float, false>), grid, block, 0, c->stream, c->g, (const float*)src,
                         (float*)dst, c->words, a, bc);
  } else {
    if (c->walls)
      hipLaunchKernelGGL((lbm::k_couple<double, true>), grid, block, 0, c->stream, c->g, (const double*)src,
                         (double*)dst, c->words, a, bc);
    else
      hipLaunchKernelGGL((lbm::k_couple<double, false>), grid, block, 0, c->stream, c->g, (const double*)src,
                         (double*)dst, c->words, a, bc);
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? LBM_OK : cfail(c, LBM_ERR_HIP, std::string("coupling kernel: ") + hipGetErrorString(e));
}

// In-place layout: save the coupled cells' populations before the step (layout of the step's start).
int lbm_enqueue_coupling_save(lbm_ctx* c) {
  if (!c->couple_on || c->ncouple == 0) return LBM_OK;
  const dim3 grid((unsigned)((c->ncouple + 255) / 256)), block(256);
  const lbm::DenseGeom g = lbm_state_geom(c);
  if (c->fp32)
    hipLaunchKernelGGL(lbm::k_couple_save<float>, grid, block, 0, c->stream, g, (const float*)c->buf[0], c->words,
                       c->d_ccells, c->ncouple, c->d_csave);
  else
    hipLaunchKernelGGL(lbm::k_couple_save<double>, grid, block, 0, c->stream, g, (const double*)c->buf[0], c->words,
                       c->d_ccells, c->ncouple, c->d_csave);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? LBM_OK : cfail(c, LBM_ERR_HIP, std::string("coupling save kernel: ") + hipGetErrorString(e));
}

// In-place layout: re-collide the coupled cells after the step of parity `parity` (0 even, 1 odd).
int lbm_enqueue_coupling_aa(lbm_ctx* c, int parity) {
  if (!c->couple_on || c->ncouple == 0) return LBM_OK;
  lbm::CoupleArgs a{c->d_ccells, c->ncouple, c->d_cu, c->couple_rho ? c->d_crho : nullptr, c->couple_mode,
                    c->couple_dt, c->cfg.tau, c->p64.omega};
  const lbm::BoundaryArgs bc{c->d_bct, c->d_bcu};
  const dim3 grid((unsigned)((c->ncouple + 255) / 256)), block(256);
#define LBM_COUPLE_AA(T, W, O)                                                                                   \
  hipLaunchKernelGGL((lbm::k_couple_aa<T, W, O>), grid, block, 0, c->stream, c->g, c->d_csave, (T*)c->buf[0],     \
                     c->words, a, bc)
  const int key = (c->fp32 ? 4 : 0) | (c->walls ? 2 : 0) | (parity ? 1 : 0);
  switch (key) {
    case 0: LBM_COUPLE_AA(double, false, false); break;
    case 1: LBM_COUPLE_AA(double, false, true); break;
    case 2: LBM_COUPLE_AA(double, true, false); break;
    case 3: LBM_COUPLE_AA(double, true, true); break;
    case 4: LBM_COUPLE_AA(float, false, false); break;
    case 5: LBM_COUPLE_AA(float, false, true); break;
    case 6: LBM_COUPLE_AA(float, true, false); break;
    default: LBM_COUPLE_AA(float, true, true); break;
  }
#undef LBM_COUPLE_AA
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? LBM_OK : cfail(c, LBM_ERR_HIP, std::string("coupling kernel: ") + hipGetErrorString(e));
}

void lbm_coupling_release(lbm_ctx* c) {
  // captured steps hold the buffers' addresses: drop them before the buffers go
  if (c->d_ccells) lbm_graphs_release(c);
  void* ptrs[] = {c->d_ccells, c->d_cu, c->d_crho, c->d_cout, c->d_csave};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  c->d_ccells = nullptr;
  c->d_cu = c->d_crho = c->d_cout = nullptr;
  c->d_csave = nullptr;
  c->ncouple = 0;
  c->couple_on = false;
}

extern "C" {

int lbm_set_coupling_cells(lbm_ctx* c, const int64_t* cells, int64_t n) {
  if (!c) return LBM_ERR_ARG;
  if (n < 0 || (n > 0 && !cells)) return cfail(c, LBM_ERR_ARG, "coupling cell list is null or negative");
  if (hipSetDevice(c->cfg.device) != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess)
    return cfail(c, LBM_ERR_HIP, "device synchronisation failed");
  lbm_coupling_release(c);
  if (n == 0) return LBM_OK;
  std::vector<uint8_t> seen((size_t)c->g.Nu, 0);
  for (int64_t k = 0; k < n; k++) {
    if (cells[k] < 0 || cells[k] >= c->g.Nu) return cfail(c, LBM_ERR_ARG, "coupling cell out of range");
    if (seen[(size_t)cells[k]]++) return cfail(c, LBM_ERR_ARG, "coupling cells must be distinct");
  }
  if (lbm::device_alloc(&c->d_ccells, (size_t)n * sizeof(int64_t)) != hipSuccess ||
      lbm::device_alloc(&c->d_cu, (size_t)n * 3 * sizeof(double)) != hipSuccess ||
      lbm::device_alloc(&c->d_crho, (size_t)n * sizeof(double)) != hipSuccess ||
      lbm::device_alloc(&c->d_cout, (size_t)n * 4 * sizeof(double)) != hipSuccess ||
      (c->aa && lbm::device_alloc(&c->d_csave, (size_t)n * Q * sizeof(double)) != hipSuccess)) {
    lbm_coupling_release(c);
    return cfail(c, LBM_ERR_NOMEM, "coupling buffers allocation failed");
  }
  // the kernels address the stored (bricked) cells
  const lbm::DenseGeom& g = c->g;
  std::vector<int64_t> stored((size_t)n);
  for (int64_t k = 0; k < n; k++) {
    const int64_t r = cells[k];
    const int x = (int)(r % g.nx), y = (int)((r / g.nx) % g.ny), z = (int)(r / ((int64_t)g.nx * g.ny));
    stored[(size_t)k] = lbm::stored_index(g, x, y, z);
  }
  if (hipMemcpy(c->d_ccells, stored.data(), (size_t)n * sizeof(int64_t), hipMemcpyHostToDevice) != hipSuccess) {
    lbm_coupling_release(c);
    return cfail(c, LBM_ERR_HIP, "coupling cell upload failed");
  }
  c->ncouple = n;
  return LBM_OK;
}

int lbm_set_coupling_data(lbm_ctx* c, int mode, const double* u, const double* rho, double dt) {
  if (!c) return LBM_ERR_ARG;
  if (!u) {
    c->couple_on = false;
    return LBM_OK;
  }
  if (c->ncouple == 0) return cfail(c, LBM_ERR_STATE, "no coupling cells (lbm_set_coupling_cells)");
  if (mode != LBM_COUPLE_OVERWRITE && mode != LBM_COUPLE_DIRICHLET) return cfail(c, LBM_ERR_ARG, "unknown coupling mode");
  if (mode == LBM_COUPLE_DIRICHLET && !(dt > 0.0 && std::isfinite(dt)))
    return cfail(c, LBM_ERR_ARG, "coupling time step must be positive");
  if (c->force || c->alpha)
    return cfail(c, LBM_ERR_STATE, "coupled cells use the plain collision: no body force or porosity");
  for (int64_t k = 0; k < 3 * c->ncouple; k++)
    if (!std::isfinite(u[k])) return cfail(c, LBM_ERR_ARG, "coupling velocities must be finite");
  if (rho)
    for (int64_t k = 0; k < c->ncouple; k++)
      if (!(rho[k] > 0.0) || !std::isfinite(rho[k])) return cfail(c, LBM_ERR_ARG, "coupling densities must be positive");
  if (hipSetDevice(c->cfg.device) != hipSuccess ||
      hipMemcpyAsync(c->d_cu, u, (size_t)c->ncouple * 3 * sizeof(double), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      (rho && hipMemcpyAsync(c->d_crho, rho, (size_t)c->ncouple * sizeof(double), hipMemcpyHostToDevice, c->stream) !=
                  hipSuccess) ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return cfail(c, LBM_ERR_HIP, "coupling data upload failed");
  c->couple_mode = mode;
  c->couple_dt = dt;
  c->couple_rho = rho != nullptr;
  c->couple_on = true;
  return LBM_OK;
}

int lbm_get_coupling_data(lbm_ctx* c, double* u, double* rho) {
  if (!c) return LBM_ERR_ARG;
  if (c->ncouple == 0) return cfail(c, LBM_ERR_STATE, "no coupling cells (lbm_set_coupling_cells)");
  if (!c->flags_set || !c->state_set) return cfail(c, LBM_ERR_STATE, "flags and state must be set first");
  if (hipSetDevice(c->cfg.device) != hipSuccess) return cfail(c, LBM_ERR_HIP, "hipSetDevice failed");
  const dim3 grid((unsigned)((c->ncouple + 255) / 256)), block(256);
  const lbm::CoupleArgs a{c->d_ccells, c->ncouple, c->d_cu, c->couple_rho ? c->d_crho : nullptr, c->couple_mode,
                          c->couple_dt, c->cfg.tau, c->p64.omega};
  const int on = c->couple_on ? 1 : 0;
  if (c->fp32)
    hipLaunchKernelGGL(lbm::k_couple_gather<float>, grid, block, 0, c->stream, lbm_state_geom(c), (const float*)c->buf[c->cur],
                       c->words, a, on, c->d_cout);
  else
    hipLaunchKernelGGL(lbm::k_couple_gather<double>, grid, block, 0, c->stream, lbm_state_geom(c), (const double*)c->buf[c->cur],
                       c->words, a, on, c->d_cout);
  std::vector<double> out((size_t)c->ncouple * 4);
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(out.data(), c->d_cout, out.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return cfail(c, LBM_ERR_HIP, "coupling gather failed");
  for (int64_t k = 0; k < c->ncouple; k++) {
    if (rho) rho[k] = out[4 * k];
    if (u)
      for (int d = 0; d < 3; d++) u[3 * k + d] = out[4 * k + 1 + d];
  }
  return LBM_OK;
}

}  // extern "C"
