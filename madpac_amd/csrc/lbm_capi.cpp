// C-ABI of the dense D3Q19 solver context (include/lbm_capi.h).  Host code only: it owns the
// device buffers, validates every argument and enqueues the kernels of lbm_dense.hip on the
// context's stream.  Citations: LB/ = madpac/peano/applications/latticeboltzmann/,
// BLB/ = LB/blocklatticeboltzmann/.
#include <algorithm>
#include "lbm_capi.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "lbm_ctx.h"
#include "lbm_internal.h"

using lbm::Q;

namespace {
thread_local std::string g_create_error;

int fail(lbm_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  else g_create_error = msg;
  return code;
}

int hipfail(lbm_ctx* c, hipError_t e, const char* what) {
  return fail(c, LBM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCHK(ctx, call)                              \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) return hipfail(ctx, e_, #call); \
  } while (0)

int ensure_scratch(lbm_ctx* c, size_t bytes) {
  if (c->scratch_bytes >= bytes) return LBM_OK;
  if (c->scratch) (void)hipFree(c->scratch);
  c->scratch = nullptr;
  c->scratch_bytes = 0;
  if (lbm::device_alloc(&c->scratch, bytes) != hipSuccess) return fail(c, LBM_ERR_NOMEM, "scratch allocation failed");
  c->scratch_bytes = bytes;
  return LBM_OK;
}

int rebuild_words(lbm_ctx* c) {
  HIPCHK(c, lbm::launch_build_words(c->g, c->d_inner, c->has_cls ? c->d_cls : nullptr, c->words, c->stream));
  HIPCHK(c, lbm::launch_count_fluid(c->g, c->words, c->partials, c->small, c->stream));
  double n = 0.0;
  HIPCHK(c, hipMemcpyAsync(&n, c->small, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->nfluid = (int64_t)n;
  // the stored state is a permutation that depends on the flags: it must be set again
  const size_t bytes = (size_t)Q * c->g.N * c->esize;
  HIPCHK(c, hipMemsetAsync(c->buf[0], 0, bytes, c->stream));
  if (c->buf[1]) HIPCHK(c, hipMemsetAsync(c->buf[1], 0, bytes, c->stream));
  c->state_set = false;
  c->parity = 0;
  return LBM_OK;
}

// Cells fetched per fluid cell when rows are read in 128-byte lines (`per_line` cells): the number
// of line-sized row segments holding any fluid cell, times their length, over the fluid cells.
// 1 for an all-fluid box or walls that fill whole rows; 1.2 (FP64 lines) for the 20 %-solid sphere
// packing.  Bricks cut that waste (1.10 there) but a shifted read touches up to 4x more lines,
// which costs 9 % on an all-fluid FP64 lattice; interpolating the measured gains and losses puts
// the break-even near 1.12.
double row_line_waste(const lbm::DenseGeom& g, const uint8_t* inner, int per_line) {
  int64_t fetched = 0, fluid = 0;
  for (int64_t row = 0; row < (int64_t)g.ny * g.nz; row++) {
    const uint8_t* r = inner + row * g.nx;
    for (int x0 = 0; x0 < g.nx; x0 += per_line) {
      const int len = std::min(per_line, g.nx - x0);
      int nf = 0;
      for (int x = x0; x < x0 + len; x++) nf += r[x] ? 1 : 0;
      fluid += nf;
      if (nf) fetched += len;
    }
  }
  return fluid ? (double)fetched / (double)fluid : 1.0;
}

// Switch the device cell order; the per-cell arrays already on the device (porosity, inlet
// velocities, coupling cell ids) follow.  The state is rebuilt by the caller (rebuild_words).
int switch_cell_order(lbm_ctx* c, bool bricks) {
  const lbm::DenseGeom from = c->g, to = bricks ? c->g_bricks : c->g_rows;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (c->alpha || c->d_bcu) {
    int rc = ensure_scratch(c, (size_t)std::max(from.N, to.N) * 3 * sizeof(double));
    if (rc) return rc;
  }
  if (c->alpha) {
    HIPCHK(c, lbm::launch_relayout(from, to, c->alpha, c->scratch, 1, (int)c->esize, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->alpha, c->scratch, (size_t)to.N * c->esize, hipMemcpyDeviceToDevice, c->stream));
  }
  if (c->d_bcu) {
    HIPCHK(c, lbm::launch_relayout(from, to, c->d_bcu, c->scratch, 3, 8, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_bcu, c->scratch, (size_t)to.N * 3 * sizeof(double), hipMemcpyDeviceToDevice,
                             c->stream));
  }
  if (c->ncouple) HIPCHK(c, lbm::launch_convert_ids(from, to, c->d_ccells, c->ncouple, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->g = to;
  return LBM_OK;
}

int upload_boundary_table(lbm_ctx* c) {
  HIPCHK(c, hipMemcpyAsync(c->d_bct, c->bct, sizeof(c->bct), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->walls = false;
  for (int j = 1; j < lbm::MAX_CLASSES; j++) c->walls = c->walls || c->bct[j].mask != 0u;
  return LBM_OK;
}

int need_state(lbm_ctx* c) {
  if (!c->flags_set) return fail(c, LBM_ERR_STATE, "boundary flags not set (lbm_set_flags)");
  if (!c->state_set) return fail(c, LBM_ERR_STATE, "state not initialised (lbm_init_equilibrium / lbm_set_populations)");
  return LBM_OK;
}
}  // namespace

// LB/configurations/LatticeBoltzmannConfiguration.cpp:269 and LB/MultiLevelSimData.h:160
double lbm_tau_effective(double tau) {
  double nu = 1.0 / 3.0 * (tau - 0.5);
  const double fac = 1.0;
  nu = nu * fac;
  return 3.0 * nu + 0.5;
}

extern "C" {

int lbm_create(lbm_ctx** out, const lbm_config* cfg) {
  if (!out || !cfg) return fail(nullptr, LBM_ERR_ARG, "null argument");
  *out = nullptr;
  if (cfg->nx < 1 || cfg->ny < 1 || cfg->nz < 1) return fail(nullptr, LBM_ERR_ARG, "lattice extent must be >= 1");
  if ((int64_t)cfg->nx * cfg->ny * cfg->nz > (int64_t)1 << 33)
    return fail(nullptr, LBM_ERR_ARG, "lattice too large");
  if (cfg->precision != LBM_FP64 && cfg->precision != LBM_FP32_SHIFTED && cfg->precision != LBM_FP32_SHIFTED_F32)
    return fail(nullptr, LBM_ERR_ARG, "unknown precision");
  if (cfg->layout != LBM_LAYOUT_TWO_BUFFER && cfg->layout != LBM_LAYOUT_AA)
    return fail(nullptr, LBM_ERR_ARG, "unknown layout");
  if (cfg->cell_order < LBM_CELL_ORDER_AUTO || cfg->cell_order > LBM_CELL_ORDER_BRICKS)
    return fail(nullptr, LBM_ERR_ARG, "unknown cell order");
  if (cfg->populations < LBM_POPULATIONS_AUTO || cfg->populations > LBM_POPULATIONS_INTERLEAVED)
    return fail(nullptr, LBM_ERR_ARG, "unknown population arrangement");
  if (!(cfg->tau > 0.5) || !(cfg->tau < 1e6)) return fail(nullptr, LBM_ERR_ARG, "tau must be > 0.5");
  for (int k = 0; k < 3; k++)
    if (!std::isfinite(cfg->force[k])) return fail(nullptr, LBM_ERR_ARG, "force must be finite");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(nullptr, LBM_ERR_HIP, "no HIP device");
  if (cfg->device < 0 || cfg->device >= ndev) return fail(nullptr, LBM_ERR_ARG, "device ordinal out of range");

  lbm_ctx* c = new lbm_ctx();
  c->cfg = *cfg;
  c->fp32 = cfg->precision != LBM_FP64;
  c->f32math = cfg->precision == LBM_FP32_SHIFTED_F32;
  c->aa = cfg->layout == LBM_LAYOUT_AA;
  c->esize = c->fp32 ? sizeof(float) : sizeof(double);
  // interleaved unless planes are asked for: measured faster in every case of the porous 256^3
  // channel and the all-fluid 256^3 / 512^3 lattices but FP32 in bricks (-2 %), which AUTO's cell
  // order does not pick
  const bool aos = cfg->populations != LBM_POPULATIONS_PLANES;
  c->g_rows = lbm::make_geom(cfg->nx, cfg->ny, cfg->nz, cfg->periodic, false, c->esize, aos);
  c->g_bricks = lbm::make_geom(cfg->nx, cfg->ny, cfg->nz, cfg->periodic, true, c->esize, aos);
  c->g = cfg->cell_order == LBM_CELL_ORDER_BRICKS ? c->g_bricks : c->g_rows;
  c->force = cfg->force[0] != 0.0 || cfg->force[1] != 0.0 || cfg->force[2] != 0.0;
  c->p64.tau_eff = lbm_tau_effective(cfg->tau);
  c->p64.omega = 1.0 / c->p64.tau_eff;
  for (int k = 0; k < 3; k++) c->p64.g[k] = cfg->force[k];
  c->p32.omega = (float)c->p64.omega;
  c->p32.fpref = (float)(1.0 - 0.5 * c->p64.omega);
  for (int k = 0; k < 3; k++) c->p32.g[k] = (float)cfg->force[k];
  for (int k = 0; k < lbm::MAX_CLASSES; k++) c->bct[k] = lbm::BoundaryClass{{0.0, 0.0, 0.0}, 1.0, 0u, 0u};

  auto bail = [&](int code, const char* msg) {
    g_create_error = msg;
    lbm_destroy(c);
    return code;
  };
  if (hipSetDevice(cfg->device) != hipSuccess) return bail(LBM_ERR_HIP, "hipSetDevice failed");
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(LBM_ERR_HIP, "stream creation failed");
  if (lbm_noalloc_setup(cfg->device) != hipSuccess)
    return bail(LBM_ERR_HIP, "device address of the no-alloc reduction array not found");
  if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess)
    return bail(LBM_ERR_HIP, "event creation failed");
  // sized for the larger (bricked) extent, so that lbm_set_flags can switch the cell order
  const size_t pbytes = (size_t)Q * c->g_bricks.N * c->esize;
  c->npartials = std::max(lbm::dense_moment_partials(c->g_rows), lbm::dense_moment_partials(c->g_bricks));
  if (lbm::device_alloc(&c->buf[0], pbytes) != hipSuccess || (!c->aa && lbm::device_alloc(&c->buf[1], pbytes) != hipSuccess) ||
      lbm::device_alloc(&c->words, (size_t)c->g_bricks.N * sizeof(uint32_t)) != hipSuccess ||
      lbm::device_alloc(&c->d_inner, (size_t)c->g.Nu) != hipSuccess || lbm::device_alloc(&c->d_cls, (size_t)c->g.Nu) != hipSuccess ||
      lbm::device_alloc(&c->partials, (size_t)(c->npartials + 1024) * 4 * sizeof(double)) != hipSuccess ||
      lbm::device_alloc(&c->small, 8 * sizeof(double)) != hipSuccess ||
      lbm::device_alloc(&c->d_bct, sizeof(c->bct)) != hipSuccess)
    return bail(LBM_ERR_NOMEM, "device allocation failed");
  if (hipMemsetAsync(c->buf[0], 0, pbytes, c->stream) != hipSuccess ||
      (c->buf[1] && hipMemsetAsync(c->buf[1], 0, pbytes, c->stream) != hipSuccess) ||
      hipMemsetAsync(c->words, 0, (size_t)c->g_bricks.N * sizeof(uint32_t), c->stream) != hipSuccess ||
      hipMemsetAsync(c->d_cls, 0, (size_t)c->g.Nu, c->stream) != hipSuccess ||
      hipMemcpyAsync(c->d_bct, c->bct, sizeof(c->bct), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return bail(LBM_ERR_HIP, "initial clear failed");
  *out = c;
  return LBM_OK;
}

int lbm_destroy(lbm_ctx* c) {
  if (!c) return LBM_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  lbm_adjoint_release(c);
  lbm_coupling_release(c);
  lbm_graphs_release(c);
  void* ptrs[] = {c->buf[0], c->buf[1], c->words,   c->d_inner, c->d_cls, c->alpha,
                  c->partials, c->small,  c->scratch, c->d_bct,   c->d_bcu};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return LBM_OK;
}

const char* lbm_last_error(const lbm_ctx* c) { return c ? c->err.c_str() : g_create_error.c_str(); }

int64_t lbm_fluid_cells(const lbm_ctx* c) { return c ? c->nfluid : -1; }

int lbm_cell_order(const lbm_ctx* c) {
  return c ? (c->g.bricked ? LBM_CELL_ORDER_BRICKS : LBM_CELL_ORDER_ROWS) : -1;
}

int lbm_set_flags(lbm_ctx* c, const uint8_t* inner) {
  if (!c) return LBM_ERR_ARG;
  if (!inner) return fail(c, LBM_ERR_ARG, "inner flags are null");
  HIPCHK(c, hipSetDevice(c->cfg.device));
  bool bricks = c->cfg.cell_order == LBM_CELL_ORDER_BRICKS;
  // FP32 keeps rows: with whole-line writes they measured faster than bricks on the 256^3 sphere
  // packing too (27.3k vs 26.7k MLUPS two-buffer, 22.3k vs 20.4k AA)
  if (c->cfg.cell_order == LBM_CELL_ORDER_AUTO) bricks = !c->fp32 && row_line_waste(c->g, inner, 16) >= 1.12;
  if (bricks != (c->g.bricked != 0)) {
    int rc = switch_cell_order(c, bricks);
    if (rc) return rc;
  }
  HIPCHK(c, hipMemcpyAsync(c->d_inner, inner, (size_t)c->g.Nu, hipMemcpyHostToDevice, c->stream));
  c->flags_set = true;
  return rebuild_words(c);
}

int lbm_set_boundary_class(lbm_ctx* c, int k, int type, double rho, const double u[3], uint32_t dir_mask) {
  if (!c) return LBM_ERR_ARG;
  if (k < 1 || k >= lbm::MAX_CLASSES) return fail(c, LBM_ERR_ARG, "boundary class must be in 1..31");
  const int ty = type & (int)lbm::BC_TYPE_MASK;
  if ((type & ~(int)(lbm::BC_TYPE_MASK | lbm::BC_CELL_VELOCITY)) || ty > (int)lbm::BC_VELOCITY_INLET)
    return fail(c, LBM_ERR_ARG, "unknown boundary type");
  if ((type & LBM_BC_CELL_VELOCITY) && ty != LBM_BC_PRESSURE_INLET && ty != LBM_BC_VELOCITY_INLET)
    return fail(c, LBM_ERR_ARG, "per-cell velocities apply to inlet classes only");
  if (!u) return fail(c, LBM_ERR_ARG, "boundary velocity is null");
  for (int d = 0; d < 3; d++)
    if (!std::isfinite(u[d])) return fail(c, LBM_ERR_ARG, "boundary velocity must be finite");
  if (!(rho > 0.0) || !std::isfinite(rho)) return fail(c, LBM_ERR_ARG, "boundary density must be positive");
  HIPCHK(c, hipSetDevice(c->cfg.device));
  lbm::BoundaryClass& b = c->bct[k];
  for (int d = 0; d < 3; d++) b.u[d] = u[d];
  b.rho = rho;
  b.type = (uint32_t)type;
  b.mask = dir_mask & ((1u << Q) - 1u) & ~(1u << lbm::REST);
  return upload_boundary_table(c);
}

int lbm_set_moving_wall_class(lbm_ctx* c, int k, const double uw[3], uint32_t dir_mask) {
  return lbm_set_boundary_class(c, k, LBM_BC_MOVING_WALL, 1.0, uw, dir_mask);
}

int lbm_set_boundary_velocity(lbm_ctx* c, const double* u) {
  if (!c) return LBM_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->cfg.device));
  if (!u) {
    if (c->d_bcu) {
      HIPCHK(c, hipStreamSynchronize(c->stream));
      (void)hipFree(c->d_bcu);
    }
    c->d_bcu = nullptr;
    return LBM_OK;
  }
  const lbm::DenseGeom& g = c->g;
  const size_t n = 3 * (size_t)g.Nu;
  for (size_t i = 0; i < n; i++)
    if (!std::isfinite(u[i])) return fail(c, LBM_ERR_ARG, "boundary velocities must be finite");
  // the kernels read the velocities by stored (bricked) cell
  std::vector<double> dev(3 * (size_t)g.N, 0.0);
  for (int z = 0; z < g.nz; z++)
    for (int y = 0; y < g.ny; y++)
      for (int x = 0; x < g.nx; x++) {
        const int64_t r = (int64_t)x + (int64_t)g.nx * ((int64_t)y + (int64_t)g.ny * z);
        const int64_t p = lbm::stored_index(g, x, y, z);
        for (int d = 0; d < 3; d++) dev[(size_t)(3 * p + d)] = u[3 * r + d];
      }
  // sized for either cell order (lbm_set_flags may switch it)
  if (!c->d_bcu && lbm::device_alloc(&c->d_bcu, 3 * (size_t)c->g_bricks.N * sizeof(double)) != hipSuccess)
    return fail(c, LBM_ERR_NOMEM, "boundary velocity allocation failed");
  HIPCHK(c, hipMemcpyAsync(c->d_bcu, dev.data(), dev.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return LBM_OK;
}

// The reference's default channel (SLB/AbstractBoundaryData4Channel.cpp:17-50 with
// SLB/BoundaryData4ChannelWithNoSlipBoundaries.cpp:17-27), classified link by link: a link
// whose target leaves the box through a y or z face is a no-slip wall; otherwise a link
// leaving through x- is a velocity inlet (rho_L = 1, LB/BoundaryData.cpp:5; u_L = (u_mean,0,0)
// or the parabolic profile of :72-77 at the cell centre, in lattice units), one leaving through
// x+ a pressure outlet with rho_L = 1 (:85-92); links into solid cells stay walls.  The inlet
// and outlet links of one cell form a class per distinct link set (at most 9 per face).
int lbm_set_channel(lbm_ctx* c, double u_mean, int parabolic) {
  if (!c) return LBM_ERR_ARG;
  const lbm::DenseGeom& g = c->g;
  if (c->cfg.periodic[0] || c->cfg.periodic[1] || c->cfg.periodic[2])
    return fail(c, LBM_ERR_ARG, "a channel needs a non-periodic box");
  if (g.nx < 2) return fail(c, LBM_ERR_ARG, "a channel needs nx >= 2");
  if (!std::isfinite(u_mean)) return fail(c, LBM_ERR_ARG, "inlet velocity must be finite");
  std::vector<uint8_t> cls((size_t)g.Nu, 0);
  std::vector<double> prof;
  if (parabolic) prof.assign(3 * (size_t)g.Nu, 0.0);
  lbm::BoundaryClass table[lbm::MAX_CLASSES];
  for (int k = 0; k < lbm::MAX_CLASSES; k++) table[k] = lbm::BoundaryClass{{0.0, 0.0, 0.0}, 1.0, 0u, 0u};
  int ncls = 1;
  for (int z = 0; z < g.nz; z++)
    for (int y = 0; y < g.ny; y++)
      for (int x : {0, g.nx - 1}) {
        const int64_t i = (int64_t)x + (int64_t)g.nx * ((int64_t)y + (int64_t)g.ny * z);
        uint32_t m = 0;
        for (int q = 0; q < Q; q++) {
          const int tx = x + lbm::CX(q), ty = y + lbm::CY(q), tz = z + lbm::CZ(q);
          if (ty < 0 || ty >= g.ny || tz < 0 || tz >= g.nz) continue;  // wall first
          if ((x == 0 && tx < 0) || (x == g.nx - 1 && tx >= g.nx)) m |= 1u << q;
        }
        if (!m) continue;
        const bool in = x == 0;
        const uint32_t type = in ? (lbm::BC_VELOCITY_INLET | (parabolic ? lbm::BC_CELL_VELOCITY : 0u))
                                 : lbm::BC_PRESSURE_OUTLET;
        int k = 1;
        while (k < ncls && !(table[k].mask == m && table[k].type == type)) k++;
        if (k == ncls) {
          if (ncls == lbm::MAX_CLASSES) return fail(c, LBM_ERR_ARG, "too many boundary classes");
          ncls++;
          table[k].mask = m;
          table[k].type = type;
          if (in) table[k].u[0] = u_mean;
        }
        cls[i] = (uint8_t)k;
        if (in && parabolic) {
          // vel(0) = vel(0)*6/L^2 * a (L - a) for d = y, z at the cell centre a = index + 1/2
          double v = u_mean;
          const double L[2] = {(double)g.ny, (double)g.nz};
          const double a[2] = {y + 0.5, z + 0.5};
          for (int e = 0; e < 2; e++) v = v * 6.0 / (L[e] * L[e]) * a[e] * (L[e] - a[e]);
          prof[3 * i] = v;
        }
      }
  HIPCHK(c, hipSetDevice(c->cfg.device));
  int rc = parabolic ? lbm_set_boundary_velocity(c, prof.data()) : LBM_OK;
  if (rc) return rc;
  std::memcpy(c->bct, table, sizeof(table));
  rc = upload_boundary_table(c);
  if (rc) return rc;
  return lbm_set_wall_classes(c, cls.data());
}

int lbm_set_wall_classes(lbm_ctx* c, const uint8_t* cls) {
  if (!c) return LBM_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->cfg.device));
  if (cls) {
    for (int64_t i = 0; i < c->g.Nu; i++)
      if (cls[i] >= lbm::MAX_CLASSES) return fail(c, LBM_ERR_ARG, "boundary class id must be < 32");
    HIPCHK(c, hipMemcpyAsync(c->d_cls, cls, (size_t)c->g.Nu, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->has_cls = true;
  } else {
    c->has_cls = false;
  }
  return c->flags_set ? rebuild_words(c) : LBM_OK;
}

int lbm_set_cavity_lid(lbm_ctx* c, double ulid) {
  if (!c) return LBM_ERR_ARG;
  const double uw[3] = {ulid, 0.0, 0.0};
  uint32_t mask = 0;
  for (int q = 0; q < Q; q++)
    if (lbm::CZ(q) > 0) mask |= 1u << q;
  int rc = lbm_set_moving_wall_class(c, 1, uw, mask);
  if (rc) return rc;
  std::vector<uint8_t> cls((size_t)c->g.Nu, 0);
  const int64_t plane = (int64_t)c->g.nx * c->g.ny;
  for (int64_t i = (int64_t)(c->g.nz - 1) * plane; i < c->g.Nu; i++) cls[i] = 1;
  return lbm_set_wall_classes(c, cls.data());
}

int lbm_set_porosity(lbm_ctx* c, const double* alpha) {
  if (!c) return LBM_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->cfg.device));
  if (!alpha) {
    if (c->alpha) {
      HIPCHK(c, hipStreamSynchronize(c->stream));  // queued steps may still read it
      (void)hipFree(c->alpha);
    }
    c->alpha = nullptr;
    return LBM_OK;
  }
  for (int64_t i = 0; i < c->g.Nu; i++)
    if (!(alpha[i] >= 0.0 && alpha[i] <= 1.0)) return fail(c, LBM_ERR_ARG, "porosity must lie in [0,1]");
  if (c->couple_on) return fail(c, LBM_ERR_STATE, "coupled cells use the plain collision: no porosity");
  if (!c->alpha && lbm::device_alloc(&c->alpha, (size_t)c->g_bricks.N * c->esize) != hipSuccess)
    return fail(c, LBM_ERR_NOMEM, "porosity allocation failed");
  // the kernels index the porosity by stored (bricked) cell; padding cells get 0
  const lbm::DenseGeom& g = c->g;
  std::vector<double> a64;
  std::vector<float> a32;
  if (c->fp32) a32.assign((size_t)g.N, 0.f);
  else a64.assign((size_t)g.N, 0.0);
  for (int z = 0; z < g.nz; z++)
    for (int y = 0; y < g.ny; y++)
      for (int x = 0; x < g.nx; x++) {
        const int64_t u = (int64_t)x + (int64_t)g.nx * ((int64_t)y + (int64_t)g.ny * z);
        const int64_t p = lbm::stored_index(g, x, y, z);
        if (c->fp32) a32[p] = (float)alpha[u];
        else a64[p] = alpha[u];
      }
  if (c->fp32) HIPCHK(c, hipMemcpy(c->alpha, a32.data(), a32.size() * sizeof(float), hipMemcpyHostToDevice));
  else HIPCHK(c, hipMemcpy(c->alpha, a64.data(), a64.size() * sizeof(double), hipMemcpyHostToDevice));
  return LBM_OK;
}

int lbm_init_equilibrium(lbm_ctx* c, const double* rho, const double* u) {
  if (!c) return LBM_ERR_ARG;
  if (!c->flags_set) return fail(c, LBM_ERR_STATE, "boundary flags not set (lbm_set_flags)");
  HIPCHK(c, hipSetDevice(c->cfg.device));
  const size_t n = (size_t)c->g.Nu;
  int rc = ensure_scratch(c, (rho ? n : 0) * sizeof(double) + (u ? 3 * n : 0) * sizeof(double) + 8);
  if (rc) return rc;
  double* d_rho = rho ? (double*)c->scratch : nullptr;
  double* d_u = u ? (double*)c->scratch + (rho ? n : 0) : nullptr;
  if (rho) HIPCHK(c, hipMemcpyAsync(d_rho, rho, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (u) HIPCHK(c, hipMemcpyAsync(d_u, u, 3 * n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  c->parity = 0;  // AA: written NATURAL
  HIPCHK(c, lbm::launch_dense_init_eq(lbm_state_geom(c), c->fp32, d_rho, d_u, c->buf[c->cur], c->words, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->state_set = true;
  return LBM_OK;
}

int lbm_set_populations(lbm_ctx* c, const double* f) {
  if (!c) return LBM_ERR_ARG;
  if (!f) return fail(c, LBM_ERR_ARG, "populations are null");
  if (!c->flags_set) return fail(c, LBM_ERR_STATE, "boundary flags not set (lbm_set_flags)");
  HIPCHK(c, hipSetDevice(c->cfg.device));
  const size_t bytes = (size_t)Q * c->g.Nu * sizeof(double);
  int rc = ensure_scratch(c, bytes);
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(c->scratch, f, bytes, hipMemcpyHostToDevice, c->stream));
  c->parity = 0;  // AA: written NATURAL
  HIPCHK(c, lbm::launch_dense_unpull(lbm_state_geom(c), c->fp32, (const double*)c->scratch, c->buf[c->cur], c->words,
                                     c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->state_set = true;
  return LBM_OK;
}

int lbm_get_populations(lbm_ctx* c, double* f) {
  if (!c) return LBM_ERR_ARG;
  if (!f) return fail(c, LBM_ERR_ARG, "output is null");
  int rc = need_state(c);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->cfg.device));
  const size_t bytes = (size_t)Q * c->g.Nu * sizeof(double);
  rc = ensure_scratch(c, bytes);
  if (rc) return rc;
  HIPCHK(c, lbm::launch_dense_pull(lbm_state_geom(c), c->fp32, c->buf[c->cur], (double*)c->scratch, c->words,
                                   c->stream));
  HIPCHK(c, hipMemcpyAsync(f, c->scratch, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return LBM_OK;
}

}  // extern "C"

int lbm_enqueue_step_from(lbm_ctx* c, const void* src, void* dst) {
  HIPCHK(c, lbm::launch_dense_step(c->g, c->fp32, c->f32math, src, dst, c->words, c->alpha, c->force, c->walls,
                                   c->p64, c->p32, lbm::BoundaryArgs{c->d_bct, c->d_bcu}, c->spec, c->stream));
  return LBM_OK;
}

int lbm_enqueue_steps(lbm_ctx* c, int nsteps) {
  if (c->aa) {
    for (int s = 0; s < nsteps; s++) {
      int rc = lbm_enqueue_coupling_save(c);
      if (rc) return rc;
      HIPCHK(c, lbm::launch_aa_step(c->g, c->fp32, c->f32math, c->parity, c->buf[0], c->words, c->alpha, c->force,
                                    c->walls, c->p64, c->p32, lbm::BoundaryArgs{c->d_bct, c->d_bcu}, c->stream));
      rc = lbm_enqueue_coupling_aa(c, c->parity);
      if (rc) return rc;
      c->parity ^= 1;
    }
    return LBM_OK;
  }
  for (int s = 0; s < nsteps; s++) {
    HIPCHK(c, lbm::launch_dense_step(c->g, c->fp32, c->f32math, c->buf[c->cur], c->buf[c->cur ^ 1], c->words,
                                     c->alpha, c->force, c->walls, c->p64, c->p32,
                                     lbm::BoundaryArgs{c->d_bct, c->d_bcu}, c->spec, c->stream));
    const int rc = lbm_enqueue_coupling(c, c->buf[c->cur], c->buf[c->cur ^ 1]);
    if (rc) return rc;
    c->cur ^= 1;
  }
  return LBM_OK;
}

namespace {
lbm_graph_key graph_key(const lbm_ctx* c) {
  lbm_graph_key k;
  k.bricked = c->g.bricked;
  k.remap = c->g.remap;
  k.by = c->g.by;
  k.walls = c->walls;
  k.force = c->force;
  k.spec = c->spec;
  k.couple_on = c->couple_on && c->ncouple > 0;
  if (k.couple_on) {
    k.couple_mode = c->couple_mode;
    k.couple_rho = c->couple_rho;
    k.couple_dt = c->couple_dt;
    k.ccells = c->d_ccells;
    k.cu = c->d_cu;
    k.crho = c->d_crho;
    k.csave = c->d_csave;
    k.ncouple = c->ncouple;
  }
  k.alpha = c->alpha;
  k.bcu = c->d_bcu;
  return k;
}

void graph_release(lbm_ctx* c, int slot) {
  if (c->gexec[slot]) (void)hipGraphExecDestroy(c->gexec[slot]);
  if (c->graph[slot]) (void)hipGraphDestroy(c->graph[slot]);
  c->gexec[slot] = nullptr;
  c->graph[slot] = nullptr;
  c->gkey[slot] = lbm_graph_key{};
}
}  // namespace

// Steps per captured graph (even).  Measured per step on MI355X / ROCm 7.2 (cavity 16^3 / 32^3 /
// 64^3, plain launches 6.1 / 6.2 / 18.4 us): 2-step graphs 8.6 / 8.7 / 21.1 us, 8-step 6.6 / 6.8 /
// 19.1 us, 32-step 7.0 / 6.4 / 18.6 us, 128-step 6.8 / 6.3 / 18.7 us.  Dependent kernels are bound by
// the dispatch-to-dispatch latency either way; replays never beat plain launches here.
constexpr int GRAPH_STEPS = 32;

void lbm_graphs_release(lbm_ctx* c) {
  for (int s = 0; s < 2; s++) graph_release(c, s);
}

extern "C" {

int lbm_step(lbm_ctx* c, int nsteps) {
  if (!c) return LBM_ERR_ARG;
  if (nsteps < 0) return fail(c, LBM_ERR_ARG, "nsteps must be >= 0");
  int rc = need_state(c);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->cfg.device));
  return lbm_enqueue_steps(c, nsteps);
}

int lbm_step_graph(lbm_ctx* c, int nsteps) {
  if (!c) return LBM_ERR_ARG;
  if (nsteps < 0) return fail(c, LBM_ERR_ARG, "nsteps must be >= 0");
  int rc = need_state(c);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->cfg.device));
  const int slot = c->aa ? c->parity : c->cur;
  const int K = GRAPH_STEPS;
  if (nsteps >= K) {
    const lbm_graph_key key = graph_key(c);
    if (!c->gexec[slot] || !(c->gkey[slot] == key)) {
      graph_release(c, slot);
      HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      rc = lbm_enqueue_steps(c, K);  // K even: cur (or parity) ends at this slot's phase
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(c->stream, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (e != hipSuccess) return fail(c, LBM_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
      c->graph[slot] = g;
      const hipError_t ei = hipGraphInstantiate(&c->gexec[slot], g, nullptr, nullptr, 0);
      if (ei != hipSuccess) {
        graph_release(c, slot);
        return fail(c, LBM_ERR_HIP, std::string("graph instantiation: ") + hipGetErrorString(ei));
      }
      c->gkey[slot] = key;
    }
    for (int k = 0; k < nsteps / K; k++) HIPCHK(c, hipGraphLaunch(c->gexec[slot], c->stream));
  }
  return lbm_enqueue_steps(c, nsteps % K);
}

int lbm_step_timed(lbm_ctx* c, int nsteps, float* ms) {
  if (!c) return LBM_ERR_ARG;
  if (!ms) return fail(c, LBM_ERR_ARG, "ms is null");
  if (nsteps < 0) return fail(c, LBM_ERR_ARG, "nsteps must be >= 0");
  int rc = need_state(c);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->cfg.device));
  HIPCHK(c, hipEventRecord(c->ev0, c->stream));
  rc = lbm_enqueue_steps(c, nsteps);
  if (rc) return rc;
  HIPCHK(c, hipEventRecord(c->ev1, c->stream));
  HIPCHK(c, hipEventSynchronize(c->ev1));
  HIPCHK(c, hipEventElapsedTime(ms, c->ev0, c->ev1));
  return LBM_OK;
}

int lbm_measure_copy(lbm_ctx* c, int mode, int nsteps, float* ms) {
  if (!c) return LBM_ERR_ARG;
  if (!ms) return fail(c, LBM_ERR_ARG, "ms is null");
  if (nsteps < 1) return fail(c, LBM_ERR_ARG, "nsteps must be >= 1");
  if (mode != 0 && mode != 1) return fail(c, LBM_ERR_ARG, "mode must be 0 (copy) or 1 (pull)");
  if (c->aa) return fail(c, LBM_ERR_STATE, "the copy reference needs the two-buffer layout");
  int rc = need_state(c);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->cfg.device));
  // into the buffer the next step overwrites: the state in buf[cur] is left as it is
  const void* src = c->buf[c->cur];
  void* dst = c->buf[c->cur ^ 1];
  HIPCHK(c, hipEventRecord(c->ev0, c->stream));
  for (int k = 0; k < nsteps; k++) HIPCHK(c, lbm::launch_dense_copy(c->g, c->fp32, mode == 1, src, dst, c->stream));
  HIPCHK(c, hipEventRecord(c->ev1, c->stream));
  HIPCHK(c, hipEventSynchronize(c->ev1));
  HIPCHK(c, hipEventElapsedTime(ms, c->ev0, c->ev1));
  return LBM_OK;
}

// Tuning variant: bit 0 the load schedule, bit 1 the workgroup order, bits 2-3 the workgroup
// shape (0: the default, 1/2/3: 2/4/8 rows or bricks).
static const int kShapeRows[4] = {0, 2, 4, 8};

static void apply_variant(lbm_ctx* c, int v) {
  c->spec = !c->aa && (v & 1);
  c->g.remap = (v >> 1) & 1;
  const int s = (v >> 2) & 3;
  lbm::set_workgroup_rows(c->g, s ? kShapeRows[s] : lbm::BY_DEFAULT);
}

static int current_variant(const lbm_ctx* c) {
  int s = 0;
  for (int k = 1; k < 4; k++)
    if (kShapeRows[k] == c->g.by) s = k;
  return (c->spec ? 1 : 0) | (c->g.remap << 1) | (s << 2);
}

int lbm_tune(lbm_ctx* c, int nsteps, int* variant) {
  if (!c) return LBM_ERR_ARG;
  if (nsteps < 0) return fail(c, LBM_ERR_ARG, "nsteps must be >= 0");
  if (nsteps == 0) {  // explicit choice
    if (!variant || *variant < 0 || *variant > 15) return fail(c, LBM_ERR_ARG, "variant must be in 0..15");
    apply_variant(c, *variant);
    return LBM_OK;
  }
  int rc = need_state(c);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->cfg.device));
  // time each variant on a copy of the state (both buffers hold the same start, so two-buffer and
  // in-place steps begin from it alike), then restore it
  const size_t bytes = (size_t)Q * c->g.N * c->esize;
  if (ensure_scratch(c, bytes) != LBM_OK) {  // no room for the copy (e.g. a 1024^3 AA state): keep the defaults
    c->err.clear();
    if (variant) *variant = current_variant(c);
    return LBM_OK;
  }
  const int cur0 = c->cur, par0 = c->parity;
  HIPCHK(c, hipMemcpyAsync(c->scratch, c->buf[cur0], bytes, hipMemcpyDeviceToDevice, c->stream));
  float best = 0.f;
  int best_v = -1;
  for (int v = 4; v < 16; v++) {  // the three explicit shapes
    if (c->aa && (v & 1)) continue;  // the AA kernels have one load schedule
    apply_variant(c, v);
    HIPCHK(c, hipMemcpyAsync(c->buf[cur0], c->scratch, bytes, hipMemcpyDeviceToDevice, c->stream));
    c->cur = cur0;
    c->parity = par0;
    rc = lbm_enqueue_steps(c, 2);  // warm-up
    if (rc) return rc;
    float ms = 0.f;
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    rc = lbm_enqueue_steps(c, nsteps);
    if (rc) return rc;
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (best_v < 0 || ms < best) {
      best = ms;
      best_v = v;
    }
  }
  apply_variant(c, best_v);
  HIPCHK(c, hipMemcpyAsync(c->buf[cur0], c->scratch, bytes, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->cur = cur0;
  c->parity = par0;
  if (variant) *variant = best_v;
  return LBM_OK;
}

int lbm_synchronize(lbm_ctx* c) {
  if (!c) return LBM_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return LBM_OK;
}

int lbm_get_fields(lbm_ctx* c, double* rho, double* u) {
  if (!c) return LBM_ERR_ARG;
  if (!rho || !u) return fail(c, LBM_ERR_ARG, "output is null");
  int rc = need_state(c);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->cfg.device));
  const size_t n = (size_t)c->g.Nu;
  rc = ensure_scratch(c, 4 * n * sizeof(double));
  if (rc) return rc;
  double* d_rho = (double*)c->scratch;
  double* d_u = d_rho + n;
  HIPCHK(c, lbm::launch_dense_fields(lbm_state_geom(c), c->fp32, c->buf[c->cur], c->words, d_rho, d_u, c->stream));
  HIPCHK(c, hipMemcpyAsync(rho, d_rho, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(u, d_u, 3 * n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return LBM_OK;
}

int lbm_reduce_moments(lbm_ctx* c, double out[4]) {
  if (!c) return LBM_ERR_ARG;
  if (!out) return fail(c, LBM_ERR_ARG, "output is null");
  int rc = need_state(c);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->cfg.device));
  double h[4];
  HIPCHK(c, lbm::launch_dense_moments(lbm_state_geom(c), c->fp32, c->buf[c->cur], c->words, c->partials, c->small,
                                      c->stream));
  HIPCHK(c, hipMemcpyAsync(h, c->small, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  // the kernel sums deviations from rest (f - w); the rest mass is one per fluid cell
  out[0] = (double)c->nfluid + h[0];
  out[1] = h[1];
  out[2] = h[2];
  out[3] = h[3];
  for (int k = 0; k < 4; k++)
    if (!std::isfinite(out[k])) return fail(c, LBM_ERR_NUMERIC, "non-finite population detected");
  return LBM_OK;
}

int lbm_kinetic_energy(lbm_ctx* c, double* ke) {
  if (!c) return LBM_ERR_ARG;
  if (!ke) return fail(c, LBM_ERR_ARG, "output is null");
  int rc = need_state(c);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->cfg.device));
  HIPCHK(c, lbm::launch_dense_energy(lbm_state_geom(c), c->fp32, c->buf[c->cur], c->words, c->partials, c->small,
                                     c->stream));
  HIPCHK(c, hipMemcpyAsync(ke, c->small, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (!std::isfinite(*ke)) return fail(c, LBM_ERR_NUMERIC, "non-finite kinetic energy");
  return LBM_OK;
}

int lbm_pressure_drop(lbm_ctx* c, int xin, int xout, double* J) {
  if (!c) return LBM_ERR_ARG;
  if (!J) return fail(c, LBM_ERR_ARG, "output is null");
  if (xin < 0 || xin >= c->g.nx || xout < 0 || xout >= c->g.nx) return fail(c, LBM_ERR_ARG, "plane out of range");
  int rc = need_state(c);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->cfg.device));
  double h[4];
  HIPCHK(c, lbm::launch_dense_plane_rho(lbm_state_geom(c), c->fp32, c->buf[c->cur], c->words, xin, xout, c->partials,
                                         c->small, c->stream));
  HIPCHK(c, hipMemcpyAsync(h, c->small, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (h[1] < 1.0 || h[3] < 1.0) return fail(c, LBM_ERR_ARG, "a plane holds no fluid cell");
  *J = (1.0 / 3.0) * (h[0] / h[1] - h[2] / h[3]);
  if (!std::isfinite(*J)) return fail(c, LBM_ERR_NUMERIC, "non-finite pressure drop");
  return LBM_OK;
}

}  // extern "C"
