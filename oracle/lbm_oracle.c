/*
 * CPU oracle for the D3Q19 BGK block-LB hot path (TEST INFRASTRUCTURE ONLY).
 *
 * Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py use
 * this file, and only to check results against.  See lbm_oracle.h for the pinning summary.
 * Reference citations: LB/ = madpac/peano/applications/latticeboltzmann/,
 * BLB/ = LB/blocklatticeboltzmann/, SLB/ = madpac/scenario/latticeboltzmann/blocklatticeboltzmann/.
 *
 * Build: gcc -O2 -ffp-contract=off -fopenmp -fPIC -shared (oracle/Makefile).
 */
#include "lbm_oracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* D3Q19 velocities, z-major then y then x; rest population q = 9, inverse q' = 18 - q.
 * LB/LatticeBoltzmannDefinitions.cpp:113-131. */
const int LBO_C[LBO_Q][3] = {
  { 0,-1,-1}, {-1, 0,-1}, { 0, 0,-1}, { 1, 0,-1}, { 0, 1,-1},
  {-1,-1, 0}, { 0,-1, 0}, { 1,-1, 0}, {-1, 0, 0}, { 0, 0, 0},
  { 1, 0, 0}, {-1, 1, 0}, { 0, 1, 0}, { 1, 1, 0}, { 0,-1, 1},
  {-1, 0, 1}, { 0, 0, 1}, { 1, 0, 1}, { 0, 1, 1}};

/* LB/LatticeBoltzmannDefinitions.cpp:149-151 */
const double LBO_W[LBO_Q] = {
  1.0/36.0, 1.0/36.0, 1.0/18.0, 1.0/36.0, 1.0/36.0,
  1.0/36.0, 1.0/18.0, 1.0/36.0, 1.0/18.0, 1.0/3.0, 1.0/18.0, 1.0/36.0, 1.0/18.0, 1.0/36.0,
  1.0/36.0, 1.0/36.0, 1.0/18.0, 1.0/36.0, 1.0/36.0};

/* The reference keeps the velocities as doubles; products c*u are exact either way. */
static double cd(int q, int d) { return (double)LBO_C[q][d]; }

/* nu_L = 1.0/3.0*(tau-0.5) (LB/configurations/LatticeBoltzmannConfiguration.cpp:269);
 * tau_eff = 3.0*nu_L + 0.5 (LB/MultiLevelSimData.h:160, level factor pow(3,0) = 1). */
double lbo_tau_eff(double tau) {
  double nu = 1.0 / 3.0 * (tau - 0.5);
  double fac = 1.0;
  nu = nu * fac;
  return 3.0 * nu + 0.5;
}

/* BLB/BlockComputeCellValues.cpp:26-43: rho = sum_q f_q in q order, from 0.0. */
double lbo_density(const double* f) {
  double rho = 0.0;
  for (int q = 0; q < LBO_Q; q++) rho += f[q];
  return rho;
}

/* BLB/BlockComputeCellValues.cpp:47-68: buffer += f_q*c_q, then (1.0/rho)*buffer. */
void lbo_velocity(const double* f, double rho, double* u) {
  double b[3] = {0.0, 0.0, 0.0};
  for (int q = 0; q < LBO_Q; q++)
    for (int d = 0; d < 3; d++) b[d] += cd(q, d) * f[q];
  double inv = 1.0 / rho;
  for (int d = 0; d < 3; d++) u[d] = b[d] * inv;
}

/* tarch::la::dot (tarch/la/VectorVectorOperations.cpph:194-209): l0*r0 + l1*r1 + l2*r2 in order. */
static double dot3(const double* a, const double* b) {
  double r = a[0] * b[0];
  r += a[1] * b[1];
  r += a[2] * b[2];
  return r;
}

/* BLB/BlockComputeCellValues.cpp:9-22 */
void lbo_feq(double rho, const double* u, double* feq) {
  for (int q = 0; q < LBO_Q; q++) {
    double c[3] = {cd(q, 0), cd(q, 1), cd(q, 2)};
    double cu = dot3(c, u);
    feq[q] = LBO_W[q] * rho * (1.0 + cu / (LBO_CS2) + (cu * cu) / (2.0 * LBO_CS2 * LBO_CS2)
                               - (dot3(u, u) / (2.0 * LBO_CS2)));
  }
}

/* BLB/collisionmodels/BlockBGK.cpp:43: f -= 1.0/tau*(f - feq) */
void lbo_bgk(double* f, double tau_eff, const double* feq) {
  for (int q = 0; q < LBO_Q; q++) f[q] -= 1.0 / tau_eff * (f[q] - feq[q]);
}

/* BLB/boundaryconditions/BlockMovingWall.cpp:23-28, with q the outgoing direction:
 * ciVel = u_w . c_q ; ciVel = ciVel*2.0*w_q*rho/cs2 ; f_q -= ciVel */
double lbo_moving_wall_term(const double* uw, int q, double rho) {
  double c[3] = {cd(q, 0), cd(q, 1), cd(q, 2)};
  double civel = dot3(uw, c);
  civel = civel * 2.0 * LBO_W[q] * rho / (LBO_CS2);
  return civel;
}

/* Value stored into the outgoing slot q of a boundary link (cell i, class k), after the
 * collision; the swap streaming later moves it into the incoming slot 18-q.
 *  moving wall      BLB/boundaryconditions/BlockMovingWall.cpp:23-28: f_q - ciVel;
 *  pressure in/out  BLB/boundaryconditions/BlockAbstractPressureCondition.cpp:24-37:
 *                   -f_q + feq_{18-q} + feq_q with feq of (rho_t, u_t); the outlet takes
 *                   rho_t = rho_L and u_t = the cell's velocity (BlockPressureOutletCondition.h:32-50),
 *                   the inlet u_t = u_L and rho_t = the cell's density (BlockPressureInletCondition.h:32-50);
 *  velocity inlet   BLB/boundaryconditions/BlockVelocityInletEquilibrium.cpp:19-23: feq_{18-q}(rho_L, u_L).
 * The cell's velocity is the one the collision used before any porosity scaling (with a
 * body force, u + g/2). */
double lbo_boundary_value(const lbo_params* p, unsigned k, int64_t i, int q, double fq, double rho,
                          const double* u) {
  const int t = p->bc_type[k];
  const int ty = t & LBO_BC_TYPE_MASK;
  if (ty == LBO_BC_MOVING_WALL) return fq - lbo_moving_wall_term(p->uw[k], q, rho);
  const double* ub = ((t & LBO_BC_CELL_VELOCITY) && p->bc_u) ? p->bc_u + 3 * i : p->uw[k];
  const int qin = LBO_Q - 1 - q;
  double feq[LBO_Q];
  if (ty == LBO_BC_VELOCITY_INLET) {
    lbo_feq(p->bc_rho[k], ub, feq);
    return feq[qin];
  }
  if (ty == LBO_BC_PRESSURE_OUTLET) lbo_feq(p->bc_rho[k], u, feq);
  else lbo_feq(rho, ub, feq);
  return -fq + feq[qin] + feq[q];
}

/* Collision of one fluid cell, in place on f[19].  Without force and porosity this is
 * the reference sequence computeDensity -> computeVelocity -> BGK (mapping
 * BLB/mappings/RegularGrid2RegularBlockSolver.cpp:126-137).  With a body acceleration g
 * it applies Guo forcing (equilibrium velocity u + g/2, source (1-1/(2 tau)) w rho
 * [(c-u).g/cs2 + (c.u)(c.g)/cs2^2]); the reference's only forcing routine
 * (madpac/peanocoupling/PeanoDataExchangeService.cpp:161-186) squares c.u in the second
 * term, which is not Guo's scheme, so the published scheme is used here.  With a
 * porosity alpha the equilibrium velocity is scaled by (1-alpha) (Brinkman penalisation). */
static double collide_cell(double* f, double tau_eff, const lbo_params* p, double alpha, double* uout) {
  double rho = lbo_density(f);
  double u[3];
  lbo_velocity(f, rho, u);
  if (p->use_force) {
    for (int d = 0; d < 3; d++) u[d] = u[d] + 0.5 * p->g[d];
  }
  for (int d = 0; d < 3; d++) uout[d] = u[d];
  if (p->alpha) {
    for (int d = 0; d < 3; d++) u[d] = (1.0 - alpha) * u[d];
  }
  double feq[LBO_Q];
  lbo_feq(rho, u, feq);
  lbo_bgk(f, tau_eff, feq);
  if (p->use_force) {
    /* Guo's source with cs2 = 1/3 folded into the factors: 3 (c - u).g + 9 (c.u)(c.g) */
    double pref = (1.0 - 0.5 / tau_eff) * rho;
    double ug = dot3(u, p->g);
    for (int q = 0; q < LBO_Q; q++) {
      double c[3] = {cd(q, 0), cd(q, 1), cd(q, 2)};
      double cu = dot3(c, u);
      double cg = dot3(c, p->g);
      f[q] += pref * LBO_W[q] * (3.0 * (cg - ug) + 9.0 * cu * cg);
    }
  }
  return rho;
}

/* computeGi (madpac/peanocoupling/PeanoDataExchangeService.cpp:161-186), as written there:
 * g_i = (1 - 0.5/tau) w_i rho [ (c_i - u).g / cs2 + (c_i.u)^2 (c_i.g) / cs2^2 ].  The squared
 * c.u differs from Guo's scheme and makes sum_i g_i = -(1 - 0.5/tau) rho u.g / cs2 (not 0);
 * it is kept because this is the coupling the reference runs. */
void lbo_couple_gi(double rho, double tau, const double* u, const double* g, double* gi) {
  for (int q = 0; q < LBO_Q; q++) {
    double c[3] = {cd(q, 0), cd(q, 1), cd(q, 2)};
    double cmu[3] = {c[0] - u[0], c[1] - u[1], c[2] - u[2]};
    double cu = dot3(c, u);
    gi[q] = (1.0 - 0.5 / tau) * LBO_W[q] * rho *
            (1 / (LBO_CS2) * dot3(cmu, g) + (cu * cu) / (LBO_CS2 * LBO_CS2) * dot3(c, g));
  }
}

/* Collision of a coupled cell (getAndSetV :69-158 before BlockBGK, adjustF :189-218 after it):
 * overwrite: the equilibrium velocity is the MD velocity; Dirichlet: g = (u_MD - u)/dt,
 * u <- u + dt/2 g, then f += dt g_i after the BGK relaxation.  With density coupling rho_MD
 * replaces the density of the equilibrium (and of g_i).  The reference's adjustF applies its
 * term to cell index `_level` of each block only; here it applies to every coupled cell. */
/* The GET block of getAndSetV (madpac/peanocoupling/PeanoDataExchangeService.cpp:100-129) for
 * coupled row `row`: overwrite sets u = u_MD (:121-122); Dirichlet computes g = (u_MD - u)/dt and
 * shifts u += dt/2 g (:110-117); with density coupling rho = rho_MD (:124-129).  g is 0 in the
 * overwrite mode. */
static void couple_get(const lbo_params* p, int64_t row, double* rho, double* u, double* g) {
  const double* um = p->couple_u + 3 * row;
  for (int d = 0; d < 3; d++) g[d] = 0.0;
  if (p->couple_mode == LBO_COUPLE_DIRICHLET) {
    for (int d = 0; d < 3; d++) {
      g[d] = (um[d] - u[d]) / p->couple_dt;
      u[d] += p->couple_dt / 2.0 * g[d];
    }
  } else {
    for (int d = 0; d < 3; d++) u[d] = um[d];
  }
  if (p->couple_rho) *rho = p->couple_rho[row];
}

/* Collision of a coupled cell (getAndSetV :69-158 before BlockBGK, adjustF :189-218 after it):
 * overwrite: the equilibrium velocity is the MD velocity; Dirichlet: g = (u_MD - u)/dt,
 * u <- u + dt/2 g, then f += dt g_i after the BGK relaxation.  With density coupling rho_MD
 * replaces the density of the equilibrium (and of g_i).  The reference's adjustF applies its
 * term to cell index `_level` of each block only; here it applies to every coupled cell. */
static double collide_coupled(double* f, double tau_eff, const lbo_params* p, int64_t row, double* uout) {
  double rho = lbo_density(f);
  double u[3], g[3];
  lbo_velocity(f, rho, u);
  couple_get(p, row, &rho, u, g);
  for (int d = 0; d < 3; d++) uout[d] = u[d];
  double feq[LBO_Q];
  lbo_feq(rho, u, feq);
  lbo_bgk(f, tau_eff, feq);
  if (p->couple_mode == LBO_COUPLE_DIRICHLET) {
    double gi[LBO_Q];
    lbo_couple_gi(rho, p->tau, u, g, gi);
    for (int q = 0; q < LBO_Q; q++) f[q] += p->couple_dt * gi[q];
  }
  return rho;
}

static int64_t ncells(const lbo_dims* d) { return (int64_t)d->nx * d->ny * d->nz; }

/* neighbour of (x,y,z) along c_q, or -1 if it leaves a non-periodic domain */
static int64_t neighbour(const lbo_dims* d, int x, int y, int z, int q) {
  int n[3] = {x + LBO_C[q][0], y + LBO_C[q][1], z + LBO_C[q][2]};
  int e[3] = {d->nx, d->ny, d->nz};
  for (int k = 0; k < 3; k++) {
    if (n[k] < 0 || n[k] >= e[k]) {
      if (!d->periodic[k]) return -1;
      n[k] = (n[k] + e[k]) % e[k];
    }
  }
  return (int64_t)n[0] + (int64_t)d->nx * ((int64_t)n[1] + (int64_t)d->ny * n[2]);
}

void lbo_build_words(const lbo_dims* d, const uint8_t* solid, uint32_t* words) {
  for (int z = 0; z < d->nz; z++)
    for (int y = 0; y < d->ny; y++)
      for (int x = 0; x < d->nx; x++) {
        int64_t i = (int64_t)x + (int64_t)d->nx * ((int64_t)y + (int64_t)d->ny * z);
        if (solid[i]) { words[i] = 0; continue; }
        uint32_t w = LBO_FLUID;
        for (int q = 0; q < LBO_Q; q++) {
          if (q == LBO_REST) continue;
          int64_t j = neighbour(d, x, y, z, q);
          if (j < 0 || solid[j]) w |= (1u << q);
        }
        words[i] = w;
      }
}

/* Lid-driven cavity (SLB/BoundaryData4LidDrivenCavity.cpp:17-55 with the link scan of
 * SLB/services/ReceiveBoundaryDataService.cpp:107-185): every fluid cell of the n^3 box;
 * a link leaving the box is a moving wall iff c_q,z > 0 and the cell is in the top layer
 * (its neighbour centre lies above the lid), including edge and corner links; all other
 * leaving links are halfway bounce-back.  Class 1 carries u_w = (ulid, 0, 0). */
void lbo_build_cavity(int n, uint32_t* words, lbo_params* p, double ulid) {
  lbo_dims d = {n, n, n, {0, 0, 0}};
  uint8_t* solid = (uint8_t*)calloc((size_t)n * n * n, 1);
  lbo_build_words(&d, solid, words);
  free(solid);
  uint32_t lidmask = 0;
  for (int q = 0; q < LBO_Q; q++)
    if (LBO_C[q][2] > 0) lidmask |= 1u << q;
  for (int y = 0; y < n; y++)
    for (int x = 0; x < n; x++) {
      int64_t i = (int64_t)x + (int64_t)n * ((int64_t)y + (int64_t)n * (n - 1));
      words[i] |= 1u << LBO_CLASS_SHIFT;
    }
  memset(p->uw, 0, sizeof(p->uw));
  memset(p->moving_mask, 0, sizeof(p->moving_mask));
  p->uw[1][0] = ulid;
  p->moving_mask[1] = lidmask;
}

/* The reference's default channel (SLB/AbstractBoundaryData4Channel.cpp:17-50 with
 * SLB/BoundaryData4ChannelWithNoSlipBoundaries.cpp:17-27), link by link: a link whose target
 * leaves the box through a y or z face is a no-slip wall (halfway bounce-back); otherwise one
 * leaving through x- is a velocity inlet (rho_L = 1, the BoundaryData default,
 * LB/BoundaryData.cpp:5; u_L = (umean, 0, 0), or the parabolic profile of :72-77 evaluated
 * at the cell centre in lattice units), one leaving through x+ a pressure outlet with
 * rho_L = 1 (:85-92), and any other (an obstacle) a wall.  The links of one cell with the
 * same treatment form one class; classes are numbered in first-appearance order. */
void lbo_build_channel(const lbo_dims* d, uint32_t* words, lbo_params* p, double umean, int parabolic,
                       double* bc_u) {
  memset(p->uw, 0, sizeof(p->uw));
  memset(p->moving_mask, 0, sizeof(p->moving_mask));
  memset(p->bc_type, 0, sizeof(p->bc_type));
  for (int k = 0; k < LBO_MAX_CLASSES; k++) p->bc_rho[k] = 1.0;
  p->bc_u = parabolic ? bc_u : NULL;
  int ncls = 1;
  for (int z = 0; z < d->nz; z++)
    for (int y = 0; y < d->ny; y++)
      for (int x = 0; x < d->nx; x++) {
        int64_t i = (int64_t)x + (int64_t)d->nx * ((int64_t)y + (int64_t)d->ny * z);
        uint32_t w = words[i];
        words[i] = w & ~(LBO_CLASS_MASK << LBO_CLASS_SHIFT);
        if (!(w & LBO_FLUID)) continue;
        uint32_t inm = 0, outm = 0;
        for (int q = 0; q < LBO_Q; q++) {
          if (!(w & (1u << q))) continue;
          int nxx = x + LBO_C[q][0], nyy = y + LBO_C[q][1], nzz = z + LBO_C[q][2];
          if (nyy < 0 || nyy >= d->ny || nzz < 0 || nzz >= d->nz) continue;
          if (nxx < 0) inm |= 1u << q;
          else if (nxx >= d->nx) outm |= 1u << q;
        }
        if (bc_u) {
          double v = 0.0;
          if (inm) {
            v = umean;
            if (parabolic) {
              const double L[2] = {(double)d->ny, (double)d->nz};
              const double a[2] = {y + 0.5, z + 0.5};
              for (int e = 0; e < 2; e++) v = v * 6.0 / (L[e] * L[e]) * a[e] * (L[e] - a[e]);
            }
          }
          bc_u[3 * i] = v;
          bc_u[3 * i + 1] = 0.0;
          bc_u[3 * i + 2] = 0.0;
        }
        if (!inm && !outm) continue;
        if (inm && outm) { p->bc_type[0] = -1; return; } /* nx == 1: not a channel */
        const uint32_t m = inm ? inm : outm;
        const int t = inm ? (LBO_BC_VELOCITY_INLET | (parabolic ? LBO_BC_CELL_VELOCITY : 0))
                          : LBO_BC_PRESSURE_OUTLET;
        int k = 1;
        while (k < ncls && !(p->moving_mask[k] == m && p->bc_type[k] == t)) k++;
        if (k == ncls) {
          if (ncls == LBO_MAX_CLASSES) { p->bc_type[0] = -1; return; }
          ncls++;
          p->moving_mask[k] = m;
          p->bc_type[k] = t;
          if (inm) p->uw[k][0] = umean;
        }
        words[i] |= (uint32_t)k << LBO_CLASS_SHIFT;
      }
}

/* One step in push form.  Equivalent to the reference's swap streaming
 * (BLB/nonlocaltreatment/AABlockStreaming.cpp:19-92): f*_q(x) goes to x+c_q when that cell
 * is fluid, otherwise it is reflected into q' = 18-q of x (halfway bounce-back is implicit,
 * BLB/boundaryconditions/BlockHalfwayBounceBack.cpp:20-22), after the moving-wall term
 * (BLB/boundaryconditions/BlockMovingWall.cpp:23-28) for moving links.  Non-fluid cells are
 * neither collided nor written (BLB/collisionmodels/CollisionModelWrapper.cpp:36). */
void lbo_step(const lbo_dims* d, const uint32_t* words, const lbo_params* p,
              const double* src, double* dst, int nthreads) {
  const int64_t N = ncells(d);
  const double tau_eff = lbo_tau_eff(p->tau);
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
  #pragma omp parallel for num_threads(nthreads) schedule(static) collapse(2)
#endif
  for (int z = 0; z < d->nz; z++)
    for (int y = 0; y < d->ny; y++) {
      for (int x = 0; x < d->nx; x++) {
        int64_t i = (int64_t)x + (int64_t)d->nx * ((int64_t)y + (int64_t)d->ny * z);
        uint32_t w = words[i];
        if (!(w & LBO_FLUID)) continue;
        double f[LBO_Q];
        for (int q = 0; q < LBO_Q; q++) f[q] = src[(int64_t)q * N + i];
        double alpha = p->alpha ? p->alpha[i] : 0.0;
        double uc[3];
        const double te = p->tau_cell ? lbo_tau_eff(p->tau_cell[i]) : tau_eff;
        const int64_t row = p->couple_row ? p->couple_row[i] : -1;
        double rho = row >= 0 ? collide_coupled(f, te, p, row, uc) : collide_cell(f, te, p, alpha, uc);
        unsigned cls = (w >> LBO_CLASS_SHIFT) & LBO_CLASS_MASK;
        for (int q = 0; q < LBO_Q; q++) {
          if (w & (1u << q)) {
            double v = f[q];
            if (cls && ((p->moving_mask[cls] >> q) & 1u)) v = lbo_boundary_value(p, cls, i, q, v, rho, uc);
            dst[(int64_t)(LBO_Q - 1 - q) * N + i] = v;
          } else {
            int64_t j = neighbour(d, x, y, z, q);
            dst[(int64_t)q * N + j] = f[q];
          }
        }
      }
    }
}

void lbo_stream(const lbo_dims* d, const uint32_t* words, const double* src, double* dst) {
  const int64_t N = ncells(d);
  for (int z = 0; z < d->nz; z++)
    for (int y = 0; y < d->ny; y++)
      for (int x = 0; x < d->nx; x++) {
        int64_t i = (int64_t)x + (int64_t)d->nx * ((int64_t)y + (int64_t)d->ny * z);
        uint32_t w = words[i];
        if (!(w & LBO_FLUID)) continue;
        for (int q = 0; q < LBO_Q; q++) {
          if (w & (1u << q)) dst[(int64_t)(LBO_Q - 1 - q) * N + i] = src[(int64_t)q * N + i];
          else dst[(int64_t)q * N + neighbour(d, x, y, z, q)] = src[(int64_t)q * N + i];
        }
      }
}

/* The reference's streaming exactly as its three phases do it, on a dense AoS lattice:
 * collide every fluid cell (phase A, mapping :126-147), swap f[x][q] <-> f[x+c_q][q'] for
 * each fluid-fluid link once (phases A/B, AABlockStreaming.cpp:19-53, link enumerated
 * from the cell whose partner has not claimed it, :111-186), then swap q <-> q' inside
 * every cell (phase C, :80-92). */
void lbo_step_swap_aos(const lbo_dims* d, const uint32_t* words, const lbo_params* p, double* f) {
  const int64_t N = ncells(d);
  const double tau_eff = lbo_tau_eff(p->tau);
  for (int64_t i = 0; i < N; i++) {
    uint32_t w = words[i];
    if (!(w & LBO_FLUID)) continue;
    double* fi = f + i * LBO_Q;
    double alpha = p->alpha ? p->alpha[i] : 0.0;
    double uc[3];
    const double te = p->tau_cell ? lbo_tau_eff(p->tau_cell[i]) : tau_eff;
    const int64_t row = p->couple_row ? p->couple_row[i] : -1;
    double rho = row >= 0 ? collide_coupled(fi, te, p, row, uc) : collide_cell(fi, te, p, alpha, uc);
    unsigned cls = (w >> LBO_CLASS_SHIFT) & LBO_CLASS_MASK;
    if (cls)
      for (int q = 0; q < LBO_Q; q++)
        if ((w & (1u << q)) && ((p->moving_mask[cls] >> q) & 1u))
          fi[q] = lbo_boundary_value(p, cls, i, q, fi[q], rho, uc);
  }
  for (int z = 0; z < d->nz; z++)
    for (int y = 0; y < d->ny; y++)
      for (int x = 0; x < d->nx; x++) {
        int64_t i = (int64_t)x + (int64_t)d->nx * ((int64_t)y + (int64_t)d->ny * z);
        if (!(words[i] & LBO_FLUID)) continue;
        /* directions q > 9 point "forward"; each link is swapped once from its lower end */
        for (int q = LBO_REST + 1; q < LBO_Q; q++) {
          if (words[i] & (1u << q)) continue;
          int64_t j = neighbour(d, x, y, z, q);
          double t = f[i * LBO_Q + q];
          f[i * LBO_Q + q] = f[j * LBO_Q + (LBO_Q - 1 - q)];
          f[j * LBO_Q + (LBO_Q - 1 - q)] = t;
        }
      }
  for (int64_t i = 0; i < N; i++)
    for (int q = 0; q < LBO_Q / 2; q++) {
      double t = f[i * LBO_Q + q];
      f[i * LBO_Q + q] = f[i * LBO_Q + (LBO_Q - 1 - q)];
      f[i * LBO_Q + (LBO_Q - 1 - q)] = t;
    }
}

void lbo_fields(const lbo_dims* d, const uint32_t* words, const double* f, double* rho, double* u) {
  const int64_t N = ncells(d);
  for (int64_t i = 0; i < N; i++) {
    if (!(words[i] & LBO_FLUID)) {
      rho[i] = 0.0;
      u[3 * i] = u[3 * i + 1] = u[3 * i + 2] = 0.0;
      continue;
    }
    double fi[LBO_Q];
    for (int q = 0; q < LBO_Q; q++) fi[q] = f[(int64_t)q * N + i];
    rho[i] = lbo_density(fi);
    lbo_velocity(fi, rho[i], u + 3 * i);
  }
}

/* The export to MD of getAndSetV (madpac/peanocoupling/PeanoDataExchangeService.cpp:69-151):
 * the SET block (:136-151) sends the block's velocity and density after the GET block (:100-129)
 * has changed them, i.e. for the listed cells rho and u = (sum f c)/rho of state f
 * (BLB/BlockComputeCellValues.cpp:26-68) passed through couple_get when the cell has coupling data
 * (p->couple_row non-NULL and row >= 0), before the scaling by _lb2ms / _densityLB2MD.  Non-fluid
 * cells give 0. */
void lbo_couple_export(const lbo_dims* d, const uint32_t* words, const lbo_params* p, const double* f,
                       const int64_t* cells, int64_t n, double* rho, double* u) {
  const int64_t N = ncells(d);
  for (int64_t k = 0; k < n; k++) {
    const int64_t i = cells[k];
    double r = 0.0, v[3] = {0.0, 0.0, 0.0}, g[3];
    if (words[i] & LBO_FLUID) {
      double fi[LBO_Q];
      for (int q = 0; q < LBO_Q; q++) fi[q] = f[(int64_t)q * N + i];
      r = lbo_density(fi);
      lbo_velocity(fi, r, v);
      const int64_t row = p->couple_row ? p->couple_row[i] : -1;
      if (row >= 0) couple_get(p, row, &r, v, g);
    }
    rho[k] = r;
    for (int dd = 0; dd < 3; dd++) u[3 * k + dd] = v[dd];
  }
}

/* BLB/services/GridManagementService.cpp:121-148 initialises f = w (rho = 1, u = 0) in
 * fluid cells and 0 elsewhere; this generalises it to f = feq(rho, u). */
void lbo_init_equilibrium(const lbo_dims* d, const uint32_t* words, const double* rho,
                          const double* u, double* f) {
  const int64_t N = ncells(d);
  for (int64_t i = 0; i < N; i++) {
    double feq[LBO_Q];
    if (words[i] & LBO_FLUID) {
      lbo_feq(rho[i], u + 3 * i, feq);
    } else {
      for (int q = 0; q < LBO_Q; q++) feq[q] = 0.0;
    }
    for (int q = 0; q < LBO_Q; q++) f[(int64_t)q * N + i] = feq[q];
  }
}

void lbo_moments(const lbo_dims* d, const uint32_t* words, const double* f, double out[4]) {
  const int64_t N = ncells(d);
  double m[4] = {0, 0, 0, 0};
  for (int64_t i = 0; i < N; i++) {
    if (!(words[i] & LBO_FLUID)) continue;
    for (int q = 0; q < LBO_Q; q++) {
      double v = f[(int64_t)q * N + i];
      m[0] += v;
      m[1] += cd(q, 0) * v;
      m[2] += cd(q, 1) * v;
      m[3] += cd(q, 2) * v;
    }
  }
  for (int k = 0; k < 4; k++) out[k] = m[k];
}

double lbo_pressure_drop(const lbo_dims* d, const uint32_t* words, const double* f, int xin, int xout) {
  const int64_t N = ncells(d);
  double s[2] = {0, 0};
  int64_t n[2] = {0, 0};
  int xs[2] = {xin, xout};
  for (int k = 0; k < 2; k++)
    for (int z = 0; z < d->nz; z++)
      for (int y = 0; y < d->ny; y++) {
        int64_t i = (int64_t)xs[k] + (int64_t)d->nx * ((int64_t)y + (int64_t)d->ny * z);
        if (!(words[i] & LBO_FLUID)) continue;
        /* density deviations rho - 1 = sum (f_q - w_q): the rest density cancels in J, and summing
         * the small deviations keeps J's relative precision (a sum of rho ~ 1 would leave ~1e-16
         * absolute error in J) */
        double r = 0.0;
        for (int q = 0; q < LBO_Q; q++) r += f[(int64_t)q * N + i] - LBO_W[q];
        s[k] += r;
        n[k]++;
      }
  if (n[0] == 0 || n[1] == 0) return 0.0;
  return LBO_CS2 * (s[0] / (double)n[0] - s[1] / (double)n[1]);
}

void lbo_pressure_drop_grad(const lbo_dims* d, const uint32_t* words, int xin, int xout, double* lam) {
  const int64_t N = ncells(d);
  memset(lam, 0, sizeof(double) * (size_t)N * LBO_Q);
  int xs[2] = {xin, xout};
  for (int k = 0; k < 2; k++) {
    int64_t n = 0;
    for (int z = 0; z < d->nz; z++)
      for (int y = 0; y < d->ny; y++) {
        int64_t i = (int64_t)xs[k] + (int64_t)d->nx * ((int64_t)y + (int64_t)d->ny * z);
        if (words[i] & LBO_FLUID) n++;
      }
    if (n == 0) continue;
    double v = (k == 0 ? LBO_CS2 : -LBO_CS2) / (double)n;
    for (int z = 0; z < d->nz; z++)
      for (int y = 0; y < d->ny; y++) {
        int64_t i = (int64_t)xs[k] + (int64_t)d->nx * ((int64_t)y + (int64_t)d->ny * z);
        if (!(words[i] & LBO_FLUID)) continue;
        for (int q = 0; q < LBO_Q; q++) lam[(int64_t)q * N + i] += v;
      }
  }
}

/* Reverse mode of lbo_step (a hand-derived tape of one step; no reference counterpart).
 * Forward per fluid cell: rho = sum f, j = sum c f, u = j/rho, v = (1-a)(u + g/2),
 * post_q = f_q - w(f_q - feq_q(rho,v)) + S_q(rho,v), then push/bounce with the
 * moving-wall term -2 w_q rho (c_q.uw)/cs2 on moving links.
 * Reverse: mu_q = adjoint of post_q (gathered from where post_q was pushed),
 * Phi(rho,v) = sum_k mu_k (w feq_k + S_k), lam_q = (1-w) mu_q + dPhi/drho
 * + sum_d dPhi/dv_d (1-a)(c_qd - u_d)/rho - sum_moving mu_wall * 2 w rho' ...,
 * dJ/da = -sum_d dPhi/dv_d (u_d + g_d/2). */
void lbo_step_adjoint(const lbo_dims* d, const uint32_t* words, const lbo_params* p,
                      const double* src, const double* lam_out, double* lam_in, double* galpha,
                      int nthreads) {
  const int64_t N = ncells(d);
  const double tau_eff = lbo_tau_eff(p->tau);
  const double om = 1.0 / tau_eff;
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
  #pragma omp parallel for num_threads(nthreads) schedule(static) collapse(2)
#endif
  for (int z = 0; z < d->nz; z++)
    for (int y = 0; y < d->ny; y++) {
      for (int x = 0; x < d->nx; x++) {
        int64_t i = (int64_t)x + (int64_t)d->nx * ((int64_t)y + (int64_t)d->ny * z);
        uint32_t w = words[i];
        if (!(w & LBO_FLUID)) {
          for (int q = 0; q < LBO_Q; q++) lam_in[(int64_t)q * N + i] = 0.0;
          continue;
        }
        double f[LBO_Q], mu[LBO_Q];
        for (int q = 0; q < LBO_Q; q++) f[q] = src[(int64_t)q * N + i];
        unsigned cls = (w >> LBO_CLASS_SHIFT) & LBO_CLASS_MASK;
        double drho_wall = 0.0;
        for (int q = 0; q < LBO_Q; q++) {
          if (w & (1u << q)) {
            mu[q] = lam_out[(int64_t)(LBO_Q - 1 - q) * N + i];
            if (cls && ((p->moving_mask[cls] >> q) & 1u))
              drho_wall -= mu[q] * lbo_moving_wall_term(p->uw[cls], q, 1.0);
          } else {
            int64_t j = neighbour(d, x, y, z, q);
            mu[q] = lam_out[(int64_t)q * N + j];
          }
        }
        double rho = 0.0, j3[3] = {0, 0, 0};
        for (int q = 0; q < LBO_Q; q++) {
          rho += f[q];
          for (int k = 0; k < 3; k++) j3[k] += cd(q, k) * f[q];
        }
        double u[3], g[3] = {0, 0, 0};
        if (p->use_force) for (int k = 0; k < 3; k++) g[k] = p->g[k];
        double a = p->alpha ? p->alpha[i] : 0.0;
        double s = p->alpha ? (1.0 - a) : 1.0;
        double v[3];
        for (int k = 0; k < 3; k++) {
          u[k] = j3[k] / rho;
          v[k] = s * (u[k] + 0.5 * g[k]);
        }
        double vv = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        double fpref = p->use_force ? (1.0 - 0.5 * om) : 0.0;
        double dPhi_drho = 0.0, dPhi_dv[3] = {0, 0, 0};
        for (int q = 0; q < LBO_Q; q++) {
          double c[3] = {cd(q, 0), cd(q, 1), cd(q, 2)};
          double cv = c[0] * v[0] + c[1] * v[1] + c[2] * v[2];
          double cg = c[0] * g[0] + c[1] * g[1] + c[2] * g[2];
          double vg = v[0] * g[0] + v[1] * g[1] + v[2] * g[2];
          double feq_over_rho = LBO_W[q] * (1.0 + 3.0 * cv + 4.5 * cv * cv - 1.5 * vv);
          double S_over_rho = fpref * LBO_W[q] * (3.0 * (cg - vg) + 9.0 * cv * cg);
          dPhi_drho += mu[q] * (om * feq_over_rho + S_over_rho);
          for (int k = 0; k < 3; k++) {
            double dfeq = LBO_W[q] * rho * (3.0 * c[k] + 9.0 * cv * c[k] - 3.0 * v[k]);
            double dS = fpref * LBO_W[q] * rho * (-3.0 * g[k] + 9.0 * c[k] * cg);
            dPhi_dv[k] += mu[q] * (om * dfeq + dS);
          }
        }
        for (int q = 0; q < LBO_Q; q++) {
          double l = (1.0 - om) * mu[q] + dPhi_drho + drho_wall;
          for (int k = 0; k < 3; k++) l += dPhi_dv[k] * s * (cd(q, k) - u[k]) / rho;
          lam_in[(int64_t)q * N + i] = l;
        }
        if (p->alpha && galpha) {
          double ga = 0.0;
          for (int k = 0; k < 3; k++) ga -= dPhi_dv[k] * (u[k] + 0.5 * g[k]);
          galpha[i] += ga;
        }
      }
    }
}

/* ---- restriction and prolongation between grid levels (BLB/BlockRestrictionProlongation.cpp) ----
 * Blocks are the reference's 6^3 vertex blocks, populations AoS: pdf[19 * cell + q] with
 * cell = x + 6 y + 36 z.  A refined coarse cell holds 4^3 fine vertices at child positions
 * pos in {0..3}^3; the fine block at pos overlaps the 2^3 coarse cells of one sub-block of
 * its father vertex's block, each coarse cell covering 3^3 fine cells. */
#define LBO_BS 6
#define LBO_BC (LBO_BS * LBO_BS * LBO_BS)

/* getFatherVertex (:17-28): pos / 2 per axis. */
void lbo_amr_father(const int pos[3], int out[3]) {
  for (int d = 0; d < 3; d++) out[d] = pos[d] / 2;
}

/* getSubBlock with the table of createSubBlockVector (:264-288): positions 0..3 -> 1, 2, 0, 1. */
void lbo_amr_subblock(const int pos[3], int out[3]) {
  static const int t[4] = {1, 2, 0, 1};
  for (int d = 0; d < 3; d++) out[d] = t[pos[d]];
}

/* Whether entry (cell, q) of a block is filled from neighbour block nb (nb = bx + 3 by + 9 bz,
 * 0..26 without 13) by the three streaming steps of createMinusOneEntriesFromNeighbourBlock
 * (:293-449): the reference marks the neighbour of a 3^3-block domain with -1, streams three
 * times (populations leaving the domain are lost, empty slots become 0) and lists the centre
 * entries that hold -1.  Streaming moves population q by c_q per step, so the entry holds -1
 * exactly when its position 6 + x - 3 c_q lies inside the neighbour along every axis. */
int lbo_amr_from_neighbour(int nb, int cell, int q) {
  const int b[3] = {nb % 3, (nb / 3) % 3, nb / 9};
  const int x[3] = {cell % LBO_BS, (cell / LBO_BS) % LBO_BS, cell / (LBO_BS * LBO_BS)};
  for (int d = 0; d < 3; d++) {
    const int s = LBO_BS + x[d] - 3 * LBO_C[q][d];
    if (s < LBO_BS * b[d] || s >= LBO_BS * (b[d] + 1)) return 0;
  }
  return 1;
}

/* findFineOverlapCells (:32-62) with neighbourIsTemporaryVertex (:66-79): a neighbour is a
 * temporary vertex if any entry it fills holds exactly -1.0; then every cell it fills is an
 * overlap cell. */
void lbo_amr_find_overlap(const double* fine, uint8_t* overlap) {
  for (int c = 0; c < LBO_BC; c++) overlap[c] = 0;
  for (int nb = 0; nb < 27; nb++) {
    if (nb == 13) continue;
    int temporary = 0;
    for (int c = 0; c < LBO_BC && !temporary; c++)
      for (int q = 0; q < LBO_Q; q++)
        if (lbo_amr_from_neighbour(nb, c, q) && fine[LBO_Q * c + q] == -1.0) {
          temporary = 1;
          break;
        }
    if (!temporary) continue;
    for (int c = 0; c < LBO_BC; c++)
      for (int q = 0; q < LBO_Q; q++)
        if (lbo_amr_from_neighbour(nb, c, q)) overlap[c] = 1;
  }
}

/* prolongateToFineLevel (:83-160): every overlap cell of the fine block takes the populations of
 * the coarse cell above it; its density is their sum (q order) and its velocity the momentum
 * (accumulated per axis in q order) divided by the density. */
void lbo_amr_prolongate(const int pos[3], const double* coarse, double* fine, double* density, double* velocity,
                        const uint8_t* overlap) {
  int sb[3];
  lbo_amr_subblock(pos, sb);
  const int h = LBO_BS / 3;
  for (int z = 0; z < h; z++)
    for (int y = 0; y < h; y++)
      for (int x = 0; x < h; x++) {
        const int cc = sb[0] * h + x + (sb[1] * h + y) * LBO_BS + (sb[2] * h + z) * LBO_BS * LBO_BS;
        for (int zz = 0; zz < 3; zz++)
          for (int yy = 0; yy < 3; yy++)
            for (int xx = 0; xx < 3; xx++) {
              const int fc = x * 3 + xx + (y * 3 + yy) * LBO_BS + (z * 3 + zz) * LBO_BS * LBO_BS;
              if (!overlap[fc]) continue;
              double rho = 0.0, m[3] = {0.0, 0.0, 0.0};
              for (int q = 0; q < LBO_Q; q++) {
                const double v = coarse[LBO_Q * cc + q];
                fine[LBO_Q * fc + q] = v;
                rho += v;
                for (int d = 0; d < 3; d++) m[d] += v * cd(q, d);
              }
              density[fc] = rho;
              for (int d = 0; d < 3; d++) velocity[3 * fc + d] = m[d] / rho;
            }
      }
}

/* restrictToCoarseLevel (:164-259): each coarse cell of the sub-block takes, per population, the
 * mean (sum in z, y, x order times 1/27) of its 3^3 fine cells, unless one of them holds the
 * marker -1.0 for that population; the selection is the reference's arithmetic blend
 * b * coarse + (1 - b) * mean with b = 1.0 or 0.0. */
void lbo_amr_restrict(const int pos[3], double* coarse, const double* fine) {
  int sb[3];
  lbo_amr_subblock(pos, sb);
  const int h = LBO_BS / 3;
  for (int z = 0; z < h; z++)
    for (int y = 0; y < h; y++)
      for (int x = 0; x < h; x++) {
        const int cc = sb[0] * h + x + (sb[1] * h + y) * LBO_BS + (sb[2] * h + z) * LBO_BS * LBO_BS;
        double r[LBO_Q];
        int marked[LBO_Q];
        for (int q = 0; q < LBO_Q; q++) {
          r[q] = 0.0;
          marked[q] = 0;
        }
        for (int zz = 0; zz < 3; zz++)
          for (int yy = 0; yy < 3; yy++)
            for (int xx = 0; xx < 3; xx++) {
              const int fc = x * 3 + xx + (y * 3 + yy) * LBO_BS + (z * 3 + zz) * LBO_BS * LBO_BS;
              for (int q = 0; q < LBO_Q; q++) {
                const double v = fine[LBO_Q * fc + q];
                marked[q] = marked[q] || v == -1.0;
                r[q] += v;
              }
            }
        for (int q = 0; q < LBO_Q; q++) {
          const double mean = (1.0 / 27.0) * r[q];
          const double b = (double)marked[q];
          coarse[LBO_Q * cc + q] = b * coarse[LBO_Q * cc + q] + (double)(1 - marked[q]) * mean;
        }
      }
}

/* ---- the block solver's per-block and per-Peano-cell operators on 6^3 blocks ----------------
 * Blocks as the reference's GridManagementService keeps them: pdf[19 * cell + q], density[216],
 * velocity[648], inner[216], cell = x + 6 y + 36 z. */

/* computeDensity (BLB/BlockComputeCellValues.cpp:26-43) and computeVelocity (:47-68): inner cells
 * without a -1.0 entry (minus_one NULL = none) take rho = sum of f in q order and
 * u = (1.0/rho) * sum f c; every other cell keeps its density and gets u = 0. */
void lbo_block_cell_values(const uint8_t* inner, const uint8_t* minus_one, const double* pdf, double* density,
                           double* velocity) {
  for (int c = 0; c < LBO_BC; c++) {
    const int act = inner[c] && !(minus_one && minus_one[c]);
    if (act) density[c] = lbo_density(pdf + LBO_Q * c);
    double u[3] = {0.0, 0.0, 0.0};
    if (act) lbo_velocity(pdf + LBO_Q * c, density[c], u);
    for (int d = 0; d < 3; d++) velocity[3 * c + d] = u[d];
  }
}

/* CollisionModelWrapper::collide (collisionmodels/CollisionModelWrapper.cpp:28-39) with BlockBGK
 * (BlockBGK.cpp:18-45): inner cells without a -1.0 entry relax toward the equilibrium of the stored
 * density and velocity, f -= 1.0/tau (f - feq); tau is the value the BGK model holds. */
void lbo_block_collide(const uint8_t* inner, const uint8_t* minus_one, const double* density,
                       const double* velocity, double tau, double* pdf) {
  for (int c = 0; c < LBO_BC; c++) {
    if (!inner[c] || (minus_one && minus_one[c])) continue;
    double feq[LBO_Q];
    lbo_feq(density[c], velocity + 3 * c, feq);
    lbo_bgk(pdf + LBO_Q * c, tau, feq);
  }
}

/* One boundary link after the collision, as BlockBoundaryConditionManager::treatBoundaryAfterCollision
 * (boundaryconditions/BlockBoundaryConditionManager.cpp:47-62) dispatches it by BoundaryData type
 * (LB/BoundaryData.h:44-61): 0 halfway bounce-back (no-op), 1 moving wall and 5 moving obstacle
 * (BlockMovingWall.cpp:13-33), 2 pressure outlet, 3 pressure inlet, 4 velocity inlet; q_in is the
 * incoming direction (boundaryDataIndex(1)) and the outgoing slot 18 - q_in is written, with the
 * rules of lbo_boundary_value and the cell's stored density and velocity. */
void lbo_block_boundary(int cell, int q_in, int type, const double* u_l, double rho_l, const double* density,
                        const double* velocity, double* pdf) {
  const int q = LBO_Q - 1 - q_in;
  double* f = pdf + LBO_Q * cell + q;
  const double rho = density[cell];
  if (type == 1 || type == 5) {
    *f -= lbo_moving_wall_term(u_l, q, rho);
    return;
  }
  if (type < 2 || type > 4) return;
  double feq[LBO_Q];
  if (type == 4) {
    lbo_feq(rho_l, u_l, feq);
    *f = feq[q_in];
    return;
  }
  if (type == 2) lbo_feq(rho_l, velocity + 3 * cell, feq);
  else lbo_feq(rho, u_l, feq);
  *f = -*f + feq[q_in] + feq[q];
}

/* createIndexForLocalStreaming (nonlocaltreatment/AABlockStreaming.cpp:111-186): cells in z, y, x
 * order and q in order; (cell, q) whose neighbour cell + c_q lies in the block is listed unless q is
 * the rest population or its partner (neighbour, 18 - q) is listed already.  Entry k is
 * {19 cell + q, 19 nb + 18 - q, cell, nb}.  Returns the number of entries (1440) and writes at most
 * max of them. */
int lbo_block_local_index(int* out, int max) {
  static uint8_t listed[LBO_BC * LBO_Q];
  memset(listed, 0, sizeof listed);
  int n = 0;
  for (int z = 0; z < LBO_BS; z++)
    for (int y = 0; y < LBO_BS; y++)
      for (int x = 0; x < LBO_BS; x++) {
        const int c = x + LBO_BS * (y + LBO_BS * z);
        for (int q = 0; q < LBO_Q; q++) {
          const int nx = x + LBO_C[q][0], ny = y + LBO_C[q][1], nz = z + LBO_C[q][2];
          if (nx < 0 || nx >= LBO_BS || ny < 0 || ny >= LBO_BS || nz < 0 || nz >= LBO_BS) continue;
          const int nb = nx + LBO_BS * (ny + LBO_BS * nz);
          if (q == LBO_REST || listed[LBO_Q * nb + LBO_Q - 1 - q]) continue;
          listed[LBO_Q * c + q] = 1;
          if (n < max) {
            int* e = out + 4 * n;
            e[0] = LBO_Q * c + q;
            e[1] = LBO_Q * nb + LBO_Q - 1 - q;
            e[2] = c;
            e[3] = nb;
          }
          n++;
        }
      }
  return n;
}

/* streamingLocal (:19-32): each listed pair is swapped when both cells are inner. */
void lbo_block_stream_local(const uint8_t* inner, double* pdf) {
  static int idx[4 * 2048];
  const int n = lbo_block_local_index(idx, 2048);
  for (int k = 0; k < n; k++) {
    const int* e = idx + 4 * k;
    if (!(inner[e[2]] && inner[e[3]])) continue;
    const double t = pdf[e[0]];
    pdf[e[0]] = pdf[e[1]];
    pdf[e[1]] = t;
  }
}

/* swapPdfsInsideCells (:80-92): in every cell, inner or not, f_q <-> f_{18-q} for q < 9. */
void lbo_block_swap_inside(double* pdf) {
  for (int c = 0; c < LBO_BC; c++)
    for (int q = 0; q < LBO_Q / 2; q++) {
      const double t = pdf[LBO_Q * c + q];
      pdf[LBO_Q * c + q] = pdf[LBO_Q * c + LBO_Q - 1 - q];
      pdf[LBO_Q * c + LBO_Q - 1 - q] = t;
    }
}

/* getNonLocalStreamingIndex (nonlocaltreatment/CellTopology.cpp:435-465) with the even-block-size
 * search (:317-418), stated on the Peano cell's own 6^3 sub-cells: vertex v (lexicographic, bit d =
 * side along axis d) owns the octant [3 v_d, 3 v_d + 3) of the cell, and sub-cell s of that octant is
 * cell s_d + 3 - 6 v_d of v's block.  For vertex pairs i < j, sub-cells s of i's octant in z, y, x
 * order and q in order, the entry {i, j, cell of s in block i, cell of s + c_q in block j, q} is
 * listed when s + c_q falls on j's side of the cell's centre planes along every axis (a target
 * outside the cell keeps the side of s).  Returns the number of entries (504). */
int lbo_block_nonlocal_index(int* out, int max) {
  int n = 0;
  for (int i = 0; i < 8; i++)
    for (int j = i + 1; j < 8; j++) {
      const int vi[3] = {i & 1, (i >> 1) & 1, i >> 2}, vj[3] = {j & 1, (j >> 1) & 1, j >> 2};
      for (int sz = 3 * vi[2]; sz < 3 * vi[2] + 3; sz++)
        for (int sy = 3 * vi[1]; sy < 3 * vi[1] + 3; sy++)
          for (int sx = 3 * vi[0]; sx < 3 * vi[0] + 3; sx++) {
            const int s[3] = {sx, sy, sz};
            for (int q = 0; q < LBO_Q; q++) {
              int t[3], side = 1;
              for (int d = 0; d < 3; d++) {
                t[d] = s[d] + LBO_C[q][d];
                side = side && ((t[d] >= 3) == vj[d]);
              }
              if (!side) continue;
              if (n < max) {
                int* e = out + 5 * n;
                e[0] = i;
                e[1] = j;
                e[2] = (s[0] + 3 - 6 * vi[0]) + LBO_BS * ((s[1] + 3 - 6 * vi[1]) + LBO_BS * (s[2] + 3 - 6 * vi[2]));
                e[3] = (t[0] + 3 - 6 * vj[0]) + LBO_BS * ((t[1] + 3 - 6 * vj[1]) + LBO_BS * (t[2] + 3 - 6 * vj[2]));
                e[4] = q;
              }
              n++;
            }
          }
    }
  return n;
}

/* streamingNonLocal (AABlockStreaming.cpp:37-53) on the 8 blocks of one Peano cell (lexicographic
 * vertex order): entry {i, j, c1, c2, q} swaps pdf_i[19 c1 + q] with pdf_j[19 c2 + 18 - q] when
 * c1 is inner in block i and c2 in block j. */
void lbo_block_stream_nonlocal(const uint8_t* const* inner, double* const* pdf) {
  static int idx[5 * 1024];
  const int n = lbo_block_nonlocal_index(idx, 1024);
  for (int k = 0; k < n; k++) {
    const int* e = idx + 5 * k;
    if (!(inner[e[0]][e[2]] && inner[e[1]][e[3]])) continue;
    double* a = pdf[e[0]] + LBO_Q * e[2] + e[4];
    double* b = pdf[e[1]] + LBO_Q * e[3] + LBO_Q - 1 - e[4];
    const double t = *a;
    *a = *b;
    *b = t;
  }
}

/* The vertex-overlap exchange between ranks (compiled in the reference's Parallel build only).
 * parallel::BlockBoundaryIndex4Pdfs::initBoundaryIndex (parallel/BlockBoundaryIndex4Pdfs.cpp:30-77):
 * for cells in z, y, x order and q in order, every population whose target cell + c_q lies outside
 * the 6^3 block is listed with its incoming slot 19 cell + 18 - q and its outgoing slot 19 cell + q.
 * Returns the number of entries (1,008). */
int lbo_block_pdf_boundary(int* incoming, int* outgoing, int max) {
  int n = 0;
  for (int z = 0; z < LBO_BS; z++)
    for (int y = 0; y < LBO_BS; y++)
      for (int x = 0; x < LBO_BS; x++)
        for (int q = 0; q < LBO_Q; q++) {
          const int t[3] = {x + LBO_C[q][0], y + LBO_C[q][1], z + LBO_C[q][2]};
          int out = 0;
          for (int d = 0; d < 3; d++) out = out || t[d] < 0 || t[d] > LBO_BS - 1;
          if (!out) continue;
          const int c = x + LBO_BS * (y + LBO_BS * z);
          if (n < max) {
            incoming[n] = LBO_Q * c + LBO_Q - 1 - q;
            outgoing[n] = LBO_Q * c + q;
          }
          n++;
        }
  return n;
}

/* BlockBoundaryIndex (BlockBoundaryIndex.cpp:15-88): the cells with a coordinate below 2 or above
 * 3, in increasing block index.  Returns the number of cells (208). */
int lbo_block_boundary_cells(int* cells, int max) {
  int n = 0;
  for (int z = 0; z < LBO_BS; z++)
    for (int y = 0; y < LBO_BS; y++)
      for (int x = 0; x < LBO_BS; x++) {
        if (!(x < 2 || x > LBO_BS - 3 || y < 2 || y > LBO_BS - 3 || z < 2 || z > LBO_BS - 3)) continue;
        if (n < max) cells[n] = x + LBO_BS * (y + LBO_BS * z);
        n++;
      }
  return n;
}

/* RegularGridBlockVertex::initPdfDiff (BLB/RegularGridBlockVertex.cpp:96-105): after the collision
 * and the boundary conditions, before the local streaming, diff_i = -pdf[outgoing_i]. */
void lbo_block_init_pdf_diff(const double* pdf, double* diff) {
  static int in[1024], out[1024];
  const int n = lbo_block_pdf_boundary(in, out, 1024);
  for (int i = 0; i < n; i++) diff[i] = -pdf[out[i]];
}

/* computePdfDiff (:108-120): after the in-cell swap, diff_i += pdf[incoming_i].  Where no local
 * Peano cell swapped the link, the in-cell swap has put the outgoing value back into the incoming
 * slot and the difference is zero. */
void lbo_block_compute_pdf_diff(const double* pdf, double* diff) {
  static int in[1024], out[1024];
  const int n = lbo_block_pdf_boundary(in, out, 1024);
  for (int i = 0; i < n; i++) diff[i] = diff[i] + pdf[in[i]];
}

/* mergePdfDiff (:123-141): with the difference a neighbour rank computed for its copy of the same
 * vertex, pdf[incoming_i] += neighbour_i wherever the own difference is exactly 0.0. */
void lbo_block_merge_pdf_diff(double* pdf, const double* own, const double* neighbour) {
  static int in[1024], out[1024];
  const int n = lbo_block_pdf_boundary(in, out, 1024);
  for (int i = 0; i < n; i++)
    if (own[i] == 0.0) pdf[in[i]] = pdf[in[i]] + neighbour[i];
}

/* setDensityOnBlockBoundary (:56-77): with a dynamic geometry, the density of each block-boundary
 * cell that is not inner but on the geometry boundary, 0.0 for the others; without one the buffer
 * keeps its values (resetDensityOnBlockBoundary, :51-53, zeroes it before). */
void lbo_block_set_density_on_boundary(const double* density, const uint8_t* inner, const uint8_t* boundary,
                                       int dynamic_geometry, double* out) {
  if (!dynamic_geometry) return;
  static int cells[256];
  const int n = lbo_block_boundary_cells(cells, 256);
  for (int i = 0; i < n; i++) {
    const int c = cells[i];
    out[i] = (!inner[c] && boundary[c]) ? density[c] : 0.0;
  }
}

/* mergeDensityOnBoundary (:144-160): density[c] += neighbour_i for each block-boundary cell that
 * is not inner but on the geometry boundary. */
void lbo_block_merge_density_on_boundary(double* density, const uint8_t* inner, const uint8_t* boundary,
                                         const double* neighbour) {
  static int cells[256];
  const int n = lbo_block_boundary_cells(cells, 256);
  for (int i = 0; i < n; i++) {
    const int c = cells[i];
    if (!inner[c] && boundary[c]) density[c] += neighbour[i];
  }
}
