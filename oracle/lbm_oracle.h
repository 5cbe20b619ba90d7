/*
 * CPU oracle for the D3Q19 BGK block-LB hot path (TEST INFRASTRUCTURE ONLY).
 *
 * This is a plain-C restatement of the reference Peano block-LB arithmetic, used
 * only to check results against, by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  The product path (madpac_amd/) never links, loads or calls it.
 *
 * Every function reproduces the reference operation order so that the FP64 result
 * is bit-identical to the reference on x86-64 (no FMA contraction: build with
 * -ffp-contract=off).  Citations are relative to
 * madpac/peano/applications/latticeboltzmann/ (LB/) and its blocklatticeboltzmann/
 * subfolder (BLB/).
 *
 * Pinning (see tests/test_oracle.py): the lattice tables, moments, equilibrium and
 * BGK relaxation are pinned by closed-form identities; the full step (collide,
 * swap streaming with implicit halfway bounce-back, moving wall) and the restriction /
 * prolongation between grid levels and the block solver's per-block operators (cell values,
 * collision, boundary links, local / cross-block / in-cell swap streaming and their index tables)
 * are pinned bit-for-bit against fixtures produced by
 * the compiled reference (tests/golden/), the overlap map of the latter also against a
 * brute-force run of the reference's own three-step streaming procedure.  The vertex exchange
 * between ranks (compiled by the reference only in its MPI build) has its two index tables pinned
 * to the compiled reference's tables (tests/golden/parallel_index.npz) and the difference / merge
 * rules pinned by the partition invariant: Peano cells split among ranks, each rank's copies of the
 * blocks, differences exchanged, every copy equal to the single-domain run bit for bit, with
 * planted mistakes in the rule failing it (tests/test_oracle_parallel.py).  Guo forcing, Brinkman porosity, periodic wrap, the pressure-drop
 * objective and the adjoint step go beyond the reference and are pinned by analytic
 * solutions (Poiseuille flow, Taylor-Green decay, the uniform-porosity fixed point
 * u + g/2 = g tau_eff / alpha, the objective of a linear density ramp), conservation
 * laws and central finite differences.
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBO_Q 19
#define LBO_REST 9
/* cs^2 exactly as the reference spells it (LB/LatticeBoltzmannDefinitions.h:89). */
#define LBO_CS2 (1.0 / 3.0)

/* Cell word: bit 31 fluid, bit q (0..18) = neighbour x+c_q is not fluid (bounce-back),
 * bits 24..28 = boundary class (0 = none, classes 1..31). */
#define LBO_FLUID 0x80000000u
#define LBO_CLASS_SHIFT 24
#define LBO_CLASS_MASK 0x1Fu
#define LBO_MAX_CLASSES 32

/* Boundary condition of a class's links (LB/BoundaryData.h:44-61 and the handlers
 * registered in BLB/boundaryconditions/BlockBoundaryConditionManager.cpp:15-20).  Links
 * outside a class's mask, and cells of class 0, are plain halfway bounce-back; a moving
 * wall with u_w = 0 is the same.  LBO_BC_CELL_VELOCITY or-ed into a type takes the target
 * velocity per cell from bc_u. */
#define LBO_BC_MOVING_WALL 0
#define LBO_BC_PRESSURE_OUTLET 1
#define LBO_BC_PRESSURE_INLET 2
#define LBO_BC_VELOCITY_INLET 3
#define LBO_BC_TYPE_MASK 15
#define LBO_BC_CELL_VELOCITY 16

/* MD coupling modes (madpac/peanocoupling/PeanoDataExchangeService.cpp:69-158): the MD velocity
 * replaces the LB velocity, or drives a Dirichlet-type forcing toward it. */
#define LBO_COUPLE_OVERWRITE 0
#define LBO_COUPLE_DIRICHLET 1

extern const int LBO_C[LBO_Q][3];
extern const double LBO_W[LBO_Q];

typedef struct {
  int nx, ny, nz;
  int periodic[3];
} lbo_dims;

typedef struct {
  double tau;                              /* configured tau (LB/configurations/...:269) */
  double uw[LBO_MAX_CLASSES][3];           /* per class: wall velocity / target velocity u_L */
  uint32_t moving_mask[LBO_MAX_CLASSES];   /* bit q: outgoing link q of a class-k cell takes the class's condition */
  double g[3];                             /* body acceleration (Guo forcing); 0 = off */
  int use_force;
  const double* alpha;                     /* per-cell Brinkman porosity or NULL */
  const double* tau_cell;                  /* per-cell configured tau (overrides tau) or NULL */
  int bc_type[LBO_MAX_CLASSES];            /* LBO_BC_* per class */
  double bc_rho[LBO_MAX_CLASSES];          /* per class: target density rho_L */
  const double* bc_u;                      /* per-cell target velocity [3*N] for LBO_BC_CELL_VELOCITY classes */
  const int64_t* couple_row;               /* per cell: row of the coupling data, -1 = not coupled; NULL = none */
  int couple_mode;                         /* LBO_COUPLE_* */
  const double* couple_u;                  /* [3*rows] MD velocity in lattice units */
  const double* couple_rho;                /* [rows] MD density, or NULL (no density coupling) */
  double couple_dt;                        /* LB time step of the Dirichlet forcing */
} lbo_params;

/* a1/a8 */
double lbo_tau_eff(double tau);
/* a4-a7 (per cell, AoS f[19]) */
double lbo_density(const double* f);
void lbo_velocity(const double* f, double rho, double* u);
void lbo_feq(double rho, const double* u, double* feq);
void lbo_bgk(double* f, double tau_eff, const double* feq);
/* a13: the amount subtracted from outgoing population q of a moving-wall link */
double lbo_moving_wall_term(const double* uw, int q, double rho);

/* a13 generalised: the value stored into the outgoing slot q of a boundary link of class k
 * at cell i, given the post-collision population fq and the cell's rho and velocity u. */
double lbo_boundary_value(const lbo_params* p, unsigned k, int64_t i, int q, double fq, double rho,
                          const double* u);

/* the reference's coupling force of one cell (PeanoDataExchangeService.cpp:161-186) */
void lbo_couple_gi(double rho, double tau, const double* u, const double* g, double* gi);

/* geometry: lid-driven cavity of n^3 fluid cells, lid on +z moving with (ulid,0,0). */
void lbo_build_cavity(int n, uint32_t* words, lbo_params* p, double ulid);
/* geometry: the reference's default channel on the whole box (walls on the y/z faces,
 * velocity inlet on x-, pressure outlet rho = 1 on x+), solid cells as obstacles.
 * words must come from lbo_build_words on a non-periodic box; bc_u receives the per-cell
 * inlet velocity (3*N doubles) when parabolic != 0 and is then referenced by p. */
void lbo_build_channel(const lbo_dims* d, uint32_t* words, lbo_params* p, double umean, int parabolic,
                       double* bc_u);
/* generic words from a solid mask (1 = solid); moving-wall classes left at 0. */
void lbo_build_words(const lbo_dims* d, const uint8_t* solid, uint32_t* words);

/* one time step, SoA layout f[q*N + cell], push form of the reference swap streaming. */
void lbo_step(const lbo_dims* d, const uint32_t* words, const lbo_params* p,
              const double* src, double* dst, int nthreads);
/* the reference's own streaming algorithm on a dense lattice (collide, pair swaps
 * along fluid-fluid links, then in-cell swap q <-> 18-q), in place on AoS f[cell*19+q]. */
void lbo_step_swap_aos(const lbo_dims* d, const uint32_t* words, const lbo_params* p, double* f);

/* pure streaming (no collision) in push form with implicit halfway bounce-back */
void lbo_stream(const lbo_dims* d, const uint32_t* words, const double* src, double* dst);

/* fields: rho, u (u = 0 and rho untouched = 0 on non-fluid cells) */
void lbo_fields(const lbo_dims* d, const uint32_t* words, const double* f, double* rho, double* u);
/* MD export of getAndSetV: rho[k], u[3k+d] of cells[k] after the coupling's GET block */
void lbo_couple_export(const lbo_dims* d, const uint32_t* words, const lbo_params* p, const double* f,
                       const int64_t* cells, int64_t n, double* rho, double* u);
/* equilibrium initialisation from rho/u fields (non-fluid cells: 0) */
void lbo_init_equilibrium(const lbo_dims* d, const uint32_t* words, const double* rho,
                          const double* u, double* f);
/* global mass and momentum over fluid cells (Kahan-free plain double sums, cell order) */
void lbo_moments(const lbo_dims* d, const uint32_t* words, const double* f, double out[4]);

/* pressure drop J = cs^2 (<rho>_{x=xin} - <rho>_{x=xout}) over fluid cells of two planes */
double lbo_pressure_drop(const lbo_dims* d, const uint32_t* words, const double* f, int xin, int xout);
/* reverse-mode (adjoint) of lbo_step: given src (the step's input state) and lam_out =
 * dJ/d(dst), returns lam_in = dJ/d(src) and accumulates dJ/dalpha into galpha. */
void lbo_step_adjoint(const lbo_dims* d, const uint32_t* words, const lbo_params* p,
                      const double* src, const double* lam_out, double* lam_in, double* galpha,
                      int nthreads);
/* dJ/df of the pressure drop */
void lbo_pressure_drop_grad(const lbo_dims* d, const uint32_t* words, int xin, int xout, double* lam);

/* Restriction and prolongation between grid levels on 6^3 blocks (AoS pdf[19 * cell + q]),
 * BLB/BlockRestrictionProlongation.cpp; pos is the fine vertex's child position in {0..3}^3. */
void lbo_amr_father(const int pos[3], int out[3]);
void lbo_amr_subblock(const int pos[3], int out[3]);
int lbo_amr_from_neighbour(int nb, int cell, int q);
void lbo_amr_find_overlap(const double* fine, uint8_t* overlap);
void lbo_amr_prolongate(const int pos[3], const double* coarse, double* fine, double* density, double* velocity,
                        const uint8_t* overlap);
void lbo_amr_restrict(const int pos[3], double* coarse, const double* fine);

/* The block solver's operators on 6^3 blocks (BLB/BlockComputeCellValues.cpp, BLB/collisionmodels/,
 * BLB/boundaryconditions/, BLB/nonlocaltreatment/AABlockStreaming.cpp and CellTopology.cpp). */
void lbo_block_cell_values(const uint8_t* inner, const uint8_t* minus_one, const double* pdf, double* density,
                           double* velocity);
void lbo_block_collide(const uint8_t* inner, const uint8_t* minus_one, const double* density,
                       const double* velocity, double tau, double* pdf);
void lbo_block_boundary(int cell, int q_in, int type, const double* u_l, double rho_l, const double* density,
                        const double* velocity, double* pdf);
int lbo_block_local_index(int* out, int max);
void lbo_block_stream_local(const uint8_t* inner, double* pdf);
void lbo_block_swap_inside(double* pdf);
int lbo_block_nonlocal_index(int* out, int max);
void lbo_block_stream_nonlocal(const uint8_t* const* inner, double* const* pdf);

/* The vertex-overlap exchange between ranks (BLB/RegularGridBlockVertex.cpp:51-160,
 * BLB/parallel/BlockBoundaryIndex4Pdfs.cpp, BLB/BlockBoundaryIndex.cpp): 1,008 boundary
 * populations and 208 block-boundary cells per block. */
int lbo_block_pdf_boundary(int* incoming, int* outgoing, int max);
int lbo_block_boundary_cells(int* cells, int max);
void lbo_block_init_pdf_diff(const double* pdf, double* diff);
void lbo_block_compute_pdf_diff(const double* pdf, double* diff);
void lbo_block_merge_pdf_diff(double* pdf, const double* own, const double* neighbour);
void lbo_block_set_density_on_boundary(const double* density, const uint8_t* inner, const uint8_t* boundary,
                                       int dynamic_geometry, double* out);
void lbo_block_merge_density_on_boundary(double* density, const uint8_t* inner, const uint8_t* boundary,
                                         const double* neighbour);

#ifdef __cplusplus
}
#endif
#endif
