/*
 * lbm_capi.h -- C ABI of the MI355X-native D3Q19 block-LB solver.
 *
 * Drop-in boundary for the Peano block lattice-Boltzmann hot path of MaDPaC.  Every
 * entry point is plain C (extern "C", plain pointers and sizes, int status) so that the
 * reference's C++ operator API (madpac_amd/csrc/peano_blocklb.h) and any FFI (ctypes,
 * cgo, JNI) can bind it.  Citations are relative to
 * madpac/peano/applications/latticeboltzmann/ (LB/) and LB/blocklatticeboltzmann/ (BLB/).
 *
 * Conventions
 *   - D3Q19 ordering and weights as LB/LatticeBoltzmannDefinitions.cpp:113-151
 *     (z-major, rest population q = 9, inverse direction 18 - q).
 *   - Cell index i = x + nx*(y + ny*z); populations are SoA f[q*ncells + i].
 *   - Populations exchanged through this ABI are always FP64 host arrays holding the
 *     reference's state (post-streaming distributions, as GridManagementService holds them
 *     between time steps); FP32-shifted contexts convert on the device.
 *   - Return value 0 = success, negative = error; lbm_last_error() gives the text.
 *     A context is thread-compatible (one host thread at a time), not thread-safe.
 *   - All device work of a context is ordered on one HIP stream owned by the context.
 */
#ifndef LBM_CAPI_H
#define LBM_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_Q 19

enum lbm_status {
  LBM_OK = 0,
  LBM_ERR_ARG = -1,      /* invalid argument (null pointer, size, range) */
  LBM_ERR_STATE = -2,    /* call not valid in the current state (e.g. no flags set) */
  LBM_ERR_HIP = -3,      /* a HIP runtime call failed */
  LBM_ERR_NOMEM = -4,    /* device allocation failed */
  LBM_ERR_NUMERIC = -5   /* non-finite value detected by a reduction guard */
};

enum lbm_precision {
  LBM_FP64 = 0,             /* FP64 populations, FP64 collision (reference arithmetic order) */
  LBM_FP32_SHIFTED = 1,     /* FP32 populations stored as f - w_q (FP32 state and traffic); moments,
                               equilibrium and the relaxation product in FP64, differences and the
                               store in FP32 */
  LBM_FP32_SHIFTED_F32 = 2  /* FP32 populations as above with FP32 collision arithmetic: persistent
                               rounding of the equilibrium biases the momentum at low Mach */
};

typedef struct lbm_ctx lbm_ctx;

typedef struct {
  int nx, ny, nz;        /* lattice extent in cells */
  int periodic[3];       /* 1 = periodic along x / y / z */
  int precision;         /* enum lbm_precision */
  double tau;            /* configured relaxation time; tau_eff follows LB/MultiLevelSimData.h:160 */
  double force[3];       /* body acceleration for Guo forcing; all zero disables it */
  int device;            /* HIP device ordinal */
  int layout;            /* enum lbm_layout (0 = two buffers) */
  int cell_order;        /* enum lbm_cell_order (0 = automatic) */
  int populations;       /* enum lbm_populations (0 = automatic) */
} lbm_config;

/* Order of the cells in device memory; the C-ABI always exchanges row-major arrays.  ROWS:
 * x + nx (y + ny z).  BRICKS: 4 x 4 x 4 bricks of consecutive cells, one wavefront per brick: far
 * fewer 128-byte lines straddle solid surfaces (porous 256^3 FP64: +3 %), but shifted reads touch
 * more lines (all-fluid 512^3: 9-18 % slower).  AUTO decides at each lbm_set_flags: BRICKS for an
 * FP64 context whose flags would make row reads fetch at least 1.12 cells per fluid cell in
 * 128-byte lines (solid cells sharing lines with fluid ones), ROWS otherwise (FP32 measured
 * faster in rows even on the porous lattice). */
enum lbm_cell_order { LBM_CELL_ORDER_AUTO = 0, LBM_CELL_ORDER_ROWS = 1, LBM_CELL_ORDER_BRICKS = 2 };

/* Arrangement of the 19 populations in device memory (the C-ABI exchanges SoA arrays either way;
 * results are bit-identical).  PLANES: plane q holds population q of every stored cell.
 * INTERLEAVED: the 19 populations of each group of 64 consecutive stored cells (one wavefront's
 * cells) lie together, so a wavefront's accesses stay within a few kilobytes instead of 19 planes:
 * on the porous 256^3 channel FP64 +6 % (bricks and rows), FP32 rows +1 %, FP32 bricks -2 %; on
 * all-fluid 256^3 / 512^3 lattices +3 to +5 %.  AUTO is INTERLEAVED. */
enum lbm_populations { LBM_POPULATIONS_AUTO = 0, LBM_POPULATIONS_PLANES = 1, LBM_POPULATIONS_INTERLEAVED = 2 };

/* Storage of the state.  TWO_BUFFER: pull-form steps between two buffers.  AA: one buffer updated
 * in place, alternating a step that touches only the cell's own slots with one that pulls from and
 * pushes to the neighbours; half the memory at the same traffic (e.g. a 1024^3 FP64 lattice fits
 * one MI355X).  MD coupling and the adjoint work in both layouts with the same results (in place,
 * coupled cells' populations are saved before each step; the adjoint borrows a second buffer). */
enum lbm_layout { LBM_LAYOUT_TWO_BUFFER = 0, LBM_LAYOUT_AA = 1 };

/* ---- lifetime ---------------------------------------------------------------------- */
/* Replaces the regular-grid runner's set-up (BLB/runners/...ForRegularGrid.cpp:49-121) and
 * GridManagementService::init (BLB/services/GridManagementService.cpp:26-54). */
int lbm_create(lbm_ctx** ctx, const lbm_config* cfg);
int lbm_destroy(lbm_ctx* ctx);
const char* lbm_last_error(const lbm_ctx* ctx);
/* cells with the fluid ("inner") flag */
int64_t lbm_fluid_cells(const lbm_ctx* ctx);
/* the cell order in use (LBM_CELL_ORDER_ROWS or LBM_CELL_ORDER_BRICKS; -1 for a null context).  No
 * reference counterpart: the reference stores 6^3 blocks of AoS cells. */
int lbm_cell_order(const lbm_ctx* ctx);

/* ---- boundary flags ------------------------------------------------------------------ */
/* inner[i] != 0 marks a fluid cell: the reference's per-block `inner` bitset
 * (GridManagementService.cpp:91-111).  Links to non-inner cells are halfway bounce-back
 * (BLB/boundaryconditions/BlockHalfwayBounceBack.cpp:20-22, implicit in the swap streaming
 * of BLB/nonlocaltreatment/AABlockStreaming.cpp:19-53). */
int lbm_set_flags(lbm_ctx* ctx, const uint8_t* inner);
/* Boundary-condition types of a class: the reference's BoundaryData::BoundaryType
 * (LB/BoundaryData.h:44-61) with the handlers of
 * BLB/boundaryconditions/BlockBoundaryConditionManager.cpp:15-20.  Halfway bounce-back is the
 * default of every link to a non-fluid cell and needs no class. */
enum {
  LBM_BC_MOVING_WALL = 0,     /* BlockMovingWall.cpp:13-33: f_q -= 2 w_q rho (u_w . c_q) / cs2 */
  LBM_BC_PRESSURE_OUTLET = 1, /* BlockPressureOutletCondition.h: rho_L, the cell's velocity */
  LBM_BC_PRESSURE_INLET = 2,  /* BlockPressureInletCondition.h: u_L, the cell's density */
  LBM_BC_VELOCITY_INLET = 3   /* BlockVelocityInletEquilibrium.cpp:12-24: feq(rho_L, u_L) */
};
/* or-ed into an inlet type: u_L per cell from lbm_set_boundary_velocity */
#define LBM_BC_CELL_VELOCITY 16
/* Boundary class k (1..31): condition `type` with target density rho (rho_L) and velocity u
 * (u_w for a moving wall, u_L for an inlet) on the outgoing directions q in dir_mask whose
 * links leave the fluid; the cell's other links stay halfway bounce-back.  This replaces the
 * per-link BoundaryData list of the reference (SLB/services/ReceiveBoundaryDataService.cpp:
 * 107-185) and its dispatch in BlockBoundaryConditionManager::treatBoundaryAfterCollision
 * (:51-62).  The pressure conditions apply -f_q + feq_{18-q} + feq_q
 * (BLB/boundaryconditions/BlockAbstractPressureCondition.cpp:24-37). */
int lbm_set_boundary_class(lbm_ctx* ctx, int k, int type, double rho, const double u[3], uint32_t dir_mask);
/* Moving-wall class k (1..31): lbm_set_boundary_class with LBM_BC_MOVING_WALL and rho = 1. */
int lbm_set_moving_wall_class(lbm_ctx* ctx, int k, const double uw[3], uint32_t dir_mask);
/* cls[i] selects the class of each cell (0 = none); NULL clears all classes. */
int lbm_set_wall_classes(lbm_ctx* ctx, const uint8_t* cls);
/* Per-cell target velocities u[3*i + d] read by classes typed with LBM_BC_CELL_VELOCITY (the
 * reference's position-dependent BoundaryData, e.g. a parabolic inlet); NULL frees them. */
int lbm_set_boundary_velocity(lbm_ctx* ctx, const double* u);
/* The reference's lid-driven-cavity classification (SLB/BoundaryData4LidDrivenCavity.cpp:17-55
 * with SLB/services/ReceiveBoundaryDataService.cpp:107-185): links of the top fluid layer
 * with c_z > 0 that leave the fluid are moving walls with u_w = (ulid, 0, 0). */
int lbm_set_cavity_lid(lbm_ctx* ctx, double ulid);
/* The reference's default channel (SLB/AbstractBoundaryData4Channel.cpp:17-101,
 * SLB/BoundaryData4ChannelWithNoSlipBoundaries.cpp:17-27) on a non-periodic box: links leaving
 * through a y or z face are no-slip walls; other links leaving through x- are a velocity inlet
 * with u_L = (u_mean, 0, 0), or with parabolic != 0 the profile
 * u_mean * prod_{d=y,z} 6 a_d (L_d - a_d) / L_d^2 at the cell centre (lattice units), rho_L = 1;
 * links leaving through x+ are a pressure outlet with rho_L = 1.  Uses classes 1..18. */
int lbm_set_channel(lbm_ctx* ctx, double u_mean, int parabolic);
/* ---- MD <-> LB coupling (madpac/peanocoupling/PeanoDataExchangeService.cpp:69-218) ------ */
/* At a list of coupled cells the MD side either overwrites the equilibrium velocity (and, with
 * rho != NULL, the density) of the collision (getAndSetV :110-129), or, in the Dirichlet mode,
 * forces the cell toward the MD velocity: g = (u_MD - u)/dt, u <- u + dt/2 g, and after the BGK
 * relaxation f_i += dt g_i with the reference's computeGi (:161-186, its (c.u)^2 term included).
 * Velocities are in lattice units (the caller applies the reference's _lb2ms / _densityLB2MD).
 * Coupled cells use the plain collision: contexts with a body force or porosity are rejected. */
enum { LBM_COUPLE_OVERWRITE = 0, LBM_COUPLE_DIRICHLET = 1 };
/* the coupled cells (distinct indices; n = 0 clears) */
int lbm_set_coupling_cells(lbm_ctx* ctx, const int64_t* cells, int64_t n);
/* MD data applied by every following step until changed: u[3k+d], rho[k] (NULL: no density
 * coupling); u == NULL switches the coupling off.  dt is the Dirichlet forcing's time step. */
int lbm_set_coupling_data(lbm_ctx* ctx, int mode, const double* u, const double* rho, double dt);
/* The export to MD (getAndSetV :100-151 before its scaling by _lb2ms / _densityLB2MD): rho[k] and
 * u[3k+d] = (sum f c)/rho of the current state at the coupled cells, passed through the GET block
 * while coupling data is set, as the reference's SET block sends the values GET changed: u_MD in the
 * overwrite mode, u + dt/2 (u_MD - u)/dt in the Dirichlet mode, rho_MD with density coupling.  0 at
 * non-fluid cells; either pointer may be NULL. */
int lbm_get_coupling_data(lbm_ctx* ctx, double* u, double* rho);

/* Per-cell Brinkman porosity alpha in [0,1] (NULL disables).  Equilibrium velocity is
 * scaled by (1 - alpha).  No reference counterpart (design variable of the adjoint). */
int lbm_set_porosity(lbm_ctx* ctx, const double* alpha);

/* ---- state ----------------------------------------------------------------------------- */
/* f = feq(rho, u) in fluid cells, 0 elsewhere; NULL rho means 1, NULL u means 0
 * (GridManagementService::createVertex, BLB/services/GridManagementService.cpp:121-148). */
int lbm_init_equilibrium(lbm_ctx* ctx, const double* rho, const double* u);
int lbm_set_populations(lbm_ctx* ctx, const double* f);
int lbm_get_populations(lbm_ctx* ctx, double* f);

/* ---- time stepping ---------------------------------------------------------------------- */
/* nsteps fused collide+stream steps: phases A, B and C of one reference iteration
 * (BLB/mappings/RegularGrid2RegularBlockSolver.cpp:83-248) per step. */
int lbm_step(lbm_ctx* ctx, int nsteps);
/* Same steps, replayed from a hipGraph of 32 steps captured on the first call (one graph per
 * starting buffer or AA parity; the remainder runs as plain steps).  Anything the launches bake in
 * (cell order, porosity, per-cell velocities, wall classes present, body force, load schedule, MD
 * coupling) is checked on every call and a change re-captures.  Results are bit-identical to
 * lbm_step, whose launches are capture-safe (no memsets, copies or synchronisation).  On MI355X /
 * ROCm 7.2 replays are not faster than plain launches (both bound by the dispatch latency of
 * dependent kernels, ~6 us per step on small lattices).  No reference counterpart. */
int lbm_step_graph(lbm_ctx* ctx, int nsteps);
/* Same, timed with HIP events on the context stream; *ms = elapsed device time. */
int lbm_step_timed(lbm_ctx* ctx, int nsteps, float* ms);
/* Bandwidth reference of the step, no reference counterpart (measurement): nsteps launches of a
 * kernel with the step's grid, workgroup mapping and non-temporal stores that moves all 19 values
 * of every stored cell (solid ones too) from the current state buffer into the other one, with no
 * flag word and no arithmetic: mode 0 the cell's own slots (a 19-plane copy), mode 1 the slots the
 * step pulls from (the all-fluid step's access pattern).  The state is left unchanged; *ms is the
 * device time of all launches (HIP events).  Two-buffer contexts only. */
int lbm_measure_copy(lbm_ctx* ctx, int mode, int nsteps, float* ms);
/* Times the step kernel's variants for nsteps steps each on a copy of the current state (the state
 * is restored) and keeps the fastest; *variant (may be NULL) receives it.  Bit 0, the load
 * schedule: 0 = flag word first, then the 19 pulled loads (solid cells load nothing); 1 = the 19
 * neighbour loads issue alongside the flag word and bounce-back populations are re-loaded
 * afterwards (two-buffer layout only).  Bit 1, the workgroup order: 0 = the hardware's round-robin
 * dispatch over the XCDs; 1 = each XCD a contiguous range of the lattice.  Bits 2-3, the workgroup
 * shape: 0 = the default (4 rows of 64 cells, or 4 bricks), 1/2/3 = 2/4/8 rows or bricks; the
 * timing covers the three explicit shapes.  Which is
 * faster depends on lattice size, cell order and allocation.  With nsteps == 0, *variant (0..15) is
 * an input and selects directly.  When the state copy does not fit the device, the current choice
 * is kept.  All variants give bit-identical states, and the reductions run with one fixed shape.
 * No reference counterpart. */
int lbm_tune(lbm_ctx* ctx, int nsteps, int* variant);
int lbm_synchronize(lbm_ctx* ctx);

/* ---- observables ------------------------------------------------------------------------- */
/* rho[i] and u[3*i+d] per cell (BLB/BlockComputeCellValues.cpp:26-68); 0 on non-fluid cells. */
int lbm_get_fields(lbm_ctx* ctx, double* rho, double* u);
/* FP64 global mass and momentum over fluid cells: out = {sum f, sum f c_x, sum f c_y, sum f c_z}.
 * Deterministic two-pass reduction; returns LBM_ERR_NUMERIC if a non-finite value is found. */
int lbm_reduce_moments(lbm_ctx* ctx, double out[4]);
/* lbm_reduce_moments that allocates nothing.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code between HIP graph captures, tight serving loops,
 * harnesses that forbid hipMalloc).  The partial sums and results live in a fixed 16 MiB device
 * array that the library holds on each device it runs on until the process ends (its address is
 * looked up at lbm_create), instead of the context's own partials buffer; results are the same bits
 * as lbm_reduce_moments.  Lattices whose reduction needs more than 524,414 workgroups of 256 cells
 * (about 2^27 cells, e.g. beyond 512^3) are rejected with LBM_ERR_ARG.  All contexts on a device
 * share its array: calls are serialised within a process (each returns only after its results are
 * copied out), but two processes on one GPU each have their own copy. */
int lbm_reduce_moments_noalloc(lbm_ctx* ctx, double out[4]);
/* FP64 kinetic energy sum over fluid cells of |j|^2 / (2 rho) (deterministic two-pass reduction).
 * No reference counterpart (observable for the Taylor-Green validation). */
int lbm_kinetic_energy(lbm_ctx* ctx, double* ke);
/* J = cs^2 (<rho>_{x=xin} - <rho>_{x=xout}) over fluid cells of two planes, FP64. */
int lbm_pressure_drop(lbm_ctx* ctx, int xin, int xout, double* J);

/* ---- adjoint ---------------------------------------------------------------------------- */
typedef struct {
  int snapshots_used;      /* revolve checkpoints held (besides the initial state) */
  int64_t forward_steps;   /* steps of the first forward sweep (= nsteps) */
  int64_t recompute_steps; /* forward steps re-run from checkpoints during the reverse sweep */
  int64_t reverse_steps;   /* adjoint steps run (= nsteps) */
  float forward_ms;        /* device time of the first forward sweep */
  float reverse_ms;        /* device time of the reverse sweep, recomputation included */
  float recompute_ms;      /* device time of the forward steps re-run during the reverse sweep */
  float adjoint_ms;        /* device time of the adjoint (reverse) steps themselves */
} lbm_adjoint_stats;
/* Runs nsteps forward steps from the current state, evaluates J = pressure drop (FP64) and
 * returns dJ/dalpha for every cell (0 on non-fluid cells), by a hand-written reverse sweep of
 * the collide+stream kernel with revolve (binomial) checkpointing on `snapshots` device
 * checkpoints.  The state afterwards equals the state after nsteps forward steps.  On an in-place
 * (LBM_LAYOUT_AA) context it allocates a second state buffer for the call, runs in the two-buffer
 * form and leaves the final state in place again. */
int lbm_adjoint_pressure_drop(lbm_ctx* ctx, int nsteps, int snapshots, int xin, int xout,
                              double* J, double* grad_alpha, lbm_adjoint_stats* stats);

/* ---- batched uneven patches -------------------------------------------------------------- */
typedef struct lbm_patchset lbm_patchset;
typedef struct {
  int n;                   /* cubic patch edge in cells (n^3 cells, plus ghost records for the
                              populations entering across its 6 faces and 12 edges) */
  double tau;              /* per-patch relaxation time */
  int nbr[LBM_Q];          /* neighbour patch in direction c_q (q != 9); must share edge n and list
                              this patch in direction c_{18-q} */
} lbm_patch;
typedef struct {
  int tile_x, tile_y, tile_z;  /* tile extents in cells (one workgroup per tile) */
  int blocks_per_cu;           /* persistent grid = CUs * blocks_per_cu (1..64); 0 = one workgroup per tile */
  int chunked;                 /* persistent grid: 0 = workgroup w takes tiles w, w + G, w + 2G, ...;
                                  1 = workgroup w takes the contiguous range [w T / G, (w + 1) T / G) of
                                  the T tiles, so a patch's neighbouring tiles run on one CU */
} lbm_tiling;
int lbm_patches_create(lbm_patchset** ps, const lbm_patch* table, int npatches, int precision,
                       int device);
int lbm_patches_destroy(lbm_patchset* ps);
const char* lbm_patches_last_error(const lbm_patchset* ps);
int64_t lbm_patches_cells(const lbm_patchset* ps);
/* Fluid flags of the interior cells (patch order, x fastest; nonzero = fluid), the per-block
 * inner bitset of the reference's block store (BLB/services/GridManagementService.h:74-80) with its
 * boundary links discovered across patch faces and edges as ReceiveBoundaryDataService::
 * getBoundaryData does per block (SLB/services/ReceiveBoundaryDataService.cpp:107-185): every
 * fluid-to-solid link is a halfway bounce-back link.  NULL makes every cell fluid again.  The state
 * must be initialised (lbm_patches_init) after the flags change; solid cells read back as 0 and are
 * left out of the moments. */
int lbm_patches_set_flags(lbm_patchset* ps, const uint8_t* fluid);
/* initial state: f = feq(rho, u) with per-cell rho/u, interior cells of all patches in
 * patch order, x fastest (sum n^3 entries; u has 3 per cell); NULL = rest. */
int lbm_patches_init(lbm_patchset* ps, const double* rho, const double* u);
int lbm_patches_get_populations(lbm_patchset* ps, double* f);   /* 19 * sum n^3, per patch SoA */
/* rho[i] and u[3*i+d] of every interior cell, patch order, x fastest (sum n^3 cells): the block
 * store's getDensity()/getVelocity() (BLB/services/GridManagementService.cpp:295-327) computed as
 * BLB/BlockComputeCellValues.cpp:26-68 does. */
int lbm_patches_get_fields(lbm_patchset* ps, double* rho, double* u);
/* Pick a tiling by timing candidates on the device (the autotuner); *chosen may be NULL. */
int lbm_patches_autotune(lbm_patchset* ps, lbm_tiling* chosen);
int lbm_patches_set_tiling(lbm_patchset* ps, const lbm_tiling* t);
/* nsteps steps; each step is one persistent launch.  With use_graph != 0 the steps are
 * captured once into a hipGraph and replayed. */
int lbm_patches_step(lbm_patchset* ps, int nsteps, int use_graph, float* ms);
int lbm_patches_reduce_moments(lbm_patchset* ps, double out[4]);

/* ---- block store: restriction and prolongation between grid levels ----------------------
 * Replaces BLB/BlockRestrictionProlongation.{h,cpp} as the spacetree solver mapping calls it
 * (BLB/mappings/SpacetreeGrid2RegularBlockSolver.cpp:160-240, :340-366), batched: a device-resident
 * store of 6^3 vertex blocks in the reference's layout (GridManagementService::VertexData,
 * BLB/services/GridManagementService.h:74-86: pdf[19 * cell + q], density[216], velocity[648],
 * cell = x + 6 y + 36 z) and one launch per list of links.  A link names the father vertex's block
 * (coarse), the fine vertex's block (fine) and the fine vertex's child position in the refined
 * coarse cell (pos, each 0..3: getFatherVertex is pos / 2, the sub-block getSubBlock maps 0..3 to
 * 1, 2, 0, 1).  In one call a block is either fine or coarse (one level pair per call), the two
 * blocks of a link differ, fine blocks are unique (overlap search, prolongation) and coarse
 * sub-blocks are unique (restriction); violations return LBM_ERR_ARG before anything runs. */
typedef struct lbm_blockstore lbm_blockstore;
typedef struct {
  int64_t coarse;  /* block of the father vertex */
  int64_t fine;    /* block of the fine vertex */
  int pos[3];      /* child position of the fine vertex, each 0..3 */
  int new_vertex;  /* overlap search: the fine vertex is new, all populations become -1.0 and all
                      cells overlap cells (SpacetreeGrid2RegularBlockSolver.cpp:186-193) */
} lbm_amr_link;
int lbm_blocks_create(lbm_blockstore** bs, int64_t nblocks, int device);
int lbm_blocks_destroy(lbm_blockstore* bs);
const char* lbm_blocks_last_error(const lbm_blockstore* bs);
/* copy blocks first .. first + count - 1 in (set) or out (get); any pointer may be NULL */
int lbm_blocks_set(lbm_blockstore* bs, int64_t first, int64_t count, const double* pdf,
                   const double* density, const double* velocity);
int lbm_blocks_get(lbm_blockstore* bs, int64_t first, int64_t count, double* pdf, double* density,
                   double* velocity);
/* findFineOverlapCells (BlockRestrictionProlongation.cpp:32-79) for the fine block of every link;
 * the overlap cells stay with the block for lbm_blocks_prolongate and, if overlap != NULL, are
 * copied out (216 bytes per link, 1 = overlap cell). */
int lbm_blocks_find_overlap(lbm_blockstore* bs, const lbm_amr_link* links, int64_t n, uint8_t* overlap);
/* overlap cells of blocks first .. first + count - 1 given by the caller (216 flags per block,
 * nonzero = overlap cell), as the reference's prolongateToFineLevel takes them as an argument */
int lbm_blocks_set_overlap(lbm_blockstore* bs, int64_t first, int64_t count, const uint8_t* overlap);
/* prolongateToFineLevel (:83-160) on the overlap cells each fine block holds (from the last
 * lbm_blocks_find_overlap or lbm_blocks_set_overlap on it) */
int lbm_blocks_prolongate(lbm_blockstore* bs, const lbm_amr_link* links, int64_t n);
/* restrictToCoarseLevel (:164-259) */
int lbm_blocks_restrict(lbm_blockstore* bs, const lbm_amr_link* links, int64_t n);
/* device time of the last block-store operator (find/prolongate/restrict, or any of the block
 * solver operators below; lbm_blocks_iterate: the whole call) in ms (HIP events); no reference
 * counterpart (measurement) */
float lbm_blocks_last_ms(const lbm_blockstore* bs);

/* ---- block store: the block solver's per-block and per-Peano-cell operators ----------------
 * The kernels that the regular-grid solver mapping (BLB/mappings/RegularGrid2RegularBlockSolver.cpp:
 * touchVertexFirstTime :83-177, handleCell :180-220, touchVertexLastTime :223-248) and the
 * spacetree mapping (SpacetreeGrid2RegularBlockSolver.cpp:215-277, :573-577, :349) call on one block
 * or one Peano cell at a time, batched over lists: one launch works on every listed block (or cell).
 * Lists name blocks by id; a block appears at most once per list, and a block sits at one vertex
 * position at most once per cell list (the 8 blocks of a Peano cell are distinct, lexicographic:
 * vertex v = x + 2 y + 4 z), so no two workgroups touch the same values; violations return
 * LBM_ERR_ARG before anything runs.  The blocks of one Peano cell must be the geometric neighbours
 * the list says they are (as Peano's vertex enumerator gives them). */

/* the inner bitset of blocks first .. first + count - 1 (216 flags per block, nonzero = inner;
 * GridManagementService::getInner, BLB/services/GridManagementService.cpp:58-153); a new store has
 * every cell inner */
int lbm_blocks_set_inner(lbm_blockstore* bs, int64_t first, int64_t count, const uint8_t* inner);
/* BlockComputeCellValues::computeDensity and computeVelocity (BLB/BlockComputeCellValues.cpp:26-68) on
 * each listed block: inner cells (with use_overlap != 0: inner cells that are not overlap cells of
 * the block's last overlap search, the reference's hasMinusOneEntries) get rho = sum f and
 * u = (1.0/rho) sum f c; every other cell keeps its density and gets u = 0. */
int lbm_blocks_compute_cell_values(lbm_blockstore* bs, const int64_t* blocks, int64_t n, int use_overlap);
/* BlockCollisionModelManager::collide with BlockBGK(tau) (BLB/collisionmodels/CollisionModelWrapper.cpp:
 * 28-39, BlockBGK.cpp:18-45): the same cells relax toward the equilibrium of the stored density and
 * velocity, f -= 1.0/tau (f - feq); tau is the value the BGK model holds (the level's tau_eff). */
int lbm_blocks_collide(lbm_blockstore* bs, const int64_t* blocks, int64_t n, double tau, int use_overlap);
/* One boundary link: a BoundaryData record (LB/BoundaryData.h:44-61) with its boundaryDataIndex
 * (cell, incoming direction), as ReceiveBoundaryDataService::getBoundaryData lists them
 * (SLB/services/ReceiveBoundaryDataService.cpp:107-185). */
typedef struct {
  int64_t block;
  int32_t cell;   /* 0..215 */
  int32_t q_in;   /* incoming direction; the condition writes the outgoing slot 18 - q_in */
  int32_t type;   /* BoundaryData::BoundaryType: 0 halfway bounce-back (nothing to do), 1 moving wall,
                     2 pressure outlet, 3 pressure inlet, 4 velocity inlet (equilibrium), 5 moving
                     obstacle (as the moving wall) */
  int32_t pad;    /* 0 */
  double u[3];    /* velocity_L */
  double rho;     /* density_L */
} lbm_block_boundary;
/* BlockBoundaryConditionManager::treatBoundaryAfterCollision (BLB/boundaryconditions/
 * BlockBoundaryConditionManager.cpp:47-62): each link's condition (BlockMovingWall.cpp:13-33,
 * BlockAbstractPressureCondition.cpp:24-37, BlockVelocityInletEquilibrium.cpp:19-23) with the cell's
 * stored density and velocity.  A (block, cell, q_in) appears at most once.  (The reference's
 * treatBoundaryBeforeCollision is a no-op for every condition.) */
int lbm_blocks_treat_boundary(lbm_blockstore* bs, const lbm_block_boundary* links, int64_t n);
/* AABlockStreaming::streamingLocal (BLB/nonlocaltreatment/AABlockStreaming.cpp:19-32): inside each
 * listed block, f[cell][q] <-> f[cell + c_q][18 - q] for every link whose two cells are inner. */
int lbm_blocks_stream_local(lbm_blockstore* bs, const int64_t* blocks, int64_t n);
/* AABlockStreaming::streamingNonLocal (:37-53) with the index of CellTopology::getNonLocalStreamingIndex
 * (BLB/nonlocaltreatment/CellTopology.cpp:317-465): the same swap across the blocks of each listed
 * Peano cell (cells: n x 8 block ids, lexicographic vertex order), the reference's ghost exchange. */
int lbm_blocks_stream_nonlocal(lbm_blockstore* bs, const int64_t* cells, int64_t n);
/* AABlockStreaming::swapPdfsInsideCells (:80-92): f_q <-> f_{18-q} in every cell of each listed block. */
int lbm_blocks_swap_inside(lbm_blockstore* bs, const int64_t* blocks, int64_t n);
/* touchVertexFirstTime's kernels on each listed block in one launch (cell values, collision,
 * boundary links after the collision, local streaming), each block staged once through LDS; same
 * results as the four calls above in that order.  Links may name only listed blocks. */
int lbm_blocks_vertex_first(lbm_blockstore* bs, const int64_t* blocks, int64_t n, double tau, int use_overlap,
                            const lbm_block_boundary* links, int64_t nlinks);
/* nsteps iterations of the regular-grid block solver (no MD coupling, no dynamic geometry):
 * touchVertexFirstTime on every listed block, handleCell on every listed Peano cell,
 * touchVertexLastTime on every listed block.  The cross-block swaps and the in-cell swap of one
 * iteration run in the same launch as the next iteration's first touch (one launch per iteration);
 * the results equal the separate operators in the reference's order.  Every block of a listed cell
 * must be in the block list, and the cells must agree on which block neighbours which (LBM_ERR_ARG
 * otherwise). */
int lbm_blocks_iterate(lbm_blockstore* bs, const int64_t* blocks, int64_t nblocks_list, const int64_t* cells,
                       int64_t ncells, double tau, const lbm_block_boundary* links, int64_t nlinks, int nsteps);

/* ---- block store: the vertex-overlap exchange between ranks --------------------------------
 * The reference's Parallel build keeps a copy of a vertex's block on every rank whose Peano cells
 * touch it, runs the operators above on each copy, and merges the copies after the iteration
 * (BLB/RegularGridBlockVertex.cpp:51-160, called from RegularGrid2RegularBlockSolver.cpp:158-159
 * before the local streaming, :243-245 after the in-cell swap, and mergeWithNeighbour :349-364 per
 * neighbour copy).  A copy records the populations that leave the block before streaming and adds
 * those that arrive after it, so its difference is nonzero exactly where one of its own rank's
 * Peano cells swapped the link; merging adds a neighbour copy's difference where the own one is
 * 0.0.  Each block keeps its 1,008 differences (in the order of parallel::BlockBoundaryIndex4Pdfs,
 * BLB/parallel/BlockBoundaryIndex4Pdfs.cpp:30-77) and 208 boundary densities (BlockBoundaryIndex,
 * BLB/BlockBoundaryIndex.cpp:15-88) in the store, as the vertex record holds _lbPdfDiff and
 * _lbDensityOnBoundary (BLB/dastgen/BlockVertex.def:11-15); a new store has them all 0.0.  The
 * transport between GPUs is the caller's: the get and merge calls take host or device memory. */
/* the geometry-boundary bitset of blocks first .. first + count - 1 (216 flags per block, nonzero =
 * the cell is neither completely inside nor completely outside the domain;
 * GridManagementService::getBoundary, BLB/services/GridManagementService.h:74-80); a new store has
 * none set.  Only the boundary-density operators read it. */
int lbm_blocks_set_boundary(lbm_blockstore* bs, int64_t first, int64_t count, const uint8_t* boundary);
/* initPdfDiff (RegularGridBlockVertex.cpp:96-105) on each listed block: diff_i = -pdf[outgoing_i] */
int lbm_blocks_init_pdf_diff(lbm_blockstore* bs, const int64_t* blocks, int64_t n);
/* computePdfDiff (:108-120): diff_i = diff_i + pdf[incoming_i] */
int lbm_blocks_compute_pdf_diff(lbm_blockstore* bs, const int64_t* blocks, int64_t n);
/* the listed blocks' differences, 1,008 per block in list order, into out (host or device memory) */
int lbm_blocks_get_pdf_diff(lbm_blockstore* bs, const int64_t* blocks, int64_t n, double* out);
/* mergePdfDiff (:123-141) with another rank's copy of each listed block: pdf[incoming_i] =
 * pdf[incoming_i] + neighbour_i wherever the block's own diff_i is 0.0 (neighbour: 1,008 values per
 * listed block in list order, host or device memory) */
int lbm_blocks_merge_pdf_diff(lbm_blockstore* bs, const int64_t* blocks, int64_t n, const double* neighbour);
/* resetDensityOnBlockBoundary (:51-53): the listed blocks' boundary densities become 0.0 */
int lbm_blocks_reset_density_on_boundary(lbm_blockstore* bs, const int64_t* blocks, int64_t n);
/* setDensityOnBlockBoundary (:56-77): with dynamic_geometry != 0, boundary density i of each listed
 * block is the density of block-boundary cell i when that cell is not inner and on the geometry
 * boundary, else 0.0; with dynamic_geometry == 0 nothing changes */
int lbm_blocks_set_density_on_boundary(lbm_blockstore* bs, const int64_t* blocks, int64_t n, int dynamic_geometry);
/* the listed blocks' boundary densities, 208 per block in list order, into out (host or device) */
int lbm_blocks_get_density_on_boundary(lbm_blockstore* bs, const int64_t* blocks, int64_t n, double* out);
/* mergeDensityOnBoundary (:144-160): density[c_i] += neighbour_i for each block-boundary cell c_i
 * that is not inner and on the geometry boundary (neighbour: 208 values per listed block) */
int lbm_blocks_merge_density_on_boundary(lbm_blockstore* bs, const int64_t* blocks, int64_t n,
                                         const double* neighbour);

#ifdef __cplusplus
}
#endif
#endif
