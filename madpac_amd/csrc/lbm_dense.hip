// Dense-lattice kernels of the D3Q19 block-LB solver for CDNA4 (gfx950).
//
// Hot path: k_dense_step, one fused collide+stream per cell in pull form.  A thread owns one
// cell; it gathers the 19 incoming populations from the SoA state (population q of cell
// x - c_q, or the reflected population 18-q of its own cell when x - c_q is not fluid), then
// collides, applies the moving-wall term and writes its 19 post-collision populations back to
// its own slots of the other buffer.  A wavefront owns one 4 x 4 x 4 brick, stored as 64
// consecutive values per population plane (512 B in FP64, 256 B in FP32): writes are whole
// aligned lines, and a shifted read touches the brick's own lines plus the facing layer of one
// to three neighbour bricks, which their own wavefronts read at about the same time from L2.
// Nothing is staged in LDS: each value is read once and written once.
//
// The reference's own update is the three-phase swap algorithm
// (BLB/mappings/RegularGrid2RegularBlockSolver.cpp:83-248 with
// BLB/nonlocaltreatment/AABlockStreaming.cpp:19-92); storing the post-collision state and
// pulling gives the same numbers, because a moving-wall correction is applied to the outgoing
// population of a wall link, which only that cell's own bounce-back ever reads.
#include "lbm_geom.h"

namespace lbm {

int64_t plane_pad(size_t esize) { return esize == 8 ? 2112 : 64; }

DenseGeom make_geom(int nx, int ny, int nz, const int periodic[3], bool bricked, size_t esize, bool aos) {
  DenseGeom g;
  g.nx = nx; g.ny = ny; g.nz = nz;
  g.px = periodic[0] ? 1 : 0; g.py = periodic[1] ? 1 : 0; g.pz = periodic[2] ? 1 : 0;
  g.bricked = bricked ? 1 : 0;
  g.rnx = (nx + BX - 1) / BX;
  g.bnx = (nx + 3) / 4; g.bny = (ny + 3) / 4; g.bnz = (nz + 3) / 4;
  g.tnx = (g.bnx + 1) / 2; g.tny = (g.bny + 1) / 2; g.tnz = (g.bnz + 1) / 2;
  const int64_t bricks = (int64_t)g.bnx * g.bny * g.bnz;
  g.Nu = (int64_t)nx * ny * nz;
  // untuned defaults (lbm_tune times both): measured better for bricks, and for rows from 2^29 cells
  // up (1024^3: FP32 38.2k against 33.7k MLUPS, FP64 in place 19.2k against 17.8k; 256^3 and 512^3
  // rows run faster in dispatch order)
  g.remap = (bricked || g.Nu >= ((int64_t)1 << 29)) ? 1 : 0;
  g.N = (bricked ? bricks * 64 : g.Nu) + plane_pad(esize);
  set_workgroup_rows(g, BY_DEFAULT);
  g.layout = LAYOUT_PULL;
  // interleaving needs whole 64-cell groups per brick or per row segment (addresses separable per axis)
  g.aos = (aos && (bricked || nx % BX == 0)) ? 1 : 0;
  return g;
}

void set_workgroup_rows(DenseGeom& g, int by) {
  g.by = by;
  g.rny = (g.ny + by - 1) / by;
  // bricks: 8 brick slots per traversal tile, by per workgroup (by divides 8)
  g.nblocks = g.bricked ? (int64_t)g.tnx * g.tny * g.tnz * 8 / by : (int64_t)g.rnx * g.rny * g.nz;
  g.nblocks_padded = (g.nblocks + NXCD - 1) / NXCD * NXCD;
}

// the geometry the reductions launch with: BY_REDUCE rows or bricks per workgroup in dispatch
// order, so that each partial sums the same cells into the same slot whatever the step's tuning
static DenseGeom reduce_geom(const DenseGeom& g) {
  DenseGeom r = g;
  set_workgroup_rows(r, BY_REDUCE);
  r.remap = 0;
  return r;
}

// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(BX * BY_MAX) void k_build_words(DenseGeom g, const uint8_t* __restrict__ inner,
                                                         const uint8_t* __restrict__ cls,
                                                         uint32_t* __restrict__ words) {
  const CellPos c = map_cell(g);
  if (!c.valid) return;
  if (!inner[c.r]) {
    words[c.i] = 0u;
    return;
  }
  // class ids index a MAX_CLASSES-entry table; the host rejects larger ids, the mask keeps it in range
  uint32_t w = FLUID | ((cls ? (uint32_t)cls[c.r] & CLASS_BITS : 0u) << CLASS_SHIFT);
#pragma unroll
  for (int q = 0; q < Q; q++) {
    if (q == REST) continue;
    int n[3] = {c.x + CX(q), c.y + CY(q), c.z + CZ(q)};
    const int e[3] = {g.nx, g.ny, g.nz};
    const int per[3] = {g.px, g.py, g.pz};
    bool outside = false;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      if (n[k] < 0 || n[k] >= e[k]) {
        if (per[k]) n[k] = (n[k] + e[k]) % e[k];
        else outside = true;
      }
    }
    if (outside || !inner[(int64_t)n[0] + (int64_t)g.nx * ((int64_t)n[1] + (int64_t)g.ny * n[2])]) w |= 1u << q;
  }
  words[c.i] = w;
}

// ------------------------------------------------------------------------------------------
// Collision of one fluid cell plus the boundary classes of its bounce-back links (shared by the
// two-buffer and the AA steps).  T = double: reference arithmetic; T = float: shifted populations,
// hybrid arithmetic (collide_f32s_h) unless M32 (all FP32).
template <typename T, bool FORCE, bool ALPHA, bool WALLS, bool M32>
__device__ __forceinline__ void collide_cell_bc(T (&f)[Q], uint32_t w, int64_t i, const T* __restrict__ alpha,
                                                const CollideParams& p64, const CollideParams32& p32,
                                                const BoundaryArgs& bc) {
  const T a = ALPHA ? alpha[i] : T(0);
  T rho;
  T uc[sizeof(T) == 8 ? 3 : 4];
  if constexpr (sizeof(T) == 8) {
    rho = collide_f64<FORCE, ALPHA>(f, p64, a, uc);
  } else if constexpr (M32) {
    rho = collide_f32s<FORCE, ALPHA>(f, p32, a, uc);
  } else {
    rho = collide_f32s_h<FORCE, ALPHA>(f, p64, p32, a, uc);
  }
  if (WALLS) {
    const uint32_t k = (w >> CLASS_SHIFT) & CLASS_BITS;
    if (k) {
      const BoundaryClass& b = bc.table[k];
      const uint32_t m = w & b.mask;
      if (m) {
        if constexpr (sizeof(T) == 8) apply_boundary_f64(f, m, b, bc.cell_u, i, rho, uc[0], uc[1], uc[2]);
        else apply_boundary_f32(f, m, b, bc.cell_u, i, rho, uc[3], uc[0], uc[1], uc[2]);
      }
    }
  }
}

// ------------------------------------------------------------------------------------------
// Whether this lane's 128-byte line of a population plane holds a fluid lane of the wavefront
// (`fluid` = the wavefront's ballot of fluid lanes).  A solid lane in such a line stores zeros into
// its own slots (never read: pulls and readers skip non-fluid cells) together with the fluid lanes'
// stores, so the line is written whole: partial-line writes at the solid surfaces cost more.
template <typename T>
__device__ __forceinline__ bool line_has_fluid(unsigned long long fluid) {
  constexpr int LANES = 128 / (int)sizeof(T);
  const int first = (int)threadIdx.x & ~(LANES - 1);
  const unsigned long long line = (LANES >= 64) ? ~0ull : (((1ull << LANES) - 1ull) << first);
  return (fluid & line) != 0;
}

// The hot kernel.  T = double: reference arithmetic; T = float: shifted populations.
// SPEC = false: the flag word is read first, then the 19 loads at the pulled addresses (two
// dependent memory round trips; solid lanes load nothing).  SPEC = true: the 19 neighbour loads
// issue together with the word, and the few bounce-back populations are re-loaded from the cell's
// own reflected slots once the word has arrived (one round trip on most lanes; solid lanes load).
template <typename T, bool FORCE, bool ALPHA, bool WALLS, bool SPEC, bool M32, bool AOS>
__global__ __launch_bounds__(BX * BY_MAX) void k_dense_step(DenseGeom g, const T* __restrict__ src,
                                                        T* __restrict__ dst, const uint32_t* __restrict__ words,
                                                        const T* __restrict__ alpha, CollideParams p64,
                                                        CollideParams32 p32, BoundaryArgs bc) {
  const CellPos c = map_cell(g);
  if (!c.valid) return;
  const CellAddr<AOS> A = cell_addr<AOS>(g, c);
  T f[Q];
  uint32_t w;
  if (SPEC) {
#pragma unroll
    for (int q = 0; q < Q; q++) f[q] = src[A.pull(q)];
    w = words[c.i];
    const bool fl = (w & FLUID) != 0;
    const unsigned long long fluid = __ballot(fl);  // every lane votes before any returns
    if (!fl && !line_has_fluid<T>(fluid)) return;
    if (fl) {
      if (w & 0x7FFFFu) {
#pragma unroll
        for (int q = 0; q < Q; q++)
          if ((w >> (Q - 1 - q)) & 1u) f[q] = src[A.own(Q - 1 - q)];
      }
      collide_cell_bc<T, FORCE, ALPHA, WALLS, M32>(f, w, c.i, alpha, p64, p32, bc);
    } else {
#pragma unroll
      for (int q = 0; q < Q; q++) f[q] = T(0);
    }
  } else {
    w = words[c.i];
    const bool fl = (w & FLUID) != 0;
    const unsigned long long fluid = __ballot(fl);  // every lane votes before any returns
    if (!fl && !line_has_fluid<T>(fluid)) return;
    if (fl) {
#pragma unroll
      for (int q = 0; q < Q; q++) f[q] = src[((w >> (Q - 1 - q)) & 1u) ? A.own(Q - 1 - q) : A.pull(q)];
      collide_cell_bc<T, FORCE, ALPHA, WALLS, M32>(f, w, c.i, alpha, p64, p32, bc);
    } else {
#pragma unroll
      for (int q = 0; q < Q; q++) f[q] = T(0);
    }
  }
#pragma unroll
  for (int q = 0; q < Q; q++) __builtin_nontemporal_store(f[q], &dst[A.own(q)]);
}

// ------------------------------------------------------------------------------------------
// AA in-place steps on one buffer a (half the memory of two buffers, the same traffic).
// Even (ODD = false): a holds the reference state NATURAL, a_q(x) = f_q(x); the cell reads its own
// 19 slots and writes f*_q(x) into its own slot 18 - q (SWAPPED): no neighbour access at all.
// Odd: the cell reads f_q(x) = a_{18-q}(x - c_q), or a_q(x) across a bounce-back link, and writes
// f*_q(x) to a_q(x + c_q), or to a_{18-q}(x) across a bounce-back link (NATURAL again).  In both, a
// cell writes exactly the slots it read, and no other cell touches them, so the update is in place
// without races.  Boundary classes act on the outgoing population of a bounce-back link, as in the
// two-buffer step.
template <typename T, bool FORCE, bool ALPHA, bool WALLS, bool M32, bool ODD, bool AOS>
__global__ __launch_bounds__(BX * BY_MAX) void k_aa_step(DenseGeom g, T* __restrict__ a, const uint32_t* __restrict__ words,
                                                     const T* __restrict__ alpha, CollideParams p64,
                                                     CollideParams32 p32, BoundaryArgs bc) {
  const CellPos c = map_cell(g);
  if (!c.valid) return;
  const uint32_t w = words[c.i];
  const bool fl = (w & FLUID) != 0;
  const unsigned long long fluid = __ballot(fl);  // every lane votes before any returns
  // Solid lanes complete the lines of fluid neighbours with zeros, as in k_dense_step: a solid
  // cell's slots are dead in both steps (no fluid cell reads them, and an odd-step push toward a
  // solid cell goes to the pusher's own reflected slot instead), so writing them races with nothing.
  if (!fl && !line_has_fluid<T>(fluid)) return;
  CellAddr<AOS> A = cell_addr<AOS>(g, c);
  T f[Q];
  if (fl) {
    // odd: f_q(x) = a_{18-q}(x - c_q), or a_q(x) across a bounce-back link
#pragma unroll
    for (int q = 0; q < Q; q++)
      f[q] = ODD ? a[((w >> (Q - 1 - q)) & 1u) ? A.own(q) : A.own(Q - 1 - q) + src_off(q, A.o)] : a[A.own(q)];
    collide_cell_bc<T, FORCE, ALPHA, WALLS, M32>(f, w, c.i, alpha, p64, p32, bc);
  } else {
#pragma unroll
    for (int q = 0; q < Q; q++) f[q] = T(0);
  }
  // The stores go to the slots just loaded.  Left to itself the compiler keeps the 19 load
  // addresses live through the collision (38 more VGPRs, one wave less per SIMD); an opaque copy
  // of the cell's base address makes it recompute them here instead.
  asm volatile("" : "+v"(A.base));
#pragma unroll
  for (int q = 0; q < Q; q++) {
    // x + c_q is not fluid (a solid lane stores its zeros to its own slots)
    const bool bounce = ((fl ? w : 0x7FFFFu) >> q) & 1u;
    const int64_t di = (!ODD || bounce) ? A.own(Q - 1 - q) : A.push(q);
    // An even step writes whole lines of its own slots.  An odd step's pushes fill each line from
    // up to four wavefronts: ordinary stores let L2 merge the parts before write-back (measured
    // 5-13 % faster on a porous 256^3 channel than streaming stores, which write each part out).
    if (ODD)
      a[di] = f[q];
    else
      __builtin_nontemporal_store(f[q], &a[di]);
  }
}

// ------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(BX * BY_MAX) void k_dense_pull(DenseGeom g, const T* __restrict__ src,
                                                        double* __restrict__ f, const uint32_t* __restrict__ words) {
  const CellPos c = map_cell(g);
  if (!c.valid) return;
  const uint32_t w = words[c.i];
  const int64_t Nu = g.Nu;
  if (!(w & FLUID)) {
#pragma unroll
    for (int q = 0; q < Q; q++) f[(int64_t)q * Nu + c.r] = 0.0;
    return;
  }
  const NbrOff o = nbr_offsets(g, c);
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const T v = src[state_index(g, q, w, c.i, o)];
    f[(int64_t)q * Nu + c.r] = (sizeof(T) == 8) ? (double)v : (double)v + W64(q);
  }
}

// Scatter of a reference state into the stored representation (inverse of the pull).
template <typename T>
__global__ __launch_bounds__(BX * BY_MAX) void k_dense_unpull(DenseGeom g, const double* __restrict__ f,
                                                          T* __restrict__ dst, const uint32_t* __restrict__ words) {
  const CellPos c = map_cell(g);
  if (!c.valid) return;
  const uint32_t w = words[c.i];
  if (!(w & FLUID)) return;
  const NbrOff o = nbr_offsets(g, c);
  const int64_t Nu = g.Nu;
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const double v = f[(int64_t)q * Nu + c.r];
    dst[state_index(g, q, w, c.i, o)] = (sizeof(T) == 8) ? (T)v : (T)(v - W64(q));
  }
}

// Bandwidth reference of the step (measurement, not part of a time step): every stored cell of the
// lattice (fluid or not) moves its 19 values from src to dst with the step's grid, workgroup mapping
// and non-temporal stores, and no flag word or arithmetic.  PULL = false: each cell copies its own
// slots (the 19-plane copy); PULL = true: it reads the slots the step pulls from (x - c_q, wrapped on
// periodic axes, clamped otherwise), i.e. the all-fluid step's memory pattern.
template <typename T, bool PULL, bool AOS>
__global__ __launch_bounds__(BX * BY_MAX) void k_dense_copy(DenseGeom g, const T* __restrict__ src, T* __restrict__ dst) {
  const CellPos c = map_cell(g);
  if (!c.valid) return;
  const CellAddr<AOS> A = cell_addr<AOS>(g, c);
  T f[Q];
#pragma unroll
  for (int q = 0; q < Q; q++) f[q] = src[PULL ? A.pull(q) : A.own(q)];
#pragma unroll
  for (int q = 0; q < Q; q++) __builtin_nontemporal_store(f[q], &dst[A.own(q)]);
}

// The same state in another layout of the same cell order (e.g. pull form <-> AA NATURAL): each
// fluid cell copies its 19 values of the reference state from where `from` keeps them to where `to`
// does.  src and dst are distinct buffers; the stored values are copied as they are (exact in both
// precisions) and non-fluid cells are left alone.
template <typename T>
__global__ __launch_bounds__(BX * BY_MAX) void k_state_relayout(DenseGeom from, DenseGeom to, const T* __restrict__ src,
                                                            T* __restrict__ dst, const uint32_t* __restrict__ words) {
  const CellPos c = map_cell(to);
  if (!c.valid) return;
  const uint32_t w = words[c.i];
  if (!(w & FLUID)) return;
  const NbrOff o = nbr_offsets(to, c);
#pragma unroll
  for (int q = 0; q < Q; q++) dst[state_index(to, q, w, c.i, o)] = src[state_index(from, q, w, c.i, o)];
}

// feq(rho, u) exactly as BLB/BlockComputeCellValues.cpp:9-22 evaluates it, scattered like unpull.
template <typename T>
__global__ __launch_bounds__(BX * BY_MAX) void k_dense_init_eq(DenseGeom g, const double* __restrict__ rho,
                                                           const double* __restrict__ u, T* __restrict__ dst,
                                                           const uint32_t* __restrict__ words) {
  const CellPos c = map_cell(g);
  if (!c.valid) return;
  const uint32_t w = words[c.i];
  if (!(w & FLUID)) return;
  const NbrOff o = nbr_offsets(g, c);
  const double r = rho ? rho[c.r] : 1.0;
  const double u0 = u ? u[3 * c.r] : 0.0, u1 = u ? u[3 * c.r + 1] : 0.0, u2 = u ? u[3 * c.r + 2] : 0.0;
  double uu = u0 * u0;
  uu += u1 * u1;
  uu += u2 * u2;
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const double cu = cdot64(q, u0, u1, u2);
    const double feq = W64(q) * r *
        (1.0 + cu / (1.0 / 3.0) + (cu * cu) / (2.0 * (1.0 / 3.0) * (1.0 / 3.0)) - (uu / (2.0 * (1.0 / 3.0))));
    dst[state_index(g, q, w, c.i, o)] = (sizeof(T) == 8) ? (T)feq : (T)(feq - W64(q));
  }
}

// rho and u of the reference state, in the reference's operation order for FP64.
template <typename T>
__global__ __launch_bounds__(BX * BY_MAX) void k_dense_fields(DenseGeom g, const T* __restrict__ src,
                                                          const uint32_t* __restrict__ words,
                                                          double* __restrict__ rho, double* __restrict__ u) {
  const CellPos c = map_cell(g);
  if (!c.valid) return;
  const uint32_t w = words[c.i];
  if (!(w & FLUID)) {
    rho[c.r] = 0.0;
    u[3 * c.r] = u[3 * c.r + 1] = u[3 * c.r + 2] = 0.0;
    return;
  }
  const NbrOff o = nbr_offsets(g, c);
  double r = 0.0, b0 = 0.0, b1 = 0.0, b2 = 0.0;
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const T v = src[state_index(g, q, w, c.i, o)];
    const double fv = (sizeof(T) == 8) ? (double)v : (double)v + W64(q);
    r += fv;
    if (CX(q) == 1) b0 += fv;
    if (CX(q) == -1) b0 -= fv;
    if (CY(q) == 1) b1 += fv;
    if (CY(q) == -1) b1 -= fv;
    if (CZ(q) == 1) b2 += fv;
    if (CZ(q) == -1) b2 -= fv;
  }
  const double inv = 1.0 / r;
  rho[c.r] = r;
  u[3 * c.r] = b0 * inv;
  u[3 * c.r + 1] = b1 * inv;
  u[3 * c.r + 2] = b2 * inv;
}

// ------------------------------------------------------------------------------------------
// Deterministic FP64 reductions: each block reduces its cells with 64-lane wavefront shuffles
// then an LDS step over its wavefronts, writes one partial, and a single block sums the
// partials in a fixed order.
template <int K>
__device__ __forceinline__ void block_reduce_store(double (&v)[K], double* __restrict__ out) {
  __shared__ double lds[K][BX * BY_MAX / 64];
#pragma unroll
  for (int k = 0; k < K; k++) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
  }
  const int tid = threadIdx.x + threadIdx.y * BX;
  const int wave = tid / 64, lane = tid % 64;
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) lds[k][wave] = v[k];
  }
  __syncthreads();
  if (tid == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) {
      double s = 0.0;
      for (int j = 0; j < (int)blockDim.y; j++) s += lds[k][j];  // one wavefront per workgroup row
      out[k] = s;
    }
  }
}

template <typename T>
__global__ __launch_bounds__(BX * BY_MAX) void k_dense_moments(DenseGeom g, const T* __restrict__ src,
                                                           const uint32_t* __restrict__ words,
                                                           double* __restrict__ partials) {
  const CellPos c = map_cell(g);
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  if (c.valid) {
    const uint32_t w = words[c.i];
    if (w & FLUID) {
      const NbrOff o = nbr_offsets(g, c);
#pragma unroll
      for (int q = 0; q < Q; q++) {
        const T s = src[state_index(g, q, w, c.i, o)];
        // deviation from the rest state: f - w (FP64) or the shifted value itself (FP32)
        const double fv = (sizeof(T) == 8) ? (double)s : (double)s + W64(q);
        const double dv = (sizeof(T) == 8) ? (double)s - W64(q) : (double)s;
        v[0] += dv;
        if (CX(q) == 1) v[1] += fv;
        if (CX(q) == -1) v[1] -= fv;
        if (CY(q) == 1) v[2] += fv;
        if (CY(q) == -1) v[2] -= fv;
        if (CZ(q) == 1) v[3] += fv;
        if (CZ(q) == -1) v[3] -= fv;
      }
    }
  }
  block_reduce_store<4>(v, partials + 4 * (int64_t)blockIdx.x);
}

// kinetic energy sum over fluid cells of (j . j) / (2 rho), FP64
template <typename T>
__global__ __launch_bounds__(BX * BY_MAX) void k_dense_energy(DenseGeom g, const T* __restrict__ src,
                                                          const uint32_t* __restrict__ words,
                                                          double* __restrict__ partials) {
  const CellPos c = map_cell(g);
  double v[1] = {0.0};
  if (c.valid) {
    const uint32_t w = words[c.i];
    if (w & FLUID) {
      const NbrOff o = nbr_offsets(g, c);
      double r = 0.0, j0 = 0.0, j1 = 0.0, j2 = 0.0;
#pragma unroll
      for (int q = 0; q < Q; q++) {
        const T s = src[state_index(g, q, w, c.i, o)];
        const double fv = (sizeof(T) == 8) ? (double)s : (double)s + W64(q);
        r += fv;
        if (CX(q) == 1) j0 += fv;
        if (CX(q) == -1) j0 -= fv;
        if (CY(q) == 1) j1 += fv;
        if (CY(q) == -1) j1 -= fv;
        if (CZ(q) == 1) j2 += fv;
        if (CZ(q) == -1) j2 -= fv;
      }
      v[0] = (j0 * j0 + j1 * j1 + j2 * j2) / (2.0 * r);
    }
  }
  block_reduce_store<1>(v, partials + (int64_t)blockIdx.x);
}

template <int K>
__global__ __launch_bounds__(256) void k_sum_partials(const double* __restrict__ partials, int64_t n,
                                                      double* __restrict__ out) {
  __shared__ double lds[K][256];
  double s[K];
#pragma unroll
  for (int k = 0; k < K; k++) s[k] = 0.0;
  for (int64_t j = threadIdx.x; j < n; j += 256) {
#pragma unroll
    for (int k = 0; k < K; k++) s[k] += partials[K * j + k];
  }
#pragma unroll
  for (int k = 0; k < K; k++) lds[k][threadIdx.x] = s[k];
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
#pragma unroll
      for (int k = 0; k < K; k++) lds[k][threadIdx.x] += lds[k][threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) out[k] = lds[k][0];
  }
}

// density sums over the fluid cells of two x-planes; one thread per (y, z) row of each plane
template <typename T>
__global__ __launch_bounds__(256) void k_plane_rho(DenseGeom g, const T* __restrict__ src,
                                                   const uint32_t* __restrict__ words, int xin, int xout,
                                                   double* __restrict__ partials) {
  const int64_t rows = (int64_t)g.ny * g.nz;
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  if (t < rows) {
    const int y = (int)(t % g.ny), z = (int)(t / g.ny);
#pragma unroll
    for (int k = 0; k < 2; k++) {
      CellPos c;
      c.x = k == 0 ? xin : xout;
      c.y = y;
      c.z = z;
      c.i = stored_index(g, c.x, y, z);
      const uint32_t w = words[c.i];
      if (w & FLUID) {
        const NbrOff o = nbr_offsets(g, c);
        // density deviation rho - 1 = sum (f_q - w_q) (the FP32 state stores f_q - w_q): the rest
        // density cancels in the pressure drop, and the small deviations keep its relative precision
        double r = 0.0;
#pragma unroll
        for (int q = 0; q < Q; q++) {
          const T s = src[state_index(g, q, w, c.i, o)];
          r += (sizeof(T) == 8) ? (double)s - W64(q) : (double)s;
        }
        v[2 * k] = r;
        v[2 * k + 1] = 1.0;
      }
    }
  }
  // 256 threads = 4 wavefronts; reuse the 2-D block reducer layout
  __shared__ double lds[4][4];
#pragma unroll
  for (int k = 0; k < 4; k++)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
  if (threadIdx.x % 64 == 0)
#pragma unroll
    for (int k = 0; k < 4; k++) lds[k][threadIdx.x / 64] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < 4; k++) partials[4 * (int64_t)blockIdx.x + k] = lds[k][0] + lds[k][1] + lds[k][2] + lds[k][3];
}

__global__ __launch_bounds__(BX * BY_MAX) void k_count_fluid(DenseGeom g, const uint32_t* __restrict__ words,
                                                         double* __restrict__ partials) {
  const CellPos c = map_cell(g);
  double v[1] = {(c.valid && (words[c.i] & FLUID)) ? 1.0 : 0.0};
  block_reduce_store<1>(v, partials + (int64_t)blockIdx.x);
}

// Cell-order conversion of per-cell arrays (one thread per cell of the target order).
template <typename E>
__global__ __launch_bounds__(BX * BY_MAX) void k_relayout(DenseGeom from, DenseGeom to, const E* __restrict__ in,
                                                      E* __restrict__ out, int comps) {
  const CellPos c = map_cell(to);
  if (!c.valid) return;
  const int64_t j = stored_index(from, c.x, c.y, c.z);
  for (int k = 0; k < comps; k++) out[comps * c.i + k] = in[comps * j + k];
}

__global__ __launch_bounds__(256) void k_convert_ids(DenseGeom from, DenseGeom to, int64_t* __restrict__ ids,
                                                     int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  const CellPos c = pos_of(from, ids[k]);
  ids[k] = stored_index(to, c.x, c.y, c.z);
}

// ------------------------------------------------------------------------------------------
static dim3 dense_block(const DenseGeom& g) { return dim3(BX, g.by, 1); }
static dim3 dense_grid(const DenseGeom& g) { return dim3((unsigned)g.nblocks_padded, 1, 1); }

hipError_t launch_build_words(const DenseGeom& g, const uint8_t* inner, const uint8_t* cls, uint32_t* words,
                              hipStream_t s) {
  hipLaunchKernelGGL(k_build_words, dense_grid(g), dense_block(g), 0, s, g, inner, cls, words);
  return hipGetLastError();
}

template <typename T, bool SPEC, bool M32, bool AOS>
static hipError_t step_t(const DenseGeom& g, const void* src, void* dst, const uint32_t* words, const void* alpha,
                         bool force, bool walls, const CollideParams& p64, const CollideParams32& p32,
                         const BoundaryArgs& bc, hipStream_t s) {
  const T* a = (const T*)alpha;
  const T* sp = (const T*)src;
  T* dp = (T*)dst;
#define LBM_STEP(F, A, W)                                                                                        \
  hipLaunchKernelGGL((k_dense_step<T, F, A, W, SPEC, M32, AOS>), dense_grid(g), dense_block(g), 0, s, g, sp, dp, words, a, \
                     p64, p32, bc)
  const int key = (force ? 4 : 0) | (alpha ? 2 : 0) | (walls ? 1 : 0);
  switch (key) {
    case 0: LBM_STEP(false, false, false); break;
    case 1: LBM_STEP(false, false, true); break;
    case 2: LBM_STEP(false, true, false); break;
    case 3: LBM_STEP(false, true, true); break;
    case 4: LBM_STEP(true, false, false); break;
    case 5: LBM_STEP(true, false, true); break;
    case 6: LBM_STEP(true, true, false); break;
    default: LBM_STEP(true, true, true); break;
  }
#undef LBM_STEP
  return hipGetLastError();
}

template <bool AOS>
static hipError_t step_aos(const DenseGeom& g, bool fp32, bool f32math, const void* src, void* dst,
                           const uint32_t* words, const void* alpha, bool force, bool walls, const CollideParams& p64,
                           const CollideParams32& p32, const BoundaryArgs& bc, bool spec, hipStream_t s) {
  if (fp32 && f32math)
    return spec ? step_t<float, true, true, AOS>(g, src, dst, words, alpha, force, walls, p64, p32, bc, s)
                : step_t<float, false, true, AOS>(g, src, dst, words, alpha, force, walls, p64, p32, bc, s);
  if (fp32)
    return spec ? step_t<float, true, false, AOS>(g, src, dst, words, alpha, force, walls, p64, p32, bc, s)
                : step_t<float, false, false, AOS>(g, src, dst, words, alpha, force, walls, p64, p32, bc, s);
  return spec ? step_t<double, true, false, AOS>(g, src, dst, words, alpha, force, walls, p64, p32, bc, s)
              : step_t<double, false, false, AOS>(g, src, dst, words, alpha, force, walls, p64, p32, bc, s);
}

hipError_t launch_dense_step(const DenseGeom& g, bool fp32, bool f32math, const void* src, void* dst,
                             const uint32_t* words, const void* alpha, bool force, bool walls,
                             const CollideParams& p64, const CollideParams32& p32, const BoundaryArgs& bc, bool spec,
                             hipStream_t s) {
  return g.aos ? step_aos<true>(g, fp32, f32math, src, dst, words, alpha, force, walls, p64, p32, bc, spec, s)
               : step_aos<false>(g, fp32, f32math, src, dst, words, alpha, force, walls, p64, p32, bc, spec, s);
}

template <typename T, bool M32, bool ODD, bool AOS>
static hipError_t aa_t(const DenseGeom& g, void* buf, const uint32_t* words, const void* alpha, bool force, bool walls,
                       const CollideParams& p64, const CollideParams32& p32, const BoundaryArgs& bc, hipStream_t s) {
  T* ap = (T*)buf;
  const T* al = (const T*)alpha;
#define LBM_AA(F, A, W)                                                                                        \
  hipLaunchKernelGGL((k_aa_step<T, F, A, W, M32, ODD, AOS>), dense_grid(g), dense_block(g), 0, s, g, ap, words, al, \
                     p64, p32, bc)
  const int key = (force ? 4 : 0) | (alpha ? 2 : 0) | (walls ? 1 : 0);
  switch (key) {
    case 0: LBM_AA(false, false, false); break;
    case 1: LBM_AA(false, false, true); break;
    case 2: LBM_AA(false, true, false); break;
    case 3: LBM_AA(false, true, true); break;
    case 4: LBM_AA(true, false, false); break;
    case 5: LBM_AA(true, false, true); break;
    case 6: LBM_AA(true, true, false); break;
    default: LBM_AA(true, true, true); break;
  }
#undef LBM_AA
  return hipGetLastError();
}

template <bool AOS>
static hipError_t aa_aos(const DenseGeom& g, bool fp32, bool f32math, bool odd, void* a, const uint32_t* words,
                         const void* alpha, bool force, bool walls, const CollideParams& p64, const CollideParams32& p32,
                         const BoundaryArgs& bc, hipStream_t s) {
  if (fp32 && f32math)
    return odd ? aa_t<float, true, true, AOS>(g, a, words, alpha, force, walls, p64, p32, bc, s)
               : aa_t<float, true, false, AOS>(g, a, words, alpha, force, walls, p64, p32, bc, s);
  if (fp32)
    return odd ? aa_t<float, false, true, AOS>(g, a, words, alpha, force, walls, p64, p32, bc, s)
               : aa_t<float, false, false, AOS>(g, a, words, alpha, force, walls, p64, p32, bc, s);
  return odd ? aa_t<double, false, true, AOS>(g, a, words, alpha, force, walls, p64, p32, bc, s)
             : aa_t<double, false, false, AOS>(g, a, words, alpha, force, walls, p64, p32, bc, s);
}

hipError_t launch_aa_step(const DenseGeom& g, bool fp32, bool f32math, int parity, void* a, const uint32_t* words,
                          const void* alpha, bool force, bool walls, const CollideParams& p64,
                          const CollideParams32& p32, const BoundaryArgs& bc, hipStream_t s) {
  const bool odd = parity != 0;
  return g.aos ? aa_aos<true>(g, fp32, f32math, odd, a, words, alpha, force, walls, p64, p32, bc, s)
               : aa_aos<false>(g, fp32, f32math, odd, a, words, alpha, force, walls, p64, p32, bc, s);
}

hipError_t launch_dense_pull(const DenseGeom& g, bool fp32, const void* src, double* f, const uint32_t* words,
                             hipStream_t s) {
  if (fp32) hipLaunchKernelGGL(k_dense_pull<float>, dense_grid(g), dense_block(g), 0, s, g, (const float*)src, f, words);
  else hipLaunchKernelGGL(k_dense_pull<double>, dense_grid(g), dense_block(g), 0, s, g, (const double*)src, f, words);
  return hipGetLastError();
}

hipError_t launch_dense_unpull(const DenseGeom& g, bool fp32, const double* f, void* dst, const uint32_t* words,
                               hipStream_t s) {
  if (fp32) hipLaunchKernelGGL(k_dense_unpull<float>, dense_grid(g), dense_block(g), 0, s, g, f, (float*)dst, words);
  else hipLaunchKernelGGL(k_dense_unpull<double>, dense_grid(g), dense_block(g), 0, s, g, f, (double*)dst, words);
  return hipGetLastError();
}

template <typename T, bool AOS>
static void copy_t(const DenseGeom& g, bool pull, const void* src, void* dst, hipStream_t s) {
  if (pull)
    hipLaunchKernelGGL((k_dense_copy<T, true, AOS>), dense_grid(g), dense_block(g), 0, s, g, (const T*)src, (T*)dst);
  else
    hipLaunchKernelGGL((k_dense_copy<T, false, AOS>), dense_grid(g), dense_block(g), 0, s, g, (const T*)src, (T*)dst);
}

hipError_t launch_dense_copy(const DenseGeom& g, bool fp32, bool pull, const void* src, void* dst, hipStream_t s) {
  if (fp32) {
    if (g.aos) copy_t<float, true>(g, pull, src, dst, s);
    else copy_t<float, false>(g, pull, src, dst, s);
  } else {
    if (g.aos) copy_t<double, true>(g, pull, src, dst, s);
    else copy_t<double, false>(g, pull, src, dst, s);
  }
  return hipGetLastError();
}

hipError_t launch_state_relayout(const DenseGeom& from, const DenseGeom& to, bool fp32, const void* src, void* dst,
                                 const uint32_t* words, hipStream_t s) {
  if (from.bricked != to.bricked || from.N != to.N) return hipErrorInvalidValue;
  if (fp32)
    hipLaunchKernelGGL(k_state_relayout<float>, dense_grid(to), dense_block(to), 0, s, from, to, (const float*)src,
                       (float*)dst, words);
  else
    hipLaunchKernelGGL(k_state_relayout<double>, dense_grid(to), dense_block(to), 0, s, from, to, (const double*)src,
                       (double*)dst, words);
  return hipGetLastError();
}

hipError_t launch_dense_init_eq(const DenseGeom& g, bool fp32, const double* rho, const double* u, void* dst,
                                const uint32_t* words, hipStream_t s) {
  if (fp32)
    hipLaunchKernelGGL(k_dense_init_eq<float>, dense_grid(g), dense_block(g), 0, s, g, rho, u, (float*)dst, words);
  else
    hipLaunchKernelGGL(k_dense_init_eq<double>, dense_grid(g), dense_block(g), 0, s, g, rho, u, (double*)dst, words);
  return hipGetLastError();
}

hipError_t launch_dense_fields(const DenseGeom& g, bool fp32, const void* src, const uint32_t* words, double* rho,
                               double* u, hipStream_t s) {
  if (fp32)
    hipLaunchKernelGGL(k_dense_fields<float>, dense_grid(g), dense_block(g), 0, s, g, (const float*)src, words, rho, u);
  else
    hipLaunchKernelGGL(k_dense_fields<double>, dense_grid(g), dense_block(g), 0, s, g, (const double*)src, words, rho,
                       u);
  return hipGetLastError();
}

int64_t dense_moment_partials(const DenseGeom& g) { return reduce_geom(g).nblocks_padded; }

hipError_t launch_relayout(const DenseGeom& from, const DenseGeom& to, const void* in, void* out, int comps,
                           int esize, hipStream_t s) {
  if (esize == 8)
    hipLaunchKernelGGL(k_relayout<double>, dense_grid(to), dense_block(to), 0, s, from, to, (const double*)in,
                       (double*)out, comps);
  else if (esize == 4)
    hipLaunchKernelGGL(k_relayout<float>, dense_grid(to), dense_block(to), 0, s, from, to, (const float*)in,
                       (float*)out, comps);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_convert_ids(const DenseGeom& from, const DenseGeom& to, int64_t* ids, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_convert_ids, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, from, to, ids, n);
  return hipGetLastError();
}

hipError_t launch_dense_moments(const DenseGeom& gs, bool fp32, const void* src, const uint32_t* words,
                                double* partials, double* out4, hipStream_t s) {
  const DenseGeom g = reduce_geom(gs);
  if (fp32)
    hipLaunchKernelGGL(k_dense_moments<float>, dense_grid(g), dense_block(g), 0, s, g, (const float*)src, words,
                       partials);
  else
    hipLaunchKernelGGL(k_dense_moments<double>, dense_grid(g), dense_block(g), 0, s, g, (const double*)src, words,
                       partials);
  hipLaunchKernelGGL(k_sum_partials<4>, dim3(1), dim3(256), 0, s, (const double*)partials, g.nblocks_padded, out4);
  return hipGetLastError();
}

hipError_t launch_dense_energy(const DenseGeom& gs, bool fp32, const void* src, const uint32_t* words,
                               double* partials, double* out1, hipStream_t s) {
  const DenseGeom g = reduce_geom(gs);
  if (fp32)
    hipLaunchKernelGGL(k_dense_energy<float>, dense_grid(g), dense_block(g), 0, s, g, (const float*)src, words,
                       partials);
  else
    hipLaunchKernelGGL(k_dense_energy<double>, dense_grid(g), dense_block(g), 0, s, g, (const double*)src, words,
                       partials);
  hipLaunchKernelGGL(k_sum_partials<1>, dim3(1), dim3(256), 0, s, (const double*)partials, g.nblocks_padded, out1);
  return hipGetLastError();
}

hipError_t launch_dense_plane_rho(const DenseGeom& g, bool fp32, const void* src, const uint32_t* words, int xin,
                                  int xout, double* partials, double* out4, hipStream_t s) {
  const int64_t rows = (int64_t)g.ny * g.nz;
  const unsigned nb = (unsigned)((rows + 255) / 256);
  if (fp32)
    hipLaunchKernelGGL(k_plane_rho<float>, dim3(nb), dim3(256), 0, s, g, (const float*)src, words, xin, xout, partials);
  else
    hipLaunchKernelGGL(k_plane_rho<double>, dim3(nb), dim3(256), 0, s, g, (const double*)src, words, xin, xout,
                       partials);
  hipLaunchKernelGGL(k_sum_partials<4>, dim3(1), dim3(256), 0, s, (const double*)partials, (int64_t)nb, out4);
  return hipGetLastError();
}

hipError_t launch_count_fluid(const DenseGeom& gs, const uint32_t* words, double* partials, double* out1,
                              hipStream_t s) {
  const DenseGeom g = reduce_geom(gs);
  hipLaunchKernelGGL(k_count_fluid, dense_grid(g), dense_block(g), 0, s, g, words, partials);
  hipLaunchKernelGGL(k_sum_partials<1>, dim3(1), dim3(256), 0, s, (const double*)partials, g.nblocks_padded, out1);
  return hipGetLastError();
}

}  // namespace lbm
