// Internal declarations shared by the HIP translation units of liblbm_hip.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "lbm_cell.h"

namespace lbm {

// One lattice stored as SoA populations s[q*N + i] in one of two cell orders.
//  * Rows: i = x + nx * (y + ny * z); a workgroup is BX x g.by cells of one z-plane, a wavefront a
//    64-cell row segment.
//  * Bricks of 4 x 4 x 4 cells: the 64 cells of a brick are consecutive (x fastest, then y, then
//    z) and bricks follow in x-fastest order, i = 64 * (bx + bnx * (by + bny * bz)) + (x & 3) +
//    4 * (y & 3) + 16 * (z & 3), padded to whole bricks (padding cells are never fluid and never
//    touched); a wavefront is one brick, a workgroup g.by bricks.  A 128-byte line holds a compact
//    4 x 4 x 1 (FP64) or 4 x 4 x 2 (FP32) cell group, so far fewer lines straddle a solid surface
//    and get fetched for a few fluid cells; but a shifted read touches up to 4x more lines than a
//    row segment does.  Bricks win for FP64 on porous lattices, rows elsewhere.
// Arrays that cross the C-ABI stay in the row-major user order r; kernels that read or write them
// index both.  The stored state is the post-collision state of the last step; the reference's
// post-streaming state is its pull (see k_dense_pull).
constexpr int BX = 64;
constexpr int BY_MAX = 8;     // workgroup rows (rows order) or bricks (bricks order): 2, 4 or 8
constexpr int BY_DEFAULT = 4;
// The reductions run with this many rows per workgroup, in dispatch order, whatever the step uses,
// so that their fixed summation order (and their bits) do not depend on the tuned variant.
constexpr int BY_REDUCE = 4;
constexpr int NXCD = 8;       // MI355X: 8 XCDs, each with its own L2

struct DenseGeom {
  int nx, ny, nz;
  int px, py, pz;            // periodic flags
  int bricked;               // 0 = rows, 1 = bricks
  int remap;                 // 1: XCD-contiguous workgroup ranges (map_cell), 0: dispatch order
  int by;                    // workgroup = BX x by cells of one plane (rows) or by bricks (bricks)
  int rnx, rny;              // rows: workgroups along x and y
  int bnx, bny, bnz;         // bricks per dimension
  int tnx, tny, tnz;         // bricks: traversal tiles of 2 x 2 x 2 bricks per dimension
  int64_t N;                 // plane stride: stored cells (rows: Nu; bricks: whole bricks) + plane_pad
  int64_t Nu;                // lattice cells nx * ny * nz (the row-major user arrays)
  int64_t nblocks;           // workgroups holding a cell
  int64_t nblocks_padded;    // rounded up to a multiple of NXCD
  int layout;                // how the stored state maps to the reference state (LAYOUT_*)
  int aos;                   // populations interleaved per group of 64 stored cells (pop_at)
};

// Address of population q of stored cell i in a state buffer.  Planes (aos = 0): plane q holds
// population q of every cell, q * N + i.  Interleaved (aos = 1): the 19 populations of each group
// of 64 consecutive stored cells (one wavefront's cells: a row segment or a brick) lie together,
// 19 x 64 values per group, so the loads and stores of a wavefront stay within a few kilobytes of
// memory instead of spreading over 19 planes N cells apart.  A neighbour's address is the same
// function of the neighbour's stored index.  Both fit in Q * N values (N >= whole groups).
template <bool AOS>
__host__ __device__ __forceinline__ int64_t pop_at(int64_t N, int q, int64_t i) {
  if constexpr (AOS) return ((((i >> 6) * Q) + q) << 6) + (i & 63);
  else return (int64_t)q * N + i;
}
__host__ __device__ __forceinline__ int64_t pop_at(const DenseGeom& g, int q, int64_t i) {
  return g.aos ? pop_at<true>(g.N, q, i) : pop_at<false>(g.N, q, i);
}

// hipMalloc that leaves no error behind when it fails: HIP keeps the failure as the thread's last
// error, and the next launch check (hipGetLastError) would report it against an unrelated kernel.
// The caller handles the returned error (e.g. lbm_tune falls back when a state copy does not fit).
template <typename T>
inline hipError_t device_alloc(T** p, size_t bytes) {
  const hipError_t e = hipMalloc((void**)p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
  }
  return e;
}

// stored index of cell (x, y, z)
__host__ __device__ inline int64_t stored_index(const DenseGeom& g, int x, int y, int z) {
  if (!g.bricked) return (int64_t)x + (int64_t)g.nx * ((int64_t)y + (int64_t)g.ny * z);
  return ((((int64_t)(z >> 2) * g.bny + (y >> 2)) * g.bnx + (x >> 2)) << 6) + ((z & 3) << 4) + ((y & 3) << 2) +
         (x & 3);
}

// Stored-state layouts.  PULL: two buffers, post-collision values in the cell's own slots, so the
// reference population q of cell x is s_q(x - c_q).  AA (one buffer, in place): NATURAL holds the
// reference state itself, a_q(x) = f_q(x); SWAPPED (after an even AA step) holds post-collision
// values reflected in the cell's own slots, a_{18-q}(x) = f*_q(x).
enum : int { LAYOUT_PULL = 0, LAYOUT_AA_NATURAL = 1, LAYOUT_AA_SWAPPED = 2 };

// Population planes are N apart, N = stored cells + plane_pad(esize): with a power-of-two stride
// (2^24 cells at 256^3) the 19 plane streams of a wavefront alias on the same HBM channels and the
// step time varies by 15 % from one allocation to the next.  Pads measured on the 256^3 porous
// lattice (bricks): FP64 +6 % with 2112 cells (16.5 KiB), FP32 stable and +1 % with 64 cells
// (256 B); pads of 1M cells were 20 % slower.
int64_t plane_pad(size_t esize);
DenseGeom make_geom(int nx, int ny, int nz, const int periodic[3], bool bricked, size_t esize, bool aos);
// Workgroup shape: `by` rows of BX cells (rows order) or `by` bricks (bricks order), by in {2, 4, 8};
// it fixes the grid (rny, nblocks) but not the stored order.  The untuned default, BY_DEFAULT = 4,
// was the best shape or within 1 % of it in a sweep of all tuned variants on the porous 256^3
// channel and the all-fluid 512^3 lattice in both precisions (8 rows led FP32 on the porous
// channel by 1 %).
void set_workgroup_rows(DenseGeom& g, int by);
inline bool valid_workgroup_rows(int by) { return by == 2 || by == 4 || by == 8; }
// per-cell array of `comps` values of `esize` bytes from the cell order of `from` to that of `to`
hipError_t launch_relayout(const DenseGeom& from, const DenseGeom& to, const void* in, void* out, int comps,
                           int esize, hipStream_t stream);
// stored cell ids of `from` -> those of `to`, in place
hipError_t launch_convert_ids(const DenseGeom& from, const DenseGeom& to, int64_t* ids, int64_t n, hipStream_t stream);

// ---- launchers (all enqueue on `stream`; return hipError_t) ----
hipError_t launch_build_words(const DenseGeom& g, const uint8_t* inner, const uint8_t* cls, uint32_t* words,
                              hipStream_t stream);
// one step src -> dst
// spec: issue the 19 neighbour loads alongside the flag word and re-load the bounce-back
// populations afterwards, instead of reading the word first (same results either way)
// boundary classes of a context: the device table (MAX_CLASSES entries) and the optional
// per-cell target velocities [3*N] of BC_CELL_VELOCITY classes
struct BoundaryArgs {
  const BoundaryClass* table;
  const double* cell_u;
};
// fp32: FP32-shifted storage; f32math: FP32 collision arithmetic (else the hybrid arithmetic of collide_f32s_h)
hipError_t launch_dense_step(const DenseGeom& g, bool fp32, bool f32math, const void* src, void* dst,
                             const uint32_t* words,
                             const void* alpha, bool force, bool walls, const CollideParams& p64,
                             const CollideParams32& p32, const BoundaryArgs& bc, bool spec, hipStream_t stream);
// reference state f (FP64) from the stored state
// one in-place AA step on a single buffer: parity 0 reads the NATURAL layout and leaves SWAPPED,
// parity 1 the reverse
hipError_t launch_aa_step(const DenseGeom& g, bool fp32, bool f32math, int parity, void* a, const uint32_t* words,
                          const void* alpha, bool force, bool walls, const CollideParams& p64,
                          const CollideParams32& p32, const BoundaryArgs& bc, hipStream_t stream);
hipError_t launch_dense_pull(const DenseGeom& g, bool fp32, const void* src, double* f, const uint32_t* words,
                             hipStream_t stream);
// stored state from a reference state f (FP64)
hipError_t launch_dense_unpull(const DenseGeom& g, bool fp32, const double* f, void* dst, const uint32_t* words,
                               hipStream_t stream);
// bandwidth reference of the step: the 19-plane copy (pull = false) or the pull pattern (pull = true)
// of every stored cell from src to dst, no arithmetic
hipError_t launch_dense_copy(const DenseGeom& g, bool fp32, bool pull, const void* src, void* dst, hipStream_t stream);
// the stored state of `from` re-stored in the layout of `to` (same cell order and plane stride) into
// another buffer, values copied exactly, no staging memory
hipError_t launch_state_relayout(const DenseGeom& from, const DenseGeom& to, bool fp32, const void* src, void* dst,
                                 const uint32_t* words, hipStream_t stream);
// stored state from feq(rho, u); rho/u may be null (1 / 0)
hipError_t launch_dense_init_eq(const DenseGeom& g, bool fp32, const double* rho, const double* u, void* dst,
                                const uint32_t* words, hipStream_t stream);
// per-cell rho/u (FP64) of the reference state
hipError_t launch_dense_fields(const DenseGeom& g, bool fp32, const void* src, const uint32_t* words, double* rho,
                               double* u, hipStream_t stream);
// deterministic two-pass reduction: out4 = {sum(rho - base), jx, jy, jz}, base = 1 for FP32-shifted
hipError_t launch_dense_moments(const DenseGeom& g, bool fp32, const void* src, const uint32_t* words,
                                double* partials, double* out4, hipStream_t stream);
int64_t dense_moment_partials(const DenseGeom& g);
// kinetic energy sum over fluid cells of |j|^2 / (2 rho)
hipError_t launch_dense_energy(const DenseGeom& g, bool fp32, const void* src, const uint32_t* words,
                               double* partials, double* out1, hipStream_t stream);
// plane density sums for the pressure drop: out4 = {sum rho(xin), n(xin), sum rho(xout), n(xout)}
hipError_t launch_dense_plane_rho(const DenseGeom& g, bool fp32, const void* src, const uint32_t* words, int xin,
                                  int xout, double* partials, double* out4, hipStream_t stream);
// count fluid cells (exact: sums of 1.0 below 2^53); needs dense_moment_partials(g) partial slots
hipError_t launch_count_fluid(const DenseGeom& g, const uint32_t* words, double* partials, double* out1,
                              hipStream_t stream);

}  // namespace lbm
