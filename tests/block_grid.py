"""A regular grid of 6^3 blocks laid out as the reference's regular-grid block solver keeps a box
of cells (test helper).

The box's nx x ny x nz cells sit in blocks_per_dim(n) = n // 6 + 1 + (1 if n % 6 else 0) blocks per
axis (BLB/runners/BlockLatticeBoltzmannBatchJobRunnerForRegularGrid.cpp:78-92); slot s of block b
holds cell 6 b + s - 3, so the blocks overhang the box by at least three cells on every side and a
cell is inner when it lies in the box (GridManagementService::createVertex,
BLB/services/GridManagementService.cpp:58-153).  Peano cells lie between neighbouring block
vertices, their 8 blocks in lexicographic order (Peano's VertexEnumerator).  Block ids are
lexicographic (x fastest).

``iterate`` composes one time step of RegularGrid2RegularBlockSolver
(BLB/mappings/RegularGrid2RegularBlockSolver.cpp:83-248) from per-block and per-cell operators:
touchVertexFirstTime on every block (cell values, collision, boundary links, local streaming),
handleCell on every Peano cell (cross-block streaming), touchVertexLastTime on every block (in-cell
swap).  The operators come from an object with the oracle's or the device's block-operator methods.
"""
from __future__ import annotations

import itertools

import numpy as np

BS, NC, Q = 6, 216, 19
C = np.array([[0, -1, -1], [-1, 0, -1], [0, 0, -1], [1, 0, -1], [0, 1, -1], [-1, -1, 0], [0, -1, 0],
              [1, -1, 0], [-1, 0, 0], [0, 0, 0], [1, 0, 0], [-1, 1, 0], [0, 1, 0], [1, 1, 0],
              [0, -1, 1], [-1, 0, 1], [0, 0, 1], [1, 0, 1], [0, 1, 1]], dtype=np.int64)

# BoundaryData::BoundaryType (LB/BoundaryData.h:44-61)
HW_BB, MOVING_WALL, PRESSURE_OUTLET, PRESSURE_INLET, VELOCITY_INLET, MOVING_OBSTACLE = range(6)


def blocks_per_dim(n: int) -> int:
    return n // BS + 1 + (1 if n % BS else 0)


class BlockGrid:
    def __init__(self, nx, ny, nz):
        self.dims = (nx, ny, nz)
        self.nb = tuple(blocks_per_dim(n) for n in self.dims)
        self.nblocks = int(np.prod(self.nb))
        bx, by, bz = self.nb
        # lattice coordinates of every (block, slot)
        b = np.arange(self.nblocks)
        s = np.arange(NC)
        self.block_xyz = np.stack([b % bx, (b // bx) % by, b // (bx * by)], axis=1)
        slot = np.stack([s % BS, (s // BS) % BS, s // (BS * BS)], axis=1)
        self.cell_xyz = BS * self.block_xyz[:, None, :] + slot[None, :, :] - 3   # (nblocks, 216, 3)
        lo = (self.cell_xyz >= 0).all(axis=2)
        hi = (self.cell_xyz < np.array(self.dims)).all(axis=2)
        self.inner = (lo & hi).astype(np.uint8)
        x, y, z = (self.cell_xyz[..., d] for d in range(3))
        self.box_index = np.where(self.inner != 0, x + nx * (y + ny * z), -1)
        cells = []
        for cz, cy, cx in itertools.product(range(bz - 1), range(by - 1), range(bx - 1)):
            cells.append([(cx + (v & 1)) + bx * ((cy + ((v >> 1) & 1)) + by * (cz + (v >> 2))) for v in range(8)])
        self.peano_cells = np.array(cells, dtype=np.int64).reshape(-1, 8)

    # ----- state transfer between the blocks and an SoA box field f[q, z, y, x] -----
    def scatter(self, f_box, fill):
        """Blocks holding f_box in the inner cells and fill[q] (e.g. the weights) elsewhere."""
        nx, ny, nz = self.dims
        f = np.asarray(f_box, dtype=np.float64).reshape(Q, nx * ny * nz)
        pdf = np.broadcast_to(np.asarray(fill, dtype=np.float64), (self.nblocks, NC, Q)).copy()
        m = self.inner != 0
        pdf[m] = f[:, self.box_index[m]].T
        return pdf.reshape(self.nblocks, NC * Q)

    def gather(self, pdf):
        nx, ny, nz = self.dims
        p = np.asarray(pdf).reshape(self.nblocks, NC, Q)
        out = np.zeros((Q, nx * ny * nz))
        m = self.inner != 0
        out[:, self.box_index[m]] = p[m].T
        return out.reshape(-1)

    # ----- boundary links (cell, incoming direction, type, u_L, rho_L) per block -----
    def links(self, classify):
        """classify(x, y, z, q_out) -> None (no treatment) or (type, u_L, rho_L) for every inner
        cell whose neighbour along the outgoing direction q_out lies outside the box
        (ReceiveBoundaryDataService::getBoundaryData, SLB/services/ReceiveBoundaryDataService.cpp:107-185).
        Returns a list of (block, cell, q_in, type, u_L, rho_L)."""
        out = []
        dims = np.array(self.dims)
        blk, cell = np.nonzero(self.inner)
        xyz = self.cell_xyz[blk, cell]
        for q in range(Q):
            t = xyz + C[q]
            leaves = ~((t >= 0) & (t < dims)).all(axis=1)
            for b, c, (x, y, z) in zip(blk[leaves], cell[leaves], xyz[leaves]):
                r = classify(int(x), int(y), int(z), q)
                if r is not None:
                    out.append((int(b), int(c), Q - 1 - q, r[0], r[1], r[2]))
        return out

    def cavity_links(self, ulid):
        """The lid of the cavity: every leaving link of the top layer with c_z > 0 is a moving wall
        with u_w = (ulid, 0, 0) (SLB/BoundaryData4LidDrivenCavity.cpp:17-55); the rest is halfway
        bounce-back, which needs no treatment."""
        nz = self.dims[2]
        return self.links(lambda x, y, z, q: (MOVING_WALL, (ulid, 0.0, 0.0), 1.0)
                          if z == nz - 1 and C[q, 2] > 0 else None)

    def channel_links(self, inlet_u):
        """The default channel: leaving through a y or z face is a wall, else through x- a velocity
        inlet (rho_L = 1, u_L = inlet_u(x, y, z)), through x+ a pressure outlet (rho_L = 1)."""
        nx, ny, nz = self.dims

        def cls(x, y, z, q):
            t = (x + C[q, 0], y + C[q, 1], z + C[q, 2])
            if not (0 <= t[1] < ny and 0 <= t[2] < nz):
                return None
            if t[0] < 0:
                return (VELOCITY_INLET, tuple(inlet_u(x, y, z)), 1.0)
            return (PRESSURE_OUTLET, (0.0, 0.0, 0.0), 1.0)
        return self.links(cls)


def iterate(grid, ops, state, tau, links, steps=1):
    """`steps` time steps of the regular-grid block solver on state = (pdf, density, velocity)
    arrays of shape (nblocks, 4104), (nblocks, 216), (nblocks, 648), updated in place through
    ops.cell_values(b, state), ops.collide(b, state, tau), ops.boundary(b, state, links_of_b),
    ops.stream_local(b, state), ops.stream_nonlocal(cell_blocks, state), ops.swap_inside(b, state)."""
    by_block = {}
    for lk in links:
        by_block.setdefault(lk[0], []).append(lk[1:])
    for _ in range(steps):
        for b in range(grid.nblocks):
            ops.cell_values(b, state)
            ops.collide(b, state, tau)
            if b in by_block:
                ops.boundary(b, state, by_block[b])
            ops.stream_local(b, state)
        for cell in grid.peano_cells:
            ops.stream_nonlocal(cell, state)
        for b in range(grid.nblocks):
            ops.swap_inside(b, state)


def partition_cells(grid, ranks, overlap=False, seed=None):
    """The Peano cells of each rank: 2 ranks split them into a lower and an upper half in z (with
    overlap, the middle layer of cells belongs to both); 8 ranks colour them by the parity of their
    coordinates (a vertex shared by up to 8 ranks); with a seed, each cell goes to a random rank (a
    ragged partition of any rank count, as a space-filling-curve split can leave).  Returns one
    array of cell rows per rank."""
    bx, by, bz = grid.nb
    k = np.arange(len(grid.peano_cells))
    cx, cy, cz = k % (bx - 1), (k // (bx - 1)) % (by - 1), k // ((bx - 1) * (by - 1))
    if seed is not None and not overlap:
        owner = np.random.default_rng(seed).integers(ranks, size=len(k))
        return [grid.peano_cells[owner == r] for r in range(ranks)]
    if ranks == 1:
        return [grid.peano_cells]
    if ranks == 2:
        mid = (bz - 1) // 2
        lo = cz <= mid if overlap else 2 * cz < bz - 1
        hi = cz >= mid if overlap else ~lo
        return [grid.peano_cells[lo], grid.peano_cells[hi]]
    if ranks == 8 and not overlap:
        owner = (cx & 1) + 2 * (cy & 1) + 4 * (cz & 1)
        return [grid.peano_cells[owner == r] for r in range(8)]
    raise ValueError("ranks must be 1, 2 or 8 without a seed (overlap with 2 only)")


def iterate_partitioned(grid, ops, states, cells_of, tau, links, steps=1):
    """The reference's Parallel iteration (RegularGrid2RegularBlockSolver.cpp:83-248 with the
    vertex exchange of :158-159, :243-245 and mergeWithNeighbour :349-364): rank r traverses the
    Peano cells cells_of[r] and keeps its own copy states[r] of every block those cells touch.  Per
    step each rank runs touchVertexFirstTime on its blocks with ops[r].init_pdf_diff(b, st) before
    the local streaming, handleCell on its cells, touchVertexLastTime with
    ops[r].compute_pdf_diff(b, st) after the in-cell swap; then every rank merges the difference of
    every other rank's copy into each block they share (ops[r].merge_pdf_diff(b, st, neighbour)).
    Returns the blocks of each rank."""
    ranks = len(states)
    blocks_of = [np.unique(np.asarray(c).reshape(-1)) for c in cells_of]
    by_block = {}
    for lk in links:
        by_block.setdefault(lk[0], []).append(lk[1:])
    for _ in range(steps):
        for r in range(ranks):
            st, op = states[r], ops[r]
            for b in blocks_of[r]:
                op.cell_values(b, st)
                op.collide(b, st, tau)
                if b in by_block:
                    op.boundary(b, st, by_block[b])
                op.init_pdf_diff(b, st)
                op.stream_local(b, st)
            for cell in cells_of[r]:
                op.stream_nonlocal(cell, st)
            for b in blocks_of[r]:
                op.swap_inside(b, st)
                op.compute_pdf_diff(b, st)
        diffs = [{int(b): ops[r].pdf_diff(b) for b in blocks_of[r]} for r in range(ranks)]
        for r in range(ranks):
            for o in range(ranks):
                if o == r:
                    continue
                for b in np.intersect1d(blocks_of[r], blocks_of[o]):
                    ops[r].merge_pdf_diff(b, states[r], diffs[o][int(b)])
    return blocks_of
