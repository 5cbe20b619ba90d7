"""The vertex-overlap exchange between ranks of the reference's Parallel build on the oracle
(BLB/RegularGridBlockVertex.cpp:51-160, BLB/parallel/BlockBoundaryIndex4Pdfs.cpp,
BLB/BlockBoundaryIndex.cpp; the calls in BLB/mappings/RegularGrid2RegularBlockSolver.cpp:158-159,
:243-245, :349-364).

The reference compiles these functions only with MPI, so the compiled reference here (built without
it) pins the two index tables (tests/golden/parallel_index.npz, oracle/ref/parallel.cpp) and the
rest is pinned by what the exchange is for: split the Peano cells among ranks, let each rank run the
block solver on its own copy of the blocks its cells touch, exchange the differences, and every copy
of every block must equal the single-domain run, bit for bit.  Mistakes in the merge rule break that
(checked below with the mistakes planted in Python, on the CPU).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from block_grid import BS, C, NC, Q, BlockGrid, iterate, iterate_partitioned, partition_cells
from conftest import golden
from test_oracle_blocks import OracleOps, initial_state


class OracleExchangeOps(OracleOps):
    """The oracle's block operators plus the per-block pdf differences of one rank."""

    def __init__(self, grid, merge=None):
        super().__init__(grid)
        self.diff = {}
        self.merge = merge

    def init_pdf_diff(self, b, st):
        self.diff[int(b)] = O.block_init_pdf_diff(st[0][b])

    def compute_pdf_diff(self, b, st):
        self.diff[int(b)] = O.block_compute_pdf_diff(st[0][b], self.diff[int(b)])

    def pdf_diff(self, b):
        return self.diff[int(b)].copy()

    def merge_pdf_diff(self, b, st, neighbour):
        if self.merge is not None:
            st[0][b] = self.merge(st[0][b], self.diff[int(b)], neighbour)
        else:
            st[0][b] = O.block_merge_pdf_diff(st[0][b], self.diff[int(b)], neighbour)


def _cell_xyz(cell):
    cell = np.asarray(cell)
    return np.stack([cell % BS, (cell // BS) % BS, cell // (BS * BS)], axis=-1)


def test_index_tables_match_the_compiled_reference():
    d = golden("parallel_index.npz")
    inc, out = O.block_pdf_boundary()
    assert np.array_equal(inc, d["incoming"]) and np.array_equal(out, d["outgoing"])
    assert np.array_equal(O.block_boundary_cells(), d["boundary_cells"])


def test_index_tables_closed_forms():
    inc, out = O.block_pdf_boundary()
    # LB_PDFS_ON_BLOCKBOUNDARY for D3Q19 (BLB/BlockDefinitions.h): faces, edges, corners
    assert len(out) == 6 * (BS - 2) ** 2 * 5 + 12 * (BS - 2) * 9 + 8 * 12 == 1008
    cell, q = out // Q, out % Q
    assert np.array_equal(inc, cell * Q + Q - 1 - q)
    # exactly the populations whose target leaves the block, in (cell, q) order
    allc = np.repeat(np.arange(NC), Q)
    allq = np.tile(np.arange(Q), NC)
    t = _cell_xyz(allc) + C[allq]
    leaves = ~((t >= 0) & (t < BS)).all(axis=1)
    assert np.array_equal(out, (allc * Q + allq)[leaves])
    cells = O.block_boundary_cells()
    assert len(cells) == NC - (BS - 4) ** 3 == 208
    assert np.all(np.diff(cells) > 0)
    xyz = _cell_xyz(np.setdiff1d(np.arange(NC), cells))
    assert ((xyz >= 2) & (xyz <= BS - 3)).all()


def test_differences_and_merge_by_hand():
    rng = np.random.default_rng(4)
    inc, out = O.block_pdf_boundary()
    before = 0.05 + 0.01 * rng.random(NC * Q)
    d0 = O.block_init_pdf_diff(before)
    assert np.array_equal(d0, -before[out])
    after = 0.05 + 0.01 * rng.random(NC * Q)
    d1 = O.block_compute_pdf_diff(after, d0)
    assert np.array_equal(d1, d0 + after[inc])
    own = d1.copy()
    own[::3] = 0.0
    nb = rng.standard_normal(len(out))
    merged = O.block_merge_pdf_diff(after, own, nb)
    expect = after.copy()
    z = own == 0.0
    expect[inc[z]] = after[inc[z]] + nb[z]
    assert np.array_equal(merged, expect)


def perturbed_state(grid, seed=3):
    """Equilibrium at rest with 1 % seeded noise on every population, so that every link carries a
    nonzero difference from the first step on (at rest the difference across a link is exactly 0.0)."""
    pdf, rho, u = initial_state(grid)
    pdf *= 1.0 + 0.01 * np.random.default_rng(seed).standard_normal(pdf.shape)
    return pdf, rho, u


def _run(n, ranks, steps, merge=None, overlap=False, tau=0.6, ulid=0.05, seed=None):
    grid = BlockGrid(n, n, n)
    links = grid.cavity_links(ulid)
    mono = perturbed_state(grid)
    iterate(grid, OracleOps(grid), mono, O.tau_eff(tau), links, steps)
    cells_of = partition_cells(grid, ranks, overlap, seed)
    states = [perturbed_state(grid) for _ in range(ranks)]
    ops = [OracleExchangeOps(grid, merge) for _ in range(ranks)]
    blocks_of = iterate_partitioned(grid, ops, states, cells_of, O.tau_eff(tau), links, steps)
    return grid, mono, states, blocks_of


@pytest.mark.parametrize("n,ranks,steps,overlap,seed", [(12, 2, 4, False, None), (12, 8, 3, False, None),
                                                       (13, 2, 3, False, None), (18, 2, 3, True, None),
                                                       (12, 3, 3, False, 7), (13, 5, 2, False, 8)])
def test_partitioned_iterations_equal_the_single_domain(n, ranks, steps, overlap, seed):
    """With overlap the middle layer of Peano cells is traversed by both ranks: their copies then
    both carry those swaps, and only the merge rule's guard (add where the own difference is 0.0)
    keeps them from being applied twice.  With a seed the cells go to random ranks: ragged
    boundaries, blocks shared by up to 8 ranks."""
    grid, mono, states, blocks_of = _run(n, ranks, steps, overlap=overlap, seed=seed)
    shared = set()
    for r in range(ranks):
        b = blocks_of[r]
        assert np.array_equal(states[r][0][b], mono[0][b])
        assert np.array_equal(states[r][1][b], mono[1][b]) and np.array_equal(states[r][2][b], mono[2][b])
        for o in range(r + 1, ranks):
            shared |= set(np.intersect1d(b, blocks_of[o]).tolist())
    assert shared and set().union(*[set(b.tolist()) for b in blocks_of]) == set(range(grid.nblocks))
    # the state moved
    assert not np.array_equal(mono[0], perturbed_state(grid)[0])


def _always_add(pdf, own, nb):
    inc, _ = O.block_pdf_boundary()
    out = np.array(pdf)
    out[inc] = out[inc] + nb
    return out


def _never_merge(pdf, own, nb):
    return np.array(pdf)


def _subtract(pdf, own, nb):
    inc, _ = O.block_pdf_boundary()
    out = np.array(pdf)
    z = own == 0.0
    out[inc[z]] = out[inc[z]] - nb[z]
    return out


@pytest.mark.parametrize("mistake,overlap", [(_always_add, True), (_never_merge, False), (_subtract, False)])
def test_planted_merge_mistakes_break_the_partition(mistake, overlap):
    """Dropping the guard only shows where two ranks traverse the same Peano cell (with disjoint
    cells the neighbour's difference is 0.0 wherever the own one is not)."""
    grid, mono, states, blocks_of = _run(18 if overlap else 12, 2, 2, merge=mistake, overlap=overlap)
    assert any(not np.array_equal(states[r][0][blocks_of[r]], mono[0][blocks_of[r]]) for r in range(2))


def test_boundary_density_closed_forms():
    rng = np.random.default_rng(9)
    cells = O.block_boundary_cells()
    for _ in range(3):
        rho = 1.0 + 0.01 * rng.random(NC)
        inner = rng.random(NC) < 0.6
        boundary = rng.random(NC) < 0.5
        on = ~inner[cells] & boundary[cells]
        prev = rng.random(len(cells))
        # without a dynamic geometry the buffer keeps its values (setDensityOnBlockBoundary returns)
        assert np.array_equal(O.block_set_density_on_boundary(rho, inner, boundary, False, prev), prev)
        got = O.block_set_density_on_boundary(rho, inner, boundary, True, prev)
        assert np.array_equal(got, np.where(on, rho[cells], 0.0))
        nb = rng.standard_normal(len(cells))
        merged = O.block_merge_density_on_boundary(rho, inner, boundary, nb)
        expect = rho.copy()
        expect[cells[on]] += nb[on]
        assert np.array_equal(merged, expect)
