"""The oracle's restatement of the block solver's operators (cell values, BGK collision, boundary
links, local / cross-block / in-cell swap streaming on 6^3 blocks) against the compiled reference:
the streaming fixtures (index table and permutations written by the reference's own
AABlockStreaming and CellTopology), the cell-kernel fixtures, and whole time steps of the
regular-grid block solver composed from the operators against the reference's cavity and channel
trajectories.  All bit-exact."""
import ctypes

import numpy as np
import pytest

import oracle as O
from block_grid import BS, NC, Q, BlockGrid, iterate
from conftest import golden

BP = NC * Q
W = np.array([1 / 36, 1 / 36, 1 / 18, 1 / 36, 1 / 36, 1 / 36, 1 / 18, 1 / 36, 1 / 18, 1 / 3,
              1 / 18, 1 / 36, 1 / 18, 1 / 36, 1 / 36, 1 / 36, 1 / 18, 1 / 36, 1 / 36])


def _values(nblocks):
    # the harness's distinct values: 1 + k + 10000 v for entry k of block v
    return 1.0 + np.arange(BP)[None, :] + 10000.0 * np.arange(nblocks)[:, None]


def _sources(out):
    k = out - 1.0
    blk = np.floor(k / 10000.0)
    return (blk * BP + (k - 10000.0 * blk)).astype(np.int32)


@pytest.mark.parametrize("seed", [1, 2])
def test_nonlocal_index_is_the_reference_table(seed):
    d = golden(f"streaming_s{seed}.npz")
    assert np.array_equal(O.block_nonlocal_index(), d["nonlocal_index"])


def test_index_tables_closed_form():
    loc = O.block_local_index()
    assert loc.shape == (1440, 4)
    # every swap pair is a link (cell, q) <-> (cell + c_q, 18 - q) inside the block, each slot at most once
    slots = loc[:, :2].reshape(-1)
    assert np.unique(slots).size == slots.size
    nl = O.block_nonlocal_index()
    assert nl.shape == (504, 5)
    # 12 face-adjacent vertex pairs with 39 links each (9 cells x 5 directions - 6 leaving the cell) and 12
    # edge pairs with 3; no corner pair exchanges anything in D3Q19
    pairs, counts = np.unique(nl[:, 0] * 8 + nl[:, 1], return_counts=True)
    hamming = np.array([bin((p // 8) ^ (p % 8)).count("1") for p in pairs])
    assert sorted(counts[hamming == 1]) == [39] * 12 and sorted(counts[hamming == 2]) == [3] * 12
    assert not (hamming == 3).any()
    # each entry links neighbouring cells of the global lattice across a block boundary
    vxyz = np.array([[v & 1, (v >> 1) & 1, v >> 2] for v in range(8)])
    for i, j, c1, c2, q in nl:
        p1 = 6 * vxyz[i] + np.array([c1 % BS, (c1 // BS) % BS, c1 // (BS * BS)])
        p2 = 6 * vxyz[j] + np.array([c2 % BS, (c2 // BS) % BS, c2 // (BS * BS)])
        assert np.array_equal(p2 - p1, O.C[q])
    assert np.unique(nl[:, [0, 2, 4]], axis=0).shape[0] == 504


@pytest.mark.parametrize("seed", [1, 2])
def test_swap_streaming_permutations_are_the_references(seed):
    d = golden(f"streaming_s{seed}.npz")
    v = _values(1)[0]
    assert np.array_equal(_sources(O.block_stream_local(np.ones(NC, np.uint8), v)), d["local_all"])
    assert np.array_equal(_sources(O.block_stream_local(d["inner_local"], v)), d["local_masked"])
    assert np.array_equal(_sources(O.block_swap_inside(v)), d["swapped"])
    v8 = _values(8)
    assert np.array_equal(_sources(O.block_stream_nonlocal(np.ones((8, NC), np.uint8), v8)), d["nonlocal_all"])
    assert np.array_equal(_sources(O.block_stream_nonlocal(d["inner_nonlocal"], v8)), d["nonlocal_masked"])


@pytest.mark.parametrize("seed", [1, 2])
def test_block_operators_reproduce_the_cell_kernel_fixtures(seed):
    d = golden(f"cellkernels_s{seed}.npz")
    inner = d["inner"]
    rho0 = np.full(NC, 7.0)  # skipped cells keep their density
    rho, u = O.block_cell_values(inner, d["pdf_in"], density=rho0)
    m = inner != 0
    assert np.array_equal(rho[m], d["density"][m]) and (rho[~m] == 7.0).all()
    assert np.array_equal(u, d["velocity"])
    col = O.block_collide(inner, d["pdf_in"], rho, u, float(d["tau_eff"]))
    assert np.array_equal(col, d["pdf_collided"])
    links = [(c, qi, 1, d["uw"], 1.0) for c, qi in d["wall_links"]]
    wall = O.block_boundary(col, rho, u, links)
    assert np.array_equal(wall, d["pdf_after_wall"])
    out = O.block_swap_inside(O.block_stream_local(inner, wall))
    assert np.array_equal(out, d["pdf_after_local_stream"])


def test_cells_with_minus_one_entries_are_skipped():
    d = golden("cellkernels_s1.npz")
    inner = d["inner"]
    mo = np.zeros(NC, np.uint8)
    mo[::5] = 1
    rho0 = np.full(NC, 3.0)
    rho, u = O.block_cell_values(inner, d["pdf_in"], density=rho0, minus_one=mo)
    act = (inner != 0) & (mo == 0)
    assert np.array_equal(rho[act], d["density"][act]) and (rho[~act] == 3.0).all()
    assert (u.reshape(NC, 3)[~act] == 0.0).all()
    col = O.block_collide(inner, d["pdf_in"], d["density"], d["velocity"], float(d["tau_eff"]), minus_one=mo)
    ref = d["pdf_collided"].reshape(NC, Q)
    assert np.array_equal(col.reshape(NC, Q)[act], ref[act])
    assert np.array_equal(col.reshape(NC, Q)[~act], d["pdf_in"].reshape(NC, Q)[~act])


def test_boundary_rules_closed_form():
    # pressure outlet, pressure inlet and velocity inlet on one link: only the outgoing slot changes,
    # to -f_q + feq_{18-q} + feq_q (pressure) or feq_{18-q} (velocity) of the target state
    rng = np.random.default_rng(4)
    pdf = 0.05 + 0.01 * rng.random(BP)
    rho = 1.0 + 0.01 * rng.random(NC)
    u = 0.02 * rng.standard_normal(3 * NC)
    ul = np.array([0.03, -0.01, 0.02])
    for ty, rl in ((2, 1.01), (3, 0.99), (4, 1.02)):
        cell, qin = 100, 8
        out = O.block_boundary(pdf, rho, u, [(cell, qin, ty, ul, rl)])
        q = 18 - qin
        r, uc = rho[cell], u[3 * cell:3 * cell + 3]
        rt, ut = {2: (rl, uc), 3: (r, ul), 4: (rl, ul)}[ty]
        feq = np.empty(Q)
        O.lib().lbo_feq(float(rt), O._dp(np.ascontiguousarray(ut, dtype=np.float64)), O._dp(feq))
        exp = feq[qin] if ty == 4 else -pdf[19 * cell + q] + feq[qin] + feq[q]
        assert out[19 * cell + q] == exp
        changed = np.nonzero(out != pdf)[0]
        assert set(changed) <= {19 * cell + q}
    # halfway bounce-back and undefined types are no-ops
    assert np.array_equal(O.block_boundary(pdf, rho, u, [(5, 3, 0, ul, 1.0), (6, 3, 6, ul, 1.0)]), pdf)


class OracleOps:
    """The oracle's block operators on the rows of (nblocks, ...) arrays, in place."""

    def __init__(self, grid):
        self.inner = np.ascontiguousarray(grid.inner)
        self.L = O.lib()

    def cell_values(self, b, st):
        pdf, rho, u = st
        self.L.lbo_block_cell_values(O._u8p(self.inner[b]), None, O._dp(pdf[b]), O._dp(rho[b]), O._dp(u[b]))

    def collide(self, b, st, tau):
        pdf, rho, u = st
        self.L.lbo_block_collide(O._u8p(self.inner[b]), None, O._dp(rho[b]), O._dp(u[b]), tau, O._dp(pdf[b]))

    def boundary(self, b, st, links):
        pdf, rho, u = st
        for cell, qin, ty, ul, rl in links:
            uu = np.ascontiguousarray(ul, dtype=np.float64)
            self.L.lbo_block_boundary(cell, qin, ty, O._dp(uu), rl, O._dp(rho[b]), O._dp(u[b]), O._dp(pdf[b]))

    def stream_local(self, b, st):
        self.L.lbo_block_stream_local(O._u8p(self.inner[b]), O._dp(st[0][b]))

    def stream_nonlocal(self, cell, st):
        ip = (ctypes.POINTER(ctypes.c_uint8) * 8)(*[O._u8p(self.inner[b]) for b in cell])
        pp = (ctypes.POINTER(ctypes.c_double) * 8)(*[O._dp(st[0][b]) for b in cell])
        self.L.lbo_block_stream_nonlocal(ip, pp)

    def swap_inside(self, b, st):
        self.L.lbo_block_swap_inside(O._dp(st[0][b]))


def initial_state(grid):
    pdf = np.broadcast_to(np.tile(W, NC), (grid.nblocks, BP)).copy()
    return pdf, np.ones((grid.nblocks, NC)), np.zeros((grid.nblocks, 3 * NC))


@pytest.mark.parametrize("name", ["cavity_n8_t1.npz", "cavity_n8_t10.npz", "cavity_n13_t5.npz"])
def test_composed_block_solver_reproduces_cavity_trajectories(name):
    d = golden(name)
    n, steps = int(d["n"]), int(d["steps"])
    grid = BlockGrid(n, n, n)
    st = initial_state(grid)
    iterate(grid, OracleOps(grid), st, O.tau_eff(float(d["tau"])), grid.cavity_links(float(d["ulid"])), steps)
    assert np.array_equal(grid.gather(st[0]), d["f"].reshape(-1))


@pytest.mark.parametrize("name", ["channel_8x8x8_t10_uniform.npz", "channel_16x8x8_t200_parabolic.npz"])
def test_composed_block_solver_reproduces_channel_trajectories(name):
    d = golden(name)
    nx, ny, nz = (int(v) for v in d["dims"])
    lat = O.Lattice.channel(nx, ny, nz, float(d["tau"]), float(d["umean"]), parabolic=bool(d["parabolic"]))
    prof = lat._bc_u.reshape(-1, 3)
    um = float(d["umean"])
    inlet = (lambda x, y, z: prof[x + nx * (y + ny * z)]) if bool(d["parabolic"]) else (lambda x, y, z: (um, 0.0, 0.0))
    grid = BlockGrid(nx, ny, nz)
    st = initial_state(grid)
    iterate(grid, OracleOps(grid), st, O.tau_eff(float(d["tau"])), grid.channel_links(inlet), int(d["steps"]))
    assert np.array_equal(grid.gather(st[0]), d["f"].reshape(-1))


@pytest.mark.parametrize("n", [8, 13])
def test_bench_block_grid_is_the_reference_layout(n):
    """bench.py's block-solver workload builds the same blocks, inner flags, Peano cells and lid links
    as the layout the composed trajectories above pin against the reference."""
    import bench
    nblocks, inner, cells, links = bench.regular_block_grid(n, 0.05)
    grid = BlockGrid(n, n, n)
    assert nblocks == grid.nblocks
    assert np.array_equal(inner, grid.inner != 0)
    assert np.array_equal(cells, grid.peano_cells)
    key = lambda lk: (lk[0], lk[1], lk[2], lk[3])
    assert sorted(map(key, links)) == sorted(map(key, grid.cavity_links(0.05)))
