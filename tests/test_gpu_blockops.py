"""The block solver's operators on the device block store (lbm_blocks_compute_cell_values, _collide,
_treat_boundary, _stream_local, _stream_nonlocal, _swap_inside, _vertex_first, _iterate) through the
C-ABI, bit-exact against the compiled reference's fixtures (cell kernels, streaming tables and
permutations, cavity and channel trajectories of the regular-grid block solver) and against the
oracle's restatement on seeded blocks."""
import hashlib
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from block_grid import NC, Q, BlockGrid
from conftest import ROOT, golden
from test_oracle_blocks import OracleOps, initial_state

pytestmark = pytest.mark.gpu
lbm = pytest.importorskip("madpac_amd.lbm")
BP = NC * Q


def _store(pdf, rho=None, u=None, inner=None):
    pdf = np.asarray(pdf, dtype=np.float64).reshape(-1, BP)
    bs = lbm.BlockStore(pdf.shape[0])
    bs.set(0, pdf=pdf, density=rho, velocity=u)
    if inner is not None:
        bs.set_inner(0, inner)
    return bs


@pytest.mark.parametrize("seed", [1, 2])
@pytest.mark.parametrize("fused", [False, True])
def test_cell_kernel_fixtures(seed, fused):
    d = golden(f"cellkernels_s{seed}.npz")
    print(f"[gpu] block operators on cellkernels_s{seed}, fused {fused}", flush=True)
    rng = np.random.default_rng(seed)
    # the fixture block sits between two other blocks that must stay untouched
    pdf = np.stack([rng.random(BP), d["pdf_in"], rng.random(BP)])
    rho0 = np.full((3, NC), 7.0)
    inner = np.stack([np.ones(NC), d["inner"], np.ones(NC)])
    links = [(1, c, qi, lbm.BOUNDARY_MOVING_WALL, d["uw"], 1.0) for c, qi in d["wall_links"]]
    tau = float(d["tau_eff"])
    with _store(pdf, rho0, np.zeros((3, 3 * NC)), inner) as bs:
        if fused:
            bs.vertex_first([1], tau, links)
        else:
            bs.compute_cell_values([1])
            f, rho, u = bs.get()
            m = d["inner"] != 0
            assert np.array_equal(rho[1][m], d["density"][m]) and (rho[1][~m] == 7.0).all()
            assert np.array_equal(u[1], d["velocity"])
            bs.collide([1], tau)
            assert np.array_equal(bs.get()[0][1], d["pdf_collided"])
            bs.treat_boundary(links)
            assert np.array_equal(bs.get()[0][1], d["pdf_after_wall"])
            bs.stream_local([1])
        bs.swap_inside([1])
        f, rho, u = bs.get()
        assert np.array_equal(f[1], d["pdf_after_local_stream"])
        assert np.array_equal(u[1], d["velocity"])
        assert np.array_equal(f[0], pdf[0]) and np.array_equal(f[2], pdf[2])


@pytest.mark.parametrize("seed", [1, 2])
def test_streaming_fixtures(seed):
    d = golden(f"streaming_s{seed}.npz")
    print(f"[gpu] block streaming on streaming_s{seed}", flush=True)
    vals = 1.0 + np.arange(BP)[None, :] + 10000.0 * np.arange(8)[:, None]

    def sources(out):
        k = out - 1.0
        blk = np.floor(k / 10000.0)
        return (blk * BP + (k - 10000.0 * blk)).astype(np.int32)
    with _store(vals) as bs:
        bs.stream_local([0])
        bs.swap_inside([1])
        bs.set_inner(2, d["inner_local"])
        bs.stream_local([2])
        f = bs.get()[0]
    assert np.array_equal(sources(f[0]), d["local_all"])
    assert np.array_equal(sources(f[1]) - BP, d["swapped"])
    assert np.array_equal(sources(f[2]) - 2 * BP, d["local_masked"])
    for inner, key in ((np.ones((8, NC)), "nonlocal_all"), (d["inner_nonlocal"], "nonlocal_masked")):
        with _store(vals, inner=inner) as bs:
            bs.stream_nonlocal(np.arange(8))
            assert np.array_equal(sources(bs.get()[0]), d[key])


def test_many_cells_and_blocks_against_the_oracle():
    # a 4 x 3 x 3 grid of blocks with seeded values and inner flags; every Peano cell in one launch
    grid = BlockGrid(19, 13, 13)
    assert grid.nb == (5, 4, 4) and grid.peano_cells.shape[0] == 36
    rng = np.random.default_rng(11)
    pdf = rng.random((grid.nblocks, BP))
    inner = (rng.random((grid.nblocks, NC)) < 0.8).astype(np.uint8)
    print(f"[gpu] cross-block streaming over {grid.peano_cells.shape[0]} cells", flush=True)
    ref = pdf.copy()
    ops = OracleOps(grid)
    ops.inner = inner
    for cell in grid.peano_cells:
        ops.stream_nonlocal(cell, (ref, None, None))
    blocks = np.arange(grid.nblocks)
    for b in blocks:
        ops.stream_local(b, (ref, None, None))
    with _store(pdf, inner=inner) as bs:
        bs.stream_nonlocal(grid.peano_cells[::-1])  # list order does not matter
        bs.stream_local(rng.permutation(blocks))
        assert np.array_equal(bs.get()[0], ref)


def test_overlap_cells_are_skipped():
    d = golden("cellkernels_s2.npz")
    mo = np.zeros(NC, np.uint8)
    mo[3::7] = 1
    tau = float(d["tau_eff"])
    rho_o, u_o = O.block_cell_values(d["inner"], d["pdf_in"], density=np.full(NC, 2.0), minus_one=mo)
    col_o = O.block_collide(d["inner"], d["pdf_in"], rho_o, u_o, tau, minus_one=mo)
    with _store(d["pdf_in"], np.full((1, NC), 2.0), np.zeros((1, 3 * NC)), d["inner"]) as bs:
        bs.set_overlap(0, mo)
        bs.compute_cell_values([0], use_overlap=True)
        bs.collide([0], tau, use_overlap=True)
        f, rho, u = bs.get()
    assert np.array_equal(rho[0], rho_o) and np.array_equal(u[0], u_o) and np.array_equal(f[0], col_o)


def test_boundary_types_against_the_oracle():
    rng = np.random.default_rng(5)
    pdf = 0.05 + 0.01 * rng.random((2, BP))
    rho = 1.0 + 0.01 * rng.random((2, NC))
    u = 0.02 * rng.standard_normal((2, 3 * NC))
    links, seen = [], set()
    for k in range(300):
        b, cell, qin = k % 2, int(rng.integers(NC)), int(rng.integers(Q))
        if (b, cell, qin) in seen:
            continue
        seen.add((b, cell, qin))
        links.append((b, cell, qin, int(rng.integers(6)), tuple(0.03 * rng.standard_normal(3)), 1.0 + 0.02 * rng.random()))
    ref = np.stack([O.block_boundary(pdf[b], rho[b], u[b], [lk[1:] for lk in links if lk[0] == b]) for b in range(2)])
    with _store(pdf, rho, u) as bs:
        bs.treat_boundary(links)
        assert np.array_equal(bs.get()[0], ref)


def _regular(grid, bs, tau, links, steps, mode):
    blocks = np.arange(grid.nblocks)
    if mode == "iterate":
        bs.iterate(blocks, grid.peano_cells, tau, links, steps)
        return
    for _ in range(steps):
        if mode == "separate":
            bs.compute_cell_values(blocks)
            bs.collide(blocks, tau)
            bs.treat_boundary(links)
            bs.stream_local(blocks)
        else:
            bs.vertex_first(blocks, tau, links)
        bs.stream_nonlocal(grid.peano_cells)
        bs.swap_inside(blocks)


@pytest.mark.parametrize("mode", ["separate", "vertex_first", "iterate"])
@pytest.mark.parametrize("name", ["cavity_n8_t1.npz", "cavity_n8_t100.npz", "cavity_n13_t5.npz"])
def test_regular_grid_cavity_trajectories(name, mode):
    d = golden(name)
    n, steps = int(d["n"]), int(d["steps"])
    print(f"[gpu] block solver {mode} on {name}", flush=True)
    grid = BlockGrid(n, n, n)
    with _store(initial_state(grid)[0], np.ones((grid.nblocks, NC)), None, grid.inner) as bs:
        _regular(grid, bs, O.tau_eff(float(d["tau"])), grid.cavity_links(float(d["ulid"])), steps, mode)
        assert np.array_equal(grid.gather(bs.get()[0]), d["f"].reshape(-1))


@pytest.mark.parametrize("mode", ["separate", "iterate"])
@pytest.mark.parametrize("name", ["channel_8x8x8_t10_uniform.npz", "channel_16x8x8_t200_parabolic.npz"])
def test_regular_grid_channel_trajectories(name, mode):
    d = golden(name)
    nx, ny, nz = (int(v) for v in d["dims"])
    print(f"[gpu] block solver {mode} on {name}", flush=True)
    lat = O.Lattice.channel(nx, ny, nz, float(d["tau"]), float(d["umean"]), parabolic=bool(d["parabolic"]))
    prof = lat._bc_u.reshape(-1, 3)
    um = float(d["umean"])
    inlet = (lambda x, y, z: prof[x + nx * (y + ny * z)]) if bool(d["parabolic"]) else (lambda x, y, z: (um, 0.0, 0.0))
    grid = BlockGrid(nx, ny, nz)
    with _store(initial_state(grid)[0], np.ones((grid.nblocks, NC)), None, grid.inner) as bs:
        _regular(grid, bs, O.tau_eff(float(d["tau"])), grid.channel_links(inlet), int(d["steps"]), mode)
        assert np.array_equal(grid.gather(bs.get()[0]), d["f"].reshape(-1))


def test_regular_grid_cavity64_1000_steps():
    d = golden("cavity_n64_t1000.npz")
    print("[gpu] block solver iterate on the 64^3 cavity, 1000 steps", flush=True)
    grid = BlockGrid(64, 64, 64)
    assert grid.nblocks == 12 ** 3
    with _store(initial_state(grid)[0], np.ones((grid.nblocks, NC)), None, grid.inner) as bs:
        bs.iterate(np.arange(grid.nblocks), grid.peano_cells, O.tau_eff(float(d["tau"])),
                   grid.cavity_links(float(d["ulid"])), 1000)
        f = grid.gather(bs.get()[0])
    assert hashlib.sha256(f.astype("<f8").tobytes()).hexdigest() == str(d["sha256"])


def test_invalid_lists_are_rejected():
    with lbm.BlockStore(10) as bs:
        with pytest.raises(lbm.LbmError, match="twice"):
            bs.collide([1, 2, 1], 0.6)
        with pytest.raises(lbm.LbmError, match="out of range"):
            bs.stream_local([10])
        cells = np.arange(8)
        with pytest.raises(lbm.LbmError, match="two vertices"):
            bs.stream_nonlocal(np.array([0, 1, 2, 3, 4, 5, 6, 0]))
        with pytest.raises(lbm.LbmError, match="two cells"):
            bs.stream_nonlocal(np.concatenate([cells, cells]))
        with pytest.raises(lbm.LbmError, match="not in the block list"):
            bs.vertex_first([0, 1], 0.6, [(2, 0, 3, 1, (0, 0, 0), 1.0)])
        with pytest.raises(lbm.LbmError, match="same"):
            bs.treat_boundary([(0, 5, 3, 1, (0, 0, 0), 1.0), (0, 5, 3, 2, (0, 0, 0), 1.0)])
        with pytest.raises(lbm.LbmError, match="BoundaryType"):
            bs.treat_boundary([(0, 5, 3, 7, (0, 0, 0), 1.0)])
        with pytest.raises(lbm.LbmError, match="tau"):
            bs.collide([0], 0.0)


BLOCKOPS_EXE = os.path.join(ROOT, "tests", "cpp", "test_peano_blockops")


@pytest.mark.parametrize("case", ["fixture", "oracle"])
def test_cpp_block_kernels_drive_the_cavity(tmp_path, case):
    """The C++ layer's kernel classes (BlockComputeCellValues, BlockBGK, BlockBoundaryConditionManager,
    AABlockStreaming on DeviceBlocks) composed as the regular-grid mapping calls them."""
    if case == "fixture":
        d = golden("cavity_n13_t5.npz")
        n, tau, ulid, steps = int(d["n"]), float(d["tau"]), float(d["ulid"]), int(d["steps"])
        ref = d["f"].reshape(-1)
    else:
        n, tau, ulid, steps = 12, 0.7, 0.08, 40
        lat = O.Lattice.cavity(n, tau, ulid)
        ref = lat.step(lat.equilibrium(), steps)
    print(f"[gpu] C++ block kernels on the {n}^3 cavity, {steps} steps", flush=True)
    out = tmp_path / "f.bin"
    subprocess.run([BLOCKOPS_EXE, str(n), repr(tau), repr(ulid), str(steps), str(out)], check=True, timeout=120)
    f = np.fromfile(out, dtype=np.float64)
    assert np.array_equal(f, ref)


def test_iterate_with_a_partial_cell_list_equals_the_separate_operators():
    """lbm_blocks_iterate carries the cross-block swaps of one iteration into the next block pass; with a
    cell list that leaves cells out (their links are not swapped), seeded inner flags and lid links it
    must still give the state of the separate operators, iteration by iteration."""
    grid = BlockGrid(14, 11, 13)
    rng = np.random.default_rng(21)
    inner = grid.inner.copy()
    inner[rng.random(inner.shape) < 0.1] = 0
    cells = grid.peano_cells[rng.random(len(grid.peano_cells)) < 0.7]
    links = [lk for lk in grid.cavity_links(0.07) if inner[lk[0], lk[1]]]
    tau = O.tau_eff(0.57)
    pdf0 = initial_state(grid)[0] * (1.0 + 0.01 * rng.standard_normal((grid.nblocks, BP)))
    blocks = np.arange(grid.nblocks)
    print(f"[gpu] iterate with {len(cells)} of {len(grid.peano_cells)} cells", flush=True)
    inner2 = inner.copy()
    inner2[rng.random(inner.shape) < 0.05] = 0
    links2 = [lk for lk in links if inner2[lk[0], lk[1]]]
    order = rng.permutation(blocks)
    out = []
    for mode in ("separate", "iterate"):
        with _store(pdf0, np.ones((grid.nblocks, NC)), None, inner) as bs:
            # one step, three more (the same cells and flags: the swap mask is reused), then changed flags
            for steps, blk, lk in ((1, blocks, links), (3, order, links), (2, blocks, links2)):
                if steps == 2:
                    bs.set_inner(0, inner2)
                if mode == "iterate":
                    bs.iterate(blk, cells, tau, lk, steps)
                else:
                    for _ in range(steps):
                        bs.vertex_first(blocks, tau, lk)
                        bs.stream_nonlocal(cells)
                        bs.swap_inside(blocks)
                out.append(bs.get())
    for a, b in zip(out[:3], out[3:]):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("mode", ["vertex_first", "iterate"])
def test_fused_passes_take_links_on_every_direction(mode):
    """Boundary links of every type on every direction, the rest population's (q_in = 9) included, on
    seeded inner flags: the fused passes (links applied to the block in LDS, with the cell values the
    pass computed) equal the separate operators (links applied in global memory) bit for bit."""
    grid = BlockGrid(12, 12, 12)
    rng = np.random.default_rng(41)
    pdf0 = initial_state(grid)[0] * (1.0 + 0.01 * rng.standard_normal((grid.nblocks, BP)))
    inner = grid.inner.copy()
    inner[rng.random(inner.shape) < 0.1] = 0
    links, seen = [], set()
    for _ in range(4000):
        b, c, qi = int(rng.integers(grid.nblocks)), int(rng.integers(NC)), int(rng.integers(Q))
        if (b, c, qi) in seen:
            continue
        seen.add((b, c, qi))
        links.append((b, c, qi, int(rng.integers(6)), tuple(0.03 * rng.standard_normal(3)), 1.0 + 0.02 * rng.random()))
    assert sum(lk[2] == Q // 2 for lk in links) > 100
    tau = O.tau_eff(0.6)
    print(f"[gpu] {mode} with {len(links)} links on every direction", flush=True)
    out = []
    for m in ("separate", mode):
        with _store(pdf0, np.ones((grid.nblocks, NC)), None, inner) as bs:
            _regular(grid, bs, tau, links, 3, m)
            out.append(bs.get())
    assert np.isfinite(out[0][0]).all()
    for x, y in zip(*out):
        assert np.array_equal(x, y)


def test_iterate_rejects_inconsistent_cells():
    grid = BlockGrid(8, 8, 8)
    blocks = np.arange(grid.nblocks)
    with lbm.BlockStore(grid.nblocks) as bs:
        with pytest.raises(lbm.LbmError, match="not in the block list"):
            bs.iterate(blocks[:-1], grid.peano_cells, 0.6)
        bad = grid.peano_cells.copy()
        bad[0, [1, 2]] = bad[0, [2, 1]]  # two vertices exchanged: the neighbours disagree with the other cells
        with pytest.raises(lbm.LbmError, match="disagree"):
            bs.iterate(blocks, bad, 0.6)


def test_two_level_vertex_sequence_equals_the_oracle():
    """The spacetree mapping's per-vertex sequence on one refined coarse cell, batched over all 64 fine
    vertices (SpacetreeGrid2RegularBlockSolver.cpp:186-273, :349, :573-577): overlap search from the
    -1.0 markers, cell values and collision skipping the overlap cells (hasMinusOneEntries),
    prolongation from the father blocks, local streaming, cross-block streaming over the 27 fine Peano
    cells, in-cell swap, restriction into the fathers.  Every stage against the oracle, bit for bit."""
    rng = np.random.default_rng(31)
    ncoarse, nfine = 8, 64
    coarse = rng.uniform(0.5, 1.5, (ncoarse, BP))
    fine = rng.uniform(0.5, 1.5, (nfine, BP))
    fine[rng.random(fine.shape) < 0.003] = -1.0
    inner = np.ones((ncoarse + nfine, NC), np.uint8)
    inner[ncoarse:][rng.random((nfine, NC)) < 0.08] = 0
    pos = [(x, y, z) for z in range(4) for y in range(4) for x in range(4)]
    links = [((p[0] // 2) + 2 * (p[1] // 2) + 4 * (p[2] // 2), ncoarse + k, p) for k, p in enumerate(pos)]
    cells = np.array([[ncoarse + (cx + (v & 1)) + 4 * (cy + ((v >> 1) & 1)) + 16 * (cz + (v >> 2)) for v in range(8)]
                      for cz in range(3) for cy in range(3) for cx in range(3)], dtype=np.int64)
    fine_ids = np.arange(ncoarse, ncoarse + nfine)
    tau = O.tau_eff(0.7)
    # oracle, vertex by vertex
    oc, of = coarse.copy(), fine.copy()
    orho = np.ones((nfine, NC))
    ou = np.zeros((nfine, 3 * NC))
    ov = np.zeros((nfine, NC), bool)
    for k, p in enumerate(pos):
        ov[k] = O.amr_find_overlap(of[k])
        orho[k], ou[k] = O.block_cell_values(inner[ncoarse + k], of[k], density=orho[k], minus_one=ov[k])
        of[k] = O.block_collide(inner[ncoarse + k], of[k], orho[k], ou[k], tau, minus_one=ov[k])
    stage1 = (of.copy(), orho.copy(), ou.copy())
    for k, p in enumerate(pos):
        of[k], orho[k], ou[k] = O.amr_prolongate(p, oc[links[k][0]], of[k], orho[k], ou[k], ov[k])
        of[k] = O.block_stream_local(inner[ncoarse + k], of[k])
    for cell in cells:
        blk = cell - ncoarse
        of[blk] = O.block_stream_nonlocal(inner[cell], of[blk])
    for k in range(nfine):
        of[k] = O.block_swap_inside(of[k])
    for k, p in enumerate(pos):
        oc[links[k][0]] = O.amr_restrict(p, oc[links[k][0]], of[k])
    print("[gpu] two-level vertex sequence, 8 coarse and 64 fine blocks", flush=True)
    with _store(np.concatenate([coarse, fine]), np.ones((ncoarse + nfine, NC)),
                np.zeros((ncoarse + nfine, 3 * NC)), inner) as bs:
        assert np.array_equal(bs.find_overlap(links), ov)
        bs.compute_cell_values(fine_ids, use_overlap=True)
        bs.collide(fine_ids, tau, use_overlap=True)
        f, rho, u = bs.get(ncoarse)
        assert np.array_equal(f, stage1[0]) and np.array_equal(rho, stage1[1]) and np.array_equal(u, stage1[2])
        bs.prolongate(links)
        bs.stream_local(fine_ids)
        bs.stream_nonlocal(cells)
        bs.swap_inside(fine_ids)
        bs.restrict(links)
        f, rho, u = bs.get()
    assert np.array_equal(f[:ncoarse], oc) and np.array_equal(f[ncoarse:], of)
    assert np.array_equal(rho[ncoarse:], orho) and np.array_equal(u[ncoarse:], ou)


def test_block_solver_equals_the_fused_step_at_128():
    """Size-independent property at a large lattice: the reference's three sweeps on its block layout
    (lbm_blocks_iterate, 12,167 blocks) and the fused pull step on a dense lattice give the same state
    bit for bit after 300 steps of the 128^3 cavity."""
    n, tau, ulid, steps = 128, 0.6, 0.05, 300
    print(f"[gpu] block solver against the fused step, {n}^3 cavity, {steps} steps", flush=True)
    grid = BlockGrid(n, n, n)
    with _store(initial_state(grid)[0], np.ones((grid.nblocks, NC)), None, grid.inner) as bs:
        bs.iterate(np.arange(grid.nblocks), grid.peano_cells, O.tau_eff(tau), grid.cavity_links(ulid), steps)
        fb = grid.gather(bs.get()[0])
    with lbm.cavity(n, tau, ulid) as lat:
        lat.step(steps)
        fd = lat.get_populations()
    assert np.array_equal(fb, fd)
