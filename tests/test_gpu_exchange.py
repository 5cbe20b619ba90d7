"""The vertex-overlap exchange between ranks on the device block store (lbm_blocks_*_pdf_diff,
lbm_blocks_*_density_on_boundary), through the C-ABI: each operator bit for bit against the oracle
(BLB/RegularGridBlockVertex.cpp:51-160), and the partitioned block solver (one store per rank, the
differences exchanged through host or device memory) bit for bit against the single-domain run.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from block_grid import NC, Q, BlockGrid, partition_cells
from conftest import ROOT
from test_oracle_parallel import perturbed_state

pytestmark = pytest.mark.gpu
lbm = pytest.importorskip("madpac_amd.lbm")
BP = NC * Q
PB, BB = lbm.BLOCK_BOUNDARY_PDFS, lbm.BLOCK_BOUNDARY_CELLS


def _store(pdf, rho=None, u=None):
    pdf = np.asarray(pdf, dtype=np.float64).reshape(-1, BP)
    bs = lbm.BlockStore(pdf.shape[0])
    bs.set(0, pdf=pdf, density=rho, velocity=u)
    return bs


def test_pdf_difference_operators_against_the_oracle():
    rng = np.random.default_rng(31)
    before = 0.05 + 0.01 * rng.random((4, BP))
    lst = [2, 0, 3]
    print("[gpu] pdf differences of 3 of 4 blocks", flush=True)
    with _store(before) as bs:
        assert np.array_equal(bs.get_pdf_diff(lst), np.zeros((3, PB)))  # a new store holds zeros
        bs.init_pdf_diff(lst)
        d0 = np.stack([O.block_init_pdf_diff(before[b]) for b in lst])
        assert np.array_equal(bs.get_pdf_diff(lst), d0)
        # after streaming: half the incoming slots hold the outgoing value (no local swap there)
        inc, out = O.block_pdf_boundary()
        after = 0.05 + 0.01 * rng.random((4, BP))
        after[:, inc[::2]] = before[:, out[::2]]
        bs.set(0, pdf=after)
        bs.compute_pdf_diff(lst)
        d1 = np.stack([O.block_compute_pdf_diff(after[b], d0[k]) for k, b in enumerate(lst)])
        assert np.array_equal(bs.get_pdf_diff(lst), d1)
        assert (d1[:, ::2] == 0.0).all() and (d1[:, 1::2] != 0.0).all()
        nb = rng.standard_normal((3, PB))
        bs.merge_pdf_diff(lst, nb)
        merged = after.copy()
        for k, b in enumerate(lst):
            merged[b] = O.block_merge_pdf_diff(after[b], d1[k], nb[k])
        assert np.array_equal(bs.get()[0], merged)
        # the differences are the block's own and survive the merge
        assert np.array_equal(bs.get_pdf_diff(lst), d1)


def test_merge_takes_device_memory():
    # device memory from the HIP runtime the library itself is linked against
    hip = ctypes.CDLL("libamdhip64.so.7")
    rng = np.random.default_rng(32)
    pdf = 0.05 + 0.01 * rng.random((2, BP))
    inc, out = O.block_pdf_boundary()
    pdf[:, inc[::2]] = pdf[:, out[::2]]  # no swap there: the difference is 0.0 and the merge adds
    nb = rng.standard_normal((2, PB))
    print("[gpu] merge from a device buffer", flush=True)
    res = []
    for device in (False, True):
        with _store(pdf) as bs:
            bs.init_pdf_diff([0, 1])
            bs.compute_pdf_diff([0, 1])
            ids = np.array([1, 0], dtype=np.int64)
            src = nb[::-1].copy()
            if device:
                dev = ctypes.c_void_p()
                assert hip.hipMalloc(ctypes.byref(dev), ctypes.c_size_t(src.nbytes)) == 0
                try:
                    assert hip.hipMemcpy(dev, src.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(src.nbytes), 1) == 0
                    rc = bs._L.lbm_blocks_merge_pdf_diff(bs._h, ids.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), 2,
                                                         ctypes.cast(dev, ctypes.POINTER(ctypes.c_double)))
                    assert rc == 0, bs._L.lbm_blocks_last_error(bs._h).decode()
                finally:
                    assert hip.hipFree(dev) == 0
            else:
                bs.merge_pdf_diff(ids, src)
            res.append(bs.get()[0])
    assert np.array_equal(res[0], res[1])
    want = np.stack([O.block_merge_pdf_diff(pdf[b], O.block_compute_pdf_diff(pdf[b], O.block_init_pdf_diff(pdf[b])),
                                            nb[b]) for b in range(2)])
    assert np.array_equal(res[1], want) and not np.array_equal(want, pdf)


def test_boundary_density_operators_against_the_oracle():
    rng = np.random.default_rng(33)
    rho = 1.0 + 0.01 * rng.random((3, NC))
    inner = rng.random((3, NC)) < 0.6
    boundary = rng.random((3, NC)) < 0.5
    pdf = np.zeros((3, BP))
    lst = [1, 2]
    print("[gpu] boundary densities of 2 of 3 blocks", flush=True)
    with _store(pdf, rho=rho) as bs:
        bs.set_inner(0, inner)
        bs.set_boundary(0, boundary)
        bs.set_density_on_boundary(lst, dynamic_geometry=False)
        assert np.array_equal(bs.get_density_on_boundary(lst), np.zeros((2, BB)))
        bs.set_density_on_boundary(lst)
        want = np.stack([O.block_set_density_on_boundary(rho[b], inner[b], boundary[b], True) for b in lst])
        assert np.array_equal(bs.get_density_on_boundary(lst), want)
        assert (want != 0.0).any()
        nb = rng.standard_normal((2, BB))
        bs.merge_density_on_boundary(lst, nb)
        merged = rho.copy()
        for k, b in enumerate(lst):
            merged[b] = O.block_merge_density_on_boundary(rho[b], inner[b], boundary[b], nb[k])
        assert np.array_equal(bs.get()[1], merged)
        bs.reset_density_on_boundary([2])
        got = bs.get_density_on_boundary(lst)
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], np.zeros(BB))


def _single_domain(grid, pdf0, tau, links, steps):
    blocks = np.arange(grid.nblocks)
    with _store(pdf0) as bs:
        for _ in range(steps):
            bs.compute_cell_values(blocks)
            bs.collide(blocks, tau)
            bs.treat_boundary(links)
            bs.stream_local(blocks)
            bs.stream_nonlocal(grid.peano_cells)
            bs.swap_inside(blocks)
        return bs.get()


@pytest.mark.parametrize("n,ranks,overlap,steps,seed", [(12, 2, False, 6, None), (12, 8, False, 4, None),
                                                       (18, 2, True, 4, None), (25, 8, False, 3, None),
                                                       (19, 3, False, 3, 17)])
def test_partitioned_block_solver_equals_the_single_domain(n, ranks, overlap, steps, seed):
    """One store per rank, each running the block operators on the blocks of its own Peano cells with
    the pdf differences recorded around the streaming, then every rank merging every other rank's
    differences into the blocks they share: every copy of every block equals the single-domain run."""
    grid = BlockGrid(n, n, n)
    tau = O.tau_eff(0.6)
    links = grid.cavity_links(0.06)
    pdf0 = perturbed_state(grid)[0]
    print(f"[gpu] {ranks} ranks on the {n}^3 cavity (overlap {overlap}, seed {seed}), {steps} steps", flush=True)
    ref = _single_domain(grid, pdf0, tau, links, steps)
    cells_of = partition_cells(grid, ranks, overlap, seed)
    blocks_of = [np.unique(c.reshape(-1)) for c in cells_of]
    links_of = [[lk for lk in links if lk[0] in set(b.tolist())] for b in blocks_of]
    stores = [_store(pdf0) for _ in range(ranks)]
    try:
        for _ in range(steps):
            for r, bs in enumerate(stores):
                b = blocks_of[r]
                bs.compute_cell_values(b)
                bs.collide(b, tau)
                bs.treat_boundary(links_of[r])
                bs.init_pdf_diff(b)
                bs.stream_local(b)
                bs.stream_nonlocal(cells_of[r])
                bs.swap_inside(b)
                bs.compute_pdf_diff(b)
            diffs = [bs.get_pdf_diff(blocks_of[r]) for r, bs in enumerate(stores)]
            for r in range(ranks):
                for o in range(ranks):
                    if o == r:
                        continue
                    shared, _, ko = np.intersect1d(blocks_of[r], blocks_of[o], return_indices=True)
                    if len(shared):
                        stores[r].merge_pdf_diff(shared, diffs[o][ko])
        for r, bs in enumerate(stores):
            f, rho, u = bs.get()
            b = blocks_of[r]
            assert np.array_equal(f[b], ref[0][b]) and np.array_equal(rho[b], ref[1][b]) and np.array_equal(u[b], ref[2][b])
    finally:
        for bs in stores:
            bs.close()


def test_exchange_argument_errors():
    with lbm.BlockStore(3) as bs:
        with pytest.raises(lbm.LbmError, match="out of range"):
            bs.init_pdf_diff([0, 3])
        with pytest.raises(lbm.LbmError, match="twice"):
            bs.compute_pdf_diff([1, 1])
        with pytest.raises(ValueError):
            bs.merge_pdf_diff([0, 1], np.zeros(PB))
        ids = np.array([0], dtype=np.int64)
        rc = bs._L.lbm_blocks_merge_pdf_diff(bs._h, ids.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), 1, None)
        assert rc != 0 and "null" in bs._L.lbm_blocks_last_error(bs._h).decode()
        with pytest.raises(lbm.LbmError, match="bounds"):
            bs.set_boundary(2, np.zeros(2 * NC))
        # nothing above changed the store
        assert np.array_equal(bs.get_pdf_diff([0, 1, 2]), np.zeros((3, PB)))


def test_exchange_with_empty_lists_changes_nothing():
    """A rank that shares no vertex with another calls the exchange with empty lists (the reference's
    loops over an empty vertex set): every call succeeds and leaves the store as it was, and a null
    buffer is accepted when nothing is copied."""
    rng = np.random.default_rng(35)
    pdf = 0.05 + 0.01 * rng.random((2, BP))
    rho = 1.0 + 0.01 * rng.random((2, NC))
    print("[gpu] exchange calls with empty lists", flush=True)
    with _store(pdf, rho=rho) as bs:
        bs.init_pdf_diff([0, 1])
        before = bs.get_pdf_diff([0, 1])
        none = np.zeros(0, dtype=np.int64)
        bs.init_pdf_diff(none)
        bs.compute_pdf_diff(none)
        bs.merge_pdf_diff(none, np.zeros(0))
        bs.reset_density_on_boundary(none)
        bs.set_density_on_boundary(none)
        bs.merge_density_on_boundary(none, np.zeros(0))
        assert bs.get_pdf_diff(none).shape == (0, PB) and bs.get_density_on_boundary(none).shape == (0, BB)
        p = none.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
        for fn in ("lbm_blocks_get_pdf_diff", "lbm_blocks_merge_pdf_diff", "lbm_blocks_get_density_on_boundary",
                   "lbm_blocks_merge_density_on_boundary"):
            assert getattr(bs._L, fn)(bs._h, p, 0, None) == 0, bs._L.lbm_blocks_last_error(bs._h).decode()
        f, r, _ = bs.get()
        assert np.array_equal(f, pdf) and np.array_equal(r, rho)
        assert np.array_equal(bs.get_pdf_diff([0, 1]), before)


BLOCKOPS_EXE = os.path.join(ROOT, "tests", "cpp", "test_peano_blockops")


@pytest.mark.parametrize("ranks", [2, 8])
def test_cpp_vertex_exchange_drives_the_partitioned_cavity(tmp_path, ranks):
    """The C++ layer's BlockVertexExchange (initPdfDiff, computePdfDiff, getPdfDiff, mergePdfDiff on
    one DeviceBlocks per rank) around the regular-grid kernels: the same field as the single-domain
    solver (the driver also checks that the copies of every shared vertex agree)."""
    n, tau, ulid, steps = 12, 0.7, 0.08, 40
    lat = O.Lattice.cavity(n, tau, ulid)
    ref = lat.step(lat.equilibrium(), steps)
    print(f"[gpu] C++ vertex exchange, {ranks} ranks, {n}^3 cavity, {steps} steps", flush=True)
    out = tmp_path / "f.bin"
    subprocess.run([BLOCKOPS_EXE, str(n), repr(tau), repr(ulid), str(steps), str(out), str(ranks)], check=True,
                   timeout=120)
    assert np.array_equal(np.fromfile(out, dtype=np.float64), ref)
