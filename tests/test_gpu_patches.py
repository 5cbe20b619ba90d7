"""Batched uneven patches (persistent launch, fused ghost-layer exchange, hipGraph, autotuner)
against the dense periodic oracle: a torus of equal-edge patches is the same lattice as one
periodic box, so every torus must reproduce the oracle bit for bit in FP64."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
lbm = pytest.importorskip("madpac_amd.lbm")
wl = pytest.importorskip("madpac_amd.workloads")


def _dense_of(patch_arrays, dims, n):
    """Assemble per-patch arrays (..., n, n, n) of one torus into the dense (..., c n, b n, a n) box."""
    a, b, c = dims
    lead = patch_arrays[0].shape[:-3]
    out = np.zeros(lead + (c * n, b * n, a * n))
    k = 0
    for kk in range(c):
        for jj in range(b):
            for ii in range(a):
                out[..., kk * n:(kk + 1) * n, jj * n:(jj + 1) * n, ii * n:(ii + 1) * n] = patch_arrays[k]
                k += 1
    return out


def _setup(tori, seed=0):
    """tori: list of (dims, n).  Returns table, per-torus info and per-cell u (patch order)."""
    rng = np.random.default_rng(seed)
    table, info = [], []
    for dims, n in tori:
        cnt = dims[0] * dims[1] * dims[2]
        taus = rng.uniform(0.55, 1.5, cnt)
        first = len(table)
        table += wl.torus_table(n, dims, taus, first)
        info.append((first, dims, n))
    cells = sum(t[0] ** 3 for t in table)
    u = 0.01 * rng.standard_normal((cells, 3))
    return table, info, u


def _oracle_torus(table, first, dims, n, u_patch, steps):
    a, b, c = dims
    ids = list(range(first, first + a * b * c))
    offs = np.cumsum([0] + [t[0] ** 3 for t in table])
    u_p = [u_patch[offs[p]:offs[p + 1]].reshape(n, n, n, 3).transpose(3, 0, 1, 2) for p in ids]
    tau_p = [np.full((n, n, n), table[p][1]) for p in ids]
    u_d = np.moveaxis(_dense_of(u_p, dims, n), 0, -1)
    tau_d = _dense_of(tau_p, dims, n)
    lat = O.Lattice(a * n, b * n, c * n, periodic=(True, True, True))
    lat.set_solid(np.zeros(tau_d.shape, np.uint8))
    lat.set_tau_cells(tau_d)
    f = lat.step(lat.equilibrium(u=u_d.reshape(-1, 3)), steps)
    rho, u = lat.fields(f)
    return f.reshape(19, c * n, b * n, a * n), ids, rho.reshape(c * n, b * n, a * n), u.reshape(c * n, b * n, a * n, 3)


@pytest.mark.parametrize("tori", [
    [((2, 2, 2), 8)],
    [((2, 1, 1), 16), ((1, 1, 1), 8), ((3, 2, 1), 8)],
    [((1, 1, 1), 32), ((2, 1, 1), 24)],
    # edges outside {8, 16, 24, 32} take the generic index split; n = 2 puts every cell on every face
    [((1, 1, 1), 2), ((1, 2, 1), 5), ((1, 1, 2), 12), ((2, 2, 1), 3)],
])
def test_patches_fp64_bitwise_equal_dense_oracle(tori):
    table, info, u = _setup(tori, seed=len(tori))
    steps = 15
    print(f"[gpu] patches {[(d, n) for d, n in tori]} FP64, {steps} steps", flush=True)
    with lbm.PatchSet(table, precision=lbm.FP64) as ps:
        ps.init(u=u)
        ps.step(steps)
        fp = ps.get_populations()
        rho_p, u_p = ps.fields()
    offs = np.cumsum([0] + [t[0] ** 3 for t in table])
    for first, dims, n in info:
        fo, ids, rho_o, u_o = _oracle_torus(table, first, dims, n, u, steps)
        fg = _dense_of([fp[p] for p in ids], dims, n)
        assert np.array_equal(fg, fo), f"torus {dims}x{n}: max |diff| {np.abs(fg - fo).max()}"
        # lbm_patches_get_fields: the reference's density and velocity of the same state
        rg = _dense_of([rho_p[offs[p]:offs[p + 1]].reshape(n, n, n) for p in ids], dims, n)
        ug = _dense_of([u_p[offs[p]:offs[p + 1]].reshape(n, n, n, 3).transpose(3, 0, 1, 2) for p in ids], dims, n)
        assert np.array_equal(rg, rho_o)
        assert np.array_equal(ug, np.moveaxis(u_o, -1, 0))


def test_graph_replay_and_autotune_are_bit_identical():
    table, info, u = _setup([((2, 2, 1), 16), ((2, 2, 2), 8), ((3, 1, 1), 24)], seed=7)
    res = []
    for mode in ("plain", "graph", "tuned"):
        with lbm.PatchSet(table, precision=lbm.FP64) as ps:
            ps.init(u=u)
            if mode == "tuned":
                t = ps.autotune()
                assert len(t) == 5
            g = mode != "plain"
            ps.step(4, graph=g)
            mid = ps.get_populations()  # readout between replays
            ps.step(7, graph=g)  # odd count: 3 graph replays + 1 plain launch
            ps.step(6, graph=g)
            res.append((mid, ps.get_populations(), ps.moments()))
    for k in (1, 2):
        for a, b in zip(res[0][0], res[k][0]):
            assert np.array_equal(a, b)
        for a, b in zip(res[0][1], res[k][1]):
            assert np.array_equal(a, b)
    # mass is conserved (all cells fluid, periodic tori)
    m0 = sum(t[0] ** 3 for t in table)
    assert abs(res[0][2][0] - m0) < 1e-9 * m0


@pytest.mark.parametrize("prec", [lbm.FP64, lbm.FP32_SHIFTED])
def test_every_tiling_is_bit_identical(prec):
    # linear runs and box tiles, persistent grids and one workgroup per tile: same bits
    table, info, u = _setup([((2, 1, 1), 24), ((2, 2, 2), 8), ((1, 1, 1), 32), ((2, 1, 1), 16)], seed=9)
    out = []
    for tiling in ((32, 8, 1, 4), (256, 0, 0, 0), (256, 0, 0, 8), (8, 8, 4, 0), (16, 16, 1, 2), (16, 4, 4, 64),
                   (256, 0, 0, 8, 1), (256, 0, 0, 64, 1), (16, 8, 2, 4, 1)):
        with lbm.PatchSet(table, precision=prec) as ps:
            ps.init(u=u)
            ps.set_tiling(*tiling)
            ps.step(9, graph=True)
            out.append(ps.get_populations())
    for other in out[1:]:
        for a, b in zip(out[0], other):
            assert np.array_equal(a, b)


def test_patches_fp32_shifted_close_to_fp64_oracle():
    table, info, u = _setup([((2, 2, 2), 8)], seed=3)
    steps = 50
    with lbm.PatchSet(table, precision=lbm.FP32_SHIFTED) as ps:
        ps.init(u=u)
        ps.step(steps, graph=True)
        fp = ps.get_populations()
    first, dims, n = info[0]
    fo, ids, _, _ = _oracle_torus(table, first, dims, n, u, steps)
    fg = _dense_of([fp[p] for p in ids], dims, n)
    c = O.C.astype(float)
    j_o = np.tensordot(c.T, fo, axes=([1], [0]))
    j_g = np.tensordot(c.T, fg, axes=([1], [0]))
    rel = np.linalg.norm(j_g - j_o) / np.linalg.norm(j_o)
    assert rel <= 1e-5, rel
    assert np.max(np.abs(fg - fo)) <= 1e-7


def test_patch_table_validation():
    table, _, _ = _setup([((2, 1, 1), 8)])
    bad = [(16, table[0][1], table[0][2])] + table[1:]  # neighbour edges differ
    with pytest.raises(lbm.LbmError, match="edge"):
        lbm.PatchSet(bad)
    # patch 0 sees itself across +x while patch 1 still sees patch 0 across -x: two writers would
    # share patch 0's -x ghost records
    nbr = list(table[0][2])
    nbr[10] = 0
    with pytest.raises(lbm.LbmError, match="reciprocal"):
        lbm.PatchSet([(table[0][0], table[0][1], nbr)] + table[1:])
    with lbm.PatchSet(table) as ps:
        with pytest.raises(lbm.LbmError, match="tile"):
            ps.set_tiling(5, 5, 5, 4)
        with pytest.raises(lbm.LbmError, match="state"):
            ps.step(1)


def _patches_of(dense, dims, n):
    """Split a dense (c n, b n, a n) box into the per-patch (n, n, n) arrays of one torus, in the
    patch order of workloads.torus_table (inverse of _dense_of)."""
    a, b, c = dims
    return [dense[kk * n:(kk + 1) * n, jj * n:(jj + 1) * n, ii * n:(ii + 1) * n]
            for kk in range(c) for jj in range(b) for ii in range(a)]


def _torus_obstacles(info, seed):
    """Random periodic spheres (radius 1..3) in each torus: (fluid per cell in patch order, dense solid
    per torus)."""
    rng = np.random.default_rng(seed)
    fluid, solids = [], []
    for first, dims, n in info:
        a, b, c = dims
        shape = (c * n, b * n, a * n)
        z, y, x = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
        solid = np.zeros(shape, bool)
        for _ in range(max(2, (a * b * c * n ** 3) // 300)):
            ctr = [rng.integers(0, s) for s in shape]
            r = rng.integers(1, 4)
            d2 = 0
            for coord, cc, s in zip((z, y, x), ctr, shape):
                d = np.abs(coord - cc)
                d = np.minimum(d, s - d)
                d2 = d2 + d * d
            solid |= d2 <= r * r
        solids.append(solid)
        fluid += [~p.reshape(-1) for p in _patches_of(solid, dims, n)]
    return np.concatenate(fluid), solids


@pytest.mark.parametrize("tori", [
    [((2, 2, 2), 8)],
    # edge classes with and without the constant-divisor split; n = 3 puts most cells on faces and edges
    [((2, 1, 1), 16), ((1, 2, 2), 6), ((3, 1, 1), 3)],
])
def test_patches_with_obstacles_bitwise_equal_dense_oracle(tori):
    table, info, u = _setup(tori, seed=21)
    fluid, solids = _torus_obstacles(info, seed=22)
    assert 0 < (~fluid).sum() < fluid.size
    steps = 12
    print(f"[gpu] patches with obstacles {[(d, n) for d, n in tori]} FP64, {steps} steps", flush=True)
    results = []
    for tiling, graph in (((256, 0, 0, 0), False), ((16, 8, 2, 4), True)):
        with lbm.PatchSet(table, precision=lbm.FP64) as ps:
            ps.set_flags(fluid)
            ps.init(u=u)
            ps.set_tiling(*tiling)
            ps.step(steps, graph=graph)
            results.append((ps.get_populations(), ps.fields(), ps.moments()))
    for fa, fb in zip(results[0][0], results[1][0]):
        assert np.array_equal(fa, fb)
    fp, (rho_p, u_p), m = results[0]
    offs = np.cumsum([0] + [t[0] ** 3 for t in table])
    mass = 0.0
    for (first, dims, n), solid in zip(info, solids):
        a, b, c = dims
        ids = list(range(first, first + a * b * c))
        u_p0 = [u[offs[p]:offs[p + 1]].reshape(n, n, n, 3).transpose(3, 0, 1, 2) for p in ids]
        tau_d = _dense_of([np.full((n, n, n), table[p][1]) for p in ids], dims, n)
        lat = O.Lattice(a * n, b * n, c * n, periodic=(True, True, True))
        lat.set_solid(solid.astype(np.uint8))
        lat.set_tau_cells(tau_d)
        u_d = np.moveaxis(_dense_of(u_p0, dims, n), 0, -1)
        fo = lat.step(lat.equilibrium(u=u_d.reshape(-1, 3)), steps)
        rho_o, u_o = lat.fields(fo)
        mass += lat.moments(fo)[0]
        fl = ~solid
        fo = fo.reshape(19, c * n, b * n, a * n)
        fg = _dense_of([fp[p] for p in ids], dims, n)
        assert np.array_equal(fg[:, fl], fo[:, fl]), f"torus {dims}x{n}: max |diff| {np.abs(fg - fo)[:, fl].max()}"
        assert not fg[:, solid].any(), "solid cells read back as 0"
        rg = _dense_of([rho_p[offs[p]:offs[p + 1]].reshape(n, n, n) for p in ids], dims, n)
        ug = _dense_of([u_p[offs[p]:offs[p + 1]].reshape(n, n, n, 3).transpose(3, 0, 1, 2) for p in ids], dims, n)
        assert np.array_equal(rg[fl], rho_o.reshape(rg.shape)[fl])
        assert np.array_equal(ug[:, fl], np.moveaxis(u_o.reshape(c * n, b * n, a * n, 3), -1, 0)[:, fl])
        assert not rg[solid].any()
    # mass over fluid cells: conserved by bounce-back, equal to the oracle's up to summation order
    assert abs(m[0] - mass) <= 1e-12 * mass
    assert abs(m[0] - fluid.sum()) <= 1e-10 * fluid.sum()


def test_patches_with_obstacles_fp32_shifted_close_to_fp64():
    table, info, u = _setup([((2, 2, 2), 8), ((2, 1, 1), 12)], seed=31)
    fluid, _ = _torus_obstacles(info, seed=32)
    res = {}
    for prec in (lbm.FP64, lbm.FP32_SHIFTED):
        with lbm.PatchSet(table, precision=prec) as ps:
            ps.set_flags(fluid)
            ps.init(u=u)
            ps.step(60, graph=True)
            res[prec] = ps.fields()[1][fluid]
    d = res[lbm.FP32_SHIFTED] - res[lbm.FP64]
    rel = np.linalg.norm(d) / np.linalg.norm(res[lbm.FP64])
    assert rel <= 1e-5, rel


def test_patch_flags_require_init_and_can_be_cleared():
    table, info, u = _setup([((2, 2, 1), 8)], seed=41)
    fluid, _ = _torus_obstacles(info, seed=42)
    with lbm.PatchSet(table) as ps:
        ps.init(u=u)
        ps.set_flags(fluid)
        with pytest.raises(lbm.LbmError, match="state"):
            ps.step(1)
        with pytest.raises(ValueError, match="flags"):
            ps.set_flags(fluid[:-1])
        ps.init(u=u)
        ps.step(3)
        assert abs(ps.moments()[0] - fluid.sum()) < 1e-10 * fluid.sum()
        ps.set_flags(None)  # all fluid again: the plain torus
        ps.init(u=u)
        ps.step(3)
        f_clear = ps.get_populations()
    with lbm.PatchSet(table) as ps:
        ps.init(u=u)
        ps.step(3)
        for a_, b_ in zip(f_clear, ps.get_populations()):
            assert np.array_equal(a_, b_)
