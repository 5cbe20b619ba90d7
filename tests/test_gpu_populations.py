"""The interleaved population arrangement (LBM_POPULATIONS_INTERLEAVED: the 19 populations of each
group of 64 stored cells together) against the oracle and against the plane arrangement, through
the C-ABI.  Bars: FP64 bit-identical to the oracle; every other result (FP32-shifted states, the
in-place layout at both parities, MD coupling, the adjoint, graph replay, tuned variants)
bit-identical to the plane arrangement; the copy reference leaves the state unchanged."""
import numpy as np
import pytest

import oracle as O
from test_gpu_dense import _porous_solid

pytestmark = pytest.mark.gpu
lbm = pytest.importorskip("madpac_amd.lbm")
ORDERS = [lbm.CELL_ORDER_ROWS, lbm.CELL_ORDER_BRICKS]
ARR = [lbm.POPULATIONS_PLANES, lbm.POPULATIONS_INTERLEAVED]


def _sphere(nx, ny, nz, c, r):
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2 <= r * r


@pytest.mark.parametrize("order", ORDERS)
def test_fp64_interleaved_channel_equals_oracle(order):
    # odd extents (groups straddle rows and padding), a sphere, the parabolic inlet, a Guo force and
    # a Brinkman porosity field
    nx, ny, nz, steps = 21, 10, 13, 17
    solid = _sphere(nx, ny, nz, (8, 5, 6), 3)
    alpha = 0.4 * np.random.default_rng(5).random(nx * ny * nz)
    ora = O.Lattice.channel(nx, ny, nz, 0.62, 0.03, parabolic=True, solid=solid)
    ora.set_force((2e-6, 0.0, 1e-6))
    ora.set_alpha(alpha)
    fo = ora.step(ora.equilibrium(), steps)
    print(f"[gpu] interleaved populations, channel {nx}x{ny}x{nz}, cell order {order}", flush=True)
    with lbm.Lattice(nx, ny, nz, 0.62, force=(2e-6, 0.0, 1e-6), cell_order=order,
                     populations=lbm.POPULATIONS_INTERLEAVED) as lat:
        lat.set_flags(~solid)
        lat.set_channel(0.03, parabolic=True)
        lat.set_porosity(alpha)
        lat.init_equilibrium()
        lat.step(steps)
        f = lat.get_populations()
        m = lat.moments()
        J = lat.pressure_drop(1, nx - 2)
        rho, u = lat.fields()
    assert np.array_equal(f, fo), f"max |diff| = {np.abs(f - fo).max()}"
    mo = ora.moments(fo)
    assert abs(m[0] - mo[0]) <= 1e-13 * mo[0]
    assert abs(J - ora.pressure_drop(fo, 1, nx - 2)) <= 1e-13
    ro, uo = ora.fields(fo)
    assert np.array_equal(rho, ro) and np.array_equal(u, uo.reshape(u.shape))


def _porous_run(prec, order, arr, layout, steps, n=24, graph=False, variant=None):
    solid = _porous_solid(n, 3, frac=0.2)
    rng = np.random.default_rng(9)
    with lbm.Lattice(n, n, n, 0.58, precision=prec, periodic=(True, True, True), force=(1e-5, 0.0, 0.0),
                     layout=layout, cell_order=order, populations=arr) as lat:
        lat.set_flags(~solid)
        lat.init_equilibrium(u=0.02 * rng.standard_normal((n ** 3, 3)))
        if variant is not None:
            lat.tune(0, variant=variant)
        lat.step(steps, graph=graph)
        return lat.get_populations(), lat.moments(), lat.kinetic_energy()


@pytest.mark.parametrize("prec", [lbm.FP64, lbm.FP32_SHIFTED])
@pytest.mark.parametrize("order", ORDERS)
@pytest.mark.parametrize("layout", [lbm.LAYOUT_TWO_BUFFER, lbm.LAYOUT_AA])
def test_interleaved_equals_planes(prec, order, layout):
    print(f"[gpu] interleaved vs planes: precision {prec} order {order} layout {layout}", flush=True)
    for steps in (6, 7):  # both parities of the in-place layout
        a = _porous_run(prec, order, lbm.POPULATIONS_PLANES, layout, steps)
        b = _porous_run(prec, order, lbm.POPULATIONS_INTERLEAVED, layout, steps)
        assert np.array_equal(a[0], b[0])
        assert np.array_equal(a[1], b[1]) and a[2] == b[2]


def test_interleaved_graph_and_tuned_variants_are_bit_identical():
    ref = _porous_run(lbm.FP32_SHIFTED, lbm.CELL_ORDER_BRICKS, lbm.POPULATIONS_INTERLEAVED, 0, 40)
    g = _porous_run(lbm.FP32_SHIFTED, lbm.CELL_ORDER_BRICKS, lbm.POPULATIONS_INTERLEAVED, 0, 40, graph=True)
    assert np.array_equal(ref[0], g[0])
    for v in (1, 2, 3, 5, 10, 15):
        t = _porous_run(lbm.FP32_SHIFTED, lbm.CELL_ORDER_BRICKS, lbm.POPULATIONS_INTERLEAVED, 0, 40, variant=v)
        assert np.array_equal(ref[0], t[0]) and np.array_equal(ref[1], t[1])


@pytest.mark.parametrize("layout", [lbm.LAYOUT_TWO_BUFFER, lbm.LAYOUT_AA])
@pytest.mark.parametrize("mode", [0, 1])
def test_interleaved_coupling_equals_planes(layout, mode):
    n, steps = 12, 9
    rng = np.random.default_rng(3 + mode)
    cells = rng.choice(n ** 3, size=n ** 3 // 10, replace=False)
    umd = 0.03 * rng.standard_normal((cells.size, 3))
    rmd = 1.0 + 0.01 * rng.standard_normal(cells.size)
    u_init = 0.02 * rng.standard_normal((n ** 3, 3))
    out = []
    for arr in ARR:
        with lbm.Lattice(n, n, n, 0.7, layout=layout, populations=arr) as lat:
            lat.set_flags(np.ones(n ** 3, np.uint8))
            lat.set_cavity_lid(0.04)
            lat.init_equilibrium(u=u_init)
            lat.set_coupling_cells(cells)
            lat.set_coupling_data(mode, umd, rmd, dt=0.8)
            lat.step(steps)
            out.append((lat.get_populations(), lat.get_coupling_data()))
    assert np.array_equal(out[0][0], out[1][0])
    for a, b in zip(out[0][1], out[1][1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("layout", [lbm.LAYOUT_TWO_BUFFER, lbm.LAYOUT_AA])
def test_interleaved_adjoint_equals_planes(layout):
    n, steps = 16, 30
    solid = np.zeros((n, n, n), bool)
    solid[:, 0, :] = solid[:, -1, :] = solid[0] = solid[-1] = True
    alpha = 0.3 * np.random.default_rng(2).random(n ** 3)
    res = []
    for arr in ARR:
        with lbm.Lattice(n, n, n, 0.6, precision=lbm.FP32_SHIFTED, periodic=(True, False, False),
                         force=(1e-5, 0, 0), layout=layout, populations=arr) as lat:
            lat.set_flags(~solid)
            lat.set_porosity(alpha)
            lat.init_equilibrium()
            J, grad, _ = lat.adjoint_pressure_drop(steps, 4, 3, n - 4)
            res.append((J, grad, lat.get_populations()))
    assert res[0][0] == res[1][0]
    assert np.array_equal(res[0][1], res[1][1])
    assert np.array_equal(res[0][2], res[1][2])


@pytest.mark.parametrize("arr", ARR)
def test_copy_reference_leaves_the_state(arr):
    n = 20
    solid = _porous_solid(n, 4, frac=0.2)
    with lbm.Lattice(n, n, n, 0.7, precision=lbm.FP32_SHIFTED, periodic=(True, True, True), populations=arr) as lat:
        lat.set_flags(~solid)
        lat.init_equilibrium(u=0.01 * np.random.default_rng(1).standard_normal((n ** 3, 3)))
        lat.step(3)
        before = lat.get_populations()
        assert lat.measure_copy(3) > 0.0 and lat.measure_copy(3, pull=True) > 0.0
        assert np.array_equal(before, lat.get_populations())
        lat.step(2)
        after = lat.get_populations()
    with lbm.Lattice(n, n, n, 0.7, precision=lbm.FP32_SHIFTED, periodic=(True, True, True), populations=arr) as lat:
        lat.set_flags(~solid)
        lat.init_equilibrium(u=0.01 * np.random.default_rng(1).standard_normal((n ** 3, 3)))
        lat.step(5)
        assert np.array_equal(after, lat.get_populations())
    with lbm.Lattice(n, n, n, 0.7, layout=lbm.LAYOUT_AA, populations=arr) as lat:
        lat.set_flags(np.ones(n ** 3, np.uint8))
        lat.init_equilibrium()
        with pytest.raises(lbm.LbmError, match="two-buffer"):
            lat.measure_copy(1)
