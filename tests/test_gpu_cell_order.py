"""The two device cell orders (rows and 4x4x4 bricks) against the oracle, and the automatic switch
at lbm_set_flags, which carries the per-cell arrays already on the device (porosity, inlet
velocities, coupling cell ids) into the new order.  Bar: FP64 bit-identical to the oracle in both
orders; FP32-shifted bit-identical between the orders (same per-cell arithmetic)."""
import numpy as np
import pytest

import oracle as O
from test_gpu_dense import _porous_solid

pytestmark = pytest.mark.gpu
lbm = pytest.importorskip("madpac_amd.lbm")
ORDERS = [lbm.CELL_ORDER_ROWS, lbm.CELL_ORDER_BRICKS]


def _sphere(nx, ny, nz, c, r):
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2 <= r * r


def _channel_case(seed):
    # odd extents (padding to whole bricks), a sphere, the parabolic inlet (per-cell velocities),
    # a Guo force and a Brinkman porosity field
    nx, ny, nz, steps = 21, 10, 13, 17
    solid = _sphere(nx, ny, nz, (8, 5, 6), 3)
    alpha = 0.4 * np.random.default_rng(seed).random(nx * ny * nz)
    ora = O.Lattice.channel(nx, ny, nz, 0.62, 0.03, parabolic=True, solid=solid)
    ora.set_force((2e-6, 0.0, 1e-6))
    ora.set_alpha(alpha)
    return (nx, ny, nz, steps), solid, alpha, ora


@pytest.mark.parametrize("order", ORDERS)
def test_fp64_channel_obstacle_porosity_in_both_orders(order):
    (nx, ny, nz, steps), solid, alpha, ora = _channel_case(3)
    fo = ora.step(ora.equilibrium(), steps)
    print(f"[gpu] channel {nx}x{ny}x{nz} in cell order {order}", flush=True)
    with lbm.Lattice(nx, ny, nz, 0.62, force=(2e-6, 0.0, 1e-6), cell_order=order) as lat:
        lat.set_flags(~solid)
        assert lat.cell_order == order
        lat.set_channel(0.03, parabolic=True)
        lat.set_porosity(alpha)
        lat.init_equilibrium()
        lat.step(steps)
        f = lat.get_populations()
        m = lat.moments()
        J = lat.pressure_drop(1, nx - 2)
    assert np.array_equal(f, fo), f"max |diff| = {np.abs(f - fo).max()}"
    mo = ora.moments(fo)
    assert abs(m[0] - mo[0]) <= 1e-13 * mo[0]
    assert abs(J - ora.pressure_drop(fo, 1, nx - 2)) <= 1e-13


def _row_line_waste(solid, per_line):
    fl = ~solid
    nz, ny, nx = fl.shape
    fetched = 0
    for x0 in range(0, nx, per_line):
        seg = fl[:, :, x0:x0 + per_line]
        fetched += int(seg.any(axis=2).sum()) * seg.shape[2]
    return fetched / fl.sum()


def test_automatic_choice_follows_row_line_waste():
    # whole-plane walls waste no row lines (rows); a sphere packing does (bricks, FP64 only)
    n = 32
    walls = np.zeros((n, n, n), bool)
    walls[:, 0, :] = walls[:, -1, :] = walls[0] = walls[-1] = True
    spheres = _porous_solid(n, 2, frac=0.2)
    for solid, prec in ((walls, lbm.FP64), (spheres, lbm.FP64), (spheres, lbm.FP32_SHIFTED)):
        bricks = prec == lbm.FP64 and _row_line_waste(solid, 16) >= 1.12
        expect = lbm.CELL_ORDER_BRICKS if bricks else lbm.CELL_ORDER_ROWS
        with lbm.Lattice(n, n, n, 0.7, precision=prec, periodic=(True, False, False)) as lat:
            lat.set_flags(~solid)
            assert lat.cell_order == expect
    assert _row_line_waste(walls, 16) == 1.0 and _row_line_waste(spheres, 16) > 1.12


def test_automatic_switch_carries_device_arrays():
    # porosity and the inlet profile are set while the (all-fluid) flags select rows; a packing of
    # small spheres then wastes row lines and set_flags moves everything into bricks
    (nx, ny, nz, steps), _, alpha, _ = _channel_case(5)
    solid = _porous_solid(max(nx, ny, nz), 9, frac=0.2, rmin=1, rmax=1)[:nz, :ny, :nx]
    assert _row_line_waste(solid, 16) >= 1.12
    ora = O.Lattice.channel(nx, ny, nz, 0.62, 0.03, parabolic=True, solid=solid)
    ora.set_force((2e-6, 0.0, 1e-6))
    ora.set_alpha(alpha)
    fo = ora.step(ora.equilibrium(), steps)
    with lbm.Lattice(nx, ny, nz, 0.62, force=(2e-6, 0.0, 1e-6)) as lat:
        lat.set_flags(np.ones(nx * ny * nz, np.uint8))
        assert lat.cell_order == lbm.CELL_ORDER_ROWS
        lat.set_channel(0.03, parabolic=True)
        lat.set_porosity(alpha)
        lat.set_flags(~solid)
        assert lat.cell_order == lbm.CELL_ORDER_BRICKS
        lat.init_equilibrium()
        lat.step(steps)
        f = lat.get_populations()
    assert np.array_equal(f, fo), f"max |diff| = {np.abs(f - fo).max()}"


def test_coupling_ids_follow_the_switch():
    # coupling cells registered under rows; solids then switch the context to bricks
    n, steps, tau = 12, 9, 0.7
    solid = _porous_solid(n, 13, frac=0.15, rmin=1, rmax=2)
    assert _row_line_waste(solid, 16) >= 1.12
    rng = np.random.default_rng(2)
    cells = rng.choice(np.flatnonzero(~solid.reshape(-1)), size=60, replace=False)
    umd = 0.03 * rng.standard_normal((cells.size, 3))
    u0 = 0.02 * rng.standard_normal((n ** 3, 3))
    ora = O.Lattice(n, n, n, periodic=(True, True, True))
    ora.set_solid(solid)
    ora.params.tau = tau
    ora.set_coupling(cells, 0, umd, None, dt=1.0)
    f0 = ora.equilibrium(u=u0)
    fo = ora.step(f0, steps)
    ro, uo = ora.coupling_export(fo, cells)
    with lbm.Lattice(n, n, n, tau, periodic=(True, True, True)) as lat:
        lat.set_flags(np.ones(n ** 3, np.uint8))
        assert lat.cell_order == lbm.CELL_ORDER_ROWS
        lat.set_coupling_cells(cells)
        lat.set_coupling_data(0, umd, None, dt=1.0)
        lat.set_flags(~solid)
        assert lat.cell_order == lbm.CELL_ORDER_BRICKS
        lat.set_populations(f0)
        lat.step(steps)
        f = lat.get_populations()
        r_out, u_out = lat.get_coupling_data()
    assert np.array_equal(f, fo), f"max |diff| = {np.abs(f - fo).max()}"
    assert np.array_equal(r_out, ro) and np.array_equal(u_out, uo)


@pytest.mark.parametrize("prec", [lbm.FP32_SHIFTED, lbm.FP32_SHIFTED_F32])
def test_fp32_orders_and_aa_are_bit_identical(prec):
    n = 18
    solid = _porous_solid(n, 8)
    rng = np.random.default_rng(4)
    u0 = 0.01 * rng.standard_normal((n ** 3, 3))
    out = []
    for order in ORDERS:
        for layout in (lbm.LAYOUT_TWO_BUFFER, lbm.LAYOUT_AA):
            with lbm.Lattice(n, n, n, 0.6, precision=prec, periodic=(True, True, True), force=(1e-5, 0, 0),
                             layout=layout, cell_order=order) as lat:
                lat.set_flags(~solid)
                lat.init_equilibrium(u=u0)
                lat.step(11)
                out.append((lat.get_populations(), lat.kinetic_energy()))
    for f, e in out[1:]:
        assert np.array_equal(f, out[0][0])


def test_adjoint_gradient_is_the_same_in_both_orders():
    n, steps = 12, 20
    solid = np.zeros((n, n, n), bool)
    solid[:, 0, :] = solid[:, -1, :] = True
    alpha = 0.2 * np.random.default_rng(6).random(n ** 3)
    res = {}
    for o in ORDERS:
        with lbm.Lattice(n, n, n, 0.7, periodic=(True, False, True), force=(1e-5, 0, 0), cell_order=o) as lat:
            lat.set_flags(~solid)
            lat.set_porosity(alpha)
            lat.init_equilibrium()
            J, g, _ = lat.adjoint_pressure_drop(steps, 4, 2, 9)
            res[o] = (J, g)
    assert res[ORDERS[0]][0] == res[ORDERS[1]][0]
    assert np.array_equal(res[ORDERS[0]][1], res[ORDERS[1]][1])


def test_unknown_cell_order_is_rejected():
    with pytest.raises(lbm.LbmError, match="cell order"):
        lbm.Lattice(8, 8, 8, 0.7, cell_order=3)
