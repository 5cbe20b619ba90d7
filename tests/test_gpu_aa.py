"""The in-place AA layout (one state buffer, alternating own-slot and neighbour steps) against
the reference fixtures, the oracle and the two-buffer layout, for odd and even step counts (the
state sits in the NATURAL layout after an even count and in the SWAPPED one after an odd count).
Bars: FP64 bit-identical to the fixtures / oracle; FP32-shifted bit-identical to the two-buffer
FP32-shifted path (same arithmetic, same rounding at the store)."""
import numpy as np
import pytest

import oracle as O
from conftest import channel_tolerance, golden
from test_gpu_dense import _porous_solid

pytestmark = pytest.mark.gpu
lbm = pytest.importorskip("madpac_amd.lbm")
AA = lbm.LAYOUT_AA


@pytest.mark.parametrize("name", ["cavity_n8_t1.npz", "cavity_n8_t10.npz", "cavity_n10_t37.npz",
                                  "cavity_n13_t5.npz", "cavity_n8_t100.npz"])
def test_aa_cavity_bitwise_equals_reference(name):
    d = golden(name)
    n, steps = int(d["n"]), int(d["steps"])
    print(f"[gpu] AA cavity fixture {name}", flush=True)
    with lbm.cavity(n, float(d["tau"]), float(d["ulid"]), layout=AA) as lat:
        lat.step(steps)
        f = lat.get_populations()
    ref = d["f"].reshape(-1)
    assert np.array_equal(f, ref), f"max |diff| = {np.abs(f - ref).max()}"


@pytest.mark.parametrize("name", ["channel_8x8x8_t10_uniform.npz", "channel_10x7x9_t13_parabolic.npz"])
def test_aa_channel_equals_reference(name):
    d = golden(name)
    nx, ny, nz = (int(v) for v in d["dims"])
    steps, par = int(d["steps"]), bool(d["parabolic"])
    print(f"[gpu] AA channel fixture {name}", flush=True)
    with lbm.channel(nx, ny, nz, float(d["tau"]), float(d["umean"]), parabolic=par, layout=AA) as lat:
        lat.step(steps)
        f = lat.get_populations()
    ref = d["f"].reshape(-1)
    tol = channel_tolerance(d)
    assert np.max(np.abs(f - ref)) <= tol * np.max(np.abs(ref))
    ora = O.Lattice.channel(nx, ny, nz, float(d["tau"]), float(d["umean"]), parabolic=par)
    assert np.array_equal(f, ora.step(ora.equilibrium(), steps))


@pytest.mark.parametrize("order", [lbm.CELL_ORDER_ROWS, lbm.CELL_ORDER_BRICKS])
@pytest.mark.parametrize("steps", [1, 6, 13])
def test_aa_porous_force_brinkman_bitwise_equals_oracle(steps, order):
    # obstacles, a wall pair, a periodic axis, Guo force and Brinkman porosity in one run; solid
    # lanes store zeros into their dead slots at surfaces in both steps and both cell orders
    n = 14
    solid = _porous_solid(n, 31, frac=0.12)
    solid[:, 0, :] = solid[:, -1, :] = True
    rng = np.random.default_rng(steps)
    alpha = 0.3 * rng.random(n ** 3)
    g = (1e-5, 0.0, 2e-6)
    ora = O.Lattice(n, n, n, periodic=(True, False, True))
    ora.set_solid(solid)
    ora.set_force(g)
    ora.set_alpha(alpha)
    ora.params.tau = 0.6
    f0 = ora.equilibrium(u=0.01 * rng.standard_normal((n ** 3, 3)))
    fo = ora.step(f0, steps)
    print(f"[gpu] AA porous+force+Brinkman n={n} steps={steps} order={order}", flush=True)
    with lbm.Lattice(n, n, n, 0.6, periodic=(True, False, True), force=g, layout=AA, cell_order=order) as lat:
        lat.set_flags(~solid)
        assert lat.cell_order == order
        lat.set_porosity(alpha)
        lat.set_populations(f0)
        lat.step(steps)
        fg = lat.get_populations()
        m = lat.moments()
        J = lat.pressure_drop(2, 11)
    assert np.array_equal(fg, fo), f"max |diff| = {np.abs(fg - fo).max()}"
    mo = ora.moments(fo)
    assert abs(m[0] - mo[0]) <= 1e-13 * mo[0]
    assert abs(J - ora.pressure_drop(fo, 2, 11)) <= 1e-13


@pytest.mark.parametrize("prec", [lbm.FP32_SHIFTED, lbm.FP32_SHIFTED_F32])
def test_aa_fp32_bit_identical_to_two_buffer(prec):
    nx, ny, nz = 20, 12, 10
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    solid = (x - 7) ** 2 + (y - 6) ** 2 + (z - 5) ** 2 <= 9
    out = {}
    for layout in (lbm.LAYOUT_TWO_BUFFER, AA):
        with lbm.channel(nx, ny, nz, 0.62, 0.04, parabolic=True, solid=solid, precision=prec, layout=layout) as lat:
            res = []
            for steps in (7, 4):  # odd, then back to the NATURAL layout
                lat.step(steps)
                res.append((lat.get_populations(), *lat.fields(), lat.kinetic_energy()))
            out[layout] = res
    for a, b in zip(out[lbm.LAYOUT_TWO_BUFFER], out[AA]):
        for x_, y_ in zip(a, b):
            assert np.array_equal(np.asarray(x_), np.asarray(y_))


def test_aa_set_get_round_trip_across_parities():
    n = 10
    solid = _porous_solid(n, 5)
    rng = np.random.default_rng(1)
    f = (O.W[:, None] * (1.0 + 0.05 * rng.standard_normal((19, n ** 3))))
    f[:, solid.reshape(-1)] = 0.0
    f = f.reshape(-1)
    with lbm.Lattice(n, n, n, 0.7, periodic=(True, True, True), layout=AA) as lat:
        lat.set_flags(~solid)
        lat.set_populations(f)
        assert np.array_equal(lat.get_populations(), f)
        lat.step(3)  # SWAPPED
        g = lat.get_populations()
        lat.set_populations(f)  # back to NATURAL
        lat.step(3)
        assert np.array_equal(lat.get_populations(), g)


@pytest.mark.parametrize("prec", [lbm.FP64, lbm.FP32_SHIFTED])
def test_aa_adjoint_bitwise_equals_two_buffer(prec):
    # the in-place context lends itself a second buffer for the adjoint: J, the gradient and the final
    # state equal the two-buffer run; it starts from the SWAPPED parity (an odd step count before)
    nx, ny, nz, tau, g = 16, 9, 8, 0.7, (1e-4, 0.0, 0.0)
    solid = np.zeros((nz, ny, nx), bool)
    solid[0], solid[-1], solid[:, 0], solid[:, -1] = True, True, True, True
    alpha = np.clip(np.random.default_rng(6).random(nx * ny * nz) * 0.5, 0, 1)
    out = {}
    for layout in (lbm.LAYOUT_TWO_BUFFER, AA):
        print(f"[gpu] AA adjoint prec {prec} layout {layout}", flush=True)
        with lbm.Lattice(nx, ny, nz, tau, precision=prec, periodic=(True, False, False), force=g, layout=layout) as lat:
            lat.set_flags(~solid)
            lat.set_porosity(alpha)
            lat.init_equilibrium()
            lat.step(3)
            J, grad, _ = lat.adjoint_pressure_drop(10, 3, 2, 13)
            f_end = lat.get_populations()
            lat.step(5)  # the context steps on from the final state in its own layout
            out[layout] = (J, grad, f_end, lat.get_populations())
    a, b = out[lbm.LAYOUT_TWO_BUFFER], out[AA]
    assert a[0] == b[0]
    for x_, y_ in zip(a[1:], b[1:]):
        assert np.array_equal(x_, y_)


def test_aa_adjoint_argument_errors_leave_the_state_alone():
    n = 8
    with lbm.cavity(n, 0.7, 0.05, layout=AA) as lat:
        lat.step(3)
        before = lat.get_populations()
        with pytest.raises(lbm.LbmError, match="plane"):
            lat.adjoint_pressure_drop(4, 2, 1, 60)
        assert np.array_equal(lat.get_populations(), before)
        assert lat.tune(variant=0) == 0


@pytest.mark.parametrize("prec", [lbm.FP64, lbm.FP32_SHIFTED])
@pytest.mark.parametrize("order", [lbm.CELL_ORDER_ROWS, lbm.CELL_ORDER_BRICKS])
@pytest.mark.parametrize("mode,density", [(0, True), (1, False)])
def test_aa_md_coupling_bitwise_equals_two_buffer(prec, order, mode, density):
    # the in-place step saves the coupled cells' populations before it and re-collides them into the
    # slots it wrote; states and the MD export equal the two-buffer run at odd and even step counts
    n = 12
    rng = np.random.default_rng(40 + mode)
    cells = rng.choice(n ** 3, size=n ** 3 // 8, replace=False)
    umd = 0.03 * rng.standard_normal((cells.size, 3))
    rmd = 1.0 + 0.01 * rng.standard_normal(cells.size)
    rho0 = 1.0 + 0.01 * rng.standard_normal(n ** 3)
    u0 = 0.02 * rng.standard_normal((n ** 3, 3))
    out = {}
    for layout in (lbm.LAYOUT_TWO_BUFFER, AA):
        print(f"[gpu] AA coupling prec {prec} order {order} mode {mode} layout {layout}", flush=True)
        with lbm.Lattice(n, n, n, 0.7, precision=prec, layout=layout, cell_order=order) as lat:
            lat.set_flags(np.ones(n ** 3, np.uint8))
            lat.set_cavity_lid(0.04)
            lat.init_equilibrium(rho=rho0, u=u0)
            lat.set_coupling_cells(cells)
            lat.set_coupling_data(mode, umd, rmd if density else None, dt=0.8)
            res = []
            for steps in (7, 6):
                lat.step(steps)
                res.append((lat.get_populations(), *lat.get_coupling_data()))
            out[layout] = res
    for a, b in zip(out[lbm.LAYOUT_TWO_BUFFER], out[AA]):
        for x_, y_ in zip(a, b):
            assert np.array_equal(np.asarray(x_), np.asarray(y_))


def test_aa_tune_restores_the_state_and_changes_no_bits():
    n = 16
    solid = _porous_solid(n, 12)
    out = []
    for tuned in (False, True):
        with lbm.Lattice(n, n, n, 0.6, periodic=(True, True, True), force=(1e-5, 0, 0), layout=AA) as lat:
            lat.set_flags(~solid)
            lat.init_equilibrium()
            lat.step(3)  # SWAPPED parity
            before = lat.get_populations()
            if tuned:
                v = lat.tune(4)
                assert v & 1 == 0 and 4 <= v < 16  # one load schedule in place; any order and shape
                assert np.array_equal(lat.get_populations(), before)
            lat.step(9)
            out.append(lat.get_populations())
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("variant", [4, 10, 14])
def test_aa_each_workgroup_shape_bitwise_equals_two_buffer(variant):
    # 2, 4 and 8 rows or bricks per workgroup (both orders by size), in place against two buffers
    nx, ny, nz, steps = 22, 13, 10, 9
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    solid = (x - 8) ** 2 + (y - 6) ** 2 + (z - 5) ** 2 <= 12
    out = []
    for layout, order in ((lbm.LAYOUT_TWO_BUFFER, lbm.CELL_ORDER_ROWS), (AA, lbm.CELL_ORDER_ROWS),
                          (AA, lbm.CELL_ORDER_BRICKS)):
        print(f"[gpu] AA shape variant {variant} layout {layout} order {order}", flush=True)
        with lbm.Lattice(nx, ny, nz, 0.6, periodic=(True, True, True), force=(1e-5, 0, 0), layout=layout,
                         cell_order=order) as lat:
            lat.set_flags(~solid)
            lat.init_equilibrium()
            lat.tune(variant=variant)
            lat.step(steps)
            out.append((lat.get_populations(), lat.moments()))
    for f, _ in out[1:]:
        assert np.array_equal(f, out[0][0])
    assert np.array_equal(out[1][1], out[0][1])  # same cell order: the reductions sum alike
