"""GPU parity of the boundary conditions beyond bounce-back and the lid (through the C-ABI).

The reference's default channel (velocity inlet, pressure outlet, no-slip walls; uniform or
parabolic inlet) is compared with reference fixtures and with the oracle; the pressure inlet,
which no reference scenario on this path uses, with the oracle only.  Bars: FP64 bit-identical
(the 10x7x9 parabolic fixture at 1e-13 relative, see conftest.channel_tolerance); FP32-shifted
against the FP64 oracle at the L2 / Linf bars written in the test.
"""
import numpy as np
import pytest

import oracle as O
from conftest import channel_tolerance, golden

pytestmark = pytest.mark.gpu

lbm = pytest.importorskip("madpac_amd.lbm")

CHANNEL_FIXTURES = ["channel_8x8x8_t10_uniform.npz", "channel_16x8x8_t200_parabolic.npz",
                    "channel_10x7x9_t13_parabolic.npz"]


@pytest.mark.parametrize("name", CHANNEL_FIXTURES)
def test_fp64_channel_equals_reference(name):
    d = golden(name)
    nx, ny, nz = (int(v) for v in d["dims"])
    steps, par = int(d["steps"]), bool(d["parabolic"])
    print(f"[gpu] channel fixture {name}", flush=True)
    with lbm.channel(nx, ny, nz, float(d["tau"]), float(d["umean"]), parabolic=par) as lat:
        lat.step(steps)
        f = lat.get_populations()
    ref = d["f"].reshape(-1)
    tol = channel_tolerance(d)
    if tol == 0.0:
        assert np.array_equal(f, ref), f"max |diff| = {np.abs(f - ref).max()}"
    else:
        assert np.max(np.abs(f - ref)) <= tol * np.max(np.abs(ref))
    # the device evaluates the profile exactly as the oracle does
    ora = O.Lattice.channel(nx, ny, nz, float(d["tau"]), float(d["umean"]), parabolic=par)
    assert np.array_equal(f, ora.step(ora.equilibrium(), steps))


def _sphere(nx, ny, nz, c, r):
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2 <= r * r


@pytest.mark.parametrize("prec", [lbm.FP64, lbm.FP32_SHIFTED])
def test_channel_with_obstacle_and_force_matches_oracle(prec):
    nx, ny, nz, steps = 24, 10, 12, 60
    solid = _sphere(nx, ny, nz, (8, 5, 6), 3)
    g = (2e-6, 0.0, 1e-6)
    print(f"[gpu] channel {nx}x{ny}x{nz} obstacle + force, precision {prec}", flush=True)
    ora = O.Lattice.channel(nx, ny, nz, 0.65, 0.03, parabolic=True, solid=solid)
    ora.set_force(g)
    fo = ora.step(ora.equilibrium(), steps)
    with lbm.Lattice(nx, ny, nz, 0.65, precision=prec, force=g) as lat:
        lat.set_flags(~solid)
        lat.set_channel(0.03, parabolic=True)
        lat.init_equilibrium()
        lat.step(steps)
        f = lat.get_populations()
        rho, u = lat.fields()
    if prec == lbm.FP64:
        assert np.array_equal(f, fo), f"max |diff| = {np.abs(f - fo).max()}"
    else:
        rho_o, u_o = ora.fields(fo)
        fl = ~solid.reshape(-1)
        l2 = np.linalg.norm(u[fl] - u_o[fl]) / np.linalg.norm(u_o[fl])
        linf = np.max(np.abs(u[fl] - u_o[fl])) / np.max(np.abs(u_o[fl]))
        print(f"[gpu] FP32-shifted channel: u L2 {l2:.3e} Linf {linf:.3e}", flush=True)
        assert l2 <= 1e-5 and linf <= 1e-4
        assert np.max(np.abs(rho[fl] - rho_o[fl])) <= 1e-6


def _oracle_pressure_channel(nx, ny, nz, tau, u_in, rho_out):
    """Channel classified as the reference does, with the inlet switched to a pressure inlet
    (u_L given, density from the cell) and the outlet density set to rho_out."""
    ora = O.Lattice.channel(nx, ny, nz, tau, 0.0)
    classes = {}
    for k in range(1, O.MAX_CLASSES):
        m = int(ora.params.moving_mask[k])
        if not m:
            continue
        t = ora.params.bc_type[k]
        if t & 15 == O.BC_VELOCITY_INLET:
            ora.params.bc_type[k] = O.BC_PRESSURE_INLET
            for d in range(3):
                ora.params.uw[k][d] = u_in[d]
        else:
            ora.params.bc_rho[k] = rho_out
        classes[k] = (ora.params.bc_type[k], ora.params.bc_rho[k], tuple(ora.params.uw[k]), m)
    cls = ((ora.words >> O.CLASS_SHIFT) & 31).astype(np.uint8)
    return ora, classes, cls


@pytest.mark.parametrize("prec", [lbm.FP64, lbm.FP32_SHIFTED])
def test_pressure_inlet_and_outlet_classes_match_oracle(prec):
    nx, ny, nz, steps, tau = 20, 9, 8, 80, 0.7
    ora, classes, cls = _oracle_pressure_channel(nx, ny, nz, tau, (0.025, 0.0, 0.002), 0.995)
    fo = ora.step(ora.equilibrium(), steps)
    print(f"[gpu] pressure channel {nx}x{ny}x{nz}, {len(classes)} classes, precision {prec}", flush=True)
    with lbm.Lattice(nx, ny, nz, tau, precision=prec) as lat:
        lat.set_flags(np.ones(lat.n, np.uint8))
        for k, (t, rho, u, m) in classes.items():
            lat.set_boundary_class(k, t, rho, u, m)
        lat.set_wall_classes(cls)
        lat.init_equilibrium()
        lat.step(steps)
        f = lat.get_populations()
        rho, u = lat.fields()
    if prec == lbm.FP64:
        assert np.array_equal(f, fo), f"max |diff| = {np.abs(f - fo).max()}"
    else:
        rho_o, u_o = ora.fields(fo)
        l2 = np.linalg.norm(u - u_o) / np.linalg.norm(u_o)
        print(f"[gpu] FP32-shifted pressure channel: u L2 {l2:.3e}", flush=True)
        assert l2 <= 1e-5
        assert np.max(np.abs(rho - rho_o)) <= 1e-6


def test_boundary_errors():
    with lbm.Lattice(8, 8, 8, 0.7, periodic=(True, False, False)) as lat:
        with pytest.raises(lbm.LbmError, match="non-periodic"):
            lat.set_channel(0.01)
        with pytest.raises(lbm.LbmError, match="1..31"):
            lat.set_boundary_class(32, lbm.BC_PRESSURE_OUTLET, 1.0, (0, 0, 0), 0x1)
        with pytest.raises(lbm.LbmError, match="unknown boundary type"):
            lat.set_boundary_class(2, 7, 1.0, (0, 0, 0), 0x1)
        with pytest.raises(lbm.LbmError, match="inlet classes only"):
            lat.set_boundary_class(2, lbm.BC_PRESSURE_OUTLET | lbm.BC_CELL_VELOCITY, 1.0, (0, 0, 0), 0x1)
    with lbm.channel(8, 6, 6, 0.7, 0.02) as lat:
        with pytest.raises(lbm.LbmError, match="moving walls only"):
            lat.adjoint_pressure_drop(4, 2, 1, 6)
