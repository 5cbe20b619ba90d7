"""GPU parity of the dense D3Q19 path (through the C-ABI) against the oracle and the reference fixtures.

Bars: FP64 is bit-identical to the reference (fixtures) and to the oracle (same seeded inputs);
FP32-shifted is checked against the FP64 oracle with the L2 / Linf tolerances written in each test.
"""
import hashlib
import os

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN, golden

pytestmark = pytest.mark.gpu

lbm = pytest.importorskip("madpac_amd.lbm")


def _porous_solid(n, seed, frac=0.15, rmin=2, rmax=4):
    """Random periodic spheres until the solid fraction reaches `frac` (small-size variant of the
    porous-channel workload)."""
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    solid = np.zeros((n, n, n), bool)
    while solid.mean() < frac:
        c = rng.integers(0, n, 3)
        r = rng.integers(rmin, rmax + 1)
        dx = np.minimum(np.abs(x - c[0]), n - np.abs(x - c[0]))
        dy = np.minimum(np.abs(y - c[1]), n - np.abs(y - c[1]))
        dz = np.minimum(np.abs(z - c[2]), n - np.abs(z - c[2]))
        solid |= dx * dx + dy * dy + dz * dz <= r * r
    return solid


CAVITY_FIXTURES = ["cavity_n8_t1.npz", "cavity_n8_t10.npz", "cavity_n8_t100.npz",
                   "cavity_n10_t37.npz", "cavity_n13_t5.npz"]


@pytest.mark.parametrize("name", CAVITY_FIXTURES)
def test_fp64_cavity_bitwise_equals_reference(name):
    d = golden(name)
    n, steps = int(d["n"]), int(d["steps"])
    print(f"[gpu] cavity fixture {name} n={n} steps={steps}", flush=True)
    with lbm.cavity(n, float(d["tau"]), float(d["ulid"])) as lat:
        lat.step(steps)
        f = lat.get_populations()
    ref = d["f"].reshape(-1)
    assert np.array_equal(f, ref), f"max |diff| = {np.abs(f - ref).max()}"


def test_fp64_cavity64_1000_steps_equals_reference():
    path = os.path.join(GOLDEN, "cavity_n64_t1000.npz")
    if not os.path.exists(path):
        pytest.skip("64^3 fixture not generated")
    d = golden("cavity_n64_t1000.npz")
    print("[gpu] cavity 64^3 x 1000 steps, FP64", flush=True)
    with lbm.cavity(64, float(d["tau"]), float(d["ulid"])) as lat:
        lat.step(1000)
        f = lat.get_populations()
        rho, u = lat.fields()
    assert hashlib.sha256(f.astype("<f8").tobytes()).hexdigest() == str(d["sha256"])
    h = 32
    rho_p = rho.reshape(64, 64, 64)[:, h, :]
    u_p = np.moveaxis(u.reshape(64, 64, 64, 3), -1, 0)[:, :, h, :]
    # the stated bar is <= 1e-12 relative (the hash above already shows the state is bit-identical;
    # the fixture's planes are numpy moments of the reference state, in numpy's summation order)
    assert np.max(np.abs(rho_p - d["rho_mid_y"]) / np.abs(d["rho_mid_y"])) <= 1e-12
    assert np.max(np.abs(u_p - d["u_mid_y"])) <= 1e-12 * np.max(np.abs(d["u_mid_y"]))


@pytest.mark.parametrize("prec", [lbm.FP64, lbm.FP32_SHIFTED])
def test_populations_round_trip_is_exact_permutation(prec):
    n = 12
    solid = _porous_solid(n, 3)
    rng = np.random.default_rng(0)
    f = (O.W[:, None] * (1.0 + 0.1 * rng.standard_normal((19, n ** 3)))).reshape(-1)
    f.reshape(19, -1)[:, solid.reshape(-1)] = 0.0
    with lbm.Lattice(n, n, n, 0.7, precision=prec, periodic=(True, True, True)) as lat:
        lat.set_flags(~solid)
        lat.set_populations(f)
        g = lat.get_populations()
    if prec == lbm.FP64:
        assert np.array_equal(g, f)
    else:
        assert np.max(np.abs(g - f)) <= 1e-8  # f - w stored in FP32 (|f - w| < 0.02)


def test_fp64_porous_periodic_force_bitwise_equals_oracle():
    n, steps = 16, 40
    solid = _porous_solid(n, 11)
    g = (1e-5, 2e-6, 0.0)
    rng = np.random.default_rng(4)
    u0 = 0.01 * rng.standard_normal((n ** 3, 3))
    print(f"[gpu] porous periodic n={n} steps={steps} FP64 with Guo force", flush=True)
    ora = O.Lattice(n, n, n, periodic=(True, True, True))
    ora.set_solid(solid)
    ora.set_force(g)
    ora.params.tau = 0.55
    f0 = ora.equilibrium(u=u0)
    fo = ora.step(f0, steps)
    with lbm.Lattice(n, n, n, 0.55, periodic=(True, True, True), force=g) as lat:
        lat.set_flags(~solid)
        lat.set_populations(f0)
        lat.step(steps)
        fg = lat.get_populations()
    assert np.array_equal(fg, fo), f"max |diff| = {np.abs(fg - fo).max()}"


@pytest.mark.parametrize("order", [lbm.CELL_ORDER_ROWS, lbm.CELL_ORDER_BRICKS])
def test_tuned_variants_are_bit_identical(order):
    # load schedule x workgroup order x workgroup shape (2, 4, 8 rows or bricks) on a lattice whose
    # edges no shape divides; the reductions keep one shape, so the moments agree bit for bit too
    nx, ny, nz, steps = 21, 13, 11, 12
    solid = _porous_solid(16, 21)[:nz, :ny, :16]
    solid = np.concatenate([solid, solid[:, :, :5]], axis=2)
    out = []
    for v in (0, 1, 2, 3, 4, 7, 9, 14, 15, None):
        with lbm.Lattice(nx, ny, nz, 0.6, periodic=(True, True, False), force=(1e-5, 0, 0), cell_order=order) as lat:
            lat.set_flags(~solid)
            lat.set_cavity_lid(0.03)
            lat.init_equilibrium()
            chosen = lat.tune(variant=v) if v is not None else lat.tune(3)  # None: timed choice
            assert chosen == v if v is not None else 4 <= chosen < 16
            lat.step(steps)
            out.append((lat.get_populations(), lat.moments(), lat.kinetic_energy()))
    for other in out[1:]:
        assert np.array_equal(out[0][0], other[0])
        assert np.array_equal(out[0][1], other[1])
        assert out[0][2] == other[2]


def test_fp64_brinkman_bitwise_equals_oracle():
    nx, ny, nz, steps = 12, 8, 8, 25
    solid = np.zeros((nz, ny, nx), bool)
    solid[0], solid[-1], solid[:, 0], solid[:, -1] = True, True, True, True
    rng = np.random.default_rng(8)
    alpha = rng.random(nx * ny * nz)
    ora = O.Lattice(nx, ny, nz, periodic=(True, False, False))
    ora.set_solid(solid)
    ora.set_force((1e-4, 0, 0))
    ora.params.tau = 0.8
    ora.set_alpha(alpha)
    f0 = ora.equilibrium()
    fo = ora.step(f0, steps)
    with lbm.Lattice(nx, ny, nz, 0.8, periodic=(True, False, False), force=(1e-4, 0, 0)) as lat:
        lat.set_flags(~solid)
        lat.set_porosity(alpha)
        lat.init_equilibrium()
        lat.step(steps)
        fg = lat.get_populations()
    assert np.array_equal(fg, fo), f"max |diff| = {np.abs(fg - fo).max()}"


def test_moments_and_pressure_drop_match_oracle():
    n, steps = 20, 30
    solid = _porous_solid(n, 5)
    ora = O.Lattice(n, n, n, periodic=(True, True, True))
    ora.set_solid(solid)
    ora.set_force((1e-5, 0, 0))
    ora.params.tau = 0.6
    rng = np.random.default_rng(9)
    f0 = ora.equilibrium(u=0.02 * rng.standard_normal((n ** 3, 3)))
    fo = ora.step(f0, steps)
    mo = ora.moments(fo)
    Jo = ora.pressure_drop(fo, 4, 15)
    for prec, tol in ((lbm.FP64, 1e-13), (lbm.FP32_SHIFTED, 2e-6)):
        with lbm.Lattice(n, n, n, 0.6, precision=prec, periodic=(True, True, True), force=(1e-5, 0, 0)) as lat:
            lat.set_flags(~solid)
            lat.set_populations(f0)
            lat.step(steps)
            m = lat.moments()
            J = lat.pressure_drop(4, 15)
            assert lat.fluid_cells == int((~solid).sum())
        assert abs(m[0] - mo[0]) <= tol * mo[0]
        scale = np.abs(mo[1:]).max()
        assert np.max(np.abs(m[1:] - mo[1:])) <= max(tol, 1e-12) * max(scale, 1.0) * 10
        assert abs(J - Jo) <= (1e-13 if prec == lbm.FP64 else 1e-7)


def test_kinetic_energy_matches_oracle_taylor_green():
    wl = pytest.importorskip("madpac_amd.workloads")
    n, steps, tau, u0 = 16, 40, 0.6, 0.03
    rho0, u0f = wl.taylor_green(n, u0)
    ora = O.Lattice(n, n, n, periodic=(True, True, True))
    ora.set_solid(np.zeros((n, n, n), np.uint8))
    ora.params.tau = tau
    fo = ora.step(ora.equilibrium(rho=rho0, u=u0f), steps)
    rho, u = ora.fields(fo)
    ke_o = float(np.sum(0.5 * rho * np.sum(u * u, axis=1)))
    for prec, tol in ((lbm.FP64, 1e-12), (lbm.FP32_SHIFTED, 1e-5)):
        with lbm.Lattice(n, n, n, tau, precision=prec, periodic=(True, True, True)) as lat:
            lat.set_flags(np.ones(n ** 3, np.uint8))
            lat.init_equilibrium(rho=rho0, u=u0f)
            lat.step(steps)
            ke = lat.kinetic_energy()
        assert abs(ke - ke_o) <= tol * ke_o, (prec, ke, ke_o)


@pytest.mark.parametrize("prec", [lbm.FP32_SHIFTED, lbm.FP32_SHIFTED_F32])
def test_fp32_shifted_cavity_error_vs_fp64_oracle(prec):
    # FP32 shifted populations against the FP64 oracle: velocity L2 / Linf relative errors
    n, steps = 24, 600
    lat64 = O.Lattice.cavity(n, 0.6, 0.05)
    fo = lat64.step(lat64.equilibrium(), steps, nthreads=os.cpu_count() or 1)
    rho_o, u_o = lat64.fields(fo)
    with lbm.cavity(n, 0.6, 0.05, precision=prec) as lat:
        lat.step(steps)
        rho, u = lat.fields()
    l2 = np.linalg.norm(u - u_o) / np.linalg.norm(u_o)
    linf = np.max(np.abs(u - u_o)) / np.max(np.abs(u_o))
    print(f"[gpu] FP32-shifted cavity {n}^3 x {steps}: u L2 {l2:.3e} Linf {linf:.3e}", flush=True)
    assert l2 <= 1e-5 and linf <= 1e-4
    assert np.max(np.abs(rho - rho_o)) <= 1e-6


def test_fp32_shifted_has_no_momentum_drift_in_forced_porous_flow():
    # Low-Mach forced flow in a porous box: FP32 collision arithmetic repeats the same rounding of
    # the moments and equilibrium every step and drifts the momentum (~1e-5 relative after 500
    # steps here); the FP32-shifted collision keeps them and the relaxation product in FP64
    # (measured: u L2 2.1e-7, momentum offset 8.8e-8).  Bars: u L2 <= 5e-7, |offset| <= 2e-7.
    n, steps, tau, g = 24, 500, 0.55, (1e-6, 0.0, 0.0)
    solid = _porous_solid(n, 17, frac=0.2, rmin=2, rmax=4)
    ora = O.Lattice(n, n, n, periodic=(True, True, True))
    ora.set_solid(solid)
    ora.set_force(g)
    ora.params.tau = tau
    fo = ora.step(ora.equilibrium(), steps, nthreads=os.cpu_count() or 1)
    _, u_o = ora.fields(fo)
    fl = ~solid.reshape(-1)
    out = {}
    for prec in (lbm.FP32_SHIFTED, lbm.FP32_SHIFTED_F32):
        with lbm.Lattice(n, n, n, tau, precision=prec, periodic=(True, True, True), force=g) as lat:
            lat.set_flags(~solid)
            lat.init_equilibrium()
            lat.step(steps)
            _, u = lat.fields()
        out[prec] = (np.linalg.norm(u[fl] - u_o[fl]) / np.linalg.norm(u_o[fl]),
                     (u[fl, 0].sum() - u_o[fl, 0].sum()) / u_o[fl, 0].sum())
    print(f"[gpu] forced porous {n}^3 x {steps}: FP64-arith u L2 {out[lbm.FP32_SHIFTED][0]:.2e} "
          f"drift {out[lbm.FP32_SHIFTED][1]:.2e}; FP32-arith u L2 {out[lbm.FP32_SHIFTED_F32][0]:.2e} "
          f"drift {out[lbm.FP32_SHIFTED_F32][1]:.2e}", flush=True)
    assert out[lbm.FP32_SHIFTED][0] <= 5e-7 and abs(out[lbm.FP32_SHIFTED][1]) <= 2e-7


def test_error_paths_return_codes():
    with lbm.Lattice(8, 8, 8, 0.7) as lat:
        with pytest.raises(lbm.LbmError, match="flags"):
            lat.step(1)
        lat.set_flags(np.ones(512, np.uint8))
        with pytest.raises(lbm.LbmError, match="state"):
            lat.step(1)
        with pytest.raises(lbm.LbmError, match="class"):
            lat.set_moving_wall_class(32, (0.1, 0, 0), 0xFF)
        with pytest.raises(lbm.LbmError, match="porosity"):
            lat.set_porosity(np.full(512, 2.0))
        lat.init_equilibrium()
        with pytest.raises(lbm.LbmError, match="range"):
            lat.pressure_drop(0, 8)


def test_nan_is_detected_by_reduction_guard():
    n = 8
    f = np.tile(O.W[:, None], (1, n ** 3)).reshape(-1).copy()
    f[5 * n ** 3 + 17] = np.nan
    with lbm.Lattice(n, n, n, 0.7, periodic=(True, True, True)) as lat:
        lat.set_flags(np.ones(n ** 3, np.uint8))
        lat.set_populations(f)
        with pytest.raises(lbm.LbmError, match="non-finite"):
            lat.moments()


def test_failed_allocation_leaves_no_error_behind():
    # a 2048^3 lattice (the largest extent accepted) needs 1.3 TB per FP64 buffer: lbm_create's
    # device allocation fails cleanly, and the failure must not surface later as the launch error
    # of an unrelated context
    with pytest.raises(lbm.LbmError, match="allocation"):
        lbm.Lattice(2048, 2048, 2048, 0.7)
    with lbm.cavity(8, 0.6, 0.05) as lat:
        lat.step(7)
        f = lat.get_populations()
    ora = O.Lattice.cavity(8, 0.6, 0.05)
    assert np.array_equal(f, ora.step(ora.equilibrium(), 7))
