"""GPU parity of the MD <-> LB coupling hooks (madpac/peanocoupling/PeanoDataExchangeService.cpp)
through the C-ABI, against the oracle's restatement.  Bars: FP64 bit-identical to the oracle;
FP32-shifted contexts (coupled cells re-collided in FP64) at u L2 <= 1e-5; the export equals the
oracle's getAndSetV export (after its GET block) bit for bit, and the plain fields once the coupling
is off."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

lbm = pytest.importorskip("madpac_amd.lbm")


def _setup(n, seed):
    rng = np.random.default_rng(seed)
    cells = rng.choice(n ** 3, size=n ** 3 // 10, replace=False)
    umd = 0.03 * rng.standard_normal((cells.size, 3))
    rmd = 1.0 + 0.01 * rng.standard_normal(cells.size)
    rho0 = 1.0 + 0.01 * rng.standard_normal(n ** 3)
    u0 = 0.02 * rng.standard_normal((n ** 3, 3))
    return cells, umd, rmd, rho0, u0


@pytest.mark.parametrize("mode,density", [(0, False), (0, True), (1, False), (1, True)])
def test_fp64_coupled_cavity_equals_oracle(mode, density):
    n, steps, tau, dt = 12, 25, 0.7, 0.8
    cells, umd, rmd, rho0, u0 = _setup(n, 7 + mode)
    print(f"[gpu] coupling mode {mode} density {density}: {cells.size} cells, cavity {n}^3 x {steps}", flush=True)
    ora = O.Lattice.cavity(n, tau, 0.04)
    ora.set_coupling(cells, mode, umd, rmd if density else None, dt=dt)
    f0 = ora.equilibrium(rho=rho0, u=u0)
    fo = ora.step(f0, steps)
    with lbm.cavity(n, tau, 0.04) as lat:
        lat.set_populations(f0)
        lat.set_coupling_cells(cells)
        lat.set_coupling_data(mode, umd, rmd if density else None, dt=dt)
        lat.step(steps)
        f = lat.get_populations()
        r_out, u_out = lat.get_coupling_data()
        lat.set_coupling_data(mode, None)
        r_off, u_off = lat.get_coupling_data()
    assert np.array_equal(f, fo), f"max |diff| = {np.abs(f - fo).max()}"
    # the export is what getAndSetV sends after its GET block: MD-modified while coupling is on
    ro, uo = ora.coupling_export(fo, cells)
    assert np.array_equal(r_out, ro) and np.array_equal(u_out, uo)
    rf, uf = ora.fields(fo)
    assert np.array_equal(r_off, rf[cells]) and np.array_equal(u_off, uf[cells])


@pytest.mark.parametrize("mode", [0, 1])
def test_fp32_coupled_periodic_box_tracks_oracle(mode):
    n, steps = 16, 60
    cells, umd, _, rho0, u0 = _setup(n, 21)
    ora = O.Lattice(n, n, n, periodic=(True, True, True))
    ora.set_solid(np.zeros(n ** 3, np.uint8))
    ora.params.tau = 0.65
    ora.set_coupling(cells, mode, umd, None, dt=1.0)
    fo = ora.step(ora.equilibrium(rho=rho0, u=u0), steps)
    _, uo = ora.fields(fo)
    with lbm.Lattice(n, n, n, 0.65, precision=lbm.FP32_SHIFTED, periodic=(True, True, True)) as lat:
        lat.set_flags(np.ones(n ** 3, np.uint8))
        lat.init_equilibrium(rho=rho0, u=u0)
        lat.set_coupling_cells(cells)
        lat.set_coupling_data(mode, umd, None, dt=1.0)
        lat.step(steps)
        _, u = lat.fields()
    l2 = np.linalg.norm(u - uo) / np.linalg.norm(uo)
    print(f"[gpu] FP32-shifted coupling mode {mode}: u L2 {l2:.3e}", flush=True)
    assert l2 <= 1e-5


def test_coupling_off_restores_plain_steps_and_errors():
    n = 8
    with lbm.cavity(n, 0.7, 0.03) as lat:
        f0 = lat.get_populations()
        lat.set_coupling_cells([1, 2, 3])
        lat.set_coupling_data(0, np.zeros(9))
        lat.set_coupling_data(0, None)          # switched off: plain steps
        lat.step(3)
        f_off = lat.get_populations()
        with pytest.raises(lbm.LbmError, match="distinct"):
            lat.set_coupling_cells([4, 4])
        with pytest.raises(lbm.LbmError, match="out of range"):
            lat.set_coupling_cells([n ** 3])
    with lbm.cavity(n, 0.7, 0.03) as ref:
        ref.step(3)
        assert np.array_equal(f_off, ref.get_populations())
    with lbm.Lattice(n, n, n, 0.7, periodic=(True, True, True), force=(1e-5, 0, 0)) as lat:
        lat.set_flags(np.ones(n ** 3, np.uint8))
        lat.init_equilibrium()
        lat.set_coupling_cells([5])
        with pytest.raises(lbm.LbmError, match="body force"):
            lat.set_coupling_data(0, np.zeros(3))
    with lbm.cavity(n, 0.7, 0.03) as lat:
        lat.set_coupling_cells([5])
        lat.set_coupling_data(1, np.zeros(3), dt=1.0)
        with pytest.raises(lbm.LbmError, match="MD-coupled"):
            lat.adjoint_pressure_drop(2, 1, 1, 6)
