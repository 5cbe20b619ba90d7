"""lbm_step_graph (32 steps captured into a hipGraph per starting phase, replayed) against plain
lbm_step: bit-identical at every readout, with the program's own eager work and synchronisations
between replays, and across changes that must re-capture the graph (porosity, MD coupling, load
schedule, cell order).  The plain path is pinned to the oracle by the other tests; FP64 equality
with it is checked here, and one case also against the oracle directly."""
import numpy as np
import pytest

import oracle as O
from test_gpu_dense import _porous_solid

pytestmark = pytest.mark.gpu
lbm = pytest.importorskip("madpac_amd.lbm")

SEQ = [37, 4, 64, 33, 1, 70]  # replays plus odd and even remainders: both phases' graphs get used


def _run(make, graph, seq=SEQ):
    out = []
    with make() as lat:
        for k in seq:
            lat.step(k, graph=graph)
            out.append((lat.get_populations(), lat.moments()))  # eager work between replays
            lat.synchronize()
    return out


def _same(a, b):
    for (fa, ma), (fb, mb) in zip(a, b):
        assert np.array_equal(fa, fb), f"max |diff| = {np.abs(fa - fb).max()}"
        assert np.array_equal(ma, mb)


@pytest.mark.parametrize("prec", [lbm.FP64, lbm.FP32_SHIFTED])
def test_graph_replays_equal_plain_steps_cavity(prec):
    make = lambda: lbm.cavity(12, 0.6, 0.05, precision=prec)
    print(f"[gpu] graph vs plain cavity 12^3 precision {prec}", flush=True)
    g = _run(make, True)
    _same(_run(make, False), g)
    if prec == lbm.FP64:
        lat = O.Lattice.cavity(12, 0.6, 0.05)
        assert np.array_equal(g[-1][0], lat.step(lat.equilibrium(), sum(SEQ)))


@pytest.mark.parametrize("layout", [lbm.LAYOUT_TWO_BUFFER, lbm.LAYOUT_AA])
def test_graph_replays_equal_plain_steps_porous_force(layout):
    n = 16
    solid = _porous_solid(n, 23, frac=0.2, rmin=1, rmax=2)

    def make():
        lat = lbm.Lattice(n, n, n, 0.58, periodic=(True, True, True), force=(1e-5, 0, 0), layout=layout)
        lat.set_flags(~solid)
        lat.init_equilibrium(u=0.01 * np.random.default_rng(1).standard_normal((n ** 3, 3)))
        return lat

    print(f"[gpu] graph vs plain porous {n}^3 layout {layout}", flush=True)
    _same(_run(make, False), _run(make, True))


def test_graph_recaptures_after_changes():
    n = 14
    solid = _porous_solid(n, 4, frac=0.2, rmin=1, rmax=2)
    alpha = 0.3 * np.random.default_rng(3).random(n ** 3)
    cells = np.flatnonzero(~solid.reshape(-1))[::17]
    umd = 0.02 * np.random.default_rng(5).standard_normal((cells.size, 3))
    res = []
    for graph in (False, True):
        out = []
        with lbm.Lattice(n, n, n, 0.7, periodic=(True, True, True)) as lat:
            lat.set_flags(~solid)
            lat.init_equilibrium()
            lat.step(36, graph=graph)
            out.append(lat.get_populations())
            lat.set_porosity(alpha)  # a porosity field: another kernel
            lat.step(40, graph=graph)
            out.append(lat.get_populations())
            lat.set_porosity(None)
            lat.set_coupling_cells(cells)  # MD coupling: an extra kernel per step
            lat.set_coupling_data(0, umd, None, dt=1.0)
            lat.step(34, graph=graph)
            out.append(lat.get_populations())
            lat.set_coupling_data(1, umd, None, dt=0.5)  # other mode and dt: other arguments
            lat.step(33, graph=graph)
            out.append(lat.get_populations())
            lat.set_coupling_data(0, None)
            lat.tune(variant=7)  # the other load schedule and workgroup order, 2-row workgroups
            lat.step(35, graph=graph)
            out.append(lat.get_populations())
            lat.set_flags(np.ones(n ** 3, np.uint8))  # all fluid: the other cell order
            lat.init_equilibrium()
            lat.step(64, graph=graph)
            out.append(lat.get_populations())
        res.append(out)
    for a, b in zip(*res):
        assert np.array_equal(a, b), f"max |diff| = {np.abs(a - b).max()}"


def test_graph_step_errors():
    with lbm.Lattice(8, 8, 8, 0.7) as lat:
        lat.set_flags(np.ones(512, np.uint8))
        with pytest.raises(lbm.LbmError, match="state"):
            lat.step(2, graph=True)
        lat.init_equilibrium()
        with pytest.raises(lbm.LbmError, match="nsteps"):
            lat.step(-1, graph=True)


def test_graph_replays_in_place_with_md_coupling():
    # in place, each captured step holds the save and re-collision kernels of the coupled cells
    n = 12
    solid = _porous_solid(n, 8, frac=0.15, rmin=1, rmax=2)
    cells = np.flatnonzero(~solid.reshape(-1))[::11]
    umd = 0.02 * np.random.default_rng(9).standard_normal((cells.size, 3))
    res = []
    for graph in (False, True):
        print(f"[gpu] graph in place with coupling, graph={graph}", flush=True)
        with lbm.Lattice(n, n, n, 0.7, periodic=(True, True, True), layout=lbm.LAYOUT_AA) as lat:
            lat.set_flags(~solid)
            lat.init_equilibrium()
            lat.set_coupling_cells(cells)
            lat.set_coupling_data(0, umd, None, dt=1.0)
            lat.step(37, graph=graph)
            out = [lat.get_populations(), *lat.get_coupling_data()]
            lat.set_coupling_data(1, umd, None, dt=0.5)  # other mode and dt: a re-capture
            lat.step(34, graph=graph)
            out += [lat.get_populations(), *lat.get_coupling_data()]
            res.append(out)
    for a, b in zip(*res):
        assert np.array_equal(a, b)
