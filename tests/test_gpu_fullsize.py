"""Full-size runs of the benchmark configurations on one GPU, checked through properties that do
not need a CPU solve at that size.

The FP64 device path equals the CPU oracle bit for bit wherever both run (test_gpu_dense,
test_gpu_boundaries, test_gpu_patches), so at full size the FP64 device run is the FP64 reference
that the FP32-shifted path is measured against.  Bars, each stated in its test:
  * Taylor-Green vortex 512^3, u0 = 0.04, tau = 0.52, 5,000 steps, kinetic energy every 250 steps:
    FP32-shifted energy within 1e-7 relative of FP64 at every sample; energy strictly decreasing;
    mass conserved to 1e-10 relative; early decay within 10 % of the viscous rate 2 nu 3 k^2.
  * Taylor-Green 384^3 in FP64, 5,000 steps: the same conservation and decay properties, and the
    FP32-shifted curve within 1e-7.
  * porous channel 256^3 (Guo g = 1e-6, spheres to porosity 0.8), tau in {0.55, 0.6, 0.7, 0.8, 1.0},
    2,000 steps: FP32-shifted velocity L2 <= 1e-5 and Linf <= 3e-5 (relative, fluid cells) against
    FP64; x-momentum within 1e-5 relative.
  * 4,096 uneven patches, 500 graph-replayed steps: FP32-shifted u L2 and Linf <= 1e-6 against FP64;
    relative mass drift <= 1e-12 (FP64) and <= 1e-9 (FP32-shifted).
  * adjoint, 128^3 duct with Fourier porosity, 500 steps, 23 revolve snapshots: the FP64 gradient
    within 1e-6 (normwise) of Richardson-extrapolated FP64 central differences on 64 seeded cells,
    the FP32-shifted gradient within 2e-5 (normwise) of the FP64 one.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

lbm = pytest.importorskip("madpac_amd.lbm")
from madpac_amd import validate  # noqa: E402
from madpac_amd import workloads as wl  # noqa: E402

_tgv = {}


def _taylor_green(n, name):
    key = (n, name)
    if key not in _tgv:
        print(f"[gpu] Taylor-Green {n}^3, {name}, 5000 steps", flush=True)
        _tgv[key] = validate.taylor_green(n=n, steps=5000, precisions=(name,))
    return _tgv[key]


def _check_tgv_run(r, n, name):
    run = r["runs"][name]
    E = np.array(run["energy"])
    assert np.all(np.diff(E) < 0.0), "kinetic energy must decrease"
    assert abs(run["mass_drift"]) <= 1e-10 * n ** 3, run["mass_drift"]
    early = r["early_decay"]
    assert abs(early["measured_rate"] / early["viscous_rate"] - 1.0) <= 0.10, early
    return E


@pytest.mark.parametrize("name", ["fp64", "fp32s"])
def test_taylor_green_512(name):
    r = _taylor_green(512, name)
    E = _check_tgv_run(r, 512, name)
    print(f"[gpu] TGV 512 {name}: {r['runs'][name]['mlups']:.0f} MLUPS, E(T)/E0 {E[-1] / E[0]:.6f}", flush=True)
    if name == "fp32s":
        E64 = np.array(_taylor_green(512, "fp64")["runs"]["fp64"]["energy"])
        rel = np.abs(E - E64) / E64
        print(f"[gpu] TGV 512 FP32-shifted vs FP64 energy: max rel {rel.max():.2e}", flush=True)
        assert rel.max() <= 1e-7


@pytest.mark.parametrize("name", ["fp64", "fp32s"])
def test_taylor_green_384(name):
    r = _taylor_green(384, name)
    E = _check_tgv_run(r, 384, name)
    if name == "fp32s":
        E64 = np.array(_taylor_green(384, "fp64")["runs"]["fp64"]["energy"])
        assert (np.abs(E - E64) / E64).max() <= 1e-7


_porous = {}


@pytest.mark.parametrize("tau", [0.55, 0.6, 0.7, 0.8, 1.0])
def test_porous_channel_256_fp32_shifted_against_fp64(tau):
    n, steps, g = 256, 2000, 1e-6
    if "solid" not in _porous:
        _porous["solid"] = wl.porous_spheres(n)
    solid = _porous["solid"]
    fluid = ~solid.reshape(-1)
    res = {}
    for prec in (lbm.FP64, lbm.FP32_SHIFTED):
        print(f"[gpu] porous 256^3 tau {tau} precision {prec}, {steps} steps", flush=True)
        with lbm.Lattice(n, n, n, tau, precision=prec, periodic=(True, True, True), force=(g, 0.0, 0.0)) as lat:
            lat.set_flags(~solid)
            lat.init_equilibrium()
            lat.step(steps)
            res[prec] = (lat.fields()[1], lat.moments())
    u64, m64 = res[lbm.FP64]
    u32, m32 = res[lbm.FP32_SHIFTED]
    d = (u32 - u64)[fluid]
    l2 = np.linalg.norm(d) / np.linalg.norm(u64[fluid])
    linf = np.abs(d).max() / np.abs(u64[fluid]).max()
    print(f"[gpu] porous tau {tau}: u L2 {l2:.2e} Linf {linf:.2e}", flush=True)
    assert l2 <= 1e-5 and linf <= 3e-5
    assert abs(m32[1] - m64[1]) <= 1e-5 * abs(m64[1])


def test_uneven_patches_4096_fp32_shifted_against_fp64():
    print("[gpu] 4096 uneven patches, 500 graph-replayed steps", flush=True)
    r = validate.patches_precision(steps=500)
    assert abs(r["fp64"]["mass_drift_rel"]) <= 1e-12
    assert abs(r["fp32s"]["mass_drift_rel"]) <= 1e-9
    assert r["fp32s"]["u"]["L2"] <= 1e-6 and r["fp32s"]["u"]["Linf"] <= 1e-6, r["fp32s"]["u"]


def test_adjoint_128_against_richardson_differences():
    print("[gpu] adjoint 128^3 x 500 steps, FD on 64 cells", flush=True)
    r = validate.adjoint_fd(n=128, steps=500, snapshots=23, ncells=64)
    e64 = r["adjoint"]["fp64"]["vs_fd_richardson"]
    print(f"[gpu] adjoint FP64 vs Richardson FD: normwise {e64['normwise_rel_err']:.2e}", flush=True)
    assert e64["normwise_rel_err"] <= 1e-6
    g64 = np.array(r["adjoint"]["fp64"]["grad"])
    g32 = np.array(r["adjoint"]["fp32s"]["grad"])
    rel = np.linalg.norm(g32 - g64) / np.linalg.norm(g64)
    print(f"[gpu] adjoint FP32-shifted vs FP64 on the 64 cells: {rel:.2e}", flush=True)
    assert rel <= 2e-5
    assert r["adjoint"]["fp64"]["snapshots"] == 23
