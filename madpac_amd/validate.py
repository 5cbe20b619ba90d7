"""Device validation runs (one GPU): precision of the FP32-shifted path against FP64, and the
large-scale Taylor-Green vortex.

  python -m madpac_amd.validate precision [--size 256 --steps 2000]   porous-channel tau sweep
  python -m madpac_amd.validate tgv [--size 512 --steps 5000]          Taylor-Green energy decay
  python -m madpac_amd.validate tgv --size 384 --precisions fp64
  python -m madpac_amd.validate adjoint [--size 128 --steps 500]       gradient vs central differences
  python -m madpac_amd.validate patches [--steps 500]                  4,096 uneven patches, FP32 vs FP64

The FP64 device path is bit-identical to the CPU oracle (tests/test_gpu_dense.py), so it serves as
the FP64 reference at sizes the CPU oracle cannot finish in time.  Output: one JSON object.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import lbm
from . import workloads as wl


def _log(*a):
    print(*a, file=sys.stderr, flush=True)


def _errors(ref, val, fluid):
    """Relative L2 and Linf of val against ref over fluid cells (per component and vector norm)."""
    out = {}
    r, v = ref[fluid], val[fluid]
    if r.ndim == 1:
        d = v - r
        out["L2"] = float(np.linalg.norm(d) / np.linalg.norm(r))
        out["Linf"] = float(np.max(np.abs(d)) / np.max(np.abs(r)))
        return out
    d = v - r
    out["L2"] = float(np.linalg.norm(d) / np.linalg.norm(r))
    out["Linf"] = float(np.max(np.abs(d)) / np.max(np.abs(r)))
    for k, name in enumerate("xyz"):
        nr = np.linalg.norm(r[:, k])
        out[f"u{name}_L2"] = float(np.linalg.norm(d[:, k]) / nr) if nr > 0 else 0.0
    return out


def porous_precision_sweep(n=256, steps=2000, taus=(0.55, 0.6, 0.7, 0.8, 1.0), g=1e-6):
    solid = wl.porous_spheres(n)
    fluid = ~solid.reshape(-1)
    rows = []
    for tau in taus:
        res = {}
        for name, prec in (("fp64", lbm.FP64), ("fp32s", lbm.FP32_SHIFTED), ("fp32s_f32math", lbm.FP32_SHIFTED_F32)):
            with lbm.Lattice(n, n, n, tau, precision=prec, periodic=(True, True, True), force=(g, 0.0, 0.0)) as lat:
                lat.set_flags(~solid)
                lat.init_equilibrium()
                lat.tune(5)
                t0 = time.perf_counter()
                ms = lat.step_timed(steps)
                res[name] = {"ms_per_step": ms / steps, "fields": lat.fields(), "moments": lat.moments(),
                             "wall_s": time.perf_counter() - t0}
        rho64, u64 = res["fp64"]["fields"]
        m64 = res["fp64"]["moments"]
        row = {"tau": tau, "u_max_fp64": float(np.max(np.abs(u64[fluid]))), "mass_fp64": float(m64[0]),
               "momentum_x_fp64": float(m64[1]), "ms_per_step": {"fp64": res["fp64"]["ms_per_step"]}}
        for name in ("fp32s", "fp32s_f32math"):
            rho32, u32 = res[name]["fields"]
            m32 = res[name]["moments"]
            row[name] = {"u": _errors(u64, u32, fluid),
                         "rho_deviation": {"L2": float(np.linalg.norm((rho32 - rho64)[fluid]) /
                                                       np.linalg.norm((rho64 - 1.0)[fluid])),
                                           "Linf_abs": float(np.max(np.abs(rho32 - rho64)[fluid]))},
                         "mass": float(m32[0]), "momentum_x": float(m32[1]),
                         "momentum_x_rel_err": float((m32[1] - m64[1]) / m64[1])}
            row["ms_per_step"][name] = res[name]["ms_per_step"]
        _log(f"[validate] tau {tau}: FP32-shifted u L2 {row['fp32s']['u']['L2']:.3e} Linf "
             f"{row['fp32s']['u']['Linf']:.3e}; FP32 arithmetic u L2 {row['fp32s_f32math']['u']['L2']:.3e}; "
             f"|u|max {row['u_max_fp64']:.3e}")
        rows.append(row)
    return {"workload": f"porous channel {n}^3 periodic, Guo g={g} along x, spheres r 4..12 to porosity 0.8, "
                        f"{steps} steps", "fluid_cells": int(fluid.sum()), "sweep": rows}


PRECISIONS = {"fp64": lbm.FP64, "fp32s": lbm.FP32_SHIFTED, "fp32s_f32math": lbm.FP32_SHIFTED_F32}


def taylor_green(n=512, steps=5000, u0=0.04, tau=0.52, precisions=("fp32s", "fp64"), every=250,
                 layout=lbm.LAYOUT_TWO_BUFFER):
    nu = (tau - 0.5) / 3.0
    k = 2 * np.pi / n
    t0 = time.perf_counter()
    rho, u = wl.taylor_green(n, u0)
    rho = rho.reshape(-1)
    u = u.reshape(-1, 3)
    _log(f"[validate] TGV {n}^3 initial fields ({time.perf_counter() - t0:.1f} s)")
    out = {"workload": f"Taylor-Green vortex {n}^3 periodic, u0={u0}, tau={tau} (Re = u0 n / (2 pi nu) = "
                       f"{u0 * n / (2 * np.pi * nu):.0f}), {steps} steps"
                       + (", in-place AA layout (one state buffer)" if layout == lbm.LAYOUT_AA else ""),
           "runs": {}}
    curves = {}
    for name in precisions:
        prec = PRECISIONS[name]
        with lbm.Lattice(n, n, n, tau, precision=prec, periodic=(True, True, True), layout=layout) as lat:
            lat.set_flags(np.ones(n ** 3, np.uint8))
            lat.init_equilibrium(rho=rho, u=u)
            lat.tune(10)
            lat.init_equilibrium(rho=rho, u=u)
            E, ts, ms_total = [lat.kinetic_energy()], [0], 0.0
            for s in range(every, steps + 1, every):
                ms_total += lat.step_timed(every)
                E.append(lat.kinetic_energy())
                ts.append(s)
            m = lat.moments()
        esize = 8 if prec == lbm.FP64 else 4
        per = ms_total * 1e-3 / steps
        gbs = n ** 3 * 38 * esize / per / 1e9
        curves[name] = np.array(E)
        out["runs"][name] = {"ms_per_step": per * 1e3, "mlups": n ** 3 / per / 1e6, "achieved_GBs": gbs,
                             "hbm_frac": gbs / 8000.0, "energy": E, "t": ts, "mass_drift": float(m[0] - n ** 3),
                             "state_GB": 19 * n ** 3 * esize * (1 if layout == lbm.LAYOUT_AA else 2) / 1e9}
        _log(f"[validate] TGV {name}: {n ** 3 / per / 1e6:.0f} MLUPS, {gbs:.0f} GB/s, E(T)/E0 {E[-1] / E[0]:.5f}")
    # early-time viscous decay E(t) = E0 exp(-2 nu k_eff^2 t) with k_eff^2 = 3 k^2 for this vortex
    first = next(iter(curves.values()))
    t1 = out["runs"][next(iter(curves))]["t"][1]
    out["early_decay"] = {"measured_rate": float(-np.log(first[1] / first[0]) / t1),
                          "viscous_rate": float(2 * nu * 3 * k * k)}
    if "fp64" in curves and "fp32s" in curves:
        rel = np.abs(curves["fp32s"] - curves["fp64"]) / curves["fp64"]
        out["energy_rel_diff_fp32s_vs_fp64"] = {"max": float(rel.max()), "final": float(rel[-1])}
    return out


def adjoint_fd(n=128, steps=500, snapshots=23, ncells=64, h=1e-4, hr=1e-2, g=1e-5, tau=0.6, seed=wl.DEFAULT_SEED):
    """The adjoint workload: n^3 duct (walls on the y/z faces, periodic along x, body force g along x as
    the driving pressure gradient), Brinkman porosity from 8 Fourier modes, J = FP64 pressure drop
    between the planes x = n/4 and 3n/4 after `steps` steps.  The gradient dJ/dalpha from the FP32-
    shifted and FP64 adjoints (revolve with `snapshots` checkpoints) against FP64 central differences
    on `ncells` seeded fluid cells: with step h, and Richardson-extrapolated from steps hr and 2 hr,
    (4 D(hr) - D(2 hr)) / 3, whose O(hr^4) truncation and O(eps/hr) round-off are both far below the
    plain difference's O(eps/h) round-off (the forward solve's FP64 rounding, divided by 2h)."""
    solid = wl.channel_walls(n, n, n)
    alpha = wl.fourier_porosity(n, n, n).reshape(-1)
    xin, xout = n // 4, 3 * n // 4
    fluid = ~solid.reshape(-1)

    def make(prec):
        lat = lbm.Lattice(n, n, n, tau, precision=prec, periodic=(True, False, False), force=(g, 0.0, 0.0))
        lat.set_flags(~solid)
        return lat

    out = {"workload": f"{n}^3 duct, Brinkman porosity (8 Fourier modes), g={g} along x, tau={tau}, {steps} steps, "
                       f"J = FP64 pressure drop x={xin} -> x={xout}, revolve {snapshots} snapshots, "
                       f"FD h={h} (FP64) on {ncells} cells", "adjoint": {}}
    grads = {}
    for name, prec in (("fp64", lbm.FP64), ("fp32s", lbm.FP32_SHIFTED)):
        with make(prec) as lat:
            lat.set_porosity(alpha)
            lat.init_equilibrium()
            t0 = time.perf_counter()
            J, grad, st = lat.adjoint_pressure_drop(steps, snapshots, xin, xout)
            wall = time.perf_counter() - t0
        grads[name] = grad
        out["adjoint"][name] = {"J": J, "wall_s": wall, "forward_ms": st.forward_ms, "reverse_ms": st.reverse_ms,
                                "recompute_steps": st.recompute_steps, "snapshots": st.snapshots_used}
        _log(f"[validate] adjoint {name}: J = {J:.12e}, {wall:.2f} s")
    rng = np.random.default_rng(seed)
    ok = np.flatnonzero(fluid & (alpha >= 2 * hr) & (alpha <= 1.0 - 2 * hr))
    cells = np.sort(rng.choice(ok, size=ncells, replace=False))
    fd = {"h": np.empty(ncells), "richardson": np.empty(ncells)}
    with make(lbm.FP64) as lat:
        def J_of(a):
            lat.set_porosity(a)
            lat.init_equilibrium()
            lat.step(steps)
            return lat.pressure_drop(xin, xout)

        def central(i, step):
            ap, am = alpha.copy(), alpha.copy()
            ap[i] += step
            am[i] -= step
            return (J_of(ap) - J_of(am)) / (2 * step)

        t0 = time.perf_counter()
        for k, i in enumerate(cells):
            fd["h"][k] = central(i, h)
            fd["richardson"][k] = (4.0 * central(i, hr) - central(i, 2 * hr)) / 3.0
        _log(f"[validate] {ncells} cells x 3 central differences in {time.perf_counter() - t0:.1f} s")
    out["cells"] = cells.tolist()
    out["fd"] = {k: v.tolist() for k, v in fd.items()}
    for name, grad in grads.items():
        gk = grad[cells]
        out["adjoint"][name]["grad"] = gk.tolist()
        for kind, ref in fd.items():
            big = np.abs(ref) >= 1e-2 * np.abs(ref).max()
            rel = np.abs(gk - ref) / np.abs(ref)
            out["adjoint"][name][f"vs_fd_{kind}"] = {
                "normwise_rel_err": float(np.linalg.norm(gk - ref) / np.linalg.norm(ref)),
                "max_rel_err_large": float(rel[big].max()),
                "max_abs_err_over_max": float(np.abs(gk - ref).max() / np.abs(ref).max())}
            e = out["adjoint"][name][f"vs_fd_{kind}"]
            _log(f"[validate] {name} adjoint vs FD ({kind}): normwise {e['normwise_rel_err']:.2e}, "
                 f"max rel (|fd| >= 1e-2 max) {e['max_rel_err_large']:.2e}")
    return out


def patches_precision(steps=500):
    """The batched-patch workload (4,096 uneven patches, per-patch tau in [0.55, 1.5], u0 ~ N(0, 0.01^2))
    in FP64, FP32-shifted and FP32-shifted with FP32 arithmetic, `steps` graph-replayed steps each:
    rho and u errors against the FP64 run (bit-identical to the oracle on tori), mass and momentum."""
    table, _ = wl.uneven_patches()
    res = {}
    for name, prec in (("fp64", lbm.FP64), ("fp32s", lbm.FP32_SHIFTED), ("fp32s_f32math", lbm.FP32_SHIFTED_F32)):
        with lbm.PatchSet(table, precision=prec) as ps:
            rng = np.random.default_rng(wl.DEFAULT_SEED)
            u0 = 0.01 * rng.standard_normal((ps.cells, 3))
            ps.init(u=u0)
            m0 = ps.moments()
            t0 = time.perf_counter()
            ps.step(steps, graph=True)
            res[name] = {"fields": ps.fields(), "moments": ps.moments(), "moments0": m0,
                         "wall_s": time.perf_counter() - t0, "cells": ps.cells}
        _log(f"[validate] patches {name}: {steps} steps in {res[name]['wall_s']:.1f} s")
    rho64, u64 = res["fp64"]["fields"]
    every = np.ones(rho64.shape, bool)
    out = {"workload": f"{len(table)} uneven patches ({res['fp64']['cells']} cells), {steps} steps",
           "u_rms_fp64": float(np.sqrt(np.mean(u64 ** 2)))}
    for name in ("fp64", "fp32s", "fp32s_f32math"):
        m, m0 = res[name]["moments"], res[name]["moments0"]
        row = {"mass_drift_rel": float((m[0] - m0[0]) / m0[0]),
               "momentum_drift": [float(m[k] - m0[k]) for k in (1, 2, 3)]}
        if name != "fp64":
            rho, u = res[name]["fields"]
            row["u"] = _errors(u64, u, every)
            row["rho"] = {"L2_of_deviation": float(np.linalg.norm(rho - rho64) / np.linalg.norm(rho64 - 1.0)),
                          "Linf_abs": float(np.max(np.abs(rho - rho64)))}
            _log(f"[validate] patches {name}: u L2 {row['u']['L2']:.2e} Linf {row['u']['Linf']:.2e}, "
                 f"rho Linf {row['rho']['Linf_abs']:.1e}, mass drift {row['mass_drift_rel']:.1e}")
        out[name] = row
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["precision", "tgv", "adjoint", "patches"])
    ap.add_argument("--size", type=int, default=None)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--precisions", default="fp32s,fp64")
    ap.add_argument("--layout", choices=["two-buffer", "aa"], default="two-buffer")
    ap.add_argument("--every", type=int, default=250, help="tgv: steps between kinetic-energy readings")
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    if a.what == "precision":
        r = porous_precision_sweep(n=a.size or 256, steps=a.steps or 2000)
    elif a.what == "adjoint":
        r = adjoint_fd(n=a.size or 128, steps=a.steps or 500)
    elif a.what == "patches":
        r = patches_precision(steps=a.steps or 500)
    else:
        r = taylor_green(n=a.size or 512, steps=a.steps or 5000, precisions=tuple(a.precisions.split(",")),
                         layout=lbm.LAYOUT_AA if a.layout == "aa" else lbm.LAYOUT_TWO_BUFFER, every=a.every)
    s = json.dumps(r)
    if a.out:
        with open(a.out, "w") as fh:
            fh.write(s + "\n")
    print(s, flush=True)


if __name__ == "__main__":
    main()
