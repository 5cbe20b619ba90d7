"""Benchmark: fused D3Q19 collide+stream MLUPS and %HBM roofline at 256^3, FP64 vs FP32-shifted.

Workload (one "step" = one fused collide+stream pass over the lattice): the porous channel --
256^3 periodic box, Guo body force g = 1e-6 along x, random solid spheres (radius 4..12, seeded)
until porosity 0.8, bounce-back on the spheres, tau = 0.6.  Populations are resident in HBM when
the timed region starts.  MLUPS counts fluid-cell updates.  Algorithmic bytes per fluid update:
2 * 19 * sizeof(T) (each population read once and written once).

Prints ONE JSON line.  `value` is the FP64 rate; the FP32-shifted rate, the adjoint/forward cost
and the CPU baseline (the FP64 oracle, OpenMP on the host cores, bounded sample) ride along.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
METRIC = "D3Q19 MLUPS + %HBM roofline at 256^3, FP64 vs FP32-shifted; adjoint/fwd cost"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "dense_step_traffic.md")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def read_traffic(key):
    """HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (committed summary)."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        if line.startswith("|") and key in line:
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            try:
                return float(cells[-1])
            except ValueError:
                return None
    return None


def run_dense(lbm, solid, n, tau, prec, steps, warmup, g, layout=0, readout=False):
    lat = lbm.Lattice(n, n, n, tau, precision=prec, periodic=(True, True, True), force=(g, 0.0, 0.0), layout=layout)
    lat.set_flags(~solid)
    lat.init_equilibrium()
    nfluid = lat.fluid_cells
    order = "bricks" if lat.cell_order == lbm.CELL_ORDER_BRICKS else "rows"
    variant = lat.tune(40)  # load schedule, workgroup order and shape (bit-identical results either way)
    lat.step(warmup)
    lat.synchronize()
    t0 = time.perf_counter()
    ms = lat.step_timed(steps)  # HIP events on the context stream around exactly `steps` launches
    lat.synchronize()
    wall = time.perf_counter() - t0
    ceil = None
    if layout == 0:
        # the step's bandwidth references on the same buffers, grid and workgroup mapping: a 19-plane
        # copy and the pull pattern over every stored cell, no flag word, no arithmetic
        cells = n ** 3
        esize_ = 8 if prec == lbm.FP64 else 4
        ceil = {}
        for key, pull in (("plane_copy", False), ("pull_copy", True)):
            lat.measure_copy(2, pull=pull)
            cms = lat.measure_copy(steps, pull=pull) / steps
            ceil[key] = {"ms": round(cms, 4), "GBs": round(cells * 38 * esize_ / (cms * 1e-3) / 1e9, 1)}
    m = lat.moments()
    host = None
    if readout:  # what a caller pays to bring results to the host (PCIe), outside the timed steps
        t0 = time.perf_counter()
        lat.fields()
        t1 = time.perf_counter()
        lat.get_populations()
        t2 = time.perf_counter()
        host = {"fields_s": round(t1 - t0, 4), "fields_bytes": n ** 3 * 4 * 8,
                "populations_s": round(t2 - t1, 4), "populations_bytes": n ** 3 * 19 * 8}
    lat.close()
    esize = 8 if prec == lbm.FP64 else 4
    per_launch_s = ms * 1e-3 / steps
    bytes_per_launch = nfluid * 2 * 19 * esize
    achieved = bytes_per_launch / per_launch_s / 1e9
    schedule = "aa-in-place" if layout == lbm.LAYOUT_AA else ("loads-with-flag-word" if variant & 1 else "flag-word-first")
    dispatch = "xcd-contiguous" if variant & 2 else "dispatch-order"
    shape = (variant >> 2) & 3
    # the shape bits are 0 only when the tuner could not copy the state: then the default holds
    wg_rows = (2, 4, 8)[shape - 1] if shape else 4
    key = f"{'fp64' if prec == lbm.FP64 else 'fp32s'} {n}^3 {schedule} ({wg_rows} {order}/workgroup)"
    traffic = read_traffic(key)
    return {
        "mlups": nfluid * steps / (ms * 1e-3) / 1e6,
        "wall_mlups": nfluid * steps / wall / 1e6,
        "ms_per_step": ms / steps,
        "wall_s": wall,
        "fluid_cells": nfluid,
        "mass": m[0],
        "variant": schedule,
        "workgroup_order": dispatch,
        "workgroup_rows": wg_rows,
        "cell_order": order,
        "host_readout": host,
        "copy_reference": ceil,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes_per_launch": bytes_per_launch},
    }


def run_md_coupling(lbm, solid, n, tau, steps=100, box=40):
    """The MD coupling hooks (PeanoDataExchangeService.cpp:69-218) on the porous channel's geometry
    without its body force (coupled cells take the plain BGK collision, which the library enforces),
    FP64: the fluid cells of a box^3 MD region in the middle of the lattice are coupled in Dirichlet mode (after
    the step kernel, k_couple re-collides them with the MD velocity); the step with coupling against
    the same step without it, both timed with HIP events on the context stream, and the export of the
    region's density and velocity (k_couple_gather plus the copy to host memory, wall clock)."""
    lo = (n - box) // 2
    x, y, z = np.meshgrid(*(np.arange(lo, lo + box),) * 3, indexing="ij")
    cells = (x + n * (y + n * z)).reshape(-1)
    cells = np.sort(cells[~solid.reshape(-1)[cells]])
    with lbm.Lattice(n, n, n, tau, precision=lbm.FP64, periodic=(True, True, True)) as lat:
        lat.set_flags(~solid)
        lat.init_equilibrium()
        lat.tune(20)
        lat.step(10)
        plain = lat.step_timed(steps) / steps
        lat.set_coupling_cells(cells)
        u = np.tile([0.01, 0.0, 0.0], (len(cells), 1))
        lat.set_coupling_data(lbm.COUPLE_DIRICHLET, u, dt=1.0)
        lat.step(10)
        coupled = lat.step_timed(steps) / steps
        lat.synchronize()
        t0 = time.perf_counter()
        for _ in range(10):
            lat.get_coupling_data()
        export = (time.perf_counter() - t0) / 10 * 1e3
    return {"workload": f"porous channel geometry {n}^3 without body force, FP64, the fluid cells of a {box}^3 "
                        f"MD region, Dirichlet mode",
            "coupled_cells": int(len(cells)), "step_ms": round(plain, 4), "coupled_step_ms": round(coupled, 4),
            "overhead_ms": round(coupled - plain, 4),
            "overhead_ns_per_coupled_cell": round((coupled - plain) * 1e6 / len(cells), 2),
            "export_ms": round(export, 4), "export_what": "density and velocity of the coupled cells to host "
                                                         "memory, wall clock"}


def run_taylor_green(lbm, wl, n=512, steps=200, tau=0.52, u0=0.04):
    """All-fluid periodic Taylor-Green vortex (the large-scale validation workload), timed like the
    porous channel: the same kernels without solid surfaces."""
    rho, u = wl.taylor_green(n, u0)
    rho, u = rho.reshape(-1), u.reshape(-1, 3)
    out = {"workload": f"Taylor-Green vortex {n}^3 periodic, tau {tau}, u0 {u0}, {steps} steps"}
    for name, prec in (("fp64", lbm.FP64), ("fp32s", lbm.FP32_SHIFTED)):
        with lbm.Lattice(n, n, n, tau, precision=prec, periodic=(True, True, True)) as lat:
            lat.set_flags(np.ones(n ** 3, np.uint8))
            lat.init_equilibrium(rho=rho, u=u)
            lat.tune(20)
            lat.step(10)
            ms = lat.step_timed(steps) / steps
        esize = 8 if prec == lbm.FP64 else 4
        gbs = n ** 3 * 38 * esize / (ms * 1e-3) / 1e9
        out[name] = {"value": round(n ** 3 / (ms * 1e3), 1), "unit": "MLUPS", "ms_per_step": round(ms, 4),
                     "roofline_frac": round(gbs / HBM_PEAK_GBS, 4)}
    return out


def run_shard_share(lbm, solid, tau, g, planes=32, steps=300):
    """One GPU's share of an 8-way z-slab split of the porous channel: the first `planes` + 2 planes of
    the same geometry (the 2 extra planes stand in for the ghost planes a sharded run would fill from
    its neighbours; here they are updated too, with bounce-back beyond them)."""
    n = solid.shape[1]
    nz = planes + 2
    sub = np.ascontiguousarray(solid[:nz])
    out = {"workload": f"porous channel slab {n}x{n}x{nz} (1/8 of {n}^3 plus 2 ghost-plane stand-ins)"}
    for name, prec in (("fp64", lbm.FP64), ("fp32s", lbm.FP32_SHIFTED)):
        with lbm.Lattice(n, n, nz, tau, precision=prec, periodic=(True, True, False), force=(g, 0.0, 0.0)) as lat:
            lat.set_flags(~sub)
            lat.init_equilibrium()
            lat.tune(20)
            lat.step(20)
            ms = lat.step_timed(steps) / steps
            nfl = lat.fluid_cells
            order = "bricks" if lat.cell_order == lbm.CELL_ORDER_BRICKS else "rows"
        esize = 8 if prec == lbm.FP64 else 4
        out[name] = {"value": round(nfl / (ms * 1e3), 1), "unit": "MLUPS", "ms_per_step": round(ms, 4),
                     "roofline_frac": round(nfl * 38 * esize / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "cell_order": order}
    return out


def run_adjoint(lbm, wl, n=128, steps=500, snapshots=23):
    solid = wl.channel_walls(n, n, n)
    alpha = wl.fourier_porosity(n, n, n).reshape(-1)
    out = {}
    lat = lbm.Lattice(n, n, n, 0.6, precision=lbm.FP32_SHIFTED, periodic=(True, False, False), force=(1e-5, 0, 0))
    lat.set_flags(~solid)
    lat.set_porosity(alpha)
    lat.init_equilibrium()
    lat.step(10)
    lat.synchronize()
    fwd_ms = lat.step_timed(steps) / steps
    lat.init_equilibrium()
    J, grad, st = lat.adjoint_pressure_drop(steps, snapshots, n // 4, 3 * n // 4)
    lat.close()
    # all device times measured with HIP events: the forward sweep, and inside the reverse sweep the
    # recomputed forward steps and the adjoint steps separately
    adj_step_ms = st.adjoint_ms / st.reverse_steps
    fwd_sweep_step_ms = st.forward_ms / st.forward_steps
    out.update({
        "workload": f"{n}^3 duct, Brinkman porosity (8 Fourier modes), FP32-shifted, {steps} steps, "
                    f"revolve {snapshots} snapshots, J = FP64 pressure drop",
        "J": J, "forward_ms_per_step": fwd_ms, "adjoint_ms_per_step": adj_step_ms,
        "adjoint_over_forward_step": adj_step_ms / fwd_ms,
        "recompute_steps": int(st.recompute_steps),
        "recompute_ms_per_step": st.recompute_ms / max(int(st.recompute_steps), 1),
        "forward_sweep_ms_per_step": fwd_sweep_step_ms,
        "reverse_ms": st.reverse_ms, "recompute_ms": st.recompute_ms, "adjoint_ms": st.adjoint_ms,
        "gradient_over_forward_run": (st.forward_ms + st.reverse_ms) / (fwd_ms * steps),
        "budget": "adjoint step <= 3x a forward step",
        "grad_norm": float(np.linalg.norm(grad)),
    })
    return out


def run_patches(lbm, wl, prec, steps=500):
    """4096 uneven patches (edges 8/16/24/32, tau in [0.55,1.5], u0 ~ N(0, 0.01^2)), autotuned
    tiling, one persistent launch per step, hipGraph replay over `steps` steps."""
    table, tori = wl.uneven_patches()
    ps = lbm.PatchSet(table, precision=prec)
    rng = np.random.default_rng(wl.DEFAULT_SEED)
    ps.init(u=0.01 * rng.standard_normal((ps.cells, 3)))
    tiling = ps.autotune()
    ps.step(10, graph=True)
    ms = ps.step(steps, graph=True, timed=True)
    m = ps.moments()
    ps.close()
    esize = 8 if prec == lbm.FP64 else 4
    per = ms * 1e-3 / steps
    interior_bytes = ps.cells * 2 * 19 * esize
    ghost_vals = sum(6 * t[0] ** 2 * 5 + 12 * t[0] for t in table)  # written once, read once per step
    achieved = interior_bytes / per / 1e9
    return {"value": round(ps.cells * steps / (ms * 1e-3) / 1e6, 1), "unit": "MLUPS",
            "ms_per_step": round(ms / steps, 4), "patches": len(table), "cells": ps.cells,
            "tiling": {"tile": list(tiling[:3]), "blocks_per_cu": tiling[3],
                       "schedule": ("one workgroup per tile" if tiling[3] == 0 else
                                    "persistent, contiguous tile ranges" if tiling[4] else "persistent, strided")},
            "graph_steps": steps,
            "mass_drift": float(m[0] - ps.cells),
            "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBS, 4),
                         "traffic": read_traffic(f"patches {'fp64' if prec == lbm.FP64 else 'fp32s'} step"),
                         "algorithmic_bytes_per_launch": interior_bytes,
                         "ghost_exchange_GBs": round(ghost_vals * 2 * esize / per / 1e9, 1)}}


def run_block_store(lbm, ncoarse=4096, reps=5):
    """Restriction and prolongation between two grid levels on the device block store: every coarse
    block fathers 8 fine vertices (the positions of one octant of its refined cell), so one call per
    operation covers 8 * ncoarse links.  Overlap search on fine blocks with sparse -1.0 markers;
    prolongation with every cell an overlap cell (new vertices: the most data it moves); restriction
    of all sub-blocks.  Device time of each launch from HIP events around it."""
    rng = np.random.default_rng(5)
    nc, q = 216, 19
    nfine = 8 * ncoarse
    links = []
    for cb in range(ncoarse):
        f = (cb % 2, (cb // 2) % 2, (cb // 4) % 2)
        for k in range(8):
            links.append((cb, ncoarse + 8 * cb + k, tuple(2 * f[d] + ((k >> d) & 1) for d in range(3))))
    out = {"workload": f"{ncoarse} coarse and {nfine} fine 6^3 blocks, {len(links)} links per call, FP64"}
    with lbm.BlockStore(ncoarse + nfine) as bs:
        bs.set(0, pdf=rng.uniform(0.5, 1.5, (ncoarse, nc * q)))
        fine = rng.uniform(0.5, 1.5, (nfine, nc * q))
        fine[rng.random(fine.shape) < 1e-3] = -1.0
        bs.set(ncoarse, pdf=fine, density=np.ones((nfine, nc)), velocity=np.zeros((nfine, 3 * nc)))
        del fine
        per_link = {"find_overlap": nc * q * 8 + nc,
                    "prolongate": nc + 8 * q * 8 + nc * (q + 4) * 8,
                    "restrict": nc * q * 8 + 2 * 8 * q * 8}
        for name in ("find_overlap", "prolongate", "restrict"):
            ms = []
            for _ in range(reps):
                if name == "find_overlap":
                    bs.find_overlap(links)
                elif name == "prolongate":
                    bs.set_overlap(ncoarse, np.ones(nfine * nc, np.uint8))
                    bs.prolongate(links)
                else:
                    bs.restrict(links)
                ms.append(bs.last_ms())
            t = float(np.mean(ms[1:]))
            gbs = len(links) * per_link[name] / (t * 1e-3) / 1e9
            out[name] = {"ms": round(t, 4), "algorithmic_bytes_per_link": per_link[name], "GBs": round(gbs, 1),
                         "roofline_frac": round(gbs / HBM_PEAK_GBS, 4)}
    return out


def regular_block_grid(n, ulid):
    """The n^3 lid-driven cavity as the reference's regular-grid block solver lays it out: n // 6 + 1
    (+ 1 if 6 does not divide n) blocks per axis, slot s of block b holding cell 6 b + s - 3, inner
    cells inside the box, Peano cells between neighbouring vertices (8 blocks, lexicographic), and
    the lid's moving-wall links (top layer, c_z > 0)."""
    nb = n // 6 + 1 + (1 if n % 6 else 0)
    b = np.arange(nb ** 3)
    s = np.arange(216)
    bxyz = np.stack([b % nb, (b // nb) % nb, b // (nb * nb)], axis=1)
    sxyz = np.stack([s % 6, (s // 6) % 6, s // 36], axis=1)
    inner = np.ones((nb ** 3, 216), bool)
    for d in range(3):
        g = 6 * bxyz[:, None, d] + sxyz[None, :, d] - 3
        inner &= (g >= 0) & (g < n)
        if d == 2:
            top = inner & (g == n - 1)
    c0 = np.arange(nb - 1)
    cz, cy, cx = np.meshgrid(c0, c0, c0, indexing="ij")
    base = (cx + nb * (cy + nb * cz)).reshape(-1)
    cells = np.stack([base + (v & 1) + nb * ((v >> 1) & 1) + nb * nb * (v >> 2) for v in range(8)], axis=1)
    blk, cell = np.nonzero(top)
    links = [(int(bb), int(cc), 18 - q, 1, (ulid, 0.0, 0.0), 1.0) for bb, cc in zip(blk, cell) for q in range(14, 19)]
    return nb ** 3, inner, cells.astype(np.int64), links


def run_block_solver(lbm, sizes=(64, 256), steps=20, tau=0.6, ulid=0.05):
    """The reference's own three-sweep algorithm on the device block store (lbm_blocks_iterate): per
    iteration one launch over all blocks (the in-cell swap of the previous iteration, cell values,
    BGK, lid links, local streaming, each block staged once through LDS) and one over all Peano cells
    (cross-block swaps in place).  Bytes per iteration: 2 x 4104 x 8 B of populations + 864 x 8 B of
    density and velocity + 216 inner bytes per block, and 504 x 32 B per Peano cell."""
    out = {}
    w = np.array([1 / 36, 1 / 36, 1 / 18, 1 / 36, 1 / 36, 1 / 36, 1 / 18, 1 / 36, 1 / 18, 1 / 3,
                  1 / 18, 1 / 36, 1 / 18, 1 / 36, 1 / 36, 1 / 36, 1 / 18, 1 / 36, 1 / 36])
    tau_eff = 3.0 * (1.0 / 3.0 * (tau - 0.5)) + 0.5
    for n in sizes:
        nblocks, inner, cells, links = regular_block_grid(n, ulid)
        with lbm.BlockStore(nblocks) as bs:
            bs.set(0, pdf=np.broadcast_to(np.tile(w, 216), (nblocks, 216 * 19)), density=np.ones((nblocks, 216)))
            bs.set_inner(0, inner)
            blocks = np.arange(nblocks)
            bs.iterate(blocks, cells, tau_eff, links, 2)
            bs.iterate(blocks, cells, tau_eff, links, steps)
            ms = bs.last_ms() / steps
        nbytes = nblocks * (2 * 4104 * 8 + 864 * 8 + 216) + len(cells) * 504 * 32
        gbs = nbytes / (ms * 1e-3) / 1e9
        out[f"cavity {n}^3"] = {"blocks": nblocks, "peano_cells": len(cells), "links": len(links), "steps": steps,
                                "ms_per_step": round(ms, 4), "mlups": round(n ** 3 / (ms * 1e-3) / 1e6, 1),
                                "bytes_per_step": nbytes, "GBs": round(gbs, 1), "roofline_frac": round(gbs / HBM_PEAK_GBS, 4)}
        log(f"[bench] block solver {n}^3 cavity: {out[f'cavity {n}^3']['mlups']:.0f} MLUPS, {ms:.3f} ms/step, "
            f"{gbs:.0f} GB/s")
    return out


def run_vertex_exchange(lbm, n=256, reps=20):
    """The per-GPU kernels of the vertex exchange between ranks (the reference's Parallel build) at
    their real shape: the 256^3 cavity's Peano cells cut into two z halves share one plane of
    44 x 44 = 1,936 vertices.  Per iteration a GPU records the pdf differences of those vertices
    before the streaming (init: 1,008 populations read, 1,008 differences written per vertex) and
    after it (compute: 1,008 populations and differences read, 1,008 differences written), copies
    them out for the neighbour, and merges the neighbour's (1,008 own and 1,008 neighbour
    differences read, and here half the incoming populations read and written: the other half
    were swapped locally)."""
    nb = n // 6 + 1 + (1 if n % 6 else 0)
    nv = nb * nb
    rng = np.random.default_rng(7)
    w = np.array([1 / 36, 1 / 36, 1 / 18, 1 / 36, 1 / 36, 1 / 36, 1 / 18, 1 / 36, 1 / 18, 1 / 3,
                  1 / 18, 1 / 36, 1 / 18, 1 / 36, 1 / 36, 1 / 36, 1 / 18, 1 / 36, 1 / 36])
    pdf = np.tile(w, 216) * (1.0 + 0.01 * rng.standard_normal((nv, 216 * 19)))
    # incoming / outgoing slots of the populations leaving a block (cells in z, y, x order, q in order)
    c = np.array([[0, -1, -1], [-1, 0, -1], [0, 0, -1], [1, 0, -1], [0, 1, -1], [-1, -1, 0], [0, -1, 0],
                  [1, -1, 0], [-1, 0, 0], [0, 0, 0], [1, 0, 0], [-1, 1, 0], [0, 1, 0], [1, 1, 0],
                  [0, -1, 1], [-1, 0, 1], [0, 0, 1], [1, 0, 1], [0, 1, 1]])
    cell = np.repeat(np.arange(216), 19)
    q = np.tile(np.arange(19), 216)
    t = np.stack([cell % 6, (cell // 6) % 6, cell // 36], axis=1) + c[q]
    leave = ~((t >= 0) & (t < 6)).all(axis=1)
    out_slot, in_slot = (cell * 19 + q)[leave], (cell * 19 + 18 - q)[leave]
    pdf[:, in_slot[::2]] = pdf[:, out_slot[::2]]
    neighbour = 1e-3 * rng.standard_normal((nv, 1008))
    ids = np.arange(nv)
    res = {"init": [], "compute": [], "merge": [], "get": []}
    with lbm.BlockStore(nv) as bs:
        for rep in range(reps + 2):
            bs.set(0, pdf=pdf)
            bs.init_pdf_diff(ids)
            t_init = bs.last_ms()
            bs.compute_pdf_diff(ids)
            t_comp = bs.last_ms()
            t0 = time.perf_counter()
            bs.get_pdf_diff(ids)
            t_get = (time.perf_counter() - t0) * 1e3
            bs.merge_pdf_diff(ids, neighbour)
            t_merge = bs.last_ms()
            if rep >= 2:
                for k, v in (("init", t_init), ("compute", t_comp), ("merge", t_merge), ("get", t_get)):
                    res[k].append(v)
    per = {"init": 2 * 1008 * 8, "compute": 3 * 1008 * 8, "merge": 2 * 1008 * 8 + 2 * 504 * 8}
    out = {"workload": f"{n}^3 cavity cut into two z halves: {nv} shared 6^3 vertices, FP64", "shared_vertices": nv,
           "differences_bytes": nv * 1008 * 8}
    for k in ("init", "compute", "merge"):
        ms = float(np.mean(res[k]))
        gbs = nv * per[k] / (ms * 1e-3) / 1e9
        out[k] = {"ms": round(ms, 4), "algorithmic_bytes_per_vertex": per[k], "GBs": round(gbs, 1)}
    ms = float(np.mean(res["get"]))
    out["get_to_host"] = {"ms": round(ms, 3), "GBs": round(nv * 1008 * 8 / (ms * 1e-3) / 1e9, 1),
                          "what": "device gather + PCIe copy to pageable host memory, wall clock"}
    log(f"[bench] vertex exchange, {nv} shared vertices: init {out['init']['ms']:.3f} ms, compute "
        f"{out['compute']['ms']:.3f} ms, merge {out['merge']['ms']:.3f} ms, differences to host {ms:.2f} ms")
    return out


def cpu_baseline(wl, n=128, seed=None, tau=0.6, g=1e-6, budget_s=12.0):
    import oracle

    cores = len(os.sched_getaffinity(0))
    solid = wl.porous_spheres(n) if seed is None else wl.porous_spheres(n, seed=seed)
    lat = oracle.Lattice(n, n, n, periodic=(True, True, True))
    lat.set_solid(solid)
    lat.set_force((g, 0.0, 0.0))
    lat.params.tau = tau
    f = lat.equilibrium()
    nfluid = int(lat.fluid.sum())
    t0 = time.perf_counter()
    f = lat.step(f, 2, nthreads=cores)
    t1 = time.perf_counter()
    per = (t1 - t0) / 2
    steps = max(2, int(budget_s / max(per, 1e-6)))
    t0 = time.perf_counter()
    f = lat.step(f, steps, nthreads=cores)
    dt = time.perf_counter() - t0
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": round(nfluid * steps / dt / 1e6, 2), "unit": "MLUPS", "cores": cores, "kind": "port",
            "sample": f"FP64 oracle (C, OpenMP) on the porous channel at {n}^3 (porosity 0.8, same generator), "
                      f"{steps} steps in {dt:.1f} s; CPU: {model}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--tau", type=float, default=0.6)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-patches", action="store_true")
    ap.add_argument("--no-all-fluid", action="store_true")
    ap.add_argument("--all-fluid-sizes", default="512,256",
                    help="edges of the all-fluid Taylor-Green legs (the porous channel's size runs --steps steps)")
    ap.add_argument("--patch-steps", type=int, default=500)
    ap.add_argument("--block-solver-sizes", default="64,256",
                    help="edges of the lid-driven cavities run through the block solver's operators")
    ap.add_argument("--block-steps", type=int, default=20)
    ap.add_argument("--only-block-solver", action="store_true",
                    help="run only the block solver leg and print its results as one JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: madpac_amd has no multi-GPU path (the domain would shard "
              "into slabs with an xGMI halo exchange).  Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    from madpac_amd import lbm, workloads as wl

    block_sizes = tuple(int(v) for v in args.block_solver_sizes.split(",") if v)
    if args.only_block_solver:
        out = run_block_solver(lbm, sizes=block_sizes, steps=args.block_steps)
        out["vertex_exchange"] = run_vertex_exchange(lbm)
        print(json.dumps(out), flush=True)
        return

    n = args.size
    t0 = time.perf_counter()
    solid = wl.porous_spheres(n)
    log(f"[bench] porous geometry {n}^3: solid fraction {solid.mean():.4f} ({time.perf_counter() - t0:.1f} s)")
    res = {}
    for name, prec in (("fp64", lbm.FP64), ("fp32s", lbm.FP32_SHIFTED), ("fp32s_f32math", lbm.FP32_SHIFTED_F32)):
        res[name] = run_dense(lbm, solid, n, args.tau, prec, args.steps, args.warmup, 1e-6, readout=name == "fp64")
        log(f"[bench] {name}: {res[name]['mlups']:.0f} MLUPS, {res[name]['ms_per_step']:.3f} ms/step, "
            f"{res[name]['roofline']['achieved']:.0f} GB/s ({100 * res[name]['roofline']['frac']:.1f} % of peak)")
        cr = res[name]["copy_reference"]
        log(f"[bench] {name} references at {n}^3: 19-plane copy {cr['plane_copy']['GBs']:.0f} GB/s, "
            f"pull copy {cr['pull_copy']['GBs']:.0f} GB/s")
    for name, prec in (("fp64_aa", lbm.FP64), ("fp32s_aa", lbm.FP32_SHIFTED)):
        res[name] = run_dense(lbm, solid, n, args.tau, prec, args.steps, args.warmup, 1e-6, layout=lbm.LAYOUT_AA)
        log(f"[bench] {name}: {res[name]['mlups']:.0f} MLUPS, {res[name]['ms_per_step']:.3f} ms/step, "
            f"{res[name]['roofline']['achieved']:.0f} GB/s ({100 * res[name]['roofline']['frac']:.1f} % of peak)")
    patches = None
    if not args.no_patches:
        patches = {}
        for name, prec in (("fp32s", lbm.FP32_SHIFTED), ("fp64", lbm.FP64)):
            patches[name] = run_patches(lbm, wl, prec, steps=args.patch_steps)
            log(f"[bench] patches {name}: {patches[name]['value']:.0f} MLUPS, tiling {patches[name]['tiling']}")
    shard = run_shard_share(lbm, solid, args.tau, 1e-6)
    log(f"[bench] slab share {n}x{n}x34: fp64 {shard['fp64']['value']:.0f} MLUPS, "
        f"fp32s {shard['fp32s']['value']:.0f} MLUPS")
    tgv = tgv_same = None
    if not args.no_all_fluid:
        for m in [int(v) for v in args.all_fluid_sizes.split(",") if v]:
            # at the porous channel's size the same kernels without solid surfaces separate the
            # lattice size from the line waste at the surfaces
            r = run_taylor_green(lbm, wl, n=m, steps=args.steps if m == n else 200)
            if m == n:
                tgv_same = r
            else:
                tgv = r
            log(f"[bench] Taylor-Green {m}^3: fp64 {r['fp64']['value']:.0f} MLUPS, fp32s {r['fp32s']['value']:.0f} MLUPS")
    amr = run_block_store(lbm)
    log(f"[bench] block store {amr['workload']}: " + ", ".join(
        f"{k} {amr[k]['ms']:.3f} ms ({amr[k]['GBs']:.0f} GB/s)" for k in ("find_overlap", "prolongate", "restrict")))
    md = run_md_coupling(lbm, solid, n, args.tau)
    log(f"[bench] MD coupling, {md['coupled_cells']} cells: step {md['step_ms']:.3f} -> {md['coupled_step_ms']:.3f} ms, "
        f"export {md['export_ms']:.3f} ms")
    blk = run_block_solver(lbm, sizes=block_sizes, steps=args.block_steps)
    vex = run_vertex_exchange(lbm)
    adj = None
    if not args.no_adjoint:
        adj = run_adjoint(lbm, wl)
        log(f"[bench] adjoint: {adj['adjoint_over_forward_step']:.2f}x forward step, "
            f"gradient {adj['gradient_over_forward_run']:.2f}x forward run")
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(wl)
        log(f"[bench] cpu baseline: {cpu['value']} MLUPS on {cpu['cores']} cores")
    r64, r32 = res["fp64"], res["fp32s"]
    line = {
        "metric": METRIC, "value": round(r64["mlups"], 1), "unit": "MLUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(r64["ms_per_step"], 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"porous channel {n}^3 periodic, Guo g=1e-6 along x, random spheres r=4..12 "
                               f"(seed {wl.DEFAULT_SEED}) to porosity 0.8, tau={args.tau}, pull-form fused "
                               f"collide+stream", "cells": n ** 3, "fluid_cells": r64["fluid_cells"],
                   "mlups_counts": "fluid cells",
                   "step_variant": {"fp64": r64["variant"], "fp32s": r32["variant"]},
                   "cell_order": {"fp64": r64["cell_order"], "fp32s": r32["cell_order"]},
                   "workgroup_order": {"fp64": r64["workgroup_order"], "fp32s": r32["workgroup_order"]},
                   "workgroup_rows": {"fp64": r64["workgroup_rows"], "fp32s": r32["workgroup_rows"]}},
        "roofline": r64["roofline"],
        "fp32_shifted": {"value": round(r32["mlups"], 1), "unit": "MLUPS", "dtype": "f32",
                         "arithmetic": "FP32 state and traffic; FP64 moments, equilibrium and relaxation product",
                         "ms_per_step": round(r32["ms_per_step"], 4), "roofline": r32["roofline"]},
        "fp32_shifted_f32_arithmetic": {"value": round(res["fp32s_f32math"]["mlups"], 1), "unit": "MLUPS",
                                        "dtype": "f32", "ms_per_step": round(res["fp32s_f32math"]["ms_per_step"], 4),
                                        "roofline": res["fp32s_f32math"]["roofline"]},
        "aa_in_place": {k: {"value": round(res[k]["mlups"], 1), "unit": "MLUPS",
                            "ms_per_step": round(res[k]["ms_per_step"], 4), "roofline": res[k]["roofline"],
                            "cell_order": res[k]["cell_order"], "workgroup_rows": res[k]["workgroup_rows"],
                            "state_bytes": 19 * n ** 3 * (8 if k == "fp64_aa" else 4)}
                        for k in ("fp64_aa", "fp32s_aa")},
        "wall_mlups": {"fp64": round(r64["wall_mlups"], 1), "fp32s": round(r32["wall_mlups"], 1)},
        "host_readout_fp64": r64["host_readout"],
        "batched_patches": patches,
        "md_coupling": md,
        "all_fluid": tgv,
        "all_fluid_same_size": tgv_same,
        "copy_reference": {"fp64": r64["copy_reference"], "fp32s": r32["copy_reference"],
                           "what": f"19-plane copy and pull-pattern copy of all {n}^3 cells on the step's buffers "
                                   f"and workgroup mapping (lbm_measure_copy), GB/s = cells x 38 x sizeof(T) / time"},
        "slab_share": shard,
        "adjoint": adj,
        "restriction_prolongation": amr,
        "block_solver": blk,
        "vertex_exchange": vex,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
