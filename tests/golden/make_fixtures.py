"""Generate the golden fixtures from the compiled reference (run in the build container only).

The reference Peano block-LB is compiled from its own sources by ``oracle/ref/Makefile``
into ``oracle/_ref/`` (never committed, never shipped).  This script runs it and stores
its outputs as small ``.npz`` fixtures next to this file:

* ``cellkernels_s{seed}.npz`` -- one 6^3 block of seeded non-equilibrium populations
  through the reference's computeDensity / computeVelocity / computeFeq / BGK collide /
  moving wall / local swap streaming (``oracle/ref/cellkernels.cpp``).
* ``cavity_n{N}_t{steps}.npz`` -- the full FP64 population field of the lid-driven cavity
  (N^3 fluid cells) after ``steps`` time steps of the reference time-step driver
  (``oracle/ref/harness.cpp``), as SoA f[q, z, y, x] over the fluid box.
* ``cavity_n64_t1000.npz`` -- the 64^3 cavity (tau 0.6, lid 0.05) after 1000 steps:
  SHA-256 of the full SoA population field, the rho/u mid-planes and centre lines, and
  global mass and momentum.
* ``channel_*.npz`` -- the reference's default channel (no-slip y/z walls, velocity inlet on
  x-, pressure outlet on x+; uniform or parabolic inlet profile), full FP64 populations over
  the fluid box after ``steps`` time steps.
* ``geometry_*.npz`` -- runs with builtin obstacles in the ``<builtin-geometry-configuration>``
  tag (an inverted ``<sphere>``, an inverted ``<hexahedron>``) besides the domain box: the whole
  block store's ``inner`` and ``boundary`` flags (peano/geometry/builtin/Sphere.cpp,
  Hexahedron.cpp through GridManagementService::createVertex) and the populations over the box
  after ``steps`` time steps (0 in non-inner cells).
* ``vtk_*.npz`` -- the legacy-VTK field files the reference writes with its plotter tag
  (``<plotter vtk-outputname=... timesteps-per-plotting=...>``), read back with
  ``tests/vtk_io.py``: per plotted iteration the points, voxels, cell types, density, velocity
  and block ids, plus the section header lines.

Usage:  python tests/golden/make_fixtures.py [--skip-64] [--only-vtk] [--only-geometry]
"""
from __future__ import annotations

import argparse
import hashlib
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
REF_DIR = os.path.join(ROOT, "oracle", "_ref")
HARNESS = os.path.join(REF_DIR, "peano_lb_harness")
CELLK = os.path.join(REF_DIR, "peano_lb_cellkernels")

# Cavity XML for the reference's regular-grid block-LB batch job.  Tags and attributes are the
# ones its configuration parsers read (blocklatticeboltzmann/configurations/
# BlockLatticeBoltzmannBatchJobConfigurationForRegularGrid.cpp:41-91).
# Scenario attributes: lid-driven-cavity takes none beyond the velocity; default-channel also
# needs use-parabolic-profile (SLB/configurations/BlockLatticeBoltzmannScenarioConfiguration.cpp:52-109).
XML = """<?xml version="1.0"?>
<peano-configuration>
  <run-blocklatticeboltzmann-on-regular-grid>
    <log-output column-separator=" " log-time-stamp="no" log-time-stamp-human-readable="no" log-machine-name="no" log-message-type="yes" log-trace="yes" />
    <regular-grid number-of-vertices="{nx};{ny};{nz}" domain-size="{lx!r};{ly!r};{lz!r}" domain-offset="0.0;0.0;0.0" />
    <lattice-boltzmann-configuration tau="{tau!r}" dt="1.0" density="1.0" number-of-timesteps="{steps}" measure-performance="no" />
    <lattice-boltzmann-scenario scenario="{scenario}" characteristic-velocity_L="{ulid!r}"{extra} dynamic-geometry="no" compute-forces="no"></lattice-boltzmann-scenario>
    <collision-model name="BGK" hMin="0.00001" hMax="10.0" />
    <builtin-geometry-configuration>
      <hexahedron geometry-id="0" invert="false" width="{lx!r};{ly!r};{lz!r}" offset="0.0;0.0;0.0" />{objects}
    </builtin-geometry-configuration>
  </run-blocklatticeboltzmann-on-regular-grid>
</peano-configuration>
"""


def blocks_per_dim(n: int) -> int:
    # BLB/runners/BlockLatticeBoltzmannBatchJobRunnerForRegularGrid.cpp:78-92
    b = n // 6 + 1
    if n % 6:
        b += 1
    return b


def read_dump(path):
    raw = open(path, "rb").read()
    hdr = np.frombuffer(raw[:16], dtype="<i4")
    assert hdr[0] == 0x4C42524D and hdr[2] == 216 and hdr[3] == 19
    nb = int(hdr[1])
    rec = 216 * 2 + 8 * (216 + 648 + 4104)
    out = []
    for k in range(nb):
        o = 16 + k * rec
        inner = np.frombuffer(raw[o:o + 216], dtype=np.uint8)
        f = np.frombuffer(raw[o + 432 + 8 * 864:o + rec], dtype="<f8")
        out.append((inner, f))
    return out


def run_reference(dims, tau, ulid, steps, scenario="lid-driven-cavity", parabolic=None):
    """Run the reference on an nx x ny x nz box of unit cells (domain size n_d / ny) and return
    the SoA populations over the fluid box, [q, z, y, x]."""
    nx, ny, nz = dims
    nb = [blocks_per_dim(n) for n in dims]
    extra = "" if parabolic is None else f' use-parabolic-profile="{"yes" if parabolic else "no"}"'
    with tempfile.TemporaryDirectory() as td:
        xml = os.path.join(td, "run.xml")
        with open(xml, "w") as fh:
            fh.write(XML.format(nx=nx, ny=ny, nz=nz, lx=nx / ny, ly=1.0, lz=nz / ny, tau=tau, ulid=ulid,
                                steps=steps, scenario=scenario, extra=extra, objects=""))
        out = os.path.join(td, "dump.bin")
        subprocess.run([HARNESS, xml, str(nb[0] * nb[1] * nb[2]), out], check=True, cwd=td,
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        blocks = read_dump(out)
    dense = np.zeros((6 * nb[2], 6 * nb[1], 6 * nb[0], 19))
    inner = np.zeros(dense.shape[:3], np.uint8)
    # blocks are created in lexicographic vertex order, x fastest
    for k, (inn, f) in enumerate(blocks):
        bx, by, bz = k % nb[0], (k // nb[0]) % nb[1], k // (nb[0] * nb[1])
        sl = (slice(6 * bz, 6 * bz + 6), slice(6 * by, 6 * by + 6), slice(6 * bx, 6 * bx + 6))
        dense[sl] = f.reshape(6, 6, 6, 19)
        inner[sl] = inn.reshape(6, 6, 6)
    # fluid cell centres (s - 2.5) dx lie in the domain for slots s = 3 .. n+2
    box = (slice(3, 3 + nz), slice(3, 3 + ny), slice(3, 3 + nx))
    assert inner[box].all() and int(inner.sum()) == nx * ny * nz, "unexpected flag layout"
    outside = np.where(inner[..., None] == 0, dense, 0.0)
    assert not outside.any(), "non-fluid cells must keep zero populations"
    return np.ascontiguousarray(np.moveaxis(dense[box], -1, 0))  # [q, z, y, x]


def _vec(v):
    return ";".join(repr(float(x)) for x in v)


def object_xml(obj):
    """One builtin geometry object as the reference's configuration reads it
    (peano/geometry/builtin/configurations/SphereConfiguration.cpp, HexahedronConfiguration.cpp)."""
    kind, gid, invert = obj[0], obj[1], "true" if obj[2] else "false"
    if kind == "sphere":
        return (f'\n      <sphere geometry-id="{gid}" invert="{invert}" center="{_vec(obj[3])}" '
                f'radius="{float(obj[4])!r}" />')
    return (f'\n      <hexahedron geometry-id="{gid}" invert="{invert}" width="{_vec(obj[4])}" '
            f'offset="{_vec(obj[3])}" />')


def run_reference_geometry(dims, tau, u, steps, objects, scenario="lid-driven-cavity", parabolic=None):
    """Run the reference with extra builtin geometry objects; return (f over the box [q, z, y, x],
    inner and boundary flags of the whole block store [6 nbz, 6 nby, 6 nbx])."""
    nx, ny, nz = dims
    nb = [blocks_per_dim(n) for n in dims]
    extra = "" if parabolic is None else f' use-parabolic-profile="{"yes" if parabolic else "no"}"'
    xml = XML.format(nx=nx, ny=ny, nz=nz, lx=nx / ny, ly=1.0, lz=nz / ny, tau=tau, ulid=u, steps=steps,
                     scenario=scenario, extra=extra, objects="".join(object_xml(o) for o in objects))
    with tempfile.TemporaryDirectory() as td:
        with open(os.path.join(td, "run.xml"), "w") as fh:
            fh.write(xml)
        subprocess.run([HARNESS, "run.xml", str(nb[0] * nb[1] * nb[2]), "dump.bin"], check=True, cwd=td,
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        raw = open(os.path.join(td, "dump.bin"), "rb").read()
    rec = 216 * 2 + 8 * (216 + 648 + 4104)
    shape = (6 * nb[2], 6 * nb[1], 6 * nb[0])
    dense = np.zeros(shape + (19,))
    inner = np.zeros(shape, np.uint8)
    boundary = np.zeros(shape, np.uint8)
    for k in range(nb[0] * nb[1] * nb[2]):
        o = 16 + k * rec
        bx, by, bz = k % nb[0], (k // nb[0]) % nb[1], k // (nb[0] * nb[1])
        sl = (slice(6 * bz, 6 * bz + 6), slice(6 * by, 6 * by + 6), slice(6 * bx, 6 * bx + 6))
        inner[sl] = np.frombuffer(raw[o:o + 216], np.uint8).reshape(6, 6, 6)
        boundary[sl] = np.frombuffer(raw[o + 216:o + 432], np.uint8).reshape(6, 6, 6)
        dense[sl] = np.frombuffer(raw[o + 432 + 8 * 864:o + rec], "<f8").reshape(6, 6, 6, 19)
    box = (slice(3, 3 + nz), slice(3, 3 + ny), slice(3, 3 + nx))
    outside = inner.copy()
    outside[box] = 0
    assert not outside.any(), "inner cells outside the box"
    assert not np.where(inner[..., None] == 0, dense, 0.0).any(), "non-inner cells must keep zero populations"
    f = np.ascontiguousarray(np.moveaxis(dense[box], -1, 0))
    return f, inner, boundary


def save_geometry_fixture(name, dims, tau, u, steps, objects, scenario, parabolic):
    f, inner, boundary = run_reference_geometry(dims, tau, u, steps, objects, scenario, parabolic)
    kinds = np.array([0 if o[0] == "hexahedron" else 1 for o in objects], np.int64)
    inv = np.array([int(o[2]) for o in objects], np.int64)
    # per object: hexahedron offset[3], width[3]; sphere centre[3], radius, 0, 0
    params = np.array([list(o[3]) + (list(o[4]) if o[0] == "hexahedron" else [float(o[4]), 0.0, 0.0])
                       for o in objects], np.float64)
    np.savez_compressed(os.path.join(HERE, name), f=f, inner=inner, boundary=boundary,
                        dims=np.array(dims, np.int64), tau=np.float64(tau), u=np.float64(u), steps=np.int64(steps),
                        scenario=np.array(scenario), parabolic=np.int64(bool(parabolic)),
                        object_kind=kinds, object_inverted=inv, object_params=params)
    print("geometry", name, int(inner.sum()), "inner cells", file=sys.stderr)


# Obstacle cases: (file, dims, tau, velocity, steps, objects, scenario, parabolic).  The domain box
# (domain size n_d / n_y) is always the first object.  Sphere radii and centres keep every cell
# centre and voxel corner at least 1e-3 dx away from a surface.
GEOMETRY_CASES = (
    ("geometry_cavity_sphere_n12_t40.npz", (12, 12, 12), 0.6, 0.08, 40,
     [("sphere", 1, True, (0.45, 0.52, 0.40), 0.26)], "lid-driven-cavity", None),
    ("geometry_channel_16x8x8_t60.npz", (16, 8, 8), 0.7, 0.04, 60,
     [("sphere", 1, True, (0.61, 0.47, 0.53), 0.23),
      ("hexahedron", 2, True, (1.25, 0.125, 0.25), (0.25, 0.5, 0.375))], "default-channel", True),
)


def run_reference_vtk(dims, tau, u, steps, every, scenario="lid-driven-cavity", parabolic=None):
    """Run the reference with VTK plotting every `every` iterations; return {iteration: parsed file}."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from vtk_io import read_vtk
    nx, ny, nz = dims
    nb = [blocks_per_dim(n) for n in dims]
    extra = "" if parabolic is None else f' use-parabolic-profile="{"yes" if parabolic else "no"}"'
    xml = XML.format(nx=nx, ny=ny, nz=nz, lx=nx / ny, ly=1.0, lz=nz / ny, tau=tau, ulid=u, steps=steps,
                     scenario=scenario, extra=extra, objects="")
    tag = f'    <plotter vtk-outputname="ref" timesteps-per-plotting="{every}" />\n'
    xml = xml.replace("    <collision-model", tag + "    <collision-model")
    out = {}
    with tempfile.TemporaryDirectory() as td:
        with open(os.path.join(td, "run.xml"), "w") as fh:
            fh.write(xml)
        subprocess.run([HARNESS, "run.xml", str(nb[0] * nb[1] * nb[2]), "dump.bin"], check=True, cwd=td,
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        for k in range(0, steps, every):
            path = os.path.join(td, f"ref_{k}.vtk")
            d = read_vtk(path)
            d["first_line"] = open(path).readline().rstrip("\n")
            out[k] = d
    return out


def save_vtk_fixture(name, files, **meta):
    arrays = {}
    for k, d in files.items():
        for key in ("points", "cells", "types", "density", "velocity", "block-id"):
            arrays[f"{key.replace('-', '_')}_{k}"] = d[key]
    first = next(iter(files.values()))
    arrays["headers"] = np.array([first["first_line"]] + first["headers"])
    arrays["plotted"] = np.array(sorted(files), np.int64)
    np.savez_compressed(os.path.join(HERE, name), **arrays, **meta)


def run_cavity(n, tau, ulid, steps):
    return run_reference((n, n, n), tau, ulid, steps)


def cellkernel_fixture(seed, tau):
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "ck.bin")
        subprocess.run([CELLK, str(seed), repr(tau), out], check=True)
        raw = open(out, "rb").read()
    o = 0

    def take(dt, n):
        nonlocal o
        a = np.frombuffer(raw[o:o + np.dtype(dt).itemsize * n], dtype=dt)
        o += np.dtype(dt).itemsize * n
        return a.copy()

    magic, sd = take("<i4", 2)
    assert magic == 0x4C42434B and sd == seed
    d = {"seed": np.int64(seed), "tau": np.float64(tau)}
    d["tau_eff"] = take("<f8", 1)[0]
    d["uw"] = take("<f8", 3)
    d["inner"] = take(np.uint8, 216)
    d["pdf_in"] = take("<f8", 216 * 19)
    d["density"] = take("<f8", 216)
    d["velocity"] = take("<f8", 648)
    d["feq"] = take("<f8", 216 * 19)
    d["pdf_collided"] = take("<f8", 216 * 19)
    nl = int(take("<i4", 1)[0])
    d["wall_links"] = take("<i4", 2 * nl).reshape(nl, 2)
    d["pdf_after_wall"] = take("<f8", 216 * 19)
    d["pdf_after_local_stream"] = take("<f8", 216 * 19)
    assert o == len(raw)
    return d


def field_digest(f_soa):
    return hashlib.sha256(np.ascontiguousarray(f_soa, dtype="<f8").tobytes()).hexdigest()


def moments(f):
    from_c = np.array([[0, -1, -1], [-1, 0, -1], [0, 0, -1], [1, 0, -1], [0, 1, -1], [-1, -1, 0], [0, -1, 0],
                       [1, -1, 0], [-1, 0, 0], [0, 0, 0], [1, 0, 0], [-1, 1, 0], [0, 1, 0], [1, 1, 0],
                       [0, -1, 1], [-1, 0, 1], [0, 0, 1], [1, 0, 1], [0, 1, 1]], dtype=np.float64)
    rho = f.sum(axis=0)
    j = np.tensordot(from_c.T, f, axes=([1], [0]))
    return rho, j / rho


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--skip-64", action="store_true")
    ap.add_argument("--only-vtk", action="store_true")
    ap.add_argument("--only-geometry", action="store_true")
    args = ap.parse_args()
    subprocess.run(["make", "-s", "-f", os.path.join(ROOT, "oracle", "ref", "Makefile"), "-j8"], check=True)

    for case in GEOMETRY_CASES:
        save_geometry_fixture(*case)
    if args.only_geometry:
        return

    # VTK field output: cavity 8^3 every iteration; parabolic channel 10x7x9 every second one
    files = run_reference_vtk((8, 8, 8), 0.6, 0.05, 3, 1)
    save_vtk_fixture("vtk_cavity_8x8x8_t3_every1.npz", files, dims=np.array((8, 8, 8), np.int64), tau=np.float64(0.6),
                     u=np.float64(0.05), steps=np.int64(3), every=np.int64(1), scenario=np.array("cavity"))
    files = run_reference_vtk((10, 7, 9), 0.55, 0.03, 6, 2, scenario="default-channel", parabolic=True)
    save_vtk_fixture("vtk_channel_10x7x9_t6_every2.npz", files, dims=np.array((10, 7, 9), np.int64),
                     tau=np.float64(0.55), u=np.float64(0.03), steps=np.int64(6), every=np.int64(2),
                     scenario=np.array("channel-parabolic"))
    print("vtk", file=sys.stderr)
    if args.only_vtk:
        return

    for seed, tau in ((1, 0.6), (2, 0.93)):
        d = cellkernel_fixture(seed, tau)
        np.savez_compressed(os.path.join(HERE, f"cellkernels_s{seed}.npz"), **d)
        print("cellkernels", seed, file=sys.stderr)

    for n, tau, ulid, steps in ((8, 0.6, 0.05, 1), (8, 0.6, 0.05, 10), (8, 0.6, 0.05, 100),
                                (10, 0.8, 0.1, 37), (13, 0.55, 0.08, 5)):
        f = run_cavity(n, tau, ulid, steps)
        np.savez_compressed(os.path.join(HERE, f"cavity_n{n}_t{steps}.npz"), f=f, n=np.int64(n),
                            tau=np.float64(tau), ulid=np.float64(ulid), steps=np.int64(steps))
        print("cavity", n, steps, file=sys.stderr)

    # default channel: (nx, ny, nz), tau, u_mean, parabolic, steps
    for dims, tau, um, par, steps in (((8, 8, 8), 0.6, 0.05, False, 10), ((16, 8, 8), 0.7, 0.04, True, 200),
                                      ((10, 7, 9), 0.55, 0.03, True, 13)):
        f = run_reference(dims, tau, um, steps, scenario="default-channel", parabolic=par)
        name = f"channel_{dims[0]}x{dims[1]}x{dims[2]}_t{steps}_{'parabolic' if par else 'uniform'}.npz"
        np.savez_compressed(os.path.join(HERE, name), f=f, dims=np.array(dims, np.int64), tau=np.float64(tau),
                            umean=np.float64(um), parabolic=np.int64(par), steps=np.int64(steps))
        print("channel", dims, steps, file=sys.stderr)

    if not args.skip_64:
        n, tau, ulid, steps = 64, 0.6, 0.05, 1000
        f = run_cavity(n, tau, ulid, steps)
        rho, u = moments(f)
        h = n // 2
        np.savez_compressed(
            os.path.join(HERE, f"cavity_n{n}_t{steps}.npz"),
            n=np.int64(n), tau=np.float64(tau), ulid=np.float64(ulid), steps=np.int64(steps),
            sha256=np.array(field_digest(f)),
            rho_mid_y=rho[:, h, :], u_mid_y=u[:, :, h, :],
            rho_line_z=rho[:, h, h], u_line_z=u[:, :, h, h], u_line_x=u[:, h, h, :],
            mass=np.float64(f.sum()))
        print("cavity", n, steps, file=sys.stderr)


if __name__ == "__main__":
    main()
