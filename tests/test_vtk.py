"""The block VTK writer of the C++ layer (writeBlockVTK in madpac_amd/csrc/peano_blocklb.h)
against the files the reference writes with its plotter tag (fixtures tests/golden/vtk_*.npz,
BLB/mappings/RegularGrid2BlockVTKOutput.cpp:95-273).  Host only: the writer is fed the oracle's
fields of the state the reference plots in that iteration (the moments its previous iteration
computed: the state one step back, rho = 1 and u = 0 in the first iteration).

Bars: sections, voxel connectivity, cell order and block ids identical; density and velocity
identical as printed (six significant digits of the float value) on the cavity, whose oracle
state is bit-identical to the reference; on the 10x7x9 parabolic channel, whose inlet profile
rounds differently in the last bits (populations within 1e-13), within 2e-6 relative plus 1e-15
absolute (round-off-level velocities near 1e-17 print differently)."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from conftest import ROOT, golden
from vtk_io import assert_same_vtk, fixture_file, read_vtk

CASES = {
    "vtk_cavity_8x8x8_t3_every1.npz": (lambda: O.Lattice.cavity(8, 0.6, 0.05), 0.0, 0.0),
    "vtk_channel_10x7x9_t6_every2.npz": (lambda: O.Lattice.channel(10, 7, 9, 0.55, 0.03, parabolic=True),
                                         2e-6, 1e-15),
}


@pytest.fixture(scope="module")
def writer(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("vtk") / "vtk_writer")
    subprocess.run(["g++", "-std=c++17", "-O2", "-I" + os.path.join(ROOT, "include"),
                    "-I" + os.path.join(ROOT, "madpac_amd", "csrc"), os.path.join(ROOT, "tests", "cpp", "vtk_writer.cpp"),
                    "-o", exe], check=True)
    return exe


def plotted_fields(lat, steps, k):
    """rho, u the reference's block store holds when iteration k plots."""
    n = lat.n
    if k == 0:
        return np.ones(n), np.zeros((n, 3))
    f = lat.step(lat.equilibrium(), k - 1)
    return lat.fields(f)


@pytest.mark.parametrize("name", sorted(CASES))
def test_writer_matches_reference_vtk(name, writer, tmp_path):
    d = golden(name)
    make, rtol, atol = CASES[name]
    nx, ny, nz = (int(v) for v in d["dims"])
    lat = make()
    headers = [str(h) for h in d["headers"]]
    for k in d["plotted"]:
        rho, u = plotted_fields(lat, int(d["steps"]), int(k))
        fields = tmp_path / f"fields_{k}.bin"
        with open(fields, "wb") as fh:
            fh.write(np.ones(nx * ny * nz, np.uint8).tobytes())
            fh.write(np.ascontiguousarray(rho, "<f8").tobytes())
            fh.write(np.ascontiguousarray(u, "<f8").tobytes())
        out = tmp_path / f"mine_{k}.vtk"
        subprocess.run([writer, str(nx), str(ny), str(nz), repr(1.0 / ny), str(fields), str(out)], check=True)
        assert_same_vtk(read_vtk(str(out)), fixture_file(d, k), headers, rtol, atol)


def test_fixture_block_ids_follow_the_block_store_order():
    # the reference numbers blocks in its vertex order (x fastest) and counts empty ones too
    d = golden("vtk_channel_10x7x9_t6_every2.npz")
    nb = [n // 6 + 1 + (1 if n % 6 else 0) for n in (int(v) for v in d["dims"])]
    ref = fixture_file(d, 0)
    lo = ref["points"][ref["cells"]].min(axis=1)
    g = np.rint(lo * int(d["dims"][1])).astype(int)
    b = (g + 3) // 6
    assert np.array_equal(ref["block-id"], b[:, 0] + nb[0] * (b[:, 1] + nb[1] * b[:, 2]))
