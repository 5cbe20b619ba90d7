"""The block store's geometry flags (madpac_amd/csrc/peano_geometry.h) against the compiled
reference's, and the oracle against the reference's obstacle runs.

The fixtures tests/golden/geometry_*.npz come from the reference with builtin obstacles (an inverted
sphere, an inverted hexahedron) beside the domain box: the inner and boundary flag of every block-store
slot (padding blocks included) and the populations after the run.  Host code only."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from conftest import ROOT, golden

EXE = os.path.join(ROOT, "tests", "cpp", "test_geometry")
CASES = ["geometry_cavity_sphere_n12_t40.npz", "geometry_channel_16x8x8_t60.npz"]


def object_args(d):
    nx, ny, nz = (int(v) for v in d["dims"])
    args = ["0", "0", "0.0", "0.0", "0.0", repr(nx / ny), "1.0", repr(nz / ny)]  # the domain box
    for k, inv, p in zip(d["object_kind"], d["object_inverted"], d["object_params"]):
        args += [str(int(k)), str(int(inv))] + [repr(float(v)) for v in p]
    return args


def _exe(tmp_path):
    if os.path.exists(EXE):
        return EXE
    exe = str(tmp_path / "test_geometry")
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"), "-I",
                    os.path.join(ROOT, "madpac_amd", "csrc"), os.path.join(ROOT, "tests", "cpp", "test_geometry.cpp"),
                    "-o", exe], check=True)
    return exe


@pytest.mark.parametrize("name", CASES)
def test_block_store_flags_equal_the_reference(name, tmp_path):
    d = golden(name)
    nx, ny, nz = (int(v) for v in d["dims"])
    out = tmp_path / "flags.bin"
    subprocess.run([_exe(tmp_path), str(out), str(nx), str(ny), str(nz), repr((nx / ny) / nx), "0.0", "0.0", "0.0"]
                   + object_args(d), check=True)
    raw = out.read_bytes()
    s = np.frombuffer(raw[:12], np.int32)
    n = int(np.prod(s))
    inner = np.frombuffer(raw[12:12 + n], np.uint8).reshape(s[2], s[1], s[0])
    boundary = np.frombuffer(raw[12 + n:12 + 2 * n], np.uint8).reshape(s[2], s[1], s[0])
    assert np.array_equal(inner, d["inner"]), int((inner != d["inner"]).sum())
    assert np.array_equal(boundary, d["boundary"]), int((boundary != d["boundary"]).sum())
    # the obstacles remove cells: the fixture is not the empty box
    box = d["inner"][3:3 + nz, 3:3 + ny, 3:3 + nx]
    assert 0 < int((box == 0).sum()) < box.size // 4


def oracle_run(d):
    """The oracle on a fixture's flags (inner cells of the lattice) for its scenario."""
    nx, ny, nz = (int(v) for v in d["dims"])
    solid = (d["inner"][3:3 + nz, 3:3 + ny, 3:3 + nx] == 0).astype(np.uint8)
    if str(d["scenario"]) == "lid-driven-cavity":
        lat = O.Lattice.cavity(nx, float(d["tau"]), float(d["u"]), solid=solid)
    else:
        lat = O.Lattice.channel(nx, ny, nz, float(d["tau"]), float(d["u"]), parabolic=bool(int(d["parabolic"])),
                                solid=solid)
    return lat


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_reference_obstacle_runs(name):
    d = golden(name)
    lat = oracle_run(d)
    f = lat.step(lat.equilibrium(), int(d["steps"]))
    assert np.array_equal(f, d["f"].reshape(-1))
