"""The reference-shaped C++ API (madpac_amd/csrc/peano_blocklb.h) driven from a C++ program:
runner -> repository -> iterate, block-store view; results bit-identical to the reference fixtures."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, channel_tolerance, golden

pytestmark = pytest.mark.gpu
EXE = os.path.join(ROOT, "tests", "cpp", "test_peano_api")


@pytest.mark.parametrize("layout", ["two-buffer", "in-place"])
@pytest.mark.parametrize("name", ["cavity_n13_t5.npz", "cavity_n8_t100.npz"])
def test_cpp_runner_matches_reference(name, layout, tmp_path):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (run __graft_entry__.build())")
    d = golden(name)
    out = tmp_path / "f.bin"
    print(f"[gpu] C++ API cavity {name} {layout}", flush=True)
    n = str(int(d["n"]))
    env = dict(os.environ, LBM_TEST_LAYOUT=layout)
    r = subprocess.run([EXE, "cavity", n, n, n, repr(float(d["tau"])), repr(float(d["ulid"])), str(int(d["steps"])),
                        str(out)], capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    f = np.fromfile(out, dtype="<f8")
    assert np.array_equal(f, d["f"].reshape(-1))


@pytest.mark.parametrize("name", ["channel_16x8x8_t200_parabolic.npz", "channel_10x7x9_t13_parabolic.npz"])
def test_cpp_runner_channel_matches_reference(name, tmp_path):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (run __graft_entry__.build())")
    d = golden(name)
    out = tmp_path / "f.bin"
    print(f"[gpu] C++ API channel {name}", flush=True)
    scen = "channel-parabolic" if int(d["parabolic"]) else "channel"
    r = subprocess.run([EXE, scen] + [str(int(v)) for v in d["dims"]] +
                       [repr(float(d["tau"])), repr(float(d["umean"])), str(int(d["steps"])), str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    f = np.fromfile(out, dtype="<f8")
    ref = d["f"].reshape(-1)
    tol = channel_tolerance(d)
    assert np.max(np.abs(f - ref)) <= tol * np.max(np.abs(ref))


@pytest.mark.parametrize("name", ["vtk_cavity_8x8x8_t3_every1.npz", "vtk_channel_10x7x9_t6_every2.npz"])
def test_cpp_runner_vtk_output_matches_reference(name, tmp_path):
    # the runner's plotter path end to end (triggerPlotting, VTK adapter, the moments the block
    # store holds when the output mapping runs) against the reference's own VTK files; bars as in
    # tests/test_vtk.py
    from vtk_io import assert_same_vtk, fixture_file, read_vtk
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (run __graft_entry__.build())")
    d = golden(name)
    print(f"[gpu] C++ API VTK output {name}", flush=True)
    scen = "cavity" if str(d["scenario"]) == "cavity" else "channel-parabolic"
    base = str(tmp_path / "mine")
    r = subprocess.run([EXE, scen] + [str(int(v)) for v in d["dims"]] +
                       [repr(float(d["tau"])), repr(float(d["u"])), str(int(d["steps"])), str(tmp_path / "f.bin"),
                        base, str(int(d["every"]))], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    written = sorted(int(p.name[5:-4]) for p in tmp_path.glob("mine_*.vtk"))
    assert written == [int(k) for k in d["plotted"]]
    exact = scen == "cavity"
    for k in d["plotted"]:
        assert_same_vtk(read_vtk(f"{base}_{int(k)}.vtk"), fixture_file(d, k), [str(h) for h in d["headers"]],
                        0.0 if exact else 2e-6, 0.0 if exact else 1e-15)


# ---- builtin obstacles, mapping events and writes through the block store (test_peano_events)
EVENTS = os.path.join(ROOT, "tests", "cpp", "test_peano_events")
GEOMETRY = ["geometry_cavity_sphere_n12_t40.npz", "geometry_channel_16x8x8_t60.npz"]


def _object_args(d):
    args = []
    for k, inv, p in zip(d["object_kind"], d["object_inverted"], d["object_params"]):
        args += [str(int(k)), str(int(inv))] + [repr(float(v)) for v in p]
    return args


def _scenario(d):
    if str(d["scenario"]) == "lid-driven-cavity":
        return "cavity"
    return "channel-parabolic" if int(d["parabolic"]) else "channel"


def _run_events(d, mode, tmp_path, layout="two-buffer", steps=None):
    if not os.path.exists(EVENTS):
        pytest.fail(f"{EVENTS} is not built (run __graft_entry__.build())")
    out = tmp_path / f"f_{mode.replace(':', '_')}_{layout}.bin"
    steps = int(d["steps"]) if steps is None else steps
    env = dict(os.environ, LBM_TEST_LAYOUT=layout)
    r = subprocess.run([EVENTS, _scenario(d)] + [str(int(v)) for v in d["dims"]] +
                       [repr(float(d["tau"])), repr(float(d["u"])), str(steps), str(out), mode] + _object_args(d),
                       capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    raw = out.read_bytes()
    n = int(np.prod(d["dims"]))
    f = np.frombuffer(raw[:8 * 19 * n], "<f8")
    o = 8 * 19 * n
    s = np.frombuffer(raw[o:o + 12], np.int32)
    m = int(np.prod(s))
    inner = np.frombuffer(raw[o + 12:o + 12 + m], np.uint8).reshape(s[2], s[1], s[0])
    boundary = np.frombuffer(raw[o + 12 + m:o + 12 + 2 * m], np.uint8).reshape(s[2], s[1], s[0])
    return f, inner, boundary


@pytest.mark.parametrize("layout", ["two-buffer", "in-place"])
@pytest.mark.parametrize("mode", ["plain", "roundtrip"])
@pytest.mark.parametrize("name", GEOMETRY)
def test_cpp_runner_with_obstacles_matches_reference(name, mode, layout, tmp_path):
    # builtin sphere / hexahedron obstacles through the C++ configuration: the block store's flags and
    # the populations are the reference's bit for bit, with or without every mapping event firing
    # and every block's pdf written back through GridManagementService
    d = golden(name)
    print(f"[gpu] C++ API obstacles {name} {mode} {layout}", flush=True)
    f, inner, boundary = _run_events(d, mode, tmp_path, layout)
    assert np.array_equal(inner, d["inner"]) and np.array_equal(boundary, d["boundary"])
    assert np.array_equal(f, d["f"].reshape(-1)), np.abs(f - d["f"].reshape(-1)).max()


def _oracle_for(d):
    import oracle as O
    nx, ny, nz = (int(v) for v in d["dims"])
    solid = (d["inner"][3:3 + nz, 3:3 + ny, 3:3 + nx] == 0).astype(np.uint8)
    if _scenario(d) == "cavity":
        return O.Lattice.cavity(nx, float(d["tau"]), float(d["u"]), solid=solid), solid
    return O.Lattice.channel(nx, ny, nz, float(d["tau"]), float(d["u"]), parabolic=bool(int(d["parabolic"])),
                             solid=solid), solid


@pytest.mark.parametrize("layout", ["two-buffer", "in-place"])
def test_block_store_writes_reach_the_device(layout, tmp_path):
    # touchVertexFirstTime of iteration T scales the populations of one cell through getPdf(); the
    # oracle applies the same change to its state after T steps
    import oracle as O
    d = golden(GEOMETRY[0])
    n, T, steps = 12, 7, 15
    fluid = np.flatnonzero(d["inner"][3:15, 3:15, 3:15].reshape(-1))
    cell = int(fluid[len(fluid) // 3])
    print(f"[gpu] block-store write: cell {cell} at iteration {T}, {layout}", flush=True)
    f, _, _ = _run_events(d, f"perturb:{T}:{cell}", tmp_path, layout, steps=steps)
    lat, _ = _oracle_for(d)
    fo = lat.step(lat.equilibrium(), T).reshape(O.Q, -1)
    fo[:, cell] *= 1.0 + np.arange(O.Q) / 1000.0
    fo = lat.step(fo.reshape(-1), steps - T)
    assert np.array_equal(f, fo), np.abs(f - fo).max()
    plain, _, _ = _run_events(d, "plain", tmp_path, layout, steps=steps)
    assert not np.array_equal(f, plain)


def test_block_store_inner_flag_write_reflags_the_lattice(tmp_path):
    # a fluid cell next to the sphere made non-inner through getInner() at iteration T: from then on
    # it is solid (populations dropped, its links bounce back), as in the oracle re-flagged at T
    import oracle as O
    d = golden(GEOMETRY[0])
    n, T, steps = 12, 5, 12
    lat, solid = _oracle_for(d)
    fluid = np.flatnonzero(solid.reshape(-1) == 0)
    cell = int(fluid[len(fluid) // 2])
    print(f"[gpu] block-store re-flag: cell {cell} at iteration {T}", flush=True)
    f, inner, _ = _run_events(d, f"solidify:{T}:{cell}", tmp_path, steps=steps)
    fo = lat.step(lat.equilibrium(), T)
    solid2 = solid.reshape(-1).copy()
    solid2[cell] = 1
    lat2 = O.Lattice.cavity(n, float(d["tau"]), float(d["u"]), solid=solid2)
    fo = fo.reshape(O.Q, -1)
    fo[:, cell] = 0.0
    fo = lat2.step(fo.reshape(-1), steps - T)
    assert np.array_equal(f, fo), np.abs(f - fo).max()
    z, y, x = cell // (n * n), (cell // n) % n, cell % n
    assert inner[z + 3, y + 3, x + 3] == 0
