"""The reference-shaped C++ API (madpac_amd/csrc/peano_blocklb.h) driven from a C++ program:
runner -> repository -> iterate, block-store view; results bit-identical to the reference fixtures."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, channel_tolerance, golden

pytestmark = pytest.mark.gpu
EXE = os.path.join(ROOT, "tests", "cpp", "test_peano_api")


@pytest.mark.parametrize("name", ["cavity_n13_t5.npz", "cavity_n8_t100.npz"])
def test_cpp_runner_matches_reference(name, tmp_path):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (run __graft_entry__.build())")
    d = golden(name)
    out = tmp_path / "f.bin"
    print(f"[gpu] C++ API cavity {name}", flush=True)
    n = str(int(d["n"]))
    r = subprocess.run([EXE, "cavity", n, n, n, repr(float(d["tau"])), repr(float(d["ulid"])), str(int(d["steps"])),
                        str(out)], capture_output=True, text=True, timeout=120)
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
