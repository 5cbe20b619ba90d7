import os
import sys

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C-ABI on the device)")
    config.addinivalue_line("markers", "slow: longer CPU runs (still part of the default suite)")


@pytest.fixture(scope="session")
def golden_dir():
    return GOLDEN


def golden(name):
    import numpy as np
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def channel_tolerance(d):
    """Relative bar of a channel fixture: 0 (bitwise) unless the inlet is parabolic on a box whose
    side ratios are not powers of two.  The reference evaluates the parabolic profile at physical
    cell centres (domain size n_d / n_y), madpac_amd in lattice units; the two round alike when the
    ratios are powers of two and to a few ulps of the profile otherwise."""
    import numpy as np
    dims = [int(v) for v in d["dims"]]
    pow2 = all((v / dims[1]) == 2.0 ** round(np.log2(v / dims[1])) for v in dims)
    return 0.0 if (not int(d["parabolic"]) or pow2) else 1e-13
