"""The C-ABI library on a host without a GPU: it loads, exports every function include/*.h declares,
and rejects invalid arguments with an error code and message before touching a device."""
import ctypes
import glob
import os
import re

import pytest

from conftest import ROOT

lbm = pytest.importorskip("madpac_amd.lbm")


def _declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(lbm_[a-z0-9_]+)\s*\(", text, re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_function():
    L = lbm.load_library()
    declared = _declared_functions()
    assert len(declared) >= 25
    missing = sorted(n for n in declared if not hasattr(L, n))
    assert not missing, missing
    # the Python binding knows them all too
    assert declared == set(lbm.EXPORTS)


def test_create_rejects_bad_arguments_without_a_device():
    L = lbm.load_library()
    h = ctypes.c_void_p()
    cfg = lbm.Config(0, 4, 4, (ctypes.c_int * 3)(0, 0, 0), 0, 0.7, (ctypes.c_double * 3)(0, 0, 0), 0)
    assert L.lbm_create(ctypes.byref(h), ctypes.byref(cfg)) == -1
    assert b"extent" in L.lbm_last_error(None)
    cfg.nx, cfg.tau = 4, 0.4
    assert L.lbm_create(ctypes.byref(h), ctypes.byref(cfg)) == -1
    assert b"tau" in L.lbm_last_error(None)
    cfg.tau, cfg.precision = 0.7, 7
    assert L.lbm_create(ctypes.byref(h), ctypes.byref(cfg)) == -1
    assert b"precision" in L.lbm_last_error(None)
    cfg.precision = 0
    cfg.force[0] = float("nan")
    assert L.lbm_create(ctypes.byref(h), ctypes.byref(cfg)) == -1
    assert b"force" in L.lbm_last_error(None)
    assert L.lbm_create(None, ctypes.byref(cfg)) == -1


def test_patch_table_rejected_before_device_use():
    L = lbm.load_library()
    h = ctypes.c_void_p()
    arr = (lbm.Patch * 1)()
    arr[0].n = 8
    arr[0].tau = 0.7
    for q in range(19):
        arr[0].nbr[q] = 3  # out of range
    assert L.lbm_patches_create(ctypes.byref(h), arr, 1, 0, 0) == -1
    assert b"range" in L.lbm_patches_last_error(None)
    arr[0].nbr[0] = 0
    for q in range(19):
        arr[0].nbr[q] = 0
    arr[0].tau = 0.5
    assert L.lbm_patches_create(ctypes.byref(h), arr, 1, 0, 0) == -1
    assert b"tau" in L.lbm_patches_last_error(None)


def test_null_handles_are_rejected():
    L = lbm.load_library()
    assert L.lbm_step(None, 1) == -1
    assert L.lbm_destroy(None) == 0
    assert L.lbm_fluid_cells(None) == -1
    assert L.lbm_patches_step(None, 1, 0, None) == -1
    assert L.lbm_patches_destroy(None) == 0
    assert L.lbm_patches_get_fields(None, None, None) == -1
    assert L.lbm_patches_get_populations(None, None) == -1
    assert L.lbm_get_fields(None, None, None) == -1
    assert L.lbm_tune(None, 0, None) == -1


def test_binding_rejects_short_arrays_before_calling_the_library():
    # the C side reads n (or 3n, 19n) values from these pointers: a short array must not get there
    class Never:
        def __getattr__(self, name):
            raise AssertionError(f"{name} called with a wrongly sized array")

    lat = object.__new__(lbm.Lattice)
    lat._L, lat._h, lat.n, lat._ncouple = Never(), None, 27, 4
    import numpy as np
    for call in (lambda: lat.set_flags(np.ones(26)), lambda: lat.set_porosity(np.zeros(26)),
                 lambda: lat.set_wall_classes(np.zeros(28, np.uint8)), lambda: lat.init_equilibrium(rho=np.ones(26)),
                 lambda: lat.init_equilibrium(u=np.zeros(3 * 27 - 1)), lambda: lat.set_populations(np.zeros(19 * 26)),
                 lambda: lat.set_boundary_velocity(np.zeros(3 * 26)), lambda: lat.set_coupling_data(0, np.zeros(11)),
                 lambda: lat.set_coupling_data(0, np.zeros(12), rho=np.ones(3)),
                 lambda: lat.set_moving_wall_class(1, (0.1, 0.0), 0)):
        with pytest.raises(ValueError, match="expected"):
            call()
