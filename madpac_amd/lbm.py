"""ctypes binding of the C-ABI in include/lbm_capi.h (liblbm_hip.so).

This is the Python face of the solver used by the tests and the benchmark: every call goes
through the C-ABI to the HIP kernels.  There is no CPU fallback -- if the shared library is
missing or cannot be loaded this module raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblbm_hip.so")

Q = 19
FP64 = 0
FP32_SHIFTED = 1       # FP32 state; FP64 moments, equilibrium and relaxation product
FP32_SHIFTED_F32 = 2   # FP32 state, FP32 collision arithmetic
LAYOUT_TWO_BUFFER, LAYOUT_AA = 0, 1
CELL_ORDER_AUTO, CELL_ORDER_ROWS, CELL_ORDER_BRICKS = 0, 1, 2
POPULATIONS_AUTO, POPULATIONS_PLANES, POPULATIONS_INTERLEAVED = 0, 1, 2
# boundary-condition types of a class (include/lbm_capi.h)
BC_MOVING_WALL, BC_PRESSURE_OUTLET, BC_PRESSURE_INLET, BC_VELOCITY_INLET = 0, 1, 2, 3
BC_CELL_VELOCITY = 16
COUPLE_OVERWRITE, COUPLE_DIRICHLET = 0, 1

# D3Q19 ordering (madpac/peano/applications/latticeboltzmann/LatticeBoltzmannDefinitions.cpp:113-131)
C = np.array([
    [0, -1, -1], [-1, 0, -1], [0, 0, -1], [1, 0, -1], [0, 1, -1],
    [-1, -1, 0], [0, -1, 0], [1, -1, 0], [-1, 0, 0], [0, 0, 0],
    [1, 0, 0], [-1, 1, 0], [0, 1, 0], [1, 1, 0], [0, -1, 1],
    [-1, 0, 1], [0, 0, 1], [1, 0, 1], [0, 1, 1]], dtype=np.int64)

# every symbol include/lbm_capi.h declares
EXPORTS = (
    "lbm_create", "lbm_destroy", "lbm_last_error", "lbm_fluid_cells", "lbm_cell_order", "lbm_step_graph",
    "lbm_set_flags",
    "lbm_set_boundary_class", "lbm_set_moving_wall_class", "lbm_set_wall_classes", "lbm_set_boundary_velocity",
    "lbm_set_cavity_lid", "lbm_set_channel", "lbm_set_porosity",
    "lbm_set_coupling_cells", "lbm_set_coupling_data", "lbm_get_coupling_data",
    "lbm_init_equilibrium", "lbm_set_populations", "lbm_get_populations", "lbm_step", "lbm_step_timed", "lbm_measure_copy",
    "lbm_tune", "lbm_synchronize", "lbm_get_fields", "lbm_reduce_moments", "lbm_reduce_moments_noalloc",
    "lbm_kinetic_energy",
    "lbm_pressure_drop",
    "lbm_adjoint_pressure_drop",
    "lbm_patches_create", "lbm_patches_destroy", "lbm_patches_last_error", "lbm_patches_cells",
    "lbm_patches_init", "lbm_patches_get_populations", "lbm_patches_get_fields", "lbm_patches_autotune",
    "lbm_patches_set_tiling", "lbm_patches_step", "lbm_patches_reduce_moments", "lbm_patches_set_flags",
    "lbm_blocks_create", "lbm_blocks_destroy", "lbm_blocks_last_error", "lbm_blocks_set", "lbm_blocks_get",
    "lbm_blocks_find_overlap", "lbm_blocks_set_overlap", "lbm_blocks_prolongate", "lbm_blocks_restrict", "lbm_blocks_last_ms",
    "lbm_blocks_set_inner", "lbm_blocks_compute_cell_values", "lbm_blocks_collide", "lbm_blocks_treat_boundary",
    "lbm_blocks_stream_local", "lbm_blocks_stream_nonlocal", "lbm_blocks_swap_inside", "lbm_blocks_vertex_first",
    "lbm_blocks_iterate", "lbm_blocks_set_boundary", "lbm_blocks_init_pdf_diff", "lbm_blocks_compute_pdf_diff",
    "lbm_blocks_get_pdf_diff", "lbm_blocks_merge_pdf_diff", "lbm_blocks_reset_density_on_boundary",
    "lbm_blocks_set_density_on_boundary", "lbm_blocks_get_density_on_boundary", "lbm_blocks_merge_density_on_boundary",
)


class LbmError(RuntimeError):
    pass


class Config(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("periodic", ctypes.c_int * 3), ("precision", ctypes.c_int), ("tau", ctypes.c_double),
                ("force", ctypes.c_double * 3), ("device", ctypes.c_int), ("layout", ctypes.c_int),
                ("cell_order", ctypes.c_int), ("populations", ctypes.c_int)]


class AdjointStats(ctypes.Structure):
    _fields_ = [("snapshots_used", ctypes.c_int), ("forward_steps", ctypes.c_int64),
                ("recompute_steps", ctypes.c_int64), ("reverse_steps", ctypes.c_int64),
                ("forward_ms", ctypes.c_float), ("reverse_ms", ctypes.c_float),
                ("recompute_ms", ctypes.c_float), ("adjoint_ms", ctypes.c_float)]


class Patch(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("tau", ctypes.c_double), ("nbr", ctypes.c_int * Q)]


class AmrLink(ctypes.Structure):
    _fields_ = [("coarse", ctypes.c_int64), ("fine", ctypes.c_int64), ("pos", ctypes.c_int * 3),
                ("new_vertex", ctypes.c_int)]


class BlockBoundary(ctypes.Structure):
    """lbm_block_boundary: a BoundaryData record with its (cell, incoming direction) index."""
    _fields_ = [("block", ctypes.c_int64), ("cell", ctypes.c_int32), ("q_in", ctypes.c_int32),
                ("type", ctypes.c_int32), ("pad", ctypes.c_int32), ("u", ctypes.c_double * 3), ("rho", ctypes.c_double)]


# BoundaryData::BoundaryType
BOUNDARY_HW_BB, BOUNDARY_MOVING_WALL, BOUNDARY_PRESSURE_OUTLET, BOUNDARY_PRESSURE_INLET, BOUNDARY_VELOCITY_INLET, \
    BOUNDARY_MOVING_OBSTACLE = range(6)


class Tiling(ctypes.Structure):
    _fields_ = [("tile_x", ctypes.c_int), ("tile_y", ctypes.c_int), ("tile_z", ctypes.c_int),
                ("blocks_per_cu", ctypes.c_int), ("chunked", ctypes.c_int)]


_lib = None


def load_library(path: str = LIB_PATH):
    """Load liblbm_hip.so (raises if it is absent: there is no fallback path)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise LbmError(f"{path} is missing: build it with `python -m madpac_amd.build` (hipcc, gfx950)")
    L = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    u8p = ctypes.POINTER(ctypes.c_uint8)
    L.lbm_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(Config)]
    L.lbm_destroy.argtypes = [vp]
    L.lbm_last_error.argtypes = [vp]
    L.lbm_last_error.restype = ctypes.c_char_p
    L.lbm_fluid_cells.argtypes = [vp]
    L.lbm_fluid_cells.restype = ctypes.c_int64
    L.lbm_cell_order.argtypes = [vp]
    L.lbm_cell_order.restype = ctypes.c_int
    L.lbm_set_flags.argtypes = [vp, u8p]
    L.lbm_set_boundary_class.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, dp, ctypes.c_uint32]
    L.lbm_set_moving_wall_class.argtypes = [vp, ctypes.c_int, dp, ctypes.c_uint32]
    L.lbm_set_wall_classes.argtypes = [vp, u8p]
    L.lbm_set_boundary_velocity.argtypes = [vp, dp]
    L.lbm_set_cavity_lid.argtypes = [vp, ctypes.c_double]
    L.lbm_set_channel.argtypes = [vp, ctypes.c_double, ctypes.c_int]
    L.lbm_set_coupling_cells.argtypes = [vp, ctypes.POINTER(ctypes.c_int64), ctypes.c_int64]
    L.lbm_set_coupling_data.argtypes = [vp, ctypes.c_int, dp, dp, ctypes.c_double]
    L.lbm_get_coupling_data.argtypes = [vp, dp, dp]
    L.lbm_set_porosity.argtypes = [vp, dp]
    L.lbm_init_equilibrium.argtypes = [vp, dp, dp]
    L.lbm_set_populations.argtypes = [vp, dp]
    L.lbm_get_populations.argtypes = [vp, dp]
    L.lbm_step.argtypes = [vp, ctypes.c_int]
    L.lbm_step_graph.argtypes = [vp, ctypes.c_int]
    L.lbm_step_timed.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]
    L.lbm_measure_copy.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]
    L.lbm_tune.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.lbm_synchronize.argtypes = [vp]
    L.lbm_get_fields.argtypes = [vp, dp, dp]
    L.lbm_reduce_moments.argtypes = [vp, dp]
    L.lbm_reduce_moments_noalloc.argtypes = [vp, dp]
    L.lbm_kinetic_energy.argtypes = [vp, dp]
    L.lbm_pressure_drop.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
    L.lbm_adjoint_pressure_drop.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp,
                                            ctypes.POINTER(AdjointStats)]
    if hasattr(L, "lbm_patches_create"):
        L.lbm_patches_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(Patch), ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int]
        L.lbm_patches_destroy.argtypes = [vp]
        L.lbm_patches_last_error.argtypes = [vp]
        L.lbm_patches_last_error.restype = ctypes.c_char_p
        L.lbm_patches_cells.argtypes = [vp]
        L.lbm_patches_cells.restype = ctypes.c_int64
        L.lbm_patches_init.argtypes = [vp, dp, dp]
        L.lbm_patches_set_flags.argtypes = [vp, u8p]
        L.lbm_patches_get_populations.argtypes = [vp, dp]
        L.lbm_patches_get_fields.argtypes = [vp, dp, dp]
        L.lbm_patches_autotune.argtypes = [vp, ctypes.POINTER(Tiling)]
        L.lbm_patches_set_tiling.argtypes = [vp, ctypes.POINTER(Tiling)]
        L.lbm_patches_step.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]
        L.lbm_patches_reduce_moments.argtypes = [vp, dp]
    lp = ctypes.POINTER(AmrLink)
    L.lbm_blocks_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int64, ctypes.c_int]
    L.lbm_blocks_destroy.argtypes = [vp]
    L.lbm_blocks_last_error.argtypes = [vp]
    L.lbm_blocks_last_error.restype = ctypes.c_char_p
    L.lbm_blocks_set.argtypes = [vp, ctypes.c_int64, ctypes.c_int64, dp, dp, dp]
    L.lbm_blocks_get.argtypes = [vp, ctypes.c_int64, ctypes.c_int64, dp, dp, dp]
    L.lbm_blocks_find_overlap.argtypes = [vp, lp, ctypes.c_int64, u8p]
    L.lbm_blocks_set_overlap.argtypes = [vp, ctypes.c_int64, ctypes.c_int64, u8p]
    L.lbm_blocks_prolongate.argtypes = [vp, lp, ctypes.c_int64]
    L.lbm_blocks_restrict.argtypes = [vp, lp, ctypes.c_int64]
    L.lbm_blocks_last_ms.argtypes = [vp]
    L.lbm_blocks_last_ms.restype = ctypes.c_float
    i64p = ctypes.POINTER(ctypes.c_int64)
    bp = ctypes.POINTER(BlockBoundary)
    L.lbm_blocks_set_inner.argtypes = [vp, ctypes.c_int64, ctypes.c_int64, u8p]
    L.lbm_blocks_compute_cell_values.argtypes = [vp, i64p, ctypes.c_int64, ctypes.c_int]
    L.lbm_blocks_collide.argtypes = [vp, i64p, ctypes.c_int64, ctypes.c_double, ctypes.c_int]
    L.lbm_blocks_treat_boundary.argtypes = [vp, bp, ctypes.c_int64]
    L.lbm_blocks_stream_local.argtypes = [vp, i64p, ctypes.c_int64]
    L.lbm_blocks_stream_nonlocal.argtypes = [vp, i64p, ctypes.c_int64]
    L.lbm_blocks_swap_inside.argtypes = [vp, i64p, ctypes.c_int64]
    L.lbm_blocks_vertex_first.argtypes = [vp, i64p, ctypes.c_int64, ctypes.c_double, ctypes.c_int, bp, ctypes.c_int64]
    L.lbm_blocks_iterate.argtypes = [vp, i64p, ctypes.c_int64, i64p, ctypes.c_int64, ctypes.c_double, bp,
                                     ctypes.c_int64, ctypes.c_int]
    dpp = ctypes.POINTER(ctypes.c_double)
    L.lbm_blocks_set_boundary.argtypes = [vp, ctypes.c_int64, ctypes.c_int64, u8p]
    for name in ("init_pdf_diff", "compute_pdf_diff", "reset_density_on_boundary"):
        getattr(L, "lbm_blocks_" + name).argtypes = [vp, i64p, ctypes.c_int64]
    for name in ("get_pdf_diff", "merge_pdf_diff", "get_density_on_boundary", "merge_density_on_boundary"):
        getattr(L, "lbm_blocks_" + name).argtypes = [vp, i64p, ctypes.c_int64, dpp]
    L.lbm_blocks_set_density_on_boundary.argtypes = [vp, i64p, ctypes.c_int64, ctypes.c_int]
    _lib = L
    return L


def _dp(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _u8(a):
    assert a.dtype == np.uint8 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))


def _sized(a, size, what):
    """a, after checking that it holds exactly `size` elements: the C side reads that many."""
    if a.size != size:
        raise ValueError(f"{what}: expected {size} values, got {a.size}")
    return a


class Lattice:
    """A dense D3Q19 lattice on one GPU (an lbm_ctx)."""

    def __init__(self, nx, ny, nz, tau, precision=FP64, periodic=(False, False, False), force=(0.0, 0.0, 0.0),
                 device=0, layout=LAYOUT_TWO_BUFFER, cell_order=CELL_ORDER_AUTO, populations=POPULATIONS_AUTO):
        self._L = load_library()
        cfg = Config(nx, ny, nz, (ctypes.c_int * 3)(*[int(p) for p in periodic]), int(precision), float(tau),
                     (ctypes.c_double * 3)(*[float(v) for v in force]), int(device), int(layout), int(cell_order),
                     int(populations))
        h = ctypes.c_void_p()
        rc = self._L.lbm_create(ctypes.byref(h), ctypes.byref(cfg))
        if rc != 0:
            raise LbmError(f"lbm_create failed ({rc}): {self._L.lbm_last_error(None).decode()}")
        self._h = h
        self.shape = (nz, ny, nx)
        self.n = nx * ny * nz
        self.precision = precision

    def _check(self, rc, what):
        if rc != 0:
            raise LbmError(f"{what} failed ({rc}): {self._L.lbm_last_error(self._h).decode()}")

    def close(self):
        if getattr(self, "_h", None):
            self._L.lbm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- geometry
    def set_flags(self, inner):
        a = _sized(np.ascontiguousarray(np.asarray(inner).reshape(-1) != 0, dtype=np.uint8), self.n, "flags")
        self._check(self._L.lbm_set_flags(self._h, _u8(a)), "lbm_set_flags")

    def set_cavity_lid(self, ulid):
        self._check(self._L.lbm_set_cavity_lid(self._h, float(ulid)), "lbm_set_cavity_lid")

    def set_channel(self, u_mean, parabolic=False):
        self._check(self._L.lbm_set_channel(self._h, float(u_mean), int(bool(parabolic))), "lbm_set_channel")

    def set_boundary_class(self, k, bc_type, rho=1.0, u=(0.0, 0.0, 0.0), mask=0):
        u = _sized(np.ascontiguousarray(u, dtype=np.float64).reshape(-1), 3, "boundary velocity")
        self._check(self._L.lbm_set_boundary_class(self._h, int(k), int(bc_type), float(rho), _dp(u),
                                                   ctypes.c_uint32(mask)), "lbm_set_boundary_class")

    def set_boundary_velocity(self, u):
        if u is None:
            self._check(self._L.lbm_set_boundary_velocity(self._h, None), "lbm_set_boundary_velocity")
            return
        a = _sized(np.ascontiguousarray(np.asarray(u, dtype=np.float64).reshape(-1)), 3 * self.n,
                   "boundary velocities")
        self._check(self._L.lbm_set_boundary_velocity(self._h, _dp(a)), "lbm_set_boundary_velocity")

    # ---- MD coupling
    def set_coupling_cells(self, cells):
        if cells is None or len(cells) == 0:
            self._check(self._L.lbm_set_coupling_cells(self._h, None, 0), "lbm_set_coupling_cells")
            self._ncouple = 0
            return
        a = np.ascontiguousarray(np.asarray(cells, dtype=np.int64).reshape(-1))
        self._check(self._L.lbm_set_coupling_cells(self._h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), a.size),
                    "lbm_set_coupling_cells")
        self._ncouple = a.size

    def set_coupling_data(self, mode, u, rho=None, dt=1.0):
        if u is None:
            self._check(self._L.lbm_set_coupling_data(self._h, int(mode), None, None, float(dt)),
                        "lbm_set_coupling_data")
            return
        n = getattr(self, "_ncouple", 0)
        u = _sized(np.ascontiguousarray(np.asarray(u, dtype=np.float64).reshape(-1)), 3 * n, "coupling velocities")
        r = None if rho is None else _sized(np.ascontiguousarray(np.asarray(rho, dtype=np.float64).reshape(-1)), n,
                                            "coupling densities")
        self._check(self._L.lbm_set_coupling_data(self._h, int(mode), _dp(u), None if r is None else _dp(r), float(dt)),
                    "lbm_set_coupling_data")

    def get_coupling_data(self):
        n = getattr(self, "_ncouple", 0)
        rho, u = np.empty(n), np.empty((n, 3))
        self._check(self._L.lbm_get_coupling_data(self._h, _dp(u.reshape(-1)), _dp(rho)), "lbm_get_coupling_data")
        return rho, u

    def set_moving_wall_class(self, k, uw, mask):
        uw = _sized(np.ascontiguousarray(uw, dtype=np.float64).reshape(-1), 3, "wall velocity")
        self._check(self._L.lbm_set_moving_wall_class(self._h, int(k), _dp(uw), ctypes.c_uint32(mask)),
                    "lbm_set_moving_wall_class")

    def set_wall_classes(self, cls):
        if cls is None:
            self._check(self._L.lbm_set_wall_classes(self._h, None), "lbm_set_wall_classes")
            return
        a = _sized(np.ascontiguousarray(np.asarray(cls).reshape(-1), dtype=np.uint8), self.n, "wall classes")
        self._check(self._L.lbm_set_wall_classes(self._h, _u8(a)), "lbm_set_wall_classes")

    def set_porosity(self, alpha):
        if alpha is None:
            self._check(self._L.lbm_set_porosity(self._h, None), "lbm_set_porosity")
            return
        a = _sized(np.ascontiguousarray(np.asarray(alpha, dtype=np.float64).reshape(-1)), self.n, "porosity")
        self._check(self._L.lbm_set_porosity(self._h, _dp(a)), "lbm_set_porosity")

    @property
    def fluid_cells(self):
        return int(self._L.lbm_fluid_cells(self._h))

    @property
    def cell_order(self):
        """CELL_ORDER_ROWS or CELL_ORDER_BRICKS, as chosen at the last set_flags."""
        return int(self._L.lbm_cell_order(self._h))

    # ---- state
    def init_equilibrium(self, rho=None, u=None):
        r = None if rho is None else _sized(np.ascontiguousarray(np.asarray(rho, dtype=np.float64).reshape(-1)),
                                            self.n, "density")
        v = None if u is None else _sized(np.ascontiguousarray(np.asarray(u, dtype=np.float64).reshape(-1)),
                                          3 * self.n, "velocity")
        self._check(self._L.lbm_init_equilibrium(self._h, None if r is None else _dp(r), None if v is None else _dp(v)),
                    "lbm_init_equilibrium")

    def set_populations(self, f):
        a = _sized(np.ascontiguousarray(np.asarray(f, dtype=np.float64).reshape(-1)), Q * self.n, "populations")
        self._check(self._L.lbm_set_populations(self._h, _dp(a)), "lbm_set_populations")

    def get_populations(self):
        f = np.empty(Q * self.n)
        self._check(self._L.lbm_get_populations(self._h, _dp(f)), "lbm_get_populations")
        return f

    def step(self, n=1, graph=False):
        if graph:
            self._check(self._L.lbm_step_graph(self._h, int(n)), "lbm_step_graph")
        else:
            self._check(self._L.lbm_step(self._h, int(n)), "lbm_step")

    def step_timed(self, n):
        ms = ctypes.c_float()
        self._check(self._L.lbm_step_timed(self._h, int(n), ctypes.byref(ms)), "lbm_step_timed")
        return float(ms.value)

    def tune(self, nsteps=10, variant=None):
        """Pick the fastest step-kernel variant on the current state (returns 4..15: bit 0 the load
        schedule, bit 1 the XCD-contiguous workgroup order, bits 2-3 the workgroup shape 2/4/8 rows
        or bricks), or with `variant` given (0..15; shape bits 0 = the default), select it."""
        v = ctypes.c_int(0 if variant is None else int(variant))
        n = 0 if variant is not None else int(nsteps)
        self._check(self._L.lbm_tune(self._h, n, ctypes.byref(v)), "lbm_tune")
        return int(v.value)

    def synchronize(self):
        self._check(self._L.lbm_synchronize(self._h), "lbm_synchronize")

    def fields(self):
        rho = np.empty(self.n)
        u = np.empty(3 * self.n)
        self._check(self._L.lbm_get_fields(self._h, _dp(rho), _dp(u)), "lbm_get_fields")
        return rho, u.reshape(-1, 3)

    def moments(self):
        out = np.empty(4)
        self._check(self._L.lbm_reduce_moments(self._h, _dp(out)), "lbm_reduce_moments")
        return out

    def moments_noalloc(self):
        """moments() through lbm_reduce_moments_noalloc (no device allocation; same bits)."""
        out = np.empty(4)
        self._check(self._L.lbm_reduce_moments_noalloc(self._h, _dp(out)), "lbm_reduce_moments_noalloc")
        return out

    def measure_copy(self, steps, pull=False):
        """Device ms of `steps` launches of the step's bandwidth reference (19-plane copy, or the
        pull pattern with pull=True) over every stored cell; the state is unchanged."""
        ms = ctypes.c_float()
        self._check(self._L.lbm_measure_copy(self._h, 1 if pull else 0, int(steps), ctypes.byref(ms)),
                    "lbm_measure_copy")
        return ms.value

    def kinetic_energy(self):
        ke = ctypes.c_double()
        self._check(self._L.lbm_kinetic_energy(self._h, ctypes.byref(ke)), "lbm_kinetic_energy")
        return ke.value

    def pressure_drop(self, xin, xout):
        J = ctypes.c_double()
        self._check(self._L.lbm_pressure_drop(self._h, int(xin), int(xout), ctypes.byref(J)), "lbm_pressure_drop")
        return J.value

    def adjoint_pressure_drop(self, nsteps, snapshots, xin, xout):
        J = ctypes.c_double()
        grad = np.empty(self.n)
        st = AdjointStats()
        self._check(self._L.lbm_adjoint_pressure_drop(self._h, int(nsteps), int(snapshots), int(xin), int(xout),
                                                      ctypes.byref(J), _dp(grad), ctypes.byref(st)),
                    "lbm_adjoint_pressure_drop")
        return J.value, grad, st


class PatchSet:
    """Thousands of cubic patches with ghost layers, one persistent launch per step (an lbm_patchset).

    `table` is a sequence of (n, tau, nbr) with nbr a length-19 sequence of neighbour patch ids per
    D3Q19 direction (entry 9 is ignored)."""

    def __init__(self, table, precision=FP64, device=0):
        self._L = load_library()
        arr = (Patch * len(table))()
        for k, (n, tau, nbr) in enumerate(table):
            arr[k].n = int(n)
            arr[k].tau = float(tau)
            for q in range(Q):
                arr[k].nbr[q] = int(nbr[q])
        h = ctypes.c_void_p()
        rc = self._L.lbm_patches_create(ctypes.byref(h), arr, len(table), int(precision), int(device))
        if rc != 0:
            raise LbmError(f"lbm_patches_create failed ({rc}): {self._L.lbm_patches_last_error(None).decode()}")
        self._h = h
        self.edges = [int(t[0]) for t in table]
        self.cells = int(self._L.lbm_patches_cells(h))
        self.precision = precision

    def _check(self, rc, what):
        if rc != 0:
            raise LbmError(f"{what} failed ({rc}): {self._L.lbm_patches_last_error(self._h).decode()}")

    def close(self):
        if getattr(self, "_h", None):
            self._L.lbm_patches_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_flags(self, fluid):
        """Per-cell fluid flags (patch order, x fastest; None = all fluid); init() must follow."""
        if fluid is None:
            self._check(self._L.lbm_patches_set_flags(self._h, None), "lbm_patches_set_flags")
            return
        a = _sized(np.ascontiguousarray(np.asarray(fluid).reshape(-1) != 0, dtype=np.uint8), self.cells, "flags")
        self._check(self._L.lbm_patches_set_flags(self._h, _u8(a)), "lbm_patches_set_flags")

    def init(self, rho=None, u=None):
        r = None if rho is None else _sized(np.ascontiguousarray(np.asarray(rho, dtype=np.float64).reshape(-1)),
                                            self.cells, "density")
        v = None if u is None else _sized(np.ascontiguousarray(np.asarray(u, dtype=np.float64).reshape(-1)),
                                          3 * self.cells, "velocity")
        self._check(self._L.lbm_patches_init(self._h, None if r is None else _dp(r), None if v is None else _dp(v)),
                    "lbm_patches_init")

    def get_populations(self):
        """Reference state per patch: a list of arrays shaped (19, n, n, n) (x fastest)."""
        f = np.empty(Q * self.cells)
        self._check(self._L.lbm_patches_get_populations(self._h, _dp(f)), "lbm_patches_get_populations")
        out, o = [], 0
        for n in self.edges:
            out.append(f[o:o + Q * n ** 3].reshape(Q, n, n, n))
            o += Q * n ** 3
        return out

    def fields(self):
        """rho (cells,) and u (cells, 3) of every interior cell, patch order, x fastest."""
        rho = np.empty(self.cells)
        u = np.empty(3 * self.cells)
        self._check(self._L.lbm_patches_get_fields(self._h, _dp(rho), _dp(u)), "lbm_patches_get_fields")
        return rho, u.reshape(-1, 3)

    def autotune(self):
        t = Tiling()
        self._check(self._L.lbm_patches_autotune(self._h, ctypes.byref(t)), "lbm_patches_autotune")
        return (t.tile_x, t.tile_y, t.tile_z, t.blocks_per_cu, t.chunked)

    def set_tiling(self, tx, ty, tz, blocks_per_cu, chunked=0):
        t = Tiling(int(tx), int(ty), int(tz), int(blocks_per_cu), int(chunked))
        self._check(self._L.lbm_patches_set_tiling(self._h, ctypes.byref(t)), "lbm_patches_set_tiling")

    def step(self, n=1, graph=False, timed=False):
        ms = ctypes.c_float()
        self._check(self._L.lbm_patches_step(self._h, int(n), 1 if graph else 0, ctypes.byref(ms) if timed else None),
                    "lbm_patches_step")
        return float(ms.value) if timed else None

    def moments(self):
        out = np.empty(4)
        self._check(self._L.lbm_patches_reduce_moments(self._h, _dp(out)), "lbm_patches_reduce_moments")
        return out


def channel(nx, ny, nz, tau, u_mean, parabolic=False, solid=None, precision=FP64, device=0, layout=LAYOUT_TWO_BUFFER):
    """The reference's default channel: no-slip y/z walls, velocity inlet on x- (uniform or
    parabolic), pressure outlet rho = 1 on x+, optional solid obstacles; f = w."""
    lat = Lattice(nx, ny, nz, tau, precision=precision, device=device, layout=layout)
    lat.set_flags(np.ones(lat.n, np.uint8) if solid is None else ~np.asarray(solid, bool).reshape(-1))
    lat.set_channel(u_mean, parabolic)
    lat.init_equilibrium()
    return lat


def cavity(n, tau, ulid, precision=FP64, device=0, layout=LAYOUT_TWO_BUFFER):
    """The reference's lid-driven cavity: n^3 fluid cells, lid (ulid, 0, 0) on +z, f = w."""
    lat = Lattice(n, n, n, tau, precision=precision, device=device, layout=layout)
    lat.set_flags(np.ones(lat.n, np.uint8))
    lat.set_cavity_lid(ulid)
    lat.init_equilibrium()
    return lat


BLOCK_CELLS = 216
BLOCK_BOUNDARY_PDFS = 1008   # populations leaving a block (BLB/parallel/BlockBoundaryIndex4Pdfs.cpp)
BLOCK_BOUNDARY_CELLS = 208   # cells near the block boundary (BLB/BlockBoundaryIndex.cpp)


class BlockStore:
    """Device-resident 6^3 vertex blocks in the reference's layout (an lbm_blockstore) with batched
    restriction and prolongation between grid levels.  Links are sequences of (coarse, fine, pos) or
    (coarse, fine, pos, new_vertex)."""

    def __init__(self, nblocks, device=0):
        self._L = load_library()
        h = ctypes.c_void_p()
        rc = self._L.lbm_blocks_create(ctypes.byref(h), int(nblocks), int(device))
        if rc != 0:
            raise LbmError(f"lbm_blocks_create failed ({rc}): {self._L.lbm_blocks_last_error(None).decode()}")
        self._h = h
        self.nblocks = int(nblocks)

    def _check(self, rc, what):
        if rc != 0:
            raise LbmError(f"{what} failed ({rc}): {self._L.lbm_blocks_last_error(self._h).decode()}")

    def close(self):
        if getattr(self, "_h", None):
            self._L.lbm_blocks_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set(self, first, pdf=None, density=None, velocity=None):
        arrs, count = [], None
        for a, per, what in ((pdf, BLOCK_CELLS * Q, "pdf"), (density, BLOCK_CELLS, "density"),
                             (velocity, 3 * BLOCK_CELLS, "velocity")):
            if a is None:
                arrs.append(None)
                continue
            a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))
            if a.size % per:
                raise ValueError(f"{what}: size {a.size} is not a whole number of blocks")
            if count is not None and a.size // per != count:
                raise ValueError("pdf, density and velocity cover different block counts")
            count = a.size // per
            arrs.append(a)
        if count is None:
            return
        self._check(self._L.lbm_blocks_set(self._h, int(first), int(count), *[None if a is None else _dp(a) for a in arrs]),
                    "lbm_blocks_set")

    def get(self, first=0, count=None):
        count = self.nblocks - first if count is None else int(count)
        pdf = np.empty(count * BLOCK_CELLS * Q)
        rho = np.empty(count * BLOCK_CELLS)
        u = np.empty(count * 3 * BLOCK_CELLS)
        self._check(self._L.lbm_blocks_get(self._h, int(first), count, _dp(pdf), _dp(rho), _dp(u)), "lbm_blocks_get")
        return pdf.reshape(count, BLOCK_CELLS * Q), rho.reshape(count, BLOCK_CELLS), u.reshape(count, 3 * BLOCK_CELLS)

    @staticmethod
    def _links(links):
        arr = (AmrLink * max(1, len(links)))()
        for k, l in enumerate(links):
            arr[k].coarse = int(l[0])
            arr[k].fine = int(l[1])
            for d in range(3):
                arr[k].pos[d] = int(l[2][d])
            arr[k].new_vertex = int(bool(l[3])) if len(l) > 3 else 0
        return arr

    def find_overlap(self, links):
        arr = self._links(links)
        out = np.zeros(len(links) * BLOCK_CELLS, np.uint8)
        self._check(self._L.lbm_blocks_find_overlap(self._h, arr, len(links), _u8(out)), "lbm_blocks_find_overlap")
        return out.reshape(len(links), BLOCK_CELLS).astype(bool)

    def set_overlap(self, first, overlap):
        m = np.ascontiguousarray(np.asarray(overlap).reshape(-1) != 0, dtype=np.uint8)
        if m.size % BLOCK_CELLS:
            raise ValueError("overlap: not a whole number of blocks")
        self._check(self._L.lbm_blocks_set_overlap(self._h, int(first), m.size // BLOCK_CELLS, _u8(m)),
                    "lbm_blocks_set_overlap")

    def prolongate(self, links):
        self._check(self._L.lbm_blocks_prolongate(self._h, self._links(links), len(links)), "lbm_blocks_prolongate")

    def restrict(self, links):
        self._check(self._L.lbm_blocks_restrict(self._h, self._links(links), len(links)), "lbm_blocks_restrict")

    def last_ms(self):
        return float(self._L.lbm_blocks_last_ms(self._h))

    # ----- the block solver's operators -----
    @staticmethod
    def _ids(ids, per=1):
        a = np.ascontiguousarray(np.asarray(ids, dtype=np.int64).reshape(-1))
        if a.size % per:
            raise ValueError(f"id list: size {a.size} is not a multiple of {per}")
        return a, a.size // per, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))

    LINK_DTYPE = np.dtype([("block", "<i8"), ("cell", "<i4"), ("q_in", "<i4"), ("type", "<i4"), ("pad", "<i4"),
                           ("u", "<f8", (3,)), ("rho", "<f8")])

    @classmethod
    def boundary_links(cls, links):
        """links: iterable of (block, cell, q_in, type, (u0, u1, u2), rho_L), or an array of LINK_DTYPE
        -> (pointer to lbm_block_boundary records, count, owner)."""
        if isinstance(links, np.ndarray) and links.dtype == cls.LINK_DTYPE:
            arr = np.ascontiguousarray(links)
        else:
            arr = np.array([(int(b), int(c), int(q), int(t), 0, tuple(float(x) for x in u), float(r))
                            for b, c, q, t, u, r in links], dtype=cls.LINK_DTYPE)
        assert cls.LINK_DTYPE.itemsize == ctypes.sizeof(BlockBoundary)
        keep = arr if arr.size else np.zeros(1, cls.LINK_DTYPE)
        return keep.ctypes.data_as(ctypes.POINTER(BlockBoundary)), int(arr.size), keep

    def set_inner(self, first, inner):
        m = np.ascontiguousarray(np.asarray(inner).reshape(-1) != 0, dtype=np.uint8)
        if m.size % BLOCK_CELLS:
            raise ValueError("inner: not a whole number of blocks")
        self._check(self._L.lbm_blocks_set_inner(self._h, int(first), m.size // BLOCK_CELLS, _u8(m)), "lbm_blocks_set_inner")

    def compute_cell_values(self, blocks, use_overlap=False):
        a, n, p = self._ids(blocks)
        self._check(self._L.lbm_blocks_compute_cell_values(self._h, p, n, int(bool(use_overlap))),
                    "lbm_blocks_compute_cell_values")

    def collide(self, blocks, tau, use_overlap=False):
        a, n, p = self._ids(blocks)
        self._check(self._L.lbm_blocks_collide(self._h, p, n, float(tau), int(bool(use_overlap))), "lbm_blocks_collide")

    def treat_boundary(self, links):
        arr, n, _keep = self.boundary_links(links)
        self._check(self._L.lbm_blocks_treat_boundary(self._h, arr, n), "lbm_blocks_treat_boundary")

    def stream_local(self, blocks):
        a, n, p = self._ids(blocks)
        self._check(self._L.lbm_blocks_stream_local(self._h, p, n), "lbm_blocks_stream_local")

    def stream_nonlocal(self, cells):
        a, n, p = self._ids(cells, 8)
        self._check(self._L.lbm_blocks_stream_nonlocal(self._h, p, n), "lbm_blocks_stream_nonlocal")

    def swap_inside(self, blocks):
        a, n, p = self._ids(blocks)
        self._check(self._L.lbm_blocks_swap_inside(self._h, p, n), "lbm_blocks_swap_inside")

    def vertex_first(self, blocks, tau, links=(), use_overlap=False):
        a, n, p = self._ids(blocks)
        arr, nl, _keep = self.boundary_links(links)
        self._check(self._L.lbm_blocks_vertex_first(self._h, p, n, float(tau), int(bool(use_overlap)), arr, nl),
                    "lbm_blocks_vertex_first")

    def iterate(self, blocks, cells, tau, links=(), steps=1):
        a, n, p = self._ids(blocks)
        c, nc, pc = self._ids(cells, 8)
        arr, nl, _keep = self.boundary_links(links)
        self._check(self._L.lbm_blocks_iterate(self._h, p, n, pc, nc, float(tau), arr, nl, int(steps)), "lbm_blocks_iterate")

    # ----- the vertex-overlap exchange between ranks (BLB/RegularGridBlockVertex.cpp:51-160) -----
    def set_boundary(self, first, boundary):
        m = np.ascontiguousarray(np.asarray(boundary).reshape(-1) != 0, dtype=np.uint8)
        if m.size % BLOCK_CELLS:
            raise ValueError("boundary: not a whole number of blocks")
        self._check(self._L.lbm_blocks_set_boundary(self._h, int(first), m.size // BLOCK_CELLS, _u8(m)),
                    "lbm_blocks_set_boundary")

    def init_pdf_diff(self, blocks):
        a, n, p = self._ids(blocks)
        self._check(self._L.lbm_blocks_init_pdf_diff(self._h, p, n), "lbm_blocks_init_pdf_diff")

    def compute_pdf_diff(self, blocks):
        a, n, p = self._ids(blocks)
        self._check(self._L.lbm_blocks_compute_pdf_diff(self._h, p, n), "lbm_blocks_compute_pdf_diff")

    def _rows_out(self, fn, blocks, width):
        a, n, p = self._ids(blocks)
        out = np.empty((n, width))
        self._check(getattr(self._L, fn)(self._h, p, n, _dp(out)), fn)
        return out

    def _rows_in(self, fn, blocks, values, width):
        a, n, p = self._ids(blocks)
        v = np.ascontiguousarray(np.asarray(values, dtype=np.float64).reshape(-1))
        if v.size != n * width:
            raise ValueError(f"{fn}: expected {n * width} values, got {v.size}")
        self._check(getattr(self._L, fn)(self._h, p, n, _dp(v)), fn)

    def get_pdf_diff(self, blocks):
        """(n, 1008) differences of the listed blocks"""
        return self._rows_out("lbm_blocks_get_pdf_diff", blocks, BLOCK_BOUNDARY_PDFS)

    def merge_pdf_diff(self, blocks, neighbour):
        self._rows_in("lbm_blocks_merge_pdf_diff", blocks, neighbour, BLOCK_BOUNDARY_PDFS)

    def reset_density_on_boundary(self, blocks):
        a, n, p = self._ids(blocks)
        self._check(self._L.lbm_blocks_reset_density_on_boundary(self._h, p, n), "lbm_blocks_reset_density_on_boundary")

    def set_density_on_boundary(self, blocks, dynamic_geometry=True):
        a, n, p = self._ids(blocks)
        self._check(self._L.lbm_blocks_set_density_on_boundary(self._h, p, n, int(bool(dynamic_geometry))),
                    "lbm_blocks_set_density_on_boundary")

    def get_density_on_boundary(self, blocks):
        """(n, 208) boundary densities of the listed blocks"""
        return self._rows_out("lbm_blocks_get_density_on_boundary", blocks, BLOCK_BOUNDARY_CELLS)

    def merge_density_on_boundary(self, blocks, neighbour):
        self._rows_in("lbm_blocks_merge_density_on_boundary", blocks, neighbour, BLOCK_BOUNDARY_CELLS)
