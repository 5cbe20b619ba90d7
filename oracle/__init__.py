"""CPU oracle for the D3Q19 block-LB hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only to check results against.  The product
package ``madpac_amd`` never imports it.

The numerics live in ``lbm_oracle.c`` (a plain-C restatement of the reference's
FP64 arithmetic, see the header for citations); this module is a thin ctypes
binding over ``liblbm_oracle.so`` plus numpy helpers for layouts.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblbm_oracle.so")

Q = 19
REST = 9
FLUID = 0x80000000
CLASS_SHIFT = 24
MAX_CLASSES = 32
# boundary-condition types of a class (lbm_oracle.h)
BC_MOVING_WALL, BC_PRESSURE_OUTLET, BC_PRESSURE_INLET, BC_VELOCITY_INLET = 0, 1, 2, 3
BC_CELL_VELOCITY = 16
COUPLE_OVERWRITE, COUPLE_DIRICHLET = 0, 1
CS2 = 1.0 / 3.0

# LB/LatticeBoltzmannDefinitions.cpp:113-131 (z-major, then y, then x)
C = np.array([
    [0, -1, -1], [-1, 0, -1], [0, 0, -1], [1, 0, -1], [0, 1, -1],
    [-1, -1, 0], [0, -1, 0], [1, -1, 0], [-1, 0, 0], [0, 0, 0],
    [1, 0, 0], [-1, 1, 0], [0, 1, 0], [1, 1, 0], [0, -1, 1],
    [-1, 0, 1], [0, 0, 1], [1, 0, 1], [0, 1, 1]], dtype=np.int64)
# LB/LatticeBoltzmannDefinitions.cpp:149-151
W = np.array([1 / 36, 1 / 36, 1 / 18, 1 / 36, 1 / 36, 1 / 36, 1 / 18, 1 / 36, 1 / 18, 1 / 3,
              1 / 18, 1 / 36, 1 / 18, 1 / 36, 1 / 36, 1 / 36, 1 / 18, 1 / 36, 1 / 36])


class Dims(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("periodic", ctypes.c_int * 3)]


class Params(ctypes.Structure):
    _fields_ = [("tau", ctypes.c_double),
                ("uw", (ctypes.c_double * 3) * MAX_CLASSES),
                ("moving_mask", ctypes.c_uint32 * MAX_CLASSES),
                ("g", ctypes.c_double * 3),
                ("use_force", ctypes.c_int),
                ("alpha", ctypes.POINTER(ctypes.c_double)),
                ("tau_cell", ctypes.POINTER(ctypes.c_double)),
                ("bc_type", ctypes.c_int * MAX_CLASSES),
                ("bc_rho", ctypes.c_double * MAX_CLASSES),
                ("bc_u", ctypes.POINTER(ctypes.c_double)),
                ("couple_row", ctypes.POINTER(ctypes.c_int64)),
                ("couple_mode", ctypes.c_int),
                ("couple_u", ctypes.POINTER(ctypes.c_double)),
                ("couple_rho", ctypes.POINTER(ctypes.c_double)),
                ("couple_dt", ctypes.c_double)]


_lib = None


def build() -> str:
    """Compile liblbm_oracle.so from oracle/lbm_oracle.c (gcc)."""
    subprocess.run(["make", "-s", "-f", os.path.join(_HERE, "Makefile")], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        up = ctypes.POINTER(ctypes.c_uint32)
        L.lbo_tau_eff.restype = ctypes.c_double
        L.lbo_tau_eff.argtypes = [ctypes.c_double]
        L.lbo_density.restype = ctypes.c_double
        L.lbo_density.argtypes = [dp]
        L.lbo_velocity.argtypes = [dp, ctypes.c_double, dp]
        L.lbo_feq.argtypes = [ctypes.c_double, dp, dp]
        L.lbo_bgk.argtypes = [dp, ctypes.c_double, dp]
        L.lbo_moving_wall_term.restype = ctypes.c_double
        L.lbo_moving_wall_term.argtypes = [dp, ctypes.c_int, ctypes.c_double]
        L.lbo_build_cavity.argtypes = [ctypes.c_int, up, ctypes.POINTER(Params), ctypes.c_double]
        L.lbo_build_channel.argtypes = [ctypes.POINTER(Dims), up, ctypes.POINTER(Params), ctypes.c_double,
                                        ctypes.c_int, dp]
        L.lbo_boundary_value.restype = ctypes.c_double
        L.lbo_boundary_value.argtypes = [ctypes.POINTER(Params), ctypes.c_uint, ctypes.c_int64, ctypes.c_int,
                                         ctypes.c_double, ctypes.c_double, dp]
        L.lbo_couple_gi.argtypes = [ctypes.c_double, ctypes.c_double, dp, dp, dp]
        L.lbo_build_words.argtypes = [ctypes.POINTER(Dims), ctypes.POINTER(ctypes.c_uint8), up]
        L.lbo_step.argtypes = [ctypes.POINTER(Dims), up, ctypes.POINTER(Params), dp, dp, ctypes.c_int]
        L.lbo_stream.argtypes = [ctypes.POINTER(Dims), up, dp, dp]
        L.lbo_step_swap_aos.argtypes = [ctypes.POINTER(Dims), up, ctypes.POINTER(Params), dp]
        L.lbo_fields.argtypes = [ctypes.POINTER(Dims), up, dp, dp, dp]
        L.lbo_couple_export.argtypes = [ctypes.POINTER(Dims), up, ctypes.POINTER(Params), dp,
                                        ctypes.POINTER(ctypes.c_int64), ctypes.c_int64, dp, dp]
        L.lbo_init_equilibrium.argtypes = [ctypes.POINTER(Dims), up, dp, dp, dp]
        L.lbo_moments.argtypes = [ctypes.POINTER(Dims), up, dp, dp]
        L.lbo_pressure_drop.restype = ctypes.c_double
        L.lbo_pressure_drop.argtypes = [ctypes.POINTER(Dims), up, dp, ctypes.c_int, ctypes.c_int]
        L.lbo_pressure_drop_grad.argtypes = [ctypes.POINTER(Dims), up, ctypes.c_int, ctypes.c_int, dp]
        L.lbo_step_adjoint.argtypes = [ctypes.POINTER(Dims), up, ctypes.POINTER(Params), dp, dp, dp, dp,
                                       ctypes.c_int]
        ip = ctypes.POINTER(ctypes.c_int)
        u8 = ctypes.POINTER(ctypes.c_uint8)
        L.lbo_amr_father.argtypes = [ip, ip]
        L.lbo_amr_subblock.argtypes = [ip, ip]
        L.lbo_amr_from_neighbour.restype = ctypes.c_int
        L.lbo_amr_from_neighbour.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.lbo_amr_find_overlap.argtypes = [dp, u8]
        L.lbo_amr_prolongate.argtypes = [ip, dp, dp, dp, dp, u8]
        L.lbo_amr_restrict.argtypes = [ip, dp, dp]
        L.lbo_block_cell_values.argtypes = [u8, u8, dp, dp, dp]
        L.lbo_block_collide.argtypes = [u8, u8, dp, dp, ctypes.c_double, dp]
        L.lbo_block_boundary.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, ctypes.c_double, dp, dp, dp]
        L.lbo_block_local_index.restype = ctypes.c_int
        L.lbo_block_local_index.argtypes = [ip, ctypes.c_int]
        L.lbo_block_stream_local.argtypes = [u8, dp]
        L.lbo_block_swap_inside.argtypes = [dp]
        L.lbo_block_nonlocal_index.restype = ctypes.c_int
        L.lbo_block_nonlocal_index.argtypes = [ip, ctypes.c_int]
        L.lbo_block_stream_nonlocal.argtypes = [ctypes.POINTER(u8), ctypes.POINTER(dp)]
        L.lbo_block_pdf_boundary.restype = ctypes.c_int
        L.lbo_block_pdf_boundary.argtypes = [ip, ip, ctypes.c_int]
        L.lbo_block_boundary_cells.restype = ctypes.c_int
        L.lbo_block_boundary_cells.argtypes = [ip, ctypes.c_int]
        L.lbo_block_init_pdf_diff.argtypes = [dp, dp]
        L.lbo_block_compute_pdf_diff.argtypes = [dp, dp]
        L.lbo_block_merge_pdf_diff.argtypes = [dp, dp, dp]
        L.lbo_block_set_density_on_boundary.argtypes = [dp, u8, u8, ctypes.c_int, dp]
        L.lbo_block_merge_density_on_boundary.argtypes = [dp, u8, u8, dp]
        _lib = L
    return _lib


def _dp(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _up(a):
    assert a.dtype == np.uint32 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


class Lattice:
    """A dense D3Q19 lattice with SoA populations f[q, z, y, x] and per-cell words."""

    def __init__(self, nx, ny, nz, periodic=(False, False, False)):
        self.dims = Dims(nx, ny, nz, (ctypes.c_int * 3)(*[int(p) for p in periodic]))
        self.shape = (nz, ny, nx)
        self.n = nx * ny * nz
        self.words = np.zeros(self.n, dtype=np.uint32)
        self.params = Params()
        self.params.tau = 1.0
        self._alpha = None

    # ----- geometry -----
    @classmethod
    def cavity(cls, n, tau, ulid, solid=None):
        """The reference's lid-driven cavity on n^3 cells; `solid` (n^3, 1 = solid) adds obstacles,
        whose links stay halfway bounce-back while the top layer's c_z > 0 links are the lid."""
        lat = cls(n, n, n)
        lib().lbo_build_cavity(n, _up(lat.words), ctypes.byref(lat.params), ulid)
        if solid is not None:
            lat.set_solid(solid)
            top = np.zeros((n, n, n), bool)
            top[n - 1] = True
            lat.words[top.reshape(-1) & lat.fluid] |= np.uint32(1 << CLASS_SHIFT)
        lat.params.tau = tau
        return lat

    @classmethod
    def channel(cls, nx, ny, nz, tau, umean, parabolic=False, solid=None):
        """The reference's default channel: no-slip y/z walls, velocity inlet at x-, pressure
        outlet rho = 1 at x+, optional solid obstacles."""
        lat = cls(nx, ny, nz)
        lat.set_solid(np.zeros(lat.n, np.uint8) if solid is None else solid)
        lat._bc_u = np.zeros(3 * lat.n)
        lib().lbo_build_channel(ctypes.byref(lat.dims), _up(lat.words), ctypes.byref(lat.params), float(umean),
                                int(bool(parabolic)), _dp(lat._bc_u))
        if lat.params.bc_type[0] != 0:
            raise ValueError("not a channel (nx < 2 or more than 31 boundary classes)")
        lat.params.tau = tau
        return lat

    def set_coupling(self, cells, mode, u, rho=None, dt=1.0):
        """MD coupling at the listed cells: mode COUPLE_OVERWRITE or COUPLE_DIRICHLET, MD velocity
        u[k] (lattice units) and optional density rho[k] for cells[k].  cells=None clears."""
        if cells is None:
            self.params.couple_row = None
            self._couple = None
            return
        cells = np.asarray(cells, dtype=np.int64).reshape(-1)
        row = np.full(self.n, -1, dtype=np.int64)
        row[cells] = np.arange(cells.size)
        u = np.ascontiguousarray(np.asarray(u, dtype=np.float64).reshape(-1))
        r = None if rho is None else np.ascontiguousarray(np.asarray(rho, dtype=np.float64).reshape(-1))
        self._couple = (row, u, r)
        self.params.couple_row = row.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
        self.params.couple_mode = int(mode)
        self.params.couple_u = _dp(u)
        self.params.couple_rho = None if r is None else _dp(r)
        self.params.couple_dt = float(dt)

    def set_solid(self, solid):
        solid = np.ascontiguousarray(solid.reshape(-1), dtype=np.uint8)
        lib().lbo_build_words(ctypes.byref(self.dims), solid.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                              _up(self.words))

    def set_force(self, g):
        self.params.use_force = 1
        for k in range(3):
            self.params.g[k] = float(g[k])

    def set_alpha(self, alpha):
        self._alpha = np.ascontiguousarray(alpha.reshape(-1), dtype=np.float64)
        self.params.alpha = _dp(self._alpha)

    def set_tau_cells(self, tau):
        """Per-cell configured relaxation time (the batched-patch case: tau per patch)."""
        self._tau_cell = np.ascontiguousarray(np.asarray(tau, dtype=np.float64).reshape(-1))
        assert self._tau_cell.size == self.n
        self.params.tau_cell = _dp(self._tau_cell)

    @property
    def fluid(self):
        return (self.words & FLUID) != 0

    # ----- state -----
    def equilibrium(self, rho=None, u=None):
        rho = np.ones(self.n) if rho is None else np.ascontiguousarray(rho.reshape(-1), dtype=np.float64)
        u = np.zeros((self.n, 3)) if u is None else np.ascontiguousarray(u.reshape(-1, 3), dtype=np.float64)
        f = np.empty(Q * self.n)
        lib().lbo_init_equilibrium(ctypes.byref(self.dims), _up(self.words), _dp(rho), _dp(u), _dp(f))
        return f

    def step(self, f, steps=1, nthreads=1):
        a = np.ascontiguousarray(f, dtype=np.float64).copy()
        b = np.zeros_like(a)
        b[:] = a  # non-fluid cells keep their values
        for _ in range(steps):
            lib().lbo_step(ctypes.byref(self.dims), _up(self.words), ctypes.byref(self.params), _dp(a), _dp(b),
                           int(nthreads))
            a, b = b, a
        return a

    def stream(self, f):
        src = np.ascontiguousarray(f, dtype=np.float64)
        dst = np.zeros_like(src)
        lib().lbo_stream(ctypes.byref(self.dims), _up(self.words), _dp(src), _dp(dst))
        return dst

    def step_swap(self, f_soa, steps=1):
        aos = np.ascontiguousarray(f_soa.reshape(Q, self.n).T)
        for _ in range(steps):
            lib().lbo_step_swap_aos(ctypes.byref(self.dims), _up(self.words), ctypes.byref(self.params), _dp(aos))
        return np.ascontiguousarray(aos.T).reshape(-1)

    def fields(self, f):
        rho = np.empty(self.n)
        u = np.empty((self.n, 3))
        lib().lbo_fields(ctypes.byref(self.dims), _up(self.words), _dp(np.ascontiguousarray(f)), _dp(rho), _dp(u))
        return rho, u

    def coupling_export(self, f, cells):
        """What getAndSetV sends to MD for `cells` from state f: (rho[k], u[k]) after its GET
        block (the coupling data set by set_coupling, if any, applied)."""
        cells = np.ascontiguousarray(np.asarray(cells, dtype=np.int64).reshape(-1))
        rho = np.empty(cells.size)
        u = np.empty((cells.size, 3))
        lib().lbo_couple_export(ctypes.byref(self.dims), _up(self.words), ctypes.byref(self.params),
                                _dp(np.ascontiguousarray(f)), cells.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                cells.size, _dp(rho), _dp(u))
        return rho, u

    def moments(self, f):
        out = np.empty(4)
        lib().lbo_moments(ctypes.byref(self.dims), _up(self.words), _dp(np.ascontiguousarray(f)), _dp(out))
        return out

    def pressure_drop(self, f, xin, xout):
        return lib().lbo_pressure_drop(ctypes.byref(self.dims), _up(self.words), _dp(np.ascontiguousarray(f)),
                                       int(xin), int(xout))

    def pressure_drop_grad(self, xin, xout):
        lam = np.empty(Q * self.n)
        lib().lbo_pressure_drop_grad(ctypes.byref(self.dims), _up(self.words), int(xin), int(xout), _dp(lam))
        return lam

    def step_adjoint(self, f_in, lam_out, galpha, nthreads=1):
        lam_in = np.empty(Q * self.n)
        lib().lbo_step_adjoint(ctypes.byref(self.dims), _up(self.words), ctypes.byref(self.params),
                               _dp(np.ascontiguousarray(f_in)), _dp(np.ascontiguousarray(lam_out)), _dp(lam_in),
                               _dp(galpha), int(nthreads))
        return lam_in


def tau_eff(tau):
    return lib().lbo_tau_eff(float(tau))


# ---- restriction / prolongation between grid levels on 6^3 blocks (AoS pdf[19 * cell + q]) ----
BLOCK_CELLS = 216


def _pos(pos):
    a = (ctypes.c_int * 3)(*[int(v) for v in pos])
    return ctypes.cast(a, ctypes.POINTER(ctypes.c_int)), a


def _u8p(a):
    assert a.dtype == np.uint8 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))


def amr_father(pos):
    p, _keep = _pos(pos)
    out = (ctypes.c_int * 3)()
    lib().lbo_amr_father(p, ctypes.cast(out, ctypes.POINTER(ctypes.c_int)))
    return tuple(out)


def amr_subblock(pos):
    p, _keep = _pos(pos)
    out = (ctypes.c_int * 3)()
    lib().lbo_amr_subblock(p, ctypes.cast(out, ctypes.POINTER(ctypes.c_int)))
    return tuple(out)


def amr_from_neighbour(nb, cell, q):
    return bool(lib().lbo_amr_from_neighbour(int(nb), int(cell), int(q)))


def amr_find_overlap(fine):
    fine = np.ascontiguousarray(fine, dtype=np.float64).reshape(-1)
    assert fine.size == BLOCK_CELLS * Q
    out = np.zeros(BLOCK_CELLS, np.uint8)
    lib().lbo_amr_find_overlap(_dp(fine), _u8p(out))
    return out.astype(bool)


def amr_prolongate(pos, coarse, fine, density, velocity, overlap):
    """Returns (fine, density, velocity) after the prolongation (inputs are not modified)."""
    p, _keep = _pos(pos)
    coarse = np.ascontiguousarray(coarse, dtype=np.float64).reshape(-1)
    fine = np.array(fine, dtype=np.float64).reshape(-1)
    density = np.array(density, dtype=np.float64).reshape(-1)
    velocity = np.array(velocity, dtype=np.float64).reshape(-1)
    ov = np.ascontiguousarray(np.asarray(overlap).reshape(-1) != 0, dtype=np.uint8)
    assert coarse.size == fine.size == BLOCK_CELLS * Q and density.size == BLOCK_CELLS and velocity.size == 3 * BLOCK_CELLS
    lib().lbo_amr_prolongate(p, _dp(coarse), _dp(fine), _dp(density), _dp(velocity), _u8p(ov))
    return fine, density, velocity


def amr_restrict(pos, coarse, fine):
    """Returns the coarse block after the restriction (inputs are not modified)."""
    p, _keep = _pos(pos)
    coarse = np.array(coarse, dtype=np.float64).reshape(-1)
    fine = np.ascontiguousarray(fine, dtype=np.float64).reshape(-1)
    assert coarse.size == fine.size == BLOCK_CELLS * Q
    lib().lbo_amr_restrict(p, _dp(coarse), _dp(fine))
    return coarse


# ---- the block solver's per-block and per-Peano-cell operators (6^3 blocks, AoS pdf[19 * cell + q]) ----
BLOCK_PDFS = BLOCK_CELLS * Q


def _block(a, n, dtype=np.float64):
    a = np.ascontiguousarray(a, dtype=dtype).reshape(-1)
    assert a.size == n
    return a


def _flags(a):
    return np.ascontiguousarray(np.asarray(a).reshape(-1) != 0, dtype=np.uint8)


def block_cell_values(inner, pdf, density=None, minus_one=None):
    """computeDensity + computeVelocity; returns (density, velocity).  density holds the values of
    the cells that are skipped (zeros if None)."""
    inn = _flags(inner)
    mo = _flags(minus_one) if minus_one is not None else None
    pdf = _block(pdf, BLOCK_PDFS)
    rho = np.zeros(BLOCK_CELLS) if density is None else np.array(density, dtype=np.float64).reshape(-1)
    u = np.zeros(3 * BLOCK_CELLS)
    lib().lbo_block_cell_values(_u8p(inn), _u8p(mo) if mo is not None else None, _dp(pdf), _dp(rho), _dp(u))
    return rho, u


def block_collide(inner, pdf, density, velocity, tau, minus_one=None):
    """BGK collision of the inner cells; returns the new pdf (input not modified)."""
    inn = _flags(inner)
    mo = _flags(minus_one) if minus_one is not None else None
    out = np.array(pdf, dtype=np.float64).reshape(-1)
    lib().lbo_block_collide(_u8p(inn), _u8p(mo) if mo is not None else None, _dp(_block(density, BLOCK_CELLS)),
                            _dp(_block(velocity, 3 * BLOCK_CELLS)), float(tau), _dp(out))
    return out


def block_boundary(pdf, density, velocity, links):
    """Boundary links after the collision; links: iterable of (cell, q_in, type, (u0, u1, u2), rho_l).
    Returns the new pdf (input not modified)."""
    out = np.array(pdf, dtype=np.float64).reshape(-1)
    rho = _block(density, BLOCK_CELLS)
    vel = _block(velocity, 3 * BLOCK_CELLS)
    for cell, q_in, ty, ul, rl in links:
        uu = np.ascontiguousarray(ul, dtype=np.float64)
        lib().lbo_block_boundary(int(cell), int(q_in), int(ty), _dp(uu), float(rl), _dp(rho), _dp(vel), _dp(out))
    return out


def block_local_index():
    buf = (ctypes.c_int * (4 * 2048))()
    n = lib().lbo_block_local_index(ctypes.cast(buf, ctypes.POINTER(ctypes.c_int)), 2048)
    return np.frombuffer(buf, np.int32)[:4 * n].reshape(n, 4).copy()


def block_nonlocal_index():
    buf = (ctypes.c_int * (5 * 1024))()
    n = lib().lbo_block_nonlocal_index(ctypes.cast(buf, ctypes.POINTER(ctypes.c_int)), 1024)
    return np.frombuffer(buf, np.int32)[:5 * n].reshape(n, 5).copy()


def block_stream_local(inner, pdf):
    out = np.array(pdf, dtype=np.float64).reshape(-1)
    assert out.size == BLOCK_PDFS
    lib().lbo_block_stream_local(_u8p(_flags(inner)), _dp(out))
    return out


def block_swap_inside(pdf):
    out = np.array(pdf, dtype=np.float64).reshape(-1)
    assert out.size == BLOCK_PDFS
    lib().lbo_block_swap_inside(_dp(out))
    return out


def block_stream_nonlocal(inner, pdfs):
    """streamingNonLocal on the 8 blocks of one Peano cell; inner: (8, 216), pdfs: (8, 4104) arrays.
    Returns the 8 new blocks (inputs not modified)."""
    inn = [_flags(inner[v]) for v in range(8)]
    out = [np.array(pdfs[v], dtype=np.float64).reshape(-1) for v in range(8)]
    ip = (ctypes.POINTER(ctypes.c_uint8) * 8)(*[_u8p(a) for a in inn])
    pp = (ctypes.POINTER(ctypes.c_double) * 8)(*[_dp(a) for a in out])
    lib().lbo_block_stream_nonlocal(ip, pp)
    return np.stack(out)


# the vertex-overlap exchange between ranks (BLB/RegularGridBlockVertex.cpp:51-160)
BLOCK_BOUNDARY_PDFS = 1008
BLOCK_BOUNDARY_CELLS = 208


def block_pdf_boundary():
    """(incoming, outgoing) slots of the populations that leave a block, in the reference's order."""
    a = (ctypes.c_int * 2048)()
    b = (ctypes.c_int * 2048)()
    n = lib().lbo_block_pdf_boundary(ctypes.cast(a, ctypes.POINTER(ctypes.c_int)),
                                     ctypes.cast(b, ctypes.POINTER(ctypes.c_int)), 2048)
    return np.frombuffer(a, np.int32)[:n].copy(), np.frombuffer(b, np.int32)[:n].copy()


def block_boundary_cells():
    a = (ctypes.c_int * 256)()
    n = lib().lbo_block_boundary_cells(ctypes.cast(a, ctypes.POINTER(ctypes.c_int)), 256)
    return np.frombuffer(a, np.int32)[:n].copy()


def block_init_pdf_diff(pdf):
    diff = np.zeros(BLOCK_BOUNDARY_PDFS)
    lib().lbo_block_init_pdf_diff(_dp(_block(pdf, BLOCK_PDFS)), _dp(diff))
    return diff


def block_compute_pdf_diff(pdf, diff):
    out = np.array(diff, dtype=np.float64).reshape(-1)
    assert out.size == BLOCK_BOUNDARY_PDFS
    lib().lbo_block_compute_pdf_diff(_dp(_block(pdf, BLOCK_PDFS)), _dp(out))
    return out


def block_merge_pdf_diff(pdf, own, neighbour):
    """Returns the merged block (input not modified)."""
    out = np.array(pdf, dtype=np.float64).reshape(-1)
    assert out.size == BLOCK_PDFS
    lib().lbo_block_merge_pdf_diff(_dp(out), _dp(_block(own, BLOCK_BOUNDARY_PDFS)),
                                   _dp(_block(neighbour, BLOCK_BOUNDARY_PDFS)))
    return out


def block_set_density_on_boundary(density, inner, boundary, dynamic_geometry, previous=None):
    out = np.zeros(BLOCK_BOUNDARY_CELLS) if previous is None else np.array(previous, dtype=np.float64).reshape(-1)
    lib().lbo_block_set_density_on_boundary(_dp(_block(density, BLOCK_CELLS)), _u8p(_flags(inner)),
                                            _u8p(_flags(boundary)), 1 if dynamic_geometry else 0, _dp(out))
    return out


def block_merge_density_on_boundary(density, inner, boundary, neighbour):
    out = np.array(density, dtype=np.float64).reshape(-1)
    assert out.size == BLOCK_CELLS
    lib().lbo_block_merge_density_on_boundary(_dp(out), _u8p(_flags(inner)), _u8p(_flags(boundary)),
                                              _dp(_block(neighbour, BLOCK_BOUNDARY_CELLS)))
    return out
