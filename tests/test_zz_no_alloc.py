"""lbm_reduce_moments_noalloc (the allocation-free moment reduction) against lbm_reduce_moments,
through the C-ABI: the same bits on two calls in a row, the second on a smaller lattice, in both
precisions and both cell orders; lattices beyond the fixed array are rejected with an error."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

lbm = pytest.importorskip("madpac_amd.lbm")


def _porous(n, seed):
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    solid = np.zeros((n, n, n), bool)
    for _ in range(n // 3):
        c = rng.integers(0, n, 3)
        r = rng.integers(2, 5)
        solid |= (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2 <= r * r
    return solid


def _check(lat):
    a = lat.moments()
    b = lat.moments_noalloc()
    assert np.array_equal(a, b), (a, b)
    # the fixed array keeps the last call's contents; a repeat must not depend on them
    assert np.array_equal(lat.moments_noalloc(), a)
    return a


@pytest.mark.parametrize("prec", [lbm.FP64, lbm.FP32_SHIFTED])
def test_noalloc_moments_equal_the_allocating_reduction(prec):
    n1, n2 = 72, 20
    print(f"[gpu] noalloc moments: porous {n1}^3 then cavity {n2}^3, precision {prec}", flush=True)
    rng = np.random.default_rng(5)
    with lbm.Lattice(n1, n1, n1, 0.6, precision=prec, periodic=(True, True, True), force=(1e-5, 0, 0)) as lat:
        lat.set_flags(~_porous(n1, 3))
        lat.init_equilibrium(u=0.02 * rng.standard_normal((n1 ** 3, 3)))
        lat.step(7)
        m1 = _check(lat)
        assert m1[1] != 0.0
    with lbm.cavity(n2, 0.7, 0.05, precision=prec) as lat:
        lat.step(13)
        m2 = _check(lat)
        assert abs(m2[0] - n2 ** 3) < 1e-9 * n2 ** 3
    # alternating with a context that is still alive on the same device
    with lbm.cavity(n2, 0.7, 0.05, precision=prec) as a, lbm.Lattice(n1, n1, n1, 0.6, precision=prec,
                                                                    periodic=(True, True, True)) as b:
        b.set_flags(np.ones(n1 ** 3, np.uint8))
        b.init_equilibrium(u=0.01 * rng.standard_normal((n1 ** 3, 3)))
        a.step(13)
        assert np.array_equal(a.moments_noalloc(), m2)
        _check(b)
        assert np.array_equal(a.moments_noalloc(), m2)


def test_noalloc_moments_in_place_layout_and_bricks():
    n = 40
    solid = _porous(n, 9)
    with lbm.Lattice(n, n, n, 0.8, periodic=(True, True, True), layout=lbm.LAYOUT_AA,
                     cell_order=lbm.CELL_ORDER_BRICKS) as lat:
        lat.set_flags(~solid)
        lat.init_equilibrium(u=0.02 * np.random.default_rng(1).standard_normal((n ** 3, 3)))
        for s in (1, 2):  # both parities of the in-place layout
            lat.step(s)
            _check(lat)


def test_noalloc_rejects_lattices_beyond_the_array():
    # 1 x 1 x nz rows: one reduction workgroup per cell, so a small lattice needs many partials
    nz = 530000
    with lbm.Lattice(1, 1, nz, 0.7, precision=lbm.FP32_SHIFTED, periodic=(True, True, True)) as lat:
        lat.set_flags(np.ones(nz, np.uint8))
        lat.init_equilibrium()
        with pytest.raises(lbm.LbmError, match="too large"):
            lat.moments_noalloc()
        m = lat.moments()
        assert abs(m[0] - nz) < 1e-6 * nz
