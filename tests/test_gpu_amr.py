"""Restriction and prolongation between grid levels on the device block store, through the C-ABI,
against the reference's own outputs (tests/golden/amr_s*.npz, from BLB/BlockRestrictionProlongation.cpp
compiled here) and the oracle: bit for bit in FP64, for all 64 child positions in one batched call
per operation, plus a large batch checked against the oracle on sampled links."""
import hashlib
import os

import numpy as np
import pytest

import amr_inputs
import oracle as O

pytestmark = pytest.mark.gpu
lbm = pytest.importorskip("madpac_amd.lbm")

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NC, Q = 216, 19


def _fixture(seed):
    return np.load(os.path.join(GOLDEN, f"amr_s{seed}.npz"), allow_pickle=False)


@pytest.mark.parametrize("seed", [7, 8, 9])
def test_block_store_bitwise_equals_reference(seed):
    d = _fixture(seed)
    cases = amr_inputs.inputs(seed, float(d["pmark"]))
    print(f"[gpu] block store: 64 child positions, seed {seed}", flush=True)
    # blocks 0..63 coarse, 64..127 fine: one call per operation covers every position
    with lbm.BlockStore(128) as bs:
        bs.set(0, pdf=np.stack([c["coarse"] for c in cases]))
        bs.set(64, pdf=np.stack([c["fine"] for c in cases]), density=np.stack([c["density"] for c in cases]),
               velocity=np.stack([c["velocity"] for c in cases]))
        links = [(p, 64 + p, c["pos"], p == 21) for p, c in enumerate(cases)]
        ov = bs.find_overlap(links)
        assert np.array_equal(ov, d["overlap"])
        bs.prolongate(links)
        pdf, rho, u = bs.get()
    for p in range(64):
        h = hashlib.sha256(pdf[64 + p].tobytes() + rho[64 + p].tobytes() + u[64 + p].tobytes()).hexdigest()
        assert h == str(d["prolongation_sha256"][p]), p
    # restriction on the untouched fine blocks (markers as drawn); restriction writes one sub-block
    # per link, so two positions with the same father sub-block go in separate calls
    rest = {}
    with lbm.BlockStore(128) as bs:
        bs.set(0, pdf=np.stack([c["coarse"] for c in cases]))
        bs.set(64, pdf=np.stack([c["fine"] for c in cases]))
        bs.restrict([(p, 64 + p, c["pos"]) for p, c in enumerate(cases)])
        rest = bs.get(0, 64)[0]
    for p in range(64):
        assert hashlib.sha256(rest[p].tobytes()).hexdigest() == str(d["restriction_sha256"][p]), p


def test_block_store_large_batch_against_oracle():
    """Two refined levels' worth of links in one launch each (8,000 fine blocks under 1,000 coarse
    blocks, every child position used), sampled links against the oracle."""
    rng = np.random.default_rng(11)
    ncoarse, per = 1000, 8
    nfine = ncoarse * per
    n = ncoarse + nfine
    coarse = rng.uniform(0.5, 1.5, (ncoarse, NC * Q))
    fine = rng.uniform(0.5, 1.5, (nfine, NC * Q))
    fine[rng.random(fine.shape) < 0.001] = -1.0
    dens = rng.uniform(0.5, 1.5, (nfine, NC))
    vel = rng.uniform(-0.5, 0.5, (nfine, 3 * NC))
    # each coarse block fathers 8 fine vertices at the positions of one octant of the refined cell
    # (father (fx, fy, fz), positions 2 f + {0, 1} per axis: distinct sub-blocks)
    links = []
    for cb in range(ncoarse):
        f = (cb % 2, (cb // 2) % 2, (cb // 4) % 2)
        for k in range(per):
            pos = tuple(2 * f[d] + ((k >> d) & 1) for d in range(3))
            links.append((cb, ncoarse + cb * per + k, pos, (cb * per + k) % 97 == 0))
    print(f"[gpu] block store: {len(links)} links over {n} blocks", flush=True)
    with lbm.BlockStore(n) as bs:
        bs.set(0, pdf=coarse)
        bs.set(ncoarse, pdf=fine, density=dens, velocity=vel)
        ov = bs.find_overlap(links)
        bs.prolongate(links)
        bs.restrict(links)
        pdf, rho, u = bs.get()
    for k in rng.choice(len(links), 40, replace=False):
        cb, fb, pos, new = links[k]
        fi = fb - ncoarse
        f0 = -np.ones(NC * Q) if new else fine[fi]
        o = np.ones(NC, bool) if new else O.amr_find_overlap(f0)
        assert np.array_equal(ov[k], o)
        fo, ro, uo = O.amr_prolongate(pos, coarse[cb], f0, dens[fi], vel[fi], o)
        assert np.array_equal(pdf[fb], fo) and np.array_equal(rho[fb], ro) and np.array_equal(u[fb], uo)
    # restriction: each coarse block received its 8 sub-blocks from the prolongated fine blocks
    for cb in rng.choice(ncoarse, 20, replace=False):
        c = coarse[cb].copy()
        for k in range(per):
            _, fb, pos, new = links[cb * per + k]
            fi = fb - ncoarse
            f0 = -np.ones(NC * Q) if new else fine[fi]
            o = np.ones(NC, bool) if new else O.amr_find_overlap(f0)
            fo, _, _ = O.amr_prolongate(pos, coarse[cb], f0, dens[fi], vel[fi], o)
            c = O.amr_restrict(pos, c, fo)
        assert np.array_equal(pdf[cb], c), cb


def test_block_store_rejects_conflicting_links():
    with lbm.BlockStore(4) as bs:
        with pytest.raises(lbm.LbmError, match="out of range"):
            bs.restrict([(0, 9, (0, 0, 0))])
        with pytest.raises(lbm.LbmError, match="same"):
            bs.restrict([(1, 1, (0, 0, 0))])
        with pytest.raises(lbm.LbmError, match="0..3"):
            bs.prolongate([(0, 1, (0, 4, 0))])
        with pytest.raises(lbm.LbmError, match="two links"):
            bs.prolongate([(0, 1, (0, 0, 0)), (2, 1, (1, 0, 0))])
        # positions 0 and 3 share sub-block 1 (of different fathers): one coarse block cannot take both
        with pytest.raises(lbm.LbmError, match="sub-block"):
            bs.restrict([(0, 1, (0, 0, 0)), (0, 2, (3, 3, 3))])
        with pytest.raises(lbm.LbmError, match="level pair"):
            bs.restrict([(0, 1, (0, 0, 0)), (1, 2, (1, 0, 0))])
        with pytest.raises(ValueError):
            bs.set(0, pdf=np.zeros(100))


# ---- the reference-shaped C++ class (peano_blocklb.h BlockRestrictionProlongation) ----
import subprocess  # noqa: E402

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
AMR_EXE = os.path.join(ROOT, "tests", "cpp", "test_peano_amr")


def _run_cpp(tmp_path, cases):
    """cases: dicts with pos, coarse, fine, density, velocity and optionally mask (given overlap)."""
    inp, outp = tmp_path / "amr_in.bin", tmp_path / "amr_out.bin"
    with open(inp, "wb") as f:
        f.write(np.int32(len(cases)).tobytes())
        for c in cases:
            given = "mask" in c
            f.write(np.asarray(c["pos"], np.int32).tobytes() + np.int32(given).tobytes())
            f.write((np.asarray(c["mask"]) if given else np.zeros(NC, bool)).astype(np.uint8).tobytes())
            for k in ("coarse", "fine", "density", "velocity"):
                f.write(np.ascontiguousarray(c[k], np.float64).tobytes())
    subprocess.run([AMR_EXE, str(inp), str(outp)], check=True, timeout=120)
    raw = open(outp, "rb").read()
    res, off = [], 0
    for _ in cases:
        ints = np.frombuffer(raw[off:off + 24], np.int32)
        off += 24
        ov = np.frombuffer(raw[off:off + NC], np.uint8).astype(bool)
        off += NC
        arrs = []
        for size in (NC * Q, NC, 3 * NC, NC * Q):
            arrs.append(np.frombuffer(raw[off:off + 8 * size], "<f8"))
            off += 8 * size
        res.append(dict(father=tuple(ints[:3]), sub=tuple(ints[3:]), overlap=ov, fine=arrs[0], density=arrs[1],
                        velocity=arrs[2], coarse=arrs[3]))
    assert off == len(raw)
    return res


def test_cpp_block_restriction_prolongation_equals_reference(tmp_path):
    seed = 7
    d = _fixture(seed)
    cases = amr_inputs.inputs(seed, float(d["pmark"]))
    print("[gpu] C++ BlockRestrictionProlongation: 64 child positions", flush=True)
    res = _run_cpp(tmp_path, cases)
    for p, r in enumerate(res):
        assert r["father"] == tuple(d["father"][p]) and r["sub"] == tuple(d["subblock"][p])
        assert np.array_equal(r["overlap"], d["overlap"][p])
        h = hashlib.sha256(r["fine"].tobytes() + r["density"].tobytes() + r["velocity"].tobytes()).hexdigest()
        assert h == str(d["prolongation_sha256"][p]), p
        assert hashlib.sha256(r["coarse"].tobytes()).hexdigest() == str(d["restriction_sha256"][p]), p


def test_cpp_prolongation_reference_unit_case(tmp_path):
    """BLB/tests/BlockRestrictionProlongationTest.cpp:41-146 through the C++ class: a given overlap
    set (cells x, y, z < 3) that no overlap search produces, child position (1, 1, 1)."""
    mask = np.zeros(NC, bool)
    for z in range(3):
        for y in range(3):
            for x in range(3):
                mask[x + 6 * y + 36 * z] = True
    r = _run_cpp(tmp_path, [dict(pos=(1, 1, 1), mask=mask, coarse=np.ones(NC * Q), fine=-np.ones(NC * Q),
                                 density=np.ones(NC), velocity=np.full(3 * NC, 0.05))])[0]
    f = r["fine"].reshape(NC, Q)
    u = r["velocity"].reshape(NC, 3)
    for c in range(NC):
        if mask[c]:
            assert r["density"][c] == 19.0 and not u[c].any() and (f[c] == 1.0).all()
        else:
            assert r["density"][c] == 1.0 and (u[c] == 0.05).all() and (f[c] == -1.0).all()
