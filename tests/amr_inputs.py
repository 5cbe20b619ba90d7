"""Seeded inputs of the restriction/prolongation fixtures (tests/golden/amr_s*.npz).

The same 64-bit LCG (Knuth MMIX constants, top 53 bits to [0, 1)) and the same draw order as the
reference harness that produced the fixtures: per child position p in 0..63, pos = (p % 4,
(p // 4) % 4, p // 16), the coarse block (0.5 + u), the fine block (-1.0 with probability pmark,
else 0.5 + u; every entry -1.0 at p = 21, a new persistent vertex), the fine density (0.5 + u) and
velocity (u - 0.5).  Blocks are 6^3 cells, populations AoS pdf[19 * cell + q]."""
import numpy as np

MASK = (1 << 64) - 1
NC, Q = 216, 19


class Lcg:
    def __init__(self, seed):
        self.s = (seed * 0x9E3779B97F4A7C15 + 1) & MASK

    def u(self):
        self.s = (self.s * 6364136223846793005 + 1442695040888963407) & MASK
        return (self.s >> 11) * (1.0 / 9007199254740992.0)


def positions():
    return [(p % 4, (p // 4) % 4, p // 16) for p in range(64)]


def inputs(seed, pmark):
    """List of 64 dicts (pos, coarse, fine, density, velocity) in position order."""
    g = Lcg(seed)
    out = []
    for p, pos in enumerate(positions()):
        coarse = np.array([0.5 + g.u() for _ in range(NC * Q)])
        fine = np.empty(NC * Q)
        for k in range(NC * Q):
            r = g.u()
            v = 0.5 + g.u()
            fine[k] = -1.0 if (p == 21 or r < pmark) else v
        density = np.array([0.5 + g.u() for _ in range(NC)])
        velocity = np.array([g.u() - 0.5 for _ in range(3 * NC)])
        out.append(dict(pos=pos, coarse=coarse, fine=fine, density=density, velocity=velocity))
    return out
