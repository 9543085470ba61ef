"""The seeded rounded-box stone wall (geometry.rounded_box_wall) against a direct restatement
that recomposes the whole mask after every stone removal: identical masks and info.  The
library keeps a coverage count over the wall's x window and updates the void fraction from
the removed stone's own cells, so generating an H = 40 wall no longer scales with the
number of removals times the lattice size."""
import time

import numpy as np
import pytest

from stonewall3d import geometry


def recomposed_wall(nx, ny, nz, X0, T, H, void_fraction, seed, size=(0.3, 0.6)):
    """The same draws; the mask rebuilt from the kept stones after each removal."""
    rng = np.random.default_rng(seed)
    lo, hi = size
    courses, z = [], 0.0
    while z < H - 0.5 * lo * H:
        h = rng.uniform(lo, hi) * H
        courses.append((z, min(h, H - z)))
        z += h
    stones = []
    for ci, (z0, h) in enumerate(courses):
        y = rng.uniform(0, 0.5) * hi * H + (0.5 * hi * H if ci % 2 else 0.0)
        start = y
        while y < start + ny - 0.5 * lo * H:
            L = min(rng.uniform(lo, hi) * H, start + ny - y)
            stones.append(dict(yc=y + L / 2, L=L, zc=z0 + h / 2, h=h))
            y += L
    xc, hx = X0 + T / 2, T / 2
    masks = []
    for st in stones:
        r = 0.25 * min(hx, st["L"] / 2, st["h"] / 2)
        m = np.zeros((nx, ny, nz), bool)
        for i in range(max(0, X0 - 1), min(nx, X0 + T + 1)):
            for k in range(max(1, int(st["zc"] - st["h"] / 2) - 1), min(nz, int(st["zc"] + st["h"] / 2) + 2)):
                for j in range(int(st["yc"] - st["L"] / 2) - 1, int(st["yc"] + st["L"] / 2) + 2):
                    dy = ((j % ny) + 0.5 - st["yc"] + ny / 2) % ny - ny / 2
                    q = [max(abs(v) - (hw - r), 0.0) for v, hw in
                         ((i + 0.5 - xc, hx), (dy, st["L"] / 2), (k - 0.5 - st["zc"], st["h"] / 2))]
                    if q[0] * q[0] + q[1] * q[1] + q[2] * q[2] <= r * r:
                        m[i, j % ny, k] = True
        masks.append(m)

    def compose(keep):
        m = np.zeros((nx, ny, nz), bool)
        for k in keep:
            m |= masks[k]
        return m

    def void(m):
        return 1.0 - m[X0:X0 + T, :, 1:H + 1].sum() / (T * ny * H)

    order = list(rng.permutation(len(stones)))
    keep = set(range(len(stones)))
    m = compose(keep)
    while void(m) < void_fraction and order:
        keep.discard(order.pop(0))
        m = compose(keep)
    return m, dict(stones=len(stones), removed=len(stones) - len(keep), void_fraction=float(void(m)))


@pytest.mark.parametrize("args", [
    (96, 48, 24, 24, 4, 6, 0.2, 0),            # 16H x 8H x 4H at H = 6
    (96, 48, 24, 24, 4, 6, 0.45, 5),
    (80, 40, 20, 20, 3, 5, 0.0, 1),            # nothing removed
    (64, 37, 18, 10, 5, 8, 0.3, 3),            # ragged width, thick wall
    (40, 20, 12, 0, 4, 6, 0.25, 2),            # wall at the inlet (x window clipped at 0)
    (40, 20, 12, 37, 5, 6, 0.25, 4),           # wall at the outlet (x window clipped at nx)
    (60, 30, 10, 20, 3, 9, 0.6, 7),            # wall as tall as the lattice
])
def test_rounded_box_wall_matches_recomposition(args):
    nx, ny, nz, X0, T, H, void, seed = args
    m, info = geometry.rounded_box_wall(nx, ny, nz, X0, T, H, void, seed)
    r, rinfo = recomposed_wall(nx, ny, nz, X0, T, H, void, seed)
    assert np.array_equal(m, r)
    assert info["stones"] == rinfo["stones"] and info["removed"] == rinfo["removed"]
    assert info["void_fraction"] == rinfo["void_fraction"]
    if void > 0:
        assert info["void_fraction"] >= min(void, 1.0) or info["removed"] == info["stones"]


def test_rounded_box_wall_size_variant_matches_recomposition():
    m, info = geometry.rounded_box_wall(48, 13, 18, 12, 5, 12, 0.3, 3, size=(0.4, 0.8))
    r, rinfo = recomposed_wall(48, 13, 18, 12, 5, 12, 0.3, 3, size=(0.4, 0.8))
    assert np.array_equal(m, r) and info["removed"] == rinfo["removed"]


def test_rounded_box_wall_h40_time():
    """The 640x320x160 wall (H = 40, 30 % void: about 10 stones removed) within 0.1 s."""
    t = time.perf_counter()
    m, info = geometry.rounded_box_wall(640, 320, 161, 160, 24, 40, 0.3, 1)
    el = time.perf_counter() - t
    assert info["removed"] > 5 and abs(info["void_fraction"] - 0.3) < 0.05
    assert el < 0.1, f"{el:.3f} s"
