"""Golden vectors for build_wall at sizes the published silhouettes do not cover.

The reference's published data pins only the 2D front silhouette of
build_wall(p, 40, 800, 160, 321, 160, 12, seed=7) for p in {0, 0.2, 0.4}.  The full 3D
mask, other wall heights and other seeds are pinned by nothing the reference publishes,
so this script imports the reference's own stonewall3d/geometry.py (numpy only; run it in
a container that holds the reference) and records its output for a few small cases:
the wall's bounding box x in [X0-2, X0+T+2) as packed bits, plus the frontal porosity.
Grids follow lbm3d.py:270-271 with the default --up 4 --down 16 --span 4 --height 8
--thick 0.3.  Output: tests/golden/geometry_golden.npz (+ a JSON index).
"""
import importlib.util
import json
import os
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
CASES = [(H, p, seed) for H in (10, 12, 16) for p in (0.0, 0.2, 0.4) for seed in (0, 7)] + [(24, 0.3, 3)]


def grid(H):
    nx, ny, nz = int(round(20 * H)), int(round(4 * H)), int(round(8 * H)) + 1
    X0, T = int(round(4 * H)), max(2, int(round(0.3 * H)))
    return nx, ny, nz, X0, T


def main():
    spec = importlib.util.spec_from_file_location("ref_geometry", os.path.join(REF, "stonewall3d/geometry.py"))
    g = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(g)
    arrays, index = {}, []
    for H, p, seed in CASES:
        nx, ny, nz, X0, T = grid(H)
        m, phi = g.build_wall(p, H, nx, ny, nz, X0, T, seed=seed, verbose=False)
        x0, x1 = X0 - 2, X0 + T + 2
        assert not m[:x0].any() and not m[x1:].any()
        key = f"H{H}_p{int(round(p * 100)):02d}_s{seed}"
        arrays[key] = np.packbits(m[x0:x1].ravel())
        index.append(dict(key=key, H=H, porosity=p, seed=seed, grid=[nx, ny, nz], X0=X0, T=T, x_range=[x0, x1],
                          phi=float(phi), solid_cells=int(m.sum())))
    np.savez_compressed(os.path.join(HERE, "geometry_golden.npz"), **arrays)
    with open(os.path.join(HERE, "geometry_golden.json"), "w") as f:
        json.dump(index, f, indent=1)
    print("wrote", len(index), "cases")


if __name__ == "__main__":
    main()
