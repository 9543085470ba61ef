"""The geometry restatement against the reference's own published data and golden masks.

Pins (no import of reference code at test time):
  - results/3d/stone_wall_3d.html:150: front silhouettes of build_wall(p, 40, 800, 160,
    321, 160, 12, seed=7) for p = 0, 0.2, 0.4 (bit-exact), their porosity and top line;
  - tests/golden/geometry_golden.npz: full 3D masks written by tests/golden/
    make_geometry_golden.py from the reference's geometry.py at H = 10, 12, 16, 24.
"""
import base64
import gzip
import json
import os

import numpy as np
import pytest

from stonewall3d import geometry
from stonewall3d.inflow import reference_setup

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF = json.load(open(os.path.join(GOLD, "reference_results.json")))


def unpack_front(front):
    raw = gzip.decompress(base64.b64decode(front["bits"]))
    bits = np.unpackbits(np.frombuffer(raw, np.uint8))[: front["ny"] * front["nz"]]
    return bits.reshape(front["nz"], front["ny"])[::-1].T.astype(bool)      # (ny, nz)


@pytest.fixture(scope="module")
def h40_walls():
    s = reference_setup(H=40)
    out = {}
    for case in REF["cases"]:
        p = {"solid": 0.0, "p20": 0.2, "p40": 0.4}[case["key"]]
        out[case["key"]] = geometry.build_wall(p, 40, s.nx, s.ny, s.nz, s.X0, s.T, seed=7, verbose=False)
    return out


def test_reference_info_matches_setup():
    s = reference_setup(H=40)
    info = REF["info"]
    assert [s.nx, s.ny, s.nz] == info["grid"]
    assert (s.n_avg / s.t_unit) == pytest.approx(info["tAvg"], abs=1e-9)
    assert s.n_spin == 11699 and s.n_avg == 17548 and s.snap_every == 146
    assert 3 * s.nu0 + 0.5 == pytest.approx(0.500547, abs=1e-6)


@pytest.mark.parametrize("key", ["solid", "p20", "p40"])
def test_front_silhouette_bit_exact(h40_walls, key):
    case = next(c for c in REF["cases"] if c["key"] == key)
    m, phi = h40_walls[key]
    wf = m.any(axis=0)
    front = wf[:, : int(1.25 * 40)]
    ref = unpack_front(case["front"])
    assert front.shape == ref.shape
    assert np.array_equal(front, ref), f"{(front != ref).sum()} silhouette cells differ"
    assert round(float(phi), 3) == case["porosity"]
    tops = np.array([np.nonzero(wf[j])[0].max() - 0.5 if wf[j].any() else 0.0 for j in range(wf.shape[0])]) / 40
    assert tops.mean() == pytest.approx(case["topMean"], abs=1e-12)
    assert tops.max() == pytest.approx(case["topMax"], abs=1e-12)


def test_silhouette_decoder_counts():
    # solid cells in the published silhouettes (decoded independently of the generator)
    counts = {c["key"]: int(unpack_front(c["front"]).sum()) for c in REF["cases"]}
    assert counts == {"solid": 6400, "p20": 5202, "p40": 3913}


IDX = json.load(open(os.path.join(GOLD, "geometry_golden.json")))


@pytest.mark.parametrize("case", IDX, ids=[c["key"] for c in IDX])
def test_build_wall_matches_reference_masks(case):
    gold = np.load(os.path.join(GOLD, "geometry_golden.npz"))
    nx, ny, nz = case["grid"]
    m, phi = geometry.build_wall(case["porosity"], case["H"], nx, ny, nz, case["X0"], case["T"], seed=case["seed"],
                                 verbose=False)
    x0, x1 = case["x_range"]
    assert not m[:x0].any() and not m[x1:].any()
    ref = np.unpackbits(gold[case["key"]])[: (x1 - x0) * ny * nz].reshape(x1 - x0, ny, nz).astype(bool)
    assert np.array_equal(m[x0:x1], ref)
    assert phi == case["phi"]
    assert int(m.sum()) == case["solid_cells"]


def test_vortex_frame_dims_match_published():
    """The Q-frame crop of lbm3d.py:308-310 at H=40 gives the published vortex dims."""
    s = reference_setup(H=40)
    ds3 = max(1, int(round(40 / 13)))
    q_x0 = max(1, s.X0 - int(round(0.5 * 40)))
    dims = [min(int(round(9.0 * 40)), s.nx - 1 - q_x0) // ds3, s.ny // ds3, min(int(round(2.5 * 40)), s.nz - 1) // ds3]
    assert dims == REF["cases"][0]["vortex_dims"] == [120, 53, 33]
