"""Statistics restated from postprocess.py:65-82 and export3d.py:169-184, pinned by
closed forms on synthetic fields (no GPU)."""
import numpy as np
import pytest

from stonewall3d import stats


def test_coefficients_closed_form():
    # two heights with constant force rows: Fx = 3, Fz = -1, moment = sum f_x(z) * (z - 1/2)
    F = np.zeros((5, 4, 2))
    F[:, 1, 0], F[:, 2, 0], F[:, 2, 1] = 1.0, 2.0, -1.0
    Cd, Cl, Cm = stats.coefficients(F, UH=0.1, H=2, ny=3)
    q = 0.5 * 0.01 * 2 * 3
    assert np.allclose(Cd, 3.0 / q) and np.allclose(Cl, -1.0 / q)
    assert np.allclose(Cm, (1.0 * 0.5 + 2.0 * 1.5) / (q * 2))


def test_force_stats_window():
    Cd = np.array([9.0, 9.0, 1.0, 2.0, 3.0])
    s = stats.force_stats(Cd, Cd * 0.5, -Cd, n_spin=2)
    assert s["Cd"] == 2.0 and s["CdMax"] == 3.0 and s["CdStd"] == pytest.approx(np.sqrt(2 / 3))
    assert s["Cm"] == 1.0 and s["CmMax"] == 1.5 and s["Cl"] == -2.0


def test_blur_preserves_constants_and_mean_interior():
    a = np.full((5, 6, 7), 2.5, np.float32)
    assert np.allclose(stats.blur3(a), 2.5)
    b = np.zeros((9, 9, 9), np.float32)
    b[4, 4, 4] = 27.0
    out = stats.blur3(b)
    assert out[4, 4, 4] == pytest.approx(1.0) and out[3, 3, 3] == pytest.approx(1.0)
    assert out.sum() == pytest.approx(27.0)


def test_recirculation_area_slab():
    """A reversed-flow slab z < h over the whole crop: the one-pass blur of a step keeps the
    step, so the area is the slab volume (8 cells per downsampled cell) / (H^2 ny)."""
    H, X0, UH = 8, 16, 0.05
    nx, ny, nz = 16 * H, 8, 4 * H
    ux = np.full((nx, ny, nz), UH, np.float32)
    h = 6                                   # full-resolution cells 0..5 reversed
    ux[:, :, :h] = -UH
    wall = np.zeros((nx, ny, nz), bool)
    area = stats.recirculation_area(ux, wall, H, X0, UH, ny)
    i0, i1 = X0 // 2, min(nx // 2, (X0 + 10 * H) // 2)
    k1 = min(nz // 2, H)
    expect = (i1 - i0) * (ny // 2) * min(h // 2, k1) * 8 / (H ** 2 * ny)
    assert area == pytest.approx(expect)
    assert stats.recirculation_area(np.full_like(ux, UH), wall, H, X0, UH, ny) == 0.0
    # solids count as forward flow
    wall[:, :, :h] = True
    assert stats.recirculation_area(ux, wall, H, X0, UH, ny) == 0.0


def test_published_cases_fixture():
    """The validation driver reads the reference's published statistics for its three
    H = 40 cases from the golden fixture."""
    import os
    from stonewall3d.validate import published_cases
    path = os.path.join(os.path.dirname(__file__), "golden", "reference_results.json")
    refs, info = published_cases(path)
    assert set(refs) == {"solid", "p20", "p40"} and int(info["H"]) == 40
    assert refs["p20"]["porosity_target"] == 0.2
    assert abs(refs["solid"]["Cd"] - 1.4029) < 1e-4 and abs(refs["p40"]["area"] - 0.5776) < 1e-3
