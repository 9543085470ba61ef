"""Force statistics and the recirculation area, as the reference's post-processing
defines them.

force_stats        stonewall3d/postprocess.py:65-82: over the averaging window t >= n_spin,
                   Cd mean, CdMax (the "gust-peak" Cd), CdStd (force fluctuation), Cm mean,
                   CmMax, Cl mean.
blur3              stonewall3d/export3d.py:23-35: separable 3-point box blur, edge-padded.
recirculation_area stonewall3d/export3d.py:169-184: the time-mean u_x downsampled by 2
                   (as lbm3d.py:412 stores it, float16), cropped to x in [X0, X0+10H) and
                   z < 2H, solids set to 0.05, normalised by U_H, the ground row copied from
                   row 1, blurred once; area = (cells < 0) * 8 / (H^2 * ny) in H^2 per unit
                   wall length.
"""
import numpy as np


def coefficients(force, UH, H, ny, zeta=None):
    """Cd, Cl, Cm per step from force rows (t, nz, 2) (lbm3d.py:397-407)."""
    force = np.asarray(force, np.float64)
    nz = force.shape[1]
    if zeta is None:
        zeta = np.arange(nz) - 0.5
    qref = 0.5 * UH ** 2
    A = H * ny
    Fx, Fz = force[:, :, 0].sum(1), force[:, :, 1].sum(1)
    M = (force[:, :, 0] * zeta[None, :]).sum(1)
    return Fx / (qref * A), Fz / (qref * A), M / (qref * A * H)


def force_stats(Cd, Cm, Cl, n_spin):
    cd, cm, cl = np.asarray(Cd)[n_spin:], np.asarray(Cm)[n_spin:], np.asarray(Cl)[n_spin:]
    return dict(Cd=float(cd.mean()), CdMax=float(cd.max()), CdStd=float(cd.std()), Cm=float(cm.mean()),
                CmMax=float(cm.max()), Cl=float(cl.mean()))


def blur3(a, n=1):
    a = a.astype(np.float32)
    for _ in range(n):
        for ax in range(3):
            pad = [(1, 1) if k == ax else (0, 0) for k in range(3)]
            p = np.pad(a, pad, mode="edge")
            lo = [slice(None)] * 3
            mid = [slice(None)] * 3
            hi = [slice(None)] * 3
            lo[ax], mid[ax], hi[ax] = slice(0, -2), slice(1, -1), slice(2, None)
            a = (p[tuple(lo)] + p[tuple(mid)] + p[tuple(hi)]) / 3.0
    return a


def recirculation_area(mean_ux, wall, H, X0, UH, ny):
    """mean_ux, wall: full-resolution (nx, ny, nz) time-mean u_x and solid mask."""
    return recirculation_area_ds2(np.asarray(mean_ux)[::2, ::2, ::2].astype(np.float16), np.asarray(wall)[::2, ::2, ::2],
                                  H, X0, UH, ny)


def recirculation_area_ds2(ux_ds2, wall_ds2, H, X0, UH, ny):
    """The same from the stored half-resolution fields (the .npz keys mean3d_ds2[0] and
    wall_ds2, as export3d.py reads them)."""
    ux = np.asarray(ux_ds2).astype(np.float32)
    solid = np.asarray(wall_ds2)
    i0, i1 = X0 // 2, min(ux.shape[0], (X0 + 10 * H) // 2)
    k1 = min(ux.shape[2], int(2 * H) // 2)
    sub = np.where(solid[i0:i1, :, :k1], 0.05, ux[i0:i1, :, :k1]) / UH
    sub[:, :, 0] = sub[:, :, 1]
    sub = np.pad(blur3(sub, 1), 1, constant_values=1.0)
    return float((sub[1:-1, 1:-1, 1:-1] < 0).sum() * 8) / (H ** 2 * ny)
