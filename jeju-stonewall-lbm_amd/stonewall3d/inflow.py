"""Inlet profiles, the gust factor and the derived run constants.

reference_setup() restates the grid / inlet / viscosity / step-count set-up of
stonewall3d/lbm3d.py:270-282; power_law_profile() and gust_table() are the inflow of the
explicit workloads (power law 0.05 (z/H)^(1/7) capped at 0.06, and a gust factor
1 + 0.25 sum_k a_k sin(2 pi f_k t + phi_k)), which the reference does not have.
"""
from dataclasses import dataclass

import numpy as np


@dataclass
class RefSetup:
    nx: int
    ny: int
    nz: int
    X0: int
    T: int
    zeta: np.ndarray
    uin: np.ndarray
    UH: float
    nu0: float
    t_unit: float
    n_spin: int
    n_avg: int
    n_total: int
    snap_every: int
    x_sponge: int


def reference_setup(H=40, Re=15000.0, U0=0.08, up=4.0, down=16.0, span=4.0, height=8.0, thick=0.3,
                    t_spinup=20.0, t_avg=30.0, snap_dt=0.25):
    """Grid, inlet and counts exactly as lbm3d.py:270-282 computes them."""
    nx, ny, nz = int(round((up + down) * H)), int(round(span * H)), int(round(height * H)) + 1
    X0, T = int(round(up * H)), max(2, int(round(thick * H)))
    zeta = np.arange(nz) - 0.5
    delta = 3.0 * H
    uin = (U0 * np.clip(zeta / delta, 0, 1) ** (1 / 7)).astype(np.float32)
    uin[0] = 0.0
    UH = float(U0 * (H / delta) ** (1 / 7))
    nu0 = UH * H / Re
    t_unit = H / UH
    n_spin, n_avg = int(t_spinup * t_unit), int(t_avg * t_unit)
    snap_every = max(1, int(round(snap_dt * t_unit)))
    return RefSetup(nx, ny, nz, X0, T, zeta, uin, UH, nu0, t_unit, n_spin, n_avg, n_spin + n_avg, snap_every,
                    int(nx - 3 * H))


def power_law_profile(nz, H, u_ref=0.05, cap=0.06):
    """u(zeta) = u_ref (zeta/H)^(1/7), capped at `cap`; zeta = k - 1/2; the ground cell gets 0."""
    zeta = np.arange(nz) - 0.5
    u = u_ref * np.clip(zeta / H, 0, None) ** (1 / 7)
    u = np.minimum(u, cap).astype(np.float32)
    u[0] = 0.0
    return u


@dataclass
class Gust:
    amp: float
    a: np.ndarray       # (K,) mode weights, sum 1
    f: np.ndarray       # (K,) frequencies per step
    phi: np.ndarray     # (K,) phases

    def factor(self, t):
        t = np.asarray(t, dtype=np.float64)
        s = np.zeros_like(t)
        for ak, fk, pk in zip(self.a, self.f, self.phi):
            s = s + ak * np.sin(2.0 * np.pi * fk * t + pk)
        return 1.0 + self.amp * s


def make_gust(seed, K=8, amp=0.25, fmin=1e-4, fmax=2e-3):
    """Gust modes from one seed: a_k ~ U[0,1) normalised to sum 1 (so the factor stays in
    [1-amp, 1+amp]), f_k ~ U[fmin, fmax) per step, phi_k ~ U[0, 2 pi)."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.0, 1.0, K)
    a = a / a.sum()
    f = rng.uniform(fmin, fmax, K)
    phi = rng.uniform(0.0, 2.0 * np.pi, K)
    return Gust(amp, a, f, phi)


def gust_table(gust, n_steps):
    """float32 factor per step t = 0..n_steps-1 (evaluated in FP64, rounded once)."""
    if gust is None:
        return None
    return gust.factor(np.arange(n_steps)).astype(np.float32)
