"""Concrete workloads: the reference's own run set-up and the explicit wall cases.

  reference(H, porosity, seed)  lbm3d.py's default case (grid 20H x 4H x (8H+1), wall at
                                x = 4H, thickness 0.3H, periodic y, U0 = 0.08, Re = 15000,
                                sponge over the last 3H) with the reference's stone wall
  box_wall_flow()               160 x 80 x 40, box wall H = 10, thickness 6 at x = 40,
                                inlet 0.05 (z/H)^(1/7) capped at 0.06, nu = 2e-4, Cs = 0.16
  porous_wall(H, porosity, ...) 16H x 8H x 4H, wall at x = 4H, thickness 0.6H, rounded-box
                                stones opened to a void fraction, gusting inlet, nu = 2e-4
                                (gust_time_scale s evaluates the gust at s*t: a lattice at
                                half the resolution uses s = 2 to keep physical time)
                                (H = 40, 20 %: the 640 x 320 x 160 headline lattice)
  batch_instances / serving_requests  the uneven-lattice mixes of the batched and serving forms
Boundaries of the explicit cases: no-slip floor, free-slip top and sides, zero-gradient
outlet; a sponge over the last 3H as in the reference (lbm3d.py:282) unless disabled.
"""
from dataclasses import dataclass, field

import numpy as np

from . import geometry
from .inflow import power_law_profile, gust_table, make_gust, reference_setup
from .lattice import flag_from_mask


@dataclass
class Workload:
    name: str
    flag: np.ndarray
    uin: np.ndarray
    H: int
    UH: float
    nu0: float
    cs: float = 0.16
    x_sponge: int = None
    nu_sponge: float = 0.05
    ybc: str = "freeslip"
    gust: np.ndarray = None
    n_steps: int = 1000
    n_spin: int = 0
    info: dict = field(default_factory=dict)

    @property
    def shape(self):
        return self.flag.shape

    @property
    def cells(self):
        return int(np.prod(self.flag.shape))

    def sim_kwargs(self):
        return dict(nu0=self.nu0, cs=self.cs, x_sponge=self.x_sponge, nu_sponge=self.nu_sponge, ybc=self.ybc,
                    gust=self.gust, n_steps=self.n_steps)


def reference(H=40, porosity=0.2, seed=7, Re=15000.0, U0=0.08, Cs=0.16, n_steps=None, verbose=False):
    s = reference_setup(H=H, Re=Re, U0=U0)
    wall, phi = geometry.build_wall(porosity, H, s.nx, s.ny, s.nz, s.X0, s.T, seed=seed, verbose=verbose)
    return Workload(f"reference H={H} p={porosity} seed={seed}", flag_from_mask(wall), s.uin, H, s.UH, s.nu0, Cs,
                    s.x_sponge, 0.05, "periodic", None, s.n_total if n_steps is None else n_steps, s.n_spin,
                    dict(phi=phi, X0=s.X0, T=s.T, t_unit=s.t_unit, zeta=s.zeta))


def box_wall_flow(n_steps=2000, ybc="freeslip", sponge=True, nx=160, ny=80, nz=40, H=10, X0=40, T=6, nu0=2e-4):
    wall = geometry.box_wall(nx, ny, nz, X0, T, H)
    uin = power_law_profile(nz, H)
    return Workload(f"box wall H={H} {nx}x{ny}x{nz}", flag_from_mask(wall), uin, H, 0.05, nu0, 0.16,
                    nx - 3 * H if sponge else None, 0.05, ybc, None, n_steps, 0, dict(X0=X0, T=T))


def porous_wall(H=40, porosity=0.2, seed=0, gust_seed=None, n_steps=20000, nu0=2e-4, sponge=True, ybc="freeslip",
                dims=(16, 8, 4), gust_time_scale=1.0):
    nx, ny, nz = dims[0] * H, dims[1] * H, dims[2] * H
    X0, T = 4 * H, max(2, int(round(0.6 * H)))
    wall, info = geometry.rounded_box_wall(nx, ny, nz, X0, T, H, porosity, seed)
    uin = power_law_profile(nz, H)
    g = make_gust(seed if gust_seed is None else gust_seed)
    info.update(X0=X0, T=T, gust=dict(a=g.a.tolist(), f=g.f.tolist(), phi=g.phi.tolist(), amp=g.amp))
    gt = gust_table(g, n_steps) if gust_time_scale == 1.0 else \
        g.factor(np.arange(n_steps) * float(gust_time_scale)).astype(np.float32)
    return Workload(f"porous wall H={H} void={porosity} {nx}x{ny}x{nz}", flag_from_mask(wall), uin, H, 0.05, nu0,
                    0.16, nx - 3 * H if sponge else None, 0.05, ybc, gt, n_steps, 0, info)


def batch_instances(n=64, seed=0, heights=(8, 12, 16, 20, 24), n_steps=5000):
    """(H, porosity, seed) of the batched form: H from `heights`, porosity U[0, 0.45]."""
    rng = np.random.default_rng(seed)
    return [(int(rng.choice(heights)), float(rng.uniform(0.0, 0.45)), int(rng.integers(1 << 30)), n_steps)
            for _ in range(n)]


def serving_requests(n, rate, seed=0, h_range=(10, 40), steps=(2000, 20000), dtypes=("f32", "f16")):
    """A finite Poisson trace: (arrival time s, H, porosity, gust seed, steps, dtype)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    return [dict(arrival=float(t[k]), H=int(rng.integers(h_range[0], h_range[1] + 1)),
                 porosity=float(rng.uniform(0.0, 0.45)), seed=int(rng.integers(1 << 30)),
                 steps=int(rng.integers(steps[0], steps[1] + 1)), dtype=str(rng.choice(dtypes))) for k in range(n)]
