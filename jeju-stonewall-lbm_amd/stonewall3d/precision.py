"""Precision study: the storage forms against the FP64 run on the same seeded wall.

For one porous wall (default H = 40, 40 % void, 640 x 320 x 160, gusting inlet) the same
case runs with FP64, FP32, FP16 and BF16 population storage (collision arithmetic FP32
for the 32/16-bit forms, FP64 for FP64; force sums FP64 in every form); each 16-bit form
runs with each storage offset (--shifts: the inlet equilibrium per height, the weights).  Each run spins
up for `spin` H/U, then averages for `avg` H/U: force statistics as postprocess.py:65-82
(mean Cd, gust-peak Cd = max, Cd std = force RMS, Cm, Cl) and the recirculation area of
export3d.py:169-184 from the time-mean field (sampled every 5 steps as lbm3d.py:365-370).
Errors are reported against the FP64 run at the same resolution, and against an FP64 run
at half the resolution (H/2, 8x fewer cells, steps halved, gust time-scaled by 2).
"""
import argparse
import json
import sys
import time

import numpy as np
import torch

from . import stats, workloads
from .lbm3d import Simulation


def run_case(H, porosity, seed, dtype, spin, avg, acc_every=5, log=None, noise_seed=None, cs=None, shift=None):
    UH = 0.05
    t_unit = H / UH
    n_spin, n_avg = int(spin * t_unit), int(avg * t_unit)
    n = n_spin + n_avg
    w = workloads.porous_wall(H=H, porosity=porosity, seed=seed, n_steps=n, gust_time_scale=40.0 / H)
    if cs is not None:
        w.cs = float(cs)
    sim = Simulation(w.flag, w.uin, dtype=dtype, shift=shift, **w.sim_kwargs())
    sim.init(noise=0.02, seed=seed if noise_seed is None else noise_seed)
    t0 = time.perf_counter()
    sim.step(n_spin)
    k = 0
    while sim.t < n:
        if k % acc_every == 0:
            sim.accumulate()
        sim.step(1)
        k += 1
        if log and sim.t % 4000 == 0:
            torch.cuda.synchronize()
            log(f"  [{dtype} H={H}] step {sim.t}/{n} {time.perf_counter() - t0:.1f}s")
    torch.cuda.synchronize()
    wall_s = time.perf_counter() - t0
    Cd, Cl, Cm = stats.coefficients(sim.force_rows(), UH, H, w.shape[1])
    st = stats.force_stats(Cd, Cm, Cl, n_spin)
    mean = sim.mean_field()
    area = stats.recirculation_area(mean[1], w.flag == 1, H, w.info["X0"], UH, w.shape[1])
    out = dict(dtype=dtype, shift=sim.shift_kind if dtype in ("f16", "bf16") else None, H=H, porosity=porosity,
               seed=seed, noise_seed=noise_seed, cs=w.cs, cells=w.cells, steps=n, n_spin=n_spin, recirculation_area=area, finite=bool(np.isfinite(Cd).all()), seconds=wall_s, **st)
    del sim
    torch.cuda.empty_cache()
    return out


def compare(res, ref):
    keys = ("Cd", "CdMax", "CdStd", "Cm", "CmMax", "recirculation_area")
    return {k: dict(value=res[k], ref=ref[k], abs_err=res[k] - ref[k],
                    rel_err=(res[k] - ref[k]) / ref[k] if ref[k] else None) for k in keys}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--H", type=int, default=40)
    ap.add_argument("--porosity", type=float, default=0.4)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--spin", type=float, default=10.0, help="spin-up [H/U]")
    ap.add_argument("--avg", type=float, default=20.0, help="averaging window [H/U]")
    ap.add_argument("--dtypes", default="f64,f32,f16,bf16")
    ap.add_argument("--shifts", default="inlet,weights",
                    help="storage offsets to run each 16-bit dtype with (inlet equilibrium, weights)")
    ap.add_argument("--half-res", action="store_true", help="also run FP64 at H/2")
    ap.add_argument("--cs", type=float, default=None, help="Smagorinsky constant (default 0.16)")
    ap.add_argument("--noise-seeds", default="", help="extra runs that differ only in the init-noise seed, "
                                                     "as dtype:seed[:shift],... (run-to-run spread of the statistics)")
    a = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    runs = {}
    for dt in a.dtypes.split(","):
        for sh in (a.shifts.split(",") if dt in ("f16", "bf16") else [None]):
            key = dt if sh is None else f"{dt}/{sh}"
            runs[key] = run_case(a.H, a.porosity, a.seed, dt, a.spin, a.avg, log=log, cs=a.cs, shift=sh)
            log(json.dumps(runs[key]))
    out = dict(runs=runs)
    if "f64" in runs:
        out["vs_f64"] = {dt: compare(r, runs["f64"]) for dt, r in runs.items() if dt != "f64"}
    if a.noise_seeds:
        out["noise_runs"] = []
        for spec in a.noise_seeds.split(","):
            dt, ns = spec.split(":")[:2]
            sh = spec.split(":")[2] if spec.count(":") > 1 else None
            r = run_case(a.H, a.porosity, a.seed, dt, a.spin, a.avg, log=log, noise_seed=int(ns), cs=a.cs, shift=sh)
            log(json.dumps(r))
            out["noise_runs"].append(r)
    if a.half_res:
        half = run_case(a.H // 2, a.porosity, a.seed, "f64", a.spin, a.avg, log=log, cs=a.cs)
        out["f64_half_res"] = half
        out["vs_f64_half_res"] = {dt: compare(r, half) for dt, r in runs.items()}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
