"""Benchmark of the fused D3Q19 step on one MI355X.

Workload (default): the H=40 porous stone wall at 20 % void fraction, 640 x 320 x 160
cells (32.8 M), gusting inlet, nu = 2e-4, Cs = 0.16, free-slip top and sides, FP32
populations.  A step is one update of the whole lattice by the fused step; FP32 runs two
steps per launch (k_step2: one HBM pass for two updates, bit-identical to two k_step
launches), the 16-bit forms one step per launch (k_step); every 10 steps the FP64
force-row reduction runs inside the timed region.  Inputs are resident in HBM before
timing starts.

Prints ONE JSON line:
  value / ms_per_step   MLUPS over all nx*ny*nz cells (as lbm3d.py:384 counts them) from
                        the HIP-event time of the whole timed region
  roofline              achieved = algorithmic bytes (152 B per FP32 cell update: 19 loads
                        + 19 stores of 4 B; 76 B for the 16-bit forms) / ms_per_step, against
                        8 TB/s; kernel_avg_ms is the main kernel's mean launch duration
                        (events around each launch, steps_per_launch updates each); traffic
                        is the HBM bytes per launch from the rocprofv3 PMC summary in
                        profiles/ (FETCH_SIZE and WRITE_SIZE, scaled by the same counters on a
                        copy of known size)
  fp32_one_step         the one-step-per-launch FP32 kernel timed the same way
  copy_ceiling          the same process's streaming copy of one population array (the
                        read + write pattern of a step without arithmetic), and the step's
                        fraction of it
  fp16_storage          the FP16-storage form timed the same way on the same lattice, with
                        its ratio to the FP32 MLUPS (headline and one-step)
  cd_err_vs_fp64        mean Cd of FP32, FP16 and BF16 storage against the FP64 run on the
                        40 %-void gusting wall at H = 20 (320 x 160 x 80, 12,000 steps each)
  serving               p50 / p99 latency and aggregate rate of a short served trace (8
                        requests, FP32 + FP16, step counts scaled by 0.05)
  cpu_baseline          the FP32 oracle (a C restatement of the reference kernel) on the
                        host: a bounded sample of the same lattice on all cores (value) and
                        on one thread, plus the 160 x 80 x 40 box-wall case
  box_wall_160x80x40_gpu  that box-wall case's 2000 steps on the GPU (a lattice of 0.5 M
                        cells: bound by launch latency and its few tiles, not by HBM)
The comparison legs (copy, one-step FP32, FP16) run before the headline so that it is timed
at the GPU's settled clocks; the Cd and serving legs after it; the CPU baseline runs last.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "jeju-stonewall-lbm_amd"))
sys.path.insert(0, ROOT)

METRIC = "MLUPS and % HBM roofline (FP32 vs FP16-storage, Cd err vs FP64); p50/p99 case latency"
HBM_PEAK_GBS = 8000.0


def traffic_from_profile(dtype, kernel="k_step"):
    """HBM bytes per launch of `kernel` and the summary they come from (rocprofv3 PMC
    counters FETCH_SIZE / WRITE_SIZE on the 640x320x160 lattice), if the summary exists."""
    name = f"{kernel}_porous_wall_640x320x160_{dtype}_traffic.json"
    p = os.path.join(ROOT, "profiles", name)
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f).get("bytes_per_launch"), f"profiles/{name} (rocprofv3 PMC)"
    return None, None


def _oracle_rate(co, w, threads, budget_s, max_steps):
    """Steps of the FP32 oracle on the host until budget_s or max_steps: (MLUPS, steps, s)."""
    import numpy as np
    co.set_threads(threads)
    f = co.init(w.flag, w.uin, "f32")
    g = np.empty_like(f)
    yb = 1 if w.ybc == "freeslip" else 0
    steps, t0 = 0, time.perf_counter()
    while steps < max_steps:
        gust = 1.0 if w.gust is None else float(w.gust[steps % len(w.gust)])
        co.step(f, w.flag, w.uin, w.nu0, w.cs, w.x_sponge, w.nu_sponge, gust=gust, ybc=yb, f_out=g)
        f, g = g, f
        steps += 1
        if time.perf_counter() - t0 > budget_s:
            break
    el = time.perf_counter() - t0
    nx, ny, nz = w.flag.shape
    return steps * nx * ny * nz / el / 1e6, steps, el


def _cpu_entry(co, w, threads, budget_s, max_steps, what):
    v, steps, el = _oracle_rate(co, w, threads, budget_s, max_steps)
    nx, ny, nz = w.flag.shape
    return dict(value=round(v, 2), unit="MLUPS", gbs=round(v * 1e6 * 152 / 1e9, 2), cores=threads,
                sample=f"{steps} steps of {what} ({nx}x{ny}x{nz}), FP32 C oracle, {threads} thread(s), {el:.1f} s")


def cpu_baseline(w, budget_s=10.0):
    """The FP32 oracle (a C restatement of the reference kernel) on the host cores, on
    bounded samples: the benchmarked lattice on all cores (at most 16 OpenMP threads; the
    headline value) and on one thread, and the CPU-sized box-wall case (160x80x40: up to its
    full 2000 steps on all cores, a sample on one thread).  GB/s counts the same 152
    algorithmic bytes per update as the GPU figure."""
    from oracle import c_oracle as co
    from stonewall3d import workloads
    threads = min(16, os.cpu_count() or 1)
    head = _cpu_entry(co, w, threads, budget_s, 40, "the benchmarked lattice")
    one = _cpu_entry(co, w, 1, budget_s / 2, 2, "the benchmarked lattice")
    c0 = workloads.box_wall_flow(n_steps=2000)
    c0_all = _cpu_entry(co, c0, threads, budget_s, 2000, "the box-wall case")
    c0_one = _cpu_entry(co, c0, 1, budget_s / 2, 2000, "the box-wall case")
    return dict(head, kind="port", single_thread=one,
                box_wall_160x80x40=dict(all_cores=c0_all, single_thread=c0_one))


def two_step_ok(w, dtype):
    """The two-steps-per-launch kernel takes FP32 / FP16 / BF16 arrays below 4 GiB."""
    return dtype != "f64" and 19 * w.cells * (4 if dtype == "f32" else 2) < (1 << 32)


def timed_steps(w, dtype, shift, warmup, steps, ring=10, two_step=False):
    """Build the lattice in HBM, take `warmup` untimed steps, then time exactly `steps` steps:
    HIP events around the whole run and around every launch, on the launch stream (the
    force-row reduction runs every `ring` steps inside the timed region).  With two_step,
    each launch advances two steps (an odd count ends with one single step).
    Returns (simulation, total ms, mean ms per launch of the main kernel, steps per launch)."""
    import numpy as np
    import torch
    from stonewall3d.lbm3d import Simulation
    sim = Simulation(w.flag, w.uin, dtype=dtype, shift=shift, ring=ring, two_step=two_step, **w.sim_kwargs())
    sim.init(noise=0.02, seed=0)
    sim.step(warmup)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    per = 2 if two_step else 1
    chunks = [per] * (steps // per) + ([steps % per] if steps % per else [])
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in chunks]
    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_start.record(stream)
    for (a, b), n in zip(ev, chunks):
        a.record(stream)
        sim.step(n)
        b.record(stream)
    t_end.record(stream)
    torch.cuda.synchronize()
    main = [a.elapsed_time(b) for (a, b), n in zip(ev, chunks) if n == per] or [a.elapsed_time(b) for a, b in ev]
    return sim, t_start.elapsed_time(t_end), float(np.mean(main)), per


def box_wall_gpu(steps=2000, warmup=20):
    """The CPU baseline's box-wall case (160 x 80 x 40, H = 10, FP32) for `steps` steps on the
    GPU, as the CLI runs a lattice: two steps per launch, the force rows reduced every 64
    steps; HIP events around the whole run.  Returns (MLUPS, ms per step)."""
    import numpy as np
    import torch
    from stonewall3d import workloads
    from stonewall3d.lbm3d import Simulation
    w = workloads.box_wall_flow(n_steps=steps + warmup)
    sim = Simulation(w.flag, w.uin, dtype="f32", ring=64, two_step=True, **w.sim_kwargs())
    sim.init()
    sim.step(warmup)
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(s)
    sim.step(steps)
    b.record(s)
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / steps
    if not np.isfinite(sim.force_rows()).all():
        raise SystemExit("box-wall force rows are not finite")
    del sim
    return w.cells / (ms * 1e-3) / 1e6, ms


def cd_error_leg(H=20, porosity=0.4, seed=0, spin=10.0, avg=20.0):
    """The metric's Cd error against FP64, on a sample of the precision study's case: the
    40 %-void gusting wall at H = 20 (320 x 160 x 80), the same seed in every storage form,
    spin-up and averaging window in H/U as at H = 40 (12,000 steps, one per launch, the
    time mean sampled every 5 steps).  Mean Cd of FP32, FP16 and BF16 (inlet offsets)
    relative to the FP64 run's."""
    from stonewall3d import precision
    t0 = time.perf_counter()
    ref = precision.run_case(H, porosity, seed, "f64", spin, avg)
    out = {"what": f"mean Cd over the averaging window, porous wall H={H} (void {porosity}, seed {seed}, "
                   f"{spin:g}+{avg:g} H/U = {ref['steps']} steps) against the FP64 run",
           "cells": ref["cells"], "fp64_mean_cd": round(ref["Cd"], 5)}
    for d in ("f32", "f16", "bf16"):
        r = precision.run_case(H, porosity, seed, d, spin, avg, shift="inlet" if d != "f32" else None)
        if not r["finite"]:
            raise SystemExit(f"{d} Cd series is not finite")
        out[d] = {"mean_cd": round(r["Cd"], 5), "rel_err": round((r["Cd"] - ref["Cd"]) / ref["Cd"], 5)}
    out["seconds"] = round(time.perf_counter() - t0, 1)
    return out


def serving_leg(n=8, rate=8.0, steps_scale=0.05):
    """The metric's p50 / p99 case latency on a short served trace: n requests of the serving
    workload (H ~ U{10..40}, porosity U[0, 0.45], FP32 and FP16, Poisson arrivals at `rate`
    per second) with their U[2000, 20000] step counts scaled by `steps_scale`, through the
    continuous-batching server with the library's default tiles (no autotuning)."""
    import numpy as np
    from stonewall3d import serving
    reqs = serving.trace(n, rate, seed=0, steps_scale=steps_scale)
    srv = serving.Server(budget_bytes=250e9, K=32, policy="sjf", fill_cells=64e6)
    try:
        results, rep = srv.serve(reqs)
    finally:
        srv.close()
    if not all(np.isfinite(r.cd_mean) for r in results):
        raise SystemExit("served Cd is not finite")
    return {"requests": rep["requests"], "p50_latency_s": round(rep["p50_latency_s"], 3),
            "p99_latency_s": round(rep["p99_latency_s"], 3), "aggregate_mlups": round(rep["aggregate_mlups"], 1),
            "utilization": round(rep["utilization"], 4), "wall_s": round(rep["wall_s"], 2),
            "what": f"{n} requests at {rate:g}/s, steps U[2000, 20000] x {steps_scale:g}, H U{{10..40}}, "
                    "FP32 + FP16, SJF admission, 32-step chunks"}


def copy_ceiling(nbytes, reps=10):
    """(GB/s of bytes read + bytes written, ms per copy) of the library's streaming copy of
    nbytes, timed with HIP events over `reps` launches after two untimed ones."""
    import torch
    from stonewall3d import _native as nat
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    s = torch.cuda.current_stream()
    for _ in range(2):
        nat.call("sw_stream_copy", src.data_ptr(), dst.data_ptr(), nbytes, s.cuda_stream)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(s)
    for _ in range(reps):
        nat.call("sw_stream_copy", src.data_ptr(), dst.data_ptr(), nbytes, s.cuda_stream)
    b.record(s)
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / reps
    del src, dst
    torch.cuda.empty_cache()
    return 2 * nbytes / (ms * 1e-3) / 1e9, ms


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--dtype", default="f32", choices=["f32", "f16", "bf16", "f64"])
    ap.add_argument("--shift", default="inlet", choices=["inlet", "weights"],
                    help="16-bit forms: offsets from the inlet equilibrium per height (inlet) or the weights")
    ap.add_argument("--workload", default="porous_wall", choices=["porous_wall", "reference"],
                    help="porous_wall: 16H x 8H x 4H rounded-box stone wall with a gusting inlet; reference: "
                         "lbm3d.py's own default case (20H x 4H x (8H+1), its stone wall, seed 7, periodic y)")
    ap.add_argument("--H", type=int, default=40)
    ap.add_argument("--porosity", type=float, default=0.2)
    ap.add_argument("--no-cpu-baseline", action="store_true",
                    help="time the steps of --dtype only: no CPU baseline, no copy ceiling and no FP16-storage "
                         "comparison leg (the form the rocprofv3 runs use)")
    ap.add_argument("--cpu-budget", type=float, default=10.0)
    ap.add_argument("--one-step", action="store_true",
                    help="time the one-step-per-launch kernel (k_step) instead of two steps per launch (k_step2)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only the streaming copy of one FP32 population array (19 x 640 x 320 x 160 x 4 B), 12 "
                         "times: the known-size copy against which rocprofv3's FETCH_SIZE / WRITE_SIZE are scaled")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the stone-wall step runs on a single MI355X here "
              "(multi-GPU x-slab sharding is a design note, not a measured path). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    if a.calibrate_copy:
        n = 19 * 640 * 320 * 160 * 4
        gbs, ms = copy_ceiling(n)
        print(json.dumps({"calibration": "sw_stream_copy", "bytes_read": n, "bytes_written": n, "launches": 12,
                          "gbs": round(gbs, 1), "ms": round(ms, 4)}))
        return
    from stonewall3d import workloads
    from stonewall3d.lbm3d import bytes_per_update

    n_total = a.warmup + a.steps
    if a.workload == "reference":
        w = workloads.reference(H=a.H, porosity=a.porosity, seed=7, n_steps=n_total)
        desc = (f"lbm3d.py default case H={a.H}, porosity {a.porosity}, seed 7, {{}} cells, U0=0.08, Re=15000, "
                f"periodic y, no gust")
    else:
        w = workloads.porous_wall(H=a.H, porosity=a.porosity, seed=0, n_steps=n_total)
        desc = (f"porous stone wall H={a.H}, void {a.porosity}, {{}} cells, gusting inlet, nu=2e-4, Cs=0.16, "
                f"free-slip top/sides")
    nx, ny, nz = w.shape
    N = nx * ny * nz
    # the fastest kernel of each storage form on this lattice: two steps per launch for FP32
    # (0.90 against 0.64 of 8 TB/s), one step per launch for the 16-bit forms (the two-step
    # form of those runs at about two thirds of their one-step rate)
    two = (a.dtype == "f32" and not a.one_step and two_step_ok(w, a.dtype))
    bpu = bytes_per_update(a.dtype)
    headline = a.workload == "porous_wall" and a.H == 40 and a.porosity == 0.2
    legs = {}
    if not a.no_cpu_baseline:
        # The comparison legs run before the headline: the GPU's clocks settle during the
        # first few hundred milliseconds of sustained load, and the headline measures the
        # steady rate that a run of thousands of steps sees.  Each leg frees its lattice.
        pop_bytes = 19 * N * 4
        legs["copy"] = copy_ceiling(pop_bytes)
        if a.dtype == "f32" and two:
            s1, t1, k1, _ = timed_steps(w, "f32", a.shift, a.warmup, a.steps)
            legs["one"] = (t1, k1)
            del s1
            torch.cuda.empty_cache()
        if a.dtype == "f32":
            s16, t16, k16, _ = timed_steps(w, "f16", a.shift, a.warmup, a.steps)
            if not np.isfinite(s16.force_rows()).all():
                raise SystemExit("FP16 force rows are not finite")
            legs["f16"] = (t16, k16, s16.shift_kind)
            del s16
            torch.cuda.empty_cache()
        # The small box-wall case runs the two-step kernel for ~50 ms right before the
        # headline: after the HBM-bound legs the first launches of the two-step kernel run
        # up to 8 % slower than its settled rate (launch by launch in the kernel trace), and
        # the same kernel's load just before the headline shortens that ramp.
        legs["box_wall"] = box_wall_gpu()

    sim, total_ms, kern_ms, per = timed_steps(w, a.dtype, a.shift, a.warmup, a.steps, two_step=two)
    F = sim.force_rows()
    if not np.isfinite(F).all():
        raise SystemExit("force rows are not finite")
    fluid = int((w.flag == 0).sum())
    ms_step = total_ms / a.steps
    mlups = N / (ms_step * 1e-3) / 1e6
    achieved = N * bpu / (ms_step * 1e-3) / 1e9
    traffic, traffic_src = traffic_from_profile(a.dtype, "k_step2" if per == 2 else "k_step") if headline \
        else (None, None)

    line = {
        "metric": METRIC,
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": a.dtype,
        "data": "synthetic (seeded rounded-box stone wall, seeded gust)" if a.workload == "porous_wall"
                else "synthetic (the reference's stone-wall generator, seed 7)",
        "config": {"workload": desc.format(f"{nx}x{ny}x{nz}"),
                   "cells": N, "fluid_cells": fluid, "bytes_per_update": bpu,
                   "storage_offset": sim.shift_kind if a.dtype in ("f16", "bf16") else None},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": "k_step2" if per == 2 else "k_step", "steps_per_launch": per,
                     "kernel_avg_ms": round(kern_ms, 4),
                     "kernel_frac": round(per * N * bpu / (kern_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "fluid_mlups": round(fluid / (ms_step * 1e-3) / 1e6, 1),
    }
    del sim
    torch.cuda.empty_cache()
    if "copy" in legs:
        gbs, cms = legs["copy"]
        line["copy_ceiling"] = {"gbs": round(gbs, 1), "frac_of_peak": round(gbs / HBM_PEAK_GBS, 4),
                                "step_frac_of_copy": round(achieved / gbs, 4), "ms": round(cms, 4),
                                "what": f"sw_stream_copy of one FP32 population array ({pop_bytes / 1e9:.2f} GB), "
                                        "bytes read + written per second"}
    fp32_one = None
    if "one" in legs:
        # the one-step-per-launch FP32 kernel on the same lattice, for comparison
        t1, k1 = legs["one"]
        fp32_one = N / (t1 / a.steps * 1e-3) / 1e6
        line["fp32_one_step"] = {"value": round(fp32_one, 1), "unit": "MLUPS", "kernel": "k_step",
                                 "ms_per_step": round(t1 / a.steps, 4), "kernel_avg_ms": round(k1, 4),
                                 "roofline_frac": round(N * bpu / (t1 / a.steps * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    if "f16" in legs:
        # the metric's FP16-storage side, on the same lattice and step count
        t16, k16, kind16 = legs["f16"]
        ms16 = t16 / a.steps
        m16 = N / (ms16 * 1e-3) / 1e6
        a16 = N * bytes_per_update("f16") / (ms16 * 1e-3) / 1e9
        tr16, src16 = traffic_from_profile("f16") if headline else (None, None)
        line["fp16_storage"] = {"value": round(m16, 1), "unit": "MLUPS", "ratio_vs_fp32": round(m16 / mlups, 3),
                                "ratio_vs_fp32_one_step": round(m16 / fp32_one, 3) if fp32_one else None,
                                "kernel": "k_step",
                                "ms_per_step": round(ms16, 4), "kernel_avg_ms": round(k16, 4),
                                "bytes_per_update": bytes_per_update("f16"), "achieved_gbs": round(a16, 1),
                                "roofline_frac": round(a16 / HBM_PEAK_GBS, 4),
                                "traffic": tr16, "traffic_source": src16, "storage_offset": kind16}
    if not a.no_cpu_baseline:
        line["cd_err_vs_fp64"] = cd_error_leg()
        line["serving"] = serving_leg()
        line["cpu_baseline"] = cpu_baseline(w, a.cpu_budget)
        # the same box-wall case as the CPU baseline's second sample, on the GPU
        gm, gms = legs["box_wall"]
        cpu_bw = line["cpu_baseline"]["box_wall_160x80x40"]["all_cores"]["value"]
        line["box_wall_160x80x40_gpu"] = {"value": round(gm, 1), "unit": "MLUPS", "steps": 2000,
                                          "ms_per_step": round(gms, 5), "kernel": "k_step2",
                                          "vs_cpu_all_cores": round(gm / cpu_bw, 1),
                                          "what": "160x80x40 box wall H=10, FP32, 2000 steps, two per launch"}
    print(json.dumps(line), flush=True)

if __name__ == "__main__":
    main()
