"""Serving form: a request loop over wall cases of uneven size with continuous batching.

The reference runs fixed cases one process at a time (stonewall3d/run_all.sh).  Here
requests (H, porosity, gust seed, step count, storage dtype) arrive over time; each is
admitted when the memory budget and a batch slot allow, its lattice is prepared (stone
geometry on the host, link masks and equilibrium init on a side stream, overlapping the
running chunk), and it joins the resident batch of its dtype.  Every loop iteration
replays one K-step chunk per dtype batch (a hipGraph), then retires finished lattices
and records their results.  Requests that arrive while others are on the GPU are
admitted as slots and memory free up; their work shares the same persistent launches.

A lattice's results do not depend on its batch-mates: the step arithmetic and the force
reduction are per lattice and in a fixed order (bit-identical alone and in the mix).
"""
import argparse
import json
import sys
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _native as nat
from . import workloads
from .batch import Batch, Lattice


@dataclass
class Request:
    rid: int
    arrival: float          # seconds after the start of the trace
    H: int
    porosity: float
    seed: int
    steps: int
    dtype: str = "f32"


@dataclass
class Result:
    rid: int
    dtype: str
    H: int
    cells: int
    steps: int
    arrival: float
    admitted: float
    finished: float
    cd_mean: float
    cd_max: float
    checksum: float
    force: np.ndarray = field(default=None, repr=False)
    state: object = field(default=None, repr=False)     # final population buffer (device), if kept

    @property
    def latency(self):
        return self.finished - self.arrival


def request_work(req):
    """Cell updates a request asks for: 16H x 8H x 4H cells times its steps."""
    return 512 * req.H ** 3 * req.steps


def nat_name(dtype_code):
    return {v: k for k, v in nat.DTYPES.items()}[dtype_code]


def lattice_bytes(H, dtype, steps, ring):
    """HBM bytes of one request's lattice (populations x2, links, partials, force, inlet)."""
    nx, ny, nz = 16 * H, 8 * H, 4 * H
    N = nx * ny * nz
    e = nat.DTYPE_BYTES[nat.DTYPES[dtype]]
    links = N + ny * nz * ((nx + 2047) // 2048)
    partials = 8 * ring * nat.lib().sw_batch_partials_per_step(nx, ny, nz)   # either step form's layout
    return 2 * 19 * N * e + 4 * links + partials + 16 * steps * nz + 4 * (nz + steps)


class Server:
    """budget_bytes: HBM admitted across all resident lattices; K: steps per chunk;
    policy: "fifo" or "sjf" admission order; fill_cells: stop admitting once the resident
    lattices hold this many cells (None: admit whatever the budget allows)."""

    policy = "fifo"
    fill_cells = None
    keep_state = False

    def __init__(self, budget_bytes=250e9, K=32, max_instances=64, graph=True, keep_force=False, noise=0.02,
                 device="cuda", variants=None, policy="fifo", fill_cells=None, keep_state=False, two_step=None):
        if policy not in ("fifo", "sjf"):
            raise ValueError("policy must be 'fifo' or 'sjf'")
        self.policy = policy
        self.fill_cells = None if fill_cells is None else float(fill_cells)
        self.budget = float(budget_bytes)
        self.K = int(K)
        self.max_instances = int(max_instances)
        self.graph = bool(graph)
        self.keep_force = keep_force
        self.keep_state = keep_state
        self.noise = float(noise)
        self.device = torch.device(device)
        self.variants = dict(variants or {})
        self.two_step = two_step           # None: the Batch default (two steps per launch for FP32)
        self.blocks = {}
        self.batches = {}
        self.streams = {}
        self.used = 0.0
        self.prep_stream = torch.cuda.Stream(device=self.device)

    def _batch(self, dtype):
        """The resident batch of a dtype and its own stream (batches of different dtypes
        run concurrently on the device)."""
        if dtype not in self.batches:
            self.streams[dtype] = torch.cuda.Stream(device=self.device)
            with torch.cuda.stream(self.streams[dtype]):
                self.batches[dtype] = Batch(dtype, self.max_instances, self.K, blocks=self.blocks.get(dtype),
                                            graph=self.graph, device=self.device, variant=self.variants.get(dtype),
                                            two_step=self.two_step)
        return self.batches[dtype]

    def make_lattice(self, req):
        w = workloads.porous_wall(H=req.H, porosity=req.porosity, seed=req.seed, n_steps=req.steps)
        lat = Lattice(w.flag, w.uin, dtype=req.dtype, ring=self.K, device=self.device, tag=req, **w.sim_kwargs())
        lat.init(noise=self.noise, seed=req.seed)
        lat.workload = w
        return lat

    def _finish(self, lat, t0):
        req = lat.tag
        F = lat.force_rows()
        w = lat.workload
        cd = F[:, :, 0].sum(1) / (0.5 * w.UH ** 2 * w.H * lat.ny)
        cs = float(np.sum(F))
        return Result(req.rid, req.dtype, req.H, lat.N, req.steps, req.arrival, lat.admitted - t0,
                      time.perf_counter() - t0, float(cd.mean()), float(cd.max()), cs,
                      F if self.keep_force else None, lat.current().clone() if self.keep_state else None)

    def tune(self, dtypes=("f32", "f16"), heights=(16, 24, 32), tuner=None):
        """Pick each dtype's tile variant and persistent grid with the autotuner, on a
        throwaway batch of representative lattices (tuning advances its lattices, so real
        requests never take part in it).  Returns {dtype: (variant, workgroups)}."""
        from .autotune import Autotuner
        tuner = tuner or Autotuner()
        out = {}
        for d in dtypes:
            b = Batch(d, max_instances=len(heights), K=self.K, graph=self.graph, device=self.device,
                      two_step=self.two_step)
            for h in heights:
                w = workloads.porous_wall(H=h, porosity=0.2, seed=h, n_steps=4 * self.K)
                b.add(Lattice(w.flag, w.uin, dtype=d, ring=self.K, device=self.device, **w.sim_kwargs())
                      .init(noise=self.noise, seed=h))
            cand = [1, 3, 7] + ([8, 9, 10] if d != "f64" else [])
            tuner.tune_batch(b, variants=cand, per_cu=(4, 8, 16, 32), chunks=2)
            self.variants[d], self.blocks[d] = b.variant, b.blocks
            out[d] = (b.variant, b.blocks)
            b.close()
            del b
        torch.cuda.synchronize(self.device)
        return out

    def _admit(self, queue, t0):
        """Admit queued requests while the budget, the batch slots and the compute cap allow.

        Order: "fifo" takes the oldest request and a request that does not fit blocks those
        behind it; "sjf" takes the request with the least work (cells x steps) first.  With
        `fill_cells` set, admission also stops once the resident lattices hold that many
        cells (enough to keep the device busy): later requests wait in the queue rather
        than slow every resident lattice down, which shortens the typical time-to-result
        when the offered load exceeds capacity.  Geometry is built on the host and the
        lattice prepared on the side stream, so this overlaps chunks already in flight;
        joining a batch waits for that batch's chunk to end."""
        while queue:
            if self.fill_cells and self._resident_cells() >= self.fill_cells:
                return
            k = 0
            if self.policy == "sjf":
                k = min(range(len(queue)), key=lambda j: (request_work(queue[j]), queue[j].arrival))
            req = queue[k]
            need = lattice_bytes(req.H, req.dtype, req.steps, self.K)
            b = self._batch(req.dtype)
            if self.used + need > self.budget or len(b.members) >= self.max_instances:
                if self.used == 0 and not b.members:
                    raise RuntimeError(f"request {req.rid} needs {need/1e9:.1f} GB > budget")
                return
            queue.pop(k)
            with torch.cuda.stream(self.prep_stream):
                lat = self.make_lattice(req)
            self.prep_stream.synchronize()
            lat.admitted = time.perf_counter()
            lat.reserved = need
            self.used += need
            with torch.cuda.stream(self.streams[req.dtype]):
                b.add(lat)

    def _resident_cells(self):
        return sum(m.N for b in self.batches.values() for m in b.members)

    def serve(self, requests, realtime=True, log=None):
        """Run a finite trace; returns (results, report).

        Each iteration launches one K-step chunk per non-empty dtype batch (asynchronous,
        one stream per dtype, so the batches overlap on the device), admits the requests
        that have arrived while those chunks run, then retires the lattices that
        finished."""
        pending = sorted(requests, key=lambda r: r.arrival)
        queue, results = [], []
        t0 = time.perf_counter()
        busy = 0.0
        cell_updates = 0
        while pending or queue or any(b.members for b in self.batches.values()):
            now = time.perf_counter() - t0 if realtime else float("inf")
            while pending and pending[0].arrival <= now:
                queue.append(pending.pop(0))
            active = [b for b in self.batches.values() if b.members]
            if not active:
                self._admit(queue, t0)
                if not any(b.members for b in self.batches.values()) and pending:
                    time.sleep(max(0.0, min(0.01, pending[0].arrival - (time.perf_counter() - t0))))
                continue
            ts = time.perf_counter()
            for b in active:
                for m in b.members:
                    cell_updates += m.N * max(0, min(self.K, m.n_steps - m.t))
                with torch.cuda.stream(self.streams[nat_name(b.dtype)]):
                    b.run_chunk()
            self._admit(queue, t0)
            for b in active:
                with torch.cuda.stream(self.streams[nat_name(b.dtype)]):
                    finished = b.retire_finished()
                for lat in finished:
                    results.append(self._finish(lat, t0))
                    self.used -= lat.reserved
                    if log:
                        log(results[-1])
                    del lat
            busy += time.perf_counter() - ts
        wall = time.perf_counter() - t0
        lat_s = np.array([r.latency for r in results])
        report = dict(requests=len(results), wall_s=wall, busy_s=busy, cell_updates=int(cell_updates),
                      aggregate_mlups=cell_updates / wall / 1e6 if wall > 0 else 0.0,
                      p50_latency_s=float(np.percentile(lat_s, 50)) if len(lat_s) else None,
                      p99_latency_s=float(np.percentile(lat_s, 99)) if len(lat_s) else None,
                      utilization=busy / wall if wall > 0 else 0.0)
        return results, report

    def close(self):
        for b in self.batches.values():
            b.close()


def trace(n, rate, seed=0, steps_scale=1.0, h_range=(10, 40), dtypes=("f32", "f16")):
    """Finite Poisson trace of n requests (rate per second), steps scaled by steps_scale."""
    reqs = []
    for k, r in enumerate(workloads.serving_requests(n, rate, seed=seed, h_range=h_range, dtypes=dtypes)):
        reqs.append(Request(k, r["arrival"], r["H"], r["porosity"], r["seed"],
                            max(1, int(round(r["steps"] * steps_scale))), r["dtype"]))
    return reqs


def main(argv=None):
    ap = argparse.ArgumentParser(description="serve a finite Poisson trace of stone-wall cases")
    ap.add_argument("--requests", type=int, default=16)
    ap.add_argument("--rate", type=float, default=2.0)
    ap.add_argument("--steps-scale", type=float, default=0.05, help="scale of the U[2000, 20000] step counts")
    ap.add_argument("--hmin", type=int, default=10)
    ap.add_argument("--hmax", type=int, default=40)
    ap.add_argument("--budget-gb", type=float, default=250.0)
    ap.add_argument("--K", type=int, default=32)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-tune", action="store_true", help="keep the library's default tiles and 8 workgroups/CU")
    ap.add_argument("--policy", default="sjf", choices=["fifo", "sjf"], help="admission order")
    ap.add_argument("--fill-mcells", type=float, default=64.0,
                    help="stop admitting once resident lattices hold this many million cells (0: budget only)")
    a = ap.parse_args(argv)
    reqs = trace(a.requests, a.rate, a.seed, a.steps_scale, (a.hmin, a.hmax))
    srv = Server(budget_bytes=a.budget_gb * 1e9, K=a.K, policy=a.policy,
                 fill_cells=a.fill_mcells * 1e6 if a.fill_mcells > 0 else None)
    tuned = None if a.no_tune else srv.tune(sorted({r.dtype for r in reqs}))
    results, rep = srv.serve(reqs, log=lambda r: print(
        f"[serve] request {r.rid} done: {r.dtype} H={r.H} {r.steps} steps, latency {r.latency:.2f} s, "
        f"Cd {r.cd_mean:.4f}", file=sys.stderr, flush=True))
    work = sum(request_work(r) for r in reqs)
    rep.update(rate=a.rate, steps_scale=a.steps_scale, K=a.K, budget_gb=a.budget_gb, policy=a.policy,
               fill_mcells=a.fill_mcells,
               offered_mlups=work / max(reqs[-1].arrival, 1e-9) / 1e6,
               arrival_span_s=reqs[-1].arrival, tuned=tuned,
               by_dtype={d: dict(requests=sum(r.dtype == d for r in results),
                                 p50_latency_s=float(np.percentile([r.latency for r in results if r.dtype == d], 50)),
                                 p99_latency_s=float(np.percentile([r.latency for r in results if r.dtype == d], 99)))
                         for d in sorted({r.dtype for r in results})},
               all_finite=bool(all(np.isfinite(r.cd_mean) for r in results)))
    print(json.dumps(rep), flush=True)
    srv.close()


if __name__ == "__main__":
    main()
