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

    @property
    def latency(self):
        return self.finished - self.arrival


def nat_name(dtype_code):
    return {v: k for k, v in nat.DTYPES.items()}[dtype_code]


def lattice_bytes(H, dtype, steps, ring):
    """HBM bytes of one request's lattice (populations x2, links, partials, force, inlet)."""
    nx, ny, nz = 16 * H, 8 * H, 4 * H
    N = nx * ny * nz
    e = nat.DTYPE_BYTES[nat.DTYPES[dtype]]
    links = N + ny * nz * ((nx + 2047) // 2048)
    nyb = (ny + 3) // 4
    return 2 * 19 * N * e + 4 * links + 8 * ring * nz * nyb * 2 + 16 * steps * nz + 4 * (nz + steps)


class Server:
    """budget_bytes: HBM admitted across all resident lattices; K: steps per chunk."""

    def __init__(self, budget_bytes=250e9, K=32, max_instances=64, graph=True, keep_force=False, noise=0.02,
                 device="cuda", variants=None):
        self.budget = float(budget_bytes)
        self.K = int(K)
        self.max_instances = int(max_instances)
        self.graph = bool(graph)
        self.keep_force = keep_force
        self.noise = float(noise)
        self.device = torch.device(device)
        self.variants = variants or {}
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
                self.batches[dtype] = Batch(dtype, self.max_instances, self.K, graph=self.graph, device=self.device,
                                            variant=self.variants.get(dtype))
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
                      F if self.keep_force else None)

    def serve(self, requests, realtime=True, log=None):
        """Run a finite trace; returns (results, report)."""
        pending = sorted(requests, key=lambda r: r.arrival)
        queue, results = [], []
        t0 = time.perf_counter()
        busy = 0.0
        cell_updates = 0
        while pending or queue or any(b.members for b in self.batches.values()):
            now = time.perf_counter() - t0
            if not realtime:
                now = float("inf")
            while pending and pending[0].arrival <= now:
                queue.append(pending.pop(0))
            # admission (FIFO; a request that does not fit blocks those behind it)
            while queue:
                req = queue[0]
                need = lattice_bytes(req.H, req.dtype, req.steps, self.K)
                b = self._batch(req.dtype)
                if self.used + need > self.budget or len(b.members) >= self.max_instances:
                    if self.used == 0 and not b.members:
                        raise RuntimeError(f"request {req.rid} needs {need/1e9:.1f} GB > budget")
                    break
                queue.pop(0)
                with torch.cuda.stream(self.prep_stream):
                    lat = self.make_lattice(req)
                self.prep_stream.synchronize()
                lat.admitted = time.perf_counter()
                lat.reserved = need
                self.used += need
                with torch.cuda.stream(self.streams[req.dtype]):
                    b.add(lat)
            active = [b for b in self.batches.values() if b.members]
            if not active:
                if pending:
                    time.sleep(max(0.0, min(0.01, pending[0].arrival - (time.perf_counter() - t0))))
                continue
            ts = time.perf_counter()
            for b in active:
                for m in b.members:
                    cell_updates += m.N * max(0, min(self.K, m.n_steps - m.t))
                with torch.cuda.stream(self.streams[nat_name(b.dtype)]):
                    b.run_chunk()
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
    a = ap.parse_args(argv)
    reqs = trace(a.requests, a.rate, a.seed, a.steps_scale, (a.hmin, a.hmax))
    srv = Server(budget_bytes=a.budget_gb * 1e9, K=a.K)
    results, rep = srv.serve(reqs)
    offered = sum(512 * r.H ** 3 * r.steps for r in reqs) / max(reqs[-1].arrival, 1e-9) / 1e6
    rep.update(rate=a.rate, steps_scale=a.steps_scale, offered_mlups=offered)
    print(json.dumps(rep), flush=True)
    srv.close()


if __name__ == "__main__":
    main()
