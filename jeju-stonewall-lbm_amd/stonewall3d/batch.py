"""Batched form: many independent lattices advanced together on one GPU.

The reference runs one lattice per process (stonewall3d/run_all.sh:12-18).  Here every
lattice of a batch lives in HBM (two population buffers, link masks, inlet, gust table,
force-partials ring, force record) and is described by one entry of a device-resident
instance table (struct sw_instance).  One persistent kernel per step strides over the
row-groups of all lattices; K steps plus the force reduction and the step advance are
captured once as a hipGraph and replayed.  The composition of the batch lives only in
the table, so lattices can join and leave between replays without recapture.

Each lattice's arithmetic is that of the single-lattice step, so its populations and
force rows are bit-identical alone and in any mix (given the same tile variant).

FP32 batches run the two-step form by default: each launch advances every lattice by two
steps (sw_batch_step2: 64 x 10-cell tiles of all lattices taken by persistent workgroups
from a device counter), a tail launch takes the odd last step of lattices that end inside
the chunk, and the force rows come from the two-step per-tile partials.  A lattice's
populations are bit-identical to single steps; its force rows are bit-identical to a
Simulation(two_step=True) run with sw_set_step2_rows(sw_batch_step2_rows()).
"""
import ctypes

import numpy as np
import torch

from . import _native as nat
from .lattice import Q, flag_from_mask
from .lbm3d import DEFAULT_SCALE, DEFAULT_SHIFT, TORCH_STORAGE, _stream, storage_shift


class Lattice:
    """Device resources of one lattice of a batch (the per-instance part of Simulation)."""

    def __init__(self, flag, uin, *, nu0, cs=0.16, x_sponge=None, nu_sponge=0.05, ybc="freeslip", dtype="f32",
                 scale=None, gust=None, n_steps=1000, ring=32, device="cuda", tag=None, shift=None):
        lib = nat.lib()
        flag = np.asarray(flag)
        if flag.dtype == bool:
            flag = flag_from_mask(flag)
        flag = np.ascontiguousarray(flag, np.uint8)
        self.nx, self.ny, self.nz = flag.shape
        self.N = self.nx * self.ny * self.nz
        self.dtype = nat.DTYPES[dtype] if isinstance(dtype, str) else int(dtype)
        self.scale = float(DEFAULT_SCALE[self.dtype] if scale is None else scale)
        self.ybc = {"periodic": nat.SW_Y_PERIODIC, "freeslip": nat.SW_Y_FREESLIP}[ybc] if isinstance(ybc, str) else ybc
        self.nu0, self.cs, self.nu_sponge = float(nu0), float(cs), float(nu_sponge)
        self.x_sponge = self.nx if x_sponge is None else int(x_sponge)
        self.n_steps, self.ring, self.tag = int(n_steps), int(ring), tag
        self.device = torch.device(device)
        dev = self.device
        self.links = torch.empty(lib.sw_links_words(self.nx, self.ny, self.nz), dtype=torch.int32, device=dev)
        flag_d = torch.from_numpy(flag).to(dev)
        nat.call("sw_prepare_links", flag_d.data_ptr(), self.links.data_ptr(), self.nx, self.ny, self.nz, self.ybc,
                 _stream())
        del flag_d
        st = TORCH_STORAGE[self.dtype]
        self.f = [torch.empty(Q * self.N, dtype=st, device=dev), torch.zeros(Q * self.N, dtype=st, device=dev)]
        self.uin = torch.from_numpy(np.ascontiguousarray(uin, np.float32)).to(dev)
        self.shift_kind = DEFAULT_SHIFT if shift is None else shift
        self.shift = storage_shift(self.uin, self.dtype, self.shift_kind)
        self.gust = None
        if gust is not None:
            g = np.ascontiguousarray(gust, np.float32)
            if g.ndim != 1 or g.shape[0] < self.n_steps:
                raise ValueError(f"gust table has {g.shape[0] if g.ndim else 0} entries; the lattice's "
                                 f"{self.n_steps} steps read one per step")
            self.gust = torch.from_numpy(g[:self.n_steps].copy()).to(dev)
        # room for either partials layout (one-step row groups or two-step tiles)
        self.partials = torch.zeros(self.ring * lib.sw_batch_partials_per_step(self.nx, self.ny, self.nz),
                                    dtype=torch.float64, device=dev)
        self.force = torch.zeros((self.n_steps, self.nz, 2), dtype=torch.float64, device=dev)
        self.t = 0
        self.parity = 0

    @property
    def nbytes(self):
        return sum(x.numel() * x.element_size() for x in (self.links, self.f[0], self.f[1], self.uin, self.partials,
                                                          self.force)) + \
            (self.gust.numel() * 4 if self.gust is not None else 0) + \
            (self.shift.numel() * 4 if self.shift is not None else 0)

    def init(self, noise=0.0, seed=0):
        nat.call("sw_init", self.f[0].data_ptr(), self.dtype, self.links.data_ptr(), self.uin.data_ptr(), self.nx,
                 self.ny, self.nz, float(noise), int(seed) & 0xFFFFFFFF, self.scale,
                 self.shift.data_ptr() if self.shift is not None else None, _stream())
        self.t, self.parity = 0, 0
        return self

    def instance(self):
        e = nat.SwInstance()
        e.f0, e.f1 = self.f[0].data_ptr(), self.f[1].data_ptr()
        e.links, e.uin = self.links.data_ptr(), self.uin.data_ptr()
        e.gust = self.gust.data_ptr() if self.gust is not None else None
        e.partials, e.force = self.partials.data_ptr(), self.force.data_ptr()
        e.shift = self.shift.data_ptr() if self.shift is not None else None
        e.nu0, e.cs, e.nu_sponge, e.scale = self.nu0, self.cs, self.nu_sponge, self.scale
        e.nx, e.ny, e.nz = self.nx, self.ny, self.nz
        e.ring, e.x_sponge, e.ybc, e.dtype = self.ring, self.x_sponge, self.ybc, self.dtype
        e.t, e.t_end, e.parity = self.t, self.n_steps, self.parity
        return e

    @property
    def done(self):
        return self.t >= self.n_steps

    def current(self):
        """Population buffer holding step t."""
        return self.f[self.parity]

    def populations(self):
        """Raw stored populations of step t in the reference layout (19, nx, ny, nz)."""
        src = self.current()
        out = torch.empty_like(src)
        nat.call("sw_to_ref_layout", src.data_ptr(), out.data_ptr(), nat.DTYPE_BYTES[self.dtype], Q, self.nx, self.ny,
                 self.nz, _stream())
        a = out.cpu()
        if self.dtype in (nat.SW_F16, nat.SW_BF16):
            a = a.view(torch.int16).numpy().view(np.uint16)
        else:
            a = a.numpy()
        return a.reshape(Q, self.nx, self.ny, self.nz)

    def force_rows(self):
        return self.force[:self.t].cpu().numpy()


def default_blocks(device=None, per_cu=8):
    props = torch.cuda.get_device_properties(device or torch.cuda.current_device())
    return int(props.multi_processor_count) * per_cu


def two_step_eligible(lattice):
    """The two-step form takes FP32 / FP16 / BF16 population arrays below 4 GiB."""
    return lattice.dtype != nat.SW_F64 and Q * lattice.N * nat.DTYPE_BYTES[lattice.dtype] < (1 << 32)


class Batch:
    """A set of lattices of one dtype advanced K steps per launch sequence.

    K        : steps per chunk (one graph replay); every lattice's ring must be >= K.
    blocks   : persistent workgroups per step launch (default: 8 per compute unit, enough
               for the step kernels' 6-7 resident wavefronts per SIMD).
    variant  : tile variant (sw_set_step_variant numbering); default: 8 (pairs, one wavefront
               per row) for the 16-bit forms, the batch kernel's fastest at every measured
               lattice size (rows of 128-640 cells), else the library's current choice for
               this dtype.
    graph    : capture the chunk as a hipGraph (else launch it eagerly).
    two_step : advance two steps per launch (sw_batch_step2; K even, FP32 / FP16 / BF16
               lattices below 4 GiB per population array); default: on for FP32, whose
               two-step kernel is the faster one (the 16-bit forms are faster one step per
               launch).  `blocks` then defaults to one workgroup per compute unit.
    """

    def __init__(self, dtype="f32", max_instances=64, K=16, blocks=None, variant=None, graph=True, device="cuda",
                 two_step=None):
        lib = nat.lib()
        self.dtype = nat.DTYPES[dtype] if isinstance(dtype, str) else int(dtype)
        self.max_instances, self.K = int(max_instances), int(K)
        self.device = torch.device(device)
        self.two_step = (self.dtype == nat.SW_F32) if two_step is None else bool(two_step)
        if self.two_step and (self.dtype == nat.SW_F64 or self.K % 2):
            raise ValueError("the two-step batch needs FP32 / FP16 / BF16 storage and an even K")
        self.blocks = int(blocks) if blocks else default_blocks(self.device, per_cu=1 if self.two_step else 8)
        if variant is None:
            variant = 8 if self.dtype in (nat.SW_F16, nat.SW_BF16) else lib.sw_get_step_variant(self.dtype)
        self.variant = int(variant)
        self.table = torch.zeros(lib.sw_batch_table_bytes(self.max_instances), dtype=torch.uint8, device=self.device)
        self.members = []
        self.big = False
        self.use_graph = bool(graph)
        self._graph = None
        self._graph_key = None
        self._upload()

    # ---------------------------------------------------------------- composition
    def _upload(self):
        n = len(self.members)
        arr = (nat.SwInstance * max(n, 1))()
        for i, m in enumerate(self.members):
            arr[i] = m.instance()
        nat.call("sw_batch_upload", ctypes.cast(arr, ctypes.c_void_p), n, self.max_instances, self.table.data_ptr(),
                 _stream())
        esize = nat.DTYPE_BYTES[self.dtype]
        self.big = any(Q * m.N * esize >= (1 << 32) for m in self.members)

    def sync(self):
        """Pull t / parity of every member back from the device table."""
        n = len(self.members)
        if n == 0:
            return
        arr = (nat.SwInstance * n)()
        nat.call("sw_batch_download", self.table.data_ptr(), ctypes.cast(arr, ctypes.c_void_p), n, _stream())
        for m, e in zip(self.members, arr):
            m.t, m.parity = int(e.t), int(e.parity)

    def add(self, lattice):
        if lattice.dtype != self.dtype:
            raise ValueError("a batch holds lattices of one dtype")
        if lattice.ring < self.K:
            raise ValueError(f"lattice ring {lattice.ring} < K={self.K}")
        if len(self.members) >= self.max_instances:
            raise RuntimeError("batch is full")
        if self.two_step and not two_step_eligible(lattice):
            raise ValueError("the two-step batch takes population arrays below 4 GiB")
        self.sync()
        self.members.append(lattice)
        self._upload()

    def remove(self, lattice):
        self.sync()
        self.members = [m for m in self.members if m is not lattice]
        self._upload()

    def retire_finished(self):
        """Remove and return the members that reached their end step."""
        self.sync()
        done = [m for m in self.members if m.done]
        if done:
            self.members = [m for m in self.members if not m.done]
            self._upload()
        return done

    # ---------------------------------------------------------------- stepping
    def _ensure_graph(self):
        key = (self.dtype, self.K, self.blocks, self.variant, self.big, self.two_step)
        if self._graph is not None and self._graph_key == key:
            return
        self.close()
        h = ctypes.c_void_p()
        if self.two_step:
            nat.call("sw_batch_graph_create2", self.table.data_ptr(), self.max_instances, self.dtype, self.K,
                     self.blocks, ctypes.byref(h))
        else:
            nat.call("sw_batch_graph_create", self.table.data_ptr(), self.max_instances, self.dtype, self.K,
                     self.blocks, self.variant, int(self.big), ctypes.byref(h))
        self._graph, self._graph_key = h, key

    def run_chunk(self):
        """Advance every unfinished member by up to K steps (asynchronous on the current stream)."""
        if not self.members:
            return
        s = _stream()
        if self.use_graph:
            self._ensure_graph()
            nat.call("sw_graph_launch", self._graph, s)
        elif self.two_step:
            for p in range(self.K // 2):
                nat.call("sw_batch_step2", self.table.data_ptr(), self.dtype, p, self.blocks, s)
            nat.call("sw_batch_step2_tail", self.table.data_ptr(), self.dtype, self.K, self.blocks, s)
            nat.call("sw_batch_force_reduce2", self.table.data_ptr(), self.max_instances, self.K, s)
            nat.call("sw_batch_advance2", self.table.data_ptr(), self.max_instances, self.K, s)
        else:
            for k in range(self.K):
                nat.call("sw_batch_step", self.table.data_ptr(), self.dtype, k, self.blocks, self.variant,
                         int(self.big), s)
            nat.call("sw_batch_force_reduce", self.table.data_ptr(), self.max_instances, self.K, s)
            nat.call("sw_batch_advance", self.table.data_ptr(), self.max_instances, self.K, s)

    def run(self, chunks):
        for _ in range(int(chunks)):
            self.run_chunk()
        self.sync()

    def run_to_completion(self):
        while True:
            self.sync()
            if all(m.done for m in self.members):
                return
            self.run_chunk()

    def close(self):
        if self._graph is not None:
            nat.lib().sw_graph_destroy(self._graph)
            self._graph = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def main(argv=None):
    """Batched run: n lattices 16H x 8H x 4H (H from {8,...,24}, porosity U[0, 0.45],
    per-instance seeds), `steps` each, one persistent kernel per step over all of them,
    K-step chunks replayed as a hipGraph, tile variant and grid picked by the autotuner."""
    import argparse
    import json
    import sys
    import time

    from . import workloads
    from .autotune import Autotuner

    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--instances", type=int, default=64)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--dtype", default="f32", choices=["f32", "f16", "bf16", "f64"])
    ap.add_argument("--K", type=int, default=32)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--one-step", action="store_true",
                    help="one step per launch (default for FP32: two, sw_batch_step2)")
    a = ap.parse_args(argv)
    spec = workloads.batch_instances(a.instances, seed=a.seed, n_steps=a.steps)
    cells = 0
    b = Batch(a.dtype, max_instances=a.instances, K=a.K, two_step=False if a.one_step else None)
    for H, por, seed, steps in spec:
        w = workloads.porous_wall(H=H, porosity=por, seed=seed, n_steps=steps)
        b.add(Lattice(w.flag, w.uin, dtype=a.dtype, ring=a.K, **w.sim_kwargs()).init(noise=0.02, seed=seed))
        cells += w.cells
    tuned = None
    if not a.no_tune:
        cand = [1, 3, 7] + ([8, 9, 10] if a.dtype != "f64" else [])
        tuned = Autotuner().tune_batch(b, variants=cand, per_cu=(4, 8, 16, 32), chunks=2)
        print(f"[batch] tuned: variant {b.variant}, {b.blocks} workgroups", file=sys.stderr, flush=True)
        # restart every lattice from its initial state after tuning
        for (H, por, seed, steps), m in zip(spec, b.members):
            m.init(noise=0.02, seed=seed)
        b._upload()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    chunks = 0
    while True:
        b.run_chunk()
        chunks += 1
        if chunks % 16 == 0:
            b.sync()
            if all(m.done for m in b.members):
                break
            print(f"[batch] step {min(m.t for m in b.members)}/{a.steps} {time.perf_counter() - t0:.1f}s",
                  file=sys.stderr, flush=True)
    b.sync()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    assert all(m.done for m in b.members)
    finite = all(bool(torch.isfinite(m.force[:m.t]).all()) for m in b.members)
    bpu = {"f32": 152, "f16": 76, "bf16": 76, "f64": 304}[a.dtype]
    out = dict(form="batched", dtype=a.dtype, instances=a.instances, cells=cells, steps=a.steps, seconds=el,
               aggregate_mlups=cells * a.steps / el / 1e6, achieved_gbs=cells * a.steps * bpu / el / 1e9,
               variant=None if b.two_step else b.variant, two_step=b.two_step, workgroups=b.blocks, K=a.K,
               graph=b.use_graph, finite=finite,
               tuning_ms_per_step={f"{k[0]}/{k[1]}": v for k, v in tuned.items()} if tuned else None)
    print(json.dumps(out), flush=True)
    b.close()


if __name__ == "__main__":
    main()
