"""Project autotuner: picks the step kernel's tile variant (and, for batches, the
persistent grid size) by timing candidates on the actual lattice shapes.

The reference has a single Warp kernel and no tuning; this module is new.  Choices are
cached in memory (and optionally in a JSON file) by a key made of the form (single
lattice / batch), the storage dtype and a size class, so a serving loop tunes once per
shape class.  Every variant is bit-identical in the populations; force rows differ only
in FP64 summation order between variants with different wavefronts per workgroup.
"""
import json
import os

import numpy as np
import torch

from . import _native as nat


def _time_ms(fn, reps):
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    fn()
    torch.cuda.synchronize()
    s.record()
    for _ in range(reps):
        fn()
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / reps


def size_class(cells):
    """Power-of-two bucket of the lattice size (cells)."""
    return int(np.ceil(np.log2(max(int(cells), 1))))


class Autotuner:
    def __init__(self, path=None):
        self.path = path
        self.cache = {}
        if path and os.path.exists(path):
            with open(path) as f:
                self.cache = json.load(f)

    def save(self):
        if self.path:
            with open(self.path, "w") as f:
                json.dump(self.cache, f, indent=1)

    # ---------------------------------------------------------------- single lattice
    def tune_step(self, sim, variants=None, steps=10, apply=True):
        """Time sim.step for each tile variant (the state of `sim` advances) and select the
        fastest for sim's dtype.  Returns {variant: ms_per_step}."""
        lib = nat.lib()
        key = f"step/{sim.dtype_name}/{size_class(sim.N)}"
        if key in self.cache:
            best = self.cache[key]["best"]
            if apply:
                nat.call("sw_set_step_variant", sim.dtype, best)
            return self.cache[key]["ms"]
        variants = list(range(lib.sw_step_variants())) if variants is None else list(variants)
        old = lib.sw_get_step_variant(sim.dtype)
        res = {}
        for v in variants:
            if sim.t + steps + 1 > sim.n_steps:
                break
            nat.call("sw_set_step_variant", sim.dtype, v)
            res[v] = _time_ms(lambda: sim.step(1), steps)
        best = min(res, key=res.get) if res else old
        self.cache[key] = {"best": int(best), "ms": {str(k): float(v) for k, v in res.items()}}
        nat.call("sw_set_step_variant", sim.dtype, int(best) if apply else old)
        self.save()
        return res

    # ---------------------------------------------------------------- batches
    def tune_batch(self, batch, variants=None, per_cu=(2, 4, 6, 8, 16), chunks=2, apply=True):
        """Time batch chunks for each (variant, persistent grid) pair and keep the fastest
        on `batch`.  The members advance during tuning: tune a throwaway batch of the
        same shapes, or re-initialise the members afterwards."""
        lib = nat.lib()
        cells = sum(m.N for m in batch.members)
        two = getattr(batch, "two_step", False)
        key = f"batch/{batch.dtype}/{size_class(cells)}/{len(batch.members)}" + ("/two-step" if two else "")
        cu = torch.cuda.get_device_properties(batch.device).multi_processor_count
        if two:
            # the two-step batch kernel has no tile variants, and its workgroup takes the
            # LDS of a whole compute unit: one per compute unit is the only useful grid
            variants, per_cu = [batch.variant], (1,)
        if key in self.cache:
            b = self.cache[key]
            if apply:
                batch.variant, batch.blocks = b["variant"], b["blocks"]
            return b["ms"]
        variants = list(range(lib.sw_step_variants())) if variants is None else list(variants)
        old = (batch.variant, batch.blocks)
        if two:
            variants = [batch.variant]
        res = {}
        for v in variants:
            for p in per_cu:
                batch.variant, batch.blocks = v, cu * p
                for m in batch.members:
                    m.t, m.parity = 0, 0
                batch._upload()
                res[(v, cu * p)] = _time_ms(batch.run_chunk, chunks) / batch.K
        best = min(res, key=res.get)
        self.cache[key] = {"variant": int(best[0]), "blocks": int(best[1]),
                           "ms": {f"{k[0]}/{k[1]}": float(v) for k, v in res.items()}}
        batch.variant, batch.blocks = best if apply else old
        for m in batch.members:
            m.t, m.parity = 0, 0
        batch._upload()
        self.save()
        return res
