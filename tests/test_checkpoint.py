"""Checkpoint directory format, atomic swap and signature guard (host logic, no GPU):
lbm3d.py:320-355.  A stand-in object provides the Simulation methods the module uses."""
import json
import os

import numpy as np
import pytest

from stonewall3d import checkpoint


class FakeSim:
    def __init__(self, nx=4, ny=3, nz=5, n_steps=10, seed=0):
        rng = np.random.default_rng(seed)
        self.nx, self.ny, self.nz, self.n_steps = nx, ny, nz, n_steps
        self.f = rng.random((19, nx, ny, nz)).astype(np.float32)
        self.F = rng.random((6, nz, 2))
        self.acc = rng.random((4, nx, ny, nz)).astype(np.float32)
        self.n_acc = 3
        self.t = 6

    def populations(self):
        return self.f.copy()

    def force_rows(self):
        return self.F[:self.t].copy()

    def accumulator(self):
        return self.acc.copy()

    def load_populations(self, a):
        self.f = a.copy()

    def set_force_rows(self, r):
        self.F = np.array(r)

    def set_accumulator(self, a, n):
        self.acc, self.n_acc = a.copy(), n


SIG = dict(grid=[4, 3, 5], H=2, porosity=0.2, seed=7, dtype="f32")


def test_round_trip_and_layout(tmp_path):
    s = FakeSim()
    p = str(tmp_path / "case.ckpt")
    checkpoint.save(s, p, 6, 3, SIG)
    assert sorted(os.listdir(p)) == ["f.npy", "force.npy", "snaps_q.npy", "snaps_xy.npy", "snaps_xz.npy",
                                     "state.json", "uacc.npy"]
    st = json.load(open(os.path.join(p, "state.json")))
    assert st == {"t": 6, "n_acc": 3, "sig": SIG}
    assert np.load(os.path.join(p, "f.npy")).shape == (19, 4, 3, 5)
    assert np.load(os.path.join(p, "force.npy")).shape == (10, 5, 2)
    r = FakeSim(seed=1)
    t, n_acc, snaps = checkpoint.load(r, p, SIG)
    assert (t, n_acc) == (6, 3) and r.t == 6
    assert np.array_equal(r.f, s.f) and np.array_equal(r.acc, s.acc)
    assert np.array_equal(r.F, s.F[:6])


def test_signature_mismatch_refused(tmp_path):
    s = FakeSim()
    p = str(tmp_path / "case.ckpt")
    checkpoint.save(s, p, 6, 3, SIG)
    with pytest.raises(SystemExit):
        checkpoint.load(FakeSim(), p, dict(SIG, porosity=0.4))
    with pytest.raises(SystemExit):
        checkpoint.load(FakeSim(), p, dict(SIG, dtype="f16"))


def test_interrupted_save_keeps_previous(tmp_path, monkeypatch):
    s = FakeSim()
    p = str(tmp_path / "case.ckpt")
    checkpoint.save(s, p, 6, 3, SIG)
    s2 = FakeSim(seed=5)

    def boom(src, dst):
        raise KeyboardInterrupt("killed between tmp and replace")

    monkeypatch.setattr(checkpoint.os, "replace", boom)
    monkeypatch.setattr(checkpoint.shutil, "rmtree", lambda *a, **k: None)
    with pytest.raises(KeyboardInterrupt):
        checkpoint.save(s2, p, 8, 4, SIG)
    monkeypatch.undo()
    r = FakeSim(seed=9)
    t, n_acc, _ = checkpoint.load(r, p, SIG)
    assert (t, n_acc) == (6, 3)
    assert np.array_equal(r.f, s.f)
