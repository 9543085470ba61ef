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
    assert sorted(os.listdir(p)) == ["f.npy", "force.npy", "force64.npy", "snaps_q.npy", "snaps_xy.npy",
                                     "snaps_xz.npy", "state.json", "uacc.npy"]
    st = json.load(open(os.path.join(p, "state.json")))
    assert st == {"t": 6, "n_acc": 3, "sig": SIG}
    # the files the reference reads (lbm3d.py:331-335) have its dtypes and shapes:
    # f (19, nx, ny, nz) float32, force (n_total, nz, 2) float32, uacc (4, nx, ny, nz) float32
    f = np.load(os.path.join(p, "f.npy"))
    assert f.shape == (19, 4, 3, 5) and f.dtype == np.float32
    F32 = np.load(os.path.join(p, "force.npy"))
    assert F32.shape == (10, 5, 2) and F32.dtype == np.float32
    assert np.array_equal(F32[:6], s.F[:6].astype(np.float32)) and not F32[6:].any()
    U = np.load(os.path.join(p, "uacc.npy"))
    assert U.shape == (4, 4, 3, 5) and U.dtype == np.float32
    assert np.load(os.path.join(p, "force64.npy")).dtype == np.float64
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


def test_reference_written_directory_resumes(tmp_path):
    """A directory as the reference writes it (no force64.npy, force.npy float32) resumes:
    the force rows come from the float32 file."""
    s = FakeSim()
    p = str(tmp_path / "case.ckpt")
    checkpoint.save(s, p, 6, 3, SIG)
    os.remove(os.path.join(p, "force64.npy"))
    r = FakeSim(seed=1)
    t, n_acc, _ = checkpoint.load(r, p, SIG)
    assert (t, n_acc) == (6, 3)
    assert r.F.dtype == np.float64
    assert np.array_equal(r.F, s.F[:6].astype(np.float32).astype(np.float64))


class _Args:
    H, porosity, seed, Re, U0, Cs, snap3d_every = 40, 0.2, 7, 15000.0, 0.08, 0.16, 2
    dtype, shift, ybc, gust_seed = "f32", "inlet", "periodic", None


def test_signature_reference_keys_at_reference_settings():
    """At the reference's settings the signature has exactly the eleven keys of
    lbm3d.py:324-325, so a checkpoint written by either side is accepted by the other."""
    from stonewall3d.lbm3d import REFERENCE_SIG_KEYS, run_signature
    a = _Args()
    sig = run_signature(a, [800, 160, 321], 11699, 29247, 146, 3)
    assert tuple(sig) == REFERENCE_SIG_KEYS
    assert sig == dict(grid=[800, 160, 321], H=40, porosity=0.2, seed=7, Re=15000.0, U0=0.08, Cs=0.16, n_spin=11699,
                       n_total=29247, snap_every=146, snap3d=[2, 3])
    # a JSON round trip (state.json) compares equal
    assert json.loads(json.dumps(sig)) == sig
    for field, value, extra in (("dtype", "f16", {"dtype", "shift"}), ("dtype", "f64", {"dtype"}),
                                ("ybc", "freeslip", {"ybc"}), ("gust_seed", 3, {"gust_seed"})):
        b = _Args()
        setattr(b, field, value)
        s2 = run_signature(b, [800, 160, 321], 11699, 29247, 146, 3)
        assert set(s2) - set(REFERENCE_SIG_KEYS) == extra, field


def test_divergence_guard():
    """The NaN check of lbm3d.py:386-391 stops the run on a non-finite field value."""
    from stonewall3d.lbm3d import divergence_guard
    U = np.ones((4, 3, 2, 2), np.float32)
    divergence_guard(U, "c", 5000)
    for bad in (np.nan, np.inf, -np.inf):
        V = U.copy()
        V[2, 1, 0, 1] = bad
        with pytest.raises(SystemExit, match="diverged at step 5000"):
            divergence_guard(V, "c", 5000)
