"""The benchmark configurations at their own sizes, and the CLI's file surface (needs an MI355X).

  c2  H = 40 porous wall at 40 % void (640 x 320 x 160) in FP16 and BF16 storage with the
      inlet offsets, against the oracle at full size; the FP64 form at H = 20
      (320 x 160 x 80); mean Cd of the 16-bit forms within 1 % of the FP64 run.
  c3  the 64-lattice batch mix (H from {8, 12, 16, 20, 24}, porosity U[0, 0.45]) in FP32
      and FP16: every lattice identical in the mix and alone.
  c4  a serving trace drawn from the c4 distribution (H U[10, 40], FP32 and FP16, step
      counts scaled down): every request's populations and force rows identical to its
      lattice run alone.
  CLI checkpoints in the reference's format resume; the run's NaN check stops a run whose
      state is not finite; every .npz key has the reference's dtype and shape.
"""
import json
import os

import numpy as np
import pytest
import torch

from oracle import c_oracle as co
from stonewall3d import _native as nat
from stonewall3d import workloads

pytestmark = pytest.mark.gpu

SCALE = {"f32": 1.0, "f16": 1024.0, "bf16": 1.0, "f64": 1.0}


# ------------------------------------------------------------------------------------------- c2
@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_c2_porous_wall_40pct_full_size_16bit(gpu, storage):
    """640x320x160, 40 % void, gusting inlet, 16-bit storage with the per-height inlet
    offsets: 3 steps bit-identical to the oracle's 16-bit form at full size."""
    from stonewall3d.lbm3d import Simulation
    w = workloads.porous_wall(H=40, porosity=0.4, seed=0, n_steps=3)
    assert w.shape == (640, 320, 160)
    print(f"c2 {storage}: {w.name}", flush=True)
    co.set_threads(16)
    sim = Simulation(w.flag, w.uin, dtype=storage, shift="inlet", **w.sim_kwargs())
    sim.init(noise=0.02, seed=0)
    sim.step(3)
    sh = co.shift_table(w.uin, "inlet")
    f = co.init(w.flag, w.uin, storage, noise=0.02, seed=0, scale=SCALE[storage], shift=sh)
    F = []
    for t in range(3):
        f, Ft = co.step(f, w.flag, w.uin, w.nu0, w.cs, w.x_sponge, w.nu_sponge, gust=float(w.gust[t]), ybc=1,
                        storage=storage, scale=SCALE[storage], shift=sh)
        F.append(Ft)
    g = sim.populations()
    fl = w.flag == 0
    differ = sum(int(((g[i] != f[i]) & fl).sum()) for i in range(19))
    assert differ == 0, f"{differ} populations differ"
    F = np.array(F)
    assert np.abs(F).max() > 0
    assert np.allclose(sim.force_rows(), F, rtol=1e-11, atol=1e-14 * np.abs(F).max())
    del sim
    torch.cuda.empty_cache()


def test_c2_fp64_half_resolution(gpu):
    """The FP64 form on the H = 20 lattice of the c2 comparison (320x160x80, 40 % void,
    gust evaluated at twice the step index): 2 steps within 1e-14 of the FP64 oracle."""
    from stonewall3d.lbm3d import Simulation
    w = workloads.porous_wall(H=20, porosity=0.4, seed=0, n_steps=2, gust_time_scale=2.0)
    assert w.shape == (320, 160, 80)
    sim = Simulation(w.flag, w.uin, dtype="f64", **w.sim_kwargs())
    sim.init(noise=0.02, seed=0)
    sim.step(2)
    f = co.init(w.flag, w.uin, "f64", noise=0.02, seed=0)
    F = []
    for t in range(2):
        f, Ft = co.step(f, w.flag, w.uin, w.nu0, w.cs, w.x_sponge, w.nu_sponge, gust=float(w.gust[t]), ybc=1,
                        storage="f64")
        F.append(Ft)
    g = sim.populations()
    fl = (w.flag == 0)[None]
    rel = np.abs(np.where(fl, g - f, 0)).max() / np.abs(f).max()
    print("c2 fp64 H=20 max rel diff", rel)
    assert rel < 1e-14
    assert np.allclose(sim.force_rows(), np.array(F), rtol=1e-12, atol=1e-16)


@pytest.fixture(scope="module")
def c2_fp64_run(gpu):
    from stonewall3d import precision
    return precision.run_case(40, 0.4, 0, "f64", 10.0, 20.0)


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_c2_mean_cd_within_1pct_of_fp64(gpu, c2_fp64_run, storage):
    """The c2 precision claim: 640x320x160 at 40 % void, gusting inlet, 10 H/U spin-up and
    20 H/U averaging (24 000 steps); mean Cd of the 16-bit storage (inlet offsets, FP32
    collision) within 1 % of the FP64 run on the same lattice and seed."""
    from stonewall3d import precision
    r = precision.run_case(40, 0.4, 0, storage, 10.0, 20.0)
    ref = c2_fp64_run
    err = (r["Cd"] - ref["Cd"]) / ref["Cd"]
    print(f"c2 {storage}: mean Cd {r['Cd']:.5f} vs FP64 {ref['Cd']:.5f} ({100 * err:+.3f} %), gust-peak Cd "
          f"{r['CdMax']:.4f} vs {ref['CdMax']:.4f}", flush=True)
    assert r["finite"] and ref["finite"]
    assert abs(err) < 0.01


# ------------------------------------------------------------------------------------------- c3
@pytest.mark.parametrize("dtype", ["f32", "f16"])
def test_c3_batch_of_64_alone_equals_mix(gpu, dtype):
    """The batched configuration's own mix (workloads.batch_instances(64): 64 lattices
    16H x 8H x 4H, H from {8, 12, 16, 20, 24}, 210.9 M cells): two 16-step graph chunks of
    the persistent batch kernel (FP32: the two-step form, FP16: one step per launch), then
    every lattice's population buffer and force rows equal those of the same lattice
    stepped alone with the same kernels and tiles."""
    from stonewall3d.batch import Batch, Lattice
    from stonewall3d.lbm3d import Simulation
    K, steps = 16, 32
    spec = workloads.batch_instances(64, n_steps=steps)
    b = Batch(dtype, max_instances=64, K=K, graph=True)
    lats = []
    for H, por, seed, n in spec:
        w = workloads.porous_wall(H=H, porosity=por, seed=seed, n_steps=n)
        lat = Lattice(w.flag, w.uin, dtype=dtype, ring=K, **w.sim_kwargs()).init(noise=0.02, seed=seed)
        b.add(lat)
        lats.append((lat, w, seed))
    print(f"c3 {dtype}: {sum(l.N for l, _, _ in lats) / 1e6:.1f} M cells, two-step {b.two_step}, variant {b.variant}",
          flush=True)
    b.run(2)
    torch.cuda.synchronize()
    code = nat.DTYPES[dtype]
    old = nat.lib().sw_get_step_variant(code)
    old_rows = nat.lib().sw_get_step2_rows()
    nat.call("sw_set_step_variant", code, b.variant)
    nat.call("sw_set_step2_rows", nat.lib().sw_batch_step2_rows())
    try:
        for lat, w, seed in lats:
            assert lat.t == steps
            sim = Simulation(w.flag, w.uin, dtype=dtype, two_step=b.two_step, **w.sim_kwargs())
            sim.init(noise=0.02, seed=seed)
            sim.step(steps)
            assert torch.equal(lat.current(), sim.f[0]), f"lattice {w.name} differs in the mix"
            assert np.array_equal(lat.force_rows(), sim.force_rows())
            del sim
    finally:
        nat.call("sw_set_step_variant", code, old)
        nat.call("sw_set_step2_rows", old_rows)
    b.close()
    del b, lats
    torch.cuda.empty_cache()


# ------------------------------------------------------------------------------------------- c4
def test_c4_serving_trace_alone_equals_mix(gpu):
    """A trace from the serving configuration's distribution (8 requests, H U[10, 40] with an
    H = 40 case, porosity U[0, 0.45], FP32 and FP16, step counts x 0.004), served with the
    default admission (shortest job first, 64 M-cell fill cap) so that requests join while
    others run: each request's final populations and force rows equal its lattice run alone."""
    from stonewall3d.lbm3d import Simulation
    from stonewall3d.serving import Server, trace
    reqs = trace(8, 8.0, seed=7, steps_scale=0.004, h_range=(10, 40))
    assert {r.dtype for r in reqs} == {"f32", "f16"} and max(r.H for r in reqs) == 40
    srv = Server(budget_bytes=250e9, K=8, keep_force=True, keep_state=True, policy="sjf", fill_cells=64e6)
    results, rep = srv.serve(reqs, realtime=False)
    srv.close()
    assert rep["requests"] == len(reqs)
    for r in results:
        req = reqs[r.rid]
        w = workloads.porous_wall(H=req.H, porosity=req.porosity, seed=req.seed, n_steps=req.steps)
        code = nat.DTYPES[req.dtype]
        bt = srv.batches[req.dtype]
        old = nat.lib().sw_get_step_variant(code)
        old_rows = nat.lib().sw_get_step2_rows()
        nat.call("sw_set_step_variant", code, bt.variant)
        nat.call("sw_set_step2_rows", nat.lib().sw_batch_step2_rows())
        try:
            sim = Simulation(w.flag, w.uin, dtype=req.dtype, two_step=bt.two_step, **w.sim_kwargs())
            sim.init(noise=0.02, seed=req.seed)
            sim.step(req.steps)
            assert torch.equal(r.state, sim.f[0]), f"request {r.rid} (H={req.H}, {req.dtype}) differs in the mix"
            assert np.array_equal(r.force, sim.force_rows())
            del sim
        finally:
            nat.call("sw_set_step_variant", code, old)
            nat.call("sw_set_step2_rows", old_rows)
    del results
    torch.cuda.empty_cache()


# ------------------------------------------------------------------------------------------- CLI
class _Stop(Exception):
    pass


def _stop_at(monkeypatch, t_stop):
    """Make Simulation.step raise once the run reaches step t_stop (a killed run)."""
    from stonewall3d import lbm3d
    real = lbm3d.Simulation.step

    def step(self, n=1, **kw):
        if self.t >= t_stop:
            raise _Stop()
        return real(self, n, **kw)

    monkeypatch.setattr(lbm3d.Simulation, "step", step)


CLI = ["--porosity", "0.2", "--H", "8", "--t-spinup", "0.1", "--t-avg", "0.5"]


def test_cli_resumes_reference_format_checkpoint(gpu, tmp_path, monkeypatch):
    """A run killed after its checkpoint at step 25 leaves a directory in the reference's
    format (state.json with exactly the reference's signature keys, f.npy float32
    (19, nx, ny, nz), force.npy float32 (n_total, nz, 2), uacc.npy float32).  With this
    package's float64 force file removed (a directory the reference itself wrote), the
    resumed run's outputs equal an uninterrupted run's: means and snapshots exactly (the
    populations are), force-derived series to rounding (before step 25: the float32 rows the
    reference stores)."""
    from stonewall3d import lbm3d
    straight, resumed = str(tmp_path / "a"), str(tmp_path / "b")
    lbm3d.main(["--case", "c", "--out", straight, "--ckpt-every", "0"] + CLI)
    _stop_at(monkeypatch, 30)
    with pytest.raises(_Stop):
        lbm3d.main(["--case", "c", "--out", resumed, "--ckpt-every", "25"] + CLI)
    monkeypatch.undo()
    ck = os.path.join(resumed, "c.ckpt")
    st = json.load(open(os.path.join(ck, "state.json")))
    assert tuple(st["sig"]) == lbm3d.REFERENCE_SIG_KEYS and st["t"] == 25
    nx, ny, nz = st["sig"]["grid"]
    f = np.load(os.path.join(ck, "f.npy"))
    assert f.dtype == np.float32 and f.shape == (19, nx, ny, nz)
    F = np.load(os.path.join(ck, "force.npy"))
    assert F.dtype == np.float32 and F.shape == (st["sig"]["n_total"], nz, 2)
    U = np.load(os.path.join(ck, "uacc.npy"))
    assert U.dtype == np.float32 and U.shape == (4, nx, ny, nz)
    os.remove(os.path.join(ck, "force64.npy"))
    lbm3d.main(["--case", "c", "--out", resumed, "--ckpt-every", "25"] + CLI)
    assert not os.path.isdir(ck)
    a, b = np.load(os.path.join(straight, "c.npz")), np.load(os.path.join(resumed, "c.npz"))
    assert a.files == b.files
    for k in ("Cd", "Cl", "Cm"):
        # from step 25 on the force rows are computed alike, but the runs split their steps into
        # launches differently (the checkpoint is an event of one of them), and one- and
        # two-step launches sum the force partials in different fixed orders
        assert np.allclose(a[k][25:], b[k][25:], rtol=1e-6, atol=1e-9), k
        assert np.allclose(a[k][:25], b[k][:25], rtol=1e-5, atol=1e-7), k
    assert np.allclose(a["fx_profile"], b["fx_profile"], rtol=1e-5, atol=1e-8)
    for k in ("mean_xz_mid", "mean_xy_half", "mean3d_ds2", "snaps_xz", "snaps_xy", "snaps_q", "wall_ds2"):
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(a["mean_xz_spanavg"], b["mean_xz_spanavg"], equal_nan=True)


def test_cli_pairs_equal_single_steps(gpu, tmp_path, monkeypatch):
    """The CLI's FP32 default (two steps per launch, the time-mean samples taken inside the
    pairs at the offsets of each event group) writes the same means, snapshots and Q frames as
    a run of single steps (--one-step), bit for bit; the force-derived series agree to
    rounding (the two kernels sum the force partials in different fixed orders)."""
    from stonewall3d import lbm3d
    calls = []
    real = lbm3d.Simulation.step

    def step(self, n=1, accumulate=False, samples=None):
        calls.append((self.two_step, n, sorted(samples or ())))
        return real(self, n, accumulate=accumulate, samples=samples)

    monkeypatch.setattr(lbm3d.Simulation, "step", step)
    pairs, single = str(tmp_path / "pairs"), str(tmp_path / "single")
    lbm3d.main(["--case", "p", "--out", pairs, "--ckpt-every", "0"] + CLI)
    sampled = [k for two, n, ks in calls if two for k in ks]
    # samples at both positions of a pair (sw_step2_accumulate's which = 1 and 2)
    assert all(two for two, _, _ in calls) and {k % 2 for k in sampled} == {0, 1}
    calls.clear()
    lbm3d.main(["--case", "p", "--out", single, "--ckpt-every", "0", "--one-step"] + CLI)
    assert not any(two for two, _, _ in calls)
    assert sum(len(ks) for _, _, ks in calls) == len(sampled)
    a, b = np.load(os.path.join(pairs, "p.npz")), np.load(os.path.join(single, "p.npz"))
    assert a.files == b.files
    for k in ("mean_xz_mid", "mean_xy_half", "mean3d_ds2", "snaps_xz", "snaps_xy", "snaps_q", "wall_ds2"):
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(a["mean_xz_spanavg"], b["mean_xz_spanavg"], equal_nan=True)
    for k in ("Cd", "Cl", "Cm", "fx_profile"):
        assert np.allclose(a[k], b[k], rtol=1e-6, atol=1e-9), k


def test_cli_divergence_guard_stops_the_run(gpu, tmp_path, monkeypatch):
    """The run's NaN check (lbm3d.py:386-391, every 5000 steps): a checkpoint whose state
    holds one NaN population in a fluid cell, resumed at step 4999, makes the run stop with
    SystemExit at the check of step 5000."""
    from stonewall3d import lbm3d
    out = str(tmp_path)
    args = ["--case", "n", "--out", out, "--ckpt-every", "25", "--t-avg", "50", "--max-steps", "5001",
            "--porosity", "0.2", "--H", "8", "--t-spinup", "0.1"]
    _stop_at(monkeypatch, 30)
    with pytest.raises(_Stop):
        lbm3d.main(args)
    monkeypatch.undo()
    ck = os.path.join(out, "n.ckpt")
    f = np.load(os.path.join(ck, "f.npy"))
    nx, ny, nz = f.shape[1:]
    f[3, nx // 2, ny // 2, nz - 5] = np.nan                      # a fluid cell well above the wall
    np.save(os.path.join(ck, "f.npy"), f)
    st = json.load(open(os.path.join(ck, "state.json")))
    st["t"] = 4999
    json.dump(st, open(os.path.join(ck, "state.json"), "w"))
    with pytest.raises(SystemExit, match="diverged at step 5000"):
        lbm3d.main(args)


def _reference_npz_spec(nx, ny, nz, n_total, n_snap, n_q, q_dims):
    """dtype and shape of every array lbm3d.py:404-423 writes (Warp arrays are float32; the
    snapshots float16 stacks; wall arrays bool; meta a JSON str)."""
    h = lambda n: (n + 1) // 2  # noqa: E731  (length of a [::2] slice)
    f32, f16 = np.dtype(np.float32), np.dtype(np.float16)
    return dict(Cd=(f32, (n_total,)), Cl=(f32, (n_total,)), Cm=(f32, (n_total,)), fx_profile=(f32, (nz,)),
                mean_xz_mid=(f32, (4, nx, nz)), mean_xz_spanavg=(f32, (4, nx, nz)), mean_xy_half=(f32, (4, nx, ny)),
                mean3d_ds2=(f16, (3, h(nx), h(ny), h(nz))), snaps_xz=(f16, (n_snap, 3, nx, nz)),
                snaps_xy=(f16, (n_snap, 3, nx, ny)), wall_xz=(np.dtype(bool), (nx, nz)),
                wall_xy=(np.dtype(bool), (nx, ny)), wall_front=(np.dtype(bool), (ny, nz)),
                wall_top=(np.dtype(bool), (nx, ny)), wall_ds2=(np.dtype(bool), (h(nx), h(ny), h(nz))),
                snaps_q=(np.dtype(np.uint8), (n_q,) + tuple(q_dims)), meta=("str", ()))


def test_cli_npz_keys_match_the_reference(gpu, tmp_path):
    """Every key of the .npz the CLI writes has the dtype and shape the reference's writer
    (lbm3d.py:404-423) gives it for the same grid, step count and snapshot count."""
    from stonewall3d import lbm3d
    from stonewall3d.inflow import reference_setup
    out = str(tmp_path)
    lbm3d.main(["--case", "k", "--out", out, "--ckpt-every", "0"] + CLI)
    d = np.load(os.path.join(out, "k.npz"))
    meta = json.loads(str(d["meta"]))
    su = reference_setup(H=8, t_spinup=0.1, t_avg=0.5)
    nx, ny, nz = su.nx, su.ny, su.nz
    assert (meta["nx"], meta["ny"], meta["nz"]) == (nx, ny, nz)
    n_snap = len(range(0, su.n_total - su.n_spin, su.snap_every))
    n_q = len(range(0, n_snap, 2))
    spec = _reference_npz_spec(nx, ny, nz, su.n_total, n_snap, n_q, meta["q3d"]["dims"])
    assert sorted(d.files) == sorted(spec)
    for k, (dt, shape) in spec.items():
        if dt == "str":
            assert d[k].dtype.kind == "U" and d[k].shape == shape, k
        else:
            assert d[k].dtype == dt and d[k].shape == shape, (k, d[k].dtype, d[k].shape, dt, shape)
    for k in ("n_spin", "n_avg", "snap_every", "steps_per_unit", "UH", "nu", "q3d", "z_half", "y_mid", "X0", "T"):
        assert k in meta, k
