"""Batched and serving forms: each lattice's results are bit-identical alone and in any
mix, graph replays equal eager launches, lattices join mid-run (needs an MI355X)."""
import numpy as np
import pytest
import torch

from stonewall3d import _native as nat
from stonewall3d import workloads
from stonewall3d.autotune import Autotuner
from stonewall3d.batch import Batch, Lattice
from stonewall3d.lbm3d import Simulation

pytestmark = pytest.mark.gpu

CASES = [(6, 0.2, 11, 37), (8, 0.35, 12, 40), (5, 0.0, 13, 29)]     # H, porosity, seed, steps
VARIANT = 7


def wl(H, por, seed, steps):
    return workloads.porous_wall(H=H, porosity=por, seed=seed, n_steps=steps)


def lat_flag(lat):
    """The lattice's cell flags in the reference layout (nx, ny, nz), from its link masks."""
    m = lat.links[:lat.N].cpu().numpy().view(np.uint32).reshape(lat.nz, lat.ny, lat.nx)
    return ((m >> 31) & 1).astype(np.uint8).transpose(2, 1, 0)


def alone(w, dtype, variant, seed, shift=None):
    code = nat.DTYPES[dtype]
    old = nat.lib().sw_get_step_variant(code)
    nat.call("sw_set_step_variant", code, variant)
    try:
        sim = Simulation(w.flag, w.uin, dtype=dtype, shift=shift, **w.sim_kwargs()).init(noise=0.02, seed=seed)
        sim.step(w.n_steps)
        return sim.populations(), sim.force_rows()
    finally:
        nat.call("sw_set_step_variant", code, old)


def alone2(w, dtype, seed, shift=None):
    """The lattice stepped alone with the two-step kernels on the two-step batch's tiles
    (sw_batch_step2_rows()): pairs by sw_step2, an odd last step by sw_step2_one."""
    old = nat.lib().sw_get_step2_rows()
    nat.call("sw_set_step2_rows", nat.lib().sw_batch_step2_rows())
    try:
        sim = Simulation(w.flag, w.uin, dtype=dtype, shift=shift, two_step=True, **w.sim_kwargs())
        sim.init(noise=0.02, seed=seed)
        sim.step(w.n_steps)
        return sim.populations(), sim.force_rows()
    finally:
        nat.call("sw_set_step2_rows", old)


def lattice(w, dtype, seed, K, shift=None):
    return Lattice(w.flag, w.uin, dtype=dtype, ring=K, shift=shift, **w.sim_kwargs()).init(noise=0.02, seed=seed)


@pytest.mark.parametrize("variant", [VARIANT, 8])
@pytest.mark.parametrize("dtype,shift", [("f32", None), ("f16", "inlet"), ("f16", "weights"), ("bf16", "inlet"),
                                         ("bf16", "weights")])
def test_alone_equals_mix(gpu, dtype, shift, variant):
    K = 8
    ws = [wl(*c) for c in CASES]
    ref = [alone(w, dtype, variant, c[2], shift) for w, c in zip(ws, CASES)]
    b = Batch(dtype, max_instances=8, K=K, variant=variant, graph=True, two_step=False)
    lats = [lattice(w, dtype, c[2], K, shift) for w, c in zip(ws, CASES)]
    for lat in lats:
        b.add(lat)
    b.run_to_completion()
    for lat, (p, F) in zip(lats, ref):
        assert lat.t == lat.n_steps
        assert np.array_equal(lat.populations(), p)
        assert np.array_equal(lat.force_rows(), F)
    b.close()


def ragged_workloads():
    """Lattices of very different shapes in one batch: the smallest accepted (2x1x2), odd
    rows with a stone, a partial last 128-cell chunk, and a porous wall; step counts differ
    so the lattices finish in different chunks."""
    from stonewall3d.inflow import gust_table, make_gust, power_law_profile
    from stonewall3d.lattice import flag_from_mask
    out = []
    for k, (dims, steps) in enumerate((((2, 1, 2), 13), ((131, 2, 3), 21), ((257, 3, 4), 17))):
        wall = np.zeros(dims, bool)
        if dims[0] > 8:
            wall[dims[0] // 2, 0, 1:dims[2] - 1] = True
        nz = dims[2]
        out.append(workloads.Workload(f"ragged {dims}", flag_from_mask(wall), power_law_profile(nz, max(1, nz - 2)),
                                      1, 0.05, 2e-4, x_sponge=dims[0] - 10, ybc="periodic" if k % 2 else "freeslip",
                                      gust=gust_table(make_gust(k), steps), n_steps=steps))
    out.append(wl(6, 0.2, 11, 25))
    return out


@pytest.mark.parametrize("dtype", ["f32", "f16", "bf16"])
def test_alone_equals_mix_ragged_default_tiles(gpu, dtype):
    """The library's default tiles (pairs over two wavefronts per row, chunk edges carried)
    in a batch of ragged lattices: each result identical to the lattice run alone."""
    K = 8
    variant = nat.lib().sw_get_step_variant(nat.DTYPES[dtype])
    ws = ragged_workloads()
    seeds = [21, 22, 23, 24]
    ref = [alone(w, dtype, variant, s) for w, s in zip(ws, seeds)]
    b = Batch(dtype, max_instances=8, K=K, variant=variant, graph=True, two_step=False)
    lats = [lattice(w, dtype, s, K) for w, s in zip(ws, seeds)]
    for lat in lats:
        b.add(lat)
    b.run_to_completion()
    for lat, (p, F) in zip(lats, ref):
        assert lat.t == lat.n_steps
        assert np.array_equal(lat.populations(), p)
        assert np.array_equal(lat.force_rows(), F)
    b.close()


@pytest.mark.parametrize("dtype,shift", [("f32", None), ("f16", "inlet"), ("bf16", "inlet")])
def test_two_step_batch_alone_equals_mix(gpu, dtype, shift):
    """The two-step batch (pairs of steps per launch, the odd last step in the tail launch):
    each lattice's populations equal its single steps (the one-step run alone) and its force
    rows equal the two-step run alone on the same tiles; step counts 37 / 40 / 29 end
    inside chunks of 8 with odd and even remainders."""
    K = 8
    ws = [wl(*c) for c in CASES]
    ref1 = [alone(w, dtype, 9, c[2], shift) for w, c in zip(ws, CASES)]
    ref2 = [alone2(w, dtype, c[2], shift) for w, c in zip(ws, CASES)]
    b = Batch(dtype, max_instances=8, K=K, graph=True, two_step=True)
    lats = [lattice(w, dtype, c[2], K, shift) for w, c in zip(ws, CASES)]
    for lat in lats:
        b.add(lat)
    b.run_to_completion()
    for lat, (p1, F1), (p2, F2) in zip(lats, ref1, ref2):
        assert lat.t == lat.n_steps
        pb = lat.populations()
        assert np.array_equal(pb, p2)
        fluid = (lat_flag(lat) == 0)[None]
        assert np.array_equal(np.where(fluid, pb, 0), np.where(fluid, p1, 0))
        assert np.array_equal(lat.force_rows(), F2)
        scale = np.abs(F1).max()
        assert np.allclose(lat.force_rows(), F1, rtol=1e-12, atol=1e-15 * scale)
    b.close()


@pytest.mark.parametrize("dtype", ["f32", "f16"])
def test_two_step_batch_ragged(gpu, dtype):
    """Ragged lattices (2x1x2 up to a porous wall, both side boundaries) in the two-step
    batch: identical to the two-step run alone."""
    K = 8
    ws = ragged_workloads()
    seeds = [21, 22, 23, 24]
    ref = [alone2(w, dtype, s) for w, s in zip(ws, seeds)]
    b = Batch(dtype, max_instances=8, K=K, graph=True, two_step=True)
    lats = [lattice(w, dtype, s, K) for w, s in zip(ws, seeds)]
    for lat in lats:
        b.add(lat)
    b.run_to_completion()
    for lat, (p, F) in zip(lats, ref):
        assert lat.t == lat.n_steps
        assert np.array_equal(lat.populations(), p)
        assert np.array_equal(lat.force_rows(), F)
    b.close()


@pytest.mark.parametrize("two_step", [False, True])
def test_graph_replays_equal_eager(gpu, two_step):
    K = 8
    ws = [wl(*c) for c in CASES]
    outs = []
    for graph in (True, False):
        b = Batch("f32", max_instances=8, K=K, variant=VARIANT, graph=graph, two_step=two_step)
        lats = [lattice(w, "f32", c[2], K) for w, c in zip(ws, CASES)]
        for lat in lats:
            b.add(lat)
        snaps = []
        for _ in range(3):
            b.run(1)                                       # replay, then host work in between
            torch.cuda.synchronize()
            snaps.append([lat.populations() for lat in lats])
        b.run_to_completion()
        outs.append((snaps, [lat.populations() for lat in lats], [lat.force_rows() for lat in lats]))
        b.close()
    (s1, p1, f1), (s2, p2, f2) = outs
    for a, c in zip(s1, s2):
        for x, y in zip(a, c):
            assert np.array_equal(x, y)
    for x, y in zip(p1 + f1, p2 + f2):
        assert np.array_equal(x, y)


def test_batch_with_lattice_beyond_4gib(gpu):
    """A batch holding an FP64 lattice whose population array exceeds 4 GiB (320x320x280:
    4.36 GB) runs on 64-bit addressing for every member; both members stay bit-identical
    to their single-lattice runs."""
    K = 2
    big = workloads.box_wall_flow(n_steps=2, nx=320, ny=320, nz=280, H=40, X0=80, T=12)
    small = wl(6, 0.2, 11, 2)
    assert 19 * big.cells * 8 >= 1 << 32
    ref = [alone(w, "f64", 1, seed) for w, seed in ((big, 5), (small, 11))]
    b = Batch("f64", max_instances=4, K=K, graph=True)
    lats = [lattice(big, "f64", 5, K), lattice(small, "f64", 11, K)]
    for lat in lats:
        b.add(lat)
    assert b.big
    b.run_to_completion()
    for lat, (p, F) in zip(lats, ref):
        assert lat.t == 2
        assert np.array_equal(lat.populations(), p)
        assert np.array_equal(lat.force_rows(), F)
    b.close()


@pytest.mark.parametrize("two_step", [False, True])
def test_join_mid_run(gpu, two_step):
    K = 8
    ws = [wl(*c) for c in CASES]
    ref = [alone2(w, "f32", c[2]) if two_step else alone(w, "f32", VARIANT, c[2]) for w, c in zip(ws, CASES)]
    b = Batch("f32", max_instances=8, K=K, variant=VARIANT, graph=True, two_step=two_step)
    lats = [lattice(w, "f32", c[2], K) for w, c in zip(ws, CASES)]
    b.add(lats[0])
    b.run(2)
    b.add(lats[1])
    b.run(1)
    b.add(lats[2])
    done = []
    while b.members:
        b.run_chunk()
        done += b.retire_finished()
    assert len(done) == 3
    for lat, (p, F) in zip(lats, ref):
        assert np.array_equal(lat.populations(), p)
        assert np.array_equal(lat.force_rows(), F)
    b.close()


def test_server_results_alone_equal_mix(gpu):
    """Served results equal single runs: FP32 requests take the two-step batch (the Batch
    default), FP16 requests the one-step batch with the given tile."""
    from stonewall3d.serving import Request, Server
    reqs = [Request(0, 0.0, 6, 0.2, 21, 33, "f32"), Request(1, 0.0, 5, 0.4, 22, 41, "f16"),
            Request(2, 0.0, 7, 0.1, 23, 26, "f32"), Request(3, 0.0, 6, 0.3, 24, 50, "f16")]
    srv = Server(budget_bytes=4e9, K=8, keep_force=True, variants={"f32": VARIANT, "f16": VARIANT})
    results, rep = srv.serve(reqs, realtime=False)
    srv.close()
    assert rep["requests"] == 4 and rep["p99_latency_s"] >= rep["p50_latency_s"]
    for r in results:
        req = reqs[r.rid]
        w = workloads.porous_wall(H=req.H, porosity=req.porosity, seed=req.seed, n_steps=req.steps)
        _, F = alone2(w, req.dtype, req.seed) if req.dtype == "f32" else alone(w, req.dtype, VARIANT, req.seed)
        assert np.array_equal(r.force, F), f"request {r.rid} differs in the mix"


def test_server_respects_budget(gpu):
    from stonewall3d.serving import Request, Server, lattice_bytes
    reqs = [Request(k, 0.0, 6, 0.2, 30 + k, 24, "f32") for k in range(4)]
    one = lattice_bytes(6, "f32", 24, 8)
    srv = Server(budget_bytes=2.5 * one, K=8)
    peak = []
    orig = srv._batch

    def spy(dtype):
        b = orig(dtype)
        peak.append(len(b.members))
        return b

    srv._batch = spy
    results, rep = srv.serve(reqs, realtime=False)
    srv.close()
    assert rep["requests"] == 4
    assert max(peak) <= 2                 # never more than two resident at once


def test_autotuner_selects_variant(gpu):
    w = wl(8, 0.2, 3, 200)
    sim = Simulation(w.flag, w.uin, **w.sim_kwargs()).init()
    old = nat.lib().sw_get_step_variant(0)
    try:
        res = Autotuner().tune_step(sim, variants=[0, 1, 7], steps=5)
        assert set(res) == {0, 1, 7}
        assert nat.lib().sw_get_step_variant(0) == min(res, key=res.get)
    finally:
        nat.call("sw_set_step_variant", 0, old)
