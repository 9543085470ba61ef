"""HIP kernels vs the CPU oracle, through the C ABI (needs an MI355X).

Bar: FP32 / FP16 / BF16 storage bit-identical populations to the oracle (both compiled
without FMA contraction, same operand order; the 16-bit forms with both storage offsets,
the per-height inlet equilibrium and the weights); force rows (FP64 sums in different fixed
orders) within 1e-12 relative; FP64 within 1e-14.  The workload-level bar (populations
within 1e-5 relative after 1000 steps, mean Cd within 0.5 %) is asserted on the 160x80x40
box-wall lattice in both y-boundary modes.
"""
import os

import numpy as np
import pytest
import torch

from oracle import c_oracle as co
from stonewall3d import geometry, workloads
from stonewall3d.inflow import power_law_profile, gust_table, make_gust
from stonewall3d.lattice import flag_from_mask
from stonewall3d.lbm3d import Simulation

pytestmark = pytest.mark.gpu

SCALE = {"f32": 1.0, "f16": 1024.0, "bf16": 1.0, "f64": 1.0}


def small_case(nx=48, ny=13, nz=18, H=8, seed=3, X0=12, lone=True):
    wall, _ = geometry.rounded_box_wall(nx, ny, nz, X0, 5, H, 0.3, seed, size=(0.4, 0.8))
    if lone:
        wall[30, 4, 3] = True                            # a lone stone
    flag = flag_from_mask(wall)
    uin = power_law_profile(nz, H)
    return flag, uin


def fluid_mask(flag):
    return (flag == 0)[None]


def oracle_shift(uin, storage, shift):
    """The oracle's offset table for a storage form (None: FP32 / FP64, or the weights)."""
    if storage not in ("f16", "bf16") or shift == "weights":
        return None
    return co.shift_table(uin, shift or "inlet")


def run_both(flag, uin, steps, storage, ybc, gust=None, noise=0.02, seed=11, nu0=2e-4, x_sponge=None, shift=None):
    nx = flag.shape[0]
    xs = nx - 10 if x_sponge is None else x_sponge
    sc = SCALE[storage]
    sim = Simulation(flag, uin, nu0=nu0, cs=0.16, x_sponge=xs, ybc=ybc, dtype=storage, scale=sc, gust=gust,
                     n_steps=steps, ring=7, shift=shift)
    sim.init(noise=noise, seed=seed)
    sh = oracle_shift(uin, storage, shift)
    f = co.init(flag, uin, storage, noise=noise, seed=seed, scale=sc, shift=sh)
    g0 = sim.populations()
    assert np.array_equal(np.where(fluid_mask(flag), g0, 0), np.where(fluid_mask(flag), f, 0)), "init differs"
    forces = []
    for t in range(steps):
        g = 1.0 if gust is None else float(gust[t])
        f, F = co.step(f, flag, uin, nu0, 0.16, xs, 0.05, gust=g, ybc=ybc, storage=storage, scale=sc, shift=sh)
        forces.append(F)
    sim.step(steps)
    return sim, f, np.array(forces)


@pytest.mark.parametrize("storage,shift", [("f32", None), ("f16", "inlet"), ("f16", "weights"), ("bf16", "inlet"),
                                           ("bf16", "weights")])
@pytest.mark.parametrize("ybc", [0, 1])
def test_step_bit_identical(gpu, storage, shift, ybc):
    flag, uin = small_case()
    gust = gust_table(make_gust(5), 40)
    sim, f, F = run_both(flag, uin, 40, storage, ybc, gust=gust, shift=shift)
    g = sim.populations()
    fl = fluid_mask(flag)
    diff = np.where(fl, g != f, False)
    assert not diff.any(), f"{diff.sum()} populations differ ({storage}, {shift}, ybc={ybc})"
    Fg = sim.force_rows()
    assert Fg.shape == F.shape
    assert np.abs(F).max() > 0
    assert np.allclose(Fg, F, rtol=1e-12, atol=1e-15 * np.abs(F).max())


def test_shift_table_bit_identical(gpu):
    """sw_shift_table (both kinds) equals the oracle's table bit for bit."""
    from stonewall3d import _native as nat
    uin = power_law_profile(37, 9)
    d = torch.from_numpy(uin).cuda()
    for kind in ("weights", "inlet"):
        t = torch.empty((37, 4), dtype=torch.float32, device="cuda")
        nat.call("sw_shift_table", d.data_ptr(), 37, nat.SHIFTS[kind], t.data_ptr(),
                 torch.cuda.current_stream().cuda_stream)
        assert np.array_equal(t.cpu().numpy().view(np.uint32), co.shift_table(uin, kind).view(np.uint32)), kind


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_macro_16bit_inlet_offsets_bit_identical(gpu, storage):
    flag, uin = small_case()
    sim, f, _ = run_both(flag, uin, 6, storage, 1, shift="inlet")
    U = sim.macro()
    Uo = co.macro(f, flag, storage, SCALE[storage], shift=co.shift_table(uin, "inlet"))
    assert np.array_equal(U, Uo)
    # the device checksum decodes with the same offsets
    v = co.decode(f, storage, SCALE[storage], shift=co.shift_table(uin, "inlet")).astype(np.float64)
    fl = fluid_mask(flag)
    s0 = np.where(fl, v, 0).sum()
    assert sim.checksum()[0] == pytest.approx(s0, rel=1e-12)


def test_division_shortcut_rounds_like_ieee(gpu):
    """The step divides with a Newton-refined hardware reciprocal and an FMA-corrected
    quotient; the library's self-test compares them with IEEE division on the device: every
    normal divisor's reciprocal (all 2^32 bit patterns) and 2^32 hashed quotients."""
    from stonewall3d import _native as nat
    counts = torch.zeros(2, dtype=torch.int64, device="cuda")
    nat.call("sw_selftest_division", counts.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    rcp_bad, div_bad = counts.tolist()
    assert rcp_bad == 0 and div_bad == 0, (rcp_bad, div_bad)


def test_fp64_matches_oracle(gpu):
    flag, uin = small_case()
    sim, f, F = run_both(flag, uin, 25, "f64", 1, gust=gust_table(make_gust(2), 25))
    g = sim.populations()
    fl = fluid_mask(flag)
    rel = np.abs(np.where(fl, g - f, 0)).max() / np.abs(f).max()
    print("fp64 max rel diff", rel, "bitwise", not np.where(fl, g != f, False).any())
    assert rel < 1e-14
    assert np.allclose(sim.force_rows(), F, rtol=1e-12, atol=1e-16)


def test_fp64_beyond_4gib_matches_oracle(gpu):
    """A 320x320x280 FP64 lattice: 19 * N * 8 B = 4.36 GB per population array, past the
    4 GiB reach of 32-bit buffer offsets, so the step takes the 64-bit-pointer kernel.
    Two steps against the FP64 oracle within 1e-14, force rows within 1e-12."""
    w = workloads.box_wall_flow(n_steps=2, nx=320, ny=320, nz=280, H=40, X0=80, T=12)
    assert 19 * w.cells * 8 >= 1 << 32
    sim = Simulation(w.flag, w.uin, dtype="f64", **w.sim_kwargs())
    sim.init(noise=0.02, seed=5)
    f = co.init(w.flag, w.uin, "f64", noise=0.02, seed=5)
    yb = 1 if w.ybc == "freeslip" else 0
    forces = []
    for t in range(2):
        f, F = co.step(f, w.flag, w.uin, w.nu0, w.cs, w.x_sponge, w.nu_sponge, ybc=yb, storage="f64")
        forces.append(F)
    sim.step(2)
    g = sim.populations()
    fl = (w.flag == 0)[None]
    np.subtract(g, f, out=g)
    np.abs(g, out=g)
    g *= fl
    rel = g.max() / np.abs(f).max()
    print("fp64 > 4 GiB: max rel diff", rel)
    assert rel < 1e-14
    assert np.allclose(sim.force_rows(), np.array(forces), rtol=1e-12, atol=1e-16)


@pytest.mark.parametrize("dims", [(2, 1, 2), (3, 1, 3), (2, 2, 5), (5, 1, 4), (8, 1, 3), (130, 1, 3), (131, 2, 3),
                                  (132, 2, 3), (257, 3, 4)])
@pytest.mark.parametrize("storage", ["f32", "f16", "bf16", "f64"])
def test_degenerate_and_ragged_lattices(gpu, dims, storage):
    """The smallest lattices the ABI accepts (nx = 2, ny = 1, nz = 2: one inlet and one
    outlet cell above the ground), a single y row (periodic and mirrored onto itself), odd
    row lengths (the pair tiles' in-kernel fallback), rows of 8 and 132 cells (16-bit
    quads), a partial last 128-cell chunk and a stone in a one-row lattice: every tile
    variant the storage form maps to, both y boundaries, 6 steps bit-identical to the
    oracle (FP64 within 1e-14)."""
    from stonewall3d import _native as nat
    nx, ny, nz = dims
    wall = np.zeros(dims, bool)
    if nx > 8:
        wall[nx // 2, 0, 1:nz - 1] = True                 # a stone column in the first y row
    flag = flag_from_mask(wall)
    uin = power_law_profile(nz, max(1, nz - 2))
    old = nat.lib().sw_get_step_variant(nat.DTYPES[storage])
    try:
        for variant in ([1] if storage == "f64" else [0, 7, 8, 9, 12]):
            nat.call("sw_set_step_variant", nat.DTYPES[storage], variant)
            for ybc in (0, 1):
                sim, f, F = run_both(flag, uin, 6, storage, ybc, gust=gust_table(make_gust(3), 6))
                g = sim.populations()
                fl = fluid_mask(flag)
                if storage == "f64":
                    assert np.abs(np.where(fl, g - f, 0)).max() <= 1e-14 * np.abs(f).max()
                else:
                    assert not np.where(fl, g != f, False).any(), (dims, storage, variant, ybc)
                assert np.allclose(sim.force_rows(), F, rtol=1e-12, atol=1e-15 * max(np.abs(F).max(), 1e-300))
    finally:
        nat.call("sw_set_step_variant", nat.DTYPES[storage], old)


@pytest.mark.parametrize("dims,wide", [((608, 320, 288), False), ((640, 320, 280), True)])
def test_fp32_at_the_4gib_offset_limit(gpu, dims, wide):
    """FP32 on both sides of the 4 GiB reach of 32-bit buffer offsets: 608x320x288 (19 * N *
    4 B = 4.26 GB, the raw-buffer pair tile with row offsets close to 2^32) and 640x320x280
    (4.36 GB, the 64-bit-pointer kernel).  Two steps bit-identical to the FP32 oracle."""
    nx, ny, nz = dims
    w = workloads.box_wall_flow(n_steps=2, nx=nx, ny=ny, nz=nz, H=40, X0=80, T=12)
    assert (19 * w.cells * 4 >= 1 << 32) == wide
    print(f"fp32 {nx}x{ny}x{nz}: {19 * w.cells * 4 / 1e9:.2f} GB per population array", flush=True)
    sim = Simulation(w.flag, w.uin, dtype="f32", **w.sim_kwargs())
    sim.init(noise=0.02, seed=6)
    f = co.init(w.flag, w.uin, "f32", noise=0.02, seed=6)
    yb = 1 if w.ybc == "freeslip" else 0
    forces = []
    for t in range(2):
        f, F = co.step(f, w.flag, w.uin, w.nu0, w.cs, w.x_sponge, w.nu_sponge, ybc=yb)
        forces.append(F)
    sim.step(2)
    g = sim.populations()
    fl = (w.flag == 0)[None]
    differ = 0
    for i in range(19):                                  # one direction at a time: bounded host memory
        differ += int(((g[i] != f[i]) & fl[0]).sum())
    assert differ == 0
    assert np.allclose(sim.force_rows(), np.array(forces), rtol=1e-12, atol=1e-15 * np.abs(forces).max())


def test_macro_and_accumulate_bit_identical(gpu):
    flag, uin = small_case()
    sim, f, _ = run_both(flag, uin, 5, "f32", 0)
    U = sim.macro()
    Uo = co.macro(f, flag)
    assert np.array_equal(U, Uo)
    sim.accumulate()
    sim.accumulate()
    acc = np.zeros_like(Uo)
    co.macro(f, flag, acc=acc)
    co.macro(f, flag, acc=acc)
    assert np.array_equal(sim.mean_field() * 2, acc)


def test_layout_round_trip(gpu):
    flag, uin = small_case()
    sim = Simulation(flag, uin, nu0=2e-4, n_steps=4)
    sim.init(noise=0.02, seed=1)
    sim.step(3)
    p = sim.populations()
    sim2 = Simulation(flag, uin, nu0=2e-4, n_steps=4)
    sim2.load_populations(p)
    assert np.array_equal(sim2.populations(), p)


def test_prepare_rejects_fluid_ground(gpu):
    from stonewall3d._native import SwError
    flag = np.zeros((8, 4, 6), np.uint8)
    with pytest.raises(SwError) as e:
        Simulation(flag, np.zeros(6, np.float32), nu0=1e-3)
    assert e.value.code == 1002


def test_deterministic_and_ring_independent(gpu):
    flag, uin = small_case()
    out = []
    for ring in (1, 3, 64):
        sim = Simulation(flag, uin, nu0=2e-4, x_sponge=38, ybc="freeslip", n_steps=30, ring=ring,
                         gust=gust_table(make_gust(1), 30))
        sim.init(noise=0.02, seed=4)
        sim.step(30)
        out.append((sim.populations(), sim.force_rows()))
    for p, F in out[1:]:
        assert np.array_equal(p, out[0][0])
        assert np.array_equal(F, out[0][1])


@pytest.mark.parametrize("ybc", ["freeslip", "periodic"])
def test_box_wall_1000_steps(gpu, ybc):
    """The 160x80x40 box-wall workload, 1000 steps: populations within 1e-5 relative of
    the oracle (bit-identical expected) and mean Cd within 0.5 %."""
    w = workloads.box_wall_flow(n_steps=1000, ybc=ybc)
    flag, uin = w.flag, w.uin
    sim = Simulation(flag, uin, **w.sim_kwargs())
    sim.init()
    f = co.init(flag, uin, "f32")
    F = []
    yb = 1 if ybc == "freeslip" else 0
    for t in range(1000):
        f, Ft = co.step(f, flag, uin, w.nu0, w.cs, w.x_sponge, w.nu_sponge, ybc=yb)
        F.append(Ft)
    sim.step(1000)
    g = sim.populations()
    fl = fluid_mask(flag)
    rel = np.abs(np.where(fl, g - f, 0)) / np.abs(f).max()
    n_diff = int(np.where(fl, g != f, False).sum())
    print(f"box wall {ybc}: max rel diff {rel.max():.3e}, differing populations {n_diff}")
    assert rel.max() <= 1e-5
    cd_o = np.array(F)[:, :, 0].sum(1) / (0.5 * w.UH ** 2 * w.H * flag.shape[1])
    cd_g = sim.coefficients(w.UH, w.H)["Cd"]
    assert abs(cd_g[100:].mean() - cd_o[100:].mean()) <= 0.005 * abs(cd_o[100:].mean())
    assert n_diff == 0


def test_headline_lattice_first_steps_bit_identical(gpu):
    """640x320x160 porous wall (20 % void, gusting inlet): 3 steps bit-identical to the
    oracle at full size, and two GPU runs identical."""
    w = workloads.porous_wall(H=40, porosity=0.2, seed=0, n_steps=3)
    co.set_threads(16)
    sims = []
    for _ in range(2):
        sim = Simulation(w.flag, w.uin, **w.sim_kwargs())
        sim.init(noise=0.02, seed=0)
        sim.step(3)
        sims.append(sim)
    f = co.init(w.flag, w.uin, "f32", noise=0.02, seed=0)
    yb = 1
    F = []
    for t in range(3):
        f, Ft = co.step(f, w.flag, w.uin, w.nu0, w.cs, w.x_sponge, w.nu_sponge, gust=float(w.gust[t]), ybc=yb)
        F.append(Ft)
    g = sims[0].populations()
    fl = fluid_mask(w.flag)
    assert not np.where(fl, g != f, False).any()
    assert np.array_equal(g, sims[1].populations())
    Fg = sims[0].force_rows()
    assert np.allclose(Fg, np.array(F), rtol=1e-11, atol=1e-14)
    assert np.array_equal(Fg, sims[1].force_rows())
    del sims
    torch.cuda.empty_cache()


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_reference_default_case_full_size(gpu, storage):
    """The reference's own default lattice (lbm3d.py at H = 40: 800x160x321, its stone
    generator at 20 %, seed 7, periodic y, no gust; odd nz): 2 steps bit-identical to the
    oracle at full size."""
    w = workloads.reference(H=40, porosity=0.2, seed=7, n_steps=2)
    assert w.shape == (800, 160, 321)
    co.set_threads(16)
    sc = SCALE[storage]
    sim = Simulation(w.flag, w.uin, dtype=storage, scale=sc, **w.sim_kwargs())
    sim.init(noise=0.02, seed=7)
    sim.step(2)
    sh = oracle_shift(w.uin, storage, None)
    f = co.init(w.flag, w.uin, storage, noise=0.02, seed=7, scale=sc, shift=sh)
    yb = 1 if w.ybc == "freeslip" else 0
    F = []
    for t in range(2):
        f, Ft = co.step(f, w.flag, w.uin, w.nu0, w.cs, w.x_sponge, w.nu_sponge, ybc=yb, storage=storage, scale=sc,
                        shift=sh)
        F.append(Ft)
    g = sim.populations()
    fl = (w.flag == 0)
    differ = sum(int(((g[i] != f[i]) & fl).sum()) for i in range(19))
    assert differ == 0
    assert np.allclose(sim.force_rows(), np.array(F), rtol=1e-11, atol=1e-14 * np.abs(F).max())
    del sim
    torch.cuda.empty_cache()


@pytest.mark.parametrize("dtype", ["f32", "f16", "bf16", "f64"])
def test_resume_is_bitwise_equal_to_uninterrupted(gpu, tmp_path, dtype):
    """Run 30 steps straight vs 12 steps, checkpoint, a fresh Simulation resumes and runs
    18 more: populations, force rows and the accumulator agree bit for bit (the
    reference's claim, stonewall3d/README.md:44); in every storage form (the 16-bit ones
    with the inlet offsets)."""
    from stonewall3d import checkpoint
    flag, uin = small_case()
    kw = dict(nu0=2e-4, x_sponge=38, ybc="periodic", n_steps=30, ring=5, gust=gust_table(make_gust(3), 30),
              dtype=dtype)
    sig = dict(case="resume")
    a = Simulation(flag, uin, **kw).init(noise=0.02, seed=2)
    for t in range(30):
        a.step(1)
        if t % 5 == 0:
            a.accumulate()
    b = Simulation(flag, uin, **kw).init(noise=0.02, seed=2)
    for t in range(12):
        b.step(1)
        if t % 5 == 0:
            b.accumulate()
    p = str(tmp_path / "r.ckpt")
    checkpoint.save(b, p, 12, b.n_acc, sig)
    del b
    c = Simulation(flag, uin, **kw)
    t0, n_acc, _ = checkpoint.load(c, p, sig)
    assert t0 == 12
    for t in range(12, 30):
        c.step(1)
        if t % 5 == 0:
            c.accumulate()
    assert np.array_equal(a.populations(), c.populations())
    assert np.array_equal(a.force_rows(), c.force_rows())
    assert np.array_equal(a.accumulator(), c.accumulator()) and a.n_acc == c.n_acc


def test_cli_short_run(gpu, tmp_path):
    """lbm3d.main with the reference's flags: 60 steps at H=8, checkpoint, resume, .npz keys."""
    from stonewall3d import lbm3d
    out = str(tmp_path)
    args = ["--case", "q", "--porosity", "0.2", "--H", "8", "--t-spinup", "0.3", "--t-avg", "0.5",
            "--out", out, "--ckpt-every", "25", "--max-steps", "60"]
    lbm3d.main(args)
    d = np.load(os.path.join(out, "q.npz"))
    for k in ("Cd", "Cl", "Cm", "fx_profile", "mean_xz_mid", "mean_xz_spanavg", "mean_xy_half", "mean3d_ds2",
              "snaps_xz", "snaps_xy", "wall_xz", "wall_xy", "wall_front", "wall_top", "wall_ds2", "snaps_q", "meta"):
        assert k in d.files, k
    assert d["Cd"].shape == (60,) and np.isfinite(d["Cd"]).all()
    assert not os.path.isdir(os.path.join(out, "q.ckpt"))


@pytest.mark.parametrize("variant", range(14))
@pytest.mark.parametrize("storage,nx", [("f32", 200), ("f16", 200), ("bf16", 200), ("f16", 256), ("bf16", 77),
                                        ("f32", 256), ("f32", 77), ("f16", 202)])
def test_every_tile_variant_bit_identical(gpu, variant, storage, nx):
    """Every kernel variant the autotuner may pick is bit-identical to the oracle (row
    lengths with a partial last chunk, whole 128-cell chunks, an odd length, and an even
    length not divisible by 4, where the quad tiles fall back)."""
    from stonewall3d import _native as nat
    old = nat.lib().sw_get_step_variant(nat.DTYPES[storage])
    nat.call("sw_set_step_variant", nat.DTYPES[storage], variant)
    try:
        flag, uin = small_case(nx=nx, ny=9, nz=14)
        sim, f, F = run_both(flag, uin, 12, storage, 1, gust=gust_table(make_gust(7), 12))
        g = sim.populations()
        assert not np.where(fluid_mask(flag), g != f, False).any()
        assert np.allclose(sim.force_rows(), F, rtol=1e-12, atol=1e-15 * np.abs(F).max())
    finally:
        nat.call("sw_set_step_variant", nat.DTYPES[storage], old)


@pytest.mark.parametrize("variant", [8, 9, 10, 11, 12, 13])
@pytest.mark.parametrize("storage", ["f16", "bf16", "f32"])
def test_pair_path_inlet_chunk(gpu, variant, storage):
    """Pair and quad tiles when the first chunk is free of solids: the inlet cell is written
    inside the packed 32-bit (16-bit pairs) or 64-bit (FP32 pairs, 16-bit quads) store."""
    from stonewall3d import _native as nat
    old = nat.lib().sw_get_step_variant(nat.DTYPES[storage])
    nat.call("sw_set_step_variant", nat.DTYPES[storage], variant)
    try:
        flag, uin = small_case(nx=384, ny=7, nz=12, X0=200, lone=False)
        sim, f, F = run_both(flag, uin, 10, storage, 1, gust=gust_table(make_gust(4), 10))
        assert not np.where(fluid_mask(flag), sim.populations() != f, False).any()
        assert np.allclose(sim.force_rows(), F, rtol=1e-12, atol=1e-15 * np.abs(F).max())
    finally:
        nat.call("sw_set_step_variant", nat.DTYPES[storage], old)


@pytest.mark.parametrize("variant", [8, 9, 10])
@pytest.mark.parametrize("storage", ["f16", "bf16", "f32"])
def test_pair_edge_carry_across_chunks(gpu, variant, storage):
    """Pair tiles take the left chunk edge of the c_x = +1 links from the previous fast
    chunk.  Rows of five 128-cell chunks with the wall in chunk 1: within the wall's height
    the carry restarts after the special chunk (chunk 2 loads its edge, chunks 3 and 4
    carry), above it every chunk after the first carries; with rows split over 2 or 4
    wavefronts each range starts with a load."""
    from stonewall3d import _native as nat
    old = nat.lib().sw_get_step_variant(nat.DTYPES[storage])
    nat.call("sw_set_step_variant", nat.DTYPES[storage], variant)
    try:
        flag, uin = small_case(nx=640, ny=5, nz=12, X0=200, lone=False)
        sim, f, F = run_both(flag, uin, 10, storage, 1, gust=gust_table(make_gust(5), 10))
        assert not np.where(fluid_mask(flag), sim.populations() != f, False).any()
        assert np.allclose(sim.force_rows(), F, rtol=1e-12, atol=1e-15 * np.abs(F).max())
    finally:
        nat.call("sw_set_step_variant", nat.DTYPES[storage], old)


def test_q_frame_bit_identical(gpu):
    """Max-pooled Q-criterion frames (k_qcoarse, lbm3d.py:203-238) equal the oracle's on the
    same macroscopic field."""
    w = workloads.reference(H=10, porosity=0.2, seed=7, n_steps=40)
    sim = Simulation(w.flag, w.uin, **w.sim_kwargs()).init(noise=0.02, seed=7)
    sim.step(40)
    U = sim.macro()
    nx, ny, nz = w.shape
    ds, x0 = 1, max(1, w.info["X0"] - 5)
    dims = (min(90, nx - 1 - x0) // ds, ny // ds, min(25, nz - 1) // ds)
    sim_q = torch.empty(int(np.prod(dims)), dtype=torch.float32, device="cuda")
    from stonewall3d import _native as nat
    nat.call("sw_qcoarse", sim.U.data_ptr(), sim.links.data_ptr(), nx, ny, nz, x0, ds, *dims, sim_q.data_ptr(),
             torch.cuda.current_stream().cuda_stream)
    q_gpu = sim_q.cpu().numpy().reshape(dims)
    q_ref = co.qcoarse(U, w.flag, x0, ds, dims)
    assert np.array_equal(q_gpu, q_ref)
    assert q_ref.max() > 0
    frame = sim.q_frame(x0, 2, (dims[0] // 2, dims[1] // 2, dims[2] // 2), (10 / w.UH) ** 2)
    assert frame.dtype == np.uint8 and frame.max() > 0
