"""Two fused steps per launch (sw_step2, Simulation(two_step=True)) against the oracle's single
steps, through the C ABI (needs an MI355X).

Bar: populations of fluid cells bit-identical to the oracle in FP32, FP16 and BF16 storage
(the kernel's arithmetic is the single step's; only the intermediate state stays on chip);
force rows within 1e-12 relative (the partials are summed per 64-cell tile instead of per
row-group, a different fixed order).  Cases: stones and a lone stone (general path, wall
force), both side boundaries, the gusting inlet and the sponge, x tiles cut through a stone,
partial x and y tiles, one-row and two-plane lattices (top mirror on the ground's plane), an
odd number of steps (pairs plus a single step, the two partials layouts in one force ring),
and the 640x320x160 headline lattice at full size.
"""
import numpy as np
import pytest
import torch

from oracle import c_oracle as co
from stonewall3d import geometry, workloads
from stonewall3d.inflow import gust_table, make_gust, power_law_profile
from stonewall3d.lattice import flag_from_mask
from stonewall3d.lbm3d import Simulation

pytestmark = pytest.mark.gpu

SCALE = {"f32": 1.0, "f16": 1024.0, "bf16": 1.0}


def stones(nx, ny, nz, H, X0, seed=3, lone=True, T=5):
    wall, _ = geometry.rounded_box_wall(nx, ny, nz, X0, T, H, 0.3, seed, size=(0.4, 0.8))
    if lone and nx > 31 and ny > 4 and nz > 4:
        wall[nx - 18, ny // 2, 3] = True
    return flag_from_mask(wall), power_law_profile(nz, H)


def compare(flag, uin, steps, storage, ybc, shift="inlet", gust=None, seed=11, ring=8, chunks=None):
    """Steps on the GPU with the two-step kernel (in `chunks` calls of step()) and with the
    oracle; returns (differing fluid populations, GPU force rows, oracle force rows)."""
    nx = flag.shape[0]
    xs = nx - 10
    sc = SCALE[storage]
    sim = Simulation(flag, uin, nu0=2e-4, cs=0.16, x_sponge=xs, ybc=ybc, dtype=storage, scale=sc, gust=gust,
                     n_steps=steps, ring=ring, shift=shift, two_step=True)
    sim.init(noise=0.02, seed=seed)
    for n in (chunks or [steps]):
        sim.step(n)
    sh = co.shift_table(uin, shift) if storage in ("f16", "bf16") else None
    f = co.init(flag, uin, storage, noise=0.02, seed=seed, scale=sc, shift=sh)
    F = []
    for t in range(steps):
        g = 1.0 if gust is None else float(gust[t])
        f, Ft = co.step(f, flag, uin, 2e-4, 0.16, xs, 0.05, gust=g, ybc=ybc, storage=storage, scale=sc, shift=sh)
        F.append(Ft)
    g = sim.populations()
    fl = (flag == 0)[None]
    return int(np.where(fl, g != f, False).sum()), sim.force_rows(), np.array(F)


def check(flag, uin, steps, storage, ybc, **kw):
    differ, Fg, Fo = compare(flag, uin, steps, storage, ybc, **kw)
    assert differ == 0, f"{differ} populations differ ({storage}, ybc={ybc}, {flag.shape})"
    assert Fg.shape == Fo.shape
    scale = max(np.abs(Fo).max(), 1e-300)
    assert np.allclose(Fg, Fo, rtol=1e-12, atol=1e-15 * scale)
    return Fo


@pytest.mark.parametrize("storage,shift", [("f32", None), ("f16", "inlet"), ("f16", "weights"), ("bf16", "inlet")])
@pytest.mark.parametrize("ybc", [0, 1])
def test_two_step_small_lattice(gpu, storage, shift, ybc):
    flag, uin = stones(48, 13, 18, 8, 12)
    Fo = check(flag, uin, 12, storage, ybc, shift=shift or "inlet", gust=gust_table(make_gust(5), 12))
    assert np.abs(Fo).max() > 0


@pytest.mark.parametrize("dims,X0", [((128, 29, 14), 60), ((200, 31, 12), 120), ((257, 17, 9), 190),
                                     ((64, 14, 10), 30), ((130, 3, 6), 64)])
@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_two_step_tiles_and_edges(gpu, dims, X0, storage):
    """Several 64-cell x tiles with a wall crossing a tile edge (halo cells next to stones),
    partial last x tiles, several y tiles with a partial last one, both side boundaries."""
    nx, ny, nz = dims
    flag, uin = stones(nx, ny, nz, max(2, nz - 4), X0)
    for ybc in (0, 1):
        check(flag, uin, 8, storage, ybc, gust=gust_table(make_gust(2), 8))


@pytest.mark.parametrize("dims", [(2, 1, 2), (5, 1, 4), (70, 1, 3), (66, 2, 2), (131, 2, 3)])
@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_two_step_degenerate_lattices(gpu, dims, storage):
    """One-row lattices, two planes (the only fluid plane is also the top), rows just past one
    x tile: bit-identical to the oracle in both side boundaries."""
    nx, ny, nz = dims
    wall = np.zeros(dims, bool)
    if nx > 8:
        wall[nx // 2, 0, 1:nz - 1] = True
    flag = flag_from_mask(wall)
    uin = power_law_profile(nz, max(1, nz - 2))
    for ybc in (0, 1):
        check(flag, uin, 6, storage, ybc, gust=gust_table(make_gust(3), 6))


@pytest.mark.parametrize("chunks", [[7], [1, 6], [3, 1, 4], [5, 5]])
def test_two_step_odd_counts_mix_layouts(gpu, chunks):
    """step() with odd counts: pairs plus single steps, so the force ring holds partials of both
    kernels; a ring of 4 forces reductions between them."""
    flag, uin = stones(96, 9, 12, 8, 40)
    steps = sum(chunks)
    check(flag, uin, steps, "f32", 1, gust=gust_table(make_gust(6), steps), ring=4, chunks=chunks)


def test_two_step_equals_single_steps_on_device(gpu):
    """Fluid populations of two-step launches equal the library's own single steps exactly."""
    flag, uin = stones(160, 40, 30, 10, 60)
    kw = dict(nu0=2e-4, x_sponge=130, ybc="periodic", n_steps=10, gust=gust_table(make_gust(9), 10))
    a = Simulation(flag, uin, **kw).init(noise=0.02, seed=4).step(10)
    b = Simulation(flag, uin, two_step=True, **kw).init(noise=0.02, seed=4).step(10)
    fl = torch.from_numpy(flag == 0).cuda().permute(2, 1, 0).reshape(-1)
    pa, pb = a.f[0].view(19, -1), b.f[0].view(19, -1)
    assert torch.equal(pa[:, fl], pb[:, fl])
    assert np.allclose(a.force_rows(), b.force_rows(), rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_two_step_headline_lattice(gpu, storage):
    """The 640x320x160 porous wall (20 % void, gusting inlet): 4 steps in two launches
    bit-identical to the oracle at full size."""
    w = workloads.porous_wall(H=40, porosity=0.2, seed=0, n_steps=4)
    co.set_threads(16)
    sc = SCALE[storage]
    sim = Simulation(w.flag, w.uin, dtype=storage, scale=sc, two_step=True, ring=8, **w.sim_kwargs())
    sim.init(noise=0.02, seed=0)
    sim.step(4)
    sh = co.shift_table(w.uin, "inlet") if storage != "f32" else None
    f = co.init(w.flag, w.uin, storage, noise=0.02, seed=0, scale=sc, shift=sh)
    F = []
    for t in range(4):
        f, Ft = co.step(f, w.flag, w.uin, w.nu0, w.cs, w.x_sponge, w.nu_sponge, gust=float(w.gust[t]), ybc=1,
                        storage=storage, scale=sc, shift=sh)
        F.append(Ft)
    g = sim.populations()
    fl = w.flag == 0
    differ = sum(int(((g[i] != f[i]) & fl).sum()) for i in range(19))
    assert differ == 0
    F = np.array(F)
    assert np.allclose(sim.force_rows(), F, rtol=1e-11, atol=1e-14 * np.abs(F).max())
    del sim
    torch.cuda.empty_cache()


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_two_step_tile_order(gpu, storage):
    """The banded tile order (one contiguous band of tiles per XCD, the default) and the
    plain order give the same populations and force rows; 30 tiles (not a multiple of the 8
    XCDs), odd steps included."""
    from stonewall3d import _native as nat
    flag, uin = stones(330, 60, 12, 8, 60)
    gust = gust_table(make_gust(4), 5)
    out = []
    try:
        for banded in (1, 0):
            nat.call("sw_set_step2_tile_order", banded)
            differ, Fg, Fo = compare(flag, uin, 5, storage, 1, gust=gust)
            assert differ == 0, f"{differ} populations differ (banded={banded})"
            out.append(Fg)
    finally:
        nat.call("sw_set_step2_tile_order", 1)
    assert np.array_equal(out[0], out[1])


class step2_segments:
    """Fixes the two-step kernels' z segments per tile for a `with` block (0: automatic)."""

    def __init__(self, n):
        self.n = n

    def __enter__(self):
        from stonewall3d import _native as nat
        nat.call("sw_set_step2_segments", self.n)

    def __exit__(self, *exc):
        from stonewall3d import _native as nat
        nat.call("sw_set_step2_segments", 0)


@pytest.mark.parametrize("rows", [12, 13])
@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
@pytest.mark.parametrize("ybc", [0, 1])
def test_two_step_z_segments(gpu, rows, storage, ybc):
    """Tiles cut into z segments (each recomputes the first step of the plane below it):
    1 (the uncut kernel), 2, 3, 5 and 16 segments of a 17-plane sweep (down to one plane per
    segment, the last ending at the free top), 7 steps (the single-step launch cut the same
    way), with 12-row tiles (the next plane's loads in flight across each barrier: the
    default for this lattice) and 13-row tiles (loads issued after the barrier).
    Populations bit-identical to the oracle; force rows bit-identical to the uncut sweep."""
    from stonewall3d import _native as nat
    flag, uin = stones(200, 31, 18, 12, 60)
    assert nat.lib().sw_get_step2_segments(200, 31, 18) > 1       # fewer tiles than CUs: cut by default
    gust = gust_table(make_gust(8), 7)
    ref = None
    nat.call("sw_set_step2_rows", rows)
    try:
        for n in (1, 2, 3, 5, 16):
            with step2_segments(n):
                differ, Fg, Fo = compare(flag, uin, 7, storage, ybc, gust=gust)
            assert differ == 0, f"{differ} populations differ with {n} segments ({rows} rows)"
            if ref is None:
                ref = Fg
                assert np.allclose(Fg, Fo, rtol=1e-12, atol=1e-15 * max(np.abs(Fo).max(), 1e-300))
            assert np.array_equal(Fg, ref), f"force rows with {n} segments differ from the uncut sweep ({rows} rows)"
    finally:
        nat.call("sw_set_step2_rows", 0)


@pytest.mark.parametrize("rows", [15, 16])
@pytest.mark.parametrize("dims,X0", [((128, 29, 14), 60), ((257, 17, 9), 190), ((130, 3, 6), 64)])
@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_two_step_uncut_small_lattices(gpu, rows, dims, X0, storage):
    """The uncut kernel (15 or 16 rows, one workgroup per tile) on lattices that by default
    take 12-row tiles cut into segments (fewer tiles than compute units): partial x and y
    tiles, both side boundaries."""
    from stonewall3d import _native as nat
    nx, ny, nz = dims
    flag, uin = stones(nx, ny, nz, max(2, nz - 4), X0)
    nat.call("sw_set_step2_rows", rows)
    try:
        for ybc in (0, 1):
            check(flag, uin, 6, storage, ybc, gust=gust_table(make_gust(2), 6))
    finally:
        nat.call("sw_set_step2_rows", 0)


def test_two_step_reference_lattice_segments(gpu):
    """The reference's default lattice (800x160x321, 169 tiles: cut into segments by default)
    gives the populations and force rows of the uncut sweep after 3 steps (FP32)."""
    from stonewall3d import _native as nat
    w = workloads.reference(H=40, porosity=0.2, seed=7, n_steps=3)
    assert nat.lib().sw_get_step2_segments(800, 160, 321) > 1
    out = []
    for n in (0, 1):
        with step2_segments(n):
            sim = Simulation(w.flag, w.uin, dtype="f32", two_step=True, ring=4, **w.sim_kwargs())
            sim.init(noise=0.02, seed=0)
            sim.step(3)
            out.append((sim.f[0].clone(), sim.force_rows()))
            del sim
    fl = torch.from_numpy(w.flag == 0).cuda().permute(2, 1, 0).reshape(-1)
    assert torch.equal(out[0][0].view(19, -1)[:, fl], out[1][0].view(19, -1)[:, fl])
    assert np.array_equal(out[0][1], out[1][1])
    del out
    torch.cuda.empty_cache()


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_two_step_launch_shapes_agree_on_headline_lattice(gpu, storage):
    """The 640x320x160 porous wall (250 tiles of 15 rows: uncut by default) stepped 4 times with
    12-row tiles (next-plane loads in flight) cut into 4 segments, uncut 12-row tiles (two
    rounds of workgroups) and 13-row tiles cut into 5 segments: fluid populations identical
    to the default launch's, force rows equal to rounding (another per-tile summation)."""
    from stonewall3d import _native as nat
    w = workloads.porous_wall(H=40, porosity=0.2, seed=0, n_steps=4)
    assert nat.lib().sw_get_step2_segments(640, 320, 160) == 1
    fl = torch.from_numpy(w.flag == 0).cuda().permute(2, 1, 0).reshape(-1)
    ref = None
    try:
        for rows, segs in ((0, 0), (12, 4), (12, 1), (13, 5)):
            nat.call("sw_set_step2_rows", rows)
            with step2_segments(segs):
                sim = Simulation(w.flag, w.uin, dtype=storage, two_step=True, ring=4, **w.sim_kwargs())
                sim.init(noise=0.02, seed=0)
                sim.step(4)
                f, F = sim.f[0].view(19, -1)[:, fl].clone(), sim.force_rows()
                del sim
            if ref is None:
                ref = (f, F)
                continue
            assert torch.equal(ref[0], f), f"populations differ with {rows} rows, {segs} segments"
            assert np.allclose(ref[1], F, rtol=1e-12, atol=1e-15 * np.abs(ref[1]).max())
            del f
    finally:
        nat.call("sw_set_step2_rows", 0)
    del ref
    torch.cuda.empty_cache()
