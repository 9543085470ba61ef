"""The step's accumulate epilogue (sw_step_accumulate, Simulation.step(accumulate=True))
against the separate macro pass (sw_macro after sw_step, as k_macro + k_accum follow k_step
in lbm3d.py:363-370), through the C ABI (needs an MI355X).

Bar: the accumulator bit-identical and the populations unchanged, in every storage form, on
the one-cell and pair tiles, with stones, odd rows, a first chunk free of solids (the inlet
cell written inside a pair) and a solid cell in the inlet column; and with two-step
launches before the sampled step."""
import numpy as np
import pytest
import torch

from stonewall3d import _native as nat
from stonewall3d import geometry
from stonewall3d.inflow import gust_table, make_gust, power_law_profile
from stonewall3d.lattice import flag_from_mask
from stonewall3d.lbm3d import Simulation

pytestmark = pytest.mark.gpu

SCALE = {"f32": 1.0, "f16": 1024.0, "bf16": 1.0, "f64": 1.0}


def lattice(nx, X0, inlet_stone=False, ny=9, nz=14):
    wall, _ = geometry.rounded_box_wall(nx, ny, nz, X0, 5, 8, 0.3, 3, size=(0.4, 0.8))
    if inlet_stone:
        wall[0:2, 2:5, 1:4] = True                     # solid cells in the inlet column
    return flag_from_mask(wall), power_law_profile(nz, 8)


def pair(flag, uin, storage, steps, two_step=False):
    kw = dict(nu0=2e-4, x_sponge=flag.shape[0] - 10, ybc="freeslip", dtype=storage, scale=SCALE[storage],
              n_steps=steps, ring=6, gust=gust_table(make_gust(4), steps))
    a = Simulation(flag, uin, **kw).init(noise=0.02, seed=3)
    b = Simulation(flag, uin, two_step=two_step, **kw).init(noise=0.02, seed=3)
    return a, b


@pytest.mark.parametrize("variant", [1, 8, 9])
@pytest.mark.parametrize("storage", ["f32", "f16", "bf16", "f64"])
def test_fused_accumulate_equals_macro_pass(gpu, storage, variant):
    code = nat.DTYPES[storage]
    old = nat.lib().sw_get_step_variant(code)
    nat.call("sw_set_step_variant", code, variant)
    try:
        for nx, X0, inlet_stone in ((48, 12, False), (77, 30, True), (384, 200, False), (130, 60, True)):
            flag, uin = lattice(nx, X0, inlet_stone)
            a, b = pair(flag, uin, storage, 12)
            for t in range(12):
                a.step(1)
                if t % 3 == 0:
                    a.accumulate()
                b.step(1, accumulate=(t % 3 == 0))
            assert a.n_acc == b.n_acc == 4
            assert torch.equal(a.Uacc, b.Uacc), (storage, variant, nx)
            assert np.array_equal(a.populations(), b.populations())
            assert np.array_equal(a.force_rows(), b.force_rows())
    finally:
        nat.call("sw_set_step_variant", code, old)


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_fused_accumulate_after_two_step_launches(gpu, storage):
    """step(n, accumulate=True) with two steps per launch: pairs, then the sampled step alone."""
    flag, uin = lattice(200, 90)
    a, b = pair(flag, uin, storage, 20, two_step=True)
    for n in (5, 5, 4, 6):
        a.step(n)
        a.accumulate()
        b.step(n, accumulate=True)
    assert torch.equal(a.Uacc, b.Uacc)
    fl = torch.from_numpy(flag == 0).cuda().permute(2, 1, 0).reshape(-1)
    assert torch.equal(a.f[0].view(19, -1)[:, fl], b.f[0].view(19, -1)[:, fl])
    assert np.allclose(a.force_rows(), b.force_rows(), rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("rows,segs", [(0, 0), (15, 0), (16, 0), (13, 3), (12, 1), (12, 4)])
def test_two_step_sampled_pairs_equal_macro_pass(gpu, rows, segs):
    """FP32, two steps per launch, with samples at the first and at the second step of pairs
    (sw_step2_accumulate), two adjacent samples (one of them in a launch of its own) and the
    last step: the accumulator bit-identical to sw_macro after every sampled single step,
    the populations identical and the force rows equal to rounding; every launch shape (the
    automatic one, 15 and 16 rows uncut, 13 rows cut into 3 segments, 12 rows uncut and cut)."""
    flag, uin = lattice(200, 90, inlet_stone=True, ny=31, nz=18)
    nat.call("sw_set_step2_rows", rows)
    nat.call("sw_set_step2_segments", segs)
    try:
        a, b = pair(flag, uin, "f32", 24, two_step=True)
        sampled = {0, 3, 4, 9, 12, 17, 20, 21, 23}
        for t in range(24):
            a.step(1)
            if t in sampled:
                a.accumulate()
        b.step(10, samples=[k for k in sampled if k < 10])
        b.step(14, samples=[k - 10 for k in sampled if k >= 10])
        assert a.n_acc == b.n_acc == len(sampled)
        assert torch.equal(a.Uacc, b.Uacc)
        fl = torch.from_numpy(flag == 0).cuda().permute(2, 1, 0).reshape(-1)
        assert torch.equal(a.f[0].view(19, -1)[:, fl], b.f[0].view(19, -1)[:, fl])
        assert np.allclose(a.force_rows(), b.force_rows(), rtol=1e-12, atol=1e-15)
    finally:
        nat.call("sw_set_step2_rows", 0)
        nat.call("sw_set_step2_segments", 0)


def test_two_step_accumulate_argument_errors(gpu):
    """The accumulating pair takes FP32 only, a non-null accumulator and which = 1 or 2."""
    flag, uin = lattice(48, 12)
    _, b = pair(flag, uin, "f16", 4, two_step=True)
    with pytest.raises(ValueError):
        b.step(2, samples=[2])
    acc = torch.zeros(4 * b.N, dtype=torch.float32, device="cuda")
    args = (b.f[0].data_ptr(), b.f[1].data_ptr(), b.dtype, b.links.data_ptr(), b.uin.data_ptr(), None, 0, None,
            b.nx, b.ny, b.nz, b.nu0, b.cs, b.x_sponge, b.nu_sponge, b.ybc, b.scale, b.shift.data_ptr(),
            b.partials2.data_ptr(), b.ring)
    assert nat.lib().sw_step2_accumulate(*args, acc.data_ptr(), 1, None) == 1001          # FP16
    assert b"FP32" in nat.lib().sw_last_error()
    # FP16 two-step runs keep sampled steps alone and stay correct
    b.step(4, samples=[1, 2])
    assert b.n_acc == 2
