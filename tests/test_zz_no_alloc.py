"""sw_step_noalloc (the allocation-free step: force partials in the library's fixed device
array) against sw_step + sw_force_reduce, through the C ABI (needs an MI355X).

Two lattices in a row, the second smaller, several steps each: populations and force rows
identical to the allocating entry points; a lattice whose partials exceed the fixed array
is refused before anything is launched."""
import pytest
import torch

from stonewall3d import _native as nat
from stonewall3d import workloads
from stonewall3d.lbm3d import Simulation

pytestmark = pytest.mark.gpu


def _stream():
    return torch.cuda.current_stream().cuda_stream


def _ptr(t):
    return None if t is None else t.data_ptr()


def _compare_steps(w, dtype, steps, seed):
    """`steps` steps from one initial state through both entry points; returns the number
    of differing population words and whether the force records agree."""
    lib = nat.lib()
    sim = Simulation(w.flag, w.uin, dtype=dtype, **w.sim_kwargs()).init(noise=0.02, seed=seed)
    nz = sim.nz
    a = [sim.f[0].clone(), torch.zeros_like(sim.f[0])]        # allocating path
    b = [sim.f[0].clone(), torch.zeros_like(sim.f[0])]        # allocation-free path
    per = lib.sw_partials_per_step(sim.ny, sim.nz)
    partials = torch.zeros(per, dtype=torch.float64, device="cuda")
    fa = torch.zeros((steps, nz, 2), dtype=torch.float64, device="cuda")
    fb = torch.full((steps, nz, 2), float("nan"), dtype=torch.float64, device="cuda")
    s = _stream()
    for t in range(steps):
        common = (sim.dtype, sim.links.data_ptr(), sim.uin.data_ptr(), _ptr(sim.gust), t, None, sim.nx, sim.ny,
                  sim.nz, sim.nu0, sim.cs, sim.x_sponge, sim.nu_sponge, sim.ybc, sim.scale, _ptr(sim.shift))
        nat.check(lib.sw_step(a[0].data_ptr(), a[1].data_ptr(), *common, partials.data_ptr(), 1, s), "sw_step")
        nat.call("sw_force_reduce", partials.data_ptr(), 1, t, None, 1, sim.ny, sim.nz, fa.data_ptr(), s)
        nat.check(lib.sw_step_noalloc(b[0].data_ptr(), b[1].data_ptr(), *common, fb.data_ptr(), s),
                  "sw_step_noalloc")
        a.reverse()
        b.reverse()
    torch.cuda.synchronize()
    differ = int((a[0] != b[0]).sum())
    same_force = torch.equal(fa, fb)
    assert float(fa.abs().max()) > 0
    del sim
    return differ, same_force


@pytest.mark.parametrize("dtype", ["f32", "f16"])
def test_noalloc_matches_allocating_step_twice(gpu, dtype):
    big = workloads.porous_wall(H=10, porosity=0.3, seed=4, n_steps=4)          # 160 x 80 x 40
    small = workloads.porous_wall(H=6, porosity=0.2, seed=5, n_steps=3)         # 96 x 48 x 24
    assert small.cells < big.cells
    for w, steps, seed in ((big, 4, 1), (small, 3, 2)):
        differ, same_force = _compare_steps(w, dtype, steps, seed)
        assert differ == 0, f"{differ} population words differ ({w.name}, {dtype})"
        assert same_force, f"force rows differ ({w.name}, {dtype})"


def test_noalloc_refuses_oversized_partials(gpu):
    lib = nat.lib()
    assert lib.sw_noalloc_max_pairs() == 262144
    dummy = torch.zeros(16, dtype=torch.float64, device="cuda")
    p = dummy.data_ptr()
    nx, ny, nz = 2, 4096, 300                                    # nz * ceil(ny/4) = 307,200 pairs
    rc = lib.sw_step_noalloc(p, p + 64, 0, p, p, None, 0, None, nx, ny, nz, 2e-4, 0.16, nx, 0.05, 0, 1.0, None, p,
                             _stream())
    assert rc == 1003
    assert b"262,144" in lib.sw_last_error()
    torch.cuda.synchronize()
    assert not dummy.any()
