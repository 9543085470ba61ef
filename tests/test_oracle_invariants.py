"""Pins of the CPU oracle by closed forms, invariants and an independent restatement.

Each check is chosen so that a plausible slip in the oracle (a wrong sign, a dropped term,
a transposed index, a wrong opposite/mirror entry) fails it.
"""
import numpy as np
import pytest

from oracle import c_oracle as co
from oracle import np_lbm as nl
from stonewall3d import lattice

NU0 = 2e-4


def small_flag(nx=12, ny=6, nz=7, wall=True):
    flag = np.zeros((nx, ny, nz), np.uint8)
    if wall:
        flag[5:7, :, 1:4] = 1
        flag[8, 2, 2] = 1
    flag[:, :, 0] = 2
    return flag


def profile(nz, H=3.0, u=0.05):
    z = np.arange(nz) - 0.5
    p = (u * np.clip(z / H, 0, 1) ** (1 / 7)).astype(np.float32)
    p[0] = 0
    return p


# ------------------------------------------------------------------ tables
def test_tables_closed_form(oracle_lib):
    t = co.tables()
    C = np.stack([t["cx"], t["cy"], t["cz"]], 1)
    assert np.array_equal(C, lattice.C)
    assert np.array_equal(t["w"], lattice.W)
    for i in range(19):
        assert np.array_equal(C[t["opp"][i]], -C[i])
        assert np.array_equal(C[t["mirz"][i]], C[i] * (1, 1, -1))
        assert np.array_equal(C[t["miry"][i]], C[i] * (1, -1, 1))
    # D3Q19 isotropy: sum w = 1, sum w c_a c_b = delta/3, sum w c_a^4 = 1/3 ... in FP64
    w = nl.W
    assert w.sum() == pytest.approx(1.0, abs=1e-15)
    assert np.allclose(np.einsum("i,ia,ib->ab", w, C, C), np.eye(3) / 3, atol=1e-15)
    assert np.array_equal(lattice.OPP, [0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17])
    assert np.array_equal(lattice.MIRZ, [0, 1, 2, 3, 4, 6, 5, 7, 8, 9, 10, 13, 14, 11, 12, 17, 18, 15, 16])


def test_feq_moments_closed_form():
    rng = np.random.default_rng(0)
    rho = 1 + 0.1 * rng.standard_normal(50)
    u = 0.05 * rng.standard_normal((3, 50))
    f = nl.feq(rho, u)
    C = nl.C
    assert np.allclose(f.sum(0), rho, rtol=1e-14)
    assert np.allclose(np.tensordot(C.T, f, 1), rho * u, atol=1e-15)
    Pi = np.einsum("ia,ib,in->abn", C, C, f)
    expect = np.eye(3)[:, :, None] * rho / 3 + rho * np.einsum("an,bn->abn", u, u)
    assert np.allclose(Pi, expect, atol=1e-15)


# ------------------------------------------------------------------ init
def test_init_noise_free_closed_form(oracle_lib):
    flag = small_flag()
    uin = profile(flag.shape[2])
    for st in ("f32", "f64"):
        f = co.init(flag, uin, st)
        ref = nl.init(flag, uin)
        assert np.allclose(f, ref, rtol=2e-7 if st == "f32" else 1e-15, atol=0)
        solid = flag != 0
        assert np.allclose(f[:, solid], nl.W[:, None], rtol=1e-7)


def test_init_noise_bounds(oracle_lib):
    flag = small_flag(nx=10, ny=8, nz=8, wall=False)
    uin = profile(8)
    f = co.init(flag, uin, "f64", noise=0.02, seed=5)
    U = nl.macro(f, flag)
    fl = flag == 0
    u0 = uin[None, None, :] * np.ones(flag.shape)
    rel = np.abs(U[1][fl] - u0[fl]) / np.maximum(u0[fl], 1e-12)
    assert rel.max() <= 0.02 + 1e-9 and rel.max() > 0.01
    assert np.abs(U[2][fl]).max() <= 0.02 * uin.max() + 1e-12
    g = co.init(flag, uin, "f64", noise=0.02, seed=6)
    assert not np.array_equal(f, g)


# ------------------------------------------------------------------ streaming / boundaries (tracer, exact)
def tracer_run(flag, i, pos, ybc=0):
    """Rest state plus a marked population i at pos; returns the f_out deviation from rest
    after one collision-free... we use the streamed field directly via np_lbm.pull and the
    C oracle with Cs=0 and tau=1 so that moments reveal where mass went."""
    f = np.tile(nl.W[:, None, None, None], (1,) + flag.shape).astype(np.float64)
    f[(i,) + pos] += 1e-3
    return f


@pytest.mark.parametrize("i", range(1, 19))
def test_streaming_tracer_interior(oracle_lib, i):
    flag = small_flag(nx=10, ny=7, nz=9, wall=False)
    pos = (4, 3, 4)
    f = tracer_run(flag, i, pos)
    fs, _ = nl.pull(f, flag)
    dst = tuple(np.array(pos) + nl.C[i])
    d = fs - nl.W[:, None, None, None]
    nz_idx = np.argwhere(np.abs(d) > 1e-12)
    assert nz_idx.tolist() == [[i, *dst]]


def test_streaming_periodic_y_and_mirror_top(oracle_lib):
    flag = small_flag(nx=8, ny=5, nz=6, wall=False)
    # y wrap: population 4 (0,-1,0) leaving y=0 arrives at y=ny-1
    f = tracer_run(flag, 4, (3, 0, 3))
    fs, _ = nl.pull(f, flag)
    assert abs(fs[4, 3, 4, 3] - nl.W[4] - 1e-3) < 1e-15
    # free-slip sides: it is reflected into population 3 at the same cell
    fs, _ = nl.pull(f, flag, ybc=1)
    assert abs(fs[3, 3, 0, 3] - nl.W[3] - 1e-3) < 1e-15
    # top mirror: population 5 (0,0,+1) at the top cell returns as 6 in place
    f = tracer_run(flag, 5, (3, 2, 5))
    fs, _ = nl.pull(f, flag)
    assert abs(fs[6, 3, 2, 5] - nl.W[6] - 1e-3) < 1e-15
    # 11 (1,0,1) at the top moves in x and mirrors to 13 (1,0,-1)
    f = tracer_run(flag, 11, (3, 2, 5))
    fs, _ = nl.pull(f, flag)
    assert abs(fs[13, 4, 2, 5] - nl.W[13] - 1e-3) < 1e-15
    # outlet: the last column pulls its cx=-1 populations from itself
    f = tracer_run(flag, 2, (7, 2, 3))
    fs, _ = nl.pull(f, flag)
    assert abs(fs[2, 7, 2, 3] - nl.W[2] - 1e-3) < 1e-15


def test_bounce_back_and_force_single_link(oracle_lib):
    flag = small_flag(nx=8, ny=5, nz=6, wall=False)
    flag[4, 2, 3] = 1                                  # one stone cell
    # population 1 (+1,0,0) at the fluid cell (3,2,3) runs into the stone at (4,2,3): it
    # returns as population 2 at (3,2,3), and the stone feels -2 (f_1 - w) c_2 = +2e-3 in x
    f = tracer_run(flag, 1, (3, 2, 3))
    fs, fc = nl.pull(f, flag)
    assert abs(fs[2, 3, 2, 3] - nl.W[2] - 1e-3) < 1e-15
    assert fc[0, 3, 2, 3] == pytest.approx(2e-3, abs=1e-15)
    assert fc[1, 3, 2, 3] == 0.0
    # the C oracle's force row agrees (FP64, collision irrelevant for the force)
    uin = np.zeros(6, np.float32)
    _, force = co.step(f, flag, uin, NU0, storage="f64")
    assert force[3, 0] == pytest.approx(2e-3, rel=1e-12)
    assert force[3, 1] == 0.0
    assert np.all(force[np.arange(6) != 3] == 0.0)
    # a link into the ground gives bounce-back but no force
    f = tracer_run(flag, 6, (2, 2, 1))
    fs, fc = nl.pull(f, flag)
    assert abs(fs[5, 2, 2, 1] - nl.W[5] - 1e-3) < 1e-15
    _, force = co.step(f, flag, uin, NU0, storage="f64")
    assert np.all(force == 0.0)


# ------------------------------------------------------------------ collision
def test_rest_state_is_fixed_point_and_zero_force(oracle_lib):
    """f = w is a fixed point (to rounding: the float32 weights do not sum to exactly 1, so
    the FP32 step moves populations by a few ulp) and the wall force is exactly zero."""
    flag = small_flag()
    uin = np.zeros(flag.shape[2], np.float32)
    fl = (flag == 0)[None]
    for st, tol in (("f64", 1e-15), ("f32", 1e-6)):
        f = co.init(flag, uin, st)
        g, force = co.step(f, flag, uin, NU0, storage=st)
        rel = np.abs(np.where(fl, g - f, 0)) / f
        assert rel.max() <= tol
        assert np.all(force == 0.0)


def test_conservation_and_second_moment(oracle_lib):
    """In a fully periodic interior (no walls touched), collision keeps rho and rho u and
    sets Pi = rho/3 I + rho u u + (1 - 1/tau) N."""
    rng = np.random.default_rng(3)
    flag = small_flag(nx=9, ny=6, nz=8, wall=False)
    uin = profile(8)
    f = co.init(flag, uin, "f64") + 1e-3 * rng.standard_normal((19,) + flag.shape) * (flag == 0)
    fs, _ = nl.pull(f, flag)
    out, _ = co.step(f, flag, uin, NU0, cs=0.16, storage="f64")
    sel = (slice(None), slice(2, 7), slice(None), slice(2, 6))    # away from inlet / outlet / top / ground
    a, b = fs[sel], out[sel]
    assert np.allclose(a.sum(0), b.sum(0), rtol=1e-14)
    assert np.allclose(np.tensordot(nl.C.T, a, 1), np.tensordot(nl.C.T, b, 1), atol=1e-16)
    rho = a.sum(0)
    u = np.tensordot(nl.C.T, a, 1) / rho
    Pi_in = np.einsum("ia,ib,i...->ab...", nl.C, nl.C, a)
    eq = np.einsum("ab,...->ab...", np.eye(3), rho / 3) + rho * np.einsum("a...,b...->ab...", u, u)
    N = Pi_in - eq
    Pi_out = np.einsum("ia,ib,i...->ab...", nl.C, nl.C, b)
    keep = (Pi_out - eq) / np.where(np.abs(N) > 1e-9, N, np.nan)
    k = keep[np.isfinite(keep)]
    assert np.all((k > -1) & (k < 1))           # keep = 1 - 1/tau with tau > 1/2
    # one keep per cell: all stress components relax by the same factor
    kk = np.nanmax(keep, axis=(0, 1)) - np.nanmin(keep, axis=(0, 1))
    assert np.nanmax(kk) < 1e-6


def test_tau_closed_form_and_cs_zero(oracle_lib):
    """With Cs = 0 the relaxation factor is 1 - 1/tau0, tau0 = 3 nu + 1/2 (plain
    regularized BGK); with Cs > 0 it matches the scalar Smagorinsky formula."""
    rng = np.random.default_rng(4)
    flag = small_flag(nx=9, ny=6, nz=8, wall=False)
    uin = profile(8)
    f = co.init(flag, uin, "f64") + 1e-3 * rng.standard_normal((19,) + flag.shape) * (flag == 0)
    fs, _ = nl.pull(f, flag)
    cell = (4, 3, 4)
    v = fs[(slice(None),) + cell]
    rho = v.sum()
    u = nl.C.T @ v / rho
    N = np.einsum("ia,ib,i->ab", nl.C, nl.C, v) - np.eye(3) * rho / 3 - rho * np.outer(u, u)
    for cs in (0.0, 0.16):
        out, _ = co.step(f, flag, uin, 1e-3, cs=cs, storage="f64")
        o = out[(slice(None),) + cell]
        tau0 = 3e-3 + 0.5
        tau = 0.5 * (tau0 + np.sqrt(tau0 ** 2 + 18 * np.sqrt(2) * cs ** 2 * np.sqrt((N * N).sum()) / rho))
        Pi_out = np.einsum("ia,ib,i->ab", nl.C, nl.C, o) - np.eye(3) * rho / 3 - rho * np.outer(u, u)
        assert np.allclose(Pi_out, (1 - 1 / tau) * N, atol=1e-15)


def test_shear_wave_decay_rate(oracle_lib):
    """u_x = A sin(2 pi y / ny), y periodic, decays as exp(-nu k^2 t) with
    nu = (tau0 - 1/2)/3 (analytic; Cs = 0).  Information moves one cell per step, so after
    8 steps the cells at least 9 cells from the inlet, outlet, ground and top still see
    the unbounded shear wave; the decay rate is fitted there over steps 2..8."""
    nx, ny, nz = 21, 32, 21
    flag = np.zeros((nx, ny, nz), np.uint8)
    flag[:, :, 0] = 2
    nu, A = 0.1, 1e-4
    y = np.arange(ny)
    sw = np.sin(2 * np.pi * y / ny)
    u = np.zeros((3, nx, ny, nz))
    u[0] = A * sw[None, :, None]
    g = nl.feq(np.ones((nx, ny, nz)), u)
    uin = np.zeros(nz, np.float32)
    amps = []
    for t in range(1, 9):
        g, _ = co.step(g, flag, uin, nu, cs=0.0, storage="f64")
        U = nl.macro(g, flag)
        core = U[1][9:12, :, 9:12]
        amps.append((core * sw[None, :, None]).sum() / ((sw ** 2).sum() * 9))
    k = 2 * np.pi / ny
    slope = np.polyfit(np.arange(2, 9), np.log(amps[1:]), 1)[0]
    assert slope == pytest.approx(-nu * k * k, rel=0.02)


# ------------------------------------------------------------------ whole step: two restatements
@pytest.mark.parametrize("ybc", [0, 1])
def test_c_oracle_matches_numpy_restatement(oracle_lib, ybc):
    rng = np.random.default_rng(1 + ybc)
    flag = small_flag()
    nx, ny, nz = flag.shape
    uin = profile(nz)
    f = co.init(flag, uin, "f64") + rng.normal(0, 1e-3, (19, nx, ny, nz)) * (flag == 0)
    a, fa = co.step(f, flag, uin, NU0, 0.16, nx - 3, 0.05, gust=1.1, ybc=ybc, storage="f64")
    b, fb = nl.step(f, flag, uin, NU0, 0.16, nx - 3, 0.05, gust=float(np.float32(1.1)), ybc=ybc)
    fl = (flag == 0)[None]
    assert np.abs(np.where(fl, a - b, 0)).max() < 1e-15
    assert np.allclose(fa, fb, rtol=1e-13, atol=1e-17)
    assert np.abs(fa[:, 0]).max() > 1e-4                     # the wall does feel a force


def test_fp32_oracle_tracks_fp64(oracle_lib):
    flag = small_flag()
    nx, ny, nz = flag.shape
    uin = profile(nz)
    f64 = co.init(flag, uin, "f64")
    f32 = co.init(flag, uin, "f32")
    for _ in range(20):
        f64, F64 = co.step(f64, flag, uin, NU0, 0.16, nx - 3, storage="f64")
        f32, F32 = co.step(f32, flag, uin, NU0, 0.16, nx - 3)
    fl = (flag == 0)[None]
    rel = np.abs(np.where(fl, f32 - f64, 0)).max() / np.abs(f64).max()
    assert 0 < rel < 1e-6
    assert np.abs(F32 - F64).max() < 1e-4 * np.abs(F64).max()    # f - w cancels: FP32 force is ~1e-5 relative


# ------------------------------------------------------------------ 16-bit codecs
def test_f16_codec_exhaustive(oracle_lib):
    L = co.lib()
    allh = np.arange(65536, dtype=np.uint16)
    vals = allh.view(np.float16).astype(np.float32)
    back = np.array([L.oracle_f16_to_f32(int(h)) for h in allh], np.float32)
    fin = np.isfinite(vals)
    assert np.array_equal(back[fin].view(np.uint32), vals[fin].view(np.uint32))
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(20000).astype(np.float32) * s for s in (1e-7, 1e-5, 1e-3, 1, 1e3, 7e4)])
    x = np.concatenate([x, vals[fin], (vals[fin] + np.float32(0)).astype(np.float32) * np.float32(1.0004)])
    ours = np.array([L.oracle_f32_to_f16(float(v)) for v in x], np.uint16)
    with np.errstate(over="ignore"):                           # values past 65504 go to inf
        ref = x.astype(np.float16).view(np.uint16)
    assert np.array_equal(ours, ref)


def test_bf16_codec(oracle_lib):
    """BF16 storage rounding (dithered): values already in BF16 and +-0 are kept; every other
    value goes to one of its two BF16 neighbours; and for values a fraction p of an ulp above
    a BF16 number the share rounded up is p (round to nearest would give 0 or 1, a dither
    correlated with the discarded bits would miss p), so the mean rounding error is ~0."""
    L = co.lib()
    enc = lambda v: np.array([L.oracle_bf16_encode(float(t)) for t in v], np.uint16)   # noqa: E731
    rng = np.random.default_rng(1)
    # kept: BF16 values (random upper halves, zero lower halves) and signed zeros
    hb = rng.integers(0x0080, 0x7F00, 4000).astype(np.uint32) | (rng.integers(0, 2, 4000).astype(np.uint32) << 15)
    exact = (hb << 16).view(np.float32)
    assert np.array_equal(enc(exact), hb.astype(np.uint16))
    assert L.oracle_bf16_encode(0.0) == 0 and L.oracle_bf16_encode(-0.0) == 0x8000
    # one of the two neighbours: the truncated upper half or the next one up in magnitude
    x = np.concatenate([rng.standard_normal(20000).astype(np.float32) * s for s in (1e-30, 1e-5, 1, 1e30)])
    d = enc(x).astype(np.int64) - (x.view(np.uint32) >> 16).astype(np.int64)
    assert set(np.unique(d)) <= {0, 1}
    # the share rounded up follows the discarded fraction
    for p in (0.1, 0.3, 0.5, 0.7, 0.9):
        n = 20000
        up16 = rng.integers(0x3000, 0x4F00, n).astype(np.uint32)               # |v| in [2^-31, 2^31)
        low = np.clip(np.round(p * 65536) + rng.integers(-64, 65, n), 0, 65535).astype(np.uint32)
        v = ((up16 << 16) | low).view(np.float32)
        share = (enc(v).astype(np.uint32) != up16).mean()
        assert abs(share - p) < 0.02, (p, share)
        err = (enc(v).astype(np.uint32) << 16).view(np.float32).astype(np.float64) - v.astype(np.float64)
        ulp = ((up16 + 1) << 16).view(np.float32).astype(np.float64) - (up16 << 16).view(np.float32)
        assert abs((err / ulp).mean()) < 0.02, (p, (err / ulp).mean())
    assert L.oracle_bf16_to_f32(0x3F80) == 1.0


def test_shifted_storage_rest_state(oracle_lib):
    """f = w is stored as exact zeros; one step keeps it within FP32 rounding of w."""
    flag = small_flag()
    uin = np.zeros(flag.shape[2], np.float32)
    for st in ("f16", "bf16"):
        scale = 1024.0 if st == "f16" else 1.0
        f = co.init(flag, uin, st, scale=scale)
        assert np.all(f == 0)                                     # w_i - w_i = +0 stored
        g, force = co.step(f, flag, uin, NU0, storage=st, scale=scale)
        v = co.decode(g, st, scale)
        w = nl.W32.reshape(19, 1, 1, 1)
        rel = np.abs(np.where((flag == 0)[None], v - w, 0)) / w
        assert rel.max() < 1e-6
        assert np.all(force == 0.0)


# ------------------------------------------------------------------ storage offsets (16-bit forms)
def test_shift_table_closed_form(oracle_lib):
    """sigma_c(z) = feq_i(1, (U, 0, 0)) / w_i for c_x = 0, +1, -1: 1 + 3cU + 4.5 c^2 U^2 - 1.5 U^2
    to FP32 rounding (FP64 closed form); kind weights (U = 0) and uin[0] = 0 give 1."""
    uin = profile(9, H=4.0, u=0.06)
    w = co.shift_table(uin, "weights")
    assert w.shape == (9, 4) and np.array_equal(w[:, :3], np.ones((9, 3), np.float32)) and np.all(w[:, 3] == 0)
    s = co.shift_table(uin, "inlet")
    u = uin.astype(np.float64)[:, None]
    c = np.array([0.0, 1.0, -1.0])[None, :]
    closed = 1 + 3 * c * u + 4.5 * c ** 2 * u ** 2 - 1.5 * u ** 2
    assert np.allclose(s[:, :3], closed, rtol=2e-7, atol=0)
    assert np.array_equal(s[0, :3], np.ones(3, np.float32))
    # per direction: the equilibrium of the inlet velocity divided by the weight (FP64)
    sig = co.sigma_per_direction(s, 9)
    cu = 3 * nl.C[:, 0][:, None] * uin.astype(np.float64)[None, :]
    feq_w = 1 + cu + 0.5 * cu * cu - 1.5 * uin.astype(np.float64)[None, :] ** 2
    assert np.allclose(sig, feq_w, rtol=2e-7, atol=0)


def _inlet_case(nx=14, ny=5, nz=8):
    flag = small_flag(nx, ny, nz)
    uin = profile(nz, H=3.0, u=0.06)
    return flag, uin, co.shift_table(uin, "inlet")


def test_inlet_shift_stores_inlet_state_exactly(oracle_lib):
    """Without noise, every fluid cell starts at feq(1, (uin[z], 0, 0)): its stored value is
    exactly zero with the inlet offsets, for both 16-bit forms, and decodes to the FP32
    form's populations up to the rounding of the two operand orders (2 ulp)."""
    flag, uin, sh = _inlet_case()
    fluid = (flag == 0)[None]
    f32 = co.init(flag, uin, "f32")
    for st, scale in (("f16", 1024.0), ("bf16", 1.0)):
        f = co.init(flag, uin, st, scale=scale, shift=sh)
        assert np.all(np.where(fluid, f, 0) == 0)
        v = co.decode(f, st, scale, shift=sh)
        assert np.all(np.where(fluid, np.abs(v - f32), 0) <= 2 * np.spacing(np.abs(f32)))


def _half_ulp(q, st):
    """Half a unit in the last place of stored 16-bit values (as float32 magnitudes)."""
    if st == "f16":
        h = q.view(np.float16)
        return np.spacing(np.abs(h)).astype(np.float32) / 2
    x = (q.astype(np.uint32) << 16).view(np.float32)
    e = np.floor(np.log2(np.maximum(np.abs(x), np.float32(1e-30))))
    return np.float32(2.0) ** (e - 7) / 2


@pytest.mark.parametrize("st,scale", [("f16", 1024.0), ("bf16", 1.0)])
@pytest.mark.parametrize("ybc", [0, 1])
def test_16bit_step_matches_fp64_restatement(oracle_lib, st, scale, ybc):
    """One 16-bit step (inlet offsets) from a noisy state equals the independent FP64 NumPy
    step of the decoded inputs within the store rounding of each output (half an ulp of the
    stored value for FP16, one ulp for the dithered BF16) plus FP32 arithmetic (1.5e-7): every streamed link (offset of the source
    height), the top and side mirrors, the bounce-backs and the inlet decode and collide
    right.  Decoding any link with another height's offset misses by ~1e-4."""
    flag, uin, sh = _inlet_case()
    fluid = (flag == 0)[None]
    f = co.init(flag, uin, st, noise=0.03, seed=5, scale=scale, shift=sh)
    fin = co.decode(f, st, scale, shift=sh).astype(np.float64)
    g, F = co.step(f, flag, uin, NU0, 0.16, 10, 0.05, gust=1.1, ybc=ybc, storage=st, scale=scale, shift=sh)
    got = co.decode(g, st, scale, shift=sh).astype(np.float64)
    ref, Fref = nl.step(fin, flag, uin, NU0, 0.16, 10, 0.05, gust=np.float32(1.1), ybc=ybc)
    # FP16 rounds to nearest (half an ulp); BF16 rounds dithered (up to one ulp)
    ulps = 1.0 if st == "f16" else 2.0
    bound = nl.W.reshape(19, 1, 1, 1) * ulps * _half_ulp(g, st) / scale * 1.01 + 1.5e-7
    err = np.where(fluid, np.abs(got - ref) - bound, -1)
    assert err.max() <= 0, (st, float(err.max()))
    assert np.abs(F - Fref).max() <= 1e-6 * np.abs(Fref).max()


def test_inlet_shift_error_below_weight_shift_error(oracle_lib):
    """Near the inlet state, storing f - feq(uin) loses much less than storing f - w: after
    40 steps of a gently perturbed channel flow (ground only) the FP16 populations stay
    several times closer to the FP32 run (RMS over the fluid populations)."""
    flag = small_flag(16, 6, 8, wall=False)
    uin = profile(8, H=3.0, u=0.06)
    sh = co.shift_table(uin, "inlet")
    fluid = (flag == 0)[None]
    f32 = co.init(flag, uin, "f32", noise=0.01, seed=3)
    fa = co.init(flag, uin, "f16", noise=0.01, seed=3, scale=1024.0, shift=sh)
    fb = co.init(flag, uin, "f16", noise=0.01, seed=3, scale=1024.0)
    for _ in range(40):
        f32, _ = co.step(f32, flag, uin, NU0, 0.16, 12, 0.05, ybc=1)
        fa, _ = co.step(fa, flag, uin, NU0, 0.16, 12, 0.05, ybc=1, storage="f16", scale=1024.0, shift=sh)
        fb, _ = co.step(fb, flag, uin, NU0, 0.16, 12, 0.05, ybc=1, storage="f16", scale=1024.0)
    n = fluid.sum() * 19
    ea = np.sqrt((np.where(fluid, co.decode(fa, "f16", 1024.0, shift=sh) - f32, 0) ** 2).sum() / n)
    eb = np.sqrt((np.where(fluid, co.decode(fb, "f16", 1024.0) - f32, 0) ** 2).sum() / n)
    assert ea < 1e-6 and eb < 1e-5
    assert ea < eb / 3, (ea, eb)
