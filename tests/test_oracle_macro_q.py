"""Pins of the oracle's macroscopic readout (k_macro + k_accum, lbm3d.py:158-188) and of its
Q-criterion frames (q_at + k_qcoarse, lbm3d.py:203-238), independent of the HIP kernels."""
import numpy as np
import pytest

from oracle import c_oracle as co
from oracle import np_lbm as nl


def small_flag(nx=10, ny=6, nz=7):
    flag = np.zeros((nx, ny, nz), np.uint8)
    flag[:, :, 0] = 2                                    # ground
    flag[4:6, 2:4, 1:4] = 1                              # a stone
    return flag


def noisy_state(flag, seed=0, amp=1e-3):
    rng = np.random.default_rng(seed)
    uin = np.linspace(0.0, 0.05, flag.shape[2]).astype(np.float32)
    f = co.init(flag, uin, "f64") + amp * rng.standard_normal((19,) + flag.shape) * (flag == 0)
    return f, uin


def test_macro_matches_numpy_restatement(oracle_lib):
    """rho and u of the C oracle (FP64 and FP32) against the NumPy restatement
    (rho = sum f, u = sum c f / rho; solids (1, 0, 0, 0))."""
    flag = small_flag()
    f, _ = noisy_state(flag)
    ref = nl.macro(f, flag)
    U64 = co.macro(f, flag, "f64")
    assert np.allclose(U64, ref, rtol=0, atol=1e-14)
    f32 = f.astype(np.float32)
    U32 = co.macro(f32, flag)
    ref32 = nl.macro(f32.astype(np.float64), flag)
    assert np.allclose(U32, ref32, rtol=0, atol=2e-6)
    solid = flag != 0
    assert np.all(U32[0][solid] == 1.0) and np.all(U32[1:, solid] == 0.0)


def test_macro_closed_form_and_accumulate(oracle_lib):
    """The noise-free initial state has rho = 1 and u = (uin(z), 0, 0) at every fluid cell;
    two accumulations of the same field give exactly twice the field."""
    flag = small_flag()
    uin = np.linspace(0.0, 0.05, flag.shape[2]).astype(np.float32)
    f = co.init(flag, uin, "f32")
    U = co.macro(f, flag)
    fl = flag == 0
    assert np.allclose(U[0][fl], 1.0, rtol=0, atol=2e-6)
    assert np.allclose(U[1], np.where(fl, uin[None, None, :], 0.0), rtol=0, atol=5e-7)
    assert np.allclose(U[2:], 0.0, rtol=0, atol=5e-7)
    acc = np.zeros_like(U)
    co.macro(f, flag, acc=acc)
    co.macro(f, flag, acc=acc)
    assert np.array_equal(acc, 2 * U)


@pytest.mark.parametrize("st,scale", [("f16", 1024.0), ("bf16", 1.0)])
def test_macro_16bit_equals_macro_of_decoded(oracle_lib, st, scale):
    """The 16-bit macro reads the stored values through the storage offsets: it equals the
    NumPy macro of the decoded populations to FP32 rounding."""
    flag = small_flag()
    _, uin = noisy_state(flag)
    sh = co.shift_table(uin, "inlet")
    q = co.init(flag, uin, st, noise=0.03, seed=4, scale=scale, shift=sh)
    U = co.macro(q, flag, st, scale, shift=sh)
    ref = nl.macro(co.decode(q, st, scale, shift=sh).astype(np.float64), flag)
    assert np.allclose(U, ref, rtol=0, atol=2e-6)


def q_numpy(U, flag):
    """Q = (|Omega|^2 - |S|^2) / 2 of the central-difference velocity gradient
    (S, Omega: symmetric and antisymmetric parts), 0 at solids and next to them, y periodic;
    defined for 1 <= x < nx - 1, 1 <= z < nz - 1 (NaN elsewhere)."""
    u = U[1:4].astype(np.float64)
    nx, ny, nz = flag.shape
    A = np.full((3, 3, nx, ny, nz), np.nan)
    for i in range(3):
        A[i, 0, 1:-1] = 0.5 * (u[i, 2:] - u[i, :-2])
        A[i, 1] = 0.5 * (np.roll(u[i], -1, 1) - np.roll(u[i], 1, 1))
        A[i, 2, :, :, 1:-1] = 0.5 * (u[i, :, :, 2:] - u[i, :, :, :-2])
    S = 0.5 * (A + A.transpose(1, 0, 2, 3, 4))
    W = 0.5 * (A - A.transpose(1, 0, 2, 3, 4))
    Q = 0.5 * ((W ** 2).sum((0, 1)) - (S ** 2).sum((0, 1)))
    s = flag != 0
    near = s.copy()
    near[1:] |= s[:-1]
    near[:-1] |= s[1:]
    near |= np.roll(s, 1, 1) | np.roll(s, -1, 1)
    near[:, :, 1:] |= s[:, :, :-1]
    near[:, :, :-1] |= s[:, :, 1:]
    Q[near] = 0.0
    Q[0] = Q[-1] = np.nan
    Q[:, :, 0] = Q[:, :, -1] = np.nan
    return Q


def pool(Q, x0, ds, dims):
    out = np.zeros(dims)
    for i in range(dims[0]):
        for j in range(dims[1]):
            for k in range(dims[2]):
                blk = Q[x0 + i * ds:x0 + (i + 1) * ds, j * ds:(j + 1) * ds, k * ds:(k + 1) * ds]
                out[i, j, k] = max(0.0, np.nanmax(blk)) if np.isfinite(blk).any() else 0.0
    return out


def test_q_rotation_strain_and_solids(oracle_lib):
    """Rigid rotation in the x-z plane, u = (w (z - zc), 0, -w (x - xc)), has Q = w^2; pure
    strain u = (a (x - xc), 0, -a (z - zc)) has Q = -a^2 < 0, pooled to 0; cells at or next
    to a solid give 0."""
    nx, ny, nz = 12, 4, 10
    flag = np.zeros((nx, ny, nz), np.uint8)
    x, z = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    U = np.zeros((4, nx, ny, nz), np.float32)
    U[0] = 1.0
    U[1] = (0.5 * (z - 5))[:, None, :]
    U[3] = (-0.5 * (x - 6))[:, None, :]
    q = co.qcoarse(U, flag, 1, 1, (nx - 2, ny, nz - 2))
    assert np.all(q[:, :, 1:] == 0.25)                   # w = 0.5; k = 0 is the z = 0 row (outside)
    flag[6, 1, 5] = 1
    q = co.qcoarse(U, flag, 1, 1, (nx - 2, ny, nz - 2))
    for (x_, y_, z_) in [(6, 1, 5), (5, 1, 5), (7, 1, 5), (6, 0, 5), (6, 2, 5), (6, 1, 4), (6, 1, 6)]:
        assert q[x_ - 1, y_, z_] == 0.0
    assert q[3, 1, 5] == 0.25
    S = U.copy()
    S[1] = (0.25 * (x - 6))[:, None, :]
    S[3] = (-0.25 * (z - 5))[:, None, :]
    assert np.all(co.qcoarse(S, np.zeros_like(flag), 1, 2, (5, 2, 4)) == 0.0)


def test_q_pooled_matches_numpy_definition(oracle_lib):
    """Max-pooled Q of a random smooth field with a stone and the ground equals the NumPy
    restatement from the symmetric / antisymmetric split (a transposed gradient or a
    dropped cross term would miss)."""
    rng = np.random.default_rng(3)
    nx, ny, nz = 20, 8, 12
    flag = np.zeros((nx, ny, nz), np.uint8)
    flag[:, :, 0] = 2
    flag[9:11, 3:5, 1:5] = 1
    x, y, z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    U = np.zeros((4, nx, ny, nz), np.float32)
    U[0] = 1.0
    for c in range(1, 4):
        a = rng.standard_normal(6) * 0.02
        U[c] = (a[0] * np.sin(0.3 * x + a[1] * 10) * np.cos(2 * np.pi * y / ny + a[2] * 10)
                + a[3] * np.sin(0.4 * z + a[4] * 10) * np.cos(0.2 * x) + a[5] * np.sin(2 * np.pi * y / ny) * z / nz)
    ref = pool(q_numpy(U, flag), 2, 2, (9, 4, 6))
    q = co.qcoarse(U, flag, 2, 2, (9, 4, 6))
    assert np.abs(ref).max() > 1e-5
    assert np.allclose(q, ref, rtol=1e-5, atol=1e-9)
