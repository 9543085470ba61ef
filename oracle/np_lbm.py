"""Independent NumPy restatement of the D3Q19 regularized-BGK + Smagorinsky step (CPU
oracle; test infrastructure only).

Written from the equations the reference implements (stonewall3d/lbm3d.py:33-155), not
from the C oracle, so that the two restatements do not share a misreading:

  pull        f*_i(x) = f_i(x - c_i), with the outlet copying x+1 -> x (:82-83), the top
              mirroring z (:84-86), y periodic (:79) or mirrored (free-slip), and
              f*_i(x) = f_opp(i)(x) where x - c_i is solid (half-way bounce-back, :88-90)
  force       F = sum over wall links of -2 (f_opp(i)(x) - w_i) c_i           (:91-94)
  moments     rho = sum f*, rho u = sum c f*, Pi = sum c c f*                 (:103-130)
  stress      N = Pi - rho/3 I - rho u u                                      (:133-138)
  viscosity   nu = nu0 (+ nu_s s^2 in the sponge), tau0 = 3 nu + 1/2,
              tau = (tau0 + sqrt(tau0^2 + 18 sqrt(2) Cs^2 |N| / rho)) / 2      (:140-147)
  collision   f_i = feq_i + (1 - 1/tau) w_i 9/2 (c_i c_i - I/3) : N           (:149-155)
  inlet       x = 0 fluid cells: f_i = feq_i(1, (u_in(z) g, 0, 0))             (:67-71)

Arrays are float64 in the reference layout f[i, x, y, z].  Meant for lattices of a few
thousand cells.
"""
import numpy as np

C = np.array([(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
              (1, 1, 0), (-1, -1, 0), (1, -1, 0), (-1, 1, 0), (1, 0, 1), (-1, 0, -1), (1, 0, -1), (-1, 0, 1),
              (0, 1, 1), (0, -1, -1), (0, 1, -1), (0, -1, 1)])
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
W32 = W.astype(np.float32)


def _find(v):
    return int(np.flatnonzero((C == v).all(1))[0])


OPP = np.array([_find(-c) for c in C])
MIRZ = np.array([_find(c * (1, 1, -1)) for c in C])
MIRY = np.array([_find(c * (1, -1, 1)) for c in C])


def feq(rho, u):
    """Second-order equilibrium for rho (...) and u (3, ...)."""
    cu = np.tensordot(C, u, axes=(1, 0))              # (19, ...)
    usq = (u * u).sum(0)
    return W.reshape((19,) + (1,) * rho.ndim) * rho * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * usq)


def init(flag, uin):
    """Noise-free k_init: u = (uin[z], 0, 0) in fluid, 0 in solids."""
    nx, ny, nz = flag.shape
    u = np.zeros((3, nx, ny, nz))
    u[0] = np.where(flag == 0, np.asarray(uin, np.float64)[None, None, :], 0.0)
    return feq(np.ones((nx, ny, nz)), u)


def pull(f, flag, ybc=0):
    """Streamed populations f* and, per cell, the wall-force contribution (2, nx, ny, nz)."""
    nx, ny, nz = flag.shape
    X, Y, Z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    fs = np.empty_like(f)
    force = np.zeros((2,) + flag.shape)
    for i in range(19):
        xs, ys, zs = X - C[i, 0], Y - C[i, 1], Z - C[i, 2]
        j = np.full(flag.shape, i)
        xs = np.where(xs >= nx, X, xs)
        if ybc == 0:
            ys = ys % ny
        else:
            out = (ys < 0) | (ys >= ny)
            ys = np.where(out, Y, ys)
            j = np.where(out, MIRY[j], j)
        top = zs >= nz
        zs = np.where(top, Z, zs)
        j = np.where(top, MIRZ[j], j)
        xs_c, zs_c = np.clip(xs, 0, nx - 1), np.clip(zs, 0, nz - 1)
        src_flag = flag[xs_c, ys, zs_c]
        streamed = f[j, xs_c, ys, zs_c]
        bounced = f[OPP[i]]
        solid_src = src_flag != 0
        fs[i] = np.where(solid_src, bounced, streamed)
        wall = src_flag == 1
        contrib = -2.0 * (bounced - W[i])
        force[0] += np.where(wall, contrib * C[i, 0], 0.0)
        force[1] += np.where(wall, contrib * C[i, 2], 0.0)
    return fs, force


def collide(fs, x_index, nu0, cs, x_sponge, nu_sponge):
    rho = fs.sum(0)
    u = np.tensordot(C.T, fs, axes=(1, 0)) / rho
    Pi = np.einsum("ia,ib,i...->ab...", C, C, fs)
    N = Pi - np.einsum("ab,...->ab...", np.eye(3), rho / 3) - rho * np.einsum("a...,b...->ab...", u, u)
    nx = fs.shape[1]
    s = np.clip((x_index - x_sponge) / (nx - x_sponge), 0, None) if x_sponge < nx else np.zeros_like(x_index, float)
    nu = nu0 + np.where(x_index > x_sponge, nu_sponge * s * s, 0.0)
    tau0 = 3 * nu + 0.5
    qn = np.sqrt((N * N).sum((0, 1)))
    tau = 0.5 * (tau0 + np.sqrt(tau0 ** 2 + 18 * np.sqrt(2) * cs * cs * qn / rho))
    Qi = np.einsum("ia,ib->iab", C, C) - np.eye(3)[None] / 3
    neq = 4.5 * W.reshape(19, 1, 1, 1) * np.einsum("iab,ab...->i...", Qi, N)
    return feq(rho, u) + (1 - 1 / tau) * neq


def step(f, flag, uin, nu0, cs=0.16, x_sponge=None, nu_sponge=0.05, gust=1.0, ybc=0):
    """One step; returns (f_out with solid cells zero, force per z (nz, 2))."""
    nx, ny, nz = flag.shape
    if x_sponge is None:
        x_sponge = nx
    fs, fcell = pull(f, flag, ybc)
    X = np.broadcast_to(np.arange(nx)[:, None, None], flag.shape).astype(float)
    out = collide(fs, X, nu0, cs, x_sponge, nu_sponge)
    uin64 = np.asarray(uin, np.float64) * gust
    u_in = np.zeros((3, ny, nz))
    u_in[0] = uin64[None, :]
    out[:, 0] = feq(np.ones((ny, nz)), u_in)
    fluid = flag == 0
    out = np.where(fluid[None], out, 0.0)
    inner = fluid.copy()
    inner[0] = False
    force = np.stack([np.where(inner, fcell[0], 0).sum((0, 1)), np.where(inner, fcell[1], 0).sum((0, 1))], 1)
    return out, force


def macro(f, flag):
    rho = f.sum(0)
    with np.errstate(invalid="ignore", divide="ignore"):      # solid cells: rho = 0, set below
        u = np.tensordot(C.T, f, axes=(1, 0)) / rho
    U = np.concatenate([rho[None], u])
    solid = flag != 0
    U[0][solid] = 1.0
    U[1:, solid] = 0.0
    return U
