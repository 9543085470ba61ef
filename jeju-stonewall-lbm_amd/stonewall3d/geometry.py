"""Solid masks for the step: the reference's dry-stone wall, a box wall and the seeded
rounded-box stone wall of the explicit workloads.

build_wall() restates stonewall3d/geometry.py:13-161 (the Jeju dry-stone wall): the same
random draws in the same order from numpy's default_rng(seed), the same floating-point
formulas, so that it reproduces the reference's masks bit for bit.  Its steps have to
follow the reference's because the seeded geometry is part of the drop-in contract.
Coordinates: x streamwise, y along the wall (periodic), z up; cell k has its centre at
height zeta = k - 1/2 above the ground (k = 0 is the ground layer).
"""
import numpy as np


# --------------------------------------------------------------------------- reference wall
def _superellipse_half(dy, b, p):
    """Half-height of a superellipse of half-width b and exponent p at offsets dy (geometry.py:13-16)."""
    ady = np.abs(dy)
    r = 1.0 - np.clip(ady / b, 0, 1) ** p
    return np.where(ady < b, r ** (1.0 / p), 0.0)


def _periodic_offset(yg, yc, L):
    return (yg - yc + L / 2) % L - L / 2


def _draw_courses(H, Ly, rng):
    """Stone shapes course by course (geometry.py:19-41)."""
    heights = []
    z = 0.0
    while z < H * 0.97:
        h = rng.uniform(0.26, 0.38) * H
        heights.append(h)
        z += h
    courses = []
    for h in heights:
        lens = []
        total = 0.0
        while total < Ly:
            L = rng.uniform(0.45, 0.9) * H
            lens.append(L)
            total += L
        lens = np.array(lens) * Ly / total
        y = rng.uniform(0, Ly)
        course = []
        for L in lens:
            yc = (y + L / 2) % Ly
            c = 0.5 * h * rng.uniform(0.85, 1.1)
            p = rng.uniform(2.2, 3.2)
            dx = rng.uniform(-0.08, 0.08)
            ax = rng.uniform(0.85, 1.0)
            course.append({"yc": yc, "L": L, "c": c, "p": p, "dx": dx, "ax": ax})
            y += L
        courses.append(course)
    return courses


def _drop(courses, gap, H, Ly):
    """Rest each stone on the highest contact below it, then stretch to mean top H (geometry.py:44-65)."""
    ny = int(round(Ly))
    yg = np.arange(ny) + 0.5
    top = np.zeros(ny)
    stones = []
    for course in courses:
        nxt = top.copy()
        for st in course:
            b = 0.5 * st["L"] * (1.0 - gap)
            dy = _periodic_offset(yg, st["yc"], Ly)
            prof = _superellipse_half(dy, b, st["p"]) * st["c"]
            inside = np.abs(dy) < b
            zc = np.max(top[inside] + prof[inside])
            nxt = np.where(inside, np.maximum(nxt, zc + prof), nxt)
            s = dict(st)
            s["b"] = b
            s["zc"] = zc
            stones.append(s)
        top = nxt
    k = H / top.mean()
    for s in stones:
        s["zc"] *= k
        s["c"] *= k
    return stones


def _anchor(stones, ny):
    """Mark top-line and ground stones (geometry.py:105-118)."""
    yg = np.arange(ny) + 0.5
    env = np.zeros(ny)
    tops = []
    for s in stones:
        dy = _periodic_offset(yg, s["yc"], ny)
        t = np.where(np.abs(dy) < s["b"], s["zc"] + _superellipse_half(dy, s["b"], s["p"]) * s["c"], -np.inf)
        tops.append(t)
        env = np.maximum(env, t)
    for s, t in zip(stones, tops):
        s["top"] = bool(np.any(t >= env - 0.5))
        s["ground"] = bool(s["zc"] - s["c"] < 0.02 * s["c"] + 1e-6)
    return stones


def _scaled(stones, s):
    """Shrink every stone about its centre; top stones keep their top, ground stones their base (:121-132)."""
    out = []
    for st in stones:
        n = dict(st)
        n["b"] = st["b"] * s
        n["c"] = st["c"] * s
        if st["top"]:
            n["zc"] = st["zc"] + st["c"] * (1 - s)
        elif st["ground"]:
            n["zc"] = st["zc"] - st["c"] * (1 - s)
        out.append(n)
    return out


def silhouette(stones, ny, nz):
    """2D (ny, nz) view along x of the stones (geometry.py:87-97)."""
    yg = np.arange(ny) + 0.5
    kg = np.arange(nz) - 0.5
    m = np.zeros((ny, nz), bool)
    for s in stones:
        dy = _periodic_offset(yg, s["yc"], ny)
        half = _superellipse_half(dy, s["b"], s["p"]) * s["c"]
        m |= np.abs(kg[None, :] - s["zc"]) < half[:, None]
    m[:, 0] = False
    return m


def frontal_porosity(front, H):
    """Open fraction of the nominal face y in [0, Ly), heights 1..H (geometry.py:100-102)."""
    return 1.0 - front[:, 1:int(round(H)) + 1].mean()


def voxelize(stones, nx, ny, nz, X0, T):
    """Superellipsoid stones into an (nx, ny, nz) bool mask (geometry.py:68-84)."""
    m = np.zeros((nx, ny, nz), bool)
    xmid = X0 + T / 2
    for s in stones:
        a, b, c, p = s["ax"] * T / 2, s["b"], s["c"], s["p"]
        xc = xmid + s["dx"] * T
        xi = np.arange(int(np.floor(xc - a)), int(np.ceil(xc + a)) + 1)
        ki = np.arange(max(1, int(np.floor(s["zc"] - c))), min(nz, int(np.ceil(s["zc"] + c)) + 2))
        yi = np.arange(int(np.floor(s["yc"] - b)) - 1, int(np.ceil(s["yc"] + b)) + 2)
        if len(xi) == 0 or len(ki) == 0:
            continue
        rx = (np.abs(xi + 0.5 - xc) / a) ** p
        ry = (np.abs(yi + 0.5 - s["yc"]) / b) ** p
        rz = (np.abs(ki - 0.5 - s["zc"]) / c) ** p
        inside = (rx[:, None, None] + ry[None, :, None]) + rz[None, None, :] <= 1.0
        ii, jj, kk = np.nonzero(inside)
        m[xi[ii], yi[jj] % ny, ki[kk]] = True
    return m


def wall_stones(porosity, H, ny, nz, seed=7):
    """The stone list and scale build_wall rasterizes (porosity > 0 path), for inspection."""
    rng = np.random.default_rng(seed)
    base = _anchor(_drop(_draw_courses(H, ny, rng), 0.0, H, ny), ny)
    if porosity <= 0.0:
        return base, 1.0
    lo, hi = 0.4, 1.2
    for _ in range(24):
        s = 0.5 * (lo + hi)
        phi = frontal_porosity(silhouette(_scaled(base, s), ny, nz), H)
        if phi > porosity:
            lo = s
        else:
            hi = s
    return _scaled(base, s), s


def build_wall(porosity, H, nx, ny, nz, X0, T, seed=7, verbose=True):
    """Seeded Jeju dry-stone wall -> (mask (nx, ny, nz) bool, frontal porosity)  (geometry.py:135-161).

    porosity > 0: the packed stones are shrunk about their centres (24-step bisection on
    the scale) until the frontal open fraction reaches `porosity`.  porosity <= 0: the
    packed stones plus a mortar core filled up to each column's highest stone cell."""
    stones, s = wall_stones(porosity, H, ny, nz, seed)
    m = voxelize(stones, nx, ny, nz, X0, T)
    if porosity <= 0.0:
        xa, xb = X0 + int(round(0.15 * T)), X0 + int(round(0.85 * T))
        col = m.any(axis=0)
        for j in range(ny):
            ks = np.nonzero(col[j])[0]
            if len(ks):
                m[xa:max(xb, xa + 1), j, 1:ks.max() + 1] = True
    phi = frontal_porosity(m.any(axis=0), H)
    if verbose:
        print(f"[wall] stones={len(stones)} scale={s:.3f} frontal porosity={phi:.3f}", flush=True)
    return m, phi


# --------------------------------------------------------------------------- explicit workloads
def box_wall(nx, ny, nz, X0, T, H):
    """Solid box x in [X0, X0+T), all y, heights k = 1..H (a porosity-0 wall spanning the width)."""
    m = np.zeros((nx, ny, nz), bool)
    m[X0:X0 + T, :, 1:H + 1] = True
    return m


def _rounded_box(X, Y, Z, hx, hy, hz, r):
    """Points within distance r of the box with half-sizes (hx-r, hy-r, hz-r) centred at 0."""
    qx = np.maximum(np.abs(X) - (hx - r), 0.0)
    qy = np.maximum(np.abs(Y) - (hy - r), 0.0)
    qz = np.maximum(np.abs(Z) - (hz - r), 0.0)
    return qx * qx + qy * qy + qz * qz <= r * r


def rounded_box_wall(nx, ny, nz, X0, T, H, void_fraction, seed, size=(0.3, 0.6)):
    """Seeded wall of rounded-box stones in staggered courses, opened by seeded stone removal.

    Stones are boxes of length and height U[size] * H (12-24 cells at H = 40), spanning the
    thickness T, with corners rounded by a quarter of their smallest half-size, laid
    course by course (course heights drawn until the mean top reaches H; alternate courses
    shifted by half a stone: staggered).  Whole stones are then removed in a seeded random
    order until the void fraction of the wall's envelope x in [X0, X0+T), all y, k in
    [1, H] reaches `void_fraction`.  Returns (mask, info) with the achieved void fraction
    and the frontal porosity in the reference's sense (geometry.py:100-102)."""
    rng = np.random.default_rng(seed)
    lo, hi = size
    courses, z = [], 0.0                        # heights zeta above the ground surface
    while z < H - 0.5 * lo * H:
        h = rng.uniform(lo, hi) * H
        courses.append((z, min(h, H - z)))
        z += h
    stones = []
    for ci, (z0, h) in enumerate(courses):
        y = rng.uniform(0, 0.5) * hi * H + (0.5 * hi * H if ci % 2 else 0.0)
        start = y
        while y < start + ny - 0.5 * lo * H:
            L = min(rng.uniform(lo, hi) * H, start + ny - y)
            stones.append(dict(yc=y + L / 2, L=L, zc=z0 + h / 2, h=h))
            y += L
    x = np.arange(nx) + 0.5
    yv = np.arange(ny) + 0.5
    kv = np.arange(nz) - 0.5
    xc, hx = X0 + T / 2, T / 2
    # Stone cells lie in the x window [xa, xb) and heights k >= 1.  Each stone's cells are kept
    # as flat indices into a coverage count over that window (how many kept stones hold the
    # cell); removing a stone decrements its cells, so the void fraction after each removal
    # is updated from the stone's own cells instead of recomposing the whole mask.
    xa, xb = max(0, X0 - 1), min(nx, X0 + T + 1)
    cells, in_env = [], []
    for st in stones:
        r = 0.25 * min(hx, st["L"] / 2, st["h"] / 2)
        xi = np.arange(xa, xb)
        ki = np.arange(max(1, int(st["zc"] - st["h"] / 2) - 1), min(nz, int(st["zc"] + st["h"] / 2) + 2))
        yi = np.arange(int(st["yc"] - st["L"] / 2) - 1, int(st["yc"] + st["L"] / 2) + 2)
        dy = (yv[yi % ny] - st["yc"] + ny / 2) % ny - ny / 2
        ins = _rounded_box(x[xi][:, None, None] - xc, dy[None, :, None], kv[ki][None, None, :] - st["zc"],
                           hx, st["L"] / 2, st["h"] / 2, r)
        ii, jj, kk = np.nonzero(ins)
        xs, ks = xi[ii], ki[kk]
        flat = np.unique(((xs - xa) * ny + yi[jj] % ny) * nz + ks)
        cells.append(flat)
        xf, kf = flat // (ny * nz) + xa, flat % nz
        in_env.append(flat[(xf >= X0) & (xf < X0 + T) & (kf >= 1) & (kf <= H)])
    env_cells = T * ny * H
    cov = np.zeros((xb - xa) * ny * nz, np.int32)
    for flat in cells:
        cov[flat] += 1
    covered = int(np.count_nonzero(cov.reshape(xb - xa, ny, nz)[X0 - xa:X0 - xa + T, :, 1:H + 1]))

    def void():
        return 1.0 - covered / env_cells

    order = list(rng.permutation(len(stones)))
    keep = set(range(len(stones)))
    while void() < void_fraction and order:
        k = order.pop(0)
        keep.discard(k)
        cov[cells[k]] -= 1
        covered -= int(np.count_nonzero(cov[in_env[k]] == 0))
    m = np.zeros((nx, ny, nz), bool)
    win = cov.reshape(xb - xa, ny, nz) > 0
    m[xa:xb] = win
    info = dict(stones=len(stones), removed=len(stones) - len(keep), void_fraction=float(void()),
                frontal_porosity=float(frontal_porosity(win.any(axis=0), H)))
    return m, info
