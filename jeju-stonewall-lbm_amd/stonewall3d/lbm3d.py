"""Solver surface of stonewall3d/lbm3d.py over the MI355X HIP kernels.

The reference's main() (lbm3d.py:241-428) builds a solid mask, allocates device arrays,
runs k_init, loops k_step with a ping-pong swap, records the wall force per height and
step, averages the velocity field, checkpoints and writes Cd / Cl / Cm.  Simulation keeps
those pieces as methods, each a call into libstonewall_lbm.so:

  reference (lbm3d.py)                         here
  flag set-up :288-290 + per-step flag tests    Simulation.__init__ -> sw_prepare_links
  wp.launch(k_init) :340                        Simulation.init     -> sw_init
  wp.launch(k_step) :359-361, swap :362         Simulation.step     -> sw_step (+ sw_force_reduce)
  wp.launch(k_macro / k_accum) :367-369         Simulation.macro / .accumulate -> sw_macro
  force.numpy() :396                            Simulation.force_rows
  Cd / Cl / Cm / fx_profile :397-408            Simulation.coefficients
  save_ckpt / resume :320-355                   stonewall3d.checkpoint

Buffers are PyTorch-ROCm tensors (plumbing only); every kernel runs on the current torch
stream.  There is no CPU fallback: without the built library the constructor raises.
"""
import numpy as np
import torch

from . import _native as nat
from . import npz
from .lattice import FLUID, GROUND, Q, WALL, flag_from_mask

TORCH_STORAGE = {nat.SW_F32: torch.float32, nat.SW_F16: torch.float16, nat.SW_BF16: torch.bfloat16,
                 nat.SW_F64: torch.float64}
DEFAULT_SCALE = {nat.SW_F32: 1.0, nat.SW_F64: 1.0, nat.SW_F16: 1024.0, nat.SW_BF16: 1.0}
BYTES_PER_UPDATE = {nat.SW_F32: 152, nat.SW_F16: 76, nat.SW_BF16: 76, nat.SW_F64: 304}
# offsets of the 16-bit forms, which store (f_i / w_i - sigma_c(z)) * scale: sigma is the
# normalised equilibrium of the inlet velocity at each height ("inlet"), or 1 ("weights",
# i.e. (f_i - w_i) / w_i)
DEFAULT_SHIFT = "inlet"
W32 = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12, np.float32)
XCLASS = np.array([0, 1, 2, 0, 0, 0, 0, 1, 2, 1, 2, 1, 2, 1, 2, 0, 0, 0, 0])   # c_x = 0, +1, -1


def storage_shift(uin_dev, dtype, kind=None):
    """Device sigma table (nz, 4) float32 of a 16-bit storage form (sigma_c(z) for c_x = 0,
    +1, -1 and a zero pad), or None for FP32 / FP64 (the kernels then ignore it)."""
    kind = DEFAULT_SHIFT if kind is None else kind
    if kind not in nat.SHIFTS:
        raise ValueError(f"shift must be one of {sorted(nat.SHIFTS)}")
    if dtype not in (nat.SW_F16, nat.SW_BF16):
        return None
    nz = uin_dev.numel()
    t = torch.empty((nz, 4), dtype=torch.float32, device=uin_dev.device)
    nat.call("sw_shift_table", uin_dev.data_ptr(), nz, nat.SHIFTS[kind], t.data_ptr(), _stream())
    return t


def decoded(f_flat, dtype, scale, shift, nz):
    """float32 (19, N) population values f_i of a device state f[i][z][y][x]; the 16-bit
    forms decode phi_i = sigma + d / scale and return w_i * phi_i."""
    f = f_flat.view(Q, -1)
    if dtype not in (nat.SW_F16, nat.SW_BF16):
        return f
    dev = f.device
    if shift is None:
        sig = torch.ones((Q, nz, 1), dtype=torch.float32, device=dev)
    else:
        sig = shift[:, torch.from_numpy(XCLASS).to(dev)].t().contiguous()[:, :, None]
    phi = sig + f.float().view(Q, nz, -1) * np.float32(1.0 / scale)
    return (torch.from_numpy(W32).to(dev)[:, None, None] * phi).view(Q, -1)


def _ptr(t):
    return None if t is None else t.data_ptr()


def _stream():
    return torch.cuda.current_stream().cuda_stream


class Simulation:
    """One lattice: solid flags in; init, step, force rows, macroscopic fields out.

    flag      : (nx, ny, nz) uint8 FLUID/WALL/GROUND (lbm3d.py:288-290), or a bool stone
                mask (then stones -> WALL and the layer z=0 -> GROUND, as the reference).
    uin       : (nz,) float32 inlet profile (lbm3d.py:274-275).
    nu0, cs   : molecular viscosity and Smagorinsky constant (lbm3d.py:277, --Cs).
    x_sponge  : cells with x > x_sponge get nu0 + nu_sponge s^2 (lbm3d.py:140-143);
                None = nx - 3H is the caller's choice, so here None means no sponge.
    ybc       : "periodic" (the reference) or "freeslip" side walls.
    dtype     : "f32" | "f16" | "bf16" | "f64" population storage.
    shift     : 16-bit forms: "inlet" (default; store f_i / w_i - feq_i(1, (uin[z], 0, 0)) / w_i)
                or "weights" (store f_i / w_i - 1), times the scale.
    gust      : optional (n_steps,) float32 inlet factor per step.
    n_steps   : steps this run may take (sizes the force record and the gust table).
    ring      : steps per force-partials ring (reduced every `ring` steps).
    two_step  : advance pairs of steps with one launch of the two-step kernel (sw_step2:
                FP32 / FP16 / BF16 storage, population arrays below 4 GiB) and odd steps with
                its one-step form (sw_step2_one); populations are bit-identical to single
                steps, force rows equal to rounding (to the two-step batch bit for bit with
                sw_set_step2_rows(sw_batch_step2_rows())).
    """

    def __init__(self, flag, uin, *, nu0, cs=0.16, x_sponge=None, nu_sponge=0.05, ybc="periodic", dtype="f32",
                 scale=None, gust=None, n_steps=1000, ring=64, device="cuda", record_force=True, shift=None,
                 two_step=False):
        nat.lib()
        flag = np.asarray(flag)
        if flag.dtype == bool:
            flag = flag_from_mask(flag)
        flag = np.ascontiguousarray(flag, np.uint8)
        if flag.ndim != 3:
            raise ValueError("flag must be (nx, ny, nz)")
        self.nx, self.ny, self.nz = flag.shape
        self.N = self.nx * self.ny * self.nz
        self.flag = flag
        self.dtype = nat.DTYPES[dtype] if isinstance(dtype, str) else int(dtype)
        self.dtype_name = {v: k for k, v in nat.DTYPES.items()}[self.dtype]
        self.scale = float(DEFAULT_SCALE[self.dtype] if scale is None else scale)
        self.ybc = {"periodic": nat.SW_Y_PERIODIC, "freeslip": nat.SW_Y_FREESLIP}[ybc] if isinstance(ybc, str) else ybc
        self.nu0, self.cs, self.nu_sponge = float(nu0), float(cs), float(nu_sponge)
        self.x_sponge = self.nx if x_sponge is None else int(x_sponge)
        self.n_steps = int(n_steps)
        self.ring = int(ring)
        self.device = torch.device(device)
        uin = np.ascontiguousarray(uin, np.float32)
        if uin.shape != (self.nz,):
            raise ValueError(f"uin must have shape ({self.nz},)")
        dev = self.device
        self.links = torch.empty(nat.lib().sw_links_words(self.nx, self.ny, self.nz), dtype=torch.int32, device=dev)
        flag_d = torch.from_numpy(flag).to(dev)
        nat.call("sw_prepare_links", flag_d.data_ptr(), self.links.data_ptr(), self.nx, self.ny, self.nz, self.ybc,
                 _stream())
        del flag_d
        st = TORCH_STORAGE[self.dtype]
        self.f = [torch.empty(Q * self.N, dtype=st, device=dev), torch.zeros(Q * self.N, dtype=st, device=dev)]
        self.uin = torch.from_numpy(uin).to(dev)
        self.shift_kind = DEFAULT_SHIFT if shift is None else shift
        self.shift = storage_shift(self.uin, self.dtype, self.shift_kind)
        self.gust = None
        if gust is not None:
            g = np.ascontiguousarray(gust, np.float32)
            if g.shape[0] < self.n_steps:
                raise ValueError("gust table shorter than n_steps")
            self.gust = torch.from_numpy(g).to(dev)
        per = nat.lib().sw_partials_per_step(self.ny, self.nz)
        self.record_force = bool(record_force)
        self.partials = torch.zeros(self.ring * per, dtype=torch.float64, device=dev) if record_force else None
        self.two_step = bool(two_step)
        if self.two_step:
            if self.dtype == nat.SW_F64 or Q * self.N * nat.DTYPE_BYTES[self.dtype] >= (1 << 32):
                raise ValueError("two_step needs FP32 / FP16 / BF16 storage and population arrays below 4 GiB")
            if self.ring < 2:
                raise ValueError("two_step needs ring >= 2")
        self.partials2 = None
        if self.two_step and record_force:
            per2 = nat.lib().sw_step2_partials_per_step(self.nx, self.ny, self.nz)
            self.partials2 = torch.zeros(self.ring * per2, dtype=torch.float64, device=dev)
        self._layout = 1                # kernel that wrote the pending partials: 1 = sw_step, 2 = sw_step2
        self.force = torch.zeros((self.n_steps, self.nz, 2), dtype=torch.float64, device=dev) if record_force else None
        self.t = 0
        self._pending = 0
        self.U = None
        self.Uacc = None
        self.n_acc = 0

    # ------------------------------------------------------------------ state
    @property
    def f_cur(self):
        return self.f[0]

    def init(self, noise=0.0, seed=0):
        """k_init (lbm3d.py:39-54, launched at :340)."""
        nat.call("sw_init", self.f[0].data_ptr(), self.dtype, self.links.data_ptr(), self.uin.data_ptr(), self.nx,
                 self.ny, self.nz, float(noise), int(seed) & 0xFFFFFFFF, self.scale, _ptr(self.shift), _stream())
        self.t = 0
        self._pending = 0
        return self

    def step(self, n=1, accumulate=False, samples=None):
        """n fused steps (lbm3d.py:358-362); with two_step, pairs of them per launch.
        samples: offsets among the n steps (0 = the first) whose resulting state is added to
        the time-mean accumulator (k_macro + k_accum after k_step, lbm3d.py:363-370) by the
        launch that computes it; accumulate=True samples the last of them.  With two_step and
        FP32 storage a sampled step inside a pair stays in the pair (sw_step2_accumulate);
        otherwise it runs alone (sw_step_accumulate)."""
        lib = nat.lib()
        s = _stream()
        n = int(n)
        if self.t + n > self.n_steps:
            raise RuntimeError(f"step {self.t + n - 1} beyond n_steps={self.n_steps}")
        samples = {int(k) for k in (samples or ())} | ({n - 1} if accumulate and n > 0 else set())
        if any(k < 0 or k >= n for k in samples):
            raise ValueError("samples must lie in [0, n)")
        if samples and self.Uacc is None:
            self.Uacc = torch.zeros(4 * self.N, dtype=torch.float32, device=self.device)
        acc_pairs = self.two_step and self.dtype == nat.SW_F32
        i = 0
        while i < n:
            s0, s1 = i in samples, (i + 1) in samples
            pair = self.two_step and n - i >= 2 and not (s0 and s1) and (acc_pairs or not (s0 or s1))
            # with two_step, a single step other than a sampled one takes the two-step kernel's
            # tiles (sw_step2_one), so its partials share the pairs' ring layout
            single2 = self.two_step and not pair and not s0
            layout = 2 if (pair or single2) else 1
            adv = 2 if pair else 1
            if self.record_force and self._pending and (self._layout != layout or
                                                        self._pending + adv > self.ring):
                self.flush_force()
            self._layout = layout
            common = (self.f[0].data_ptr(), self.f[1].data_ptr(), self.dtype, self.links.data_ptr(),
                      self.uin.data_ptr(), _ptr(self.gust), self.t, None, self.nx, self.ny, self.nz, self.nu0,
                      self.cs, self.x_sponge, self.nu_sponge, self.ybc, self.scale, _ptr(self.shift))
            if single2:
                nat.check(lib.sw_step2_one(*common, _ptr(self.partials2), self.ring, s), "sw_step2_one")
            elif pair and (s0 or s1):
                nat.check(lib.sw_step2_accumulate(*common, _ptr(self.partials2), self.ring, self.Uacc.data_ptr(),
                                                  1 if s0 else 2, s), "sw_step2_accumulate")
                self.n_acc += 1
            elif pair:
                nat.check(lib.sw_step2(*common, _ptr(self.partials2), self.ring, s), "sw_step2")
            elif s0:
                nat.check(lib.sw_step_accumulate(*common, _ptr(self.partials), self.ring, self.Uacc.data_ptr(), s),
                          "sw_step_accumulate")
                self.n_acc += 1
            else:
                nat.check(lib.sw_step(*common, _ptr(self.partials), self.ring, s), "sw_step")
            self.f[0], self.f[1] = self.f[1], self.f[0]
            self.t += adv
            i += adv
            if self.record_force:
                self._pending += adv
                if self._pending == self.ring:
                    self.flush_force()
        return self

    def flush_force(self):
        """Reduce the pending partials into force rows (fixed summation order)."""
        if self.record_force and self._pending:
            if self._layout == 2:
                nat.call("sw_force_reduce2", self.partials2.data_ptr(), self.ring, self.t - self._pending, None,
                         self._pending, self.nx, self.ny, self.nz, self.force.data_ptr(), _stream())
            else:
                nat.call("sw_force_reduce", self.partials.data_ptr(), self.ring, self.t - self._pending, None,
                         self._pending, self.ny, self.nz, self.force.data_ptr(), _stream())
            self._pending = 0

    # ------------------------------------------------------------------ readout
    def force_rows(self):
        """(t, nz, 2) float64 wall force per step and height (lbm3d.py:301, :396)."""
        self.flush_force()
        return self.force[:self.t].cpu().numpy()

    def coefficients(self, UH, H, zeta=None, n_spin=0):
        """Cd, Cl, Cm per step and the mean fx profile, as lbm3d.py:396-408."""
        fh = self.force_rows()
        if zeta is None:
            zeta = np.arange(self.nz) - 0.5
        qref = 0.5 * UH ** 2
        Fx, Fz = fh[:, :, 0].sum(1), fh[:, :, 1].sum(1)
        M = (fh[:, :, 0] * zeta[None, :]).sum(1)
        A = H * self.ny
        return dict(Cd=Fx / (qref * A), Cl=Fz / (qref * A), Cm=M / (qref * A * H),
                    fx_profile=fh[n_spin:, :, 0].mean(0) / (qref * self.ny))

    def _to_ref(self, src, ncomp, elem_bytes):
        out = torch.empty_like(src)
        nat.call("sw_to_ref_layout", src.data_ptr(), out.data_ptr(), elem_bytes, ncomp, self.nx, self.ny, self.nz,
                 _stream())
        return out

    def populations(self):
        """Raw stored populations in the reference layout (19, nx, ny, nz) (storage dtype;
        the 16-bit forms hold (f_i / w_i - sigma) * scale, see shift_table())."""
        eb = nat.DTYPE_BYTES[self.dtype]
        out = self._to_ref(self.f[0], Q, eb).cpu()
        if self.dtype == nat.SW_BF16:
            return out.view(torch.int16).numpy().view(np.uint16).reshape(Q, self.nx, self.ny, self.nz)
        if self.dtype == nat.SW_F16:
            return out.numpy().view(np.uint16).reshape(Q, self.nx, self.ny, self.nz)
        return out.numpy().reshape(Q, self.nx, self.ny, self.nz)

    def shift_table(self):
        """(nz, 4) float32 sigma table of a 16-bit form (None for FP32 / FP64)."""
        return None if self.shift is None else self.shift.cpu().numpy()

    def load_populations(self, arr):
        """Inverse of populations(): raw storage values in the reference layout."""
        arr = np.ascontiguousarray(arr).reshape(-1)
        st = TORCH_STORAGE[self.dtype]
        if self.dtype in (nat.SW_F16, nat.SW_BF16):
            src = torch.from_numpy(arr.view(np.int16)).to(self.device).view(st)
        else:
            src = torch.from_numpy(arr).to(self.device)
        if src.numel() != Q * self.N:
            raise ValueError("population array has the wrong size")
        nat.call("sw_from_ref_layout", src.data_ptr(), self.f[0].data_ptr(), nat.DTYPE_BYTES[self.dtype], Q, self.nx,
                 self.ny, self.nz, _stream())
        torch.cuda.current_stream().synchronize()

    def macro(self):
        """(4, nx, ny, nz) float32 (rho, ux, uy, uz) in the reference layout (k_macro, lbm3d.py:158-182)."""
        if self.U is None:
            self.U = torch.empty(4 * self.N, dtype=torch.float32, device=self.device)
        nat.call("sw_macro", self.f[0].data_ptr(), self.dtype, self.links.data_ptr(), self.nx, self.ny, self.nz,
                 self.scale, _ptr(self.shift), self.U.data_ptr(), None, _stream())
        return self._to_ref(self.U, 4, 4).cpu().numpy().reshape(4, self.nx, self.ny, self.nz)

    def accumulate(self):
        """Uacc += (rho, u) of the current state (k_macro + k_accum, lbm3d.py:367-370), on the device."""
        if self.Uacc is None:
            self.Uacc = torch.zeros(4 * self.N, dtype=torch.float32, device=self.device)
        nat.call("sw_macro", self.f[0].data_ptr(), self.dtype, self.links.data_ptr(), self.nx, self.ny, self.nz,
                 self.scale, _ptr(self.shift), None, self.Uacc.data_ptr(), _stream())
        self.n_acc += 1

    def accumulator(self):
        """Raw Uacc (sum of sampled (rho, u)) in the reference layout, float32 (zeros if unused)."""
        if self.Uacc is None:
            return np.zeros((4, self.nx, self.ny, self.nz), np.float32)
        return self._to_ref(self.Uacc, 4, 4).cpu().numpy().reshape(4, self.nx, self.ny, self.nz)

    def set_accumulator(self, acc, n_acc):
        """Restore Uacc from a reference-layout (4, nx, ny, nz) float32 array."""
        acc = np.ascontiguousarray(acc, np.float32).reshape(-1)
        if acc.size != 4 * self.N:
            raise ValueError("accumulator has the wrong size")
        if self.Uacc is None:
            self.Uacc = torch.zeros(4 * self.N, dtype=torch.float32, device=self.device)
        src = torch.from_numpy(acc).to(self.device)
        nat.call("sw_from_ref_layout", src.data_ptr(), self.Uacc.data_ptr(), 4, 4, self.nx, self.ny, self.nz,
                 _stream())
        torch.cuda.current_stream().synchronize()
        self.n_acc = int(n_acc)

    def set_force_rows(self, rows):
        """Restore the force record of steps 0..len(rows)-1 (resume)."""
        if not self.record_force:
            return
        rows = np.asarray(rows, np.float64)
        self.force.zero_()
        if rows.shape[0]:
            self.force[:rows.shape[0]] = torch.from_numpy(rows).to(self.device)
        self._pending = 0

    def mean_field(self):
        """Time-mean (rho, u) = Uacc / n_acc in the reference layout (lbm3d.py:402)."""
        if self.Uacc is None:
            return None
        return self._to_ref(self.Uacc, 4, 4).cpu().numpy().reshape(4, self.nx, self.ny, self.nz) / max(self.n_acc, 1)

    def checksum(self):
        """FP64 fluid-only sums (sum f, sum (i+1) f) of the decoded populations, on the device."""
        v = decoded(self.f[0], self.dtype, self.scale, self.shift, self.nz)
        fluid = (self.links[:self.N] & (1 << 31)) == 0
        v = torch.where(fluid[None, :], v.double(), torch.zeros((), dtype=torch.float64, device=self.device))
        idx = torch.arange(1, Q + 1, dtype=torch.float64, device=self.device)[:, None]
        return float(v.sum()), float((v * idx).sum())


    def slices(self, y_mid, z_half):
        """(ux, uy, uz) on the plane y = y_mid (3, nx, nz) and z = z_half (3, nx, ny) as float16,
        from the current state (k_macro + k_slice_xz / k_slice_xy, lbm3d.py:367, :372-375).
        The slices are cut on the device; only the planes are copied to the host."""
        if self.U is None:
            self.U = torch.empty(4 * self.N, dtype=torch.float32, device=self.device)
        nat.call("sw_macro", self.f[0].data_ptr(), self.dtype, self.links.data_ptr(), self.nx, self.ny, self.nz,
                 self.scale, _ptr(self.shift), self.U.data_ptr(), None, _stream())
        U = self.U.view(4, self.nz, self.ny, self.nx)
        xz = U[1:, :, y_mid, :].transpose(1, 2).to(torch.float16).cpu().numpy()
        xy = U[1:, z_half, :, :].transpose(1, 2).to(torch.float16).cpu().numpy()
        return xz, xy


    def q_frame(self, x0, ds, dims, q_scale, qmax=400.0):
        """uint8 log-quantised max-pooled Q-criterion frame of the current state
        (k_macro + k_qcoarse + q_frame, lbm3d.py:312-318, :367, :376-377)."""
        if self.U is None:
            self.U = torch.empty(4 * self.N, dtype=torch.float32, device=self.device)
        nat.call("sw_macro", self.f[0].data_ptr(), self.dtype, self.links.data_ptr(), self.nx, self.ny, self.nz,
                 self.scale, _ptr(self.shift), self.U.data_ptr(), None, _stream())
        q = torch.empty(int(np.prod(dims)), dtype=torch.float32, device=self.device)
        nat.call("sw_qcoarse", self.U.data_ptr(), self.links.data_ptr(), self.nx, self.ny, self.nz, int(x0), int(ds),
                 int(dims[0]), int(dims[1]), int(dims[2]), q.data_ptr(), _stream())
        q = np.clip(q.cpu().numpy().reshape(dims) * q_scale, 0, qmax)
        return np.round(255 * np.log1p(q) / np.log1p(qmax)).astype(np.uint8)


def bytes_per_update(dtype):
    d = nat.DTYPES[dtype] if isinstance(dtype, str) else dtype
    return BYTES_PER_UPDATE[d]


__all__ = ["Simulation", "bytes_per_update", "FLUID", "WALL", "GROUND"]


# ---------------------------------------------------------------------------------- CLI
REFERENCE_SIG_KEYS = ("grid", "H", "porosity", "seed", "Re", "U0", "Cs", "n_spin", "n_total", "snap_every", "snap3d")


def run_signature(a, grid, n_spin, n_total, snap_every, ds3):
    """Checkpoint signature of a CLI run (lbm3d.py:324-325).  With the reference's storage
    (FP32), side boundary (periodic) and steady inlet it holds exactly the reference's eleven
    keys, so the reference and this package accept each other's checkpoints; each setting
    that departs from those adds its own key, because such populations are not
    interchangeable with the reference's."""
    sig = dict(grid=list(grid), H=a.H, porosity=a.porosity, seed=a.seed, Re=a.Re, U0=a.U0, Cs=a.Cs, n_spin=n_spin,
               n_total=n_total, snap_every=snap_every, snap3d=[a.snap3d_every, ds3])
    if a.dtype != "f32":
        sig["dtype"] = a.dtype
        if a.dtype in ("f16", "bf16"):
            sig["shift"] = a.shift
    if a.ybc != "periodic":
        sig["ybc"] = a.ybc
    if a.gust_seed is not None:
        sig["gust_seed"] = a.gust_seed
    return sig


def divergence_guard(U, case, t):
    """The run's NaN check (lbm3d.py:386-391): stops the run with SystemExit when the
    macroscopic field U (rho, ux, uy, uz) holds a value that is not finite."""
    if not np.isfinite(U).all():
        raise SystemExit(f"[{case}] diverged at step {t}")


def _parser():
    import argparse
    ap = argparse.ArgumentParser(description="D3Q19 regularized-BGK + Smagorinsky LES of wind over a dry-stone "
                                             "wall on one MI355X (flags as stonewall3d/lbm3d.py:242-264)")
    ap.add_argument("--case", required=True, help="name of the run; the outputs are <out>/<case>.npz and .ckpt")
    ap.add_argument("--porosity", type=float, required=True,
                    help="open fraction of the wall's front face; 0 fills the gaps with mortar")
    ap.add_argument("--H", type=int, default=40, help="cells across the wall height")
    ap.add_argument("--device", default="cuda", help="torch device of the buffers (a ROCm GPU)")
    ap.add_argument("--seed", type=int, default=7, help="seed of the stone layout and the initial noise")
    ap.add_argument("--Re", type=float, default=15000)
    ap.add_argument("--U0", type=float, default=0.08)
    ap.add_argument("--Cs", type=float, default=0.16)
    ap.add_argument("--up", type=float, default=4)
    ap.add_argument("--down", type=float, default=16)
    ap.add_argument("--span", type=float, default=4)
    ap.add_argument("--height", type=float, default=8)
    ap.add_argument("--thick", type=float, default=0.3)
    ap.add_argument("--t-spinup", type=float, default=20)
    ap.add_argument("--t-avg", type=float, default=30)
    ap.add_argument("--snap-dt", type=float, default=0.25)
    ap.add_argument("--snap3d-every", type=int, default=2,
                    help="snapshots between two pooled Q-criterion volumes (0 disables them)")
    ap.add_argument("--snap3d-ds", type=int, default=0)
    ap.add_argument("--out", default="results")
    ap.add_argument("--ckpt-every", type=int, default=5000)
    ap.add_argument("--no-resume", action="store_true")
    # additions
    ap.add_argument("--dtype", default="f32", choices=["f32", "f16", "bf16", "f64"], help="population storage")
    ap.add_argument("--shift", default=DEFAULT_SHIFT, choices=sorted(nat.SHIFTS),
                    help="16-bit storage offset: inlet equilibrium per height, or the weights")
    ap.add_argument("--ybc", default="periodic", choices=["periodic", "freeslip"], help="side boundaries")
    ap.add_argument("--gust-seed", type=int, default=None, help="gusting inlet from this seed (default: steady)")
    ap.add_argument("--noise", type=float, default=0.02, help="init velocity noise amplitude (lbm3d.py:340)")
    ap.add_argument("--max-steps", type=int, default=None, help="stop after this many steps (smoke runs)")
    ap.add_argument("--one-step", action="store_true",
                    help="one step per kernel launch (default for FP32 arrays below 4 GiB: two per launch, "
                         "populations identical, force rows equal to rounding; the 16-bit and FP64 forms always "
                         "take one, their fastest form)")
    return ap


def main(argv=None):
    """The reference's run (lbm3d.py:241-428) over the HIP kernels: same flags, same
    checkpoint directory and the same .npz keys."""
    import json
    import os
    import shutil
    import time

    from . import checkpoint
    from .geometry import build_wall
    from .inflow import gust_table, make_gust, reference_setup

    a = _parser().parse_args(argv)
    H = a.H
    su = reference_setup(H=H, Re=a.Re, U0=a.U0, up=a.up, down=a.down, span=a.span, height=a.height, thick=a.thick,
                         t_spinup=a.t_spinup, t_avg=a.t_avg, snap_dt=a.snap_dt)
    nx, ny, nz, X0, T = su.nx, su.ny, su.nz, su.X0, su.T
    n_total, n_spin, snap_every = su.n_total, su.n_spin, su.snap_every
    if a.max_steps is not None:
        n_total = min(n_total, a.max_steps)
    bpu = BYTES_PER_UPDATE[nat.DTYPES[a.dtype]]
    print(f"[lattice] {nx} x {ny} x {nz} ({nx*ny*nz/1e6:.1f} M cells), {a.dtype} populations "
          f"{2*nx*ny*nz*bpu/2/1e9:.1f} GB; inflow U_H {su.UH:.4f}, viscosity {su.nu0:.2e} (tau0 "
          f"{3*su.nu0+0.5:.5f}), Re {a.Re:.0f}; {n_total} steps, one H/U = {su.t_unit:.0f} steps", flush=True)
    wall, phi = build_wall(a.porosity, H, nx, ny, nz, X0, T, seed=a.seed)
    gust = gust_table(make_gust(a.gust_seed), n_total) if a.gust_seed is not None else None
    # two steps per launch where it is the faster kernel: FP32 below 4 GiB per array (0.90 against
    # 0.64 of 8 TB/s on 640x320x160); the 16-bit two-step form runs at about two thirds of the
    # rate of their one-step kernel, and FP64 has no two-step form
    two = a.dtype == "f32" and not a.one_step and Q * nx * ny * nz * nat.DTYPE_BYTES[nat.DTYPES[a.dtype]] < (1 << 32)
    sim = Simulation(wall, su.uin, nu0=su.nu0, cs=a.Cs, x_sponge=su.x_sponge, nu_sponge=0.05, ybc=a.ybc,
                     dtype=a.dtype, gust=gust, n_steps=n_total, device=a.device, shift=a.shift, two_step=two)
    y_mid, z_half = ny // 2, int(round(0.5 * H)) + 1
    ds3 = a.snap3d_ds if a.snap3d_ds > 0 else max(1, int(round(H / 13)))            # lbm3d.py:308-313
    q_x0 = max(1, X0 - int(round(0.5 * H)))
    q_dims = (min(int(round(9.0 * H)), nx - 1 - q_x0) // ds3, ny // ds3, min(int(round(2.5 * H)), nz - 1) // ds3)
    QMAX, q_scale = 400.0, (H / su.UH) ** 2
    os.makedirs(a.out, exist_ok=True)
    ckpt = os.path.join(a.out, f"{a.case}.ckpt")
    sig = run_signature(a, [nx, ny, nz], n_spin, n_total, snap_every, ds3)
    snaps_xz, snaps_xy, snaps_q, t_start = [], [], [], 0
    if os.path.isdir(ckpt) and not a.no_resume:
        t_start, n_acc, snaps = checkpoint.load(sim, ckpt, sig)
        snaps_xz, snaps_xy, snaps_q = list(snaps["snaps_xz"]), list(snaps["snaps_xy"]), list(snaps["snaps_q"])
        print(f"[{a.case}] continuing from the checkpoint at step {t_start} = {t_start/su.t_unit:.1f} H/U", flush=True)
    else:
        sim.init(noise=a.noise, seed=a.seed)

    def save_ckpt(t_next):
        checkpoint.save(sim, ckpt, t_next, sim.n_acc, sig,
                        dict(snaps_xz=np.array(snaps_xz, np.float16).reshape(-1, 3, nx, nz),
                             snaps_xy=np.array(snaps_xy, np.float16).reshape(-1, 3, nx, ny),
                             snaps_q=np.array(snaps_q, np.uint8).reshape(-1, *q_dims)))
        print(f"[{a.case}] step {t_next}: state written to {ckpt}", flush=True)

    def first_at_or_after(t, base, period):
        """Smallest step k >= max(t, base) with (k - base) % period == 0."""
        k = max(t, base)
        return k + (-(k - base)) % period

    def next_event(t):
        """The first step >= t after which the loop snapshots, checkpoints or reports."""
        ev = [n_total - 1, first_at_or_after(t, 0, 1000), first_at_or_after(t, n_spin, snap_every)]
        if a.ckpt_every > 0:
            ev.append(first_at_or_after(t + 1, 0, a.ckpt_every) - 1)
        return min(e for e in ev if e >= t)

    # The steps between two events run as one call (pairs of steps per launch where the storage
    # allows); each sampled step (every 5th of the averaging window) adds its state to the time
    # mean in the launch that computes it (lbm3d.py:363-370).
    t0 = time.time()
    t = t_start
    while t < n_total:
        e = next_event(t)
        sim.step(e - t + 1, samples=[k - t for k in range(first_at_or_after(t, n_spin, 5), e + 1, 5)])
        t = e
        if t >= n_spin:
            k = t - n_spin
            if k % snap_every == 0:
                xz, xy = sim.slices(y_mid, z_half)
                snaps_xz.append(xz)
                snaps_xy.append(xy)
                if a.snap3d_every > 0 and (k // snap_every) % a.snap3d_every == 0:
                    snaps_q.append(sim.q_frame(q_x0, ds3, q_dims, q_scale, QMAX))
        if a.ckpt_every > 0 and (t + 1) % a.ckpt_every == 0 and t + 1 < n_total:
            save_ckpt(t + 1)
        if t % 1000 == 0 or t == n_total - 1:
            torch.cuda.synchronize()
            el = time.time() - t0
            done = t + 1 - t_start
            mlups = done * nx * ny * nz / max(el, 1e-9) / 1e6
            chk = ""
            if t % 5000 == 0:
                un = sim.macro()
                divergence_guard(un, a.case, t)
                chk = (f"; density {un[0].min():.3f} .. {un[0].max():.3f}, largest |u| / U_H "
                       f"{np.abs(un[1:]).max()/su.UH:.2f}")
            print(f"[{a.case}] {t} of {n_total} steps ({t/su.t_unit:.1f} H/U), {mlups:.0f} MLUPS, "
                  f"{el/done*(n_total-t-1)/60:.1f} min to go{chk}", flush=True)
        t += 1

    fh = sim.force_rows()
    qref = 0.5 * su.UH ** 2
    Fx, Fz = fh[:, :, 0].sum(1), fh[:, :, 1].sum(1)
    M = (fh[:, :, 0] * su.zeta[None, :]).sum(1)
    A = H * ny
    mean = sim.mean_field()
    if mean is None:
        mean = np.zeros((4, nx, ny, nz), np.float32)
    npz.savez_compressed(                     # np.savez_compressed's archive, deflated on threads
        os.path.join(a.out, f"{a.case}.npz"),
        Cd=(Fx / (qref * A)).astype(np.float32), Cl=(Fz / (qref * A)).astype(np.float32),
        Cm=(M / (qref * A * H)).astype(np.float32),
        fx_profile=(fh[n_spin:, :, 0].mean(0) / (qref * ny)).astype(np.float32) if n_total > n_spin
        else np.zeros(nz, np.float32),
        mean_xz_mid=mean[:, :, y_mid, :].astype(np.float32),
        mean_xz_spanavg=np.where(wall.any(1)[None], np.nan, mean.mean(2)).astype(np.float32),
        mean_xy_half=mean[:, :, :, z_half].astype(np.float32),
        mean3d_ds2=mean[1:, ::2, ::2, ::2].astype(np.float16),
        snaps_xz=np.array(snaps_xz), snaps_xy=np.array(snaps_xy),      # as lbm3d.py:418 stacks them
        wall_xz=wall[:, y_mid, :], wall_xy=wall[:, :, z_half], wall_front=wall.any(0), wall_top=wall.any(2),
        wall_ds2=wall[::2, ::2, ::2], snaps_q=np.array(snaps_q, dtype=np.uint8).reshape(-1, *q_dims),
        meta=json.dumps(dict(case=a.case, porosity_target=a.porosity, porosity=float(phi), H=H, T=T, X0=X0,
                             q3d=dict(x0=q_x0, ds=ds3, dims=list(q_dims), qmax=QMAX,
                                      every=a.snap3d_every * snap_every),
                             nx=nx, ny=ny, nz=nz, UH=su.UH, U0=a.U0, Re=a.Re, nu=su.nu0, Cs=a.Cs, n_spin=n_spin,
                             n_avg=n_total - n_spin, steps_per_unit=su.t_unit, snap_every=snap_every, y_mid=y_mid,
                             z_half=z_half, seed=a.seed, device=a.device, dtype=a.dtype, shift=sim.shift_kind, ybc=a.ybc,
                             gust_seed=a.gust_seed, wall_seconds=time.time() - t0)))
    s = slice(n_spin, None)
    if n_total > n_spin:
        cd_avg = Fx[s] / (qref * A)
        print(f"[{a.case}] finished: frontal porosity {phi:.3f}, over the averaging window Cd {cd_avg.mean():.3f} "
              f"(peak {cd_avg.max():.3f}), Cm {(M[s] / (qref * A * H)).mean():.3f}", flush=True)
    else:
        print(f"[{a.case}] finished before the averaging window; frontal porosity {phi:.3f}", flush=True)
    if os.path.isdir(ckpt):
        shutil.rmtree(ckpt)


if __name__ == "__main__":
    main()
