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
from .lattice import FLUID, GROUND, Q, WALL, flag_from_mask

TORCH_STORAGE = {nat.SW_F32: torch.float32, nat.SW_F16: torch.float16, nat.SW_BF16: torch.bfloat16,
                 nat.SW_F64: torch.float64}
DEFAULT_SCALE = {nat.SW_F32: 1.0, nat.SW_F64: 1.0, nat.SW_F16: 1024.0, nat.SW_BF16: 1.0}
BYTES_PER_UPDATE = {nat.SW_F32: 152, nat.SW_F16: 76, nat.SW_BF16: 76, nat.SW_F64: 304}


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
    gust      : optional (n_steps,) float32 inlet factor per step.
    n_steps   : steps this run may take (sizes the force record and the gust table).
    ring      : steps per force-partials ring (reduced every `ring` steps).
    """

    def __init__(self, flag, uin, *, nu0, cs=0.16, x_sponge=None, nu_sponge=0.05, ybc="periodic", dtype="f32",
                 scale=None, gust=None, n_steps=1000, ring=64, device="cuda", record_force=True):
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
        self.links = torch.empty(self.N, dtype=torch.int32, device=dev)
        flag_d = torch.from_numpy(flag).to(dev)
        nat.call("sw_prepare_links", flag_d.data_ptr(), self.links.data_ptr(), self.nx, self.ny, self.nz, self.ybc,
                 _stream())
        del flag_d
        st = TORCH_STORAGE[self.dtype]
        self.f = [torch.empty(Q * self.N, dtype=st, device=dev), torch.zeros(Q * self.N, dtype=st, device=dev)]
        self.uin = torch.from_numpy(uin).to(dev)
        self.gust = None
        if gust is not None:
            g = np.ascontiguousarray(gust, np.float32)
            if g.shape[0] < self.n_steps:
                raise ValueError("gust table shorter than n_steps")
            self.gust = torch.from_numpy(g).to(dev)
        per = nat.lib().sw_partials_per_step(self.ny, self.nz)
        self.record_force = bool(record_force)
        self.partials = torch.zeros(self.ring * per, dtype=torch.float64, device=dev) if record_force else None
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
                 self.ny, self.nz, float(noise), int(seed) & 0xFFFFFFFF, self.scale, _stream())
        self.t = 0
        self._pending = 0
        return self

    def step(self, n=1):
        """n fused steps (lbm3d.py:358-362)."""
        lib = nat.lib()
        s = _stream()
        for _ in range(int(n)):
            if self.t >= self.n_steps:
                raise RuntimeError(f"step {self.t} beyond n_steps={self.n_steps}")
            rc = lib.sw_step(self.f[0].data_ptr(), self.f[1].data_ptr(), self.dtype, self.links.data_ptr(),
                             self.uin.data_ptr(), _ptr(self.gust), self.t, None, self.nx, self.ny, self.nz,
                             self.nu0, self.cs, self.x_sponge, self.nu_sponge, self.ybc, self.scale,
                             _ptr(self.partials), self.ring, s)
            nat.check(rc, "sw_step")
            self.f[0], self.f[1] = self.f[1], self.f[0]
            self.t += 1
            if self.record_force:
                self._pending += 1
                if self._pending == self.ring:
                    self.flush_force()
        return self

    def flush_force(self):
        """Reduce the pending partials into force rows (fixed summation order)."""
        if self.record_force and self._pending:
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
        """Raw stored populations in the reference layout (19, nx, ny, nz) (storage dtype)."""
        eb = nat.DTYPE_BYTES[self.dtype]
        out = self._to_ref(self.f[0], Q, eb).cpu()
        if self.dtype == nat.SW_BF16:
            return out.view(torch.int16).numpy().view(np.uint16).reshape(Q, self.nx, self.ny, self.nz)
        if self.dtype == nat.SW_F16:
            return out.numpy().view(np.uint16).reshape(Q, self.nx, self.ny, self.nz)
        return out.numpy().reshape(Q, self.nx, self.ny, self.nz)

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
                 self.scale, self.U.data_ptr(), None, _stream())
        return self._to_ref(self.U, 4, 4).cpu().numpy().reshape(4, self.nx, self.ny, self.nz)

    def accumulate(self):
        """Uacc += (rho, u) of the current state (k_macro + k_accum, lbm3d.py:367-370), on the device."""
        if self.Uacc is None:
            self.Uacc = torch.zeros(4 * self.N, dtype=torch.float32, device=self.device)
        nat.call("sw_macro", self.f[0].data_ptr(), self.dtype, self.links.data_ptr(), self.nx, self.ny, self.nz,
                 self.scale, None, self.Uacc.data_ptr(), _stream())
        self.n_acc += 1

    def mean_field(self):
        """Time-mean (rho, u) = Uacc / n_acc in the reference layout (lbm3d.py:402)."""
        if self.Uacc is None:
            return None
        return self._to_ref(self.Uacc, 4, 4).cpu().numpy().reshape(4, self.nx, self.ny, self.nz) / max(self.n_acc, 1)

    def checksum(self):
        """FP64 fluid-only sums (sum f, sum (i+1) f) of the decoded populations, on the device."""
        f = self.f[0].view(Q, self.N)
        w = torch.tensor([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12, dtype=torch.float32, device=self.device)
        if self.dtype in (nat.SW_F16, nat.SW_BF16):
            v = w[:, None] + f.float() * np.float32(1.0 / self.scale)
        else:
            v = f
        fluid = (self.links & (1 << 31)) == 0
        v = torch.where(fluid[None, :], v.double(), torch.zeros((), dtype=torch.float64, device=self.device))
        idx = torch.arange(1, Q + 1, dtype=torch.float64, device=self.device)[:, None]
        return float(v.sum()), float((v * idx).sum())


def bytes_per_update(dtype):
    d = nat.DTYPES[dtype] if isinstance(dtype, str) else dtype
    return BYTES_PER_UPDATE[d]


__all__ = ["Simulation", "bytes_per_update", "FLUID", "WALL", "GROUND"]
