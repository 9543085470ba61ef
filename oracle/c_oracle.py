"""ctypes binding of oracle/liblbm_oracle.so (CPU oracle; test infrastructure only).

Arrays are numpy, in the reference layout f[i, x, y, z] (lbm3d.py:293)."""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liblbm_oracle.so")
ST = {"f32": 0, "f16": 1, "bf16": 2}
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        p, i, f, d, u = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_double, ctypes.c_uint
        L.oracle_init_f32.argtypes = [p, i, p, p, i, i, i, f, u, f, p]
        L.oracle_step_f32.argtypes = [p, p, i, p, p, f, i, i, i, f, f, i, f, i, f, p, p]
        L.oracle_macro_f32.argtypes = [p, i, p, i, i, i, f, p, p, p]
        L.oracle_shift_table.argtypes = [p, i, i, p]
        L.oracle_init_f64.argtypes = [p, p, p, i, i, i, f, u]
        L.oracle_step_f64.argtypes = [p, p, p, p, d, i, i, i, d, d, i, d, i, p]
        L.oracle_macro_f64.argtypes = [p, p, i, i, i, p]
        L.oracle_f32_to_f16.argtypes = [f]
        L.oracle_f32_to_f16.restype = ctypes.c_uint16
        L.oracle_f16_to_f32.argtypes = [ctypes.c_uint16]
        L.oracle_f16_to_f32.restype = f
        L.oracle_bf16_encode.argtypes = [f]
        L.oracle_bf16_encode.restype = ctypes.c_uint16
        L.oracle_bf16_to_f32.argtypes = [ctypes.c_uint16]
        L.oracle_bf16_to_f32.restype = f
        L.oracle_qcoarse.argtypes = [p, p, i, i, i, i, i, i, i, i, p]
        L.oracle_set_threads.argtypes = [i]
        L.oracle_num_threads.restype = i
        L.oracle_tables.argtypes = [p] * 7
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def set_threads(n):
    lib().oracle_set_threads(int(n))


def num_threads():
    return lib().oracle_num_threads()


def tables():
    cx, cy, cz, opp, mirz, miry = (np.zeros(19, np.int32) for _ in range(6))
    w = np.zeros(19, np.float32)
    lib().oracle_tables(_ptr(cx), _ptr(cy), _ptr(cz), _ptr(opp), _ptr(mirz), _ptr(miry), _ptr(w))
    return dict(cx=cx, cy=cy, cz=cz, opp=opp, mirz=mirz, miry=miry, w=w)


SHIFT_KINDS = {"weights": 0, "inlet": 1}


def shift_table(uin, kind="inlet"):
    """(nz, 4) float32 offset table of the 16-bit forms: sigma_c(z) for c_x = 0, +1, -1 (and
    a zero pad), the equilibrium feq_i(1, (U, 0, 0)) / w_i of U = uin[z] ("inlet") or U = 0
    ("weights", sigma = 1)."""
    uin = np.ascontiguousarray(uin, np.float32)
    out = np.zeros((uin.shape[0], 4), np.float32)
    lib().oracle_shift_table(_ptr(uin), uin.shape[0], SHIFT_KINDS[kind], _ptr(out))
    return out


def _shift_ptr(shift):
    if shift is None:
        return None
    assert shift.dtype == np.float32 and shift.flags.c_contiguous and shift.shape[1] == 4
    return _ptr(shift)


def storage_dtype(st):
    return np.float32 if st == "f32" else np.uint16


def init(flag, uin, storage="f32", noise=0.0, seed=0, scale=1.0, shift=None):
    flag = np.ascontiguousarray(flag, np.uint8)
    nx, ny, nz = flag.shape
    uin = np.ascontiguousarray(uin, np.float32)
    if storage == "f64":
        f = np.zeros((19, nx, ny, nz), np.float64)
        lib().oracle_init_f64(_ptr(f), _ptr(flag), _ptr(uin), nx, ny, nz, float(np.float32(noise)), seed)
        return f
    f = np.zeros((19, nx, ny, nz), storage_dtype(storage))
    lib().oracle_init_f32(_ptr(f), ST[storage], _ptr(flag), _ptr(uin), nx, ny, nz, noise, seed, scale,
                          _shift_ptr(shift))
    return f


def step(f_in, flag, uin, nu0, cs=0.16, x_sponge=None, nu_sponge=0.05, gust=1.0, ybc=0, storage="f32",
         scale=1.0, f_out=None, shift=None):
    """One k_step; returns (f_out, force_z (nz, 2) float64).  f_out is allocated if None;
    its solid cells keep whatever they held (the reference never writes them).  shift: the
    (nz, 19) offset table of the 16-bit forms (None = the weights)."""
    flag = np.ascontiguousarray(flag, np.uint8)
    nx, ny, nz = flag.shape
    if x_sponge is None:
        x_sponge = nx
    uin = np.ascontiguousarray(uin, np.float32)
    if f_out is None:
        f_out = np.zeros_like(f_in)
    force = np.zeros((nz, 2), np.float64)
    if storage == "f64":
        rc = lib().oracle_step_f64(_ptr(f_in), _ptr(f_out), _ptr(flag), _ptr(uin), float(np.float32(gust)), nx, ny,
                                   nz, float(nu0), float(cs), int(x_sponge), float(nu_sponge), ybc, _ptr(force))
    else:
        rc = lib().oracle_step_f32(_ptr(f_in), _ptr(f_out), ST[storage], _ptr(flag), _ptr(uin), gust, nx, ny, nz,
                                   nu0, cs, int(x_sponge), nu_sponge, ybc, scale, _shift_ptr(shift), _ptr(force))
    assert rc == 0
    return f_out, force


def macro(f, flag, storage="f32", scale=1.0, acc=None, shift=None):
    flag = np.ascontiguousarray(flag, np.uint8)
    nx, ny, nz = flag.shape
    if storage == "f64":
        U = np.zeros((4, nx, ny, nz), np.float64)
        lib().oracle_macro_f64(_ptr(f), _ptr(flag), nx, ny, nz, _ptr(U))
        return U
    U = np.zeros((4, nx, ny, nz), np.float32)
    lib().oracle_macro_f32(_ptr(f), ST[storage], _ptr(flag), nx, ny, nz, scale, _shift_ptr(shift), _ptr(U),
                           _ptr(acc) if acc is not None else None)
    return U


def sigma_per_direction(shift, nz):
    """(19, nz) float32 sigma of each direction's c_x class at each height."""
    from . import np_lbm
    cls = np.where(np_lbm.C[:, 0] == 0, 0, np.where(np_lbm.C[:, 0] > 0, 1, 2))
    if shift is None:
        return np.ones((19, nz), np.float32)
    return np.ascontiguousarray(shift[:, cls].T)


def decode(f, storage, scale=1.0, shift=None):
    """Populations of a storage form as float32 values.  16-bit forms: f_i = w_i * phi_i with
    phi_i = sigma_c(z) + d / scale (sigma = 1 when shift is None)."""
    from . import np_lbm
    if storage in ("f32", "f64"):
        return f
    nz = f.shape[3]
    sig = sigma_per_direction(shift, nz)[:, None, None, :]
    if storage == "f16":
        d = f.view(np.float16).astype(np.float32)
    else:
        d = (f.astype(np.uint32) << 16).view(np.float32)
    phi = sig + d * np.float32(1.0 / scale)
    return np_lbm.W32.reshape(19, 1, 1, 1) * phi


def qcoarse(U, flag, x0, ds, dims):
    """k_qcoarse (lbm3d.py:224-238) on U (4, nx, ny, nz) float32."""
    U = np.ascontiguousarray(U, np.float32)
    flag = np.ascontiguousarray(flag, np.uint8)
    nx, ny, nz = flag.shape
    out = np.zeros(dims, np.float32)
    lib().oracle_qcoarse(_ptr(U), _ptr(flag), nx, ny, nz, int(x0), int(ds), int(dims[0]), int(dims[1]), int(dims[2]),
                         _ptr(out))
    return out
