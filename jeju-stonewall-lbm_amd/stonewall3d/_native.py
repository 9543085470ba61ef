"""ctypes binding of libstonewall_lbm.so (the HIP kernels' C ABI, include/stonewall_lbm.h).

The library is built in-tree (``make -C jeju-stonewall-lbm_amd``) and is the only
compute path: if it is missing, ``lib()`` raises; there is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libstonewall_lbm.so")

SW_F32, SW_F16, SW_BF16, SW_F64 = 0, 1, 2, 3
SW_Y_PERIODIC, SW_Y_FREESLIP = 0, 1
SW_SHIFT_WEIGHTS, SW_SHIFT_INLET = 0, 1
SHIFTS = {"weights": SW_SHIFT_WEIGHTS, "inlet": SW_SHIFT_INLET}
DTYPES = {"f32": SW_F32, "f16": SW_F16, "bf16": SW_BF16, "f64": SW_F64}
DTYPE_BYTES = {SW_F32: 4, SW_F16: 2, SW_BF16: 2, SW_F64: 8}

_c_int, _c_ll, _c_dbl, _c_flt, _c_uint = ctypes.c_int, ctypes.c_longlong, ctypes.c_double, ctypes.c_float, ctypes.c_uint
_p = ctypes.c_void_p

# name -> (restype, argtypes); every symbol declared in include/stonewall_lbm.h
SIGNATURES = {
    "sw_version": (_c_int, []),
    "sw_last_error": (ctypes.c_char_p, []),
    "sw_rows_per_block": (_c_int, []),
    "sw_links_words": (_c_ll, [_c_int, _c_int, _c_int]),
    "sw_partials_per_step": (_c_ll, [_c_int, _c_int]),
    "sw_prepare_links": (_c_int, [_p, _p, _c_int, _c_int, _c_int, _c_int, _p]),
    "sw_init": (_c_int, [_p, _c_int, _p, _p, _c_int, _c_int, _c_int, _c_flt, _c_uint, _c_flt, _p, _p]),
    "sw_step": (_c_int, [_p, _p, _c_int, _p, _p, _p, _c_int, _p, _c_int, _c_int, _c_int,
                         _c_dbl, _c_dbl, _c_int, _c_dbl, _c_int, _c_flt, _p, _p, _c_int, _p]),
    "sw_step_accumulate": (_c_int, [_p, _p, _c_int, _p, _p, _p, _c_int, _p, _c_int, _c_int, _c_int,
                                    _c_dbl, _c_dbl, _c_int, _c_dbl, _c_int, _c_flt, _p, _p, _c_int, _p, _p]),
    "sw_step2_partials_per_step": (_c_ll, [_c_int, _c_int, _c_int]),
    "sw_step2": (_c_int, [_p, _p, _c_int, _p, _p, _p, _c_int, _p, _c_int, _c_int, _c_int,
                          _c_dbl, _c_dbl, _c_int, _c_dbl, _c_int, _c_flt, _p, _p, _c_int, _p]),
    "sw_step2_one": (_c_int, [_p, _p, _c_int, _p, _p, _p, _c_int, _p, _c_int, _c_int, _c_int,
                              _c_dbl, _c_dbl, _c_int, _c_dbl, _c_int, _c_flt, _p, _p, _c_int, _p]),
    "sw_step2_accumulate": (_c_int, [_p, _p, _c_int, _p, _p, _p, _c_int, _p, _c_int, _c_int, _c_int,
                                     _c_dbl, _c_dbl, _c_int, _c_dbl, _c_int, _c_flt, _p, _p, _c_int, _p, _c_int,
                                     _p]),
    "sw_force_reduce2": (_c_int, [_p, _c_int, _c_int, _p, _c_int, _c_int, _c_int, _c_int, _p, _p]),
    "sw_set_step2_rows": (_c_int, [_c_int]),
    "sw_get_step2_rows": (_c_int, []),
    "sw_set_step2_tile_order": (_c_int, [_c_int]),
    "sw_set_step2_segments": (_c_int, [_c_int]),
    "sw_get_step2_segments": (_c_int, [_c_int, _c_int, _c_int]),
    "sw_step_variants": (_c_int, []),
    "sw_set_step_variant": (_c_int, [_c_int, _c_int]),
    "sw_get_step_variant": (_c_int, [_c_int]),
    "sw_force_reduce": (_c_int, [_p, _c_int, _c_int, _p, _c_int, _c_int, _c_int, _p, _p]),
    "sw_macro": (_c_int, [_p, _c_int, _p, _c_int, _c_int, _c_int, _c_flt, _p, _p, _p, _p]),
    "sw_shift_table": (_c_int, [_p, _c_int, _c_int, _p, _p]),
    "sw_qcoarse": (_c_int, [_p, _p, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _p, _p]),
    "sw_selftest_division": (_c_int, [_p, _p]),
    "sw_stream_copy": (_c_int, [_p, _p, _c_ll, _p]),
    "sw_noalloc_max_pairs": (_c_ll, []),
    "sw_step_noalloc": (_c_int, [_p, _p, _c_int, _p, _p, _p, _c_int, _p, _c_int, _c_int, _c_int,
                                 _c_dbl, _c_dbl, _c_int, _c_dbl, _c_int, _c_flt, _p, _p, _p]),
    "sw_batch_table_bytes": (_c_ll, [_c_int]),
    "sw_batch_upload": (_c_int, [_p, _c_int, _c_int, _p, _p]),
    "sw_batch_download": (_c_int, [_p, _p, _c_int, _p]),
    "sw_batch_step": (_c_int, [_p, _c_int, _c_int, _c_int, _c_int, _c_int, _p]),
    "sw_batch_force_reduce": (_c_int, [_p, _c_int, _c_int, _p]),
    "sw_batch_advance": (_c_int, [_p, _c_int, _c_int, _p]),
    "sw_batch_graph_create": (_c_int, [_p, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _p]),
    "sw_batch_partials_per_step": (_c_ll, [_c_int, _c_int, _c_int]),
    "sw_batch_step2_rows": (_c_int, []),
    "sw_batch_step2": (_c_int, [_p, _c_int, _c_int, _c_int, _p]),
    "sw_batch_step2_tail": (_c_int, [_p, _c_int, _c_int, _c_int, _p]),
    "sw_batch_force_reduce2": (_c_int, [_p, _c_int, _c_int, _p]),
    "sw_batch_advance2": (_c_int, [_p, _c_int, _c_int, _p]),
    "sw_batch_graph_create2": (_c_int, [_p, _c_int, _c_int, _c_int, _c_int, _p]),
    "sw_graph_launch": (_c_int, [_p, _p]),
    "sw_graph_destroy": (_c_int, [_p]),
    "sw_to_ref_layout": (_c_int, [_p, _p, _c_int, _c_int, _c_int, _c_int, _c_int, _p]),
    "sw_from_ref_layout": (_c_int, [_p, _p, _c_int, _c_int, _c_int, _c_int, _c_int, _p]),
}



class SwInstance(ctypes.Structure):
    """Mirror of struct sw_instance (include/stonewall_lbm.h)."""
    _fields_ = [("f0", _p), ("f1", _p), ("links", _p), ("uin", _p), ("gust", _p), ("partials", _p),
                ("force", _p), ("shift", _p), ("nu0", _c_dbl), ("cs", _c_dbl), ("nu_sponge", _c_dbl), ("scale", _c_flt),
                ("nx", _c_int), ("ny", _c_int), ("nz", _c_int), ("ring", _c_int), ("x_sponge", _c_int),
                ("ybc", _c_int), ("dtype", _c_int), ("t", _c_int), ("t_end", _c_int), ("parity", _c_int),
                ("item0", _c_int), ("tile0", _c_int)]


_lib = None


class SwError(RuntimeError):
    """A C-ABI call returned a non-zero code."""

    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def lib(path=None):
    """Load (once) and return the library with every signature declared."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    # PyTorch-ROCm ships its own HIP runtime; it has to be the one loaded when the library
    # resolves its HIP symbols, else the two runtimes each see no device
    import torch  # noqa: F401
    if not os.path.exists(p):
        raise RuntimeError(f"{p} is not built; run `make -C {os.path.dirname(p)}` "
                           "(or __graft_entry__.build()); there is no fallback compute path")
    L = ctypes.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = L
    return L


def check(rc, what=""):
    if rc != 0:
        msg = lib().sw_last_error().decode(errors="replace")
        raise SwError(rc, f"{what}: {msg}" if what else msg)
    return rc


def call(name, *args):
    return check(getattr(lib(), name)(*args), name)
