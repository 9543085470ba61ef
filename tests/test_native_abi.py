"""The C-ABI library loads, exports every symbol include/*.h declares, and rejects bad
arguments with library error codes before touching a device (runs without a GPU)."""
import ctypes
import glob
import os
import re

import pytest

from stonewall3d import _native as nat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(sw_\w+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 11
    L = ctypes.CDLL(nat.LIB_PATH)
    for n in names:
        assert hasattr(L, n), n
        assert n in nat.SIGNATURES, f"{n} has no ctypes signature"
    assert set(nat.SIGNATURES) <= set(names)


def test_version_and_sizes():
    L = nat.lib()
    assert L.sw_version() == 120
    assert L.sw_rows_per_block() == 4
    assert L.sw_partials_per_step(320, 160) == 160 * 80 * 2
    assert L.sw_partials_per_step(321, 3) == 3 * 81 * 2
    assert L.sw_partials_per_step(0, 3) == -1


def test_argument_errors_without_device():
    L = nat.lib()
    # null buffers
    rc = L.sw_step(None, None, 0, None, None, None, 0, None, 8, 8, 8, 1e-3, 0.16, 8, 0.05, 0, 1.0, None, None, 1,
                   None)
    assert rc == 1001
    assert b"non-null" in L.sw_last_error()
    fake = ctypes.c_void_p(0x1000)
    # bad dims / dtype / scale / ybc / too large
    args = dict(nx=8, ny=8, nz=8, dtype=0, scale=1.0, ybc=0)

    def step(**kw):
        a = dict(args, **kw)
        return L.sw_step(fake, ctypes.c_void_p(0x2000), a["dtype"], fake, fake, None, 0, None, a["nx"], a["ny"],
                         a["nz"], 1e-3, 0.16, 8, 0.05, a["ybc"], a["scale"], None, None, 1, None)

    assert step(nx=1) == 1001
    assert step(dtype=7) == 1001
    assert step(dtype=1, scale=3.0) == 1001          # 16-bit scale must be a power of two
    assert step(dtype=1, scale=2.0 ** 21) == 1001    # ... in [2^-20, 2^20]
    assert b"[2^-20, 2^20]" in L.sw_last_error()
    assert step(ybc=5) == 1001
    assert step(nx=2000, ny=2000, nz=100) == 1003
    assert b"113,025,455" in L.sw_last_error()
    # in-place is rejected (pull scheme needs two buffers)
    assert L.sw_step(fake, fake, 0, fake, fake, None, 0, None, 8, 8, 8, 1e-3, 0.16, 8, 0.05, 0, 1.0, None, None, 1,
                     None) == 1001
    # partials without a ring
    assert L.sw_step(fake, ctypes.c_void_p(0x2000), 0, fake, fake, None, 0, None, 8, 8, 8, 1e-3, 0.16, 8, 0.05, 0,
                     1.0, None, fake, 0, None) == 1001
    # offset tables: unknown kind, missing inlet profile
    assert L.sw_shift_table(fake, 8, 2, fake, None) == 1001
    assert L.sw_shift_table(None, 8, 1, fake, None) == 1001
    assert L.sw_shift_table(fake, 0, 1, fake, None) == 1001
    assert L.sw_force_reduce(fake, 4, 0, None, 5, 8, 8, fake, None) == 1001
    assert L.sw_to_ref_layout(fake, fake, 4, 19, 8, 8, 8, None) == 1001
    assert L.sw_to_ref_layout(fake, ctypes.c_void_p(0x2000), 3, 19, 8, 8, 8, None) == 1001


def test_check_raises_with_message():
    with pytest.raises(nat.SwError) as e:
        nat.call("sw_macro", None, 0, None, 8, 8, 8, 1.0, None, None, None, None)
    assert e.value.code == 1001


def test_two_step_sizes_and_argument_errors_without_device():
    """Host-side rules of the two-step entry points: partials sizes, the tile-rows setting,
    the batch table size, and argument errors raised before any device work."""
    L = nat.lib()
    # the two-step batch's tiles: 64 x 10 cells (12 rows with their halo rows); a lattice's
    # ring holds either layout
    assert L.sw_batch_step2_rows() == 12
    assert L.sw_batch_partials_per_step(640, 320, 160) == 160 * (10 * 32) * 2
    assert L.sw_batch_partials_per_step(128, 64, 32) == 32 * 16 * 2          # the row groups are more
    assert L.sw_batch_partials_per_step(0, 8, 8) == -1
    inst = ctypes.sizeof(nat.SwInstance)
    assert inst % 8 == 0
    assert L.sw_batch_table_bytes(4) == 64 + 4 * (inst + 2 * 136)           # + two argument records each
    # tile rows: 0 (automatic), 12, 13, 15 or 16
    assert L.sw_get_step2_rows() == 0
    assert L.sw_set_step2_rows(17) == 1001
    assert b"15 or 16" in L.sw_last_error()
    assert L.sw_set_step2_rows(14) == 1001
    assert L.sw_set_step2_rows(12) == 0 and L.sw_get_step2_rows() == 12
    assert L.sw_set_step2_rows(16) == 0 and L.sw_get_step2_rows() == 16
    assert L.sw_set_step2_rows(0) == 0 and L.sw_get_step2_rows() == 0
    fake = ctypes.c_void_p(0x1000)
    args = (0, fake, fake, None, 0, None, 8, 8, 8, 1e-3, 0.16, 8, 0.05, 0, 1.0, None, None, 2, None)
    assert L.sw_step2_one(None, ctypes.c_void_p(0x2000), *args) == 1001
    assert L.sw_step2_one(fake, fake, *args) == 1001                         # f_in == f_out
    f64 = (3,) + args[1:]
    assert L.sw_step2_one(fake, ctypes.c_void_p(0x2000), *f64) == 1001
    assert b"FP32, FP16 or BF16" in L.sw_last_error()
    acc_args = args[:-1]                                                    # up to ring; then acc, which, stream
    assert L.sw_step2_accumulate(fake, ctypes.c_void_p(0x2000), *acc_args, None, 1, None) == 1001   # acc missing
    assert L.sw_step2_accumulate(fake, ctypes.c_void_p(0x2000), *acc_args, fake, 3, None) == 1001   # which
    assert L.sw_step2_accumulate(fake, ctypes.c_void_p(0x2000), *((1,) + acc_args[1:]), fake, 1, None) == 1001
    assert b"FP32" in L.sw_last_error()
    assert L.sw_batch_step2(None, 0, 0, 256, None) == 1001
    assert L.sw_batch_step2(fake, 3, 0, 256, None) == 1001                   # FP64
    assert L.sw_batch_step2_tail(fake, 0, 7, 256, None) == 1001              # K odd
    assert L.sw_batch_advance2(fake, 4, 7, None) == 1001
    g = ctypes.c_void_p()
    assert L.sw_batch_graph_create2(fake, 4, 0, 7, 256, ctypes.byref(g)) == 1001
    assert b"even" in L.sw_last_error()
