"""A .npz writer in np.savez_compressed's format with the deflate work spread over threads.

The reference writes its results with np.savez_compressed (stonewall3d/lbm3d.py:404-423): a
ZIP archive with one deflated `<key>.npy` member per array.  For the H = 40 case that is
≈0.3 GB of snapshot stacks and mean fields, which zlib compresses on one core in ≈9 s, a
quarter of the whole run.  This writer produces the same kind of archive (np.load reads it,
and every member holds the same .npy bytes), but cuts each member into blocks deflated
independently on a thread pool (zlib releases the GIL), in the manner of pigz: every block
but a member's last ends with a sync flush (an empty stored block, byte-aligned, no final
bit), so the blocks concatenate into one valid deflate stream.  The archive is plain ZIP
(no ZIP64), so members and the whole file stay below 4 GiB; larger inputs go to
np.savez_compressed.
"""
import io
import os
import struct
import time
import zlib
from concurrent.futures import ThreadPoolExecutor

import numpy as np

BLOCK = 8 << 20            # bytes of a member deflated as one job
_LIMIT = (1 << 32) - 1


def _deflate(data, last, level):
    c = zlib.compressobj(level, zlib.DEFLATED, -15)
    return c.compress(data) + c.flush(zlib.Z_FINISH if last else zlib.Z_SYNC_FLUSH)


def _dos_time(t):
    lt = time.localtime(t)
    return ((lt.tm_hour << 11) | (lt.tm_min << 5) | (lt.tm_sec // 2),
            (((lt.tm_year - 1980) & 0x7F) << 9) | (lt.tm_mon << 5) | lt.tm_mday)


def _numpy_archive(path, arrays):
    """np.savez_compressed into `path` as given (through a file object, so that numpy does not
    append ".npz" to it)."""
    with open(path, "wb") as f:
        np.savez_compressed(f, **arrays)


def savez_compressed(path, workers=None, level=6, block=BLOCK, **arrays):
    """Write `arrays` to `path` (".npz" is not appended) as np.savez_compressed does, with
    the deflate blocks compressed on `workers` threads (default: the CPU count, at most 16)."""
    members = []
    for key, val in arrays.items():
        buf = io.BytesIO()
        np.lib.format.write_array(buf, np.asanyarray(val), allow_pickle=False)
        members.append((f"{key}.npy".encode(), buf.getbuffer()))
    if not members or sum(len(m) for _, m in members) >= _LIMIT:
        _numpy_archive(path, arrays)
        return
    workers = workers or min(16, os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=workers) as pool:
        jobs = []
        for _, data in members:
            n = len(data)
            starts = list(range(0, n, block)) or [0]
            jobs.append([pool.submit(_deflate, data[s:s + block], s + block >= n, level) for s in starts])
        crcs = [zlib.crc32(data) & 0xFFFFFFFF for _, data in members]
        comp = [[j.result() for j in js] for js in jobs]
    sizes = [sum(len(p) for p in parts) for parts in comp]
    names = sum(len(name) for name, _ in members)
    end = sum(sizes) + 30 * len(members) + names                       # local headers and data
    if max(sizes) >= _LIMIT or end + 46 * len(members) + names >= _LIMIT or len(members) >= 0xFFFF:
        _numpy_archive(path, arrays)                                    # needs ZIP64
        return
    tm, dt = _dos_time(time.time())
    central = []
    with open(path, "wb") as f:
        for (name, data), crc, parts, csize in zip(members, crcs, comp, sizes):
            off = f.tell()
            f.write(struct.pack("<IHHHHHIIIHH", 0x04034B50, 20, 0, 8, tm, dt, crc, csize, len(data), len(name), 0))
            f.write(name)
            for p in parts:
                f.write(p)
            central.append(struct.pack("<IHHHHHHIIIHHHHHII", 0x02014B50, 20, 20, 0, 8, tm, dt, crc, csize,
                                       len(data), len(name), 0, 0, 0, 0, 0o600 << 16, off) + name)
        cd_off = f.tell()
        for c in central:
            f.write(c)
        cd_size = f.tell() - cd_off
        f.write(struct.pack("<IHHHHIIH", 0x06054B50, 0, 0, len(central), len(central), cd_size, cd_off, 0))
