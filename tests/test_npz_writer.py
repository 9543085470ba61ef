"""The threaded .npz writer (stonewall3d.npz) against np.savez_compressed: np.load reads back
the same arrays (its ZIP reader checks every member's CRC-32), zipfile finds no damaged
member, and each member holds the same .npy bytes as numpy's own archive; members of many
deflate blocks (the blocks concatenate into one stream) and of one, empty and 0-d arrays,
and every dtype the CLI writes."""
import json
import zipfile

import numpy as np

from stonewall3d import npz


def arrays():
    rng = np.random.default_rng(3)
    x = np.linspace(0, 20, 400)[None, None, :, None]
    z = np.linspace(0, 5, 160)[None, None, None, :]
    smooth = (np.sin(x + np.arange(7)[:, None, None, None] * 0.1) * np.cos(z)).astype(np.float16)
    return dict(Cd=rng.standard_normal(1000).astype(np.float32), snaps=np.broadcast_to(smooth, (7, 3, 400, 160)).copy(),
                noise=rng.standard_normal((300, 1000)).astype(np.float32), wall=rng.random((50, 40, 30)) < 0.3,
                q=rng.integers(0, 255, (5, 20, 10, 8), dtype=np.uint8), empty=np.zeros((0, 3), np.float32),
                scalar=np.float64(2.5), meta=json.dumps(dict(case="p20", H=40)))


def test_npz_roundtrip_many_blocks(tmp_path):
    a = arrays()
    p = tmp_path / "fast.npz"
    npz.savez_compressed(str(p), workers=4, block=1 << 16, **a)    # small blocks: many per member
    with zipfile.ZipFile(p) as z:
        assert z.testzip() is None
        assert sorted(z.namelist()) == sorted(f"{k}.npy" for k in a)
        assert all(i.compress_type == zipfile.ZIP_DEFLATED for i in z.infolist())
    ref = tmp_path / "ref.npz"
    np.savez_compressed(str(ref), **a)
    with np.load(p) as d, np.load(ref) as r:
        for k in a:
            assert d[k].dtype == r[k].dtype and d[k].shape == r[k].shape
            assert np.array_equal(d[k], r[k])
    with zipfile.ZipFile(p) as z, zipfile.ZipFile(ref) as zr:
        for k in a:
            assert z.read(f"{k}.npy") == zr.read(f"{k}.npy")
    assert str(np.load(p)["meta"]) == a["meta"]


def test_npz_one_block_per_member(tmp_path):
    a = arrays()
    p = tmp_path / "one.npz"
    npz.savez_compressed(str(p), workers=1, **a)
    with np.load(p) as d:
        for k, v in a.items():
            assert np.array_equal(d[k], v)
    # smooth fields compress
    assert p.stat().st_size < sum(np.asarray(v).nbytes for v in a.values())


def test_npz_path_kept_on_every_path(tmp_path):
    """The archive goes to the path as given, without ".npz" appended: by the threaded writer
    and by numpy's writer, which takes the inputs the threaded one leaves to it (no arrays)."""
    a = arrays()
    p = tmp_path / "fast.out"
    npz.savez_compressed(str(p), **a)
    with np.load(p) as d:
        assert sorted(d.files) == sorted(a)
    q = tmp_path / "empty.out"
    npz.savez_compressed(str(q))
    assert q.exists() and not (tmp_path / "empty.out.npz").exists()
    with np.load(q) as d:
        assert d.files == []
