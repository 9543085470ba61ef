"""Host logic of the serving loop with stand-in batches (no GPU): trace generation,
memory estimate, FIFO admission under a byte budget, retirement and reporting."""
import numpy as np
import pytest

from stonewall3d import serving


def test_trace_poisson_and_ranges():
    reqs = serving.trace(400, rate=8.0, seed=3, steps_scale=0.1)
    t = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(t) >= 0)
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert gaps.mean() == pytest.approx(1 / 8.0, rel=0.15)           # Poisson rate
    assert all(10 <= r.H <= 40 and 0.0 <= r.porosity <= 0.45 for r in reqs)
    assert all(200 <= r.steps <= 2000 for r in reqs)                   # U[2000, 20000] scaled by 0.1
    assert {r.dtype for r in reqs} == {"f32", "f16"}


def test_lattice_bytes():
    H = 40
    N = 16 * H * 8 * H * 4 * H
    b32 = serving.lattice_bytes(H, "f32", 1000, 32)
    b16 = serving.lattice_bytes(H, "f16", 1000, 32)
    assert b32 > 2 * 19 * N * 4 and b32 < 2 * 19 * N * 4 * 1.05
    assert b16 < b32 * 0.55
    # 250 GB holds about 50 FP32 H=40 cases
    assert 45 <= 250e9 // b32 <= 51


class FakeLattice:
    def __init__(self, req, K):
        self.tag, self.N, self.n_steps, self.t, self.ny = req, 512 * req.H ** 3, req.steps, 0, 8 * req.H
        self.workload = type("W", (), {"UH": 0.05, "H": req.H})()

    @property
    def done(self):
        return self.t >= self.n_steps

    def force_rows(self):
        return np.ones((self.t, 4 * self.tag.H, 2))


class FakeBatch:
    def __init__(self, K):
        self.K, self.members, self.dtype = K, [], 0
        self.max_resident = 0

    def add(self, lat):
        self.members.append(lat)
        self.max_resident = max(self.max_resident, len(self.members))

    def run_chunk(self):
        for m in self.members:
            m.t = min(m.n_steps, m.t + self.K)

    def retire_finished(self):
        done = [m for m in self.members if m.done]
        self.members = [m for m in self.members if not m.done]
        return done

    def close(self):
        pass


def test_admission_under_budget(monkeypatch):
    K = 10
    reqs = [serving.Request(k, 0.0, 10, 0.2, k, 30 + 10 * k, "f32") for k in range(6)]
    one = serving.lattice_bytes(10, "f32", 80, K)
    srv = serving.Server.__new__(serving.Server)
    srv.budget, srv.K, srv.max_instances, srv.keep_force, srv.used = 2.5 * one, K, 64, False, 0.0
    fb = FakeBatch(K)
    srv.batches, srv.streams = {"f32": fb}, {"f32": None}
    monkeypatch.setattr(srv, "_batch", lambda dtype: fb, raising=False)
    monkeypatch.setattr(srv, "make_lattice", lambda req: FakeLattice(req, K), raising=False)

    class NoStream:
        def synchronize(self):
            pass

    srv.prep_stream = NoStream()
    monkeypatch.setattr(serving.torch.cuda, "stream", lambda s: _Null())
    monkeypatch.setattr(serving, "nat_name", lambda code: "f32")
    results, rep = srv.serve(reqs, realtime=False)
    assert rep["requests"] == 6
    assert fb.max_resident == 2                                   # the budget admits two at a time
    assert sorted(r.rid for r in results) == list(range(6))
    assert all(r.latency >= 0 for r in results)
    assert rep["cell_updates"] == sum(512 * 10 ** 3 * r.steps for r in reqs)


def _fake_server(monkeypatch, budget, K, fb, **attrs):
    srv = serving.Server.__new__(serving.Server)
    srv.budget, srv.K, srv.max_instances, srv.keep_force, srv.used = budget, K, 64, False, 0.0
    for k, v in attrs.items():
        setattr(srv, k, v)
    srv.batches, srv.streams = {"f32": fb}, {"f32": None}
    monkeypatch.setattr(srv, "_batch", lambda dtype: fb, raising=False)
    monkeypatch.setattr(srv, "make_lattice", lambda req: FakeLattice(req, K), raising=False)

    class NoStream:
        def synchronize(self):
            pass

    srv.prep_stream = NoStream()
    monkeypatch.setattr(serving.torch.cuda, "stream", lambda s: _Null())
    monkeypatch.setattr(serving, "nat_name", lambda code: "f32")
    return srv


def test_sjf_with_fill_cap_runs_shortest_first(monkeypatch):
    """With a compute cap of one lattice, "sjf" admits the request with the least work
    first; every request still completes and the cell-update count is exact."""
    K = 10
    steps = [90, 30, 60, 20, 50, 40]
    reqs = [serving.Request(k, 0.0, 10, 0.2, k, s, "f32") for k, s in enumerate(steps)]
    fb = FakeBatch(K)
    srv = _fake_server(monkeypatch, 1e12, K, fb, policy="sjf", fill_cells=512 * 10 ** 3)
    results, rep = srv.serve(reqs, realtime=False)
    assert fb.max_resident == 1
    assert [r.rid for r in results] == sorted(range(6), key=lambda k: steps[k])
    assert rep["cell_updates"] == sum(512 * 10 ** 3 * s for s in steps)


def test_fifo_keeps_arrival_order(monkeypatch):
    K = 10
    steps = [90, 30, 60]
    reqs = [serving.Request(k, 0.0, 10, 0.2, k, s, "f32") for k, s in enumerate(steps)]
    fb = FakeBatch(K)
    srv = _fake_server(monkeypatch, 1e12, K, fb, policy="fifo", fill_cells=512 * 10 ** 3)
    results, rep = srv.serve(reqs, realtime=False)
    assert [r.rid for r in results] == [0, 1, 2]


class _Null:
    def __enter__(self):
        return self

    def __exit__(self, *a):
        return False
