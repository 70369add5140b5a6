"""The serving layer's host logic: the workload generator (seeded, within
the stated ranges, Poisson arrivals with the asked rate) and the engine's
admission and continuous batching, with the device side replaced by a model
that counts iterations."""
import time
import types

import numpy as np
import pytest

from fvm2d_amd import serving


def test_workload_ranges_and_seed():
    a = serving.workload(2000, 11)
    b = serving.workload(2000, 11)
    assert [(r.nj, r.mach, r.seed) for r in a] == [(r.nj, r.mach, r.seed) for r in b]
    npv = np.array([r.np for r in a])
    assert npv.min() >= 900 and npv.max() <= 1.01e6
    # log-uniform: the median is near the geometric mean of the bounds
    assert 2e4 < np.median(npv) < 5e4
    for r in a:
        assert r.ni - 1 == 4 * (r.nj - 1)
        assert 0.0 <= r.m <= 0.04 and 0.2 <= r.p <= 0.6 and 0.09 <= r.t <= 0.18
        assert 0.3 <= r.mach <= 0.85 and -4.0 <= r.aoa <= 8.0 and 1e6 <= r.rey <= 1e7
    frac = np.mean([r.flow == "rans" for r in a])
    assert abs(frac - 0.3) < 0.04


def test_poisson_arrivals():
    reqs = serving.poisson_arrivals(serving.workload(4000, 3), rate=50.0, seed=3)
    t = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(t) >= 0)
    gaps = np.diff(t)
    assert abs(gaps.mean() - 1 / 50.0) < 0.1 / 50.0
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.1          # exponential gaps
    closed = serving.poisson_arrivals(serving.workload(5, 3), rate=0.0, seed=3)
    assert all(r.arrival == 0.0 for r in closed)


class ModelSolver:
    """Stands in for a context: counts iterations up to the cap."""
    live = 0

    def __init__(self, req, cfg):
        self.np = req.np
        self.cap = cfg.max_iters
        self.it = 0
        ModelSolver.live += 1

    def monitor(self):
        st = 2 if self.it >= self.cap else 0
        return dict(iter=self.it, fres=1.0 / (1 + self.it), cl=0.1 * self.np, cd=0.0,
                    iresi=1, fresi=0.0), st

    def close(self):
        ModelSolver.live -= 1


class ModelBatch:
    def __init__(self):
        self.m = []
        self.n = None

    def add(self, s):
        assert self.n is None
        self.m.append(s)

    def remove(self, s):
        assert self.n is None
        self.m.remove(s)

    def launch(self, n):
        assert self.n is None and n % 8 == 0
        self.n = n
        self.polls = 0

    def done(self):
        # the chunk "runs" for two polls, so host work interleaves with it
        self.polls += 1
        return self.polls > 2

    def wait(self):
        time.sleep(0.001)                       # a step takes time
        for s in self.m:
            s.it = min(s.cap, s.it + self.n)
        self.n = None
        return [s.monitor()[1] for s in self.m]

    def close(self):
        pass


@pytest.fixture
def model(monkeypatch):
    def prepare(req, cfg):
        return None, types.SimpleNamespace(dims=dict(np=req.np, nt=2 * req.np, netot=3 * req.np))
    monkeypatch.setattr(serving, "prepare", prepare)
    monkeypatch.setattr(serving, "footprint", lambda geo, req, cfg: 1000 * geo.dims["np"])
    monkeypatch.setattr(serving, "create", lambda req, g, geo, cfg: ModelSolver(req, cfg))
    monkeypatch.setattr(serving, "Batch", ModelBatch)
    ModelSolver.live = 0


@pytest.mark.parametrize("rate", [0.0, 2000.0])
def test_admission_holds_the_budget(model, rate):
    reqs = serving.poisson_arrivals(serving.workload(40, 9, np_min=600, np_max=6000), rate, 9)
    budget = 2 * 1000 * max(r.np for r in reqs)
    cfg = serving.ServeConfig(max_iters=64, step_vertex_iters=2e4,
                              size_classes=(2500,), hbm_budget=budget)
    srv = serving.Server(cfg)
    res = srv.run(reqs, timeout=60)
    assert [r.rid for r in res] == list(range(40)) and not srv.errors
    assert all(r.iters == 64 and r.status == 2 for r in res)
    assert srv.peak_reserved <= budget
    assert 1 < srv.peak_in_flight < 40                # shared, and admission had to wait
    assert srv.budget == budget and ModelSolver.live == 0
    for r in res:
        assert r.arrival <= r.prepared <= r.admitted <= r.joined <= r.finished


def test_requests_join_running_batches(model):
    """Continuous batching: late arrivals join classes that are running."""
    reqs = serving.poisson_arrivals(serving.workload(30, 4, np_min=600, np_max=2000), 400.0, 4)
    cfg = serving.ServeConfig(max_iters=400, step_vertex_iters=1e4,
                              size_classes=(10**9,))
    srv = serving.Server(cfg)
    res = srv.run(reqs, timeout=60)
    assert len(res) == 30
    # some request joined after another had started and before it finished
    assert any(a.joined < b.joined < a.finished for a in res for b in res if a is not b)


def test_engine_timeout(model):
    reqs = serving.workload(3, 1, np_min=600, np_max=700)
    cfg = serving.ServeConfig(max_iters=10**9, step_vertex_iters=1e3)
    with pytest.raises(TimeoutError):
        serving.Server(cfg).run(reqs, timeout=0.5)


def test_precision_per_flow_kind():
    """Euler and SA-RANS requests are served in split precision by default;
    each flow kind's precision can be set on its own."""
    reqs = serving.workload(40, 3)
    cfg = serving.ServeConfig()
    for q in reqs:
        assert serving.precision_for(q, cfg) == 1
        assert serving.inputs_for(q, cfg).precision == 1
    cfg2 = serving.ServeConfig(precision=0, rans_precision=1)
    assert {serving.precision_for(q, cfg2) for q in reqs} == {0, 1}
