"""Host logic of the serving loop (admission by footprint, FIFO order,
continuous batching, latency accounting) against a fake batch; no GPU."""

from __future__ import annotations

import numpy as np
import pytest

from skyrmion_solver_amd import _engine
from skyrmion_solver_amd.serving import Request, Server, poisson_arrivals


class FakeBatch:
    """Each lattice finishes after `p_i[0]` chunks; at most `slots` resident."""

    def __init__(self, theory, slots, log):
        self.theory, self.slots, self.log = theory, slots, log
        self.lat = {}

    def add(self, p_i, p_f, F0, V0, max_steps, tol, device, init=None, x_chunk=0):
        if len(self.lat) >= self.slots:
            raise _engine.SkxError(-5, "skx_batch_add", "busy")
        s = min(set(range(self.slots)) - set(self.lat))
        self.lat[s] = dict(left=int(p_i[0]), steps=0)
        self.log.append(("add", self.theory, int(p_i[1])))
        return s

    def launch(self, steps):
        for v in self.lat.values():
            if v["left"] > 0:
                v["left"] -= 1
                v["steps"] += steps

    def wait(self):
        pass

    def poll(self):
        return [dict(slot=s, done=int(v["left"] == 0), steps=v["steps"]) for s, v in self.lat.items()]

    def remove(self, slot, want_fields=True):
        v = self.lat.pop(slot)
        return dict(steps=v["steps"], converged=1, energy=1.0), None, None

    def close(self):
        pass


def make_requests(n, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        chunks = int(rng.integers(1, 6))
        p_i = np.array([chunks, k, 1, 2, 3, 4, 1000, 4000], dtype=np.int32)
        out.append(Request(rid=k, theory=["pion", "omega"][k % 2], p_i=p_i, p_f=np.zeros(10), F0=np.zeros(4),
                           arrival=0.0))
    return out


def test_admission_respects_budget_and_fifo():
    log = []
    fp = {k: 100 + 10 * (k % 3) for k in range(12)}
    srv = Server(hbm_budget_bytes=350, chunk_steps=8, batch_factory=lambda t: FakeBatch(t, 8, log),
                 footprint_fn=lambda r: fp[r.rid])
    reqs = make_requests(12)
    for r in reqs:
        srv.submit(r)
    order = []
    while not srv.idle():
        srv.step()
        assert srv.used <= 350
        order = [e[2] for e in log]
    assert order == sorted(order)  # admitted in arrival order
    assert len(srv.done) == 12 and srv.used == 0
    assert srv.peak_used <= 350


def test_busy_batch_defers_admission():
    log = []
    srv = Server(hbm_budget_bytes=10 ** 9, chunk_steps=4, batch_factory=lambda t: FakeBatch(t, 1, log),
                 footprint_fn=lambda r: 1)
    reqs = [r for r in make_requests(6) if r.theory == "pion"]
    for r in reqs:
        srv.submit(r)
    srv.step()
    assert len(srv.running) == 1 and len(srv.queue) == len(reqs) - 1
    while not srv.idle():
        srv.step()
    assert [r.rid for r in srv.done] == [r.rid for r in reqs]


def test_late_arrivals_join_running_work():
    log = []
    t = [0.0]

    def clock():  # fake time: every read advances 2 ms
        t[0] += 0.002
        return t[0]

    srv = Server(hbm_budget_bytes=10 ** 9, chunk_steps=4, batch_factory=lambda th: FakeBatch(th, 8, log),
                 footprint_fn=lambda r: 1, clock=clock)
    reqs = make_requests(8, seed=3)
    arr = poisson_arrivals(8, rate=50.0, seed=1)
    for r, a in zip(reqs, arr):
        r.arrival = float(a)
    rep = srv.serve(reqs, poll_sleep=0.0)
    assert rep["requests"] == 8
    overlapped = any(a.admitted < b.finished and b.admitted < a.finished
                     for a in srv.done for b in srv.done if a is not b)
    assert overlapped
    assert all(r.latency >= 0 for r in srv.done)
    assert rep["p99_s"] >= rep["p50_s"]


def test_request_larger_than_budget_is_rejected():
    srv = Server(hbm_budget_bytes=10, batch_factory=lambda t: FakeBatch(t, 4, []), footprint_fn=lambda r: 11)
    srv.submit(make_requests(1)[0])
    with pytest.raises(ValueError):
        srv.step()


def test_progress_lines_report_counts(capsys):
    t = [0.0]

    def clock():  # fake time: every read advances 10 ms
        t[0] += 0.01
        return t[0]

    srv = Server(hbm_budget_bytes=10 ** 9, chunk_steps=4, batch_factory=lambda th: FakeBatch(th, 2, []),
                 footprint_fn=lambda r: 1, clock=clock)
    reqs = make_requests(6, seed=5)
    rep = srv.serve(reqs, poll_sleep=0.0, progress_every=0.05)
    err = capsys.readouterr().err.splitlines()
    notes = [ln for ln in err if ln.startswith("[serve]")]
    assert rep["requests"] == 6 and notes
    for ln in notes:  # finished + running + queued never exceeds the requests that have arrived
        f = ln.split()
        done, total = map(int, f[3].split("/"))
        assert total == 6 and 0 <= done <= 6 and int(f[5]) <= 4 and int(f[7]) <= 6
    srv2 = Server(hbm_budget_bytes=10 ** 9, chunk_steps=4, batch_factory=lambda th: FakeBatch(th, 2, []),
                  footprint_fn=lambda r: 1, clock=clock)
    srv2.serve(make_requests(3, seed=6), poll_sleep=0.0)
    assert not [ln for ln in capsys.readouterr().err.splitlines() if ln.startswith("[serve]")]


class FloorBatch:
    """FP32 stand-in: max|g| halves every chunk but never goes below 1.5 tol (a rounding floor);
    FP64 stand-in (`floor=False`): halves with no floor and converges below tol."""

    def __init__(self, floor, log):
        self.floor, self.log, self.lat = floor, log, {}

    def add(self, p_i, p_f, F0, V0, max_steps, tol, device, init=None, x_chunk=0):
        s = min(set(range(8)) - set(self.lat))
        err = float(F0[1]) if F0 is not None and not self.floor else 1.0
        self.lat[s] = dict(rid=int(p_i[1]), err=err, tol=tol, steps=0, max=max_steps, F0=F0, V0=V0)
        self.log.append(("add", "fp32" if self.floor else "fp64", int(p_i[1]), None if F0 is None else F0.copy()))
        return s

    def launch(self, steps):
        for v in self.lat.values():
            if v["err"] >= v["tol"] and v["steps"] < v["max"]:
                v["steps"] += steps
                v["err"] = max(v["err"] / 2, 1.5 * v["tol"]) if self.floor else v["err"] / 2

    def wait(self):
        pass

    def poll(self):
        return [dict(slot=s, steps=v["steps"], err_abs=v["err"],
                     done=int(v["err"] < v["tol"] or v["steps"] >= v["max"])) for s, v in self.lat.items()]

    def remove(self, slot, want_fields=True):
        v = self.lat.pop(slot)
        F = np.array([v["rid"], v["err"], 7.0])  # the state handed on: rid, the force reached
        return dict(steps=v["steps"], converged=int(v["err"] < v["tol"]), energy=1.0), F, F * 2


@pytest.mark.parametrize("refine", [True, False])
def test_fp32_lattices_finish_in_fp64(refine):
    log = []
    t = [0.0]

    def clock():
        t[0] += 0.001
        return t[0]

    srv = Server(precision=_engine.FP32, hbm_budget_bytes=10 ** 9, chunk_steps=4, footprint_fn=lambda r: 1,
                 batch_factory=lambda th: FloorBatch(True, log), refine_fp64=refine, refine_factor=3.0,
                 refine_batch_factory=lambda th: FloorBatch(False, log), refine_footprint_fn=lambda r: 2, clock=clock)
    reqs = make_requests(5, seed=7)
    for r in reqs:
        r.tol, r.max_steps = 1.0e-4, 400
    rep = srv.serve(reqs, poll_sleep=0.0)
    assert rep["requests"] == 5
    if not refine:
        assert rep["refined"] == 0 and rep["converged"] == 0
        assert all(r.status["steps"] == 400 for r in reqs)
        assert not [e for e in log if e[1] == "fp64"]
        return
    assert rep["refined"] == 5 and rep["converged"] == 5
    for r in reqs:
        # handed off once its force fell below 3 tol: 1 -> 2^-12 < 3e-4 after 12 chunks of 4 steps
        assert r.refined and r.steps_fp32 == 48
        assert r.status["steps_fp32"] == 48 and r.status["steps"] > 48 and r.status["converged"]
        add64 = [e for e in log if e[1] == "fp64" and e[2] == r.rid]
        assert len(add64) == 1 and add64[0][3][0] == r.rid  # the FP32 field went on to the FP64 batch
    assert srv.used == 0


def test_refinement_skips_lattices_whose_fp64_footprint_exceeds_the_budget():
    log = []
    t = [0.0]

    def clock():
        t[0] += 0.001
        return t[0]

    srv = Server(precision=_engine.FP32, hbm_budget_bytes=10, chunk_steps=4, footprint_fn=lambda r: 6,
                 batch_factory=lambda th: FloorBatch(True, log), refine_fp64=True,
                 refine_batch_factory=lambda th: FloorBatch(False, log), refine_footprint_fn=lambda r: 11, clock=clock)
    reqs = make_requests(2, seed=8)
    for r in reqs:
        r.tol, r.max_steps = 1.0e-4, 200
    rep = srv.serve(reqs, poll_sleep=0.0)  # returns: nothing waits for an FP64 slot that can never fit
    assert rep["requests"] == 2 and rep["refined"] == 0
    assert all(r.status["steps"] == 200 for r in reqs)


def test_request_larger_than_its_batch_grid_is_rejected():
    class TooBig(FakeBatch):
        def add(self, *a, **k):
            raise _engine.SkxError(-1, "skx_batch_add", "the lattice needs 99 blocks")

    srv = Server(hbm_budget_bytes=10 ** 9, batch_factory=lambda t: TooBig(t, 4, []), footprint_fn=lambda r: 1)
    srv.submit(make_requests(1)[0])
    with pytest.raises(ValueError):
        srv.step()
