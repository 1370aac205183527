"""Continuous-batching minimisation server on one GPU.

Requests (one lattice each: theory, packed parameters, initial field,
tolerance) arrive over time.  The server

* admits queued requests in arrival order while their device footprint
  (``skx_footprint_bytes``) fits the HBM budget and their theory's batch has
  a free slot and enough free blocks (head-of-line order is kept: a request
  that does not fit blocks the ones behind it);
* keeps every admitted lattice resident on the device until it converges
  (stage-4 max|g| < tol) or hits its step limit;
* advances all resident lattices in chunks of steps, one batched launch per
  RK4 stage per theory (``skx_batch_launch``), the batches of different
  theories on their own streams so their work overlaps;
* admits newly arrived requests between chunks, so requests that arrive
  while others run join them on the device (continuous batching);
* reports time-to-converged (finish - arrival) p50 / p99 and throughput.

FP32 refinement (``refine_fp64=True``, FP32 servers): the FP32 form's force
has a rounding floor near 1e-4 on stiff solutions, so a lattice whose
max|g| has fallen below ``refine_factor * tol`` leaves its FP32 batch with
its field and velocity and finishes in an FP64 batch of its theory.  The
hand-off waits in a queue that is admitted before new requests.

A lattice's result does not depend on what it shares the device with: each
one owns its buffers, reductions and status, and the batched kernels compute
every site exactly as a solo run does.
"""

from __future__ import annotations

import sys
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from skyrmion_solver_amd import _engine

THEORY_IDS = {"pion": _engine.THEORY_PION, "omega": _engine.THEORY_OMEGA, "coulomb": _engine.THEORY_COULOMB,
              "rho": _engine.THEORY_RHO, "berger": _engine.THEORY_BERGER, "lightly": _engine.THEORY_LIGHTLY_BOUND}


@dataclass
class Request:
    rid: int
    theory: str
    p_i: np.ndarray
    p_f: np.ndarray
    F0: np.ndarray
    tol: float = 1.0e-4
    max_steps: int = 20000
    arrival: float = 0.0
    V0: np.ndarray | None = None
    init: object = None  # _engine.InitDesc: build the field on the device at admission (F0 None)
    # filled in by the server
    admitted: float | None = None
    finished: float | None = None
    status: dict = field(default_factory=dict)
    F: np.ndarray | None = None
    V: np.ndarray | None = None
    steps_fp32: int = 0  # with FP32 refinement: steps taken in the FP32 form before the hand-off
    refined: bool = False

    @property
    def sites(self) -> int:
        return int(self.p_i[6])

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


class Server:
    """Admission by HBM footprint + continuous batching over ``skx_batch`` handles."""

    def __init__(self, precision: int = _engine.FP32, hbm_budget_bytes: int = 64 << 30, chunk_steps: int = 32,
                 max_slots: int = 64, max_blocks: int = 32768, tile_z: int = 0, device: int = 0,
                 keep_fields: bool = True, batch_factory=None, footprint_fn=None, clock=time.perf_counter,
                 refine_fp64: bool = False, refine_factor: float = 3.0, refine_batch_factory=None,
                 refine_footprint_fn=None, x_chunk_target: int = 32, x_chunk_blocks: int = 256):
        self.precision = precision
        self.budget = int(hbm_budget_bytes)
        self.chunk = int(chunk_steps)
        self.max_slots, self.max_blocks, self.tile_z, self.device = max_slots, max_blocks, tile_z, device
        self.keep_fields = keep_fields
        self._batch_factory = batch_factory or (lambda theory: _engine.Batch(
            THEORY_IDS[theory], precision, max_slots=max_slots, max_blocks=max_blocks, tile_z=tile_z, device=device))
        self._footprint = footprint_fn or (lambda r: _engine.footprint_bytes(
            THEORY_IDS[r.theory], precision, r.p_i, r.p_f, tile_z=tile_z))
        # planes per block along x for each admitted lattice: about `x_chunk_blocks` blocks per lattice
        # with at least 8 planes per block (``autotune.chunk_for_blocks``), or, with x_chunk_blocks 0,
        # about `x_chunk_target` planes per block (``autotune.chunk_rule``; 0: the whole lattice).  A
        # few small lattices (the omega requests at 64^3 that run 50k steps) then still spread over the
        # CUs: on the bench's serving sample (8 requests, rate 20/s, FP32 refined in FP64) the makespan
        # was 21.5 s with 32 planes per block, 14.5 s with 16, 11.8 s with 8 and 11.4 s with 256 blocks
        # per lattice, with the same steps and results per request
        self.x_chunk_target = int(x_chunk_target)
        self.x_chunk_blocks = int(x_chunk_blocks)
        self.refine = bool(refine_fp64) and precision == _engine.FP32
        self.refine_factor = float(refine_factor)
        self._refine_factory = refine_batch_factory or (lambda theory: _engine.Batch(
            THEORY_IDS[theory], _engine.FP64, max_slots=max_slots, max_blocks=max_blocks, tile_z=tile_z,
            device=device))
        self._refine_footprint = refine_footprint_fn or (lambda r: _engine.footprint_bytes(
            THEORY_IDS[r.theory], _engine.FP64, r.p_i, r.p_f, tile_z=tile_z))
        self.clock = clock
        self.t0 = clock()
        self.batches: dict = {}  # (kind, theory) -> batch; kind 0 = the server's form, 1 = FP64 refinement
        self.queue: deque = deque()
        self.handoff: deque = deque()  # (request, F, V, steps left) waiting for an FP64 slot
        self.running: dict = {}  # (kind, theory, slot) -> (request, footprint)
        self.used = 0
        self.done: list = []
        self.site_updates = 0
        self.peak_used = 0

    def now(self) -> float:
        return self.clock() - self.t0

    def submit(self, req: Request):
        self.queue.append(req)

    def x_chunk(self, req) -> int:
        """Planes per block along x with which a request's lattice is admitted (0: the whole lattice)."""
        from skyrmion_solver_amd.autotune import chunk_for_blocks, chunk_rule

        X, Y, Z = (int(v) for v in req.p_i[:3])
        if self.x_chunk_blocks > 0:
            return chunk_for_blocks(X, Y, Z, self.x_chunk_blocks, self.tile_z)
        return chunk_rule(X, self.x_chunk_target)

    def _batch(self, theory: str, kind: int = 0):
        if (kind, theory) not in self.batches:
            make = self._batch_factory if kind == 0 else self._refine_factory
            self.batches[(kind, theory)] = make(theory)
        return self.batches[(kind, theory)]

    def _admit_handoffs(self) -> bool:
        """Admit FP32 lattices waiting to finish in FP64; False while the head does not fit."""
        while self.handoff:
            req, F, V, left = self.handoff[0]
            fp = self._refine_footprint(req)
            if self.used + fp > self.budget:
                return False
            try:
                slot = self._batch(req.theory, 1).add(req.p_i, req.p_f, F, V, max_steps=left, tol=req.tol,
                                                      device=self.device, x_chunk=self.x_chunk(req))
            except _engine.SkxError as e:
                if e.code == -5:
                    return False
                raise
            self.handoff.popleft()
            self.running[(1, req.theory, slot)] = (req, fp)
            self.used += fp
            self.peak_used = max(self.peak_used, self.used)
        return True

    def admit(self):
        if not self._admit_handoffs():
            return
        while self.queue:
            req = self.queue[0]
            fp = self._footprint(req)
            if fp > self.budget:
                raise ValueError(f"request {req.rid} needs {fp} bytes, more than the whole budget")
            if self.used + fp > self.budget:
                return
            try:
                slot = self._batch(req.theory).add(req.p_i, req.p_f, req.F0, req.V0, max_steps=req.max_steps,
                                                   tol=req.tol, device=self.device, init=req.init,
                                                   x_chunk=self.x_chunk(req))
            except _engine.SkxError as e:
                if e.code == -5:  # SKX_EBUSY: no slot / blocks free in this batch now
                    return
                if e.code == -1:  # SKX_EINVAL: e.g. more blocks than the batch's whole grid
                    raise ValueError(f"request {req.rid} cannot be admitted: {e}") from e
                raise
            self.queue.popleft()
            req.admitted = self.now()
            self.running[(0, req.theory, slot)] = (req, fp)
            self.used += fp
            self.peak_used = max(self.peak_used, self.used)

    def step(self):
        """Admit, advance every resident lattice by one chunk, retire the finished ones."""
        self.admit()
        active = sorted({(k, t) for (k, t, _) in self.running})
        for key in active:  # enqueue every batch's chunk first: their streams overlap
            self.batches[key].launch(self.chunk)
        for key in active:
            self.batches[key].wait()
        for k, t in active:
            b = self.batches[(k, t)]
            for st in b.poll():
                key = (k, t, st["slot"])
                if key not in self.running:
                    continue
                if st["done"]:
                    req, fp = self.running.pop(key)
                    final, F, V = b.remove(st["slot"], want_fields=self.keep_fields)
                    req.finished = self.now()
                    self.site_updates += req.sites * int(final["steps"])
                    if req.refined:  # the FP64 part; report the whole run
                        final = dict(final, steps_fp32=req.steps_fp32, steps=req.steps_fp32 + int(final["steps"]))
                    req.status = final
                    req.F, req.V = F, V
                    self.used -= fp
                    self.done.append(req)
                elif k == 0 and self.refine and st["steps"] > 0 and st["err_abs"] < self.refine_factor * \
                        self.running[key][0].tol and self._refine_footprint(self.running[key][0]) <= self.budget:
                    # (a lattice whose FP64 footprint exceeds the whole budget stays in FP32)
                    req, fp = self.running.pop(key)
                    final, F, V = b.remove(st["slot"], want_fields=True)
                    self.used -= fp
                    req.steps_fp32 = int(final["steps"])
                    req.refined = True
                    self.site_updates += req.sites * req.steps_fp32
                    self.handoff.append((req, F, V, max(req.max_steps - req.steps_fp32, 1)))

    def idle(self) -> bool:
        return not self.queue and not self.running and not self.handoff

    def serve(self, requests, poll_sleep: float = 1.0e-4, progress_every: float = 0.0):
        """Replay `requests` (their `arrival` in seconds from now) until all have finished.  With
        `progress_every` > 0, a line on stderr every that many seconds gives the finished, running and
        queued counts."""
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        total = len(pending)
        self.t0 = self.clock()
        next_note = progress_every
        while pending or not self.idle():
            now = self.now()
            if progress_every > 0 and now >= next_note:
                print(f"[serve] t={now:.0f}s finished {len(self.done)}/{total} running {len(self.running)} "
                      f"queued {len(self.queue)}", file=sys.stderr, flush=True)
                next_note = now + progress_every
            while pending and pending[0].arrival <= now:
                self.submit(pending.popleft())
            if self.idle():
                time.sleep(min(poll_sleep, max(0.0, pending[0].arrival - now)))
                continue
            self.step()
        return self.report()

    def report(self) -> dict:
        """Time-to-result (arrival -> result) p50 / p99 over all finished requests and over the converged
        ones, throughput, and per-theory counts (converged, stopped at max_steps, non-finite)."""
        lat = np.array([r.latency for r in self.done]) if self.done else np.zeros(1)
        conv = np.array([r.latency for r in self.done if r.status.get("converged")])
        span = max((r.finished for r in self.done), default=0.0) - min((r.arrival for r in self.done), default=0.0)
        by = {}
        for r in self.done:
            t = by.setdefault(r.theory, dict(requests=0, converged=0, nonfinite=0, steps=[]))
            t["requests"] += 1
            t["converged"] += int(bool(r.status.get("converged")))
            t["nonfinite"] += int(bool(r.status.get("nonfinite")))
            t["steps"].append(int(r.status.get("steps", 0)))
        for t in by.values():
            t["median_steps"] = int(np.median(t.pop("steps")))
        return dict(requests=len(self.done), p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)),
                    mean_s=float(lat.mean()),
                    converged_p50_s=float(np.percentile(conv, 50)) if conv.size else None,
                    converged_p99_s=float(np.percentile(conv, 99)) if conv.size else None,
                    makespan_s=float(span), site_updates_per_s=self.site_updates / span if span > 0 else 0.0,
                    converged=int(sum(int(bool(r.status.get("converged"))) for r in self.done)),
                    by_theory=by, peak_hbm_bytes=self.peak_used,
                    refined=int(sum(int(r.refined) for r in self.done)))

    def close(self):
        for b in self.batches.values():
            b.close()
        self.batches.clear()


def poisson_arrivals(n: int, rate: float, seed: int = 0) -> np.ndarray:
    """Arrival times (s) of a Poisson stream of `rate` requests per second."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))


def main(argv=None):
    """Serving sweep: Poisson request streams of the theory mix at several rates, one JSON line per rate."""
    import argparse
    import json

    from skyrmion_solver_amd import workloads

    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--rates", type=float, nargs="+", default=[5, 10, 20, 50, 100, 200])
    ap.add_argument("--n", type=int, default=600, help="requests per rate")
    ap.add_argument("--sizes", type=int, nargs="+", default=[64, 96, 128, 160, 192])
    ap.add_argument("--precision", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--tol", type=float, default=1.0e-4)
    ap.add_argument("--max-steps", type=int, default=20000)
    ap.add_argument("--chunk", type=int, default=32)
    ap.add_argument("--hbm-gib", type=float, default=200.0)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--progress", type=float, default=30.0, help="seconds between progress lines on stderr (0: none)")
    ap.add_argument("--refine", action="store_true",
                    help="FP32: finish each lattice in FP64 once its max|g| is below --refine-factor x tol")
    ap.add_argument("--refine-factor", type=float, default=3.0)
    ap.add_argument("--omega-factor", type=float, default=workloads.STEP_BUDGET_FACTOR["omega"],
                    help="step limit of omega requests as a multiple of --max-steps")
    ap.add_argument("--omega-max-size", type=int, default=0, help="largest omega lattice (0: no cap)")
    args = ap.parse_args(argv)
    factors = dict(workloads.STEP_BUDGET_FACTOR, omega=args.omega_factor)
    caps = {"omega": args.omega_max_size} if args.omega_max_size > 0 else None
    prec = _engine.FP32 if args.precision == "fp32" else _engine.FP64
    for rate in args.rates:
        reqs = workloads.c5_requests(n=args.n, rate=rate, seed=args.seed, sizes=args.sizes, tol=args.tol,
                                     max_steps=args.max_steps, device_init=True, max_size_by_theory=caps,
                                     budget_factor=factors)
        srv = Server(precision=prec, hbm_budget_bytes=int(args.hbm_gib * (1 << 30)), chunk_steps=args.chunk,
                     refine_fp64=args.refine, refine_factor=args.refine_factor)
        rep = srv.serve(reqs, progress_every=args.progress)
        srv.close()
        rep.update(rate_per_s=rate, precision=args.precision, sizes=args.sizes, tol=args.tol,
                   max_steps=args.max_steps, refine=bool(args.refine and prec == _engine.FP32))
        print(json.dumps(rep), flush=True)


if __name__ == "__main__":
    main()
