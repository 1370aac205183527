"""Headline benchmark: site-updates/s of arrested Newton flow, pion Skyrme, 128^3, B=4.

One *step* = one full ANF step of the whole lattice (4 RK4 gradient stages,
RK4 combine, S^3 renormalisation / tangent projection, energy, baryon-number
and max-force reductions); one *site-update* = one site advanced one step.

Workload (one GPU): pion-only Skyrme model, 128^3 sites, dx = 0.15, m = 1,
kappa = 1, dt = 0.2 dx, B = 4 rational map R(z) = (z^4 + 2 sqrt(3) i z^2 + 1) /
(z^4 - 2 sqrt(3) i z^2 + 1) with profile f(r) = pi exp(-r/1.5), N(0, 1e-3)
noise then renormalised, vacuum boundary band.  Fields are resident in HBM
before the timed region.  Prints one JSON line.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from skyrmion_solver_amd import _engine, initial_data  # noqa: E402
from skyrmion_solver_amd.theories import load_theory  # noqa: E402

HBM_PEAK_GBS = 8000.0
# dt = 0.2 dx, the reference's own example setting (examples/nuclear_skyrme_gl.py:13); dt = 0.5 dx is
# beyond the RK4 stability limit of this workload (the quartic term stiffens the B=4 core)
COURANT = 0.2
METRIC = "site-updates/s at 128^3 B=4 pion ANF, FP32 vs FP64; p50/p99 solve latency"
# compulsory HBM traffic of each kernel in nf-vectors of the storage type:
# stage 1 reads F, writes l1; 2 reads F, V, writes l2; 3 reads F, V, l1, writes l3;
# 4 reads F, V, l1, l2, l3, writes F', V'; energy pass reads F
STAGE_VECTORS = {0: 1, 1: 2, 2: 3, 3: 4, 4: 7}
STAGE_NAMES = {0: "energy", 1: "stage1", 2: "stage2", 3: "stage3", 4: "stage4"}


def workload(L: int = 128, dx: float = 0.15, seed: int = 0):
    theory = load_theory("Nuclear Skyrme model")
    p = theory.params.default_params(xlen=L, ylen=L, zlen=L, xsize=dx * (L - 1), ysize=dx * (L - 1),
                                     zsize=dx * (L - 1), lsx=dx, lsy=dx, lsz=dx, courant=COURANT, mpi=1.0, kappa=1.0,
                                     baryon_number=4.0, ansatz="rational_map")
    rp = p.resolved()
    p_i, p_f = theory.params.pack_device_params(rp)
    cfg = dict(mode="rational_map", baryon_number=4, profile=("exp", 1.5), noise=1.0e-3, seed=seed, vacuum_band=True)
    F = initial_data.build_field(rp, 4, cfg)
    return rp, p_i, p_f, F


def time_form(prec, p_i, p_f, F, steps, warmup, chunk, trace=True, theory=None, trace_min_steps=200):
    """Timed device-resident run (hipGraph chunks of `chunk` steps, capped at `steps` so that every
    launched step is a real one; both ping-pong parities' graphs are captured before the timed region).

    With `trace`, a traced pass runs first on a second engine: the same warm-up, then max(steps,
    `trace_min_steps`) steps from the same field with every stage kernel bracketed by HIP events on the
    engine's stream (direct launches), giving the per-kernel average durations of this workload.  Run
    right before the timed pass, it also brings the GPU to its sustained clocks, which a short timed run
    (tens of steps) would otherwise spend ramping up."""
    theory = _engine.THEORY_PION if theory is None else theory
    chunk = max(1, min(chunk, steps))
    out = {}
    eng = _engine.Engine(theory, prec, p_i, p_f)
    eng.upload(_engine.FIELD, F)
    eng.prepare(chunk)
    tr = None
    if trace:
        tr = _engine.Engine(theory, prec, p_i, p_f, flags=_engine.FLAG_TRACE)
        tr.upload(_engine.FIELD, F)
        tr.anf_run(warmup, -1.0, chunk)
        tr.trace_reset()
        tr.anf_run(max(steps, trace_min_steps), -1.0, chunk)
        ms, cnt = tr.trace_read()
        out["kernels"] = {STAGE_NAMES[k]: dict(avg_us=1000.0 * ms[k] / cnt[k], launches=int(cnt[k]))
                          for k in range(5) if cnt[k]}
    eng.anf_run(warmup, -1.0, chunk)
    eng.sync()
    t0 = time.perf_counter()
    st = eng.anf_run(steps, -1.0, chunk)
    eng.sync()
    dt = time.perf_counter() - t0
    if st.issued != steps or st.steps != steps:
        raise RuntimeError(f"timed run issued {st.issued} and took {st.steps} steps, expected {steps}")
    out.update(seconds=dt, steps=st.steps, issued=st.issued, graph_launches=st.graph_launches, energy=st.energy,
               err_abs=st.err_abs, baryon=st.baryon)
    eng.close()
    if tr is not None:
        tr.close()
    return out


def profiled_traffic(kernel: str, form: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary (FETCH_SIZE and
    WRITE_SIZE scaled by the copy-kernel calibration), or None when no summary covers it."""
    path = os.path.join(ROOT, "profiles", f"stage_kernels_pion_128_b4_{form}.md")
    key = f"traffic_bytes_per_launch {kernel} {form}:"
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key):
            return float(line.split(":")[1])
    return None


def solve_latency(p_i, p_f, seeds, tol, max_steps):
    """Time-to-converged (max|g| < tol) of both forms from each seed's initial field, with the converged E
    and B; and per seed the field rel-L2 of the FP32 solution against the FP64 one.  Each seed's field is
    built once and solved in both forms."""
    forms = (("fp32", _engine.FP32), ("fp64", _engine.FP64))
    rec = {name: dict(lat=[], steps=[], energy=[], baryon=[], conv=[]) for name, _ in forms}
    l2 = []
    for s in seeds:
        _, _, _, F = workload(seed=s)
        fields = {}
        for name, prec in forms:
            r = rec[name]
            eng = _engine.Engine(_engine.THEORY_PION, prec, p_i, p_f)
            eng.upload(_engine.FIELD, F)
            eng.sync()
            t0 = time.perf_counter()
            st = eng.anf_run(max_steps, tol, 64)
            eng.sync()
            r["lat"].append(time.perf_counter() - t0)
            r["steps"].append(st.steps)
            r["energy"].append(st.energy)
            r["baryon"].append(st.baryon)
            r["conv"].append(bool(st.converged))
            fields[name] = eng.download(_engine.FIELD)
            eng.close()
        l2.append(float(np.linalg.norm(fields["fp32"] - fields["fp64"]) / np.linalg.norm(fields["fp64"])))
    out = {}
    for name, _ in forms:
        r = rec[name]
        lat = np.array(r["lat"])
        out[name] = dict(tol=tol, seeds=len(seeds), p50_s=float(np.percentile(lat, 50)),
                         p99_s=float(np.percentile(lat, 99)), steps=r["steps"], converged=r["conv"],
                         energy=r["energy"], baryon=r["baryon"])
    return out, l2


def host_identity():
    """Host CPU identity: `nproc`, the cores this process may use, the model name and the sockets
    (from /proc/cpuinfo)."""
    model, sockets = None, set()
    try:
        for line in open("/proc/cpuinfo"):
            key, _, val = line.partition(":")
            key = key.strip()
            if key == "model name" and model is None:
                model = val.strip()
            elif key == "physical id":
                sockets.add(val.strip())
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return dict(nproc=os.cpu_count(), usable_cores=usable, model=model, sockets=len(sockets) or None)


def cpu_baseline(p_i, p_f, F, steps: int = 3, threads: int | None = None, threads_steps: int = 5):
    """FP64 NumPy oracle of the same 128^3 workload on the host: `steps` ANF steps on one thread, and
    `threads_steps` steps with the stencil, point-gradient and RK4 passes split over x-slabs on `threads`
    threads (default: the usable cores, at most 16, the host share of one GPU).  `value` is the
    all-threads figure; both runs take the same arithmetic (they agree bit for bit)."""
    from oracle import integrator as OI
    from oracle import parallel as OP

    N = int(p_i[6])
    host = host_identity()
    st = OI.OracleState("pion", p_i, p_f, F)
    E = OI.energy(st)
    t0 = time.perf_counter()
    for _ in range(steps):
        E, _ = OI.anf_step(st, E)
    t1 = time.perf_counter() - t0
    threads = threads or max(1, min(16, host["usable_cores"] or 1))
    pool = OP.SlabPool(int(p_i[0]), threads)
    try:
        st = OI.OracleState("pion", p_i, p_f, F)
        E = OP.energy(st, pool)
        t0 = time.perf_counter()
        for _ in range(threads_steps):
            E, _ = OP.anf_step(st, E, pool)
        tn = time.perf_counter() - t0
    finally:
        pool.close()
    return dict(value=N * threads_steps / tn, unit="site-updates/s", cores=threads, kind="port",
                sample=f"{threads_steps} ANF steps of the 128^3 B=4 pion workload, FP64 NumPy oracle split over "
                       f"x-slabs on {threads} threads, {tn:.1f} s",
                single_thread=dict(value=N * steps / t1, cores=1,
                                   sample=f"{steps} ANF steps, same oracle on 1 thread, {t1:.1f} s"),
                host=host)


def batch_sample(n: int, steps: int, shape=None, solo: bool = True):
    """Batched form: `n` independent pion lattices (L uniform in {48, 64, 96, 128}^3, B = 1..6 rational maps
    with random orientations, m ~ U[0, 1.5]) advanced together, FP32.  The batch's tile and x-chunk rule
    come from the autotuner (timed on one lattice per size class), or are `shape` = (tile_z, x-chunk
    target) when given; the comparison (`solo`) is the same lattices run one after another as solo engines
    with the engine's default launch shape."""
    from skyrmion_solver_amd import autotune, workloads

    inst = workloads.c4_instances(n=n, seed=0, device_init=True)
    per_class = {}
    for d in inst:
        per_class.setdefault(d["L"], d)
    if shape is None:
        tuned = autotune.tune_batch(_engine.THEORY_PION, _engine.FP32,
                                    [(d["p_i"], d["p_f"]) for d in per_class.values()], steps=8, use_cache=False)
        tz, tgt = tuned["tile_z"], tuned["target"]
    else:
        tz, tgt = shape
    dims = [(d["L"],) * 3 for d in inst]
    blocks = autotune.batch_blocks(dims, tz, tgt)
    sites = sum(d["L"] ** 3 for d in inst)
    with _engine.Batch(_engine.THEORY_PION, _engine.FP32, max_slots=len(inst), max_blocks=blocks, tile_z=tz) as b:
        for d in inst:  # fields built on the device from each lattice's initial-data descriptor
            b.add(d["p_i"], d["p_f"], init=d["init"], max_steps=1 << 30, x_chunk=autotune.chunk_rule(d["L"], tgt))
        b.run(4)
        t0 = time.perf_counter()
        b.run(steps)
        dt = time.perf_counter() - t0
    if not solo:
        return dict(instances=len(inst), sites=sites, steps=steps, site_updates_per_s=sites * steps / dt,
                    tile_z=tz, x_chunk_target=tgt, blocks=blocks)
    # solo baseline: one engine per size class, scaled by the class count
    solo_s = 0.0
    for L, d in per_class.items():
        count = sum(1 for e in inst if e["L"] == L)
        eng = _engine.Engine(_engine.THEORY_PION, _engine.FP32, d["p_i"], d["p_f"])
        eng.init_field(d["init"])
        eng.anf_run(4, -1.0, 4)
        st = eng.anf_run(steps, -1.0, steps)
        eng.close()
        solo_s += count * st.seconds
    return dict(instances=len(inst), sites=sites, steps=steps, site_updates_per_s=sites * steps / dt,
                solo_sequential_site_updates_per_s=sites * steps / solo_s, tile_z=tz, x_chunk_target=tgt,
                blocks=blocks)


THEORY_IDS = {"pion": _engine.THEORY_PION, "omega": _engine.THEORY_OMEGA, "coulomb": _engine.THEORY_COULOMB}


def c3_workload(theory: str, L: int = 160):
    """The omega / Coulomb B=8 product workload of the 160^3 sample on an L^3 lattice, as a host field
    (omega starting at -c_omega CS); returns (rp, p_i, p_f, F)."""
    from skyrmion_solver_amd import workloads

    _, rp, p_i, p_f, F = workloads.c3_b8(theory, L=L)
    if theory == "omega":
        with _engine.Engine(_engine.THEORY_OMEGA, _engine.FP64, p_i, p_f) as eng:
            eng.upload(_engine.FIELD, F)
            F = workloads.set_omega_from_chern_simons(eng, rp)
    return rp, p_i, p_f, F


def c3_sample(steps: int = 2000, theories=("omega", "coulomb")):
    """160^3 B=8 omega and Coulomb (product of two B=4 rational maps at -+1.5 along x, dx = 0.15; fields
    built on the device): `steps` ANF steps in the FP64 and FP32 forms, each timed as one device-resident
    run, and the FP32 result against the FP64 one (energy, baryon number, field rel-L2)."""
    from skyrmion_solver_amd import workloads

    tids = {"omega": _engine.THEORY_OMEGA, "coulomb": _engine.THEORY_COULOMB}
    out = {}
    for th in theories:
        _, rp, p_i, p_f, init = workloads.c3_b8(th, device_init=True)
        N = int(p_i[6])
        r = {}
        for name, prec in (("fp64", _engine.FP64), ("fp32", _engine.FP32)):
            with _engine.Engine(tids[th], prec, p_i, p_f) as eng:
                eng.init_field(init)
                B0 = eng.baryon()
                eng.prepare(100)
                t0 = time.perf_counter()
                st = eng.anf_run(steps, -1.0, 100)
                eng.sync()
                dt = time.perf_counter() - t0
                r[name] = dict(site_updates_per_s=N * st.steps / dt, ms_per_step=1000.0 * dt / st.steps,
                               steps=st.steps, energy=st.energy, baryon=st.baryon, baryon_initial=B0,
                               err_abs=st.err_abs)
                r[name + "_field"] = eng.download(_engine.FIELD)
        F64, F32 = r.pop("fp64_field"), r.pop("fp32_field")
        r["fp32_vs_fp64"] = dict(abs_dE=abs(r["fp32"]["energy"] - r["fp64"]["energy"]),
                                 rel_dE=abs(r["fp32"]["energy"] - r["fp64"]["energy"]) / abs(r["fp64"]["energy"]),
                                 abs_dB=abs(r["fp32"]["baryon"] - r["fp64"]["baryon"]),
                                 field_rel_l2=float(np.linalg.norm(F32 - F64) / np.linalg.norm(F64)))
        r["fp32_over_fp64"] = r["fp32"]["site_updates_per_s"] / r["fp64"]["site_updates_per_s"]
        r["algorithmic_GBs_fp32"] = r["fp32"]["site_updates_per_s"] * 16 * 5 * 4 / 1e9
        out[th] = r
    return out


def _serve_summary(rep: dict) -> dict:
    """Time-to-result over all requests and time-to-converged over the converged ones, with the counts
    of requests that converged, stopped at their step limit and went non-finite."""
    by = rep["by_theory"]
    nonfinite = sum(t["nonfinite"] for t in by.values())
    return dict(requests=rep["requests"], converged=rep["converged"],
                stopped_at_step_limit=rep["requests"] - rep["converged"] - nonfinite, nonfinite=nonfinite,
                converged_p50_s=rep["converged_p50_s"], converged_p99_s=rep["converged_p99_s"],
                p50_s=rep["p50_s"], p99_s=rep["p99_s"], makespan_s=rep["makespan_s"],
                site_updates_per_s=rep["site_updates_per_s"], refined=rep["refined"], by_theory=by,
                peak_hbm_bytes=rep["peak_hbm_bytes"])


def serving_sample(n: int, rates, sizes, max_steps: int, compare_rate: float, omega_max_size: int = 64,
                   seed: int = 3):
    """Serving form: Poisson streams of `n` minimisation requests (pion 60 %, omega 25 %, rho 15 %; L from
    `sizes`; B = 1..8; step limit `max_steps`) served to max|g| < 1e-4 by continuous batching.  The omega
    saddle flow relaxes slowly (up to ~80k steps at 64^3), so omega requests run on lattices of at most
    `omega_max_size`^3 with ``workloads.STEP_BUDGET_FACTOR['omega']`` times the step limit.  At every rate of
    `rates` the FP32 form with each lattice finished in FP64 once max|g| < 3 tol; at `compare_rate` also
    the plain FP32 and FP64 forms.  `seed` draws the stream (the default, 3, gives 5 pion, 2 omega and 1 rho
    request over all four sizes, the nearest 8-request stream to the 60/25/15 mix)."""
    from skyrmion_solver_amd import serving, workloads

    runs = []
    for rate in rates:
        forms = [("fp32_refined_fp64", _engine.FP32, True)]
        if rate == compare_rate:
            forms = [("fp32", _engine.FP32, False), ("fp64", _engine.FP64, False)] + forms
        for name, prec, refine in forms:
            reqs = workloads.c5_requests(n=n, rate=rate, seed=seed, sizes=sizes, max_steps=max_steps, device_init=True,
                                         max_size_by_theory={"omega": omega_max_size})
            srv = serving.Server(precision=prec, hbm_budget_bytes=64 << 30, chunk_steps=32, max_slots=64,
                                 max_blocks=32768, refine_fp64=refine)
            rep = srv.serve(reqs)
            srv.close()
            runs.append(dict(form=name, rate_per_s=rate, **_serve_summary(rep)))
            print(f"[bench] serving {name} rate {rate}: converged {rep['converged']}/{rep['requests']}, "
                  f"converged p50 {rep['converged_p50_s']} s, makespan {rep['makespan_s']:.1f} s",
                  file=sys.stderr, flush=True)
    return dict(requests_per_stream=n, seed=seed, sizes=list(sizes), omega_max_size=omega_max_size, max_steps=max_steps,
                step_budget_factor=dict(workloads.STEP_BUDGET_FACTOR), tol=1.0e-4, runs=runs)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--chunk", type=int, default=100,
                    help="steps per captured graph (capped at --steps; a remainder runs as direct launches)")
    ap.add_argument("--no-trace", action="store_true", help="time without the per-kernel HIP events")
    ap.add_argument("--solve-seeds", type=int, default=20, help="seeds of the solve-latency sample (0..n-1)")
    ap.add_argument("--solve-max-steps", type=int, default=20000)
    ap.add_argument("--tol", type=float, default=1.0e-4)
    ap.add_argument("--cpu-steps", type=int, default=3, help="oracle steps of the CPU baseline (~4.6 s each)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-n", type=int, default=48, help="lattices of the batched sample (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=40)
    ap.add_argument("--serve-n", type=int, default=8, help="requests per serving stream (0: skip)")
    ap.add_argument("--serve-rates", type=float, nargs="+", default=[5.0, 20.0, 200.0])
    ap.add_argument("--serve-compare-rate", type=float, default=20.0,
                    help="rate at which the plain FP32 and FP64 forms are served too")
    ap.add_argument("--serve-sizes", type=int, nargs="+", default=[64, 96, 128, 192])
    ap.add_argument("--serve-max-steps", type=int, default=2500,
                    help="step limit of a serving request (omega requests: STEP_BUDGET_FACTOR['omega'] times this)")
    ap.add_argument("--serve-omega-max-size", type=int, default=64,
                    help="largest lattice of an omega serving request")
    ap.add_argument("--serve-seed", type=int, default=3,
                    help="seed of the serving streams (3: 5 pion, 2 omega, 1 rho request, the 60/25/15 mix)")
    ap.add_argument("--c3-steps", type=int, default=2000, help="ANF steps of the 160^3 B=8 omega / Coulomb sample "
                                                               "(0: skip)")
    ap.add_argument("--batch-shape", type=int, nargs=2, default=None, metavar=("TILE_Z", "TARGET"),
                    help="batch tile and x-chunk target instead of the autotuner's (profiling aid)")
    ap.add_argument("--only", choices=["fp32", "fp64", "copy", "batch"], default=None,
                    help="profiling aid: run only one form (or only the copy calibration) and print no JSON")
    ap.add_argument("--size", type=int, default=128,
                    help="profiling aid with --only fp32/fp64: lattice edge of the workload")
    ap.add_argument("--theory", choices=["pion", "omega", "coulomb"], default="pion",
                    help="profiling aid with --only fp32/fp64: the pion B=4 workload, or the omega / Coulomb "
                         "B=8 product workload of the 160^3 sample")
    args = ap.parse_args()
    if (args.size != 128 or args.theory != "pion") and args.only not in ("fp32", "fp64"):
        ap.error("--size and --theory apply to --only fp32 / fp64 runs; the metric is quoted at 128^3 pion")
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus must be 1 (multi-GPU sharding is design-only)"}))
        sys.exit(2)

    if args.only == "copy":
        gbs, ms = _engine.stream_copy_bandwidth(1 << 30, 20)
        print(f"[bench] stream copy of 1 GiB: {gbs:.1f} GB/s, {ms:.4f} ms per copy", file=sys.stderr, flush=True)
        return
    if args.only == "batch":
        b = batch_sample(args.batch_n, args.batch_steps, shape=args.batch_shape, solo=False)
        print(f"[bench] batch: {b['site_updates_per_s']:.4g} site-updates/s over {b['instances']} lattices, "
              f"{1000.0 * b['sites'] / b['site_updates_per_s']:.4f} ms/step, {b}", file=sys.stderr, flush=True)
        return
    if args.only in ("fp32", "fp64") and args.theory != "pion":
        rp, p_i, p_f, F = c3_workload(args.theory, args.size)
    else:
        rp, p_i, p_f, F = workload(L=args.size)
    N = int(p_i[6])
    nf = 4
    if args.only in ("fp32", "fp64"):
        prec = _engine.FP32 if args.only == "fp32" else _engine.FP64
        r = time_form(prec, p_i, p_f, F, args.steps, args.warmup, args.chunk, trace=not args.no_trace,
                      theory=THEORY_IDS[args.theory])
        print(f"[bench] {args.only}: {N * r['steps'] / r['seconds']:.4g} site-updates/s, "
              f"{1000.0 * r['seconds'] / r['steps']:.4f} ms/step", file=sys.stderr, flush=True)
        for k, v in r.get("kernels", {}).items():
            print(f"[bench]   {k}: {v['avg_us']:.2f} us x {v['launches']}", file=sys.stderr, flush=True)
        return
    copy_gbs, _ = _engine.stream_copy_bandwidth(1 << 30, 20)
    res = {}
    for name, prec in (("fp64", _engine.FP64), ("fp32", _engine.FP32)):
        r = time_form(prec, p_i, p_f, F, args.steps, args.warmup, args.chunk, trace=not args.no_trace)
        r["site_updates_per_s"] = N * r["steps"] / r["seconds"]
        r["ms_per_step"] = 1000.0 * r["seconds"] / r["steps"]
        s = 8 if prec == _engine.FP64 else 4
        r["algorithmic_GBs"] = r["site_updates_per_s"] * 16 * nf * s / 1e9
        res[name] = r
        print(f"[bench] {name}: {r['site_updates_per_s']:.4g} site-updates/s, {r['ms_per_step']:.4f} ms/step, "
              f"{r['algorithmic_GBs']:.1f} GB/s algorithmic", file=sys.stderr, flush=True)

    head = res["fp32"]
    roofline = None
    if "kernels" in head:
        # dominant kernel of the headline form: largest time per step
        kern = head["kernels"]
        dom = max(kern, key=lambda k: kern[k]["avg_us"] * (1 if k != "energy" else 0))
        stage = [k for k, v in STAGE_NAMES.items() if v == dom][0]
        bytes_per_launch = STAGE_VECTORS[stage] * nf * 4 * N
        achieved = bytes_per_launch / (kern[dom]["avg_us"] * 1e-6) / 1e9
        roofline = dict(bound="hbm", kernel=dom, achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
                        frac=achieved / HBM_PEAK_GBS, traffic=profiled_traffic(dom, "fp32"),
                        bytes_per_launch=bytes_per_launch, measured_copy_GBs=copy_gbs,
                        frac_of_copy=achieved / copy_gbs)

    latency, l2 = {}, []
    if args.solve_seeds > 0:
        latency, l2 = solve_latency(p_i, p_f, list(range(args.solve_seeds)), args.tol, args.solve_max_steps)
        for name in ("fp32", "fp64"):
            print(f"[bench] solve {name}: {latency[name]}", file=sys.stderr, flush=True)

    accuracy = None
    if "fp32" in latency and "fp64" in latency:
        e32, e64 = np.array(latency["fp32"]["energy"]), np.array(latency["fp64"]["energy"])
        b32, b64 = np.array(latency["fp32"]["baryon"]), np.array(latency["fp64"]["baryon"])
        accuracy = dict(max_rel_dE=float(np.max(np.abs(e32 - e64) / np.abs(e64))),
                        max_abs_dB=float(np.max(np.abs(b32 - b64))), max_field_rel_l2=max(l2),
                        converged=all(latency["fp32"]["converged"]) and all(latency["fp64"]["converged"]),
                        note="FP32 form vs FP64 form at convergence (max|g| < tol), same seeds; the FP64 form "
                             "matches the FP64 oracle step for step at this size")
    c3 = c3_sample(args.c3_steps) if args.c3_steps > 0 else None
    if c3:
        print(f"[bench] c3: {c3}", file=sys.stderr, flush=True)

    batch = batch_sample(args.batch_n, args.batch_steps) if args.batch_n > 0 else None
    if batch:
        print(f"[bench] batch: {batch}", file=sys.stderr, flush=True)
    serve = serving_sample(args.serve_n, args.serve_rates, args.serve_sizes, args.serve_max_steps,
                           args.serve_compare_rate, args.serve_omega_max_size, args.serve_seed) if args.serve_n > 0 else None

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(p_i, p_f, F, args.cpu_steps)
        print(f"[bench] cpu baseline: {cpu}", file=sys.stderr, flush=True)

    line = {
        "metric": METRIC,
        "value": head["site_updates_per_s"],
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "pion Skyrme ANF, 128^3, dx=0.15, m=1, kappa=1, dt=0.03, B=4 rational map "
                               "f=pi*exp(-r/1.5), noise 1e-3, FP32 storage with FP64 reductions",
                   "sites": N, "components": nf},
        "fp64": {"value": res["fp64"]["site_updates_per_s"], "ms_per_step": res["fp64"]["ms_per_step"],
                 "algorithmic_GBs": res["fp64"]["algorithmic_GBs"], "kernels": res["fp64"]["kernels"]},
        "fp32": {"algorithmic_GBs": head["algorithmic_GBs"], "kernels": head["kernels"]},
        "fp32_over_fp64": head["site_updates_per_s"] / res["fp64"]["site_updates_per_s"],
        "solve_latency": latency,
        "fp32_accuracy": accuracy,
        "c3_160_b8": c3,
        "batch": batch,
        "serving": serve,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "device": dict(zip(("arch", "compute_units"), _engine.device_info(0))),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
