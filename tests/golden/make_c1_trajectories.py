"""Write ``c1_trajectories.npz``: the oracle's 500-step ANF trajectories of the
32^3 B=1 hedgehog workload (pion, dx = 0.2, m = 0, dt = 0.1, hedgehog
f = pi e^{-r} with a random SU(2) isorotation and a centre jitter of 0.1), seeds 0..9.

Per seed (arrays indexed [seed, ...]):

* ``energy[s, k]``, ``baryon[s, k]`` (k = 0..500): energy and baryon number of
  the field entering step k (k = 500: the final field);
* ``err[s, k]``, ``max_force[s, k]`` (k = 0..499): the reference's signed
  maximum of step k's stage-4 energy gradient, and its maximum modulus;
* ``arrested[s, k]``: step k ended with the velocity zeroed (E rose);
* ``sites[s]`` (512 site indices drawn by ``numpy.random.default_rng(1000 + s)``)
  and ``final_sample[s, a, j]``: component a of the final field at those sites;
* ``final_sum[s, a]``, ``final_sumsq[s, a]``: per component, the sum and the
  sum of squares of the final field over the whole lattice;
* ``input_sha[s]``: SHA-256 (hex) of the input field's FP64 bytes.

The oracle (``oracle.integrator``, ``core/integrator.py:284-410``) computes
every number.  The input fields come from ``skyrmion_solver_amd.workloads.c1_hedgehog``
(the host builder, which tests/test_oracle_initial_config.py holds equal to the
oracle's initial data); the stored hashes make a change of that builder show
as a fixture mismatch rather than as a parity failure.

Usage: python tests/golden/make_c1_trajectories.py [--jobs 8]   (~85 s of one core per seed)
"""

from __future__ import annotations

import argparse
import hashlib
import sys
from concurrent.futures import ProcessPoolExecutor
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
OUT = Path(__file__).resolve().parent / "c1_trajectories.npz"
SEEDS = tuple(range(10))
STEPS = 500
NSAMPLE = 512


def sample_sites(seed: int, n_sites: int) -> np.ndarray:
    return np.sort(np.random.default_rng(1000 + seed).choice(n_sites, NSAMPLE, replace=False)).astype(np.int64)


def input_hash(F: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(F, dtype=np.float64).tobytes()).hexdigest()


def trajectory(seed: int) -> dict:
    sys.path.insert(0, str(ROOT))
    from oracle import integrator as OI
    from skyrmion_solver_amd import workloads

    _th, _rp, p_i, p_f, F = workloads.c1_hedgehog(seed=seed)
    sha = input_hash(F)
    st = OI.OracleState("pion", p_i, p_f, F)
    E = OI.energy(st)
    energy, baryon, err, fmax, arrested = [E], [OI.baryon_number(st)], [], [], []
    for _ in range(STEPS):
        E, e = OI.anf_step(st, E)
        energy.append(E)
        baryon.append(OI.baryon_number(st))
        err.append(e)
        fmax.append(OI.max_abs_force(st))
        arrested.append(bool(np.max(np.abs(st.V)) == 0.0))
    Ff = st.F.reshape(4, -1)
    sites = sample_sites(seed, Ff.shape[1])
    return dict(energy=energy, baryon=baryon, err=err, max_force=fmax, arrested=arrested, sites=sites,
                final_sample=Ff[:, sites], final_sum=Ff.sum(axis=1), final_sumsq=(Ff * Ff).sum(axis=1), input_sha=sha)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    args = ap.parse_args()
    with ProcessPoolExecutor(max_workers=args.jobs) as ex:
        rows = list(ex.map(trajectory, SEEDS))
    out = {k: np.array([r[k] for r in rows]) for k in rows[0] if k != "input_sha"}
    out["input_sha"] = np.array([r["input_sha"] for r in rows])
    out["seeds"] = np.array(SEEDS)
    np.savez_compressed(OUT, **out)
    for s, r in zip(SEEDS, rows):
        dE = np.diff(r["energy"])
        print(f"seed {s}: E {r['energy'][0]:.10g} -> {r['energy'][-1]:.10g}, B {r['baryon'][-1]:.6f}, "
              f"arrests {sum(r['arrested'])}, min |dE|/E {np.min(np.abs(dE)) / abs(r['energy'][-1]):.3e}")
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes)")


if __name__ == "__main__":
    main()
