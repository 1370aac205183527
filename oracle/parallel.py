"""All-cores driver of the oracle's ANF step (oracle; test infrastructure and the
CPU baseline's multi-threaded leg only).

The arithmetic is the oracle's (:mod:`oracle.integrator`, restating
``core/integrator.py:284-410``); only the stencil-and-point-gradient passes
and the energy density are split into slabs of x-planes, each evaluated by a
thread of a pool (NumPy releases the GIL inside its array operations).  A slab
is evaluated on a copy with ghost planes: three below it, so that the mixed
derivatives' low-side guard ``i > halo`` (``core/derivatives.py:121``) falls on
a ghost plane, and two above; at the lattice ends the slab starts or stops at
the boundary itself, where the reference's guards apply unchanged.  Every
site's value is therefore computed exactly as in the single-threaded oracle,
so the two agree bit for bit (``tests/test_oracle_integrator.py``).
"""

from __future__ import annotations

from concurrent.futures import ThreadPoolExecutor

import numpy as np

from oracle import integrator as OI
from oracle import reductions, stencils, theories

LOW_GHOSTS, HIGH_GHOSTS = 3, 2


def slabs(X: int, n: int):
    """`n` contiguous x-ranges covering 0..X (at most X)."""
    n = max(1, min(n, X))
    edges = np.linspace(0, X, n + 1).round().astype(int)
    return [(int(a), int(b)) for a, b in zip(edges[:-1], edges[1:]) if b > a]


class SlabPool:
    """A thread pool and the slab decomposition of one lattice."""

    def __init__(self, X: int, threads: int):
        self.threads = max(1, int(threads))
        self.slabs = slabs(X, self.threads)
        self.pool = ThreadPoolExecutor(max_workers=self.threads)

    def close(self):
        self.pool.shutdown(wait=True)

    def map_slabs(self, fn, T: np.ndarray, out_shape):
        """out[:, a:b] = fn(T with ghosts)[:, real planes] for every slab (a, b); out is (k, ...) arrays."""
        X = T.shape[1]
        outs = [np.empty(s) for s in out_shape]

        def run(ab):
            a, b = ab
            lo, hi = max(a - LOW_GHOSTS, 0), min(b + HIGH_GHOSTS, X)
            res = fn(T[:, lo:hi])
            for o, r in zip(outs, res):
                o[:, a:b] = r[:, a - lo:a - lo + (b - a)]

        list(self.pool.map(run, self.slabs))
        return outs


def gradient(state: OI.OracleState, pool: SlabPool, T: np.ndarray):
    """:func:`oracle.integrator.gradient` evaluated slab by slab."""

    def fn(Ts):
        d1 = stencils.first_derivatives(Ts, state.ls, state.halo)
        d2 = stencils.second_derivatives(Ts, state.ls, state.halo)
        return theories.GRADIENTS[state.theory](Ts, d1, d2, state.p_i, state.p_f)

    EG, l = pool.map_slabs(fn, T, (T.shape, T.shape))
    return EG, l


def energy(state: OI.OracleState, pool: SlabPool, F: np.ndarray | None = None) -> float:
    """:func:`oracle.integrator.energy` with the per-site density evaluated slab by slab."""
    F = state.F if F is None else F

    def fn(Fs):
        d1 = stencils.first_derivatives(Fs, state.ls, state.halo)
        return (theories.ENERGIES[state.theory](Fs, d1, state.p_i, state.p_f)[None],)

    (dens,) = pool.map_slabs(fn, F, ((1,) + F.shape[1:],))
    return reductions.ref_sum(dens[0])


def anf_step(state: OI.OracleState, prev_energy: float, pool: SlabPool):
    """:func:`oracle.integrator.anf_step` on all threads of `pool`.

    The element-wise statements of ``rk4_step`` (``core/integrator.py:382-398``) run per slab inside the
    stage passes, with the same operands and order, and the final update of F and V runs in a pass of
    its own (stage 4 still reads F's ghost planes), so every value equals the single-threaded oracle's."""
    dt = state.p_f[7]
    unit_length = int(state.p_i[10])
    killkinen = int(state.p_i[8])
    nsk = state.nsk
    F, V = state.F, state.V
    X = F.shape[1]
    k = [np.empty_like(F) for _ in range(4)]
    l = [np.empty_like(F) for _ in range(4)]
    EG = np.empty_like(F)

    def grad(Ts):
        d1 = stencils.first_derivatives(Ts, state.ls, state.halo)
        d2 = stencils.second_derivatives(Ts, state.ls, state.halo)
        return theories.GRADIENTS[state.theory](Ts, d1, d2, state.p_i, state.p_f)

    def stage(s):
        def run(ab):
            a, b = ab
            lo, hi = max(a - LOW_GHOSTS, 0), min(b + HIGH_GHOSTS, X)
            r = slice(a - lo, a - lo + (b - a))
            if s == 0:
                Ts = F[:, lo:hi] + 0.0 * F[:, lo:hi]
                k[0][:, a:b] = dt * (V[:, a:b] + 0.0 * V[:, a:b])
            else:
                Ts = F[:, lo:hi] + (0.5 if s < 3 else 1.0) * k[s - 1][:, lo:hi]
            eg, ls_ = grad(Ts)
            l[s][:, a:b] = ls_[:, r]
            if s < 3:
                k[s + 1][:, a:b] = dt * (V[:, a:b] + (0.5 if s < 2 else 1.0) * l[s][:, a:b])
            else:
                EG[:, a:b] = eg[:, r]
        list(pool.pool.map(run, pool.slabs))

    for s in range(4):
        stage(s)

    def update(ab):
        a, b = ab
        Vs, Fs = V[:, a:b], F[:, a:b]
        Vs += (1.0 / 6.0) * (l[0][:, a:b] + 2.0 * l[1][:, a:b] + 2.0 * l[2][:, a:b] + l[3][:, a:b])
        Fs += (1.0 / 6.0) * (k[0][:, a:b] + 2.0 * k[1][:, a:b] + 2.0 * k[2][:, a:b] + k[3][:, a:b])
        if unit_length:
            theories.normalise(Fs, nsk)
            theories.project(Vs, Fs, nsk)

    list(pool.pool.map(update, pool.slabs))
    state.EG = EG
    new_energy = energy(state, pool)
    if killkinen and (new_energy > prev_energy):
        state.V[...] = 0.0
    return new_energy, reductions.ref_max(EG)
