"""Shared test helpers: seeded lattices and fields for oracle / engine parity."""

from __future__ import annotations

import numpy as np

from skyrmion_solver_amd import initial_data
from skyrmion_solver_amd.theories import load_theory

THEORY_NAMES = {"pion": "Nuclear Skyrme model", "omega": "Omega-Skyrme model",
                "coulomb": "Coulomb-Skyrme model", "rho": "Rho-Skyrme model",
                "berger": "Berger Skyrme model", "lightly": "Lightly bound Skyrme model"}


def make_params(theory: str, X: int, Y: int, Z: int, **kw):
    """Resolved params + packed arrays for a small lattice (non-uniform spacings by default)."""
    mod = load_theory(THEORY_NAMES[theory])
    # time_step ~ 0.15 * spacing keeps RK4 stable for the quartic term at these gradients
    base = dict(xlen=X, ylen=Y, zlen=Z, lsx=0.21, lsy=0.19, lsz=0.2, time_step=0.03)
    if theory == "pion":
        base.update(mpi=0.7, kappa=1.1)
    if theory == "omega":
        base.update(c_omega=0.5, mass=0.45)
    if theory == "coulomb":
        base.update(kappa=0.3, mass=0.6)
    if theory == "berger":
        base.update(mpi=0.7, kappa=1.1, alpha=0.6)  # beta = alpha^2 - 1 = -0.64
    if theory == "lightly":
        base.update(mpi=0.5, alpha=0.3)
    if theory == "rho":
        # the rho Laplacian enters with weight 8 (pion: 2): RK4 needs the smaller step
        base.update(mpi=0.5, mrho=1.2, c_alpha=0.3, time_step=0.02)
    base.update(kw)
    p = mod.params.default_params(**base)
    rp = p.resolved()
    p_i, p_f = mod.params.pack_device_params(rp)
    return mod, rp, p_i, p_f


def smooth_field(rp, nf: int, seed: int = 0, noise: float = 0.02, normalise: bool = False, band: bool = False):
    """A hedgehog-like configuration with seeded smooth + random perturbations.

    Unnormalised by default: stage fields of RK stages 2-4 are not on S^3.
    Extra components (omega, V, rho) get smooth seeded bumps.
    """
    rng = np.random.default_rng(seed)
    cfg = dict(mode="hedgehog", profile=("exp", 0.9), isorotation="haar", jitter=0.05, seed=seed)
    F = initial_data.build_field(rp, nf, cfg)
    rx, ry, rz = initial_data.coordinates(rp)
    for a in range(nf):
        k = rng.normal(size=3)
        F[a] += noise * np.sin(k[0] * rx + k[1] * ry + k[2] * rz + rng.uniform(0, 6.28))
    for a in range(4, nf):
        c = rng.normal(scale=0.3, size=3)
        F[a] += 0.3 * np.exp(-((rx - c[0]) ** 2 + (ry - c[1]) ** 2 + (rz - c[2]) ** 2))
    F += noise * 0.1 * rng.normal(size=F.shape)
    if normalise:
        initial_data.normalise_sigma(F)
    if band:
        initial_data.set_vacuum_band(F)
    return F


def rel_l2(a, b) -> float:
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))
