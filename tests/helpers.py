"""Shared test helpers: seeded lattices and fields for oracle / engine parity."""

from __future__ import annotations

import numpy as np

from oracle import initial_config as OIC
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


def oracle_field(rp, p_i, p_f, nf, cfg):
    """The oracle's field for a build request: the lumps ``initial_data.lump_specs`` draws (centres,
    rotation matrices, profiles), each evaluated by the oracle, multiplied in order, then the band."""
    lumps, extra = initial_data.lump_specs(rp, cfg)
    acc = None
    for sp in lumps:
        prof = sp["profile"]
        if prof is None:
            if sp["mode"] == "hedgehog":
                prof = ("gauss", max(max(p_f[0], p_f[1], p_f[2]) / (8.0 * max(abs(sp["B"]), 1.0)), 1.0e-6))
            else:
                prof = ("fit", 1.0)
        elif prof == "fit":
            prof = ("fit", 1.0)
        if sp["rotation"] is not None:
            iso = sp["rotation"]
        else:
            iso = OIC.euler_matrix(sp["iso"], 0.0, 0.0) if sp["mode"] == "hedgehog" else OIC.euler_matrix(0.0, 0.0,
                                                                                                    sp["iso"])
        part = OIC.lump(p_i, p_f, sp["mode"], sp["B"], tuple(prof), sp["center"], sp["space_rotation"], iso)
        acc = part if acc is None else OIC.product_ansatz(acc, part)
    F = np.zeros((nf, int(p_i[0]), int(p_i[1]), int(p_i[2])))
    F[0] = 1.0
    if acc is not None:
        F[:4] = acc
    if extra["vacuum_band"]:
        OIC.vacuum_band(F, int(p_i[3]))
    return F


def ill_conditioned_sites(rp, p_i, p_f, cfg, shift: float = 1.0e-11, tol: float = 1.0e-7, dilate: int = 0):
    """Sites where a request's ansatz is ill-conditioned, as a boolean (X, Y, Z) mask.

    Where a site lies exactly on a direction in which a rational map's denominator vanishes (e.g. four
    of the cube diagonals of the B = 4 map, hit by lattices whose lump centres sit half a spacing off
    the sites), the reference's regularised division (1e-18 added to |den|^2) returns a value set by
    rounding noise, so implementations with different rounding (device math library, operand order)
    differ there.  Detected per rational-map lump as the sites whose map direction n(R(z)) moves by
    more than `tol` when the lump centre moves by `shift` in a generic direction (the profile is left
    out: far from a lump sin f is small and would hide the jump); `dilate` widens the mask by that many
    sites along each axis (for quantities built from stencils over the field)."""
    lumps, _ = initial_data.lump_specs(rp, cfg)
    rx, ry, rz = OIC.centered_position(p_i, p_f)
    step = shift * np.array([1.0, 0.37, -0.61])  # off every symmetry axis and diagonal
    mask = np.zeros(rx.shape, dtype=bool)
    for sp in lumps:
        if sp["mode"] != "rational_map":
            continue
        S = np.eye(3) if sp["space_rotation"] is None else np.asarray(sp["space_rotation"])
        bint = max(int(abs(sp["B"]) + 0.5), 1)

        def direction(c):
            x, y, z = OIC.apply_matrix(S, rx - c[0], ry - c[1], rz - c[2])
            z0, z1, _ = OIC.stereographic(x, y, z)
            return np.array(OIC.zton(*OIC.rational_map_value(z0, z1, bint)))

        c = np.asarray(sp["center"], dtype=np.float64)
        mask |= np.max(np.abs(direction(c + step) - direction(c)), axis=0) > tol
    for _ in range(dilate):
        grown = mask.copy()
        for ax in range(3):
            grown |= np.roll(mask, 1, axis=ax) | np.roll(mask, -1, axis=ax)
        mask = grown
    return mask


def rel_l2(a, b) -> float:
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))
