"""Synthetic benchmark workloads (seeded host initial data + packed parameters).

Each single-lattice builder returns ``(theory_module, resolved_params, p_i,
p_f, F0)``.  The time step is the reference's example setting
``courant = 0.2`` (``examples/nuclear_skyrme_gl.py:13``): the larger
dt = 0.5 dx these workloads were specified with is beyond the RK4 stability
limit of the reference's own scheme (the run goes non-finite within tens of
steps).  The rho theory takes ``courant = 0.05``: its Laplacian enters the
gradient with weight 8 (the pion's with 2), and the B >= 2 cores of the
serving requests stiffen it further.  At 0.1, five of the eight rho requests
of the 48-request serving stream (seed 0) went non-finite within 100-240
steps, and one at 0.07; at 0.05 all eight stayed finite and converged or kept
relaxing.
"""

from __future__ import annotations

import numpy as np

from skyrmion_solver_amd import initial_data
from skyrmion_solver_amd.theories import load_theory

COURANT = 0.2
COURANT_BY_THEORY = {"pion": 0.2, "omega": 0.2, "coulomb": 0.2, "rho": 0.05, "berger": 0.2, "lightly": 0.2}
THEORY_NAMES = {"pion": "Nuclear Skyrme model", "omega": "Omega-Skyrme model", "coulomb": "Coulomb-Skyrme model",
                "rho": "Rho-Skyrme model", "berger": "Berger Skyrme model", "lightly": "Lightly bound Skyrme model"}
SERVE_MIX = (("pion", 0.60), ("omega", 0.25), ("rho", 0.15))
# step budget of a serving request per theory, as a multiple of the stream's max_steps: the omega saddle
# flow (pion field descending while omega ascends) relaxes slowly.  Measured at 64^3 (seed 0, FP32 /
# FP64 steps to max|g| < 1e-4): B = 1 9.6k / 10.2k, B = 2 22.6k / 23.5k, B = 3 17.3k / 20.2k,
# B = 4 48.4k / 66.4k, B = 5 23.3k / 21.9k, B = 6 73.1k / 65.3k, B = 7 79.3k / > 80k, B = 8 > 80k.
STEP_BUDGET_FACTOR = {"pion": 1, "omega": 32, "rho": 1}


def _lattice(L, dx, courant=COURANT):
    return dict(xlen=L, ylen=L, zlen=L, lsx=dx, lsy=dx, lsz=dx, xsize=dx * (L - 1), ysize=dx * (L - 1),
                zsize=dx * (L - 1), courant=courant)


def c1_hedgehog(seed: int = 0, L: int = 32):
    """Pion Skyrme, B=1 hedgehog f = pi e^{-r}, 32^3, dx = 0.2, m = 0, random SU(2) isorotation, centre jitter 0.1."""
    th = load_theory("Nuclear Skyrme model")
    rp = th.params.default_params(**_lattice(L, 0.2), mpi=0.0, kappa=1.0).resolved()
    p_i, p_f = th.params.pack_device_params(rp)
    F = initial_data.build_field(rp, 4, dict(mode="hedgehog", profile=("exp", 1.0), jitter=0.1, isorotation="haar",
                                            seed=seed, vacuum_band=True))
    return th, rp, p_i, p_f, F


def c2_b4(seed: int = 0, L: int = 128):
    """Pion Skyrme, B=4 rational map f = pi e^{-r/1.5}, 128^3, dx = 0.15, m = 1, N(0, 1e-3) noise."""
    th = load_theory("Nuclear Skyrme model")
    rp = th.params.default_params(**_lattice(L, 0.15), mpi=1.0, kappa=1.0, baryon_number=4.0,
                                  ansatz="rational_map").resolved()
    p_i, p_f = th.params.pack_device_params(rp)
    F = initial_data.build_field(rp, 4, dict(mode="rational_map", baryon_number=4, profile=("exp", 1.5),
                                            noise=1.0e-3, seed=seed, vacuum_band=True))
    return th, rp, p_i, p_f, F


def c3_b8(theory: str = "omega", seed: int = 0, L: int = 160, device_init: bool = False):
    """Omega or Coulomb Skyrme, B=8 = product of two B=4 rational maps at -+1.5 along x, 160^3, dx = 0.15.

    omega: c_omega = 0.5 (omega mass 1 in kernel units), pion mass the
    derived default m_pi / m_omega; omega starts at -c_omega * CS.
    Coulomb: kappa = 0.1, pion mass the derived default; V starts at 0.
    With `device_init` the last entry is the descriptor that builds the field on the device (omega
    included) instead of the host field (whose omega the caller sets, ``set_omega_from_chern_simons``).
    """
    name = {"omega": "Omega-Skyrme model", "coulomb": "Coulomb-Skyrme model"}[theory]
    th = load_theory(name)
    extra = dict(c_omega=0.5) if theory == "omega" else dict(kappa=0.1)
    rp = th.params.default_params(**_lattice(L, 0.15), baryon_number=8.0, **extra).resolved()
    p_i, p_f = th.params.pack_device_params(rp)
    cfg = c3_b8_config(seed)
    if device_init:
        c_cs = float(rp.c_omega) if theory == "omega" else 0.0
        return th, rp, p_i, p_f, initial_data.device_init_desc(rp, cfg, omega_cs=c_cs)
    F = initial_data.build_field(rp, 5, cfg)
    return th, rp, p_i, p_f, F


def c3_b8_config(seed: int = 0) -> dict:
    """Initial-data request of the 160^3 B=8 workload: the product of two B=4 rational maps
    f = pi e^{-r/1.5} centred at -+1.5 along x, vacuum band."""
    lump = dict(mode="rational_map", baryon_number=4, profile=("exp", 1.5))
    return dict(mode="product", seed=seed, vacuum_band=True,
                lumps=[dict(lump, center=(-1.5, 0.0, 0.0)), dict(lump, center=(1.5, 0.0, 0.0))])


def c4_instances(n: int = 48, seed: int = 0, sizes=(48, 64, 96, 128), dx: float = 0.15, device_init: bool = False):
    """Independent pion lattices: L from `sizes`, B = 1..6 rational maps with random space and
    isospin orientations, m ~ U[0, 1.5].  Each entry carries the host field ``F``, or with
    `device_init` the ``init`` descriptor that builds the same field on the device."""
    rng = np.random.default_rng(seed)
    th = load_theory("Nuclear Skyrme model")
    out = []
    for k in range(n):
        L = int(rng.choice(sizes))
        B = int(rng.integers(1, 7))
        m = float(rng.uniform(0.0, 1.5))
        rp = th.params.default_params(**_lattice(L, dx), mpi=m, kappa=1.0, baryon_number=float(B)).resolved()
        p_i, p_f = th.params.pack_device_params(rp)
        cfg = dict(mode="rational_map", baryon_number=B, profile=("exp", 1.0 + 0.15 * B), isorotation="haar",
                   space_rotation="haar", seed=int(rng.integers(1 << 30)), vacuum_band=True)
        if device_init:
            out.append(dict(L=L, B=B, m=m, rp=rp, p_i=p_i, p_f=p_f, F=None, init=initial_data.device_init_desc(rp, cfg)))
        else:
            out.append(dict(L=L, B=B, m=m, rp=rp, p_i=p_i, p_f=p_f, F=initial_data.build_field(rp, 4, cfg), init=None))
    return out


def set_omega_from_chern_simons(engine, rp):
    """omega = -c_omega * CS of the uploaded pion components (``create_omega_field_kernel``,
    ``theories/omega_skyrme/initial_config.py:1495-1532``), computed with the engine's baryon density."""
    from skyrmion_solver_amd import _engine

    F = engine.download(_engine.FIELD)
    cs_vol = engine.density(_engine.DENSITY_BARYON)
    N = cs_vol.size
    F[4 * N:5 * N] = -float(rp.c_omega) * (cs_vol / float(rp.grid_volume))
    engine.upload(_engine.FIELD, F)
    return F


def serving_request(rid: int, theory: str, L: int, B: int, rng, tol: float = 1.0e-4, max_steps: int = 20000,
                    arrival: float = 0.0, device_init: bool = False):
    """One minimisation request: a B = 1..8 rational map (random isospin orientation, aligned with the
    lattice) on an L^3 lattice with seeded parameters:

    * pion: dx = 0.15, courant 0.2, m ~ U[0.5, 1.5], kappa = 1;
    * omega: the units of examples/omega_skyrme_gl.py:11-16 (dx = 12/63, courant 0.2, f_pi = 139.8,
      m_pi = 43.91, m_omega = 249.5) with beta_omega ~ U[7.5, 8.5] (c_omega ~ 14); omega starts at
      -c_omega CS;
    * rho: dx = 0.15, courant 0.05 (the rho Laplacian has weight 8), the physical masses of
      examples/rho_skyrme_gl.py:12-13 (e = 3.65, f_pi = 129, m_pi = 138, m_rho = 775) with
      alpha ~ U[0.055, 0.068]; rho starts at 0.

    With `device_init` the request carries the descriptor that builds its
    field on the device at admission (``Request.init``) instead of a host
    field.

    Space orientation stays aligned: a misaligned B >= 4 solution relaxes through a slow lattice
    rotation mode, which is physics of the discretisation rather than of the server.
    """
    from skyrmion_solver_amd.serving import Request

    th = load_theory(THEORY_NAMES[theory])
    if theory == "pion":
        lat = _lattice(L, 0.15, COURANT_BY_THEORY[theory])
        extra = dict(mpi=float(rng.uniform(0.5, 1.5)), kappa=1.0)
    elif theory == "omega":
        lat = _lattice(L, 12.0 / 63.0, COURANT_BY_THEORY[theory])
        extra = dict(f_pi=139.8, m_pi=43.91, m_omega=249.5, beta_omega=float(rng.uniform(7.5, 8.5)))
    elif theory == "rho":
        lat = _lattice(L, 0.15, COURANT_BY_THEORY[theory])
        extra = dict(e_skyrme=3.65, f_pi=129.0, mass_pion=138.0, mass_rho=775.0, alpha=float(rng.uniform(0.055, 0.068)))
    elif theory == "berger":
        lat = _lattice(L, 0.15, COURANT_BY_THEORY[theory])
        extra = dict(mpi=float(rng.uniform(0.5, 1.5)), kappa=1.0, alpha=float(rng.uniform(0.8, 1.2)))
    elif theory == "lightly":
        lat = _lattice(L, 0.15, COURANT_BY_THEORY[theory])
        extra = dict(mpi=float(rng.uniform(0.5, 1.5)), alpha=float(rng.uniform(0.0, 0.3)))
    else:
        lat = _lattice(L, 0.15, COURANT_BY_THEORY[theory])
        extra = dict(kappa=0.1)
    rp = th.params.default_params(**lat, baryon_number=float(B), **extra).resolved()
    p_i, p_f = th.params.pack_device_params(rp)
    nf = int(p_i[5])
    cfg = dict(mode="rational_map", baryon_number=B, profile=("exp", 1.0 + 0.15 * B), isorotation="haar",
               seed=int(rng.integers(1 << 30)), vacuum_band=True)
    c_cs = float(rp.c_omega) if theory == "omega" else 0.0
    if device_init:
        return Request(rid=rid, theory=theory, p_i=p_i, p_f=p_f, F0=None, tol=tol, max_steps=max_steps,
                       arrival=arrival, init=initial_data.device_init_desc(rp, cfg, omega_cs=c_cs))
    F = initial_data.build_field(rp, nf, cfg)
    if theory == "omega":
        F[4] = -c_cs * initial_data.chern_simons_density(F, rp)
    return Request(rid=rid, theory=theory, p_i=p_i, p_f=p_f, F0=F, tol=tol, max_steps=max_steps, arrival=arrival)


def c5_requests(n: int = 600, rate: float = 20.0, seed: int = 0, sizes=(64, 96, 128, 160, 192), mix=SERVE_MIX,
                tol: float = 1.0e-4, max_steps: int = 20000, device_init: bool = False, max_size_by_theory=None,
                budget_factor=None):
    """Poisson stream of `n` requests at `rate` per second: theory drawn from `mix`, L from `sizes`, B = 1..8;
    each request's step limit is `max_steps` times its theory's factor in `budget_factor` (default
    ``STEP_BUDGET_FACTOR``).  `max_size_by_theory` ({theory: L_max}) caps a theory's lattice size without
    changing the random draws, so every other request of the stream stays the same."""
    factors = STEP_BUDGET_FACTOR if budget_factor is None else budget_factor
    caps = max_size_by_theory or {}
    from skyrmion_solver_amd.serving import poisson_arrivals

    rng = np.random.default_rng(seed)
    arrivals = poisson_arrivals(n, rate, seed)
    names = [m[0] for m in mix]
    probs = np.array([m[1] for m in mix], dtype=float)
    probs /= probs.sum()
    out = []
    for k in range(n):
        theory = str(rng.choice(names, p=probs))
        L = int(rng.choice(sizes))
        if theory in caps:
            L = min(L, int(caps[theory]))
        B = int(rng.integers(1, 9))
        out.append(serving_request(k, theory, L, B, rng, tol=tol,
                                   max_steps=int(max_steps * factors.get(theory, 1)),
                                   arrival=float(arrivals[k]), device_init=device_init))
    return out
