"""Per-site physics of the six theories (oracle; test infrastructure only).

Every function works on whole grids: ``phi`` is ``(nf, X, Y, Z)``, ``d1`` is
``(3, nf, X, Y, Z)`` and ``d2`` is ``(3, 3, nf, X, Y, Z)`` (see
:mod:`oracle.stencils`).  ``p_i`` / ``p_f`` are the reference's packed
parameter arrays, read at the indices the reference kernels use.

Each function restates one reference device function, in the reference's
loop and operand order.
"""

from __future__ import annotations

import math

import numpy as np

PION, OMEGA, COULOMB, RHO = "pion", "omega", "coulomb", "rho"
BERGER, LIGHTLY = "berger", "lightly"  # Berger-Skyrme, lightly bound Skyrme


# ---------------------------------------------------------------- helpers
def dot4(a, b):
    """``_dot4`` (``theories/nuclear_skyrme/kernels.py:103-123``)."""
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3]


def det3(a00, a01, a02, a10, a11, a12, a20, a21, a22):
    """``_det3`` (``theories/nuclear_skyrme/kernels.py:126-152``)."""
    return (a00 * (a11 * a22 - a12 * a21)
            - a01 * (a10 * a22 - a12 * a20)
            + a02 * (a10 * a21 - a11 * a20))


def det4_cols(c0, c1, c2, c3):
    """``_det4_cols`` (``theories/nuclear_skyrme/kernels.py:155-194``).

    ``c0..c3`` are the four columns (each a length-4 sequence of arrays);
    cofactor expansion along the first row.
    """
    m00, m10, m20, m30 = c0
    m01, m11, m21, m31 = c1
    m02, m12, m22, m32 = c2
    m03, m13, m23, m33 = c3
    det0 = det3(m11, m12, m13, m21, m22, m23, m31, m32, m33)
    det1 = det3(m10, m12, m13, m20, m22, m23, m30, m32, m33)
    det2 = det3(m10, m11, m13, m20, m21, m23, m30, m31, m33)
    det3_ = det3(m10, m11, m12, m20, m21, m22, m30, m31, m32)
    return m00 * det0 - m01 * det1 + m02 * det2 - m03 * det3_


def eps3(i, j, k) -> float:
    """``_eps3`` (``theories/rho_skyrme/kernels.py:59-81``)."""
    if i == j or j == k or i == k:
        return 0.0
    if (i, j, k) in ((0, 1, 2), (1, 2, 0), (2, 0, 1)):
        return 1.0
    return -1.0


def chern_simons(phi, d1):
    """``_chern_simons_current_point`` (``theories/omega_skyrme/kernels.py:443-500``).

    ``-det[phi, d_x phi, d_y phi, d_z phi] / (2 pi^2)`` over components 0..3.
    """
    det = det4_cols([phi[a] for a in range(4)], [d1[0, a] for a in range(4)],
                    [d1[1, a] for a in range(4)], [d1[2, a] for a in range(4)])
    return -det / (2.0 * math.pi * math.pi)


def _unit(mu, like):
    e = [np.zeros_like(like) for _ in range(4)]
    e[mu] = np.ones_like(like)
    return e


# ---------------------------------------------------------------- S^3 helpers
def n_sigma_fields(theory: str, p_i) -> int:
    """Number of S^3 components normalised / projected by the theory.

    nuclear: ``p_i[5]`` (``theories/nuclear_skyrme/kernels.py:225,267``);
    omega / Coulomb: ``p_i[11]`` (``theories/omega_skyrme/kernels.py:179,221``);
    rho: hard-coded 4 (``theories/rho_skyrme/kernels.py:243-288``).
    """
    if theory in (PION, BERGER, LIGHTLY):
        return int(p_i[5])  # berger / lightly bound: kernels.py compute_norm_skyrme_field, same as nuclear
    if theory in (OMEGA, COULOMB):
        return int(p_i[11])
    return 4


def normalise(F, nsk):
    """``compute_norm_skyrme_field``: divide comps 0..nsk-1 by their norm, skip 0.

    ``theories/nuclear_skyrme/kernels.py:225-234`` (s accumulated from 0.0
    in component order, ``math.sqrt``, early return when s == 0).
    """
    s = np.zeros_like(F[0])
    for a in range(nsk):
        s = s + F[a] * F[a]
    s = np.sqrt(s)
    nz = s != 0.0
    for a in range(nsk):
        F[a] = np.where(nz, F[a] / np.where(nz, s, 1.0), F[a])
    return F


def project(func, F, nsk):
    """``project_orthogonal_skyrme_field``: ``func -= (func . F) F`` on comps 0..nsk-1.

    ``theories/nuclear_skyrme/kernels.py:267-272``.
    """
    lm = np.zeros_like(F[0])
    for a in range(nsk):
        lm = lm + func[a] * F[a]
    for a in range(nsk):
        func[a] = func[a] - lm * F[a]
    return func


# ---------------------------------------------------------------- gradients
def gradient_pion(phi, d1, d2, p_i, p_f):
    """``do_gradient_step_point`` (``theories/nuclear_skyrme/kernels.py:1077-1150``).

    Returns ``(EnergyGradient, l)`` with ``l = -dt * EG`` after the projection
    onto the tangent space of the (unnormalised) ``phi``.
    """
    nf = int(p_i[5])
    dt = p_f[7]
    mpi = p_f[8]
    mpi2 = mpi * mpi
    kappa = p_f[9]
    kappa2 = kappa * kappa
    F = d1
    S = d2
    g = [np.zeros_like(phi[0]) for _ in range(nf)]
    norm = []
    for i in range(3):
        s = np.zeros_like(phi[0])
        for a in range(nf):
            s = s + F[i, a] * F[i, a]
        norm.append(s)
    for a in range(nf):
        g[a] = g[a] - (S[0, 0, a] + S[1, 1, a] + S[2, 2, a])
    for i in range(3):
        for j in range(3):
            if i == j:
                continue
            dnormj_di = np.zeros_like(phi[0])
            ddotij_di = np.zeros_like(phi[0])
            dotij = np.zeros_like(phi[0])
            for b in range(nf):
                dnormj_di = dnormj_di + 2.0 * F[j, b] * S[i, j, b]
                ddotij_di = ddotij_di + (S[i, i, b] * F[j, b] + F[i, b] * S[i, j, b])
                dotij = dotij + F[i, b] * F[j, b]
            for a in range(nf):
                g[a] = g[a] - kappa2 * (dnormj_di * F[i, a] + norm[j] * S[i, i, a]
                                        - ddotij_di * F[j, a] - dotij * S[i, j, a])
    g[0] = g[0] - mpi2
    EG = np.array(g)
    project(EG, phi, nf)
    return EG, -dt * EG


def _coupling_cofactor(mu, phi, Fj, Fk):
    """``_det4_cols(e_mu, phi, F_j, F_k)`` as used at ``omega kernels.py:717-722``."""
    e = _unit(mu, phi[0])
    return det4_cols(e, [phi[a] for a in range(4)], [Fj[a] for a in range(4)],
                     [Fk[a] for a in range(4)])


def _eps_signed_pairs():
    """(i, j, k, eps_ijk) over distinct triples in the reference's loop order."""
    out = []
    for i in range(3):
        for j in range(3):
            for k in range(3):
                if i == j or j == k or i == k:
                    continue
                out.append((i, j, k, 1.0 if (i, j, k) in ((0, 1, 2), (1, 2, 0), (2, 0, 1)) else -1.0))
    return out


def gradient_omega(phi, d1, d2, p_i, p_f):
    """``do_gradient_step_point`` (``theories/omega_skyrme/kernels.py:626-743``).

    Pion comps (projected):
    ``g_mu = -1/4 sum_i S_ii,mu + (c/4pi^2) sum eps_ijk d_i omega det[e_mu, phi, F_j, F_k] - m^2/4 delta_mu0``;
    omega: ``g_w = w - sum_i S_ii,w + c * CS`` (not projected).
    """
    nf = int(p_i[5])
    nsk = int(p_i[11])
    dt = p_f[7]
    mass = p_f[8]
    c_omega = p_f[9]
    mass2 = mass * mass
    four_pi2 = 4.0 * math.pi * math.pi
    omega = phi[4]
    g = [np.zeros_like(phi[0]) for _ in range(nf)]
    dom = [d1[0, 4], d1[1, 4], d1[2, 4]]
    triples = _eps_signed_pairs()
    for mu in range(nsk):
        for i in range(3):
            g[mu] = g[mu] - 0.25 * d2[i, i, mu]
            for (ii, j, k, eijk) in triples:
                if ii != i:
                    continue
                cof = _coupling_cofactor(mu, phi, d1[j], d1[k])
                g[mu] = g[mu] + c_omega / four_pi2 * eijk * dom[i] * cof
        if mu == 0:
            g[mu] = g[mu] - 0.25 * mass2
    EG = np.zeros((nf,) + phi.shape[1:])
    for mu in range(nsk):
        EG[mu] = g[mu]
    project(EG, phi, nsk)
    EG[4] = omega - (d2[0, 0, 4] + d2[1, 1, 4] + d2[2, 2, 4]) + c_omega * chern_simons(phi, d1)
    return EG, -dt * EG


def gradient_coulomb(phi, d1, d2, p_i, p_f):
    """``do_gradient_step_point`` (``theories/coulomb_skyrme/kernels.py:617-767``).

    Pion comps: the nuclear gradient with kappa^2 = 1 and m^2 (``:651-682``),
    then ``g_mu -= (kappa/4pi^2) sum eps_ijk d_i V det[e_mu, phi, F_j, F_k]``
    (``:697-749``), projected.  The dead store ``EG = 2 g`` (``:684-685``) is
    overwritten and has no effect.  V: ``g_V = -kappa sum_i S_ii,V - kappa CS``.
    """
    nf = int(p_i[5])
    nsk = int(p_i[11])
    dt = p_f[7]
    mass = p_f[8]
    kappa = p_f[9]
    mass2 = mass * mass
    four_pi2 = 4.0 * math.pi * math.pi
    F = d1
    S = d2
    g = [np.zeros_like(phi[0]) for _ in range(4)]
    norm = []
    for i in range(3):
        s = np.zeros_like(phi[0])
        for a in range(4):
            s = s + F[i, a] * F[i, a]
        norm.append(s)
    for a in range(4):
        g[a] = g[a] - (S[0, 0, a] + S[1, 1, a] + S[2, 2, a])
    for i in range(3):
        for j in range(3):
            if i == j:
                continue
            dnormj_di = np.zeros_like(phi[0])
            ddotij_di = np.zeros_like(phi[0])
            dotij = np.zeros_like(phi[0])
            for b in range(4):
                dnormj_di = dnormj_di + 2.0 * F[j, b] * S[i, j, b]
                ddotij_di = ddotij_di + (S[i, i, b] * F[j, b] + F[i, b] * S[i, j, b])
                dotij = dotij + F[i, b] * F[j, b]
            for a in range(4):
                g[a] = g[a] - (dnormj_di * F[i, a] + norm[j] * S[i, i, a]
                               - ddotij_di * F[j, a] - dotij * S[i, j, a])
    g[0] = g[0] - mass2
    dV = [d1[0, 4], d1[1, 4], d1[2, 4]]
    triples = _eps_signed_pairs()
    for mu in range(nsk):
        for i in range(3):
            for (ii, j, k, eijk) in triples:
                if ii != i:
                    continue
                cof = _coupling_cofactor(mu, phi, F[j], F[k])
                g[mu] = g[mu] - kappa / four_pi2 * eijk * dV[i] * cof
    EG = np.zeros((nf,) + phi.shape[1:])
    for mu in range(nsk):
        EG[mu] = g[mu]
    project(EG, phi, nsk)
    gV = np.zeros_like(phi[0])
    gV = gV - (kappa * (d2[0, 0, 4] + d2[1, 1, 4] + d2[2, 2, 4]) + kappa * chern_simons(phi, d1))
    EG[4] = gV
    return EG, -dt * EG


def idx_rho(i, a):
    """``idx_rho`` (``theories/rho_skyrme/kernels.py:84-88``): comp of rho^a_i."""
    return 4 + 3 * i + a


def _dR(S, a, j, k, l):
    """``_d_R_a_ijk`` (``theories/rho_skyrme/kernels.py:180-184``): d_l R^a_jk."""
    return S[j, l, idx_rho(k, a)] - S[k, l, idx_rho(j, a)]


def _R(F, a, i, j):
    """``_R_a_ij`` (``theories/rho_skyrme/kernels.py:172-176``)."""
    return F[i, idx_rho(j, a)] - F[j, idx_rho(i, a)]


def _Omega(F, a, i, j):
    """``_Omega_a_ij`` (``theories/rho_skyrme/kernels.py:188-196``)."""
    temp = F[i, 0] * F[j, a + 1] - F[i, a + 1] * F[j, 0]
    for b in range(3):
        for c in range(3):
            e = eps3(a, b, c)
            if e != 0.0:
                temp = temp + e * F[i, b + 1] * F[j, c + 1]
    return temp


def _dOmega(F, S, a, j, k, l):
    """``_d_Omega_a_ijk`` (``theories/rho_skyrme/kernels.py:200-212``)."""
    temp = np.zeros_like(F[0, 0])
    temp = temp + S[j, l, 0] * F[k, a + 1]
    temp = temp + F[j, 0] * S[k, l, a + 1]
    temp = temp - S[j, l, a + 1] * F[k, 0]
    temp = temp - F[j, a + 1] * S[k, l, 0]
    for b in range(3):
        for c in range(3):
            e = eps3(a, b, c)
            if e != 0.0:
                temp = temp + e * (S[j, l, b + 1] * F[k, c + 1] + F[j, b + 1] * S[k, l, c + 1])
    return temp


def gradient_rho(phi, d1, d2, p_i, p_f):
    """``do_gradient_step_point`` (``theories/rho_skyrme/kernels.py:607-682``).

    ``c_alpha`` is read from ``p_f[10]`` as the code does (the docstring's
    ``p_f[11]`` is wrong).
    """
    nf = int(p_i[5])
    dt = p_f[7]
    mpi = p_f[8]
    mpi2 = mpi * mpi
    mrho = p_f[9]
    mrho2 = mrho * mrho
    c_alpha = p_f[10]
    F = d1
    S = d2
    g = [np.zeros_like(phi[0]) for _ in range(nf)]
    for mu in range(4):
        for i in range(3):
            g[mu] = g[mu] - 2.0 * S[i, i, mu]
            for j in range(3):
                for nu in range(4):
                    g[mu] = g[mu] - 2.0 * (S[i, i, mu] * (F[j, nu] * F[j, nu])
                                           + (F[i, mu] * (S[i, j, nu] * F[j, nu] - S[j, j, nu] * F[i, nu])
                                              - S[i, j, mu] * F[i, nu] * F[j, nu]))
        if mu == 0:
            g[mu] = g[mu] - 2.0 * mpi2
    for a in range(3):
        for i in range(3):
            for j in range(3):
                dR = _dR(S, a, i, j, i)
                g[0] = g[0] + 16.0 * c_alpha * F[j, a + 1] * dR
                g[a + 1] = g[a + 1] - 16.0 * c_alpha * F[j, 0] * dR
                for b in range(3):
                    for c in range(3):
                        e = eps3(b, a, c)
                        # the reference adds the e == 0 terms too; they are exact zeros
                        if e != 0.0:
                            g[a + 1] = g[a + 1] + 16.0 * c_alpha * e * F[j, c + 1] * _dR(S, b, i, j, i)
    for a in range(3):
        for i in range(3):
            ai = idx_rho(i, a)
            val = 8.0 * mrho2 * phi[ai]
            for j in range(3):
                val = val + 8.0 * _dR(S, a, i, j, j)
                val = val - 16.0 * c_alpha * _dOmega(F, S, a, i, j, j)
            g[ai] = val
    EG = np.array(g)
    project(EG, phi, 4)
    return EG, -dt * EG


def _omega4(a, b):
    """``_omega_contract`` (``theories/berger_skyrme/kernels.py:81-111``): a0 b3 + a1 b2 - a2 b1 - a3 b0."""
    return a[0] * b[3] + a[1] * b[2] - a[2] * b[1] - a[3] * b[0]


def _omega_vec(v):
    """The vector Omega v with a.(Omega v) = omega(a, v): (v3, v2, -v1, -v0) (``berger_skyrme/kernels.py:659-677``)."""
    return [v[3], v[2], -v[1], -v[0]]


def _quartic_terms(F, S, nf, like):
    """The i != j loop of the standard quartic variation, per component, without its coupling factor."""
    norm = [sum(F[i, b] * F[i, b] for b in range(nf)) for i in range(3)]
    out = [np.zeros_like(like) for _ in range(nf)]
    for i in range(3):
        for j in range(3):
            if i == j:
                continue
            dnormj_di = np.zeros_like(like)
            ddotij_di = np.zeros_like(like)
            dotij = np.zeros_like(like)
            for b in range(nf):
                dnormj_di = dnormj_di + 2.0 * F[j, b] * S[i, j, b]
                ddotij_di = ddotij_di + (S[i, i, b] * F[j, b] + F[i, b] * S[i, j, b])
                dotij = dotij + F[i, b] * F[j, b]
            for a in range(nf):
                out[a] = out[a] + (dnormj_di * F[i, a] + norm[j] * S[i, i, a] - ddotij_di * F[j, a]
                                   - dotij * S[i, j, a])
    return out


def gradient_berger(phi, d1, d2, p_i, p_f):
    """``do_gradient_step_point`` (``theories/berger_skyrme/kernels.py:554-737``).

    Nuclear gradient plus the squashing terms (beta = alpha^2 - 1, alpha =
    ``p_f[12]``): ``-beta [2 sum_i u_i Omega F_i + (sum_i omega(S_ii, phi)) Omega phi]``
    with ``u_i = omega(F_i, phi)`` (``:642-682``) and
    ``-kappa^2 beta sum_{i<j} [d_i W_ij Omega F_j - d_j W_ij Omega F_i]`` with
    ``W_ij = omega(F_i, F_j)`` (``:705-724``); then ``- mpi^2`` on sigma and the
    tangent projection.
    """
    dt = p_f[7]
    mpi2 = p_f[8] * p_f[8]
    kappa2 = p_f[9] * p_f[9]
    alpha = p_f[12]
    beta = alpha * alpha - 1.0
    F, S = d1, d2
    g = [-(S[0, 0, a] + S[1, 1, a] + S[2, 2, a]) for a in range(4)]
    u = [_omega4(F[i], phi) for i in range(3)]
    div_u = _omega4(S[0, 0], phi) + _omega4(S[1, 1], phi) + _omega4(S[2, 2], phi)
    op = _omega_vec(phi)
    of = [_omega_vec(F[i]) for i in range(3)]
    for a in range(4):
        g[a] = g[a] - beta * (2.0 * u[0] * of[0][a] + 2.0 * u[1] * of[1][a] + 2.0 * u[2] * of[2][a]
                              + div_u * op[a])
    q = _quartic_terms(F, S, 4, phi[0])
    for a in range(4):
        g[a] = g[a] - kappa2 * q[a]
    pairs = [(0, 1), (0, 2), (1, 2)]
    for a in range(4):
        acc = np.zeros_like(phi[0])
        for i, j in pairs:
            dWi = _omega4(S[i, i], F[j]) + _omega4(F[i], S[i, j])  # d_i W_ij
            dWj = _omega4(S[i, j], F[j]) + _omega4(F[i], S[j, j])  # d_j W_ij
            acc = acc + (dWi * of[j][a] - dWj * of[i][a])
        g[a] = g[a] - kappa2 * beta * acc
    g[0] = g[0] - mpi2
    EG = np.array(g)
    project(EG, phi, 4)
    return EG, -dt * EG


def gradient_lightly(phi, d1, d2, p_i, p_f):
    """``do_gradient_step_point`` (``theories/lightly_bound_skyrme/kernels.py:567-683``).

    ``-2 (1 - alpha) Laplacian phi`` minus the quartic variation with unit
    coupling, then ``- 2 (1 - alpha) mpi^2 - 4 alpha (1 - sigma)^3`` on sigma
    (alpha = ``p_f[9]``) and the tangent projection.
    """
    nf = int(p_i[5])
    dt = p_f[7]
    mpi2 = p_f[8] * p_f[8]
    alpha = p_f[9]
    F, S = d1, d2
    g = [-2.0 * (1.0 - alpha) * (S[0, 0, a] + S[1, 1, a] + S[2, 2, a]) for a in range(nf)]
    q = _quartic_terms(F, S, nf, phi[0])
    for a in range(nf):
        g[a] = g[a] - q[a]
    s0 = phi[0]
    g[0] = g[0] - (1.0 - alpha) * 2.0 * mpi2
    g[0] = g[0] - 4.0 * alpha * (1.0 - s0) * (1.0 - s0) * (1.0 - s0)
    EG = np.array(g)
    project(EG, phi, nf)
    return EG, -dt * EG


GRADIENTS = {PION: gradient_pion, OMEGA: gradient_omega, COULOMB: gradient_coulomb, RHO: gradient_rho,
             BERGER: gradient_berger, LIGHTLY: gradient_lightly}


# ---------------------------------------------------------------- energy densities
def energy_pion(phi, d1, p_i, p_f):
    """``compute_energy_point`` (``theories/nuclear_skyrme/kernels.py:312-347``), times vol."""
    vol = p_f[6]
    mpi = p_f[8]
    mpi2 = mpi * mpi
    kappa = p_f[9]
    kappa2 = kappa * kappa
    fx = [d1[0, a] for a in range(4)]
    fy = [d1[1, a] for a in range(4)]
    fz = [d1[2, a] for a in range(4)]
    nx, ny, nz = dot4(fx, fx), dot4(fy, fy), dot4(fz, fz)
    dxy, dxz, dyz = dot4(fx, fy), dot4(fx, fz), dot4(fy, fz)
    energy = 0.5 * (nx + ny + nz)
    energy = energy + 0.5 * kappa2 * ((nx * ny - dxy * dxy) + (nx * nz - dxz * dxz) + (ny * nz - dyz * dyz))
    energy = energy + mpi2 * (1.0 - phi[0])
    return energy * vol


def energy_omega(phi, d1, p_i, p_f):
    """``compute_energy_point`` (``theories/omega_skyrme/kernels.py:266-302``), times vol.

    The coupling term is commented out in the reference (``:301``) and is
    therefore absent here too.
    """
    vol = p_f[6]
    mass = p_f[8]
    mass2 = mass * mass
    omega = phi[4]
    fx = [d1[0, a] for a in range(4)]
    fy = [d1[1, a] for a in range(4)]
    fz = [d1[2, a] for a in range(4)]
    nx, ny, nz = dot4(fx, fx), dot4(fy, fy), dot4(fz, fz)
    energy = 0.125 * (nx + ny + nz)
    energy = energy + mass2 * 0.25 * (1.0 - phi[0])
    energy = energy + 0.5 * omega * omega
    energy = energy + 0.5 * (d1[0, 4] * d1[0, 4] + d1[1, 4] * d1[1, 4] + d1[2, 4] * d1[2, 4])
    return energy * vol


def energy_coulomb(phi, d1, p_i, p_f):
    """``compute_energy_point`` (``theories/coulomb_skyrme/kernels.py:266-306``), times vol.

    Twice the nuclear form with kappa = 1; the coupling term is commented out
    in the reference (``:305``).
    """
    vol = p_f[6]
    mass = p_f[8]
    kappa = p_f[9]
    mass2 = mass * mass
    fx = [d1[0, a] for a in range(4)]
    fy = [d1[1, a] for a in range(4)]
    fz = [d1[2, a] for a in range(4)]
    nx, ny, nz = dot4(fx, fx), dot4(fy, fy), dot4(fz, fz)
    dxy, dxz, dyz = dot4(fx, fy), dot4(fx, fz), dot4(fy, fz)
    energy = nx + ny + nz
    energy = energy + ((nx * ny - dxy * dxy) + (nx * nz - dxz * dxz) + (ny * nz - dyz * dyz))
    energy = energy + 2.0 * mass2 * (1.0 - phi[0])
    energy = energy + 0.5 * kappa * (d1[0, 4] * d1[0, 4] + d1[1, 4] * d1[1, 4] + d1[2, 4] * d1[2, 4])
    return energy * vol


def energy_rho(phi, d1, p_i, p_f):
    """``compute_energy_point`` (``theories/rho_skyrme/kernels.py:302-347``), times vol."""
    vol = p_f[6]
    mpi = p_f[8]
    mpi2 = mpi * mpi
    mrho = p_f[9]
    mrho2 = mrho * mrho
    c_alpha = p_f[10]
    F = d1
    energy = 0.0 * phi[0]
    for i in range(3):
        for mu in range(4):
            energy = energy + F[i, mu] * F[i, mu]
    for i in range(3):
        for j in range(3):
            for mu in range(4):
                for nu in range(4):
                    energy = energy + 0.5 * ((F[i, mu] * F[j, nu]) * (F[i, mu] * F[j, nu])
                                             - F[i, mu] * F[j, mu] * F[i, nu] * F[j, nu])
    energy = energy + 2.0 * mpi2 * (1.0 - phi[0])
    for a in range(3):
        for i in range(3):
            rho_ai = phi[idx_rho(i, a)]
            energy = energy + 4.0 * mrho2 * rho_ai * rho_ai
            for j in range(3):
                Rij = _R(F, a, i, j)
                energy = energy + 2.0 * Rij * Rij
                energy = energy - 8.0 * c_alpha * Rij * _Omega(F, a, i, j)
    return energy * vol


def energy_berger(phi, d1, p_i, p_f):
    """``compute_energy_point`` (``theories/berger_skyrme/kernels.py:251-337``), times vol.

    Nuclear energy plus ``beta/2 sum_i u_i^2 + kappa^2 beta/2 sum_{i<j} W_ij^2``.
    """
    vol = p_f[6]
    mpi2 = p_f[8] * p_f[8]
    kappa2 = p_f[9] * p_f[9]
    alpha = p_f[12]
    beta = alpha * alpha - 1.0
    f = [[d1[i, a] for a in range(4)] for i in range(3)]
    nx, ny, nz = dot4(f[0], f[0]), dot4(f[1], f[1]), dot4(f[2], f[2])
    dxy, dxz, dyz = dot4(f[0], f[1]), dot4(f[0], f[2]), dot4(f[1], f[2])
    ux, uy, uz = (_omega4(f[i], phi) for i in range(3))
    wxy, wxz, wyz = _omega4(f[0], f[1]), _omega4(f[0], f[2]), _omega4(f[1], f[2])
    energy = 0.5 * (nx + ny + nz)
    energy = energy + 0.5 * beta * (ux * ux + uy * uy + uz * uz)
    energy = energy + 0.5 * kappa2 * ((nx * ny - dxy * dxy) + (nx * nz - dxz * dxz) + (ny * nz - dyz * dyz))
    energy = energy + 0.5 * kappa2 * beta * (wxy * wxy + wxz * wxz + wyz * wyz)
    energy = energy + mpi2 * (1.0 - phi[0])
    return energy * vol


def energy_lightly(phi, d1, p_i, p_f):
    """``compute_energy_point`` (``theories/lightly_bound_skyrme/kernels.py:275-345``), times vol.

    ``(1 - alpha)|d phi|^2 + quartic/2 + 2 (1 - alpha) mpi^2 (1 - sigma) + alpha (1 - sigma)^4``.
    """
    vol = p_f[6]
    mpi2 = p_f[8] * p_f[8]
    alpha = p_f[9]
    f = [[d1[i, a] for a in range(4)] for i in range(3)]
    nx, ny, nz = dot4(f[0], f[0]), dot4(f[1], f[1]), dot4(f[2], f[2])
    dxy, dxz, dyz = dot4(f[0], f[1]), dot4(f[0], f[2]), dot4(f[1], f[2])
    om = 1.0 - phi[0]
    energy = (1.0 - alpha) * (nx + ny + nz)
    energy = energy + 0.5 * ((nx * ny - dxy * dxy) + (nx * nz - dxz * dxz) + (ny * nz - dyz * dyz))
    energy = energy + (1.0 - alpha) * 2.0 * mpi2 * om
    energy = energy + alpha * om * om * om * om
    return energy * vol


ENERGIES = {PION: energy_pion, OMEGA: energy_omega, COULOMB: energy_coulomb, RHO: energy_rho,
            BERGER: energy_berger, LIGHTLY: energy_lightly}


def baryon_density(theory, phi, d1, p_i, p_f):
    """Per-site baryon density times the cell volume.

    nuclear / rho / Berger / lightly bound: ``-det[phi, F_x, F_y, F_z] * (vol / (2 pi^2))``
    (``theories/nuclear_skyrme/kernels.py:441-447``,
    ``theories/berger_skyrme/kernels.py:374-433``,
    ``theories/lightly_bound_skyrme/kernels.py:383-446``); omega / Coulomb:
    ``vol * CS`` (``theories/omega_skyrme/kernels.py:374-375``).
    """
    vol = p_f[6]
    if theory in (OMEGA, COULOMB):
        return vol * chern_simons(phi, d1)
    charge = det4_cols([phi[a] for a in range(4)], [d1[0, a] for a in range(4)],
                       [d1[1, a] for a in range(4)], [d1[2, a] for a in range(4)])
    return -charge * (vol / (2.0 * math.pi * math.pi))


def energy_comps(theory, nf):
    """Components whose first derivatives the energy kernel computes (all nf)."""
    return range(nf)
