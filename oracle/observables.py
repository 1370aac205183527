"""Post-convergence observables (oracle; test infrastructure only).

Restates, in the reference's order, the observable kernels of
``theories/nuclear_skyrme/kernels.py:484-1032`` and their host reductions in
``theories/nuclear_skyrme/observables.py:182-486``.  Centre of mass and RMS
radius exist for every theory with the same kernels
(``theories/omega_skyrme/kernels.py:503,545``,
``theories/coulomb_skyrme/kernels.py:501,543``,
``theories/rho_skyrme/kernels.py:514,556``); only the baryon density differs
in form.  Fields are ``(nf, X, Y, Z)``; only the four Skyrme components are
differentiated (``p_i[5]`` of them in the observable kernels); the grid is the
reference's uncentred one, ``x = lsx * i`` (``create_grid_kernel``,
``theories/nuclear_skyrme/kernels.py:50-56``).

Pinned by ``tests/test_oracle_observables.py``: exact translation covariance
of the centre of mass, isospin-rotation covariance of U, spatial-rotation
covariance of V and the quadrupole, tracelessness and symmetry, the virial
identity between the stress-tensor trace and the energy terms, and the
hedgehog's U = V = W isotropy.
"""

from __future__ import annotations

import math

import numpy as np

from oracle import stencils as OS
from oracle.theories import baryon_density, eps3


def grid(p_i, p_f):
    """Site coordinates ``(3, X, Y, Z)``: ``ls * index`` (``create_grid_kernel``)."""
    X, Y, Z = int(p_i[0]), int(p_i[1]), int(p_i[2])
    gx = p_f[3] * np.arange(X, dtype=np.float64)
    gy = p_f[4] * np.arange(Y, dtype=np.float64)
    gz = p_f[5] * np.arange(Z, dtype=np.float64)
    return np.array(np.meshgrid(gx, gy, gz, indexing="ij"))


def _d1(F, p_i, p_f):
    """D1 of the Skyrme components with the lattice's halo guard (``p_i[3]``)."""
    return OS.first_derivatives(F, (p_f[3], p_f[4], p_f[5]), int(p_i[3]), comps=range(4))


def _com_or(theory, F, p_i, p_f, com):
    return center_of_mass(theory, F, p_i, p_f) if com is None else np.asarray(com, dtype=np.float64)


def _rel(p_i, p_f, com):
    g = grid(p_i, p_f)
    return [g[i] - com[i] for i in range(3)]


def moments_sites(theory, F, p_i, p_f):
    """``compute_center_of_mass_kernel`` (``kernels.py:678-718``): per-site ``(b, x b, y b, z b)``."""
    b = baryon_density(theory, F, _d1(F, p_i, p_f), p_i, p_f)
    g = grid(p_i, p_f)
    return np.array([b, g[0] * b, g[1] * b, g[2] * b])


def moments(theory, F, p_i, p_f):
    """``_reduce_obs4`` of the centre-of-mass moments: ``(B, sum x b, sum y b, sum z b)``."""
    return moments_sites(theory, F, p_i, p_f).reshape(4, -1).sum(axis=1)


def center_of_mass(theory, F, p_i, p_f):
    """``compute_center_of_mass`` (``observables.py:182-222``): zeros when |B| < 1e-30."""
    m = moments(theory, F, p_i, p_f)
    if abs(m[0]) < 1.0e-30:
        return np.zeros(3)
    return m[1:4] / m[0]


def rms_sites(theory, F, p_i, p_f, com=None):
    """``compute_rms_radius_kernel`` (``kernels.py:720-759``): per-site ``|r|^2 |b|``."""
    com = _com_or(theory, F, p_i, p_f, com)
    b = baryon_density(theory, F, _d1(F, p_i, p_f), p_i, p_f)
    r = _rel(p_i, p_f, com)
    return (r[0] * r[0] + r[1] * r[1] + r[2] * r[2]) * np.abs(b)


def rms_radius(theory, F, p_i, p_f):
    """``compute_rms_radius`` (``observables.py:224-271``): sqrt(sum r^2 |b| / B) about the centre of mass."""
    com = center_of_mass(theory, F, p_i, p_f)
    B = baryon_density(theory, F, _d1(F, p_i, p_f), p_i, p_f).sum()
    if abs(B) < 1.0e-30:
        return 0.0
    return math.sqrt(max(rms_sites(theory, F, p_i, p_f, com).sum() / B, 0.0))


def quadrupole_sites(F, p_i, p_f, com=None):
    """``compute_quadrupole_tensor_kernel`` (``kernels.py:761-812``): per-site (3 r_i r_j - d_ij r^2) b / 2."""
    com = _com_or("pion", F, p_i, p_f, com)
    rho = 0.5 * baryon_density("pion", F, _d1(F, p_i, p_f), p_i, p_f)
    r = _rel(p_i, p_f, com)
    r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2]
    Q = np.zeros((3, 3) + F.shape[1:])
    for i in range(3):
        for j in range(3):
            val = 3.0 * r[i] * r[j]
            if i == j:
                val = val - r2
            Q[i, j] = val * rho
    return Q


def quadrupole(F, p_i, p_f, com=None):
    """``compute_quadrupole_tensor`` (``observables.py:396-438``)."""
    return quadrupole_sites(F, p_i, p_f, com).reshape(3, 3, -1).sum(axis=2)


def currents(F, d1):
    """``_build_left_and_iso_currents`` (``kernels.py:485-547``): L_k^a and T_i^a per site."""
    s, p = F[0], F[1:4]
    L = np.zeros((3, 3) + F.shape[1:])
    for k in range(3):
        ds, dp = d1[k, 0], d1[k, 1:4]
        cross = [p[1] * dp[2] - p[2] * dp[1], p[2] * dp[0] - p[0] * dp[2], p[0] * dp[1] - p[1] * dp[0]]
        for a in range(3):
            L[k, a] = s * dp[a] - p[a] * ds + cross[a]
    psq = p[0] * p[0] + p[1] * p[1] + p[2] * p[2]
    Tc = np.zeros((3, 3) + F.shape[1:])
    for i in range(3):
        for a in range(3):
            val = p[i] * p[a]
            if i == a:
                val = val - psq
            eps_term = np.zeros_like(s)
            for b in range(3):
                eps_term = eps_term + eps3(i, a, b) * p[b]
            Tc[i, a] = val + s * eps_term
    return L, Tc


def _dots(A, B):
    return np.array([[A[i, 0] * B[j, 0] + A[i, 1] * B[j, 1] + A[i, 2] * B[j, 2] for j in range(3)] for i in range(3)])


def _U0_and_A(F, d1, p_f):
    """``_compute_U0_and_A`` (``kernels.py:549-613``)."""
    k2 = p_f[9] * p_f[9]
    L, Tc = currents(F, d1)
    LL, TT = _dots(L, L), _dots(Tc, Tc)
    LT = np.array([[Tc[i, 0] * L[j, 0] + Tc[i, 1] * L[j, 1] + Tc[i, 2] * L[j, 2] for j in range(3)] for i in range(3)])
    U0, A = TT.copy(), LL.copy()
    for i in range(3):
        for j in range(3):
            qU = np.zeros_like(F[0])
            qA = np.zeros_like(F[0])
            for k in range(3):
                qU = qU + (TT[i, j] * LL[k, k] - LT[i, k] * LT[j, k])
                qA = qA + (LL[i, j] * LL[k, k] - LL[i, k] * LL[j, k])
            U0[i, j] = U0[i, j] + k2 * qU
            A[i, j] = A[i, j] + k2 * qA
    return U0, A, LL, LT


def U_sites(F, p_i, p_f, com=None):
    """``compute_U_tensor_kernel`` (``kernels.py:814-851``): per-site 2 U0_ij vol (``com`` unused)."""
    U0, _, _, _ = _U0_and_A(F, _d1(F, p_i, p_f), p_f)
    return 2.0 * U0 * p_f[6]


def U_tensor(F, p_i, p_f):
    """``compute_U_tensor`` (``observables.py:273-306``)."""
    return U_sites(F, p_i, p_f).reshape(3, 3, -1).sum(axis=2)


def V_sites(F, p_i, p_f, com=None):
    """``compute_V_tensor_kernel`` (``kernels.py:853-905``): per-site 2 eps_ilm eps_jnp r_l r_n A_mp vol."""
    com = _com_or("pion", F, p_i, p_f, com)
    _, A, _, _ = _U0_and_A(F, _d1(F, p_i, p_f), p_f)
    r = _rel(p_i, p_f, com)
    V = np.zeros((3, 3) + F.shape[1:])
    for i in range(3):
        for j in range(3):
            val = np.zeros_like(F[0])
            for l in range(3):
                for n in range(3):
                    for m in range(3):
                        for p in range(3):
                            e = eps3(i, l, m) * eps3(j, n, p)
                            if e != 0.0:
                                val = val + e * r[l] * r[n] * A[m, p]
            V[i, j] = 2.0 * val * p_f[6]
    return V


def V_tensor(F, p_i, p_f, com=None):
    """``compute_V_tensor`` (``observables.py:308-350``)."""
    return V_sites(F, p_i, p_f, com).reshape(3, 3, -1).sum(axis=2)


def W_sites(F, p_i, p_f, com=None):
    """``compute_W_tensor_kernel`` (``kernels.py:907-975``): per-site 2 eps_jlm r_l B_im vol."""
    com = _com_or("pion", F, p_i, p_f, com)
    k2 = p_f[9] * p_f[9]
    _, _, LL, LT = _U0_and_A(F, _d1(F, p_i, p_f), p_f)
    Bm = np.zeros((3, 3) + F.shape[1:])
    for i in range(3):
        for m in range(3):
            val = LT[i, m].copy()
            for k in range(3):
                val = val + k2 * (LT[i, m] * LL[k, k] - LT[i, k] * LL[m, k])
            Bm[i, m] = val
    r = _rel(p_i, p_f, com)
    W = np.zeros((3, 3) + F.shape[1:])
    for i in range(3):
        for j in range(3):
            val = np.zeros_like(F[0])
            for l in range(3):
                for m in range(3):
                    e = eps3(j, l, m)
                    if e != 0.0:
                        val = val + e * r[l] * Bm[i, m]
            W[i, j] = 2.0 * val * p_f[6]
    return W


def W_tensor(F, p_i, p_f, com=None):
    """``compute_W_tensor`` (``observables.py:352-394``)."""
    return W_sites(F, p_i, p_f, com).reshape(3, 3, -1).sum(axis=2)


def stress_tensor(F, d1, p_f):
    """``_compute_stress_tensor`` (``kernels.py:615-676``): T_ij = D_ij + k2 (trD D_ij - (D^2)_ij) - d_ij e."""
    mpi2 = p_f[8] * p_f[8]
    k2 = p_f[9] * p_f[9]
    L, _ = currents(F, d1)
    D = _dots(L, L)
    trD = D[0, 0] + D[1, 1] + D[2, 2]
    quartic = np.zeros_like(F[0])
    for i in range(3):
        for j in range(i + 1, 3):
            quartic = quartic + (D[i, i] * D[j, j] - D[i, j] * D[i, j])
    e = 0.5 * trD + 0.5 * k2 * quartic + mpi2 * (1.0 - F[0])
    T = np.zeros((3, 3) + F.shape[1:])
    for i in range(3):
        for j in range(3):
            D2 = D[i, 0] * D[0, j] + D[i, 1] * D[1, j] + D[i, 2] * D[2, j]
            T[i, j] = D[i, j] + k2 * (trD * D[i, j] - D2)
            if i == j:
                T[i, j] = T[i, j] - e
    return T, trD, quartic


def D_sites(F, p_i, p_f, com=None):
    """``compute_D_integrand_kernel`` (``kernels.py:977-1033``): per-site (r_i r_j T_ij - r^2 tr T / 3) vol."""
    com = _com_or("pion", F, p_i, p_f, com)
    T, _, _ = stress_tensor(F, _d1(F, p_i, p_f), p_f)
    r = _rel(p_i, p_f, com)
    r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2]
    tr = T[0, 0] + T[1, 1] + T[2, 2]
    s = np.zeros_like(F[0])
    for i in range(3):
        for j in range(3):
            s = s + r[i] * r[j] * T[i, j]
    s = s - (r2 * tr) / 3.0
    return s * p_f[6]


def D_integral(F, p_i, p_f, com=None):
    """Sum of the D-term integrand (``_sum_scalar_field`` in ``compute_D_term``)."""
    return D_sites(F, p_i, p_f, com).sum()


def D_term(F, p_i, p_f, energy):
    """``compute_D_term`` (``observables.py:440-486``): -(2/5) E * integral."""
    return -(2.0 / 5.0) * energy * D_integral(F, p_i, p_f)
