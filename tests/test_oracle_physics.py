"""Pin the oracle's point gradients against its own energy densities.

The reference computes each theory's energy density (``compute_energy_point``)
and its Euler-Lagrange gradient (``do_gradient_step_point``) in two separate
hand-written functions.  For a density e(phi, F) the functional derivative is

    g_a = de/dphi_a - sum_i [ sum_{jb} d2e/dF_ia dF_jb S_ij,b + sum_b d2e/dF_ia dphi_b F_ib ]

at any point (phi, F = d_i phi, S = d_i d_j phi).  This is evaluated here
with ``torch.func`` on the oracle's *own* energy functions and compared with
the oracle's gradient functions on random per-site data, so a dropped term, a
wrong sign or factor, or a transposed index in either fails the test.

Conventions the reference fixes (documented, not "fixed"):

* pion: g = E-L(e) exactly;
* omega: pion components = E-L(e - c w CS), omega = E-L(e + c w CS): a
  saddle flow; the coupling term is commented out of the coded energy
  (theories/omega_skyrme/kernels.py:301);
* Coulomb: pion components = E-L(e_pion / 2 + kappa V CS) (the coded energy
  is twice the pion form, theories/coulomb_skyrme/kernels.py:300-302, while
  the gradient uses the 1x form), V = E-L(kappa |grad V|^2 / 2 - kappa V CS);
* rho: see the rho test.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import theories as OT


def _torch():
    torch = pytest.importorskip("torch")
    return torch


def random_point_data(nf, n=64, seed=0, unit=False):
    rng = np.random.default_rng(seed)
    phi = rng.normal(size=(nf, n))
    if unit:
        phi[:4] /= np.linalg.norm(phi[:4], axis=0)
    F = rng.normal(size=(3, nf, n))
    S = rng.normal(size=(3, 3, nf, n))
    S = 0.5 * (S + S.transpose(1, 0, 2, 3))
    return phi, F, S


def params(theory):
    p_i = np.zeros(16, dtype=np.int32)
    p_f = np.zeros(16)
    p_f[6] = 1.0  # vol
    p_f[7] = 0.1  # dt
    if theory == "pion":
        p_i[5] = 4
        p_f[8], p_f[9] = 0.8, 1.3
    elif theory in ("omega", "coulomb"):
        p_i[5], p_i[11] = 5, 4
        p_f[8], p_f[9] = 0.6, 0.45
    elif theory == "berger":
        p_i[5] = 4
        p_f[8], p_f[9], p_f[12] = 0.7, 1.1, 0.6  # alpha = 0.6: beta = -0.64
    elif theory == "lightly":
        p_i[5] = 4
        p_f[8], p_f[9] = 0.8, 0.3  # alpha = 0.3
    else:
        p_i[5] = 13
        p_f[8], p_f[9], p_f[10] = 0.5, 1.2, 0.3
    return p_i, p_f


def euler_lagrange(energy_fn, phi, F, S):
    """g_a = de/dphi_a - D_i (de/dF_ia) by the chain rule, on per-site data."""
    torch = _torch()
    from torch.func import grad, jvp

    phi_t = torch.tensor(phi, dtype=torch.float64)
    F_t = torch.tensor(F, dtype=torch.float64)
    S_t = torch.tensor(S, dtype=torch.float64)

    def etot(p, f):
        return energy_fn(p, f).sum()

    dphi = grad(etot, argnums=0)(phi_t, F_t)

    def GF(p, f):
        return grad(etot, argnums=1)(p, f)

    div = torch.zeros_like(phi_t)
    for i in range(3):
        _, t1 = jvp(lambda f: GF(phi_t, f), (F_t,), (S_t[i],))
        _, t2 = jvp(lambda p: GF(p, F_t), (phi_t,), (F_t[i],))
        div = div + t1[i] + t2[i]
    return (dphi - div).numpy()


def project(g, phi, nsk=4):
    g = g.copy()
    lm = np.sum(g[:nsk] * phi[:nsk], axis=0)
    g[:nsk] -= lm * phi[:nsk]
    return g


def test_pion_gradient_is_euler_lagrange_of_energy():
    p_i, p_f = params("pion")
    phi, F, S = random_point_data(4, seed=1)
    g_ref = euler_lagrange(lambda p, f: OT.energy_pion(p, f, p_i, p_f), phi, F, S)
    EG, l = OT.gradient_pion(phi, F, S, p_i, p_f)
    np.testing.assert_allclose(EG, project(g_ref, phi), rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(l, -p_f[7] * EG, rtol=0, atol=0)


def tangent_point_data(nf, n=64, seed=0):
    """phi on S^3 and first derivatives tangent to it (what the flow keeps)."""
    phi, F, S = random_point_data(nf, n, seed)
    phi[:4] /= np.linalg.norm(phi[:4], axis=0)
    for i in range(3):
        F[i, :4] -= np.sum(F[i, :4] * phi[:4], axis=0) * phi[:4]
    return phi, F, S


def test_omega_gradient_is_a_saddle_flow_of_the_coupled_energy():
    """Pion comps: E-L of (e - c w CS); omega comp: E-L of (e + c w CS).

    The coupling term is commented out of the coded energy
    (theories/omega_skyrme/kernels.py:301); the gradient
    (:670-740) descends in the pion field and ascends in omega for
    E = e_sigma - (w^2/2 + |grad w|^2/2 + c w CS), i.e. the two equations
    carry opposite coupling signs.  Holds on tangent data (the omega-times-
    Jacobian terms are normal to S^3 and are projected away).
    """
    p_i, p_f = params("omega")
    c = p_f[9]
    phi, F, S = tangent_point_data(5, seed=2)

    def e_pion(p, f):
        return OT.energy_omega(p, f, p_i, p_f) - c * p[4] * OT.chern_simons(p, f)

    def e_omega(p, f):
        return OT.energy_omega(p, f, p_i, p_f) + c * p[4] * OT.chern_simons(p, f)

    EG, _ = OT.gradient_omega(phi, F, S, p_i, p_f)
    np.testing.assert_allclose(EG[:4], project(euler_lagrange(e_pion, phi, F, S), phi)[:4], rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(EG[4], euler_lagrange(e_omega, phi, F, S)[4], rtol=1e-12, atol=1e-12)
    # the opposite (single-functional) convention does not hold
    assert np.abs(EG[:4] - project(euler_lagrange(e_omega, phi, F, S), phi)[:4]).max() > 1e-3


def test_coulomb_gradient_is_a_saddle_flow_of_half_the_coded_energy():
    """Pion comps: E-L of (e_pion/2 + kappa V CS); V: E-L of (kappa |grad V|^2 / 2 - kappa V CS).

    The coded Coulomb energy is twice the pion form
    (theories/coulomb_skyrme/kernels.py:300-302) while its gradient uses the
    1x form (:651-682); the coupling term is commented out (:305).
    """
    p_i, p_f = params("coulomb")
    kap = p_f[9]
    phi, F, S = tangent_point_data(5, seed=3)

    def split(p, f):
        dV2 = f[0, 4] * f[0, 4] + f[1, 4] * f[1, 4] + f[2, 4] * f[2, 4]
        return OT.energy_coulomb(p, f, p_i, p_f) - 0.5 * kap * dV2, dV2

    def e_pion(p, f):
        return 0.5 * split(p, f)[0] + kap * p[4] * OT.chern_simons(p, f)

    def e_v(p, f):
        return 0.5 * kap * split(p, f)[1] - kap * p[4] * OT.chern_simons(p, f)

    EG, _ = OT.gradient_coulomb(phi, F, S, p_i, p_f)
    np.testing.assert_allclose(EG[:4], project(euler_lagrange(e_pion, phi, F, S), phi)[:4], rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(EG[4], euler_lagrange(e_v, phi, F, S)[4], rtol=1e-12, atol=1e-12)


def test_rho_gradient_matches_euler_lagrange_of_energy():
    p_i, p_f = params("rho")
    phi, F, S = random_point_data(13, n=16, seed=4)
    g_ref = euler_lagrange(lambda p, f: OT.energy_rho(p, f, p_i, p_f), phi, F, S)
    EG, _ = OT.gradient_rho(phi, F, S, p_i, p_f)
    np.testing.assert_allclose(EG, project(g_ref, phi), rtol=1e-10, atol=1e-9)


@pytest.mark.parametrize("theory", ["berger", "lightly"])
def test_berger_and_lightly_bound_gradients_are_euler_lagrange(theory):
    """Both are plain gradient flows of their coded energies
    (theories/berger_skyrme/kernels.py:251-337 vs :554-737;
    theories/lightly_bound_skyrme/kernels.py:275-345 vs :567-683)."""
    p_i, p_f = params(theory)
    phi, F, S = random_point_data(4, seed=6 + len(theory))
    g_ref = euler_lagrange(lambda p, f: OT.ENERGIES[theory](p, f, p_i, p_f), phi, F, S)
    EG, l = OT.GRADIENTS[theory](phi, F, S, p_i, p_f)
    np.testing.assert_allclose(EG, project(g_ref, phi), rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(l, -p_f[7] * EG, rtol=0, atol=0)


def test_berger_at_alpha_one_is_the_nuclear_model():
    """alpha = 1 (beta = 0) recovers the standard model (berger kernels.py:275)."""
    p_i, p_f = params("berger")
    p_f[12] = 1.0
    phi, F, S = random_point_data(4, seed=9)
    np.testing.assert_array_equal(OT.energy_berger(phi, F, p_i, p_f), OT.energy_pion(phi, F, p_i, p_f))
    np.testing.assert_allclose(OT.gradient_berger(phi, F, S, p_i, p_f)[0], OT.gradient_pion(phi, F, S, p_i, p_f)[0],
                               rtol=1e-13, atol=1e-13)


def test_berger_squashing_terms_bite():
    """A dropped squashing term (beta) in the gradient fails the identity."""
    p_i, p_f = params("berger")
    phi, F, S = random_point_data(4, seed=10)
    g_ref = project(euler_lagrange(lambda p, f: OT.energy_berger(p, f, p_i, p_f), phi, F, S), phi)
    q = p_f.copy()
    q[12] = 1.0  # the gradient without the squashing terms
    assert np.max(np.abs(OT.gradient_berger(phi, F, S, p_i, q)[0] - g_ref)) > 1e-2


def test_planted_coupling_error_fails_the_identity():
    """A 1% error in the quartic coupling of the oracle gradient must fail the identity."""
    p_i, p_f = params("pion")
    phi, F, S = random_point_data(4, seed=5)
    g_ref = euler_lagrange(lambda p, f: OT.energy_pion(p, f, p_i, p_f), phi, F, S)
    p_f_bad = p_f.copy()
    p_f_bad[9] = p_f[9] * 1.01  # kappa off by 1% in the gradient only
    EG_bad, _ = OT.gradient_pion(phi, F, S, p_i, p_f_bad)
    assert np.max(np.abs(EG_bad - project(g_ref, phi))) > 1e-3


@pytest.mark.parametrize("theory", ["pion", "omega", "coulomb", "rho", "berger", "lightly"])
def test_vacuum_is_a_fixed_point(theory):
    p_i, p_f = params(theory)
    nf = int(p_i[5])
    phi = np.zeros((nf, 3))
    phi[0] = 1.0
    F = np.zeros((3, nf, 3))
    S = np.zeros((3, 3, nf, 3))
    EG, _ = OT.GRADIENTS[theory](phi, F, S, p_i, p_f)
    assert np.max(np.abs(EG)) == 0.0


def test_baryon_density_of_hedgehog_integrates_to_one():
    from skyrmion_solver_amd import initial_data
    from skyrmion_solver_amd.theories import load_theory
    from oracle import integrator as OI

    mod = load_theory("Nuclear Skyrme model")
    p = mod.params.default_params(xlen=40, ylen=40, zlen=40, lsx=0.25, lsy=0.25, lsz=0.25, mpi=1.0)
    rp = p.resolved()
    p_i, p_f = mod.params.pack_device_params(rp)
    F = initial_data.build_field(rp, 4, dict(mode="hedgehog", profile=("exp", 1.0)))
    st = OI.OracleState("pion", p_i, p_f, F)
    B = OI.baryon_number(st)
    assert abs(B - 1.0) < 0.02, B


def test_baryon_density_of_b4_rational_map():
    from skyrmion_solver_amd import initial_data
    from skyrmion_solver_amd.theories import load_theory
    from oracle import integrator as OI

    mod = load_theory("Nuclear Skyrme model")
    p = mod.params.default_params(xlen=48, ylen=48, zlen=48, lsx=0.25, lsy=0.25, lsz=0.25, mpi=1.0)
    rp = p.resolved()
    p_i, p_f = mod.params.pack_device_params(rp)
    F = initial_data.build_field(rp, 4, dict(mode="rational_map", baryon_number=4, profile=("exp", 1.5)))
    st = OI.OracleState("pion", p_i, p_f, F)
    B = OI.baryon_number(st)
    assert abs(B - 4.0) < 0.3, B  # degree 4 (the exp(-r/1.5) tail is truncated by the box)


def test_energy_invariant_under_isorotation():
    from skyrmion_solver_amd import initial_data
    from skyrmion_solver_amd.theories import load_theory
    from oracle import integrator as OI

    mod = load_theory("Nuclear Skyrme model")
    p = mod.params.default_params(xlen=24, ylen=24, zlen=24, lsx=0.3, lsy=0.3, lsz=0.3, mpi=0.8, kappa=1.2)
    rp = p.resolved()
    p_i, p_f = mod.params.pack_device_params(rp)
    F = initial_data.build_field(rp, 4, dict(mode="hedgehog", profile=("exp", 1.0)))
    R = initial_data.haar_rotation(np.random.default_rng(3))
    G = F.copy()
    G[1:4] = np.einsum("ij,j...->i...", R, F[1:4])
    e1 = OI.energy(OI.OracleState("pion", p_i, p_f, F))
    e2 = OI.energy(OI.OracleState("pion", p_i, p_f, G))
    assert abs(e1 - e2) < 1e-11 * abs(e1)
