"""Initial configurations of the nuclear Skyrme theory, FP64 NumPy restatement.

TEST INFRASTRUCTURE ONLY (the checker of the engine's device-side initial
data and of the host builder); the engine never imports it.

Restates, per site and in the reference's operand order, the device
functions of ``theories/nuclear_skyrme/initial_config.py`` (vectorised over
a grid with NumPy, every array operation being the reference's scalar one
applied element-wise):

* ``rotate_by_y`` / ``rotate_by_z`` / ``rotate_xyz``       :21-112
* complex helpers ``cadd csub cmul cscale cdiv cpow_uint``  :114-281
* ``zton``                                                  :284-306
* ``ffun`` and the fitted radial profile ``profilefun``     :309-442
* the rational maps ``rational_map_value`` (B = 1..8, z^B)  :445-549
* ``calc_rational_map_skyrmion``                            :552-587
* ``quaternion_prod``                                       :590-627
* ``centered_position``                                     :664-698
* ``hedgehog_field``                                        :701-758
* ``create_initial_configuration_kernel`` (ansatz branch)   :880-975
* ``create_rational_map_skyrmion_kernel``                   :978-1078
* ``product_ansatz_kernel``                                 :1081-1126

with the grid of ``create_grid_kernel`` (``kernels.py:27-56``) and the S^3
renormalisation of ``compute_norm_skyrme_field`` (``kernels.py:196-234``).

Beside the reference's two entry points, :func:`lump` evaluates one lump
with its space and isospin rotations given as 3x3 matrices (the reference's
Euler angles are the special case ``euler_matrix``) and with, besides the
reference's fitted and Gaussian profiles, the exponential profile
``pi exp(-r / a)`` the benchmark configurations name.

How it is pinned (``tests/test_oracle_initial_config.py``): the profile and
map constants against a fixture extracted from the reference's source text
(``tests/golden/``); profilefun(0) = pi and its tail; the topological degree
of every map (the lattice baryon number of the fitted-profile Skyrmion is B,
B = 1..8) and of the hedgehog; the degree of a product equals the sum of the
degrees; the quaternion product against the textbook Hamilton product; the
Euler rotations against their matrices; a planted wrong map fails.
"""

from __future__ import annotations

import math

import numpy as np

PI = math.pi

# profilefun fits, initial_config.py:380-416 (m0 < 1: massless pion) and
# :417-438 / :421-438 (massive pion): coefficients (a, b, c, d, e) per B = 1..9
PROFILE_FIT_MASSLESS = {
    1: (0.630647, 0.0218873, -0.0245431, 0.00300978, -0.00010542),
    2: (0.162388, 0.297889, -0.0878554, 0.00931881, -0.000317398),
    3: (-0.000352877, 0.324957, -0.0792377, 0.00740042, -0.000232486),
    4: (-0.0469393, 0.294323, -0.0599714, 0.00457823, -0.000110399),
    5: (-0.0575367, 0.187057, -0.00254209, -0.00580857, 0.000517839),
    6: (-0.0450119, 0.104984, 0.0347624, -0.0119037, 0.000864464),
    7: (-0.0383968, 0.0703267, 0.0464174, -0.0130402, 0.000887014),
    8: (-0.013271, -0.00611001, 0.0751757, -0.0170388, 0.0010835),
    9: (0.00853506, -0.0576792, 0.0896121, -0.0181863, 0.00109042),
}
PROFILE_FIT_MASSIVE = {
    1: (0.839091, -0.148868, 0.222428, -0.0900171, 0.0107864),
    2: (0.219378, 0.485652, -0.154501, 0.0197873, -0.000774393),
    3: (0.0354001, 0.502948, -0.13181, 0.0142808, -0.000509656),
    4: (-0.0127361, 0.4427, -0.0941955, 0.00808476, -0.00024485),
    5: (-0.0568381, 0.337281, -0.0167685, -0.010309, 0.00119672),
    6: (-0.05592, 0.229393, 0.0498254, -0.0257167, 0.00248365),
    7: (-0.0575166, 0.196602, 0.056918, -0.0244016, 0.00219296),
    8: (-0.0335398, 0.0655383, 0.135378, -0.0430527, 0.00381564),
    9: (-0.0101412, -0.0285699, 0.181351, -0.0517179, 0.00441035),
}
# rational-map constants (initial_config.py:479, 494, 507-508, 518, 540)
MAP_CONSTANTS = {3: {"a": 1.0}, 4: {"a": 1.0}, 5: {"a": 3.0660, "b": 3.9334}, 6: {"a": 0.1585}, 8: {"a": 0.1352}}
MAP_EPS = 1.0e-9      # added to the B = 3, 6, 8 denominators (:476, 525, 546)
CDIV_EPS = 1.0e-18    # added to |den|^2 in cdiv (:249)


# ---------------------------------------------------------------- rotations (:21-112)
def rotate_by_y(x0, x1, x2, alpha):
    ca, sa = math.cos(alpha), math.sin(alpha)
    return ca * x0 - sa * x2, x1, sa * x0 + ca * x2


def rotate_by_z(x0, x1, x2, alpha):
    ca, sa = math.cos(alpha), math.sin(alpha)
    return ca * x0 + sa * x1, -sa * x0 + ca * x1, x2


def rotate_xyz(x0, x1, x2, a0, a1, a2):
    y0, y1, y2 = rotate_by_z(x0, x1, x2, a0)
    z0, z1, z2 = rotate_by_y(y0, y1, y2, a1)
    return rotate_by_z(z0, z1, z2, a2)


def euler_matrix(a0, a1, a2):
    """The 3x3 matrix M with M v = rotate_xyz(v, a0, a1, a2)."""
    M = np.zeros((3, 3))
    for j in range(3):
        e = [0.0, 0.0, 0.0]
        e[j] = 1.0
        M[:, j] = rotate_xyz(e[0], e[1], e[2], a0, a1, a2)
    return M


def apply_matrix(M, x0, x1, x2):
    M = np.asarray(M, dtype=np.float64)
    return (M[0, 0] * x0 + M[0, 1] * x1 + M[0, 2] * x2,
            M[1, 0] * x0 + M[1, 1] * x1 + M[1, 2] * x2,
            M[2, 0] * x0 + M[2, 1] * x1 + M[2, 2] * x2)


# ---------------------------------------------------------------- complex pairs (:114-281)
def cadd(a0, a1, b0, b1):
    return a0 + b0, a1 + b1


def csub(a0, a1, b0, b1):
    return a0 - b0, a1 - b1


def cmul(a0, a1, b0, b1):
    return a0 * b0 - a1 * b1, a0 * b1 + a1 * b0


def cscale(s, a0, a1):
    return s * a0, s * a1


def cdiv(a0, a1, b0, b1):
    den = b0 * b0 + b1 * b1 + CDIV_EPS
    return (a0 * b0 + a1 * b1) / den, (a1 * b0 - a0 * b1) / den


def cpow_uint(z0, z1, n):
    r0, r1 = 1.0, 0.0
    for _ in range(n):
        r0, r1 = cmul(r0, r1, z0, z1)
    return r0, r1


def zton(z0, z1):
    zsq = z0 * z0 + z1 * z1
    den = 1.0 + zsq
    return 2.0 * z0 / den, 2.0 * z1 / den, (1.0 - zsq) / den


# ---------------------------------------------------------------- profile (:309-442)
def ffun(r, a, b, c, d, e):
    poly = r * (a + r * (b + r * (c + r * (d + r * e))))
    return PI * (1.0 - np.tanh(np.abs(poly)))


def profilefun(r, bint: int, m0: float, m1: float = 0.0):
    """Fitted radial profile.  For B outside 1..9 the reference's coefficients stay 0: the massive
    branch then gives ffun(r, 0, ...) = pi, and in the m0 < 1 branch the massless term stays 0."""
    r = np.asarray(r, dtype=np.float64)
    zero = (0.0, 0.0, 0.0, 0.0, 0.0)
    if m0 < 1.0:
        f = ffun(r, *PROFILE_FIT_MASSLESS[bint]) if bint in PROFILE_FIT_MASSLESS else np.zeros_like(r)
        fm = ffun(r, *PROFILE_FIT_MASSIVE.get(bint, zero))
        out = (1.0 - m0) * f + m0 * fm
    else:
        out = ffun(r, *PROFILE_FIT_MASSIVE.get(bint, zero))
    out = np.where(r > 15.0, 0.0, out)
    return np.where(r < 1.0e-6, PI, out)


# ---------------------------------------------------------------- rational maps (:445-549)
def rational_map_value(z0, z1, bint: int):
    unit0, unit1 = 1.0, 0.0
    ii0, ii1 = 0.0, 1.0
    eps0, eps1 = MAP_EPS, 0.0
    s3 = math.sqrt(3.0)
    if bint == 1:
        return z0, z1
    if bint == 2:
        return cmul(z0, z1, z0, z1)
    if bint == 3:
        a = MAP_CONSTANTS[3]["a"]
        z2_0, z2_1 = cmul(z0, z1, z0, z1)
        num0, num1 = cscale(a * s3, ii0, ii1)
        num0, num1 = cmul(num0, num1, z2_0, z2_1)
        num0, num1 = csub(num0, num1, unit0, unit1)
        den0, den1 = csub(z2_0, z2_1, a * s3 * ii0, a * s3 * ii1)
        den0, den1 = cmul(z0, z1, den0, den1)
        den0, den1 = cadd(den0, den1, eps0, eps1)
        return cdiv(num0, num1, den0, den1)
    if bint == 4:
        a = MAP_CONSTANTS[4]["a"]
        z4_0, z4_1 = cpow_uint(z0, z1, 4)
        t0, t1 = cscale(2.0 * s3, ii0, ii1)
        t0, t1 = cmul(t0, t1, z0, z1)
        t0, t1 = cmul(t0, t1, z0, z1)
        num0, num1 = cadd(z4_0, z4_1, t0, t1)
        num0, num1 = cadd(num0, num1, unit0, unit1)
        den0, den1 = csub(z4_0, z4_1, t0, t1)
        den0, den1 = cadd(den0, den1, unit0, unit1)
        num0, num1 = cscale(a, num0, num1)
        return cdiv(num0, num1, den0, den1)
    if bint == 5:
        a, b = MAP_CONSTANTS[5]["a"], MAP_CONSTANTS[5]["b"]
        z4_0, z4_1 = cpow_uint(z0, z1, 4)
        z2_0, z2_1 = cmul(z0, z1, z0, z1)
        num0, num1 = cadd(z4_0, z4_1, b * z2_0, b * z2_1)
        num0, num1 = cadd(num0, num1, a, 0.0)
        num0, num1 = cmul(z0, z1, num0, num1)
        den0, den1 = cscale(a, z4_0, z4_1)
        den0, den1 = csub(den0, den1, b * z2_0, b * z2_1)
        den0, den1 = cadd(den0, den1, unit0, unit1)
        return cdiv(num0, num1, den0, den1)
    if bint == 6:
        a = MAP_CONSTANTS[6]["a"]
        z4_0, z4_1 = cpow_uint(z0, z1, 4)
        num0, num1 = cadd(z4_0, z4_1, a * ii0, a * ii1)
        den0, den1 = cmul(a * ii0, a * ii1, z4_0, z4_1)
        den0, den1 = cadd(den0, den1, unit0, unit1)
        z2_0, z2_1 = cmul(z0, z1, z0, z1)
        den0, den1 = cmul(z2_0, z2_1, den0, den1)
        den0, den1 = cadd(den0, den1, eps0, eps1)
        return cdiv(num0, num1, den0, den1)
    if bint == 7:
        z7_0, z7_1 = cpow_uint(z0, z1, 7)
        z5_0, z5_1 = cpow_uint(z0, z1, 5)
        z2_0, z2_1 = cmul(z0, z1, z0, z1)
        num0, num1 = csub(z7_0, z7_1, 7.0 * z5_0, 7.0 * z5_1)
        num0, num1 = csub(num0, num1, 7.0 * z2_0, 7.0 * z2_1)
        num0, num1 = csub(num0, num1, unit0, unit1)
        den0, den1 = cadd(z7_0, z7_1, 7.0 * z5_0, 7.0 * z5_1)
        den0, den1 = csub(den0, den1, 7.0 * z2_0, 7.0 * z2_1)
        den0, den1 = cadd(den0, den1, unit0, unit1)
        return cdiv(num0, num1, den0, den1)
    if bint == 8:
        a = MAP_CONSTANTS[8]["a"]
        z6_0, z6_1 = cpow_uint(z0, z1, 6)
        num0, num1 = csub(z6_0, z6_1, a, 0.0)
        den0, den1 = cscale(a, z6_0, z6_1)
        den0, den1 = cadd(den0, den1, unit0, unit1)
        z2_0, z2_1 = cmul(z0, z1, z0, z1)
        den0, den1 = cmul(z2_0, z2_1, den0, den1)
        den0, den1 = cadd(den0, den1, eps0, eps1)
        return cdiv(num0, num1, den0, den1)
    return cpow_uint(z0, z1, bint)


def stereographic(rx, ry, rz):
    """z = tan(theta / 2) e^{i phi} of the direction (rx, ry, rz), 0 at the origin
    (initial_config.py:941-951 and :1057-1070)."""
    r = np.sqrt(rx * rx + ry * ry + rz * rz)
    small = r < 1.0e-6
    safe = np.where(small, 1.0, r)
    theta = np.where(small, 0.0, np.arccos(np.clip(rz / safe, -1.0, 1.0)))
    phi = np.where(small, 0.0, np.arctan2(ry, rx))
    t = np.tan(0.5 * theta)
    return np.cos(phi) * t, np.sin(phi) * t, r


def calc_rational_map_skyrmion(f, bint, z0, z1, beta0, beta1, beta2):
    r0, r1 = rational_map_value(z0, z1, bint)
    nx, ny, nz = zton(r0, r1)
    nx, ny, nz = rotate_xyz(nx, ny, nz, beta0, beta1, beta2)
    sf = np.sin(f)
    return np.cos(f), nx * sf, ny * sf, nz * sf


# ---------------------------------------------------------------- SU(2) (:590-627) and S^3 norm
def quaternion_prod(a, b):
    a0, a1, a2, a3 = a
    b0, b1, b2, b3 = b
    c0 = a0 * b0 - a1 * b1 - a2 * b2 - a3 * b3
    c1 = a0 * b1 + b0 * a1 + a2 * b3 - a3 * b2
    c2 = a0 * b2 + b0 * a2 + a3 * b1 - a1 * b3
    c3 = a0 * b3 + b0 * a3 + a1 * b2 - a2 * b1
    return np.array([c0, c1, c2, c3])


def normalise(F, ncomp: int = 4):
    """compute_norm_skyrme_field (kernels.py:196-234): divide the first `ncomp` components by their norm,
    sites with norm exactly 0 untouched."""
    s = np.sqrt(sum(F[a] * F[a] for a in range(ncomp)))
    nz = s != 0.0
    for a in range(ncomp):
        F[a] = np.where(nz, F[a] / np.where(nz, s, 1.0), F[a])
    return F


# ---------------------------------------------------------------- grids and entry points
def grid(p_i, p_f):
    """create_grid_kernel (kernels.py:27-56): site (x, y, z) at (lsx x, lsy y, lsz z)."""
    X, Y, Z = int(p_i[0]), int(p_i[1]), int(p_i[2])
    gx = p_f[3] * np.arange(X, dtype=np.float64)
    gy = p_f[4] * np.arange(Y, dtype=np.float64)
    gz = p_f[5] * np.arange(Z, dtype=np.float64)
    return np.meshgrid(gx, gy, gz, indexing="ij")


def centered_position(p_i, p_f):
    """centered_position (initial_config.py:664-698): coordinates minus the midpoint of the first and
    last site."""
    gx, gy, gz = grid(p_i, p_f)
    return (gx - 0.5 * (gx[0, 0, 0] + gx[-1, -1, -1]), gy - 0.5 * (gy[0, 0, 0] + gy[-1, -1, -1]),
            gz - 0.5 * (gz[0, 0, 0] + gz[-1, -1, -1]))


def hedgehog_field(rx, ry, rz, bfloat, rotation, p_f):
    """hedgehog_field (initial_config.py:701-758): f = pi exp(-(r / r0)^2), r0 = max(size) / (8 max(|B|, 1)),
    isorotation about z; (-1, 0, 0, 0) at r < 1e-12."""
    r0 = max(max(p_f[0], p_f[1], p_f[2]) / (8.0 * max(abs(bfloat), 1.0)), 1.0e-6)
    return _hedgehog(rx, ry, rz, ("gauss", r0), euler_matrix(rotation, 0.0, 0.0))


def _hedgehog(rx, ry, rz, profile, iso):
    r = np.sqrt(rx * rx + ry * ry + rz * rz)
    at0 = r < 1.0e-12
    f = radial_profile(r, profile, 1)
    sf = np.sin(f)
    invr = 1.0 / np.where(at0, 1.0, r)
    px, py, pz = apply_matrix(iso, rx * invr, ry * invr, rz * invr)
    out = np.array([np.cos(f), sf * px, sf * py, sf * pz])
    out[:, at0] = np.array([-1.0, 0.0, 0.0, 0.0])[:, None]
    return out


def radial_profile(r, profile, bint):
    """('fit', m0[, m1]): the reference's profilefun; ('gauss', r0): pi exp(-(r/r0)^2) of its hedgehog;
    ('exp', a): pi exp(-r/a), the profile the benchmark configurations name."""
    kind = profile[0]
    if kind == "fit":
        return profilefun(r, bint, float(profile[1]), float(profile[2]) if len(profile) > 2 else 0.0)
    if kind == "gauss":
        return PI * np.exp(-(r / float(profile[1])) * (r / float(profile[1])))
    if kind == "exp":
        return PI * np.exp(-r / float(profile[1]))
    raise ValueError(f"unknown profile {profile!r}")


def initial_configuration(p_i, p_f, ansatz: str, baryon_number: float, isospin_rotation: float):
    """The Skyrme components written by create_initial_configuration_kernel (initial_config.py:880-975):
    'rational_map' (profilefun(r, B, 1, 0), isorotation rotate_xyz(n, 0, 0, iso)), 'hedgehog', or the
    vacuum; then renormalised."""
    rx, ry, rz = centered_position(p_i, p_f)
    bint = max(int(abs(baryon_number) + 0.5), 1)
    if ansatz == "rational_map":
        z0, z1, r = stereographic(rx, ry, rz)
        f = profilefun(r, bint, 1.0, 0.0)
        F = np.array(calc_rational_map_skyrmion(f, bint, z0, z1, 0.0, 0.0, isospin_rotation))
    elif ansatz == "hedgehog":
        F = hedgehog_field(rx, ry, rz, baryon_number, isospin_rotation, p_f)
    else:
        F = np.array([np.ones_like(rx), np.zeros_like(rx), np.zeros_like(rx), np.zeros_like(rx)])
    return normalise(F)


def rational_map_skyrmion(p_i, p_f, centre, alpha, beta, baryon_number, m0, m1=0.0):
    """create_rational_map_skyrmion_kernel (initial_config.py:978-1078): a rational-map Skyrmion centred
    at the physical point `centre`, space rotation rotate_xyz(r, *alpha), isorotation
    rotate_xyz(n, *beta), profile profilefun(r, B, m0, m1); renormalised."""
    gx, gy, gz = grid(p_i, p_f)
    rx, ry, rz = rotate_xyz(gx - centre[0], gy - centre[1], gz - centre[2], *alpha)
    bint = max(int(abs(baryon_number) + 0.5), 1)
    z0, z1, r = stereographic(rx, ry, rz)
    f = profilefun(r, bint, m0, m1)
    return normalise(np.array(calc_rational_map_skyrmion(f, bint, z0, z1, *beta)))


def product_ansatz(F, T):
    """product_ansatz_kernel (initial_config.py:1081-1126): F <- normalise(F T) sitewise."""
    return normalise(quaternion_prod(F, T))


def lump(p_i, p_f, kind: str, baryon_number: float, profile, centre=(0.0, 0.0, 0.0), space=None, iso=None):
    """One lump about `centre` (relative to the box centre): 'hedgehog' (radial n, no space rotation) or
    'rational_map' (n from R(z) of the space-rotated position), n isorotated by `iso`, profile from
    :func:`radial_profile`; renormalised.  With the reference's Euler angles, space = euler_matrix(*alpha)
    and iso = euler_matrix(*beta) give create_rational_map_skyrmion_kernel."""
    rx, ry, rz = centered_position(p_i, p_f)
    rx, ry, rz = rx - centre[0], ry - centre[1], rz - centre[2]
    I = np.eye(3)
    if kind == "hedgehog":
        return normalise(_hedgehog(rx, ry, rz, profile, I if iso is None else iso))
    rx, ry, rz = apply_matrix(I if space is None else space, rx, ry, rz)
    bint = max(int(abs(baryon_number) + 0.5), 1)
    z0, z1, r = stereographic(rx, ry, rz)
    f = radial_profile(r, profile, bint)
    r0, r1 = rational_map_value(z0, z1, bint)
    nx, ny, nz = apply_matrix(I if iso is None else iso, *zton(r0, r1))
    sf = np.sin(f)
    return normalise(np.array([np.cos(f), nx * sf, ny * sf, nz * sf]))


def vacuum_band(F, halo: int = 2):
    """(1, 0, 0, ...) on the outermost `halo` sites of every axis (the vacuum boundary band)."""
    band = np.zeros(F.shape[1:], dtype=bool)
    band[:halo] = band[-halo:] = True
    band[:, :halo] = band[:, -halo:] = True
    band[:, :, :halo] = band[:, :, -halo:] = True
    F[:, band] = 0.0
    F[0, band] = 1.0
    return F


def create_omega_field(F, p_i, p_f, c_omega: float):
    """create_omega_field_kernel (theories/omega_skyrme/initial_config.py:1495-1532): component 4 of `F`
    (nf, X, Y, Z) set to -c_omega times the Chern-Simons current of components 0..3."""
    from oracle import stencils, theories

    d1 = stencils.first_derivatives(F, (p_f[3], p_f[4], p_f[5]), int(p_i[3]), comps=range(4))
    F[4] = -c_omega * theories.chern_simons(F, d1)
    return F
