"""Host-side (NumPy, FP64) initial data for the engine.

Restates the reference's initial-configuration device code on whole grids:

* grid coordinates ``ls * index`` and the box centre
  (``theories/nuclear_skyrme/kernels.py:27-56``, ``initial_config.py:664-698``);
* the hedgehog ansatz (``initial_config.py:701-758``);
* the rational-map ansatz with the fitted profile ``profilefun`` and the maps
  ``R(z)`` for B = 1..9 (``initial_config.py:342-587``);
* the quaternion product ansatz (``initial_config.py:590-627``, ``:1081-1126``);

plus the controlled variants used by the benchmark workloads (exponential
profiles, full SU(2) isorotations, centre jitter, seeded noise, a vacuum
boundary band).  Arrays are ``(nf, X, Y, Z)``: the reference layout.
"""

from __future__ import annotations

import math

import numpy as np

# ---------------------------------------------------------------- geometry


def coordinates(rp):
    """Centred physical coordinates ``(rx, ry, rz)`` of every site.

    The reference places site ``i`` at ``ls * i`` and the centre halfway
    between the first and last site (``initial_config.py:690-698``).
    """
    X, Y, Z = rp.xlen, rp.ylen, rp.zlen
    gx = rp.lsx * np.arange(X, dtype=np.float64)
    gy = rp.lsy * np.arange(Y, dtype=np.float64)
    gz = rp.lsz * np.arange(Z, dtype=np.float64)
    cx, cy, cz = 0.5 * (gx[0] + gx[-1]), 0.5 * (gy[0] + gy[-1]), 0.5 * (gz[0] + gz[-1])
    rx, ry, rz = np.meshgrid(gx - cx, gy - cy, gz - cz, indexing="ij")
    return rx, ry, rz


def rotate_by_z(x0, x1, x2, alpha):
    """``rotate_by_z`` (``initial_config.py:51-78``)."""
    ca, sa = math.cos(alpha), math.sin(alpha)
    return ca * x0 + sa * x1, -sa * x0 + ca * x1, x2


def rotate_by_y(x0, x1, x2, alpha):
    """``rotate_by_y`` (``initial_config.py:21-48``)."""
    ca, sa = math.cos(alpha), math.sin(alpha)
    return ca * x0 - sa * x2, x1, sa * x0 + ca * x2


def rotate_xyz(x0, x1, x2, a0, a1, a2):
    """z-y-z Euler rotation (``initial_config.py:81-112``)."""
    y = rotate_by_z(x0, x1, x2, a0)
    z = rotate_by_y(*y, a1)
    return rotate_by_z(*z, a2)


# ---------------------------------------------------------------- profiles and maps
_PROFILE_FITS = {  # massive-pion fits of profilefun (initial_config.py:421-438)
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
_PROFILE_FITS_MASSLESS = {  # initial_config.py:380-416
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


def _ffun(r, c):
    a, b, cc, d, e = c
    poly = r * (a + r * (b + r * (cc + r * (d + r * e))))
    return np.pi * (1.0 - np.tanh(np.abs(poly)))


def profilefun(r, bint: int, m0: float = 1.0):
    """Fitted radial profile (``initial_config.py:343-442``)."""
    r = np.asarray(r, dtype=np.float64)
    if m0 < 1.0:
        # beyond B = 9 the reference leaves the massless term at 0 (and the massive coefficients at 0)
        f = _ffun(r, _PROFILE_FITS_MASSLESS[bint]) if bint in _PROFILE_FITS_MASSLESS else np.zeros_like(r)
        fm = _ffun(r, _PROFILE_FITS.get(bint, (0, 0, 0, 0, 0)))
        out = (1.0 - m0) * f + m0 * fm
    else:
        out = _ffun(r, _PROFILE_FITS.get(bint, (0, 0, 0, 0, 0)))
    out = np.where(r > 15.0, 0.0, out)
    return np.where(r < 1.0e-6, np.pi, out)


def rational_map(z, bint: int):
    """``R(z)`` for B = 1..8, ``z^B`` otherwise (``initial_config.py:446-549``).

    The reference adds 1e-9 to some denominators and 1e-18 in the complex
    division; both are kept.
    """
    eps = 1.0e-9

    def cdiv(num, den):
        # cdiv adds 1e-18 to |den|^2 (initial_config.py:249-250)
        d = den.real * den.real + den.imag * den.imag + 1.0e-18
        return ((num.real * den.real + num.imag * den.imag) + 1j * (num.imag * den.real - num.real * den.imag)) / d

    s3 = math.sqrt(3.0)
    if bint == 1:
        return z
    if bint == 2:
        return z * z
    if bint == 3:
        z2 = z * z
        return cdiv(1j * s3 * z2 - 1.0, z * (z2 - 1j * s3) + eps)
    if bint == 4:
        z4 = z ** 4
        t = 2.0j * s3 * z * z
        return cdiv(z4 + t + 1.0, z4 - t + 1.0)
    if bint == 5:
        a, b = 3.0660, 3.9334
        z4, z2 = z ** 4, z * z
        return cdiv(z * (z4 + b * z2 + a), a * z4 - b * z2 + 1.0)
    if bint == 6:
        a = 0.1585
        z4 = z ** 4
        return cdiv(z4 + 1j * a, z * z * (1j * a * z4 + 1.0) + eps)
    if bint == 7:
        z7, z5, z2 = z ** 7, z ** 5, z * z
        return cdiv(z7 - 7.0 * z5 - 7.0 * z2 - 1.0, z7 + 7.0 * z5 - 7.0 * z2 + 1.0)
    if bint == 8:
        a = 0.1352
        z6 = z ** 6
        return cdiv(z6 - a, z * z * (a * z6 + 1.0) + eps)
    return z ** bint


def stereographic(rx, ry, rz):
    """Riemann-sphere coordinate ``z = tan(theta/2) e^{i phi}`` of a direction
    (``initial_config.py:941-951``; the origin maps to 0)."""
    r = np.sqrt(rx * rx + ry * ry + rz * rz)
    safe = np.where(r < 1.0e-6, 1.0, r)
    theta = np.where(r < 1.0e-6, 0.0, np.arccos(np.clip(rz / safe, -1.0, 1.0)))
    phi = np.where(r < 1.0e-6, 0.0, np.arctan2(ry, rx))
    t = np.tan(0.5 * theta)
    return np.cos(phi) * t + 1j * np.sin(phi) * t, r


def zton(R):
    """Unit 3-vector of a Riemann-sphere point (``initial_config.py:284-306``)."""
    zsq = R.real * R.real + R.imag * R.imag
    den = 1.0 + zsq
    return 2.0 * R.real / den, 2.0 * R.imag / den, (1.0 - zsq) / den


# ---------------------------------------------------------------- SU(2) helpers


def quaternion_product(a, b):
    """Sitewise product of two unit-quaternion fields (``initial_config.py:591-627``)."""
    a0, a1, a2, a3 = a
    b0, b1, b2, b3 = b
    return np.array([
        a0 * b0 - a1 * b1 - a2 * b2 - a3 * b3,
        a0 * b1 + b0 * a1 + a2 * b3 - a3 * b2,
        a0 * b2 + b0 * a2 + a3 * b1 - a1 * b3,
        a0 * b3 + b0 * a3 + a1 * b2 - a2 * b1,
    ])


def haar_rotation(rng: np.random.Generator) -> np.ndarray:
    """A Haar-random SO(3) matrix (from a uniformly random unit quaternion)."""
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
    ])


def normalise_sigma(F, nsk: int = 4):
    """Renormalise the first ``nsk`` components onto S^3 (skipping zero norms)."""
    s = np.sqrt(np.sum(F[:nsk] * F[:nsk], axis=0))
    nz = s != 0.0
    F[:nsk] = np.where(nz, F[:nsk] / np.where(nz, s, 1.0), F[:nsk])
    return F


# ---------------------------------------------------------------- ansatz builders


def _profile(r, profile, bint, rp):
    kind = profile[0] if isinstance(profile, (tuple, list)) else profile
    if kind == "exp":  # f = pi exp(-r / a)
        return np.pi * np.exp(-r / float(profile[1]))
    if kind == "gauss":  # reference hedgehog: f = pi exp(-(r/r0)^2)
        return np.pi * np.exp(-(r / float(profile[1])) ** 2)
    if kind == "fit":  # reference rational-map fits, ("fit", m0) or "fit" (m0 = 1)
        return profilefun(r, bint, float(profile[1]) if isinstance(profile, (tuple, list)) else 1.0)
    raise ValueError(f"unknown profile {profile!r}")


def hedgehog(rp, baryon_number=1.0, isorotation=0.0, profile=None, center=(0.0, 0.0, 0.0), rotation=None):
    """Hedgehog ``(cos f, sin f * R n)`` about ``center``.

    With the defaults this is the reference hedgehog
    (``initial_config.py:732-758``: ``f = pi exp(-(r/r0)^2)``,
    ``r0 = max(size) / (8 max(|B|, 1))``, isorotation about z).  ``rotation``
    (a 3x3 matrix) replaces the z-isorotation by a general one.
    """
    rx, ry, rz = coordinates(rp)
    rx, ry, rz = rx - center[0], ry - center[1], rz - center[2]
    r = np.sqrt(rx * rx + ry * ry + rz * rz)
    if profile is None:
        rmax = max(rp.xsize, rp.ysize, rp.zsize)
        r0 = max(rmax / (8.0 * max(abs(baryon_number), 1.0)), 1.0e-6)
        profile = ("gauss", r0)
    f = _profile(r, profile, 1, rp)
    safe = np.where(r < 1.0e-12, 1.0, r)
    nx, ny, nz = rx / safe, ry / safe, rz / safe
    if rotation is not None:
        R = np.asarray(rotation)
        nx, ny, nz = (R[0, 0] * nx + R[0, 1] * ny + R[0, 2] * nz,
                      R[1, 0] * nx + R[1, 1] * ny + R[1, 2] * nz,
                      R[2, 0] * nx + R[2, 1] * ny + R[2, 2] * nz)
    else:
        nx, ny, nz = rotate_by_z(nx, ny, nz, isorotation)
    sf = np.sin(f)
    out = np.array([np.cos(f), sf * nx, sf * ny, sf * nz])
    at0 = r < 1.0e-12
    out[:, at0] = np.array([-1.0, 0.0, 0.0, 0.0])[:, None]
    return out


def rational_map_skyrmion(rp, baryon_number=1.0, isorotation=0.0, profile=None, center=(0.0, 0.0, 0.0),
                          euler=(0.0, 0.0, 0.0), rotation=None, space_rotation=None):
    """Rational-map Skyrmion ``(cos f, sin f * n_R)`` about ``center``.

    With the defaults: the reference's ``create_initial_configuration_kernel``
    rational-map branch (``initial_config.py:941-953``): fitted profile
    ``profilefun(r, B, 1, 0)``, ``R(z)`` for ``B = int(|B| + 0.5)`` clamped to
    >= 1, isorotation ``rotate_xyz(n, 0, 0, isorotation)``.
    """
    bint = max(int(abs(baryon_number) + 0.5), 1)
    rx, ry, rz = coordinates(rp)
    rx, ry, rz = rx - center[0], ry - center[1], rz - center[2]
    if space_rotation is not None:
        S = np.asarray(space_rotation)
        rx, ry, rz = (S[0, 0] * rx + S[0, 1] * ry + S[0, 2] * rz,
                      S[1, 0] * rx + S[1, 1] * ry + S[1, 2] * rz,
                      S[2, 0] * rx + S[2, 1] * ry + S[2, 2] * rz)
    zc, r = stereographic(rx, ry, rz)
    f = profilefun(r, bint, 1.0) if profile is None else _profile(r, profile, bint, rp)
    nx, ny, nz = zton(rational_map(zc, bint))
    if rotation is not None:
        R = np.asarray(rotation)
        nx, ny, nz = (R[0, 0] * nx + R[0, 1] * ny + R[0, 2] * nz,
                      R[1, 0] * nx + R[1, 1] * ny + R[1, 2] * nz,
                      R[2, 0] * nx + R[2, 1] * ny + R[2, 2] * nz)
    else:
        nx, ny, nz = rotate_xyz(nx, ny, nz, euler[0], euler[1], euler[2] + isorotation)
    sf = np.sin(f)
    return np.array([np.cos(f), sf * nx, sf * ny, sf * nz])


def vacuum(rp, nf):
    F = np.zeros((nf, rp.xlen, rp.ylen, rp.zlen))
    F[0] = 1.0
    return F


def set_vacuum_band(F, width=2):
    """Write the vacuum (1, 0, 0, 0, 0...) into the outermost ``width`` sites of every axis."""
    band = np.zeros(F.shape[1:], dtype=bool)
    band[:width] = band[-width:] = True
    band[:, :width] = band[:, -width:] = True
    band[:, :, :width] = band[:, :, -width:] = True
    F[:, band] = 0.0
    F[0, band] = 1.0
    return F


def first_derivative(f: np.ndarray, axis: int, h: float, halo: int = 2) -> np.ndarray:
    """4th-order central first derivative of the 3D array `f` along `axis`, zero within `halo` sites of
    either end (the stencil and band of ``compute_derivative_first``, core/derivatives.py:17-67)."""
    n = f.shape[axis]
    out = np.zeros_like(f)
    lo, hi = halo, n - halo
    if hi <= lo:
        return out

    def sl(a, b):
        idx = [slice(None)] * 3
        idx[axis] = slice(a, b)
        return tuple(idx)

    out[sl(lo, hi)] = (f[sl(lo - 2, hi - 2)] - 8.0 * f[sl(lo - 1, hi - 1)] + 8.0 * f[sl(lo + 1, hi + 1)]
                       - f[sl(lo + 2, hi + 2)]) / (12.0 * h)
    return out


def chern_simons_density(F: np.ndarray, rp) -> np.ndarray:
    """CS = -det[phi, d_x phi, d_y phi, d_z phi] / (2 pi^2) of the 4 pion components of `F` (nf, X, Y, Z),
    the baryon density without the cell volume (``_chern_simons_current_point``,
    theories/omega_skyrme/kernels.py:443-500)."""
    halo = int(getattr(rp, "halo", 2))
    phi = F[:4]
    d = [np.stack([first_derivative(phi[a], ax, h, halo) for a in range(4)])
         for ax, h in ((0, rp.lsx), (1, rp.lsy), (2, rp.lsz))]
    M = np.stack([phi, d[0], d[1], d[2]])  # (column, component, X, Y, Z)
    det = np.linalg.det(np.moveaxis(M, (0, 1), (-1, -2)))  # rows = components, columns = (phi, dx, dy, dz)
    return -det / (2.0 * np.pi * np.pi)


def _mode(rp, cfg) -> str:
    mode = str(cfg.get("mode", "initial")).lower().strip()
    if mode == "initial":
        if getattr(rp, "ansatz_rational_map", False):
            mode = "rational_map"
        elif getattr(rp, "ansatz_uniform", False):
            mode = "ground"
        else:
            mode = "hedgehog"
    if mode in ("rational map", "rational-map"):
        mode = "rational_map"
    return mode


def lump_specs(rp, config: dict | None = None):
    """The lumps an ``initialize(config)`` request builds, with every random draw made.

    Returns ``(lumps, extra)``: ``lumps`` is a list of dicts (``mode``,
    ``B``, ``iso``, ``profile``, ``center``, ``rotation``, ``space_rotation``)
    whose quaternion product is the field; ``extra`` holds ``noise``,
    ``seed`` and ``vacuum_band``.  :func:`build_field` evaluates them on the
    host and :func:`device_init_desc` hands them to the device."""
    cfg = dict(config or {})
    rng = np.random.default_rng(cfg.get("seed", 0))
    mode = _mode(rp, cfg)
    if mode in ("ground", "uniform", "vacuum", "gs"):
        lumps = []
    elif mode in ("hedgehog", "rational_map"):
        lumps = [_lump_spec(rp, mode, cfg, rng)]
    elif mode == "product":
        lumps = [_lump_spec(rp, _mode(rp, dict(lump, mode=lump.get("mode", "rational_map"))), lump, rng)
                 for lump in cfg.get("lumps", [])]
    elif mode in ("smorgasbord", "smorgaasbord", "smorg"):
        lumps = _smorgasbord_specs(rp, cfg, rng)
    else:
        raise ValueError(f"unknown initial-data mode {mode!r}")
    extra = dict(noise=float(cfg.get("noise", 0.0)), seed=int(cfg.get("seed", 0)),
                 vacuum_band=bool(cfg.get("vacuum_band", False)), rng=rng)
    return lumps, extra


def build_field(rp, nf: int, config: dict | None = None) -> np.ndarray:
    """Initial field ``(nf, X, Y, Z)`` for an ``initialize(config)`` request, built on the host.

    ``mode``: ``ground`` / ``uniform`` / ``vacuum``; ``hedgehog``;
    ``rational_map``; ``initial`` (the ansatz named by the parameters);
    ``product`` (quaternion product of ``lumps``, each a dict of ansatz
    options with its own ``mode``); ``smorgasbord`` (the reference's
    randomly placed rational-map lumps: ``count``, ``lump_baryon``,
    ``profile_m0``, ``position_scale``, ``seed``).  Options: ``baryon_number``,
    ``isospin_rotation``, ``profile`` (``("exp", a)``, ``("gauss", r0)`` or
    ``"fit"``), ``center``, ``jitter`` (uniform centre offset), ``isorotation``
    (``"haar"`` for a random SU(2) isorotation), ``space_rotation``
    (``"haar"``), ``noise`` (N(0, noise) per component before
    renormalisation), ``seed``, ``vacuum_band`` (bool).
    """
    lumps, extra = lump_specs(rp, config)
    F = vacuum(rp, nf)
    acc = None
    for spec in lumps:
        part = _eval_lump(rp, spec)
        acc = part if acc is None else normalise_sigma(quaternion_product(acc, part))
    if acc is not None:
        F[:4] = acc
    if extra["noise"] > 0.0:
        F[:4] += extra["rng"].normal(0.0, extra["noise"], size=F[:4].shape)
    normalise_sigma(F)
    if extra["vacuum_band"]:
        set_vacuum_band(F, int(getattr(rp, "halo", 2)))
    return F


def _smorgasbord_specs(rp, cfg, rng):
    """``_initialize_smorgasbord`` (``theories/nuclear_skyrme/initial_config.py:1267-1407``).

    ``count`` rational-map lumps of ``lump_baryon`` each (count defaults to
    round(|B|)), centres uniform in +-``position_scale`` x the box size about
    the box centre and then shifted to zero mean, space and isospin rotations
    ``rotate_xyz`` by three uniform angles each, fitted profile with
    ``profile_m0``; drawn from ``default_rng(seed)`` in the reference's order
    (positions, space angles, isospin angles), so a seed gives the
    reference's lumps.
    """
    B = float(cfg.get("baryon_number", getattr(rp, "baryon_number", 1.0)))
    count = int(cfg.get("count", max(1, int(round(abs(B))))))
    lump_baryon = int(cfg.get("lump_baryon", 1))
    m0 = float(cfg.get("profile_m0", 1.0))
    scale = float(cfg.get("position_scale", 0.18))
    d = scale * np.array([rp.xsize, rp.ysize, rp.zsize], dtype=np.float64)
    pos = rng.uniform(low=-d, high=d, size=(count, 3))
    pos -= pos.mean(axis=0, keepdims=True)
    alpha = rng.uniform(0.0, 2.0 * np.pi, size=(count, 3))
    beta = rng.uniform(0.0, 2.0 * np.pi, size=(count, 3))
    return [dict(mode="rational_map", B=float(lump_baryon), iso=0.0, profile=("fit", m0), center=pos[n],
                 rotation=_apply(rotate_xyz, *beta[n]), space_rotation=_apply(rotate_xyz, *alpha[n]))
            for n in range(count)]


def _lump_spec(rp, mode, cfg, rng):
    B = float(cfg.get("baryon_number", getattr(rp, "baryon_number", 1.0)))
    iso = float(cfg.get("isospin_rotation", getattr(rp, "isospin_rotation", 0.0)))
    center = np.array(cfg.get("center", (0.0, 0.0, 0.0)), dtype=np.float64)
    jitter = float(cfg.get("jitter", 0.0))
    if jitter > 0.0:
        center = center + rng.uniform(-jitter, jitter, size=3)
    rot = haar_rotation(rng) if cfg.get("isorotation") == "haar" else None
    srot = haar_rotation(rng) if cfg.get("space_rotation") == "haar" else None
    return dict(mode=mode, B=B, iso=iso, profile=cfg.get("profile"), center=center, rotation=rot,
                space_rotation=srot)


def _eval_lump(rp, spec):
    if spec["mode"] == "hedgehog":
        return hedgehog(rp, spec["B"], spec["iso"], spec["profile"], spec["center"], rotation=spec["rotation"])
    return rational_map_skyrmion(rp, spec["B"], spec["iso"], spec["profile"], spec["center"],
                                 rotation=spec["rotation"], space_rotation=spec["space_rotation"])


def _apply(fn, *args):
    """3x3 matrix of a linear map given as a function of three coordinates."""
    M = np.zeros((3, 3))
    for j in range(3):
        e = [0.0, 0.0, 0.0]
        e[j] = 1.0
        M[:, j] = fn(*e, *args)
    return M


def device_init_desc(rp, config: dict | None = None, omega_cs: float = 0.0):
    """The ``skx_init_desc`` of an ``initialize(config)`` request (the same lumps as :func:`build_field`).

    The device adds noise from its own counter-based generator keyed by the
    config's ``seed``, not NumPy's stream: with ``noise > 0`` the device and
    host fields differ by the noise realisation only."""
    from skyrmion_solver_amd import _engine

    lumps, extra = lump_specs(rp, config)
    if len(lumps) > _engine.MAX_LUMPS:
        raise ValueError(f"at most {_engine.MAX_LUMPS} lumps")
    d = _engine.InitDesc()
    d.n_lumps = len(lumps)
    d.vacuum_band = int(extra["vacuum_band"])
    d.noise = extra["noise"]
    d.noise_seed = extra["seed"] & ((1 << 64) - 1)
    d.omega_cs = float(omega_cs)
    for k, sp in enumerate(lumps):
        q = d.lumps[k]
        q.kind = _engine.LUMP_HEDGEHOG if sp["mode"] == "hedgehog" else _engine.LUMP_RATIONAL_MAP
        q.baryon_number = max(int(abs(sp["B"]) + 0.5), 1)
        prof = sp["profile"]
        if prof is None and sp["mode"] == "hedgehog":
            rmax = max(rp.xsize, rp.ysize, rp.zsize)
            prof = ("gauss", max(rmax / (8.0 * max(abs(sp["B"]), 1.0)), 1.0e-6))
        kind = prof[0] if isinstance(prof, (tuple, list)) else prof
        if prof is None or kind == "fit":
            q.profile = _engine.PROFILE_FIT
            q.profile_param = float(prof[1]) if isinstance(prof, (tuple, list)) else 1.0
        elif kind == "exp":
            q.profile, q.profile_param = _engine.PROFILE_EXP, float(prof[1])
        elif kind == "gauss":
            q.profile, q.profile_param = _engine.PROFILE_GAUSS, float(prof[1])
        else:
            raise ValueError(f"unknown profile {prof!r}")
        for i in range(3):
            q.center[i] = float(sp["center"][i])
        S = np.eye(3) if sp["space_rotation"] is None or sp["mode"] == "hedgehog" else np.asarray(sp["space_rotation"])
        if sp["rotation"] is not None:
            R = np.asarray(sp["rotation"])
        elif sp["mode"] == "hedgehog":
            R = _apply(rotate_by_z, sp["iso"])
        else:
            R = _apply(rotate_xyz, 0.0, 0.0, sp["iso"])
        for i in range(9):
            q.space_rot[i] = float(S.flat[i])
            q.iso_rot[i] = float(R.flat[i])
    return d


