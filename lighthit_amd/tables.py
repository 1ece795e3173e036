"""Order >=2 RTE tables: the directional radial table and the kernel pack.

One-time preparation per (medium, acceptance, configuration), on the host:
the per-event route only reads the pack. Restates, in this package's own
code:

* ``DirectionalCache.build`` (``directional.py:360-522``): the directional
  transfer coefficients ``t_(l, lambda, |mu|)(r, omega)`` of the scattered
  orders on a geometric radius grid, as a Bessel transform over wave number
  ``k`` of the angular resolvent rows;
* ``resolvent_rows`` (``angular.py:427-475``): rows ``lambda <= L_A`` of the
  single- and multiple-scattering resolvents of the truncated-scattering
  radiative transfer equation, per azimuthal block ``m``, with the exact free
  angular tail above the scattering degree;
* ``CouplingTable.build`` (``directional.py:161-190``): the Clebsch-Gordan
  weights that fold the ``(lambda, l, J)`` triples into ``(lambda, |mu|)``
  pairs;
* ``PreparedDirectionalKernel.from_cache`` (``_directional_numba.py:
  419-445``) and ``Ge2KernelStore.bundle`` (``shower_basis.py:214-286``): the
  order >=2 part, detrended by ``e^{-mu_a r} / (4 pi r^2)`` and the flight
  phase, as a not-a-knot cubic spline in ``log r``, stacked over the
  wavelength nodes into the pack layout
  ``(wavelength, interval, degree + 1, pair, power, frequency)``.

Method (the angular part). In the frame where the wave vector is the polar
axis, the transfer equation in normalised harmonics ``h_l`` of azimuthal
order ``m`` is tridiagonal:
``(d0 - gamma_l) h_l + i k (a_(l,m) h_(l-1) + a_(l+1,m) h_(l+1)) = rhs_l`` with
``d0 = mu_e - i omega / v``, ``gamma_l = mu_s g^l`` (zero above the scattering
degree ``L``) and ``a_(l,m) = sqrt((l^2 - m^2) / (4 l^2 - 1))``. Above ``L``
the equation is free, and its solution decaying in ``l`` (the minimal one)
closes the system at ``L`` exactly: ``h_(L+1) = R_L h_L``. The ratios ``R_l``
of the minimal solution come here from the downward continued fraction
``R_(l-1) = a_(l,m) / (z - a_(l+1,m) R_l)``, ``z = i d0 / k``, started deep
enough that the start's error has decayed below double precision (the
reference takes the m = 0 ratios from a Miller/forward scheme and raises them
in m; the continued fraction is the same minimal solution).
"""
from __future__ import annotations

from fractions import Fraction
from math import factorial, sqrt

import numpy as np

C_VACUUM_M_PER_NS = 0.299792458

#: Moment orders of a directional table: single scattering, two or more.
ORDERS = ("first", "multiple")


# ------------------------------------------------------------- k quadrature

def k_quadrature(k_max_per_m, k_panel_per_m, order, taper=True):
    """Wave numbers and weights on ``[0, k_max]``: Gauss-Legendre panels of
    ``order`` nodes and width at most ``k_panel``; with ``taper`` the weights
    roll off over the upper half by the smooth polynomial
    ``(1 - a)^5 (1 + 5a + 15a^2 + 35a^3 + 70a^4)``, ``a = max(0, 2k/k_max - 1)``
    (``green.py:44-54``)."""
    panels = int(np.ceil(k_max_per_m / k_panel_per_m))
    edges = np.linspace(0.0, k_max_per_m, panels + 1)
    x, w = np.polynomial.legendre.leggauss(int(order))
    half = 0.5 * np.diff(edges)
    k = (edges[:-1, None] + half[:, None] * (x + 1.0)).ravel()
    weight = (half[:, None] * w[None, :]).ravel()
    if taper:
        a = np.clip(2.0 * k / k_max_per_m - 1.0, 0.0, 1.0)
        weight = weight * (1 - a) ** 5 * (1 + a * (5 + a * (15 + a * (35 + 70 * a))))
    return k, weight


# ------------------------------------------------------------ coupling table

def clebsch_gordan(j1, m1, j2, m2, j, m):
    """``<j1 m1; j2 m2 | j m>`` for integer angular momenta, exact until the
    final square root (Racah's closed form; the squared prefactor and the sum
    are rationals)."""
    if m1 + m2 != m or not abs(j1 - j2) <= j <= j1 + j2:
        return 0.0
    if abs(m1) > j1 or abs(m2) > j2 or abs(m) > j:
        return 0.0
    f = factorial
    prefactor = Fraction((2 * j + 1) * f(j + j1 - j2) * f(j - j1 + j2) * f(j1 + j2 - j)
                         * f(j1 + m1) * f(j1 - m1) * f(j2 + m2) * f(j2 - m2)
                         * f(j + m) * f(j - m), f(j1 + j2 + j + 1))
    total = Fraction(0)
    lo = max(0, j2 - j - m1, j1 + m2 - j)
    hi = min(j1 + j2 - j, j1 - m1, j2 + m2)
    for s in range(lo, hi + 1):
        total += Fraction((-1) ** s, f(s) * f(j1 + j2 - j - s) * f(j1 - m1 - s)
                          * f(j2 + m2 - s) * f(j - j2 + m1 + s) * f(j - j1 - m2 + s))
    if total == 0:
        return 0.0
    magnitude = sqrt(float(prefactor * total * total))
    return magnitude if total > 0 else -magnitude


class Coupling:
    """The ``(lambda, l, J)`` triples with ``l + lambda + J`` even, the
    ``(lambda, |mu|)`` pairs, ``nu_weight[t, nu] = (2 - delta_nu0) (-1)^nu
    <lambda nu; l -nu | J 0>`` and ``pair_weight[t, p] = i^J (-1)^mu
    <lambda mu; l -mu | J 0>`` (``directional.py:127-190``)."""

    def __init__(self, source_degree, acceptance_degree):
        self.source_degree = int(source_degree)
        self.acceptance_degree = int(acceptance_degree)
        pairs = [(lam, mu) for lam in range(self.acceptance_degree + 1)
                 for mu in range(lam + 1)]
        self.pair_lambda = np.array([p[0] for p in pairs], np.int64)
        self.pair_mu = np.array([p[1] for p in pairs], np.int64)
        index = {p: i for i, p in enumerate(pairs)}
        triples = [(lam, l, J) for lam in range(self.acceptance_degree + 1)
                   for l in range(self.source_degree + 1)
                   for J in range(abs(l - lam), l + lam + 1) if (l + lam + J) % 2 == 0]
        self.triple_lambda = np.array([t[0] for t in triples], np.int64)
        self.triple_l = np.array([t[1] for t in triples], np.int64)
        self.triple_J = np.array([t[2] for t in triples], np.int64)
        self.nu_weight = np.zeros((len(triples), self.acceptance_degree + 1))
        self.pair_weight = np.zeros((len(triples), len(pairs)), complex)
        for t, (lam, l, J) in enumerate(triples):
            for nu in range(min(lam, l) + 1):
                c = (-1.0) ** nu * clebsch_gordan(lam, nu, l, -nu, J, 0)
                self.nu_weight[t, nu] = c if nu == 0 else 2.0 * c
                self.pair_weight[t, index[(lam, nu)]] = (1j ** J) * c

    @property
    def pairs(self):
        return len(self.pair_lambda)


# ------------------------------------------------------------- angular rows

def ladder(m, top):
    """``a_(l,m)`` for ``l = 0 .. top`` (zero below ``l = max(m, 1)``)."""
    ell = np.arange(top + 1, dtype=float)
    out = np.zeros(top + 1)
    use = ell >= max(m, 1)
    out[use] = np.sqrt((ell[use] ** 2 - m * m) / (4.0 * ell[use] ** 2 - 1.0))
    return out


def minimal_ratios(k, d0, degree, m):
    """``R_l = h_(l+1) / h_l`` of the minimal free solution of block ``m`` for
    ``l = 0 .. degree`` (zero below ``m``), by the downward continued fraction.
    ``k`` (...,) positive wave numbers, ``d0`` broadcastable complex.

    Each node starts ``degree + 1 + max(32, ceil(28 / eta))`` degrees up, ``eta``
    the decay rate of the minimal solution (``|rho| = e^-eta``): the start's
    error is damped by ``|rho|^(2 depth)``."""
    k = np.asarray(k, float)
    z = 1j * np.asarray(d0, complex) / k
    root = np.sqrt(z - 1.0) * np.sqrt(z + 1.0)
    rho = 1.0 / (z + root)
    rho = np.where(np.abs(rho) > 1.0, 1.0 / rho, rho)
    eta = np.maximum(-np.log(np.abs(rho)), 1e-300)
    depth = degree + 1 + np.maximum(32, np.ceil(28.0 / eta)).astype(np.int64)
    top = int(depth.max())
    a = ladder(m, top + 1)
    ratio = np.zeros(np.broadcast(z, k).shape, complex)
    out = np.zeros(ratio.shape + (degree + 1,), complex)
    # nodes join the recursion at their own depth (a deeper start is only more
    # accurate; the per-node start keeps the cost of the slow nodes local)
    for ell in range(top, m, -1):
        on = depth >= ell
        ratio = np.where(on, a[ell] / (z - a[ell + 1] * ratio), ratio)
        if ell - 1 <= degree:
            out[..., ell - 1] = ratio
    return out


def solve_block(k, d0, gamma, tail, rhs, m):
    """Tridiagonal solve of block ``m`` over ``l = m .. m + n - 1`` with the free
    tail folded into the last diagonal: ``(d0 - gamma_l)`` on the diagonal,
    ``i k a_(l,m)`` beside it, ``+ i k a_(m+n,m) tail`` at the end. ``rhs`` (..., n)."""
    n = rhs.shape[-1]
    a = ladder(m, m + n)[m:]                        # a_(m+i, m), i = 0..n
    off = 1j * k[..., None] * a[None, 1:n]           # (..., n-1)
    diag = np.broadcast_to(d0 - gamma, rhs.shape).astype(complex).copy()
    diag[..., -1] += 1j * k * a[n] * tail
    r = rhs.astype(complex).copy()
    for i in range(1, n):                             # elimination
        f = off[..., i - 1] / diag[..., i - 1]
        diag[..., i] -= f * off[..., i - 1]
        r[..., i] -= f * r[..., i - 1]
    x = np.empty_like(r)
    x[..., -1] = r[..., -1] / diag[..., -1]
    for i in range(n - 2, -1, -1):                    # back substitution
        x[..., i] = (r[..., i] - off[..., i] * x[..., i + 1]) / diag[..., i]
    if not np.isfinite(x).all():
        raise FloatingPointError("nonfinite angular solution")
    return x


def resolvent_rows(k, omega_per_ns, extinction_per_m, scattering_per_m, g, speed_m_per_ns,
                   scattering_degree, output_degree, acceptance_degree):
    """Rows ``lambda <= L_A`` of ``G0 Gamma G0`` (``first``) and ``G Gamma G0 Gamma
    G0`` (``multiple``): dicts keyed ``(m, lambda)`` of arrays (..., K,
    output_degree + 1) over the source degree ``l`` (``angular.py:427-475``).
    ``omega_per_ns`` (W,) gives a leading frequency axis."""
    L, Lout, A = int(scattering_degree), int(output_degree), int(acceptance_degree)
    if A > L:
        raise ValueError("acceptance_degree must not exceed scattering_degree")
    N = max(L, Lout)
    omega = np.atleast_1d(np.asarray(omega_per_ns, float))
    d0 = (extinction_per_m - 1j * omega / speed_m_per_ns)[:, None]       # (W, 1)
    kk = np.broadcast_to(np.asarray(k, float), (len(omega), len(k)))      # (W, K)
    gamma_all = scattering_per_m * g ** np.arange(L + 1)
    first, multiple = {}, {}
    for m in range(A + 1):
        ratios = minimal_ratios(kk, d0, N, m)                              # (W, K, N+1)
        tail = ratios[..., L]
        gamma = gamma_all[m:]
        zero = np.zeros_like(gamma)
        d0m = d0[..., None]                                                # (W, 1, 1)
        for lam in range(m, A + 1):
            unit = np.zeros(kk.shape + (L + 1 - m,), complex)
            unit[..., lam - m] = 1.0
            free = solve_block(kk, d0m, zero, tail, unit, m)
            one = solve_block(kk, d0m, zero, tail, free * gamma, m)
            more = solve_block(kk, d0m, gamma, tail, one * gamma, m)
            for rows, low in ((first, one), (multiple, more)):
                full = np.zeros(kk.shape + (N + 1,), complex)
                full[..., m:L + 1] = low
                if N > L:                             # free continuation above L
                    full[..., L + 1:] = low[..., -1:] * np.cumprod(ratios[..., L:N], axis=-1)
                rows[(m, lam)] = full
    return first, multiple


# --------------------------------------------------------- directional table

def directional_table(medium_band, radii_m, omega_per_ns, coupling, *, scattering_degree,
                      k_max_per_m, k_panel_per_m, k_order, taper=True):
    """``t[omega, r, l, pair, order]`` of one wavelength band
    (``directional.py:360-522``): ``t = sum_(t in triples of l) pair_weight
    (1 / 2 pi^2) int k^2 j_J(k r) g_t(k) dk`` with
    ``g_t = sum_nu nu_weight[t, nu] row_(nu, lambda_t)(k)[l_t]`` per order.

    ``medium_band``: dict with ``absorption_per_m``, ``scattering_per_m``,
    ``group_index`` and ``g``."""
    from scipy.special import spherical_jn
    radii = np.asarray(radii_m, float)
    omega = np.asarray(omega_per_ns, float)
    mu_a = float(medium_band["absorption_per_m"])
    mu_s = float(medium_band["scattering_per_m"])
    speed = C_VACUUM_M_PER_NS / float(medium_band["group_index"])
    out = np.zeros((len(omega), len(radii), coupling.source_degree + 1, coupling.pairs,
                    len(ORDERS)), complex)
    if mu_s <= 0:
        return out
    k, kw = k_quadrature(k_max_per_m, k_panel_per_m, k_order, taper)
    weight = kw * k * k / (2.0 * np.pi ** 2)
    top_J = int(coupling.triple_J.max())
    bessel = spherical_jn(np.arange(top_J + 1)[None, None, :],
                          k[None, :, None] * radii[:, None, None]) * weight[None, :, None]
    first, multiple = resolvent_rows(k, omega, mu_a + mu_s, mu_s, float(medium_band["g"]),
                                     speed, scattering_degree, coupling.source_degree,
                                     coupling.acceptance_degree)
    T = len(coupling.triple_l)
    g_t = np.zeros((len(omega), len(k), T, len(ORDERS)), complex)
    for o, rows in enumerate((first, multiple)):
        for (m, lam), block in rows.items():
            sel = np.flatnonzero(coupling.triple_lambda == lam)
            g_t[:, :, sel, o] += (coupling.nu_weight[sel, m][None, None, :]
                                  * block[:, :, coupling.triple_l[sel]])
    # moments[w, r, t, o] = sum_k bessel[r, k, J_t] g_t[w, k, t, o]
    moments = np.zeros((len(omega), len(radii), T, len(ORDERS)), complex)
    for J in range(top_J + 1):
        sel = np.flatnonzero(coupling.triple_J == J)
        if len(sel):
            moments[:, :, sel] = np.einsum("rk,wkto->wrto", bessel[:, :, J], g_t[:, :, sel],
                                           optimize=False)
    for l in range(coupling.source_degree + 1):
        sel = np.flatnonzero(coupling.triple_l == l)
        if len(sel):
            out[:, :, l] = np.einsum("wrto,tp->wrpo", moments[:, :, sel],
                                     coupling.pair_weight[sel], optimize=False)
    return out


# ------------------------------------------------------------------- pack

def radial_spline(table, radii_m, omega_per_ns, absorption_per_m, speed_m_per_ns, order=1):
    """Spline coefficients ``(interval, degree + 1, pair, power, frequency)`` of one
    order of a directional table in ``log r`` after removing ``e^{-mu_a r} /
    (4 pi r^2)`` and the flight phase ``e^{i omega r / v}`` (not-a-knot cubics,
    highest power first; ``_directional_numba.py:419-445``)."""
    from scipy.interpolate import CubicSpline
    radii = np.asarray(radii_m, float)
    omega = np.asarray(omega_per_ns, float)
    scale = np.exp(-absorption_per_m * radii) / (4.0 * np.pi * radii ** 2)
    if np.any(scale == 0) or not np.isfinite(scale).all():
        raise ValueError("radial scale underflows on this grid")
    values = table[..., order] / scale[None, :, None, None]
    values = values * np.exp(-1j * omega[:, None] * radii[None, :]
                             / speed_m_per_ns)[:, :, None, None]
    spline = CubicSpline(np.log(radii), values, axis=1)
    # spline.c: (power, interval, frequency, degree, pair)
    return np.ascontiguousarray(np.transpose(spline.c, (1, 3, 4, 0, 2)))


def build_pack(kernel, *, frequency_indices=None, dtype=np.complex64, progress=None):
    """The order >=2 pack of ``kernel`` for :meth:`lighthit_amd.fold.KernelPack.add`:
    ``(wavelengths, indices, bundle)``. ``kernel`` has ``medium`` (``sample``,
    ``g``), ``wavelength.wavelength_nm``, ``acceptance()["degree"]`` and a
    ``config`` with ``omega_per_ns``, ``radial_range_m``, ``radial_nodes``,
    ``scattering_degree``, ``source_degree``, ``k_max_per_m``, ``k_panel_per_m``
    and ``k_order`` (a reference ``TransportKernel`` has all of them)."""
    c = kernel.config
    omega_all = np.asarray(c.omega_per_ns, float)
    indices = (np.arange(len(omega_all)) if frequency_indices is None
               else np.asarray(frequency_indices, np.int64))
    omega = omega_all[indices]
    low, high = c.radial_range_m
    radii = np.geomspace(float(low), float(high), int(c.radial_nodes))
    acceptance_degree = int(kernel.acceptance()["degree"])
    coupling = Coupling(int(c.source_degree), acceptance_degree)
    wavelengths = np.asarray(kernel.wavelength.wavelength_nm, float)
    coefficients = None
    absorption, speed = [], []
    for band, wavelength in enumerate(wavelengths):
        sample = kernel.medium.sample(float(wavelength))
        medium_band = dict(absorption_per_m=float(sample["absorption_per_m"]),
                           scattering_per_m=float(sample["scattering_per_m"]),
                           group_index=float(sample["group_index"]), g=float(kernel.medium.g))
        table = directional_table(medium_band, radii, omega, coupling,
                                  scattering_degree=int(c.scattering_degree),
                                  k_max_per_m=float(c.k_max_per_m),
                                  k_panel_per_m=float(c.k_panel_per_m), k_order=int(c.k_order))
        v = C_VACUUM_M_PER_NS / medium_band["group_index"]
        spline = radial_spline(table, radii, omega, medium_band["absorption_per_m"], v)
        if coefficients is None:
            coefficients = np.empty((len(wavelengths),) + spline.shape, dtype)
        coefficients[band] = spline
        absorption.append(medium_band["absorption_per_m"])
        speed.append(v)
        if progress is not None:
            progress(band + 1, len(wavelengths))
    bundle = dict(degree=int(c.source_degree), acceptance_degree=acceptance_degree,
                  pair_lambda=coupling.pair_lambda, pair_mu=coupling.pair_mu, omega=omega,
                  phase="flight", coefficients=coefficients, log_r=np.log(radii),
                  absorption=np.asarray(absorption, float), speed=np.asarray(speed, float))
    return wavelengths, indices, bundle
