"""CPU pins of the RTE table build (lighthit_amd/tables.py): the published
demonstration totals through the whole oracle route, and the pieces against
closed forms and independent dense solves."""
import numpy as np
import pytest

from lighthit_amd import tables


@pytest.fixture(scope="module")
def demonstration_pack():
    from lighthit_amd.fold import KernelPack
    from lighthit_amd.route import RouteConfig, RouteKernel
    from lighthit_amd.synthetic import plain_small
    w = plain_small()
    kernel = RouteKernel(w.medium, w.detector, RouteConfig(relative_time_edges_ns=w.edges_ns))
    waves, indices, bundle = tables.build_pack(kernel)
    pack = KernelPack()
    pack.add(waves, indices, bundle)
    return w, kernel, pack


def test_published_demonstration_orders(demonstration_pack):
    """README.md:133: "shower 34 active OMs, [63.74 50.68 22.5 ] pe in orders 0, 1,
    >=2". With the tables built here the oracle route gives the published order
    >=2 total and active set (order 1 is the published one; the route's order 0
    is the first-entry evaluator, see tests/test_route_host.py)."""
    from lighthit_amd.route import ShowerConfig
    from oracle import route as oracle_route
    w, kernel, pack = demonstration_pack
    bundle = pack.bundle(kernel.wavelength.wavelength_nm, 8, np.arange(65))
    out = oracle_route.transport_shower(kernel, w.elements, bundle, config=ShowerConfig())
    totals = out["charge"].sum(axis=0)
    print("orders", totals, "active", int(out["active"].sum()))
    assert abs(totals[2] - 22.5) < 0.05
    assert abs(totals[1] - 50.68) < 0.005
    assert int(out["active"].sum()) == 34
    assert not out["fallback"]


def test_pack_layout_and_spline_nodes(demonstration_pack):
    """The pack has the reference's layout, and each cubic's constant term is the
    detrended table value at its interval's left node."""
    w, kernel, pack = demonstration_pack
    bundle = pack.bundle(kernel.wavelength.wavelength_nm, 8, np.arange(65))
    c = bundle["coefficients"]
    assert c.shape == (3, 47, 9, 6, 4, 65) and c.dtype == np.complex64
    assert np.array_equal(bundle["pair_lambda"], [0, 1, 1, 2, 2, 2])
    assert np.array_equal(bundle["pair_mu"], [0, 0, 1, 0, 1, 2])
    radii = np.exp(bundle["log_r"])
    np.testing.assert_allclose(radii[[0, -1]], [2.0, 150.0], rtol=1e-12)
    # re-evaluate one band's table directly and compare the splines' node values
    band = 1
    sample = w.medium.sample(float(kernel.wavelength.wavelength_nm[band]))
    medium = dict(absorption_per_m=sample["absorption_per_m"],
                  scattering_per_m=sample["scattering_per_m"],
                  group_index=sample["group_index"], g=w.medium.g)
    coupling = tables.Coupling(8, 2)
    omega = np.asarray(bundle["omega"])[[0, 20, 64]]
    table = tables.directional_table(medium, radii, omega, coupling, scattering_degree=8,
                                     k_max_per_m=4.0, k_panel_per_m=0.1, k_order=6)
    scale = np.exp(-medium["absorption_per_m"] * radii) / (4 * np.pi * radii ** 2)
    speed = tables.C_VACUUM_M_PER_NS / medium["group_index"]
    node = table[..., 1] / scale[None, :, None, None] * np.exp(
        -1j * omega[:, None] * radii[None, :] / speed)[:, :, None, None]
    spline_at_nodes = np.moveaxis(c[band, :, :, :, 3][..., [0, 20, 64]], -1, 0)  # (w, j, l, p)
    np.testing.assert_allclose(spline_at_nodes, node[:, :-1], rtol=2e-6,
                               atol=2e-6 * np.abs(node).max())


def test_clebsch_gordan_closed_forms_and_orthogonality():
    cg = tables.clebsch_gordan
    assert cg(1, 0, 1, 0, 2, 0) == pytest.approx(np.sqrt(2 / 3), abs=1e-15)
    assert cg(1, 1, 1, -1, 0, 0) == pytest.approx(1 / np.sqrt(3), abs=1e-15)
    assert cg(1, 0, 1, 0, 1, 0) == 0.0                      # parity
    assert cg(1, 1, 1, -1, 2, 0) == pytest.approx(1 / np.sqrt(6), abs=1e-15)
    # <l1 m1; l2 m2 | J M> over m1 is orthonormal in J for fixed M
    for l1, l2, M in ((2, 3, 0), (4, 2, 1), (6, 5, -2)):
        Js = range(abs(l1 - l2), l1 + l2 + 1)
        m1s = [m for m in range(-l1, l1 + 1) if abs(M - m) <= l2]
        U = np.array([[cg(l1, m, l2, M - m, J, M) for m in m1s] for J in Js if abs(M) <= J])
        np.testing.assert_allclose(U @ U.T, np.eye(len(U)), atol=1e-13)
    # large degrees stay finite (exact rationals until the last square root)
    assert np.isfinite(cg(60, 3, 55, -3, 40, 0))


def test_k_quadrature_is_exact_without_taper():
    k, w = tables.k_quadrature(4.0, 0.1, 6, taper=False)
    assert np.sum(w * k ** 5) == pytest.approx(4.0 ** 6 / 6, rel=1e-13)
    k, w = tables.k_quadrature(4.0, 0.1, 6, taper=True)
    assert np.all(w[k < 2.0] > 0) and np.all(w >= 0)
    assert w[-1] < 1e-9 * w[0]                                 # rolled off at k_max


@pytest.mark.parametrize("d0,k", [(0.3 - 0.2j, 1.0), (0.05 - 0.0j, 0.5), (0.1 - 2.0j, 3.0)])
def test_minimal_ratios_m0_are_the_legendre_moment_ratios(d0, k):
    """m = 0: the free solution is h_l = sqrt((2l+1)/2) int P_l(x) / (d0 + i k x) dx
    (a direct Gauss-Legendre integral, far from the pole here)."""
    degree = 12
    x, w = np.polynomial.legendre.leggauss(4000)
    P = np.polynomial.legendre.legvander(x, degree + 1).T           # (l, x)
    h = np.sqrt((2 * np.arange(degree + 2) + 1) / 2) * (P @ (w / (d0 + 1j * k * x)))
    want = h[1:] / h[:-1]
    got = tables.minimal_ratios(np.array([k]), d0, degree, 0)[0]
    np.testing.assert_allclose(got, want, rtol=1e-9)


@pytest.mark.parametrize("m", [0, 1, 2, 4])
def test_tail_closed_block_equals_a_deep_truncated_solve(m):
    """The tridiagonal block closed by the free tail at L equals the same system
    truncated 400 degrees higher (hard truncation), for every azimuthal order."""
    L, deep = 8, 400
    k = np.array([0.7, 2.5])
    d0 = 0.12 - 0.4j
    gamma_all = 0.05 * 0.9 ** np.arange(L + 1)
    n = L + 1 - m
    rhs = np.zeros((2, n), complex)
    rhs[:, 0] = 1.0
    rhs[:, min(2, n - 1)] += 0.5j
    tail = tables.minimal_ratios(k, d0, L, m)[:, L]
    got = tables.solve_block(k, d0, gamma_all[m:], tail, rhs, m)
    for i, kk in enumerate(k):
        size = deep - m + 1
        a = tables.ladder(m, deep + 1)
        A = np.zeros((size, size), complex)
        gamma = np.zeros(size)
        gamma[:n] = gamma_all[m:]
        A[np.arange(size), np.arange(size)] = d0 - gamma
        off = 1j * kk * a[m + 1:m + size]
        A[np.arange(size - 1), np.arange(1, size)] = off
        A[np.arange(1, size), np.arange(size - 1)] = off
        b = np.zeros(size, complex)
        b[:n] = rhs[i]
        np.testing.assert_allclose(got[i], np.linalg.solve(A, b)[:n], rtol=1e-10, atol=1e-14)
