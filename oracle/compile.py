"""Oracle of the axial-slice source compile (order >=2 source) in FP64.

Restates ``compile_axial_slices`` (``_shower_basis.py:188-645``) for the
options the shower route uses (``shower/ge2.py:49-55``): one light field,
stencils centred on the axis, no point selection, no axial clock, a linear or
nearest time lattice. The steps follow the reference: chord quadrature nodes
(``:300-303``), axial Lagrange weights on ``cell_m`` slices (``_axial_weights``,
``:129-135``), time ticks of ``time_step_ns`` (``:330-349``), one lattice row per
occupied (slice, tick) (``:356-372``), the slice geometry and thin/wide
decision from the axial weights alone (``:373-396``), the deposit of
``weight * moment * P_l(cone) Y_lm(u)`` per row (``:482-555``), the phase sum
``sum_tick D e^{i omega t_tick}`` per slice and the five-point stencil
(``:563-598``), and the point layout (``:599-614``).

Parity pins (tests/test_compile_oracle.py):
  * moments of the compiled basis equal those of the chords: total, z, z^2, z^3
    (cubic), x, y, x^2, y^2 (stencil) — ``tests/test_shower_basis.py:32-63``;
  * the time-lattice error against the exact ``e^{i omega t}`` sum of the same
    axial weights is within ``(omega dt)^2 / 8`` (``:66-82``);
  * the charge-only compile equals frequency 0 of a full compile, with one
    time row per slice (``:85-100``);
  * harmonics against the directional-harmonic oracle (itself pinned on SciPy)
    and ``P_l`` against numpy's Legendre series.
"""
from __future__ import annotations

import numpy as np


def layout(degree, max_m):
    """``starts``, kept flat positions ``l*l+l+m`` and ``l`` of every kept channel
    (``axial_fast.py:111-123``)."""
    starts, kept, ell_of = [0], [], []
    for ell in range(degree + 1):
        for m in range(-min(ell, max_m), min(ell, max_m) + 1):
            kept.append(ell * ell + ell + m)
            ell_of.append(ell)
        starts.append(len(kept))
    return np.asarray(starts), np.asarray(kept), np.asarray(ell_of)


def harmonics(vectors, degree, max_m):
    """Real orthonormal harmonics of (N, 3) directions on the kept channels
    (``axial_fast.py:21-52``), vectorised over directions."""
    v = np.asarray(vectors, float)
    x, y, z = v[:, 0], v[:, 1], v[:, 2]
    radius = np.sqrt(x * x + y * y + z * z)
    mu = z / radius
    transverse = np.sqrt(x * x + y * y)
    sine = transverse / radius
    safe = np.where(transverse == 0.0, 1.0, transverse)
    cp = np.where(transverse == 0.0, 1.0, x / safe)
    sp = np.where(transverse == 0.0, 0.0, y / safe)
    starts, _, _ = layout(degree, max_m)
    out = np.zeros((len(v), int(starts[-1])))
    diagonal = np.full(len(v), 1.0 / np.sqrt(4 * np.pi))
    cm, sm = np.ones(len(v)), np.zeros(len(v))
    for m in range(max_m + 1):
        if m:
            diagonal = diagonal * np.sqrt((2.0 * m + 1) / (2.0 * m)) * sine
            cm, sm = cm * cp - sm * sp, sm * cp + cm * sp
        prev2, prev1 = None, None
        for ell in range(m, degree + 1):
            if ell == m:
                value = diagonal
            elif ell == m + 1:
                value = np.sqrt(2 * m + 3) * mu * diagonal
            else:
                a = np.sqrt((4.0 * ell * ell - 1) / (ell * ell - m * m))
                b = np.sqrt(((2.0 * ell + 1) * ((ell - 1) ** 2 - m * m))
                            / ((2.0 * ell - 3) * (ell * ell - m * m)))
                value = a * mu * prev1 - b * prev2
            prev2, prev1 = prev1, value
            centre = starts[ell] + min(ell, max_m)
            if m == 0:
                out[:, centre] = value
            else:
                out[:, centre - m] = np.sqrt(2) * value * sm
                out[:, centre + m] = np.sqrt(2) * value * cm
    return out


def legendre(cosines, degree):
    """``P_l(c)`` for l = 0..degree, (N, degree+1) (``_shower_basis.py:82-100``)."""
    c = np.asarray(cosines, float)
    out = np.ones((len(c), degree + 1))
    if degree >= 1:
        out[:, 1] = c
    for ell in range(2, degree + 1):
        out[:, ell] = ((2 * ell - 1) * c * out[:, ell - 1] - (ell - 1) * out[:, ell - 2]) / ell
    return out


def axial_weights(r, order):
    """Slice shifts and Lagrange weights at fractional offset ``r`` (``_shower_basis.py:129-135``)."""
    if order == 1:
        return (0, 1), (1 - r, r)
    return ((-1, 0, 1, 2),
            (-r * (r - 1) * (r - 2) / 6, (r + 1) * (r - 1) * (r - 2) / 2,
             -(r + 1) * r * (r - 2) / 2, (r + 1) * r * (r - 1) / 6))


def compile_axial_slices(field, degree, omega_per_ns, *, azimuthal_degree=4, cell_m=0.25,
                         axial_order=3, transverse="stencil", stencil_m=0.2,
                         stencil_threshold_m=0.0, time_step_ns=0.02, time_deposit="linear",
                         element_order="auto", frame=None):
    """Axial-slice source of one light field (``_shower_basis.py:188-645``)."""
    from .host import AxialSource, axis_frame
    if not (cell_m > 0 and time_step_ns > 0 and stencil_m > 0 and stencil_threshold_m >= 0):
        raise ValueError("cell_m, time_step_ns and stencil_m must be positive")
    if axial_order not in (1, 3) or transverse not in ("collapse", "stencil"):
        raise ValueError("axial_order must be 1 or 3 and transverse collapse or stencil")
    if time_deposit not in ("linear", "nearest"):
        raise ValueError("the oracle restates the linear and nearest time lattices")
    frame = frame or axis_frame(field)
    omega = np.atleast_1d(np.asarray(omega_per_ns, float))
    charge_only = len(omega) == 1 and omega[0] == 0
    photons = np.asarray(field.photons, float)
    count = len(field)
    max_m = min(int(degree), int(azimuthal_degree))
    starts, kept, ell_of = layout(int(degree), max_m)
    nch = len(kept)
    if element_order == "auto":
        element_order = 1 if float(np.max(field.length_m)) <= cell_m / 10 else 2
    nodes, node_weights = np.polynomial.legendre.leggauss(int(element_order))
    nodes, node_weights = (nodes + 1) / 2, node_weights / 2
    reference = float(np.min(field.start_ns))
    margin = 1 if axial_order == 1 else 2
    z_origin = float(frame.coordinates(field)[0].min()) - margin * cell_m
    moments = 1 if transverse == "collapse" else 5
    coordinates = [frame.coordinates(field, fraction) for fraction in nodes]
    time_max = max(float(np.max(c[3])) for c in coordinates)
    kmax = 1 if charge_only else int(np.floor((time_max - reference) / time_step_ns)) + 3

    # lattice links of every chord node: (row key, axial x tick weight)
    keys, weights, factors, axial = [], [], [], []
    for (z, offset, _, time), node_weight in zip(coordinates, node_weights):
        if moments == 1:
            mom = np.ones((count, 1))
        else:
            x, y = offset @ frame.first, offset @ frame.second
            mom = np.stack([np.ones(count), x, y, x * x, y * y], axis=1)
        factors.append(photons[:, None] * mom * node_weight)
        zl = (z - z_origin) / cell_m
        zb = np.floor(zl).astype(np.int64)
        if charge_only:
            tb, ticks = np.zeros(count, np.int64), ((0, np.ones(count)),)
        elif time_deposit == "linear":
            tl = (np.asarray(time, float) - reference) / time_step_ns
            tb = np.floor(tl).astype(np.int64)
            ticks = ((0, 1 - (tl - tb)), (1, tl - tb))
        else:
            tl = (np.asarray(time, float) - reference) / time_step_ns
            tb, ticks = np.rint(tl).astype(np.int64), ((0, np.ones(count)),)
        shifts, zw = axial_weights(zl - zb, axial_order)
        axial.append([(zb + shift, w) for shift, w in zip(shifts, zw)])
        for shift, w in zip(shifts, zw):
            for dt, wt in ticks:
                keys.append((zb + shift) * kmax + tb + dt)
                weights.append(w * wt)
    links = len(keys) // len(nodes)
    keys = np.stack(keys, axis=1)                  # (count, nodes * links)
    weights = np.stack(weights, axis=1)
    unique, inverse = np.unique(keys, return_inverse=True)
    rows = inverse.reshape(keys.shape)
    cell_of, tick_of = unique // kmax, unique % kmax
    slices, row_start = np.unique(cell_of, return_index=True)
    row_start = np.append(row_start, len(unique))

    # slice geometry from the axial weights alone
    zero = np.zeros((len(slices), moments))
    for node_links, factor in zip(axial, factors):
        for cells, w in node_links:
            at = np.searchsorted(slices, cells)
            for k in range(moments):
                zero[:, k] += np.bincount(at, weights=w * factor[:, k], minlength=len(slices))
    if moments == 1:
        wide = np.zeros(len(slices), bool)
        centroid = np.zeros((len(slices), 2))
    else:
        mass = zero[:, 0]
        safe = np.where(np.abs(mass) > 0, mass, 1.0)
        wide = np.sqrt(np.abs(zero[:, 3] + zero[:, 4]) / np.abs(safe)) > stencil_threshold_m
        centroid = np.stack([np.where(wide, 0.0, zero[:, 1] / safe),
                             np.where(wide, 0.0, zero[:, 2] / safe)], axis=1)

    # deposit: weight * moment * P_l(cone) Y_lm(direction) per lattice row
    angular = (harmonics(frame.rotate(field.direction), int(degree), max_m)
               * legendre(field.cone_cosine, int(degree))[:, ell_of])
    deposit = np.zeros((len(unique), moments, nch))
    for n in range(len(nodes)):
        span = slice(n * links, (n + 1) * links)
        for j in range(links):
            contribution = (weights[:, span][:, j, None, None] * factors[n][:, :, None]
                            * angular[:, None, :])
            np.add.at(deposit, rows[:, span][:, j], contribution)

    # phase sum per slice and the five-point stencil
    h = float(stencil_m)
    count_wide = int(wide.sum())
    wide_index = np.cumsum(wide) - 1
    npoints = len(slices) + 4 * count_wide
    channels = np.zeros((npoints, nch, len(omega)), complex)
    times = tick_of * time_step_ns
    for cell in range(len(slices)):
        span = slice(row_start[cell], row_start[cell + 1])
        phase = np.exp(1j * np.outer(times[span], omega))
        values = np.einsum("rkc,rw->kcw", deposit[span], phase)
        if not wide[cell]:
            channels[cell] = values[0]
            continue
        m0, mx, my, mxx, myy = values
        stencil = (m0 - (mxx + myy) / h ** 2, 0.5 * (mxx / h ** 2 + mx / h),
                   0.5 * (mxx / h ** 2 - mx / h), 0.5 * (myy / h ** 2 + my / h),
                   0.5 * (myy / h ** 2 - my / h))
        ids = [cell] + [len(slices) + o * count_wide + wide_index[cell] for o in range(4)]
        for point, value in zip(ids, stencil):
            channels[point] = value
    z_centres = z_origin + slices * cell_m
    offsets = [(h, 0.0), (-h, 0.0), (0.0, h), (0.0, -h)]
    points_z = np.concatenate([z_centres] + [z_centres[wide]] * 4)
    points_t = np.concatenate([centroid] + [np.tile(o, (count_wide, 1)) for o in offsets])
    summary = dict(elements=count, slices=len(slices), stencil_slices=count_wide,
                   time_rows=len(unique), element_order=int(element_order),
                   z_origin_m=z_origin, kmax=kmax, photons=float(photons.sum()))
    return AxialSource(frame, points_z, points_t, channels, int(degree), int(azimuthal_degree),
                       reference, summary)
