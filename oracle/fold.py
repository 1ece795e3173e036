"""Oracle of the band fold re-fit (``fold_bundle``) in FP64 numpy.

Restates ``band_fold.py:155-208``: each radial interval, optionally
subdivided, is refitted as a cubic in log r through four equally spaced nodes
of ``sum_lambda L[l, lambda, j] e^{-(mua_lambda - mua_ref) r}
e^{i omega r (1/v_lambda - 1/v_ref)} S_lambda(t)``.

Parity pins (tests/test_fold_oracle.py):
  * at full rank the interpolation is the identity and the folded apply
    converges to the per-band apply at fourth order under subdivision
    (``tests/test_shower_route.py:102-138``), checked through the stage (b)
    oracle;
  * node bands are reproduced exactly, the residual is at most one, and the
    greedy search contains the best single band (``:36-56``).
"""
import numpy as np


def fold_bundle(bundle, fold, *, subdivide=1):
    coefficients = np.asarray(bundle["coefficients"])
    nband, nint, ndeg, npair, _, nfreq = coefficients.shape
    log_r = np.asarray(bundle["log_r"], float)
    omega = np.asarray(bundle["omega"], float)
    absorption = np.asarray(bundle["absorption"], float)
    speed = np.asarray(bundle["speed"], float)
    absorption_ref, speed_ref = float(absorption.min()), float(np.mean(speed))
    rank = len(fold.bands)
    parts = np.linspace(0.0, 1.0, int(subdivide) + 1)
    new_log_r = np.concatenate([log_r[j] + parts[:-1] * (log_r[j + 1] - log_r[j])
                                for j in range(nint)] + [log_r[-1:]])
    nodes = np.array([0.0, 1 / 3, 2 / 3, 1.0])
    out = np.zeros((rank, nint * int(subdivide), ndeg, npair, 4, nfreq), np.complex128)
    for new in range(nint * int(subdivide)):
        old = new // int(subdivide)
        low, high = new_log_r[new], new_log_r[new + 1]
        at = low + nodes * (high - low)
        radius = np.exp(at)
        inverse = np.linalg.inv(np.vander(at - low, 4))
        c = np.asarray(coefficients[:, old], np.complex128)        # (lambda, l, pair, 4, w)
        t = (at - log_r[old])[None, :, None, None, None]
        values = (((c[:, None, :, :, 0] * t + c[:, None, :, :, 1]) * t
                   + c[:, None, :, :, 2]) * t + c[:, None, :, :, 3])
        factor = (np.exp(-(absorption[:, None] - absorption_ref) * radius[None, :])[:, :, None]
                  * np.exp(1j * omega[None, None, :] * radius[None, :, None]
                           * (1.0 / speed[:, None, None] - 1.0 / speed_ref)))
        values = values * factor[:, :, None, None, :]
        folded = np.einsum("lmj,mnlpw->jnlpw", fold.interpolation, values)
        out[:, new] = np.einsum("an,jnlpw->jlpaw", inverse, folded)
    result = {key: bundle[key] for key in ("degree", "acceptance_degree", "pair_lambda",
                                          "pair_mu", "omega", "phase")}
    result.update(coefficients=out, log_r=new_log_r, absorption=np.full(rank, absorption_ref),
                  speed=np.full(rank, speed_ref), subdivide=int(subdivide))
    return result
