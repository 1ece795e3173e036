"""GPU parity of stage (c), lh_readout_bins, against the FP64 oracle."""
import numpy as np
import pytest
from scipy.special import ndtr

from oracle import readout as oracle

pytestmark = pytest.mark.gpu


def _spectra(rng, omega, count, origins):
    """Smooth order>=2-like spectra: delayed exponential pulses at the OM origins."""
    tau = rng.uniform(10., 60., count)
    delay = rng.uniform(0., 150., count)
    charge = rng.uniform(0.5, 20., count)
    return (charge / (1 - 1j * omega[:, None] * tau)
            * np.exp(1j * omega[:, None] * (origins + delay)))


@pytest.mark.parametrize("precision", ["fp32", "split"])
@pytest.mark.parametrize("count,bin_ns,seed", [(34, 5., 0), (1, 1., 1), (77, 1., 2)])
def test_readout_matches_oracle(device, count, bin_ns, seed, precision):
    import torch
    from lighthit_amd.readout import ReadoutBasis
    rng = np.random.default_rng(seed)
    omega = np.linspace(0., 1.2, 65)
    readout = np.linspace(0., 1.2, 321)
    edges = np.arange(-60., 745., bin_ns)
    origins = rng.uniform(100., 3000., count)
    spectrum = _spectra(rng, omega, count, origins)
    expected = oracle.route_readout(omega, spectrum, origins, readout, edges)
    basis = ReadoutBasis(omega, edges, readout)
    dev_spec = torch.as_tensor(spectrum.astype(np.complex64), device=device)
    dev_orig = torch.as_tensor(origins, device=device)
    got = basis.apply(dev_spec, dev_orig, precision=precision).double().cpu().numpy()
    # FP32 basis and spectra, or FP16 hi + lo pairs (about 22 bits of every
    # scaled row and observation): bin errors relative to the OM's total |bins|
    scale = np.abs(expected).sum(axis=1, keepdims=True)
    assert np.max(np.abs(got - expected) / scale) < (2e-6 if precision == "fp32" else 5e-6)
    np.testing.assert_allclose(got.sum(axis=1), expected.sum(axis=1), rtol=1e-5)


def test_inverse_bins_gaussian_delta_on_gpu(device):
    from lighthit_amd.readout import inverse_bins
    omega = np.linspace(0, 5, 2001)
    t0, sigma = 12., 2.
    edges = np.arange(-4, 30, 1.)
    spectrum = np.exp(1j * omega * t0)[:, None]
    got = inverse_bins(omega, spectrum, edges, sigma)[0]
    expected = ndtr((edges[1:] - t0) / sigma) - ndtr((edges[:-1] - t0) / sigma)
    np.testing.assert_allclose(got, expected, atol=2e-6)


def test_empty_observation_set_is_a_no_op(device):
    import torch
    from lighthit_amd.readout import ReadoutBasis
    basis = ReadoutBasis(np.linspace(0., 1.2, 65), np.arange(-60., 745., 5.))
    out = basis.apply(torch.zeros((65, 0), dtype=torch.complex64, device=device))
    assert out.shape == (0, 160)


def test_split_readout_tracks_fp32(device):
    """The split-precision product stays within a few FP32 roundings of the FP32 one."""
    import torch
    from lighthit_amd.readout import ReadoutBasis
    rng = np.random.default_rng(9)
    omega = np.linspace(0., 1.2, 65)
    edges = np.arange(-60., 1501., 1.)
    origins = rng.uniform(0., 2000., 100)
    spectrum = _spectra(rng, omega, 100, origins)
    basis = ReadoutBasis(omega, edges, np.linspace(0., 1.2, 321))
    dev_spec = torch.as_tensor(spectrum.astype(np.complex64), device=device)
    dev_orig = torch.as_tensor(origins, device=device)
    a = basis.apply(dev_spec, dev_orig).double().cpu().numpy()
    b = basis.apply(dev_spec, dev_orig, precision="split").double().cpu().numpy()
    scale = np.abs(a).sum(axis=1, keepdims=True)
    assert np.max(np.abs(a - b) / scale) < 2e-6
