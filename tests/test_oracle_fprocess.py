"""Pins of the oracle's block-wise FFT and power spectra (oracle/fprocess.py) without a GPU:
the reference's own equal-power test (fprocess tests, ``test_equal_power``: the mean of the
auto-power rows equals the Welch CSM), a bin-centred tone under 'amplitude' scaling (its
amplitude in its bin, nothing elsewhere), the row counts of ``RFFT.result(num)`` including the
reference's zero row, and the cross-power matrices (rank one per snapshot, diagonal = auto
power, equal to calcCSM's upper triangle plus its conjugate transpose)."""

import numpy as np
import pytest

from oracle import fprocess as P
from oracle import spectra as S

WINDOWS = ['Rectangular', 'Hanning', 'Hamming', 'Blackman', 'Bartlett']


@pytest.mark.parametrize('window', WINDOWS)
def test_equal_power(window):
    bs, fs = 128, 64.0
    x = np.random.RandomState(0).standard_normal(2 * bs)[:, None]
    rows = P.rfft_blocks(x, bs, window)
    scale = P.auto_power_scale(bs, fs)
    ap = P.auto_power(rows, scale)
    cp = P.cross_power(rows, bs // 2 + 1, 1, scale)
    np.testing.assert_allclose(ap, cp.real, rtol=1e-12)
    np.testing.assert_allclose(cp.imag, 0.0, atol=1e-15)
    csm = S.welch_csm(x, bs, window, 'None')
    np.testing.assert_allclose(ap.mean(axis=0), csm.reshape(-1).real, rtol=1e-10, atol=1e-18)


@pytest.mark.parametrize('k0', [1, 5, 31])
def test_tone_amplitude_scaling(k0):
    bs, amp = 64, 0.7
    n = np.arange(3 * bs)
    x = (amp * np.cos(2 * np.pi * k0 * n / bs + 0.3))[:, None]
    rows = P.rfft_blocks(x, bs, 'Rectangular', scaling='amplitude')
    mag = np.abs(rows)
    np.testing.assert_allclose(mag[:, k0], amp, rtol=1e-12)
    assert np.delete(mag, k0, axis=1).max() < 1e-12


def test_chunk_rows():
    assert P.chunks(4, 128) == [4]
    assert P.chunks(5, 2) == [2, 2, 1]
    assert P.chunks(4, 2) == [2, 2]
    assert P.chunks(0, 1) == []
    assert P.chunks(0, 8) == [1]  # the reference yields its zeroed buffer row


def test_cross_power_structure():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((300, 5))
    rows = P.rfft_blocks(x, 32, 'Hanning', '50%', 'energy')
    nf, nc = 17, 5
    scale = P.auto_power_scale(32, 100.0, 'psd')
    cp = P.cross_power(rows, nf, nc, scale).reshape(rows.shape[0], nf, nc, nc)
    ap = P.auto_power(rows, scale).reshape(rows.shape[0], nf, nc)
    np.testing.assert_allclose(np.einsum('bfii->bfi', cp).real, ap, rtol=1e-12)
    xs = rows.reshape(rows.shape[0], nf, nc)
    np.testing.assert_allclose(cp, scale * xs[..., :, None] * np.conj(xs[..., None, :]), rtol=1e-12, atol=1e-20)
    # against calcCSM's upper triangle on one snapshot
    upper = S.calc_csm_update(np.zeros((nf, nc, nc), complex), xs[3])
    np.testing.assert_allclose(np.triu(cp[3]), upper * scale, rtol=1e-12, atol=1e-20)


def test_irfft_inverts_rfft():
    """The reference's test_equal_fft_ifft: with a rectangular window and no scaling the inverse
    transform of the spectra gives back the samples of every complete block."""
    x = np.random.default_rng(8).standard_normal((64, 3))
    rows = P.rfft_blocks(x, 8, 'Rectangular')
    np.testing.assert_allclose(P.irfft_blocks(rows, 8, 3), x, atol=1e-14)
