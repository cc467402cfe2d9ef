"""Pins of the oracle's source synthesis (oracle/sources.py).

* ``PointSource``: the block shapes of the reference's doctest (sources.py:714-716: 6 samples in
  blocks of 4 -> (4, 9), (2, 9), all indices outside the signal); with small indices the sample-by-sample index equals the closed
  form ind0 + b * up, so every output equals signal[int(0.5 + ind0 + b up)] / r; 'zeros' gives
  exactly zero before the arrival and the same values after it; 'loop' reads the end of the
  signal there (numpy's negative indices).
* White noise: numpy's legacy generator's known first values for seed 0.
"""

import numpy as np
import pytest

from oracle import sources as O


def test_point_source_block_shapes_of_the_doctest():
    sig = np.sin(np.arange(6 * 16) * 0.1)
    rm = np.array([1.0, 1.2, 1.4, 1.1, 1.3, 1.5, 1.05, 1.25, 1.45])
    ind0 = (-rm / 343.0) * 51200 * 16
    out = O.point_source(sig, rm, ind0, 16, 'loop', 6, 4, 6.0)
    assert out.shape == (6, 9)
    # the delays exceed the 6-sample signal: numba reads outside it, the oracle gives NaN
    assert np.isnan(out).all()


def test_point_source_closed_form_and_prepadding():
    rng = np.random.default_rng(1)
    up, n = 4, 300
    sig = rng.standard_normal(n * up)
    rm = np.array([0.7, 1.3, 2.1])
    ind0 = (-rm / 343.0) * 8000.0 * up
    loop = O.point_source(sig, rm, ind0, up, 'loop', n, 64, float(n))
    zeros = O.point_source(sig, rm, ind0, up, 'zeros', n, 64, float(n))
    b = np.arange(n)[:, None]
    x = ind0[None, :] + b * up
    k = (0.5 + x).astype(np.int64)
    want = sig[np.where(k < 0, k + len(sig), k)] / rm[None, :]
    np.testing.assert_array_equal(loop, want)
    np.testing.assert_array_equal(zeros, np.where(x < 0, 0.0, want))
    assert np.all(zeros[:2] == 0.0) and np.all(loop[:2] != 0.0)


def test_white_noise_known_values():
    w = O.white_noise([0], 6, 1.0)[:, 0]
    np.testing.assert_allclose(w, [1.764052345967664, 0.4001572083672233, 0.9787379841057392, 2.240893199201458,
                                   1.8675579901499675, -0.977277879876411], rtol=1e-15)


@pytest.mark.parametrize(('n', 'k', 'up'), [(64, 5, 16), (51, 7, 4), (200, 99, 2), (10, 1, 1)])
def test_usignal_of_a_periodic_tone_is_the_tone(n, k, up):
    """Fourier resampling is exact for a tone that is periodic in the record and below Nyquist:
    sin(2 pi k t / n) sampled n times becomes the same sine at up n points."""
    x = np.sin(2 * np.pi * k * np.arange(n) / n + 0.4)
    y = O.usignal(x, up)
    assert y.shape == (up * n,)
    np.testing.assert_allclose(y, np.sin(2 * np.pi * k * np.arange(up * n) / (up * n) + 0.4), atol=1e-12)
