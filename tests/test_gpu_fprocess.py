"""GPU parity of the block-wise FFT and power spectra (``abf_rfft_blocks``, ``abf_auto_power``,
``abf_cross_power``; fprocess.py:35-470) against the oracle (oracle/fprocess.py).

Tolerances: spectra to 2e-6 of their largest magnitude (float32 FFT against float64), power
spectra to 5e-6 of their largest value; the reference's equal-power test (auto power averaged
over the blocks equals the Welch CSM) to 1e-5 relative.
"""

import numpy as np
import pytest

import acoular_amd as aa
from oracle import fprocess as P

pytestmark = pytest.mark.gpu

WINDOWS = ['Rectangular', 'Hanning', 'Hamming', 'Blackman', 'Bartlett']


def _ts(n, nch, fs=64.0, seed=0):
    return aa.TimeSamples(data=np.random.RandomState(seed).standard_normal((n, nch)), sample_freq=fs)


def _rows(gen):
    return np.concatenate(list(gen), axis=0)


@pytest.mark.parametrize('window', WINDOWS)
@pytest.mark.parametrize('scaling', ['none', 'energy', 'amplitude'])
@pytest.mark.parametrize(('n', 'nch', 'bs', 'overlap'), [(256, 1, 64, 'None'), (1000, 5, 128, '50%'),
                                                         (4096, 16, 1024, '75%'), (77, 3, 8, 'None')])
def test_rfft_matches_oracle(gpu, window, scaling, n, nch, bs, overlap):
    ts = _ts(n, nch)
    fft = aa.RFFT(source=ts, block_size=bs, window=window, overlap=overlap, scaling=scaling)
    got = _rows(fft.result(7))
    want = P.rfft_blocks(ts.data, bs, window, overlap, scaling)
    assert got.shape == want.shape and got.dtype == np.complex128
    assert np.abs(got - want).max() <= 2e-6 * np.abs(want).max()


@pytest.mark.parametrize('window', WINDOWS)
def test_equal_power(gpu, window):
    """The reference's test_equal_power: auto power == cross power (one channel), and their mean
    over the blocks == the Welch CSM of PowerSpectra."""
    bs = 128
    ts = _ts(2 * bs, 1)
    fft = aa.RFFT(source=ts, block_size=bs, window=window)
    ap = next(aa.AutoPowerSpectra(source=fft).result(2))
    cp = next(aa.CrossPowerSpectra(source=fft).result(2))
    np.testing.assert_allclose(ap, cp, rtol=1e-6, atol=1e-9 * np.abs(ap).max())
    ps = aa.PowerSpectra(source=ts, block_size=bs, window=window, cached=False)
    np.testing.assert_allclose(ap.mean(axis=0), ps.csm[:, :, :].reshape(-1).real, rtol=1e-5,
                               atol=1e-6 * np.abs(ap).max())


@pytest.mark.parametrize(('scaling', 'single_sided'), [('power', True), ('psd', True), ('power', False)])
def test_power_spectra_match_oracle(gpu, scaling, single_sided):
    ts = _ts(3000, 6, fs=48000.0, seed=4)
    fft = aa.RFFT(source=ts, block_size=256, window='Hanning', overlap='50%', scaling='energy')
    rows = P.rfft_blocks(ts.data, 256, 'Hanning', '50%', 'energy')
    scale = P.auto_power_scale(256, 48000.0, scaling, single_sided)
    ap = aa.AutoPowerSpectra(source=fft, scaling=scaling, single_sided=single_sided)
    got = _rows(ap.result(4))
    want = P.auto_power(rows, scale)
    assert got.dtype == np.float64 and got.shape == want.shape
    assert np.abs(got - want).max() <= 5e-6 * np.abs(want).max()
    cp = aa.CrossPowerSpectra(source=fft, scaling=scaling, single_sided=single_sided, precision='complex64')
    got = _rows(cp.result(5))
    want = P.cross_power(rows, 129, 6, scale)
    assert got.dtype == np.complex64 and got.shape == want.shape
    assert np.abs(got - want).max() <= 5e-6 * np.abs(want).max()


def test_chunks_and_precision(gpu):
    ts = _ts(127, 2)
    fft = aa.RFFT(source=ts, block_size=16)
    assert [b.shape[0] for b in fft.result(3)] == P.chunks(7, 3) == [3, 3, 1]
    assert sum(b.shape[0] for b in fft.result(128)) == fft.num_samples == 7
    fft.precision = 'complex64'
    assert next(fft.result(1)).dtype == np.complex64
    short = aa.RFFT(source=_ts(10, 2), block_size=16)
    blocks = list(short.result(4))
    assert len(blocks) == 1 and blocks[0].shape == (1, 18) and not blocks[0].any()
    assert list(short.result(1)) == []
    ap = aa.AutoPowerSpectra(source=fft, precision='float32')
    assert next(ap.result(2)).dtype == np.float32


def test_cross_power_triangle_modes(gpu):
    one = aa.RFFT(source=_ts(64, 1), block_size=16)
    full = _rows(aa.CrossPowerSpectra(source=one).result(2))
    for mode in ('upper', 'lower'):
        got = _rows(aa.CrossPowerSpectra(source=one, calc_mode=mode).result(2))
        np.testing.assert_allclose(got, full)
    many = aa.RFFT(source=_ts(64, 3), block_size=16)
    with pytest.raises(ValueError):
        next(aa.CrossPowerSpectra(source=many, calc_mode='upper').result(2))


@pytest.mark.parametrize('num', [1, 4, 7, 8, 9, 16])
def test_irfft_roundtrip_and_blocks(gpu, num):
    """The reference's IRFFT tests: the inverse of the rectangular, unscaled RFFT gives back the
    time data, in blocks of any num (test_equal_fft_ifft, test_irfft_buffer); float32 FFTs both
    ways, so to 2e-6 of the largest sample."""
    ts = _ts(64, 3, seed=3)
    ifft = aa.IRFFT(source=aa.RFFT(source=ts, block_size=8))
    assert ifft.num_samples == 64
    blocks = list(ifft.result(num))
    assert all(b.shape[0] == num for b in blocks[:-1]) and blocks[-1].shape[0] <= num
    got = np.concatenate(blocks, axis=0)
    assert got.dtype == np.float64 and got.shape == ts.data.shape
    assert np.abs(got - ts.data).max() <= 2e-6 * np.abs(ts.data).max()


def test_irfft_matches_scipy(gpu):
    """Spectra that are not the transform of real blocks (nonzero imaginary parts at DC and
    Nyquist, which a real inverse transform ignores) against scipy.fft.irfft."""
    bs, nch = 256, 5
    ts = _ts(bs * 6, nch, seed=9)
    fft = aa.RFFT(source=ts, block_size=bs, window='Hanning', scaling='amplitude')
    spec = fft.result_device()
    spec[:, 0, :] += 0.3j
    spec[:, -1, :] -= 0.2j
    rows = spec.cpu().numpy().reshape(spec.shape[0], -1)
    ifft = aa.IRFFT(source=fft, precision='float32')
    got = ifft.result_device().cpu().numpy()
    want = P.irfft_blocks(rows.astype(np.complex128), bs, nch)
    assert np.abs(got - want).max() <= 2e-6 * np.abs(want).max()
    assert next(ifft.result(bs)).dtype == np.float32
