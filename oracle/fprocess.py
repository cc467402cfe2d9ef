"""Float64 numpy restatement of Acoular's block-wise frequency-domain processing.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

* ``rfft_blocks``: ``RFFT.result`` (``fprocess.py:129-180``): the windowed float64 rFFT of every
  block of the Welch block generator (``spectra.py:171-183``), cast to ``precision``, then the
  scaling of ``RFFT._scale`` (``fprocess.py:113-118``): bins other than DC and the last doubled
  for 'amplitude' and 'energy', all bins times 1 / sqrt(w.w / bs) ('none', 'energy') or 1 / sum w
  ('amplitude').  Rows ``[B, F * M]``, frequency-major with channels interlaced.
* ``chunks``: the row counts ``RFFT.result(num)`` yields, including its single row of zeros when
  no block fits and ``num > 1`` (``fprocess.py:168-180``).
* ``auto_power``: ``AutoPowerSpectra`` (``fprocess.py:320-366``), ``|X|^2`` times
  1 / bs^2, doubled when single-sided, times bs * fs for 'psd'.
* ``irfft_blocks``: ``IRFFT.result`` (``fprocess.py:235-283``): ``scipy.fft.irfft`` of every
  snapshot along the frequency axis, the blocks concatenated in time.
* ``cross_power``: ``CrossPowerSpectra.result`` with ``calc_mode='full'``
  (``fprocess.py:430-470``): per row the upper triangle of ``calcCSM`` plus its conjugate
  transpose without the diagonal, times the auto-power scale.
"""

from __future__ import annotations

import numpy as np
from scipy import fft

from .spectra import WINDOWS, block_starts


def bin_scale(block_size: int, window: str, scaling: str) -> np.ndarray:
    """The factor per bin that ``RFFT._scale`` applies (``fprocess.py:113-118, 160-164``)."""
    w = WINDOWS[window](block_size)
    sv = 1.0 / w.sum() if scaling == 'amplitude' else 1.0 / np.sqrt(np.dot(w, w) / block_size)
    s = np.full(block_size // 2 + 1, sv)
    if scaling in ('amplitude', 'energy'):
        s[1:-1] *= 2.0
    return s


def rfft_blocks(data, block_size, window='Rectangular', overlap='None', scaling='none', precision='complex128'):
    """``RFFT.result`` rows for every block, ``[B, F * M]`` of ``precision``."""
    data = np.asarray(data, dtype=np.float64)
    w = WINDOWS[window](block_size)
    starts = block_starts(data.shape[0], block_size, overlap)
    idx = starts[:, None] + np.arange(block_size)[None, :]
    spec = fft.rfft(data[idx] * w[None, :, None], n=block_size, axis=1).astype(precision)  # [B, F, M]
    if scaling in ('amplitude', 'energy'):
        spec[:, 1:-1] *= 2.0
    spec *= 1.0 / w.sum() if scaling == 'amplitude' else 1.0 / np.sqrt(np.dot(w, w) / block_size)
    return spec.reshape(spec.shape[0], -1)


def chunks(num_blocks: int, num: int) -> list[int]:
    """Rows per block that ``RFFT.result(num)`` yields (``fprocess.py:168-180``)."""
    if num_blocks == 0:
        return [1] if num > 1 else []
    out = [num] * (num_blocks // num)
    if num_blocks % num:
        out.append(num_blocks % num)
    return out


def auto_power_scale(block_size, sample_freq, scaling='power', single_sided=True) -> float:
    """``AutoPowerSpectra._get_scaling_value`` (``fprocess.py:320-326``)."""
    scale = 1 / block_size**2
    if single_sided:
        scale *= 2
    if scaling == 'psd':
        scale *= block_size * sample_freq
    return scale


def auto_power(rows, scale, precision='float64'):
    """``AutoPowerSpectra.result`` on spectrum rows (``fprocess.py:359-366``)."""
    rows = np.asarray(rows)
    return ((rows * rows.conjugate()).real * scale).astype(precision)


def cross_power(rows, num_freqs, num_channels, scale, precision='complex128'):
    """``CrossPowerSpectra.result`` with ``calc_mode='full'`` on spectrum rows
    (``fprocess.py:446-470``): ``[n, F * M * M]``."""
    rows = np.asarray(rows)
    out = np.zeros((rows.shape[0], num_freqs * num_channels**2), dtype=precision)
    idx = np.arange(num_channels)
    for i in range(rows.shape[0]):
        x = rows[i].astype(precision).reshape(num_freqs, num_channels)
        upper = np.triu(x[:, :, None] * np.conj(x[:, None, :]))  # calcCSM: conj(X_c) X_r, r <= c
        lower = np.conj(upper).transpose(0, 2, 1).copy()
        lower[:, idx, idx] = 0
        out[i] = (lower + upper).reshape(-1)
    return out * scale


def irfft_blocks(rows, block_size, num_channels):
    """``IRFFT.result`` samples of spectrum rows ``[n, F * M]``: ``[n * block_size, M]``."""
    rows = np.asarray(rows)
    nf = block_size // 2 + 1
    out = [fft.irfft(r.reshape(nf, num_channels), n=block_size, axis=0) for r in rows]
    return np.concatenate(out, axis=0) if out else np.zeros((0, num_channels))
