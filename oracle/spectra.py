"""Float64 numpy restatement of Acoular's Welch cross-spectral-matrix estimation.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows ``PowerSpectra.calc_csm`` (``spectra.py:366-416``), the block generator
``BaseSpectra._get_source_data`` (``spectra.py:171-183``), the float ``num_blocks``
(``spectra.py:281-283``), ``fftfreq`` (``spectra.py:128-168``) and the upper-triangle update
``calcCSM`` (``fastFuncs.py:19-58``).
"""

from __future__ import annotations

import numpy as np
from scipy import fft

WINDOWS = {
    'Rectangular': np.ones,
    'Hanning': np.hanning,
    'Hamming': np.hamming,
    'Bartlett': np.bartlett,
    'Blackman': np.blackman,
}
OVERLAPS = {'None': 1, '50%': 2, '75%': 4, '87.5%': 8}


def fftfreq(block_size: int, sample_freq: float) -> np.ndarray:
    """``BaseSpectra.fftfreq`` (``spectra.py:128-168``)."""
    return abs(fft.fftfreq(block_size, 1.0 / sample_freq)[: int(block_size / 2 + 1)])


def num_blocks(num_samples: int, block_size: int, overlap: str) -> float:
    """``PowerSpectra._get_num_blocks`` (``spectra.py:281-283``) -- a float, not a block count."""
    ov = OVERLAPS[overlap]
    return ov * num_samples / block_size - ov + 1


def block_starts(num_samples: int, block_size: int, overlap: str) -> np.ndarray:
    """Start sample of every block the generator ``_get_source_data`` (``spectra.py:171-183``)
    yields: hop ``block_size / overlap`` and only blocks that end inside the data."""
    hop = block_size // OVERLAPS[overlap]
    if num_samples < block_size:
        return np.zeros(0, dtype=np.int64)
    return np.arange(0, num_samples - block_size + 1, hop, dtype=np.int64)


def block_generator(data: np.ndarray, block_size: int, overlap: str, chunk: int | None = None):
    """Literal restatement of the ring-buffer generator ``_get_source_data``
    (``spectra.py:171-183``) fed by ``TimeSamples.result(block_size)`` chunks; used to pin
    :func:`block_starts`."""
    bs = block_size
    chunk = bs if chunk is None else chunk
    temp = np.empty((2 * bs, data.shape[1]))
    pos = bs
    posinc = bs / OVERLAPS[overlap]
    for s in range(0, data.shape[0], chunk):
        data_block = data[s : s + chunk]
        ns = data_block.shape[0]
        temp[bs : bs + ns] = data_block
        while pos + bs <= bs + ns:
            yield temp[int(pos) : int(pos + bs)].copy()
            pos += posinc
        temp[0:bs] = temp[bs:]
        pos -= bs


def calc_csm_update(csm: np.ndarray, spec: np.ndarray) -> np.ndarray:
    """``calcCSM`` (``fastFuncs.py:30-58``): ``csm[f, r, c] += conj(spec[f, c]) * spec[f, r]``
    for ``r <= c`` (upper triangle and diagonal); in place, returns ``csm``."""
    nf, nm = spec.shape
    upd = spec[:, :, None] * np.conj(spec[:, None, :])  # [f, r, c] = spec_r conj(spec_c)
    mask = np.triu(np.ones((nm, nm), dtype=bool))
    csm[:, mask] += upd[:, mask].astype(csm.dtype)
    return csm


def block_spectra(data: np.ndarray, block_size: int, window: str = 'Hanning', overlap: str = '50%',
                  precision: str = 'complex128') -> np.ndarray:
    """Windowed float64 rFFT of every Welch block, cast to ``precision``
    (``spectra.py:401-409``).  Returns ``[B, F, M]``."""
    data = np.asarray(data, dtype=np.float64)
    wind = WINDOWS[window](block_size)
    starts = block_starts(data.shape[0], block_size, overlap)
    idx = starts[:, None] + np.arange(block_size)[None, :]
    blocks = data[idx] * wind[None, :, None]  # [B, bs, M]
    return fft.rfft(blocks, axis=1).astype(precision)


def welch_csm(data: np.ndarray, block_size: int, window: str = 'Hanning', overlap: str = '50%',
              precision: str = 'complex128') -> np.ndarray:
    """``PowerSpectra.calc_csm`` (``spectra.py:366-416``) with float64 accumulation.

    Returns the full Hermitian CSM ``[F, M, M]`` in ``precision``.
    """
    data = np.asarray(data, dtype=np.float64)
    wind = WINDOWS[window](block_size)
    weight = float(np.dot(wind, wind))
    spec = block_spectra(data, block_size, window, overlap, precision).astype(np.complex128)
    # upper triangle accumulated over blocks; C_rc = sum_b X_r conj(X_c)
    upper = np.triu(np.einsum('bfr,bfc->frc', spec, np.conj(spec)))
    lower = np.conj(upper).transpose(0, 2, 1).copy()
    idx = np.arange(lower.shape[1])
    lower[:, idx, idx] = 0
    csm = lower + upper
    nb = num_blocks(data.shape[0], block_size, overlap)
    return (csm * (2.0 / block_size / weight / nb)).astype(precision)


def welch_csm_bins(data: np.ndarray, block_size: int, bins, window: str = 'Hanning',
                   overlap: str = '50%') -> np.ndarray:
    """The Welch CSM of a few bins only (``spectra.py:366-416``), by a direct float64 DFT of every
    block at those bins: ``[len(bins), M, M]`` complex128, Hermitian, the same blocks, window and
    ``num_blocks`` scaling as :func:`welch_csm` (which it equals on every bin; pinned in the tests),
    for arrays too large to transform whole."""
    data = np.asarray(data)
    n, m = data.shape
    wind = WINDOWS[window](block_size)
    e = np.exp(-2j * np.pi * np.outer(np.asarray(bins), np.arange(block_size)) / block_size)  # [nb, bs]
    csm = np.zeros((len(bins), m, m), dtype=np.complex128)
    for st in block_starts(n, block_size, overlap):
        x = e @ (data[st:st + block_size].astype(np.float64) * wind[:, None])  # [nb, M]
        csm += x[:, :, None] * np.conj(x[:, None, :])
    return csm * (2.0 / block_size / float(np.dot(wind, wind)) / num_blocks(n, block_size, overlap))


def synthetic_ev(eigvals: np.ndarray, freqs: np.ndarray, freq: float, num: int = 0) -> np.ndarray:
    """``PowerSpectra.synthetic_ev`` (``spectra.py:528-588``)."""
    if num == 0:
        return eigvals[np.searchsorted(freqs, freq)]
    f1 = np.searchsorted(freqs, freq * 2.0 ** (-0.5 / num))
    f2 = np.searchsorted(freqs, freq * 2.0 ** (0.5 / num))
    if f1 == f2:
        return eigvals[f1]
    return np.sum(eigvals[f1:f2], 0)
