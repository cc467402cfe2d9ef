"""Numpy restatement of Acoular's time-domain source synthesis.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

* ``point_source``: ``PointSource.result`` (``sources.py:774-842``) with ``_fill_mic_signal_block``
  (``sources.py:77-91``), block by block with the reference's branch structure ('zeros'
  prepadding checks the sign of the index only in the first ``pre`` blocks) and the float64
  index advanced by ``up`` one sample at a time, vectorised over microphones.  Indices outside
  the upsampled signal, which numba reads from memory without a bounds check, give NaN.
* ``white_noise``: ``UncorrelatedNoiseSource.result`` (``sources.py:1911-1967``) with
  ``WNoiseGenerator.signal`` (``signals.py:255-269``): numpy's own legacy ``RandomState`` is the
  generator the reference calls, so it is called here directly.
* ``mix``: ``SourceMixer.result`` (``sources.py:2139-2185``).
"""

from __future__ import annotations

import numpy as np


def _fill(out, signal, rm, ind, blocksize, up, prepadding):
    L = len(signal)
    for b in range(blocksize):
        k = (0.5 + ind[0]).astype(np.int64)  # int() truncates toward zero
        k = np.where(k < 0, k + L, k)  # one wrap-around, as numba's negative indexing
        # numba reads indices outside the array unchecked (undefined values): NaN here
        ok = (k >= 0) & (k < L)
        vals = np.where(ok, signal[np.clip(k, 0, L - 1)], np.nan) / rm[0]
        if prepadding:
            vals = np.where(ind[0] < 0, 0.0, vals)
        out[b] = vals
        ind += up


def usignal(signal, factor: int) -> np.ndarray:
    """``SignalGenerator.usignal`` (``signals.py:93-126``): ``scipy.signal.resample`` of the signal
    to ``factor`` times its length (the reference's own call; scipy's published algorithm: the
    one-sided spectrum zero-padded, the Nyquist bin of an even length halved, scaled by
    num / n)."""
    from scipy.signal import resample

    x = np.asarray(signal, dtype=np.float64)
    return resample(x, int(factor) * x.shape[0])


def point_source(usignal, rm, ind0, up, prepadding, num_samples, num, estimate):
    """Concatenated blocks of ``PointSource.result(num)`` for the start index ``ind0 [1, M]``."""
    signal = np.asarray(usignal, dtype=np.float64)
    rm = np.asarray(rm, dtype=np.float64).reshape(1, -1)
    ind = np.array(ind0, dtype=np.float64).reshape(1, -1)
    N = int(np.ceil(estimate / num))
    out = np.empty((num, rm.shape[1]))
    blocks = []
    if prepadding == 'zeros':
        pre = -int(np.min(ind[0]) // (up * num))
        if pre >= N:
            for _ in range(N - 1):
                _fill(out, signal, rm, ind, num, up, True)
                blocks.append(out.copy())
            bsz = num_samples % num or num
            _fill(out, signal, rm, ind, bsz, up, True)
            blocks.append(out[:bsz].copy())
            return np.concatenate(blocks)
        for _ in range(pre):
            _fill(out, signal, rm, ind, num, up, True)
            blocks.append(out.copy())
    else:
        pre = 0
    for _ in range(N - pre - 1):
        _fill(out, signal, rm, ind, num, up, False)
        blocks.append(out.copy())
    bsz = num_samples % num or num
    _fill(out, signal, rm, ind, bsz, up, False)
    blocks.append(out[:bsz].copy())
    return np.concatenate(blocks)


def white_noise(seeds, num_samples, rms) -> np.ndarray:
    """``[num_samples, len(seeds)]``: channel c is ``rms * RandomState(seeds[c]).standard_normal``."""
    return np.array([rms * np.random.RandomState(int(s)).standard_normal(num_samples) for s in seeds]).T


def mix(parts, weights) -> np.ndarray:
    """Weighted sum in source order, truncated to the shortest part."""
    T = min(p.shape[0] for p in parts)
    acc = parts[0][:T] * weights[0]
    for p, w in zip(parts[1:], weights[1:]):
        acc = acc + p[:T] * w
    return acc
