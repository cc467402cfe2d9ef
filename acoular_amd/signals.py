"""Signal generators feeding the point sources (``acoular/signals.py``).

A generator describes one source signal: ``signal()`` returns ``num_samples`` float64 values at
``sample_freq`` and ``usignal(factor)`` the same signal resampled to ``factor`` times the rate
(``signals.py:93-126``, Fourier resampling).  ``signal()`` is computed on the host exactly as the
reference computes it; ``signal_device()`` (white noise generated on the GPU) and
``usignal_device()`` (the resampling on the GPU, ``abf_resample``) feed the point sources, whose
per-microphone propagation runs on the GPU as well (``acoular_amd.sources``).
"""

from __future__ import annotations

import numpy as np

from ._base import HasParams


class SignalGenerator(HasParams):
    """Base of the signal generators (``signals.py:51-126``)."""

    _params = ('sample_freq', 'num_samples')

    def __init__(self, **kwargs):
        object.__setattr__(self, 'sample_freq', 1.0)
        object.__setattr__(self, 'num_samples', 0)
        super().__init__(**kwargs)

    def __setattr__(self, name, value):
        if name == 'sample_freq':
            value = float(value)
        if name == 'num_samples':
            value = int(value)
            if value < 0:
                raise ValueError('num_samples must be >= 0')
        object.__setattr__(self, name, value)

    def signal(self) -> np.ndarray:
        raise NotImplementedError

    def usignal(self, factor: int) -> np.ndarray:
        """The signal resampled to ``factor * num_samples`` samples (``signals.py:93-126``)."""
        return self.usignal_device(factor).cpu().numpy()

    def signal_device(self):
        """``signal()`` as a float64 device tensor."""
        from . import _lib

        return _lib.to_device(np.ascontiguousarray(self.signal(), dtype=np.float64))

    def usignal_device(self, factor: int):
        """``usignal(factor)`` as a float64 device tensor: the Fourier resampling of
        ``scipy.signal.resample`` (``signals.py:93-126``) on the device (``abf_resample``)."""
        return resample_device(self.signal_device(), factor)


def resample_device(x, factor: int):
    """A float64 device signal resampled to ``factor`` times its length (``abf_resample``, the
    algorithm of ``scipy.signal.resample`` that ``usignal`` calls)."""
    from . import _lib

    factor = int(factor)
    if factor < 1:
        raise ValueError('factor must be >= 1')
    torch = _lib.torch()
    x = x.reshape(-1).to(torch.float64).contiguous()
    n = int(x.numel())
    y = torch.empty(factor * n, dtype=torch.float64, device=x.device)
    if n == 0:
        return y
    lib = _lib.lib()
    ws = _lib.workspace(lib.abf_resample_workspace_size(n, factor * n))
    _lib.check(lib.abf_resample(_lib.ptr(x), n, factor * n, _lib.ptr(y), _lib.ptr(ws), ws.numel(),
                                _lib.stream_handle()), 'usignal')
    return y


def white_noise_device(seeds, num_samples: int, rms=1.0):
    """Standard normal streams of numpy's legacy ``RandomState(seed)`` for every seed, times
    ``rms`` (``abf_white_noise``, one workgroup per stream): float64 device tensor
    ``[num_samples, len(seeds)]``."""
    from . import _lib

    torch = _lib.torch()
    seeds = np.asarray(seeds, dtype=np.int64).ravel()
    if seeds.size and (seeds.min() < 0 or seeds.max() > 2**32 - 1):
        raise ValueError('Seed must be between 0 and 2**32 - 1')
    out = torch.empty((int(num_samples), seeds.size), dtype=torch.float64, device=_lib.device())
    if num_samples == 0 or seeds.size == 0:
        return out
    sd = torch.from_numpy(seeds.astype(np.uint32).view(np.int32)).to(out.device)
    _lib.check(_lib.lib().abf_white_noise(_lib.ptr(sd), seeds.size, int(num_samples), float(rms), _lib.ptr(out),
                                          _lib.stream_handle()), 'white noise')
    return out


class NoiseGenerator(SignalGenerator):
    """Noise with an RMS amplitude and a seed (``signals.py:166-202``)."""

    _params = SignalGenerator._params + ('rms', 'seed')

    def __init__(self, **kwargs):
        object.__setattr__(self, 'rms', 1.0)
        object.__setattr__(self, 'seed', 0)
        super().__init__(**kwargs)

    def __setattr__(self, name, value):
        if name == 'rms':
            value = float(value)
        if name == 'seed':
            value = int(value)
        super().__setattr__(name, value)


class WNoiseGenerator(NoiseGenerator):
    """White noise: ``rms * RandomState(seed).standard_normal(num_samples)`` (``signals.py:204-269``)."""

    def signal(self) -> np.ndarray:
        rnd_gen = np.random.RandomState(self.seed)
        return self.rms * rnd_gen.standard_normal(self.num_samples)

    def signal_device(self):
        """The same numbers generated on the device (``abf_white_noise``: numpy's legacy MT19937
        stream and polar method; equal to ``signal()`` up to the last bit of the float64 log)."""
        return white_noise_device([self.seed], self.num_samples, self.rms).reshape(-1)


class SineGenerator(SignalGenerator):
    """``amplitude * sin(2 pi freq t + phase)`` (``signals.py:470-561``)."""

    _params = SignalGenerator._params + ('freq', 'phase', 'amplitude')

    def __init__(self, **kwargs):
        object.__setattr__(self, 'freq', 1000.0)
        object.__setattr__(self, 'phase', 0.0)
        object.__setattr__(self, 'amplitude', 1.0)
        super().__init__(**kwargs)

    def __setattr__(self, name, value):
        if name in ('freq', 'phase', 'amplitude'):
            value = float(value)
        super().__setattr__(name, value)

    def signal(self) -> np.ndarray:
        t = np.arange(self.num_samples, dtype=float) / self.sample_freq
        return self.amplitude * np.sin(2 * np.pi * self.freq * t + self.phase)
