"""Signal generators feeding the point sources (``acoular/signals.py``).

A generator describes one source signal: ``signal()`` returns ``num_samples`` float64 values at
``sample_freq`` and ``usignal(factor)`` the same signal resampled to ``factor`` times the rate
(``signals.py:93-126``, Fourier resampling).  These are single 1-D sequences computed on the host
exactly as the reference computes them; the per-microphone propagation and the per-channel noise
of whole arrays run on the GPU (``acoular_amd.sources``).
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
        from scipy.signal import resample

        return resample(self.signal(), factor * self.num_samples)


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
