"""Time-data sources (``acoular/sources.py``).

``TimeSamples`` holds a ``(num_samples, num_channels)`` array and yields it block-wise like the
reference's ``result(num)`` generator (``sources.py:417-470``), the last block possibly shorter.
``PointSource``, ``UncorrelatedNoiseSource`` and ``SourceMixer`` synthesise microphone signals on
the GPU (``abf_point_source``, ``abf_white_noise``, ``abf_mix``) with the reference's arithmetic;
their ``result_device(num)`` hands the whole signal to ``PowerSpectra`` without leaving HBM.
"""

from __future__ import annotations

import numpy as np

from ._base import HasParams, digest_of


class TimeSamples(HasParams):
    _params = ('data', 'sample_freq')

    def __init__(self, **kwargs):
        self.data = np.zeros((0, 0))
        self.sample_freq = 1.0
        super().__init__(**kwargs)

    def __setattr__(self, name, value):
        if name == 'data':
            value = np.asarray(value)
            if value.ndim != 2:
                msg = 'data must have shape (num_samples, num_channels)'
                raise ValueError(msg)
        if name == 'sample_freq':
            value = float(value)
        object.__setattr__(self, name, value)

    @property
    def digest(self) -> str:
        # the sample array is identified by object identity and shape, not hashed (it can be GBs)
        return digest_of('TimeSamples', id(self.data), self.data.shape, self.data.dtype.str, self.sample_freq)

    @property
    def num_samples(self) -> int:
        return self.data.shape[0]

    @property
    def num_channels(self) -> int:
        return self.data.shape[1]

    def result(self, num=128):
        if self.num_samples == 0:
            msg = 'no samples available'
            raise OSError(msg)
        for i in range(0, self.num_samples, num):
            yield self.data[i : i + num]


def _block_plan(num_samples: int, num_samples_estimate: float, num: int) -> int:
    """Samples a block generator of ``PointSource.result`` yields in total (``sources.py:799-842``):
    ``ceil(estimate / num)`` blocks, the last one ``num_samples % num or num`` long."""
    nblocks = int(np.ceil(num_samples_estimate / num))
    if nblocks < 1:
        return 0
    return (nblocks - 1) * num + (num_samples % num or num)


def _blocks(dev, num):
    """Host blocks of ``num`` rows of a device tensor (the last one possibly shorter)."""
    host = dev.cpu().numpy()
    for i in range(0, host.shape[0], num):
        yield host[i : i + num]


class PointSource(HasParams):
    """A stationary point source emitting ``signal`` at ``loc`` (``sources.py:637-842``).

    ``result_device(num)`` computes every microphone's signal in one ``abf_point_source`` launch as a
    float64 device tensor ``[num_samples, num_mics]``; ``result(num)`` yields it block by block like
    the reference's generator.  ``prepadding`` 'loop' reads the periodically continued signal
    before ``start_t``, 'zeros' reads zeros.  Samples whose index falls outside the upsampled
    signal even after one wrap-around (the reference reads them unchecked from memory) are NaN.
    """

    _params = ('signal', 'loc', 'mics', 'env', 'start_t', 'start', 'prepadding', 'up')

    def __init__(self, **kwargs):
        from .environments import Environment

        object.__setattr__(self, 'signal', None)
        object.__setattr__(self, 'loc', np.array([[0.0], [0.0], [1.0]]))
        object.__setattr__(self, 'mics', None)
        object.__setattr__(self, 'env', Environment())
        object.__setattr__(self, 'start_t', 0.0)
        object.__setattr__(self, 'start', 0.0)
        object.__setattr__(self, 'prepadding', 'loop')
        object.__setattr__(self, 'up', 16)
        super().__init__(**kwargs)

    def __setattr__(self, name, value):
        if name == 'loc':
            value = np.asarray(value, dtype=float).reshape((3, 1))
        if name in ('start_t', 'start'):
            value = float(value)
        if name == 'up':
            value = int(value)
        if name == 'prepadding' and value not in ('loop', 'zeros'):
            raise ValueError("prepadding must be 'loop' or 'zeros'")
        object.__setattr__(self, name, value)

    @property
    def num_channels(self) -> int:
        return self.mics.num_mics

    @property
    def num_samples(self) -> int:
        return self.signal.num_samples

    @property
    def sample_freq(self) -> float:
        return self.signal.sample_freq

    def _distances(self) -> np.ndarray:
        rm = self.env.apparent_r(self.loc, self.mics.pos).reshape(1, -1)
        if np.any(rm < 1e-7):
            import warnings

            warnings.warn('Source and microphone locations are identical.', Warning, stacklevel=3)
        return rm

    def _geometry(self):
        """Distances and start indices, computed on the host as the reference does (sources.py:804)."""
        rm = self._distances()
        ind = (-rm / self.env.c - self.start_t + self.start) * self.sample_freq * self.up
        return rm, ind

    def _host_inputs(self):
        """Upsampled signal (host copy of the device resampling), distances and start indices."""
        rm, ind = self._geometry()
        return np.ascontiguousarray(self.signal.usignal(self.up), dtype=np.float64), rm, ind

    def result_device(self, num: int = 128, host_inputs=None, raw_signal=None, geometry=None):
        """All microphone signals ``[T, num_mics]`` (float64, device), T as the block generator yields.
        ``raw_signal`` (a device copy of ``signal.signal()``) skips generating it here; ``geometry``
        (``_geometry()``'s result, computed earlier) skips computing it here."""
        from . import _lib
        from .signals import resample_device

        torch = _lib.torch()
        dev = _lib.device()
        if host_inputs is not None:
            signal, rm, ind = host_inputs
            sig = _lib.to_device(np.ascontiguousarray(signal, dtype=np.float64))
        else:
            rm, ind = geometry if geometry is not None else self._geometry()
            sig = None
        estimate = self.num_samples + (self.start_t - self.start) * self.sample_freq
        T = _block_plan(self.num_samples, estimate, int(num))
        M = self.num_channels
        out = torch.empty((T, M), dtype=torch.float64, device=dev)
        if T == 0:
            return out
        if sig is None:
            # the upsampled signal is generated and resampled on the device
            raw = raw_signal if raw_signal is not None else self.signal.signal_device()
            sig = resample_device(raw, self.up)
        L = int(sig.shape[0])
        ind0 = _lib.to_device(np.ascontiguousarray(ind[0]))
        rmd = _lib.to_device(np.ascontiguousarray(rm[0]))
        lib = _lib.lib()
        ws = _lib.workspace(lib.abf_point_source_workspace_size(M, T))
        _lib.check(lib.abf_point_source(_lib.ptr(sig), L, _lib.ptr(ind0), _lib.ptr(rmd), M, self.up,
                                        int(self.prepadding == 'zeros'), T, _lib.ptr(out), _lib.ptr(ws), ws.numel(),
                                        _lib.stream_handle()), 'PointSource')
        return out

    def result(self, num: int = 128):
        yield from _blocks(self.result_device(num), num)


class UncorrelatedNoiseSource(HasParams):
    """Independent noise at every microphone (``sources.py:1807-1967``): channel ``c`` is
    ``signal`` re-seeded with ``seed[c]`` (default ``signal.seed + c``).  White noise
    (``WNoiseGenerator``) is generated on the device by ``abf_white_noise``."""

    _params = ('signal', 'seed', 'mics', 'start_t', 'start')

    def __init__(self, **kwargs):
        object.__setattr__(self, 'signal', None)
        object.__setattr__(self, 'seed', np.array([], dtype=np.uint32))
        object.__setattr__(self, 'mics', None)
        object.__setattr__(self, 'start_t', 0.0)
        object.__setattr__(self, 'start', 0.0)
        super().__init__(**kwargs)

    def __setattr__(self, name, value):
        if name == 'seed':
            value = np.atleast_1d(np.asarray(value, dtype=np.uint32))
        object.__setattr__(self, name, value)

    @property
    def num_channels(self) -> int:
        return self.mics.num_mics

    @property
    def num_samples(self) -> int:
        return self.signal.num_samples

    @property
    def sample_freq(self) -> float:
        return self.signal.sample_freq

    def _seeds(self) -> np.ndarray:
        if not self.seed.size > 0:
            seed = np.arange(self.num_channels) + self.signal.seed
        elif self.seed.shape == (self.num_channels,):
            seed = self.seed
        else:
            msg = f'Seed array expected to be of shape ({self.num_channels:d},), but has shape {self.seed.shape}.'
            raise ValueError(msg)
        seed = np.asarray(seed, dtype=np.int64)
        if seed.min() < 0 or seed.max() > 2**32 - 1:
            raise ValueError('Seed must be between 0 and 2**32 - 1')
        return seed.astype(np.uint32)

    def result_device(self, num: int = 128):
        """All channels ``[num_samples, num_mics]`` (float64, device)."""
        from .signals import WNoiseGenerator, white_noise_device

        if type(self.signal) is not WNoiseGenerator:
            raise NotImplementedError('device noise generation supports WNoiseGenerator signals')
        return white_noise_device(self._seeds(), self.num_samples, self.signal.rms)

    def result(self, num: int = 128):
        yield from _blocks(self.result_device(num), num)


class SourceMixer(HasParams):
    """Weighted sum of sources with equal sample rate, channels and length (``sources.py:1969-2185``)."""

    _params = ('sources', 'weights')

    def _digest_items(self):
        # every source by its own digest: a changed source changes what depends on the mixer
        return (*self.sources, self.weights)

    def __init__(self, **kwargs):
        object.__setattr__(self, 'sources', [])
        object.__setattr__(self, 'weights', np.array([]))
        super().__init__(**kwargs)

    def __setattr__(self, name, value):
        if name == 'weights':
            value = np.atleast_1d(np.asarray(value, dtype=float))
        if name == 'sources':
            value = list(value)
        object.__setattr__(self, name, value)

    @property
    def sample_freq(self):
        return self.sources[0].sample_freq if self.sources else 0

    @property
    def num_channels(self):
        return self.sources[0].num_channels if self.sources else 0

    @property
    def num_samples(self):
        return self.sources[0].num_samples if self.sources else 0

    def validate_sources(self):
        """``sources.py:2116-2137``."""
        if len(self.sources) < 1:
            msg = 'Number of sources in SourceMixer should be at least 1.'
            raise ValueError(msg)
        for s in self.sources[1:]:
            if self.sample_freq != s.sample_freq:
                msg = f'Sample frequency of {s} does not fit'
                raise ValueError(msg)
            if self.num_channels != s.num_channels:
                msg = f'Channel count of {s} does not fit'
                raise ValueError(msg)
            if self.num_samples != s.num_samples:
                msg = f'Number of samples of {s} does not fit'
                raise ValueError(msg)

    def result_device(self, num: int = 128):
        """The weighted sum ``[T, num_channels]`` (float64, device), T the shortest source's length."""
        from . import _lib

        self.validate_sources()
        weights = self.weights.copy()
        if weights.size == 0:
            weights = np.array([1.0 for _ in range(len(self.sources))])
        if weights.shape[0] != len(self.sources):
            msg = f'weights must match the number of sources ({len(self.sources)})'
            raise ValueError(msg)
        parts = self._parts_device(num)
        T = min(p.shape[0] for p in parts)
        acc = _lib.torch().empty((T, self.num_channels), dtype=parts[0].dtype, device=parts[0].device)
        lib = _lib.lib()
        for j, p in enumerate(parts):
            src = p[:T].contiguous()
            _lib.check(lib.abf_mix(_lib.ptr(src), float(weights[j]), src.numel(), int(j == 0), _lib.ptr(acc),
                                   _lib.stream_handle()), 'SourceMixer')
        return acc

    def _parts_device(self, num):
        """Every source's signals on the device.  The white-noise streams of all sources (the point
        sources' signals and the uncorrelated noise channels) come from one ``abf_white_noise``
        launch, one workgroup per stream, each scaled by its rms afterwards (``abf_mix``, the same
        single product as generating it scaled)."""
        from . import _lib
        from .signals import WNoiseGenerator, white_noise_device

        torch = _lib.torch()
        # the point sources' distances come back to the host (their start indices are computed
        # there, as the reference does): done before any long kernel is queued, so that these
        # reads do not wait for it
        geo = {i: s._geometry() for i, s in enumerate(self.sources) if isinstance(s, PointSource)}
        streams = []  # (source index, seeds, rms)
        for i, s in enumerate(self.sources):
            if type(getattr(s, 'signal', None)) is not WNoiseGenerator:
                continue
            if isinstance(s, PointSource):
                streams.append((i, np.array([s.signal.seed], dtype=np.int64), s.signal.rms))
            elif isinstance(s, UncorrelatedNoiseSource):
                streams.append((i, s._seeds().astype(np.int64), s.signal.rms))
        pre = {}
        if len(streams) > 1:
            n = self.sources[streams[0][0]].num_samples
            base = white_noise_device(np.concatenate([sd for _, sd, _ in streams]), n, 1.0)
            lib = _lib.lib()
            c = 0
            for i, sd, rms in streams:
                block = base[:, c : c + sd.size].contiguous()
                scaled = torch.empty_like(block)
                _lib.check(lib.abf_mix(_lib.ptr(block), float(rms), block.numel(), 1, _lib.ptr(scaled),
                                       _lib.stream_handle()), 'white noise scaling')
                pre[i] = scaled
                c += sd.size
        parts = []
        for i, s in enumerate(self.sources):
            if i in pre and isinstance(s, PointSource):
                parts.append(s.result_device(num, raw_signal=pre[i].reshape(-1), geometry=geo[i]))
            elif i in pre:
                parts.append(pre[i])
            elif i in geo:
                parts.append(s.result_device(num, geometry=geo[i]))
            else:
                parts.append(s.result_device(num))
        return parts

    def result(self, num: int = 128):
        yield from _blocks(self.result_device(num), num)
