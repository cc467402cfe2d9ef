"""Block-wise frequency-domain processing (``acoular/fprocess.py``) on the GPU.

``RFFT`` transforms every block of the Welch block generator (``spectra.py:171-183``) with the
LDS FFT of the Welch CSM (``abf_rfft_blocks``: window, one-sided FFT, the scaling of
``RFFT._scale``), ``AutoPowerSpectra`` and ``CrossPowerSpectra`` form |X|^2 and X X^H per
snapshot (``abf_auto_power``, ``abf_cross_power``), and ``IRFFT`` turns every snapshot spectrum
back into a block of samples (``abf_irfft_blocks``).  Each stage keeps its whole result resident
on the device (``result_device``); ``result(num)`` yields it as numpy rows of ``precision`` in the
chunks the reference's generators yield.  The FFT runs in float32 on the device (the reference
transforms in float64 and then casts to ``precision``).
"""

from __future__ import annotations

import numpy as np
from scipy import fft

from . import _lib
from ._base import HasParams
from .spectra import OVERLAPS, WINDOWS


def _chunk_rows(num_blocks: int, num: int) -> list[int]:
    """Rows per block of ``RFFT.result(num)`` (``fprocess.py:168-180``): full blocks of ``num``,
    then the rest; with no spectrum at all and ``num > 1`` the reference still yields one row of
    zeros (its preallocated buffer)."""
    if num < 1:
        raise ValueError('num must be >= 1')
    if num_blocks == 0:
        return [1] if num > 1 else []
    rows = [num] * (num_blocks // num)
    if num_blocks % num:
        rows.append(num_blocks % num)
    return rows


def _split(dev, rows, precision):
    """Yield the device rows ``dev [B, ...]`` as numpy arrays of ``precision`` in ``rows`` chunks."""
    host = dev.cpu().numpy()
    b = 0
    for n in rows:
        chunk = host[b : b + n]
        if chunk.shape[0] < n:  # the zero row of an empty transform
            chunk = np.zeros((n, *host.shape[1:]), dtype=host.dtype)
        yield chunk.reshape(n, -1).astype(precision)
        b += n


class RFFT(HasParams):
    """One-sided FFT of multichannel time data, block by block (``fprocess.py:35-180``)."""

    _params = ('source', 'window', 'overlap', 'block_size', 'precision', 'scaling', 'workers')

    def __init__(self, **kwargs):
        object.__setattr__(self, 'source', None)
        object.__setattr__(self, 'window', 'Rectangular')
        object.__setattr__(self, 'overlap', 'None')
        object.__setattr__(self, 'block_size', 1024)
        object.__setattr__(self, 'precision', 'complex128')
        object.__setattr__(self, 'scaling', 'none')
        object.__setattr__(self, 'workers', None)
        super().__init__(**kwargs)

    def __setattr__(self, name, value):
        if name == 'window' and value not in WINDOWS:
            raise ValueError(f'window must be one of {list(WINDOWS)}')
        if name == 'overlap' and value not in OVERLAPS:
            raise ValueError(f'overlap must be one of {list(OVERLAPS)}')
        if name == 'block_size':
            value = int(value)
            if value % 2 != 0:
                raise ValueError('Block size must be even')
        if name == 'precision' and value not in ('complex128', 'complex64'):
            raise ValueError("precision must be 'complex128' or 'complex64'")
        if name == 'scaling' and value not in ('none', 'energy', 'amplitude'):
            raise ValueError("scaling must be 'none', 'energy' or 'amplitude'")
        super().__setattr__(name, value)

    def _digest_items(self):
        # workers only sets the host FFT's threads in the reference and is not part of its digest
        return tuple(getattr(self, p) for p in self._params if p != 'workers')

    @property
    def window_(self):
        return WINDOWS[self.window]

    @property
    def overlap_(self):
        return OVERLAPS[self.overlap]

    @property
    def sample_freq(self):
        return self.source.sample_freq

    @property
    def num_channels(self):
        return self.source.num_channels

    @property
    def num_freqs(self):
        return int(self.block_size / 2 + 1)

    @property
    def num_samples(self):
        """Snapshots as the reference counts them (``fprocess.py:91-95``): floor(N / block_size),
        whatever the overlap."""
        if self.source.num_samples >= 0:
            return int(np.floor(self.source.num_samples / self.block_size))
        return -1

    @property
    def freqs(self):
        if self.source is not None:
            return abs(fft.fftfreq(self.block_size, 1.0 / self.source.sample_freq)[: int(self.block_size / 2 + 1)])
        return np.array([])

    def bin_scale(self) -> np.ndarray:
        """The factor per bin of ``RFFT._scale`` (``fprocess.py:113-118, 160-164``)."""
        w = self.window_(self.block_size)
        if self.scaling == 'amplitude':
            sv = 1 / w.sum()
        else:
            sv = 1 / np.sqrt(np.dot(w, w) / self.block_size)
        s = np.full(self.num_freqs, sv)
        if self.scaling in ('amplitude', 'energy'):
            s[1:-1] *= 2.0
        return s

    def _time_device(self):
        src = self.source
        if getattr(src, 'num_samples', 0) == -1:
            raise ValueError('RFFT on the device requires a finite source')
        if hasattr(src, 'result_device'):
            return src.result_device(self.block_size).to(_lib.torch().float32).contiguous()
        data = getattr(src, 'data', None)
        if data is None:
            data = np.concatenate(list(src.result(self.block_size)), axis=0)
        return _lib.to_device(np.ascontiguousarray(data, dtype=np.float32))

    def _calc(self):
        torch = _lib.torch()
        bs = self.block_size
        if bs & (bs - 1) or bs > 16384:
            raise NotImplementedError('the device FFT takes power-of-two block sizes up to 16384')
        tdev = self._time_device()
        n, nch = int(tdev.shape[0]), int(tdev.shape[1])
        hop = bs // self.overlap_
        nb = (n - bs) // hop + 1 if n >= bs else 0
        out = torch.empty((nb, self.num_freqs, nch), dtype=torch.complex64, device=_lib.device())
        if nb == 0:
            return out
        lib = _lib.lib()
        wdev = _lib.to_device(self.window_(bs).astype(np.float32))
        sdev = _lib.to_device(self.bin_scale().astype(np.float32))
        ws = _lib.workspace(lib.abf_rfft_workspace_size(bs))
        _lib.check(lib.abf_rfft_blocks(_lib.ptr(tdev), n, nch, bs, hop, nb, _lib.ptr(wdev), _lib.ptr(sdev),
                                       _lib.ptr(out), _lib.ptr(ws), ws.numel(), _lib.stream_handle()), 'RFFT')
        return out

    def result_device(self):
        """Spectra of every block as a complex64 device tensor ``[blocks, num_freqs, channels]``."""
        return self._cached('spectra', self._calc)

    def result(self, num=1):
        """Rows ``(n, num_freqs * num_channels)`` of ``precision``, at most ``num`` per block
        (``fprocess.py:129-180``)."""
        dev = self.result_device()
        yield from _split(dev, _chunk_rows(int(dev.shape[0]), num), self.precision)


class AutoPowerSpectra(HasParams):
    """Real auto-power spectra |X|^2 of spectrum blocks (``fprocess.py:286-366``)."""

    _params = ('source', 'precision', 'scaling', 'single_sided')
    _precisions = ('float64', 'float32')

    def __init__(self, **kwargs):
        object.__setattr__(self, 'source', None)
        object.__setattr__(self, 'scaling', 'power')
        object.__setattr__(self, 'single_sided', True)
        object.__setattr__(self, 'precision', self._precisions[0])
        super().__init__(**kwargs)

    def __setattr__(self, name, value):
        if name == 'scaling' and value not in ('power', 'psd'):
            raise ValueError("scaling must be 'power' or 'psd'")
        if name == 'precision' and value not in self._precisions:
            raise ValueError(f'precision must be one of {self._precisions}')
        if name == 'single_sided':
            value = bool(value)
        super().__setattr__(name, value)

    @property
    def block_size(self):
        return self.source.block_size

    @property
    def num_freqs(self):
        return self.source.num_freqs

    @property
    def sample_freq(self):
        return self.source.sample_freq

    @property
    def num_samples(self):
        return self.source.num_samples

    @property
    def num_channels(self):
        return self.source.num_channels

    @property
    def freqs(self):
        return self.source.freqs

    def _get_scaling_value(self):
        """``AutoPowerSpectra._get_scaling_value`` (``fprocess.py:320-326``)."""
        scale = 1 / self.block_size**2
        if self.single_sided:
            scale *= 2
        if self.scaling == 'psd':
            scale *= self.block_size * self.source.sample_freq
        return scale

    def _spectra(self):
        if not hasattr(self.source, 'result_device'):
            raise TypeError('the source must be an RFFT (device spectra)')
        return self.source.result_device()

    def _calc(self):
        torch = _lib.torch()
        spec = self._spectra()
        out = torch.empty(spec.shape, dtype=torch.float32, device=spec.device)
        _lib.check(_lib.lib().abf_auto_power(_lib.ptr(spec), spec.numel(), float(self._get_scaling_value()),
                                             _lib.ptr(out), _lib.stream_handle()), 'AutoPowerSpectra')
        return out

    def result_device(self):
        """Auto-power spectra of every snapshot, float32 device tensor ``[blocks, num_freqs, channels]``."""
        return self._cached('auto', self._calc)

    def result(self, num=1):
        dev = self.result_device()
        yield from _split(dev, _chunk_rows(int(dev.shape[0]), num), self.precision)


class CrossPowerSpectra(AutoPowerSpectra):
    """Auto- and cross-power spectra X X^H per snapshot (``fprocess.py:369-470``)."""

    _params = AutoPowerSpectra._params + ('calc_mode',)
    _precisions = ('complex128', 'complex64')

    def __init__(self, **kwargs):
        object.__setattr__(self, 'calc_mode', 'full')
        super().__init__(**kwargs)

    def __setattr__(self, name, value):
        if name == 'calc_mode' and value not in ('full', 'upper', 'lower'):
            raise ValueError("calc_mode must be 'full', 'upper' or 'lower'")
        super().__setattr__(name, value)

    @property
    def num_channels(self):
        n = self.source.num_channels
        return n**2 if self.calc_mode == 'full' else int(n + n * (n - 1) / 2)

    def _calc(self):
        torch = _lib.torch()
        spec = self._spectra()
        nb, nf, nc = (int(x) for x in spec.shape)
        if self.calc_mode != 'full' and nc > 1:
            # the reference writes the M x M matrix of each bin into M (M + 1) / 2 slots per bin
            # (fprocess.py:460-464), which numpy rejects
            msg = (f'could not broadcast input array from shape ({nf * nc * nc},) into shape '
                   f'({nf * self.num_channels},)')
            raise ValueError(msg)
        out = torch.empty((nb, nf, nc, nc), dtype=torch.complex64, device=spec.device)
        _lib.check(_lib.lib().abf_cross_power(_lib.ptr(spec), nb, nf, nc, float(self._get_scaling_value()),
                                              _lib.ptr(out), _lib.stream_handle()), 'CrossPowerSpectra')
        return out

    def result_device(self):
        """X X^H of every snapshot and bin, complex64 device tensor ``[blocks, num_freqs, M, M]``."""
        return self._cached('cross', self._calc)


class IRFFT(HasParams):
    """Real inverse FFT of one-sided spectrum snapshots back to time blocks (``fprocess.py:182-283``)."""

    _params = ('source', 'precision', 'workers')

    def __init__(self, **kwargs):
        object.__setattr__(self, 'source', None)
        object.__setattr__(self, 'precision', 'float64')
        object.__setattr__(self, 'workers', None)
        super().__init__(**kwargs)

    def __setattr__(self, name, value):
        if name == 'precision' and value not in ('float64', 'float32'):
            raise ValueError("precision must be 'float64' or 'float32'")
        super().__setattr__(name, value)

    def _digest_items(self):
        # workers only sets the host FFT's threads in the reference and is not part of its digest
        return (self.source, self.precision)

    @property
    def sample_freq(self):
        return self.source.sample_freq

    @property
    def num_channels(self):
        return self.source.num_channels

    @property
    def num_samples(self):
        """``IRFFT._get_num_samples`` (``fprocess.py:206-209``): snapshots times block size."""
        if self.source.num_samples >= 0:
            return int(self.source.num_samples * self.source.block_size)
        return -1

    def _validate(self):
        """``IRFFT._validate`` (``fprocess.py:215-233``)."""
        bs = self.source.block_size
        if not bs or bs < 0:
            msg = (f'Source of class {type(self).__name__} has an unknown blocksize: {bs}.'
                   'This is likely due to incomplete spectral data from which the inverse FFT cannot be calculated.')
            raise ValueError(msg)
        if (self.source.num_freqs - 1) * 2 != bs:
            msg = (f'Block size must be 2*(num_freqs-1) but is {bs}.'
                   'This is likely due to incomplete spectral data from which the inverse FFT cannot be calculated.')
            raise ValueError(msg)
        if bs % 2 != 0:
            msg = f'Block size must be even but is {bs}.'
            raise ValueError(msg)

    def _calc(self):
        torch = _lib.torch()
        self._validate()
        if not hasattr(self.source, 'result_device'):
            raise TypeError('the source must be an RFFT (device spectra)')
        spec = self.source.result_device().to(torch.complex64).contiguous()
        nb, nf, nc = (int(x) for x in spec.shape)
        bs = self.source.block_size
        if bs & (bs - 1) or bs > 16384:
            raise NotImplementedError('the device FFT takes power-of-two block sizes up to 16384')
        out = torch.empty((nb * bs, nc), dtype=torch.float32, device=spec.device)
        if nb == 0:
            return out
        lib = _lib.lib()
        ws = _lib.workspace(lib.abf_irfft_workspace_size(bs))
        _lib.check(lib.abf_irfft_blocks(_lib.ptr(spec), nb, nc, bs, _lib.ptr(out), _lib.ptr(ws), ws.numel(),
                                        _lib.stream_handle()), 'IRFFT')
        return out

    def result_device(self):
        """The samples of every snapshot, float32 device tensor ``[blocks * block_size, channels]``."""
        return self._cached('time', self._calc)

    def result(self, num):
        """Blocks of ``num`` samples ``(num, num_channels)`` of ``precision``, the last one shorter
        (``fprocess.py:235-283``: one block per snapshot when ``num`` equals the block size,
        re-blocked through a buffer otherwise -- the same sample sequence either way)."""
        if num < 1:
            raise ValueError('num must be >= 1')
        host = self.result_device().cpu().numpy().astype(self.precision)
        for b in range(0, host.shape[0], num):
            yield host[b : b + num]
