"""Cross-spectral matrix estimation (``acoular/spectra.py``), computed on the GPU.

``PowerSpectra.csm`` runs the Welch estimate of ``PowerSpectra.calc_csm``
(``spectra.py:366-416``) as gfx950 kernels (windowed FFT per block, CSM update on fp32 MFMA,
Hermitian completion and scaling) through ``abf_welch_csm``.  The CSM stays resident on the
device (``csm_device``, complex64) for the beamformer; ``csm`` returns it as a numpy array of
``precision``.
"""

from __future__ import annotations

import warnings

import numpy as np
from scipy import fft

from . import _lib
from ._base import HasParams

WINDOWS = {
    'Rectangular': np.ones,
    'Hanning': np.hanning,
    'Hamming': np.hamming,
    'Bartlett': np.bartlett,
    'Blackman': np.blackman,
}
OVERLAPS = {'None': 1, '50%': 2, '75%': 4, '87.5%': 8}
BLOCK_SIZES = (1024, 128, 256, 512, 2048, 4096, 8192, 16384, 32768, 65536)
PRECISIONS = ('complex128', 'complex64')
# spectra buffer of one Welch pass (frequencies x blocks x channels complex64)
_PASS_BYTES = 512 << 20


class BaseSpectra(HasParams):
    """Shared FFT parameters (``spectra.py:52-183``)."""

    _params = ('source', 'window', 'overlap', 'block_size', 'precision')

    def __init__(self, **kwargs):
        self.source = None
        self.window = 'Rectangular'
        self.overlap = 'None'
        self.block_size = 1024
        self.precision = 'complex128'
        super().__init__(**kwargs)

    def __setattr__(self, name, value):
        if name == 'window' and value is not None and value not in WINDOWS:
            raise ValueError(f'window must be one of {list(WINDOWS)}')
        if name == 'overlap' and value is not None and value not in OVERLAPS:
            raise ValueError(f'overlap must be one of {list(OVERLAPS)}')
        if name == 'block_size' and value is not None and value not in BLOCK_SIZES:
            raise ValueError(f'block_size must be one of {BLOCK_SIZES}')
        if name == 'precision' and value not in PRECISIONS:
            raise ValueError(f'precision must be one of {PRECISIONS}')
        object.__setattr__(self, name, value)

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

    def fftfreq(self):
        """``BaseSpectra.fftfreq`` (``spectra.py:128-168``)."""
        if self.source is not None:
            return abs(fft.fftfreq(self.block_size, 1.0 / self.source.sample_freq)[: int(self.block_size / 2 + 1)])
        return None


class PowerSpectra(BaseSpectra):
    """Welch CSM of multichannel time data (``spectra.py:187-588``).

    Precision: the GPU path transforms and accumulates in complex64 (fp32 FFT, fp32 MFMA CSM
    update) whatever ``precision`` says; ``precision='complex128'`` (the default, as in
    ``spectra.py:119``) returns that CSM widened to complex128.  Against the float64 Welch
    estimate the result stays within 1e-5 of the CSM's largest element and within 1e-5 relative
    on every auto-spectrum (64 microphones, 1 s at 51.2 kHz, block size 128, Hanning, 50 %),
    and it is exactly Hermitian.
    """

    _params = ('source', 'window', 'overlap', 'block_size', 'precision')

    def __init__(self, **kwargs):
        self.cached = True  # accepted for API compatibility; results stay resident in HBM instead
        self._ind_low = 1
        self._ind_high = -1
        ind_low = kwargs.pop('ind_low', None)
        ind_high = kwargs.pop('ind_high', None)
        cached = kwargs.pop('cached', True)
        super().__init__(**kwargs)
        self.cached = cached
        if ind_low is not None:
            self._ind_low = ind_low
        if ind_high is not None:
            self._ind_high = ind_high

    @property
    def num_blocks(self):
        """``overlap * N / block_size - overlap + 1`` -- a float (``spectra.py:281-283``)."""
        return self.overlap_ * self.source.num_samples / self.block_size - self.overlap_ + 1

    @property
    def ind_low(self):
        f = self.fftfreq()
        if f is not None:
            return min(self._ind_low, f.shape[0] - 1)
        return 0

    @ind_low.setter
    def ind_low(self, v):
        self._ind_low = v

    @property
    def ind_high(self):
        f = self.fftfreq()
        if f is not None:
            if self._ind_high is None:
                return None
            return min(self._ind_high, f.shape[0] - 1)
        return None

    @ind_high.setter
    def ind_high(self, v):
        self._ind_high = v

    @property
    def indices(self):
        f = self.fftfreq()
        if f is None:
            return None
        idx = np.arange(f.shape[0], dtype=int)
        if self.ind_high is None:
            return idx[self.ind_low :]
        return idx[self.ind_low : self.ind_high]

    def _digest_items(self):
        return (self.source, self.window, self.overlap, self.block_size, self.precision)

    def block_starts(self) -> np.ndarray:
        """First sample of every Welch block the reference generator yields
        (``spectra.py:171-183``): hop ``block_size/overlap``, blocks ending inside the data."""
        n, bs = self.source.num_samples, self.block_size
        if n < bs:
            return np.zeros(0, dtype=np.int64)
        return np.arange(0, n - bs + 1, bs // self.overlap_, dtype=np.int64)

    def _time_data(self):
        src = self.source
        if getattr(src, 'num_samples', 0) == -1:
            msg = 'PowerSpectra requires a finite source'
            raise ValueError(msg)
        data = getattr(src, 'data', None)
        if data is None:
            data = np.concatenate(list(src.result(self.block_size)), axis=0)
        return data

    def _calc_csm_device(self):
        torch = _lib.torch()
        bs = self.block_size
        if hasattr(self.source, 'result_device'):
            # sources synthesised on the GPU hand their samples over in HBM
            tdev = self.source.result_device(bs).to(torch.float32).contiguous()
        else:
            tdev = _lib.to_device(np.ascontiguousarray(self._time_data(), dtype=np.float32))
        n, nch = int(tdev.shape[0]), int(tdev.shape[1])
        nb = max(0, (n - bs) // (bs // self.overlap_) + 1) if n >= bs else 0
        wind = self.window_(bs)
        weight = float(np.dot(wind, wind))
        scale = 2.0 / self.block_size / weight / self.num_blocks  # ZeroDivisionError as the reference
        nf = bs // 2 + 1
        csm = torch.empty((nf, nch, nch), dtype=torch.complex64, device=_lib.device())
        wdev = _lib.to_device(wind.astype(np.float32))
        bpp = int(max(1, min(max(nb, 1), _PASS_BYTES // max(1, nf * nch * 8))))
        lib = _lib.lib()
        ws = _lib.workspace(lib.abf_welch_csm_workspace_size(nch, bs, bpp))
        _lib.check(
            lib.abf_welch_csm(_lib.ptr(tdev), n, nch, bs, bs // self.overlap_, nb, _lib.ptr(wdev),
                              float(scale), _lib.ptr(csm), bpp, _lib.ptr(ws), ws.numel(), _lib.stream_handle()),
            'PowerSpectra.csm')
        return csm

    @property
    def csm_device(self):
        """The CSM as a resident complex64 device tensor ``[F, M, M]``."""
        return self._cached('csm_device', self._calc_csm_device)

    def csm_device_f64(self):
        """The CSM as a complex128 device tensor ``[F, M, M]`` (input of the float64 beamformer)."""
        torch = _lib.torch()
        return self._cached('csm_device_f64', lambda: self.csm_device.to(torch.complex128))

    def calc_csm(self):
        """``PowerSpectra.calc_csm`` (``spectra.py:366-416``) as a numpy array of ``precision``."""
        return self._calc_csm_device().cpu().numpy().astype(self.precision)

    @property
    def csm(self):
        return self._cached('csm', lambda: self.csm_device.cpu().numpy().astype(self.precision))

    def _eig_input(self):
        # the reference decomposes the CSM in its precision (spectra.py:356-362)
        torch = _lib.torch()
        return self.csm_device.to(torch.complex128 if self.precision == 'complex128' else torch.complex64)

    def _evd_native(self):
        # one batched abf_csm_eig call over every slice, in the precision of _eig_input
        torch = _lib.torch()
        c = self._eig_input().contiguous()
        nf, nm = int(c.shape[0]), int(c.shape[1])
        is128 = int(c.dtype == torch.complex128)
        ev = torch.empty((nf, nm), dtype=torch.float64, device=c.device)
        evec = torch.empty_like(c)
        info = torch.zeros(nf, dtype=torch.int32, device=c.device)
        lib = _lib.lib()
        ws = _lib.workspace(lib.abf_csm_eig_workspace_size(nf, nm, is128))
        _lib.check(lib.abf_csm_eig(_lib.ptr(c), nf, nm, is128, _lib.ptr(ev), _lib.ptr(evec), _lib.ptr(info),
                                   _lib.ptr(ws), ws.numel(), _lib.stream_handle()), 'CSM eigendecomposition')
        bad = torch.nonzero(info).flatten().cpu().numpy()
        if bad.size:
            msg = f'eigendecomposition did not converge for frequency indices {bad.tolist()}'
            raise np.linalg.LinAlgError(msg)
        return ev, evec

    def eig_device(self):
        """Eigendecomposition of every CSM slice, resident on the device (``spectra.py:351-364``):
        eigenvalues float64 ``[F, M]`` in ascending order and eigenvectors complex64 ``[F, M, M]``
        with eigenvector ``e`` in column ``e`` (the layout of ``numpy.linalg.eigh``), computed by
        one batched ``abf_csm_eig`` call in the CSM's precision."""
        torch = _lib.torch()

        def calc():
            ev, evec = self._cached('evd_native', self._evd_native)
            return ev, evec.to(torch.complex64).contiguous()

        return self._cached('evd_device', calc)

    def _evd(self):
        ev, evec = self._cached('evd_native', self._evd_native)
        rdt = 'float64' if self.precision == 'complex128' else 'float32'
        return ev.cpu().numpy().astype(rdt), evec.cpu().numpy().astype(self.precision)

    @property
    def eigvals(self):
        return self._cached('evd', self._evd)[0]

    @property
    def eigvecs(self):
        return self._cached('evd', self._evd)[1]

    def synthetic_ev(self, freq, num=0):
        """``PowerSpectra.synthetic_ev`` (``spectra.py:528-588``)."""
        f = self.fftfreq()
        if num == 0:
            return self.eigvals[np.searchsorted(f, freq)]
        f1 = np.searchsorted(f, freq * 2.0 ** (-0.5 / num))
        f2 = np.searchsorted(f, freq * 2.0 ** (0.5 / num))
        if f1 == f2:
            return self.eigvals[f1]
        return np.sum(self.eigvals[f1:f2], 0)


class PowerSpectraImport(PowerSpectra):
    """Injects a precomputed CSM and its frequencies (``spectra.py:591-697``)."""

    _params = ('csm_array', 'frequencies', 'precision')

    def __init__(self, csm=None, frequencies=None, **kwargs):
        object.__setattr__(self, 'csm_array', None)
        object.__setattr__(self, '_csm_dev', None)
        # inherited trait (spectra.py:119): the dtype of the host eigendecomposition
        object.__setattr__(self, 'precision', 'complex128')
        self.frequencies = frequencies
        HasParams.__init__(self, **kwargs)
        self.cached = False
        self._ind_low = 0
        self._ind_high = None
        self.source = None
        self.block_size = None
        self.window = None
        self.overlap = None
        if csm is not None:
            self.csm = csm

    def __setattr__(self, name, value):
        if name in ('window', 'overlap', 'block_size') and value is not None:
            return  # non-functional in this class, as in the reference
        if name == 'precision' and value not in PRECISIONS:
            raise ValueError(f'precision must be one of {PRECISIONS}')
        object.__setattr__(self, name, value)

    def _digest_items(self):
        return (id(self.csm_array), self.frequencies if not isinstance(self.frequencies, np.ndarray) else
                tuple(np.ravel(self.frequencies)), self.precision)

    @property
    def csm(self):
        return self.csm_array

    @csm.setter
    def csm(self, value):
        if value is None:
            return
        is_tensor = not isinstance(value, np.ndarray) and hasattr(value, 'is_cuda')
        shape = tuple(value.shape)
        if len(shape) != 3 or shape[1] != shape[2]:
            msg = ('The cross spectral matrix must have the following shape: '
                   '(number of frequencies, num_channels, num_channels)!')
            raise ValueError(msg)
        if is_tensor:
            object.__setattr__(self, '_csm_dev', value)
            object.__setattr__(self, 'csm_array', value)
        else:
            object.__setattr__(self, 'csm_array', np.asarray(value))
            object.__setattr__(self, '_csm_dev', None)
        self._cache.clear()

    @property
    def csm_device(self):
        if self._csm_dev is None:
            torch = _lib.torch()
            object.__setattr__(self, '_csm_dev', _lib.to_device(np.asarray(self.csm_array), torch.complex64))
        dev = self._csm_dev
        torch = _lib.torch()
        if dev.dtype != torch.complex64:
            dev = dev.to(torch.complex64)
        return dev

    def csm_device_f64(self):
        torch = _lib.torch()

        def calc():
            if self._csm_dev is not None:
                return self._csm_dev.to(torch.complex128)
            return _lib.to_device(np.asarray(self.csm_array), torch.complex128)

        return self._cached('csm_device_f64', calc)

    def _eig_input(self):
        # np.linalg.eigh of the imported array runs in that array's own precision
        torch = _lib.torch()
        if self._csm_dev is not None:
            dt = torch.complex128 if self._csm_dev.dtype == torch.complex128 else torch.complex64
            return self._csm_dev.to(dt)
        host = np.asarray(self.csm_array)
        dt = torch.complex128 if host.dtype in (np.complex128, np.float64) else torch.complex64
        return _lib.to_device(host, dt)

    @property
    def num_channels(self):
        return self.csm_array.shape[1]

    @property
    def num_blocks(self):
        return None

    @property
    def sample_freq(self):
        return None

    def fftfreq(self):
        """The injected frequencies (``spectra.py:681-697``)."""
        if isinstance(self.frequencies, float):
            return np.array([self.frequencies])
        if isinstance(self.frequencies, (int, np.integer)) and not isinstance(self.frequencies, bool):
            return np.array([float(self.frequencies)])
        if isinstance(self.frequencies, (list, tuple)):
            return np.asarray(self.frequencies, dtype=float)
        if self.frequencies is None:
            warnings.warn('No frequencies set for PowerSpectraImport.', Warning, stacklevel=2)
        return self.frequencies

    @property
    def csm_host(self):
        return self.csm_array
