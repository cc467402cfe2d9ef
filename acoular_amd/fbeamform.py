"""Frequency-domain beamforming (``acoular/fbeamform.py``): ``SteeringVector``,
``BeamformerBase`` with lazy per-frequency results, ``L_p`` and ``integrate``.

``BeamformerBase._calc`` batches every missing frequency of a request into one GPU job
(``abf_beamform``): the CSM stays resident on the device, steering vectors are generated inside
the kernels from grid and microphone coordinates, and only the map comes back to the host.
"""

from __future__ import annotations

from warnings import warn as _warn

import numpy as np

from . import _lib
from ._base import HasParams
from .environments import Environment
from .fastFuncs import beamformerFreq, calcTransfer
from .grids import Grid
from .microphones import MicGeom
from .spectra import PowerSpectra

STEER_TYPES = ('true level', 'true location', 'classic', 'inverse')

#: defaults of the GPU path (not reference traits)
# autotuned: MFMA jobs take their launch shape from the autotuner's cache when it holds their
# shape class (acoular_amd.autotune)
config = {'form': 'mfma_bf16x3', 'phase_mode': 'reference', 'fast_trig': True, 'autotuned': True}


class SteeringVector(HasParams):
    """Monopole steering vectors of the four formulations (``fbeamform.py:105-258``)."""

    _params = ('grid', 'mics', 'steer_type', 'env', 'ref')

    def __init__(self, **kwargs):
        self.grid = None
        self.mics = None
        self.steer_type = 'true level'
        self.env = Environment()
        self.ref = np.array([0.0, 0.0, 0.0])
        super().__init__(**kwargs)

    def __setattr__(self, name, value):
        if name == 'steer_type' and value not in STEER_TYPES:
            raise ValueError(f'steer_type must be one of {STEER_TYPES}')
        if name == 'grid' and value is not None and not isinstance(value, Grid):
            raise TypeError('grid must be a Grid')
        if name == 'mics' and value is not None and not isinstance(value, MicGeom):
            raise TypeError('mics must be a MicGeom')
        if name == 'ref':
            if np.isscalar(value):
                try:
                    value = np.abs(float(value))
                except ValueError as ve:
                    raise TypeError('ref must be a float or a 3-vector') from ve
            elif len(value) == 3:
                value = np.array(value, dtype=float)
            else:
                raise TypeError('ref must be a float or a 3-vector')
        object.__setattr__(self, name, value)

    @property
    def r0(self):
        """Distances of the grid points to the reference (``fbeamform.py:187-193``)."""

        def calc():
            if np.isscalar(self.ref):
                if self.ref > 0:
                    return np.full((self.grid.size,), self.ref)
                return self.env.apparent_r(self.grid.pos)
            return self.env.apparent_r(self.grid.pos, self.ref[:, np.newaxis])

        return self._cached('r0', calc)

    @property
    def rm(self):
        """Distances of the grid points to the microphones (``fbeamform.py:195-197``)."""
        return self._cached('rm', lambda: np.atleast_2d(self.env.apparent_r(self.grid.pos, self.mics.pos)))

    def device_geometry(self):
        """Grid and microphone coordinates resident on the device (float64, SoA)."""

        def calc():
            return (_lib.to_device(np.ascontiguousarray(self.grid.pos, dtype=np.float64)),
                    _lib.to_device(np.ascontiguousarray(self.mics.pos, dtype=np.float64)))

        return self._cached('dev_geom', calc)

    def _k(self, f):
        return 2 * np.pi * f / self.env.c

    def transfer(self, f, ind=None):
        """Transfer matrix for one frequency (``fbeamform.py:207-235``), complex128."""
        if ind is None:
            return calcTransfer(self.r0, self.rm, np.array(self._k(f)))
        if not isinstance(ind, np.ndarray):
            return calcTransfer(self.r0[ind], self.rm[ind, :][np.newaxis], np.array(self._k(f)))
        return calcTransfer(self.r0[ind], self.rm[ind, :], np.array(self._k(f)))

    def steer_vector(self, f, ind=None):
        """Steering vectors (``fbeamform.py:237-258``): the transfer function normalised per
        ``steer_type`` (``fbeamform.py:145-153``), computed on the GPU."""
        from .fastFuncs import _transfer_device

        mode = {'classic': 1, 'inverse': 2, 'true level': 3, 'true location': 4}[self.steer_type]
        if ind is None:
            r0, rm = self.r0, self.rm
        elif not isinstance(ind, np.ndarray):
            r0, rm = self.r0[ind], self.rm[ind, :][np.newaxis]
        else:
            r0, rm = self.r0[ind], self.rm[ind, :]
        return _transfer_device(np.atleast_1d(r0), rm, self._k(f), mode)


class LazyBfResult:
    """Per-frequency lazy evaluation of ``BeamformerBase.result`` (``fbeamform.py:261-286``)."""

    def __init__(self, bf):
        self.bf = bf

    def __getitem__(self, key):
        sl = np.index_exp[key][0]
        if isinstance(sl, (int, np.integer)):
            sl = slice(sl, sl + 1)
        missingind = np.arange(*sl.indices(self.bf._numfreq))[self.bf._fr[sl] == 0]
        if missingind.size:
            self.bf._calc(missingind)
        return self.bf._ac.__getitem__(key)


class BeamformerBase(HasParams):
    """Conventional (delay-and-sum) beamformer in the frequency domain (``fbeamform.py:289-685``).

    Parameters beyond the reference's: ``form`` ('mfma_bf16x3', 'mfma_f16x2', 'mfma_f32' or 'plain_f32') and
    ``phase_mode`` ('reference': the float32-rounded phase of ``fastFuncs.py:260``, or 'fp64').
    """

    _params = ('freq_data', 'steer', 'r_diag', 'r_diag_norm', 'precision', 'form', 'phase_mode')
    #: parameters a subclass pins to one value (the reference declares them ``Enum(value)``)
    _fixed: dict = {}

    def __init__(self, **kwargs):
        self.freq_data = None
        self.steer = SteeringVector()
        self.r_diag = self._fixed.get('r_diag', True)
        self.r_diag_norm = self._fixed.get('r_diag_norm', 0.0)
        self.precision = 'float64'
        self.cached = True  # accepted for API compatibility (no HDF5 cache in this build)
        self.form = config['form']
        self.phase_mode = config['phase_mode']
        cached = kwargs.pop('cached', True)
        super().__init__(**kwargs)
        self.cached = cached
        object.__setattr__(self, '_result_digest', None)

    def __setattr__(self, name, value):
        if name in self._fixed and value != self._fixed[name]:
            msg = f'{type(self).__name__}.{name} can only be {self._fixed[name]!r}'
            raise ValueError(msg)
        if name == 'precision' and value not in ('float64', 'float32'):
            raise ValueError("precision must be 'float64' or 'float32'")
        if name == 'form' and value not in _lib.FORMS:
            raise ValueError(f'form must be one of {list(_lib.FORMS)}')
        if name == 'phase_mode' and value not in _lib.PHASE_MODES:
            raise ValueError(f'phase_mode must be one of {list(_lib.PHASE_MODES)}')
        if name == 'freq_data' and value is not None and not isinstance(value, PowerSpectra):
            raise TypeError('freq_data must be a PowerSpectra')
        if name == 'steer' and value is not None and not isinstance(value, SteeringVector):
            raise TypeError('steer must be a SteeringVector')
        object.__setattr__(self, name, value)

    # -- reference API ----------------------------------------------------------------------
    def _assert_equal_channels(self):
        num_channels = self.freq_data.num_channels
        if num_channels != self.steer.mics.num_mics or num_channels == 0:
            msg = f'{num_channels:d} channels do not fit {self.steer.mics.num_mics:d} mics'
            raise ValueError(msg)

    @property
    def result(self):
        """Lazy ``[F, G]`` map (``fbeamform.py:403-436``); recomputed when parameters change."""
        d = self.digest
        if self._result_digest != d:
            self._f = self.freq_data.fftfreq()
            self._numfreq = self._f.shape[0]
            self._assert_equal_channels()
            self._ac = np.zeros((self._numfreq, self.steer.grid.size), dtype=self.precision)
            self._fr = np.zeros(self._numfreq, dtype='int8')
            object.__setattr__(self, '_result_digest', d)
        return LazyBfResult(self)

    def sig_loss_norm(self):
        """Normalisation for the removed CSM diagonal (``fbeamform.py:438-451``)."""
        if not self.r_diag:
            return 1.0
        if self.r_diag_norm == 0.0:
            nm = float(self.freq_data.num_channels)
            return nm / (nm - 1)
        return self.r_diag_norm

    def _beamformer_params(self):
        """``fbeamform.py:453-473``: the fast path only for exactly ``SteeringVector``."""
        if type(self.steer) is SteeringVector:
            param_type = self.steer.steer_type

            def param_steer_func(f):
                return (self.steer.r0, self.steer.rm, 2 * np.pi * f / self.steer.env.c)

        else:
            param_type = 'custom'
            param_steer_func = self.steer.steer_vector
        return param_type, param_steer_func

    def _calc(self, ind):
        """Compute the frequencies ``ind`` (``fbeamform.py:475-510``) in one GPU job."""
        ind = np.asarray(ind, dtype=np.int64)
        normfactor = self.sig_loss_norm()
        param_type, steer_func = self._beamformer_params()
        if param_type == 'custom':
            for i in ind:
                csm = np.array(self.freq_data.csm[i], dtype='complex128')
                out = beamformerFreq('custom', self.r_diag, normfactor, steer_func(self._f[i]), csm)[0]
                if self.r_diag:
                    out[np.sign(out) < 0] = 0.0
                self._ac[i] = out
                self._fr[i] = 1
            return
        maps = beamform_device(self.freq_data.csm_device, self._f[ind], ind, self.steer, param_type, self.r_diag,
                               normfactor, self.form, self.phase_mode, clip=self.r_diag)
        self._ac[ind] = maps.cpu().numpy().astype(self.precision)
        self._fr[ind] = 1

    def synthetic(self, f, num=0):
        """Band-summed map at ``f`` (``fbeamform.py:512-593``), reshaped to ``grid.shape``."""
        res = self.result
        freq = self.freq_data.fftfreq()
        if len(freq) == 0:
            return None
        if num == 0:
            ind = np.searchsorted(freq, f)
            if ind >= len(freq):
                _warn(f'Queried frequency ({f:g} Hz) not in resolved frequency range. Returning zeros.', Warning,
                      stacklevel=2)
                h = np.zeros_like(res[0])
            else:
                if freq[ind] != f:
                    _warn(f'Queried frequency ({f:g} Hz) not in set of discrete FFT sample frequencies. '
                          f'Using frequency {freq[ind]:g} Hz instead.', Warning, stacklevel=2)
                h = res[ind]
        else:
            if isinstance(num, list):
                f1, f2 = num[0], num[-1]
            else:
                f1 = f * 2.0 ** (-0.5 / num)
                f2 = f * 2.0 ** (+0.5 / num)
            ind1 = np.searchsorted(freq, f1)
            ind2 = np.searchsorted(freq, f2)
            if ind1 == ind2:
                _warn(f'Queried frequency band ({f1:g} to {f2:g} Hz) does not include any discrete FFT sample '
                      'frequencies. Returning zeros.', Warning, stacklevel=2)
                h = np.zeros_like(res[0])
            else:
                h = np.sum(res[ind1:ind2], 0)
        return h.reshape(self.steer.grid.shape)

    def integrate(self, sector, frange=None, num=0):
        """Sector-integrated spectrum (``fbeamform.py:595-685``): ``sector`` is a ``Sector`` or the
        arguments of the grid's ``indices``."""
        ind = _sector_indices(self.steer.grid, sector)
        gshape = self.steer.grid.shape
        num_freqs = self.freq_data.fftfreq().shape[0]
        if num == 0 or frange is None:
            if frange is None:
                ind_low = self.freq_data.ind_low
                ind_high = self.freq_data.ind_high
                if ind_low is None:
                    ind_low = 0
                if ind_high is None:
                    ind_high = num_freqs
                irange = (ind_low, ind_high)
                num = 0
            elif len(frange) == 2:
                _ = self.result
                irange = (np.searchsorted(self._f, frange[0]), np.searchsorted(self._f, frange[1]))
            else:
                msg = 'Only a tuple of length 2 is allowed for frange if num==0'
                raise TypeError(msg)
            h = np.zeros(num_freqs, dtype=float)
            sl = slice(*irange)
            r = self.result[sl]
            for i in range(num_freqs)[sl]:
                h[i] = r[i - sl.start].reshape(gshape)[ind].sum()
            if frange is None:
                return h
            return self._f[sl], h[sl]
        h = np.zeros(len(frange), dtype=float)
        for i, f in enumerate(frange):
            h[i] = self.synthetic(f, num).reshape(gshape)[ind].sum()
        return h


class BeamformerCapon(BeamformerBase):
    """Capon (minimum variance) beamforming (``fbeamform.py:777-818``): ``1 / (h^H C^-1 h)`` with
    normalisation M^2.  The reference inverts every CSM slice and runs the full-CSM form on the
    inverse; here the inverse enters through the resident eigendecomposition, h^H C^-1 h =
    sum_k |v_k^H h|^2 / l_k, one ``abf_beamform_eig_grid`` launch over all eigenpairs (fp32 MFMA)
    for all requested bins.  Only defined for the full CSM.
    """

    _fixed = {'r_diag': False, 'r_diag_norm': 1.0}

    def _calc(self, ind):
        ind = np.asarray(ind, dtype=np.int64)
        nm = self.freq_data.num_channels
        normfactor = self.sig_loss_norm() * nm**2
        if type(self.steer) is not SteeringVector:
            import scipy.linalg as spla

            _, steer_func = self._beamformer_params()
            for i in ind:
                csm = np.array(spla.inv(np.array(self.freq_data.csm[i], dtype='complex128')), order='C')
                self._ac[i] = 1.0 / beamformerFreq('custom', self.r_diag, normfactor, steer_func(self._f[i]), csm)[0]
                self._fr[i] = 1
            return
        ev, evec = self.freq_data.eig_device()
        maps = beamform_eig_device(1.0 / ev, evec, self._f[ind], ind, self.steer, self.steer.steer_type, False,
                                   normfactor, 0, nm, self.phase_mode)
        self._ac[ind] = (1.0 / maps.double()).cpu().numpy().astype(self.precision)
        self._fr[ind] = 1


class BeamformerEig(BeamformerBase):
    """Beamforming with one eigenvalue/eigenvector pair of the CSM (``fbeamform.py:821-888``):
    ``n`` selects the pair (negative values count from the largest eigenvalue, default -1).

    All requested bins run in one ``abf_beamform_eig_grid`` launch over the resident
    eigendecomposition (``PowerSpectra.eig_device``).
    """

    _params = BeamformerBase._params + ('n',)

    def __init__(self, **kwargs):
        object.__setattr__(self, 'n', -1)
        super().__init__(**kwargs)

    @property
    def na(self):
        """Index of the eigenvalue taken into account (``fbeamform.py:845-851``)."""
        na = int(self.n)
        nm = self.steer.mics.num_mics
        if na < 0:
            na = max(nm + na, 0)
        return min(nm - 1, na)

    def _eig_maps(self, ind, eig_first, num_eig, norm, eigvals=None, out_norm=False, clip=False):
        """Eigenvalue-form maps of the bins ``ind`` for eigenvectors
        ``[eig_first, eig_first + num_eig)`` as a float32 device tensor ``[len(ind), G]``."""
        ev, evec = self.freq_data.eig_device()
        if eigvals is not None:
            ev = eigvals
        return beamform_eig_device(ev, evec, self._f[ind], ind, self.steer, self.steer.steer_type, self.r_diag, norm,
                                   eig_first, num_eig, self.phase_mode, clip=clip, out_norm=out_norm)

    def _calc_custom(self, ind, eig_slice, norm, post=None):
        # explicit steering vectors: the reference's per-bin loop over beamformerFreq('custom', ...)
        _, steer_func = self._beamformer_params()
        for i in ind:
            eigvals = np.array(self.freq_data.eigvals[i], dtype='float64')
            eigvecs = np.array(self.freq_data.eigvecs[i], dtype='complex128')
            out = beamformerFreq('custom', self.r_diag, norm, steer_func(self._f[i]),
                                 (eigvals[eig_slice], eigvecs[:, eig_slice]))[0]
            self._ac[i] = post(out) if post else out
            self._fr[i] = 1

    def _calc(self, ind):
        ind = np.asarray(ind, dtype=np.int64)
        na = int(self.na)
        normfactor = self.sig_loss_norm()
        if type(self.steer) is not SteeringVector:

            def clip(out):
                if self.r_diag:
                    out[np.sign(out) < 0] = 0
                return out

            self._calc_custom(ind, slice(na, na + 1), normfactor, clip)
            return
        maps = self._eig_maps(ind, na, 1, normfactor, clip=self.r_diag)
        self._ac[ind] = maps.cpu().numpy().astype(self.precision)
        self._fr[ind] = 1


class BeamformerMusic(BeamformerEig):
    """MUSIC beamforming (``fbeamform.py:891-945``): the eigenvalue form over the ``M - na``
    smallest eigenpairs (the noise subspace for ``n`` assumed sources), mapped to
    ``4e-10 * min(B) / B`` per bin.  The CSM diagonal is always kept."""

    _fixed = {'r_diag': False, 'r_diag_norm': 1.0}

    def __init__(self, **kwargs):
        super().__init__(**kwargs)
        if 'n' not in kwargs:
            object.__setattr__(self, 'n', 1)  # assumed number of sources

    def _calc(self, ind):
        ind = np.asarray(ind, dtype=np.int64)
        nmics = self.freq_data.num_channels
        n = int(self.steer.mics.num_mics - self.na)
        normfactor = self.sig_loss_norm() * nmics**2
        if type(self.steer) is not SteeringVector:
            self._calc_custom(ind, slice(0, n), normfactor, lambda out: 4e-10 * out.min() / out)
            return
        maps = self._eig_maps(ind, 0, n, normfactor).double()
        maps = 4e-10 * maps.amin(dim=1, keepdim=True) / maps
        self._ac[ind] = maps.cpu().numpy().astype(self.precision)
        self._fr[ind] = 1


class BeamformerFunctional(BeamformerBase):
    """Functional beamforming (``fbeamform.py:688-774``): the eigenvalue form with eigenvalues
    raised to ``1/gamma``, normalised by the steering norm and raised back to ``gamma``.
    Only defined for the full CSM (``r_diag`` is always False)."""

    _params = BeamformerBase._params + ('gamma',)
    _fixed = {'r_diag': False, 'r_diag_norm': 1.0}

    def __init__(self, **kwargs):
        object.__setattr__(self, 'gamma', 1.0)
        super().__init__(**kwargs)

    def _calc(self, ind):
        ind = np.asarray(ind, dtype=np.int64)
        normfactor = self.sig_loss_norm()
        gamma = float(self.gamma)
        nm = self.freq_data.num_channels
        if type(self.steer) is not SteeringVector:
            _, steer_func = self._beamformer_params()
            for i in ind:
                eigvals = np.array(self.freq_data.eigvals[i], dtype='float64') ** (1.0 / gamma)
                eigvecs = np.array(self.freq_data.eigvecs[i], dtype='complex128')
                bo, sn = beamformerFreq('custom', self.r_diag, 1.0, steer_func(self._f[i]), (eigvals, eigvecs))
                bo = bo / sn
                self._ac[i] = (bo**gamma) * sn * normfactor
                self._fr[i] = 1
            return
        ev, evec = self.freq_data.eig_device()
        ev = ev ** (1.0 / gamma)
        bo, sn = beamform_eig_device(ev, evec, self._f[ind], ind, self.steer, self.steer.steer_type, False, 1.0, 0, nm,
                                     self.phase_mode, out_norm=True)
        bo, sn = bo.double(), sn.double()
        res = ((bo / sn) ** gamma) * sn * normfactor
        self._ac[ind] = res.cpu().numpy().astype(self.precision)
        self._fr[ind] = 1


class BeamformerOrth(BeamformerBase):
    """Orthogonal beamforming (``fbeamform.py:1334-1400``): every selected eigenvector of the CSM
    is beamformed on its own (eigenvalue 1) and its eigenvalue / M is placed at the grid point
    where that map peaks.  ``eva_list`` holds the eigenvector indices (numpy indexing, default
    the largest); setting ``n`` selects the ``n`` largest.  The maps of one eigenvector for all
    requested bins come from one ``abf_beamform_eig_grid`` launch.
    """

    _params = BeamformerBase._params + ('eva_list', 'n')

    def __init__(self, **kwargs):
        object.__setattr__(self, 'eva_list', np.array([-1]))
        object.__setattr__(self, 'n', 1)
        super().__init__(**kwargs)

    def __setattr__(self, name, value):
        if name == 'eva_list':
            value = np.atleast_1d(np.asarray(value, dtype=int))
        if name == 'n':
            value = int(value)
            super().__setattr__('n', value)
            # the observer of the reference (fbeamform.py:1364-1367)
            super().__setattr__('eva_list', np.arange(-1, -1 - value, -1))
            return
        super().__setattr__(name, value)

    def _digest_items(self):
        # n acts only through eva_list (fbeamform.py:1356-1358)
        return tuple(getattr(self, p) for p in self._params if p != 'n')

    def _calc(self, ind):
        torch = _lib.torch()
        ind = np.asarray(ind, dtype=np.int64)
        nm = self.freq_data.num_channels
        normfactor = self.sig_loss_norm()
        eigvals = self.freq_data.eigvals
        if type(self.steer) is not SteeringVector:
            _, steer_func = self._beamformer_params()
            for i in ind:
                eigvecs = np.array(self.freq_data.eigvecs[i], dtype='complex128')
                for n in self.eva_list:
                    out = beamformerFreq('custom', self.r_diag, normfactor, steer_func(self._f[i]),
                                         (np.ones(1), eigvecs[:, n].reshape((-1, 1))))[0]
                    self._ac[i, out.argmax()] += np.float64(eigvals[i][n]) / nm
                self._fr[i] = 1
            return
        _, evec = self.freq_data.eig_device()
        ones = torch.ones((evec.shape[0], nm), dtype=torch.float64, device=evec.device)
        for n in self.eva_list:
            e = int(np.arange(nm)[n])  # numpy indexing: negative counts from the end, IndexError beyond
            maps = beamform_eig_device(ones, evec, self._f[ind], ind, self.steer, self.steer.steer_type, self.r_diag,
                                       normfactor, e, 1, self.phase_mode).cpu().numpy()
            peak = maps.argmax(axis=1)
            for row, i in enumerate(ind):
                self._ac[i, peak[row]] += np.float64(eigvals[i][n]) / nm
        self._fr[ind] = 1


class BeamformerCleansc(BeamformerBase):
    """CLEAN-SC deconvolution (``fbeamform.py:1403-1492``): up to ``n_iter`` (0: 2 M) removals of
    the coherent source component at the map's peak, damped by ``damp``, stopping when the peak
    rises against the one ``stopn`` iterations earlier.  All requested bins iterate together on the
    device (``abf_beamform_f64`` for the dirty maps, ``abf_cleansc`` for the iterations) in
    float64, as the reference computes.
    """

    _params = BeamformerBase._params + ('n_iter', 'damp', 'stopn')

    def __init__(self, **kwargs):
        object.__setattr__(self, 'n_iter', 0)
        object.__setattr__(self, 'damp', 0.6)
        object.__setattr__(self, 'stopn', 3)
        super().__init__(**kwargs)

    def __setattr__(self, name, value):
        if name in ('n_iter', 'stopn'):
            value = int(value)
        if name == 'damp':
            value = float(value)
            if not 0.01 <= value <= 1.0:
                raise ValueError('damp must be in [0.01, 1.0]')
        super().__setattr__(name, value)

    def _calc(self, ind):
        ind = np.asarray(ind, dtype=np.int64)
        if type(self.steer) is not SteeringVector:
            msg = 'BeamformerCleansc on the device needs a plain SteeringVector'
            raise NotImplementedError(msg)
        normfactor = self.sig_loss_norm()
        csm = self.freq_data.csm_device_f64()
        out = cleansc_device(csm, self._f[ind], ind, self.steer, self.steer.steer_type, self.r_diag, normfactor,
                             self.n_iter, self.damp, self.stopn, self.phase_mode)
        self._ac[ind] = out.cpu().numpy().astype(self.precision)
        self._fr[ind] = 1


class BeamformerClean(BeamformerBase):
    """Hoegbom CLEAN (``fbeamform.py:1495-1598``): from the float64 dirty map (clipped at zero with
    ``r_diag``), repeatedly move ``damp`` times the peak into the clean map and subtract that
    amplitude times the peak's PSF column, until an iteration no longer lowers the sum of absolute
    values, ``n_iter`` iterations have run or no positive value is left.  All requested bins
    iterate together on the device (``abf_beamform_f64``, then ``abf_clean``); the PSF columns are
    formed per iteration on the device, so ``calcmode`` only keeps the reference's interface."""

    _params = BeamformerBase._params + ('n_iter', 'damp', 'psf_precision')

    def __init__(self, **kwargs):
        object.__setattr__(self, 'psf_precision', 'float64')
        object.__setattr__(self, 'damp', 0.6)
        object.__setattr__(self, 'n_iter', 100)
        object.__setattr__(self, 'calcmode', 'block')
        super().__init__(**kwargs)

    def __setattr__(self, name, value):
        if name == 'psf_precision' and value not in ('float64', 'float32'):
            raise ValueError("psf_precision must be 'float64' or 'float32'")
        if name == 'calcmode' and value not in ('block', 'full', 'single', 'readonly'):
            raise ValueError("calcmode must be 'block', 'full', 'single' or 'readonly'")
        if name == 'n_iter':
            value = int(value)
        if name == 'damp':
            value = float(value)
            if not 0.01 <= value <= 1.0:
                raise ValueError('damp must be in [0.01, 1.0]')
        super().__setattr__(name, value)

    def _calc(self, ind):
        ind = np.asarray(ind, dtype=np.int64)
        if type(self.steer) is not SteeringVector:
            msg = 'BeamformerClean on the device needs a plain SteeringVector'
            raise NotImplementedError(msg)
        normfactor = self.sig_loss_norm()
        dirty = beamform_f64_device(self.freq_data.csm_device_f64(), self._f[ind], ind, self.steer,
                                    self.steer.steer_type, self.r_diag, normfactor, self.phase_mode, clip=self.r_diag)
        out = clean_device(dirty, self._f[ind], self.steer, self.n_iter, self.damp, self.psf_precision,
                           self.phase_mode)
        self._ac[ind] = out.cpu().numpy().astype(self.precision)
        self._fr[ind] = 1


class PointSpreadFunction(HasParams):
    """The point spread function of a grid and microphone array (``fbeamform.py:948-1124``).

    ``psf`` is ``[G, len(grid_indices)]`` (``[G, G]`` for all points): column ``c`` is the map a
    unit monopole at grid point ``grid_indices[c]`` produces.  Computed columns are kept per
    ``(digest, freq)`` in memory, where the reference keeps them in an HDF5 cache node
    ``Hz_<freq:.2f>`` (``fbeamform.py:1005-1035``); ``calcmode`` keeps its meaning ('readonly'
    refuses to compute columns missing from a stored entry).  Every mode computes the missing
    columns in one ``abf_psf`` launch.
    """

    _params = ('steer', 'grid_indices', 'calcmode', 'precision', 'freq', 'cached', 'phase_mode')

    def __init__(self, **kwargs):
        self.steer = SteeringVector()
        self.grid_indices = np.array([], dtype=int)
        self.calcmode = 'single'
        self.precision = 'float64'
        self.freq = 1.0
        self.cached = True
        self.phase_mode = config['phase_mode']
        object.__setattr__(self, '_store', {})
        super().__init__(**kwargs)

    def __setattr__(self, name, value):
        if name == 'steer' and not isinstance(value, SteeringVector):
            raise TypeError('steer must be a SteeringVector')
        if name == 'calcmode' and value not in ('single', 'block', 'full', 'readonly'):
            raise ValueError("calcmode must be one of 'single', 'block', 'full', 'readonly'")
        if name == 'precision' and value not in ('float64', 'float32'):
            raise ValueError("precision must be 'float64' or 'float32'")
        if name == 'grid_indices':
            value = np.atleast_1d(np.asarray(value, dtype=int))
        if name == 'freq':
            value = float(value)
        object.__setattr__(self, name, value)

    def _digest_items(self):
        # the reference's digest depends on the steering vector and the precision only
        return (self.steer, self.precision)

    @property
    def psf(self):
        """``_get_psf`` (``fbeamform.py:1037-1071``)."""
        gs = self.steer.grid.size
        grid_indices = self.grid_indices if self.grid_indices.size else np.arange(gs)
        key = (self.digest, f'Hz_{self.freq:.2f}')
        entry = self._store.get(key) if self.cached else None
        from_cache = entry is not None
        if from_cache:
            ac, gp = entry
        else:
            ac = np.zeros((gs, gs), dtype=self.precision)
            gp = np.zeros((gs,), dtype='int8')
        if not gp[grid_indices].astype(bool).all():
            if from_cache and self.calcmode == 'readonly':
                msg = "Cannot calculate missing PSF (points) in 'readonly' mode."
                raise ValueError(msg)
            self.calc_psf(ac, gp, grid_indices)
            if self.cached:
                self._store[key] = (ac, gp)
        if grid_indices.size == gs and np.array_equal(grid_indices, np.arange(gs)):
            return ac
        return ac[:, grid_indices].copy()

    def calc_psf(self, ac, gp, grid_indices=None):
        """Fill the columns ``grid_indices`` of ``ac`` and mark them in ``gp``
        (``fbeamform.py:1073-1094``)."""
        if grid_indices is None:
            grid_indices = self.grid_indices if self.grid_indices.size else np.arange(self.steer.grid.size)
        if self.calcmode == 'full':
            gp[:] = 1
            ac[:] = self._psf_call(np.arange(self.steer.grid.size))
            return
        todo = grid_indices[~gp[grid_indices].astype(bool)]
        if todo.size:
            ac[:, todo] = self._psf_call(todo)
            gp[todo] = 1

    def _psf_call(self, ind):
        """``_psf_call`` (``fbeamform.py:1096-1124``): ``[G, len(ind)]`` of ``precision``."""
        ind = np.asarray(ind, dtype=np.int64)
        if type(self.steer) is SteeringVector:
            out = psf_device(self.steer, [self.freq], ind, self.precision, self.phase_mode)
            return out[0].cpu().numpy()
        product = np.dot(self.steer.steer_vector(self.freq).conj(), self.steer.transfer(self.freq, ind).T)
        return (product * product.conj()).real.astype(self.precision)


class BeamformerDamas(BeamformerBase):
    """DAMAS deconvolution (``fbeamform.py:1127-1208``): the conventional map of every requested
    bin, then ``n_iter`` relaxed Gauss-Seidel sweeps (damping ``damp``) against the full point
    spread function.  The dirty maps, the PSFs of all bins and the sweeps stay on the device:
    one ``abf_beamform`` job, one ``abf_psf`` launch and one ``abf_damas`` launch per group of
    bins whose PSFs fit the memory budget.
    """

    _params = BeamformerBase._params + ('n_iter', 'damp', 'psf_precision', 'calcmode')
    #: bytes of PSF matrices held on the device at once
    psf_budget = 8 << 30

    def __init__(self, **kwargs):
        object.__setattr__(self, 'psf_precision', 'float64')
        object.__setattr__(self, 'n_iter', 100)
        object.__setattr__(self, 'damp', 1.0)
        object.__setattr__(self, 'calcmode', 'full')
        super().__init__(**kwargs)

    def __setattr__(self, name, value):
        if name == 'psf_precision' and value not in ('float64', 'float32'):
            raise ValueError("psf_precision must be 'float64' or 'float32'")
        if name == 'calcmode' and value not in ('full', 'single', 'block', 'readonly'):
            raise ValueError("calcmode must be one of 'full', 'single', 'block', 'readonly'")
        if name == 'n_iter':
            value = int(value)
        if name == 'damp':
            value = float(value)
        super().__setattr__(name, value)

    def _digest_items(self):
        # calcmode does not change the result (fbeamform.py:1158-1160)
        return tuple(getattr(self, p) for p in self._params if p != 'calcmode')

    def _calc(self, ind):
        torch = _lib.torch()
        ind = np.asarray(ind, dtype=np.int64)
        normfactor = self.sig_loss_norm()
        G = self.steer.grid.size
        f64 = self.psf_precision == 'float64'
        if type(self.steer) is not SteeringVector:
            p = PointSpreadFunction(steer=self.steer, calcmode=self.calcmode, precision=self.psf_precision,
                                    cached=self.cached, phase_mode=self.phase_mode)
            _, steer_func = self._beamformer_params()
            for i in ind:
                csm = np.array(self.freq_data.csm[i], dtype='complex128')
                y = beamformerFreq('custom', self.r_diag, normfactor, steer_func(self._f[i]), csm)[0]
                if self.r_diag:
                    y[np.sign(y) < 0] = 0.0
                p.freq = self._f[i]
                self._ac[i] = damas_device(_lib.to_device(np.ascontiguousarray(p.psf)), _lib.to_device(y[None]),
                                           self.n_iter, self.damp)[0].cpu().numpy()
                self._fr[i] = 1
            return
        per_bin = G * G * (8 if f64 else 4)
        chunk = max(1, int(self.psf_budget // max(per_bin, 1)))
        csm = self.freq_data.csm_device_f64()
        for c0 in range(0, len(ind), chunk):
            sel = ind[c0:c0 + chunk]
            # the sweeps amplify the rounding of the dirty map by up to ~1e3 at low frequencies, so
            # it is formed in float64 like the reference's (fbeamform.py:1191-1197)
            y = beamform_f64_device(csm, self._f[sel], sel, self.steer, self.steer.steer_type, self.r_diag,
                                    normfactor, self.phase_mode, clip=self.r_diag)
            psf = psf_device(self.steer, self._f[sel], np.arange(G), self.psf_precision, self.phase_mode)
            x = damas_device(psf, y, self.n_iter, self.damp)
            del psf
            self._ac[sel] = x.cpu().numpy().astype(self.precision)
            self._fr[sel] = 1


class BeamformerDamasPlus(BeamformerDamas):
    """DAMAS through a general solver instead of Gauss-Seidel sweeps (``fbeamform.py:1211-1331``):
    ``method`` 'NNLS' or 'LP' (scipy.optimize) or 'LassoLars' or 'OMPCV' (scikit-learn), on the
    dirty map scaled by ``unit_mult``.  The float64 dirty maps and the full PSFs of a group of bins
    come from one ``abf_beamform_f64`` and one ``abf_psf`` launch; the solve of each bin is the
    reference's library call on the host (those libraries have no device form).
    """

    _params = BeamformerDamas._params + ('method', 'alpha', 'unit_mult')

    def __init__(self, **kwargs):
        object.__setattr__(self, 'method', 'NNLS')
        object.__setattr__(self, 'alpha', 0.0)
        object.__setattr__(self, 'unit_mult', 1e9)
        super().__init__(**kwargs)
        if 'n_iter' not in kwargs:
            object.__setattr__(self, 'n_iter', 500)

    def __setattr__(self, name, value):
        if name == 'method' and value not in ('NNLS', 'LP', 'LassoLars', 'OMPCV'):
            raise ValueError("method must be one of 'NNLS', 'LP', 'LassoLars', 'OMPCV'")
        if name == 'alpha':
            value = float(value)
            if not 0.0 <= value <= 1.0:
                raise ValueError('alpha must lie in [0, 1]')
        if name == 'unit_mult':
            value = float(value)
        super().__setattr__(name, value)

    def _calc(self, ind):
        ind = np.asarray(ind, dtype=np.int64)
        normfactor = self.sig_loss_norm()
        G = self.steer.grid.size
        if type(self.steer) is not SteeringVector:
            p = PointSpreadFunction(steer=self.steer, calcmode=self.calcmode, precision=self.psf_precision,
                                    cached=self.cached, phase_mode=self.phase_mode)
            _, steer_func = self._beamformer_params()
            for i in ind:
                csm = np.array(self.freq_data.csm[i], dtype='complex128')
                y = beamformerFreq('custom', self.r_diag, normfactor, steer_func(self._f[i]), csm)[0]
                if self.r_diag:
                    y[np.sign(y) < 0] = 0.0
                p.freq = self._f[i]
                self._ac[i] = damas_plus_solve(p.psf, y, self.method, self.alpha, self.n_iter, self.unit_mult,
                                               self.r_diag)
                self._fr[i] = 1
            return
        per_bin = G * G * (8 if self.psf_precision == 'float64' else 4)
        chunk = max(1, int(self.psf_budget // max(per_bin, 1)))
        csm = self.freq_data.csm_device_f64()
        for c0 in range(0, len(ind), chunk):
            sel = ind[c0:c0 + chunk]
            y = beamform_f64_device(csm, self._f[sel], sel, self.steer, self.steer.steer_type, self.r_diag,
                                    normfactor, self.phase_mode, clip=self.r_diag).cpu().numpy()
            psf = psf_device(self.steer, self._f[sel], np.arange(G), self.psf_precision, self.phase_mode)
            psf = psf.cpu().numpy()
            for j, i in enumerate(sel):
                self._ac[i] = damas_plus_solve(psf[j], y[j], self.method, self.alpha, self.n_iter, self.unit_mult,
                                               self.r_diag)
                self._fr[i] = 1


def damas_plus_solve(psf, y, method: str, alpha: float, n_iter: int, unit: float, r_diag: bool):
    """One bin of ``BeamformerDamasPlus._calc`` (``fbeamform.py:1286-1331``): solve psf x = y
    (x >= 0) for the dirty map ``y`` scaled by ``unit`` with the reference's solver and return
    x / unit."""
    y = np.asarray(y, dtype=np.float64) * unit
    if method == 'NNLS':
        from scipy.optimize import nnls
        return nnls(psf, y)[0] / unit
    if method == 'LP':
        from scipy.optimize import linprog
        if r_diag:
            _warn('Linear programming solver may fail when CSM main diagonal is removed for delay-and-sum '
                  'beamforming.', Warning, stacklevel=5)
        # maximise the summed map: minimise -psf.sum(1) . x subject to psf x <= y, x >= 0
        return linprog(c=-1 * psf.sum(1), A_ub=psf, b_ub=y).x / unit
    if method not in ('LassoLars', 'OMPCV'):
        raise NotImplementedError(f'Method {method} not implemented.')
    import warnings

    from sklearn.linear_model import LassoLars, OrthogonalMatchingPursuitCV
    if method == 'LassoLars':
        model = LassoLars(alpha=alpha * unit, max_iter=n_iter, positive=True)
    else:
        model = OrthogonalMatchingPursuitCV()
    # columns normalised here and the coefficients scaled back, independent of scikit-learn's
    # own (version-dependent) normalisation (fbeamform.py:1315-1330)
    norms = np.linalg.norm(psf, axis=0)
    with warnings.catch_warnings():
        warnings.simplefilter('ignore', category=FutureWarning)
        model.fit(psf / norms, y)
    return model.coef_[:] / norms / unit


class BeamformerAdaptiveGrid(BeamformerBase, Grid):
    """A beamformer whose result lives on positions it finds itself (``fbeamform.py:2242-2282``):
    ``pos`` holds those positions (3, ``size``), ``shape`` is ``(size,)``."""

    @property
    def shape(self):
        return (self.size,)

    @property
    def pos(self):
        return getattr(self, '_gpos', None)

    @property
    def result(self):
        d = self.digest
        if self._result_digest != d:
            self._f = self.freq_data.fftfreq()
            self._numfreq = self._f.shape[0]
            self._assert_equal_channels()
            # fbeamform.py:425-428
            self._gpos = np.zeros((3, self.size), dtype=self.precision)
            self._ac = np.zeros((self._numfreq, self.size), dtype=self.precision)
            self._fr = np.zeros(self._numfreq, dtype='int8')
            object.__setattr__(self, '_result_digest', d)
        return LazyBfResult(self)

    def integrate(self, sector):
        """Per-bin sum of the result over the found positions inside ``sector``
        (``fbeamform.py:2255-2282``)."""
        from .grids import Sector

        if not isinstance(sector, Sector):
            msg = (f'Please use a sector derived instance of type :class:`~acoular.grids.Sector` instead of type '
                   f'{type(sector)}.')
            raise NotImplementedError(msg)
        # every bin first: the positions the sector selects from exist only once computed (the
        # reference takes the subdomain before, and the bin count from the lazy result's
        # r.shape[0], which it does not have)
        r = self.result[:]
        ind = self.subdomain(sector)
        h = np.zeros(r.shape[0])
        for i in range(r.shape[0]):
            h[i] = r[i][ind].sum()
        return h


class BeamformerGridlessOrth(BeamformerAdaptiveGrid):
    """Orthogonal beamforming without a grid (``fbeamform.py:2285-2410``): for every bin and every
    selected eigenvector, scipy's ``shgo`` looks for the position inside ``bounds`` where the
    eigenvector's map (eigenvalue 1) peaks; the position and eigenvalue / M are stored.

    The objective is the drop-in ``beamformerFreq`` (eigenvalue form, one point) as in the
    reference; the sampling points of each ``shgo`` iteration are evaluated together in one
    ``beamformerFreq`` call over all of them (``shgo``'s ``workers`` map), which gives the same
    values as one call per point.
    """

    _params = BeamformerBase._params + ('eva_list', 'n', 'bounds', 'shgo')
    _fixed = {'r_diag_norm': 1.0}

    def __init__(self, **kwargs):
        object.__setattr__(self, 'eva_list', np.array([-1]))
        object.__setattr__(self, 'n', 1)
        object.__setattr__(self, 'bounds', [(-1.0, 1.0), (-1.0, 1.0), (0.01, 1.0)])
        object.__setattr__(self, 'shgo', {})
        super().__init__(**kwargs)

    def __setattr__(self, name, value):
        if name == 'eva_list':
            value = np.atleast_1d(np.asarray(value, dtype=int))
        if name == 'bounds':
            value = [(float(a), float(b)) for a, b in value]
            if len(value) != 3:
                raise ValueError('bounds must hold three (min, max) pairs')
        if name == 'shgo':
            value = dict(value)
        if name == 'n':
            value = int(value)
            super().__setattr__('n', value)
            # the observer of the reference (fbeamform.py:2331-2334)
            super().__setattr__('eva_list', np.arange(-1, -1 - value, -1))
            return
        super().__setattr__(name, value)

    def _digest_items(self):
        # fbeamform.py:2322-2324: n acts through eva_list and the size
        return tuple(getattr(self, p) for p in self._params if p not in ('r_diag_norm', 'form'))

    @property
    def size(self):
        return self.n * self.freq_data.fftfreq().shape[0]

    def _roi(self):
        return np.array([(x, y, z) for x in self.bounds[0] for y in self.bounds[1] for z in self.bounds[2]]).T

    @property
    def result(self):
        # the region of interest the reference sets inside _calc (fbeamform.py:2374-2380), set
        # before the digest is taken so that it does not invalidate the result it belongs to
        roi = self._roi()
        if self.steer.env.roi is None or not np.array_equal(np.asarray(self.steer.env.roi), roi):
            self.steer.env.roi = roi
        return super().result

    def _calc(self, ind):
        """``fbeamform.py:2337-2410``."""
        from scipy.optimize import shgo

        f = self._f
        normfactor = self.sig_loss_norm()
        num_channels = self.freq_data.num_channels
        eva_list = np.unique(self.eva_list % self.steer.mics.num_mics)[::-1]
        steer_type = self.steer.steer_type
        if steer_type == 'custom':
            raise NotImplementedError('custom steer_type is not implemented')
        mpos = self.steer.mics.pos
        env = self.steer.env
        opts = {'n': 256, 'iters': 1, 'sampling_method': 'sobol', 'minimizer_kwargs': {'method': 'Nelder-Mead'}}
        opts.update(self.shgo)
        bmin = np.array(tuple(map(min, self.bounds)))
        bmax = np.array(tuple(map(max, self.bounds)))
        for i in ind:
            eigvals = np.array(self.freq_data.eigvals[i], dtype='float64')
            eigvecs = np.array(self.freq_data.eigvecs[i], dtype='complex128')
            k = 2 * np.pi * f[i] / env.c
            for j, n in enumerate(eva_list):
                vec = (np.ones(1), eigvecs[:, n:n + 1])

                def levels(pts, vec=vec, k=k):
                    # minus the map of eigenvector n at the clipped points (3, P)
                    pts = np.clip(pts, bmin[:, None], bmax[:, None])
                    r0 = env.apparent_r(pts)
                    rm = env.apparent_r(pts, mpos)
                    return -beamformerFreq(steer_type, self.r_diag, normfactor, (np.atleast_1d(r0),
                                                                                  np.atleast_2d(rm), k), vec)[0]

                def func(xy, levels=levels):
                    return levels(np.asarray(xy, dtype=np.float64)[:, None])[0]

                def batch(fn, points, levels=levels):
                    # shgo's map over sampling points: one device call for all of them, with
                    # FieldWrapper's treatment of nan (scipy/optimize/_shgo_lib/_vertex.py)
                    pts = list(points)
                    if not pts:
                        return []
                    v = levels(np.stack([np.asarray(p, dtype=np.float64) for p in pts], axis=1))
                    return [np.inf if np.isnan(x) else float(x) for x in v]

                o = dict(opts)
                o.setdefault('workers', batch)
                res = shgo(func, self.bounds, **o)
                i1 = i * self.n + j
                self._gpos[:, i1] = res['x']
                self._ac[i, i1] = eigvals[n] / num_channels
            self._fr[i] = 1


def beamform_f64_device(csm, freqs, ind, steer: SteeringVector, steer_type: str, r_diag: bool, norm: float,
                        phase_mode: str = 'reference', clip: bool = False, out_norm: bool = False):
    """One ``abf_beamform_f64`` call: float64 maps ``[len(ind), G]`` of the complex128 device CSM
    slices ``ind`` at the frequencies ``freqs`` (and the steer normalisation if asked)."""
    torch = _lib.torch()
    ind = np.asarray(ind, dtype=np.int64)
    freqs = np.asarray(freqs, dtype=np.float64)
    if csm.dtype != torch.complex128:
        raise TypeError('csm must be a complex128 device tensor')
    dev = _lib.device()
    if len(ind) != csm.shape[0] or not np.array_equal(ind, np.arange(csm.shape[0])):
        csm = csm.index_select(0, _lib.to_device(ind))
    csm = csm.contiguous()
    gpos, mpos = steer.device_geometry()
    G, M, F = gpos.shape[1], mpos.shape[1], len(ind)
    if csm.shape[1] != M:
        msg = f'{csm.shape[1]:d} channels do not fit {M:d} mics'
        raise ValueError(msg)
    k = _lib.to_device(2 * np.pi * freqs / steer.env.c)
    out = torch.empty((F, G), dtype=torch.float64, device=dev)
    onorm = torch.empty((F, G), dtype=torch.float64, device=dev) if out_norm else None
    lib = _lib.lib()
    ws = _lib.workspace(lib.abf_beamform_f64_workspace_size(M, F))
    _lib.check(lib.abf_beamform_f64(_geometry(gpos, mpos, steer.ref), _lib.ptr(csm), _lib.ptr(k), F,
                                    _lib.STEER_CODES[steer_type], int(bool(r_diag)), float(norm),
                                    _lib.PHASE_MODES[phase_mode], int(bool(clip)), _lib.ptr(out), _lib.ptr(onorm),
                                    _lib.ptr(ws), ws.numel(), _lib.stream_handle()), 'float64 beamform')
    if out_norm:
        return out, onorm
    return out


def psf_device(steer: SteeringVector, freqs, sources, precision: str = 'float64', phase_mode: str = 'reference',
               fast_trig: bool | None = None):
    """One ``abf_psf`` launch: the PSFs of the grid points ``sources`` at every frequency of
    ``freqs`` as a device tensor ``[F, G, len(sources)]`` of ``precision``."""
    sources = np.asarray(sources, dtype=np.int64).ravel()
    G = steer.grid.size
    if sources.size and (sources.min() < 0 or sources.max() >= G):
        raise IndexError('grid index out of range')
    torch = _lib.torch()
    dev = _lib.device()
    freqs = np.atleast_1d(np.asarray(freqs, dtype=np.float64))
    if fast_trig is None:
        fast_trig = config['fast_trig']
    gpos, mpos = steer.device_geometry()
    k = _lib.to_device(2 * np.pi * freqs / steer.env.c)
    src = _lib.to_device(sources)
    dt = torch.float64 if precision == 'float64' else torch.float32
    out = torch.empty((len(freqs), G, len(sources)), dtype=dt, device=dev)
    pm = _lib.PHASE_MODES[phase_mode] | (_lib.FAST_TRIG if fast_trig else 0)
    _lib.check(_lib.lib().abf_psf(_geometry(gpos, mpos, steer.ref), _lib.ptr(k), len(freqs), _lib.ptr(src),
                                  len(sources), _lib.STEER_CODES[steer.steer_type], pm, int(dt == torch.float64),
                                  _lib.ptr(out), _lib.stream_handle()), 'point spread function')
    return out


def damas_device(psf, dirty, n_iter: int, damp: float):
    """One ``abf_damas`` launch: ``n_iter`` Gauss-Seidel sweeps of ``psf [F, G, G] x = dirty
    [F, G]`` (float64 device tensors; ``psf`` may be float32) from ``x = dirty``."""
    torch = _lib.torch()
    if psf.dtype not in (torch.float64, torch.float32) or dirty.dtype != torch.float64:
        raise TypeError('psf must be float64 or float32 and dirty float64')
    psf, dirty = psf.contiguous(), dirty.contiguous()
    F, G = dirty.shape
    if tuple(psf.shape) != (F, G, G):
        raise ValueError(f'psf of shape {tuple(psf.shape)} does not fit dirty maps of shape {(F, G)}')
    x = dirty.clone()
    _lib.check(_lib.lib().abf_damas(_lib.ptr(psf), int(psf.dtype == torch.float64), _lib.ptr(dirty), _lib.ptr(x),
                                    G, F, int(n_iter), float(damp), _lib.stream_handle()), 'DAMAS')
    return x


def cleansc_device(csm, freqs, ind, steer: SteeringVector, steer_type: str, r_diag: bool, norm: float, n_iter: int,
                   damp: float, stopn: int, phase_mode: str = 'reference'):
    """CLEAN-SC of the complex128 device CSM slices ``ind`` at the frequencies ``freqs``: the
    float64 dirty maps (``abf_beamform_f64``) and the iterations of every bin together
    (``abf_cleansc``, which works on a copy of the slices).  Returns the float32 result maps
    ``[len(ind), G]`` as a device tensor."""
    torch = _lib.torch()
    ind = np.asarray(ind, dtype=np.int64)
    dev = _lib.device()
    if csm.dtype != torch.complex128:
        raise TypeError('csm must be a complex128 device tensor')
    work = csm.index_select(0, _lib.to_device(ind)).contiguous()
    h = beamform_f64_device(work, freqs, np.arange(len(ind)), steer, steer_type, r_diag, norm, phase_mode)
    gpos, mpos = steer.device_geometry()
    F, M = len(ind), mpos.shape[1]
    k = _lib.to_device(2 * np.pi * np.asarray(freqs, dtype=np.float64) / steer.env.c)
    out = torch.empty((F, gpos.shape[1]), dtype=torch.float32, device=dev)
    lib = _lib.lib()
    ws = _lib.workspace(lib.abf_cleansc_workspace_size(M, gpos.shape[1], F, int(n_iter)))
    _lib.check(lib.abf_cleansc(_geometry(gpos, mpos, steer.ref), _lib.ptr(k), F, _lib.ptr(work), _lib.ptr(h),
                               _lib.STEER_CODES[steer_type], int(bool(r_diag)), float(norm),
                               _lib.PHASE_MODES[phase_mode], int(n_iter), float(damp), int(stopn), _lib.ptr(out),
                               _lib.ptr(ws), ws.numel(), _lib.stream_handle()), 'CLEAN-SC')
    return out


def clean_device(dirty, freqs, steer: SteeringVector, n_iter: int, damp: float, psf_precision: str = 'float64',
                 phase_mode: str = 'reference'):
    """One ``abf_clean`` call: CLEAN of the float64 device dirty maps ``[F, G]`` (updated in place)
    at the frequencies ``freqs``.  Returns the float64 clean maps ``[F, G]`` as a device tensor."""
    torch = _lib.torch()
    if dirty.dtype != torch.float64:
        raise TypeError('dirty must be a float64 device tensor')
    if psf_precision not in ('float64', 'float32'):
        raise ValueError("psf_precision must be 'float64' or 'float32'")
    dirty = dirty.contiguous()
    gpos, mpos = steer.device_geometry()
    F, G = dirty.shape
    freqs = np.asarray(freqs, dtype=np.float64)
    if G != gpos.shape[1] or len(freqs) != F:
        raise ValueError(f'dirty maps of shape {tuple(dirty.shape)} do not fit {gpos.shape[1]:d} grid points and '
                         f'{len(freqs):d} frequencies')
    dev = dirty.device
    k = _lib.to_device(2 * np.pi * freqs / steer.env.c)
    out = torch.empty((F, G), dtype=torch.float64, device=dev)
    lib = _lib.lib()
    ws = _lib.workspace(lib.abf_clean_workspace_size(mpos.shape[1], G, F))
    _lib.check(lib.abf_clean(_geometry(gpos, mpos, steer.ref), _lib.ptr(k), F, _lib.ptr(dirty),
                             _lib.STEER_CODES[steer.steer_type], _lib.PHASE_MODES[phase_mode],
                             int(psf_precision == 'float32'), int(n_iter), float(damp), _lib.ptr(out), _lib.ptr(ws),
                             ws.numel(), _lib.stream_handle()), 'CLEAN')
    return out


def _geometry(gpos, mpos, ref) -> _lib.AbfGeometry:
    """``abf_geometry`` of device coordinate tensors and a reference point or distance."""
    g = _lib.AbfGeometry()
    g.gpos = gpos.data_ptr()
    g.mpos = mpos.data_ptr()
    g.num_grid = gpos.shape[1]
    g.num_mics = mpos.shape[1]
    if np.isscalar(ref):
        g.ref_dist = float(ref) if ref > 0 else 0.0
        g.ref[:] = [0.0, 0.0, 0.0]
    else:
        g.ref[:] = [float(x) for x in ref]
        g.ref_dist = 0.0
    return g


def beamform_eig_device(eigvals, eigvecs, freqs, ind, steer: SteeringVector, steer_type: str, r_diag: bool,
                        norm: float, eig_first: int, num_eig: int, phase_mode: str = 'reference', clip: bool = False,
                        out_norm: bool = False, fast_trig: bool | None = None):
    """One ``abf_beamform_eig_grid`` call: eigenvalue-form maps of the bins ``ind`` (rows of the
    device eigendecomposition ``eigvals [F, M]`` float64, ``eigvecs [F, M, M]`` complex64) at the
    frequencies ``freqs``.  Returns a float32 device tensor ``[len(ind), G]`` (and the steer
    normalisation if asked)."""
    torch = _lib.torch()
    ind = np.asarray(ind, dtype=np.int64)
    freqs = np.asarray(freqs, dtype=np.float64)
    dev = _lib.device()
    if eigvals.dtype != torch.float64 or eigvecs.dtype != torch.complex64:
        raise TypeError('eigvals must be float64 and eigvecs complex64 device tensors')
    if len(ind) != eigvals.shape[0] or not np.array_equal(ind, np.arange(eigvals.shape[0])):
        sel = _lib.to_device(ind)
        eigvals, eigvecs = eigvals.index_select(0, sel), eigvecs.index_select(0, sel)
    eigvals, eigvecs = eigvals.contiguous(), eigvecs.contiguous()
    gpos, mpos = steer.device_geometry()
    G, M, F = gpos.shape[1], mpos.shape[1], len(ind)
    if eigvecs.shape[1:] != (M, M) or eigvals.shape[1] != M:
        msg = f'eigendecomposition of shape {tuple(eigvecs.shape[1:])} does not fit {M:d} mics'
        raise ValueError(msg)
    if fast_trig is None:
        fast_trig = config['fast_trig']
    k = _lib.to_device(2 * np.pi * freqs / steer.env.c)
    out = torch.empty((F, G), dtype=torch.float32, device=dev)
    onorm = torch.empty((F, G), dtype=torch.float32, device=dev) if out_norm else None
    geom = _geometry(gpos, mpos, steer.ref)
    pm = _lib.PHASE_MODES[phase_mode] | (_lib.FAST_TRIG if fast_trig else 0)
    _lib.check(_lib.lib().abf_beamform_eig_grid(geom, _lib.ptr(k), F, _lib.ptr(eigvals), _lib.ptr(eigvecs),
                                                int(eig_first), int(num_eig), _lib.STEER_CODES[steer_type],
                                                int(bool(r_diag)), float(norm), pm, int(bool(clip)), _lib.ptr(out),
                                                _lib.ptr(onorm), _lib.stream_handle()), 'eigenvalue beamform')
    if out_norm:
        return out, onorm
    return out


def beamform_device(csm_device, freqs, ind, steer: SteeringVector, steer_type: str, r_diag: bool, norm: float,
                    form: str = 'mfma_bf16x3', phase_mode: str = 'reference', clip: bool = False,
                    out_norm: bool = False, fast_trig: bool | None = None):
    """One ``abf_beamform`` job: the maps of CSM slices ``ind`` at frequencies ``freqs``.

    Returns a float32 device tensor ``[len(ind), G]`` (and the steer normalisation if asked).
    """
    torch = _lib.torch()
    ind = np.asarray(ind, dtype=np.int64)
    freqs = np.asarray(freqs, dtype=np.float64)
    dev = _lib.device()
    csm = csm_device
    if len(ind) != csm.shape[0] or not np.array_equal(ind, np.arange(csm.shape[0])):
        csm = csm.index_select(0, _lib.to_device(ind))
    csm = csm.contiguous()
    gpos, mpos = steer.device_geometry()
    G, M, F = gpos.shape[1], mpos.shape[1], len(ind)
    if csm.shape[1] != M:
        msg = f'{csm.shape[1]:d} channels do not fit {M:d} mics'
        raise ValueError(msg)
    if M > 1150 and form != 'plain_f32':
        form = 'plain_f32'
    k = _lib.to_device(2 * np.pi * freqs / steer.env.c)
    out = torch.empty((F, G), dtype=torch.float32, device=dev)
    onorm = torch.empty((F, G), dtype=torch.float32, device=dev) if out_norm else None
    job = make_job(csm, k, gpos, mpos, steer.ref, steer_type, r_diag, norm, form, phase_mode, out, onorm, clip,
                   fast_trig)
    lib = _lib.lib()
    ws = _lib.workspace(lib.abf_beamform_workspace_size(job))
    _lib.check(lib.abf_beamform(job, _lib.ptr(ws), ws.numel(), _lib.stream_handle()), 'beamform')
    job._keep = (csm, k, gpos, mpos, out, onorm, ws)
    if out_norm:
        return out, onorm
    return out


def make_job(csm, k, gpos, mpos, ref, steer_type, r_diag, norm, form, phase_mode, out, out_norm=None, clip=False,
             fast_trig=None) -> _lib.AbfGridJob:
    """Fill an ``abf_grid_job`` from device tensors (the tensors must outlive the call)."""
    if fast_trig is None:
        fast_trig = config['fast_trig']
    j = _lib.AbfGridJob()
    j.geom = _geometry(gpos, mpos, ref)
    j.csm = csm.data_ptr()
    j.wavenumber = k.data_ptr()
    j.num_freqs = csm.shape[0]
    j.steer_type = _lib.STEER_CODES[steer_type]
    j.r_diag = int(bool(r_diag))
    j.clip_negative = int(bool(clip))
    j.norm_factor = float(norm)
    j.form = _lib.FORMS[form]
    j.phase_mode = _lib.PHASE_MODES[phase_mode] | (_lib.FAST_TRIG if fast_trig else 0)
    j.out = out.data_ptr()
    j.out_norm = out_norm.data_ptr() if out_norm is not None else None
    if config['autotuned']:
        from . import autotune

        autotune.apply(j)
    return j


def L_p(x):  # noqa: N802
    """Sound pressure level from squared pressure (``fbeamform.py:2412-2432``)."""
    return 10 * np.log10(np.maximum(x / 4e-10, np.finfo(x.dtype).eps))


def _sector_indices(grid, sector):
    """``grid.subdomain(sector)`` for a ``Sector``, else ``grid.indices(*sector)``
    (``fbeamform.py:640-650, 2478-2490``)."""
    from .grids import Sector

    if isinstance(sector, Sector):
        return grid.subdomain(sector)
    if hasattr(grid, 'indices'):
        return grid.indices(*sector)
    msg = (f'Grid of type {grid.__class__.__name__} does not have an indices method! '
           f'Please use a sector derived instance of type :class:`~acoular.grids.Sector` '
           'instead of type numpy.array.')
    raise NotImplementedError(msg)


def integrate(data, grid, sector):
    """Integrate a map over a sector (``fbeamform.py:2435-2497``): a ``Sector`` or the arguments
    of the grid's ``indices``."""
    ind = _sector_indices(grid, sector)
    gshape = grid.shape
    gsize = grid.size
    if np.size(data) == gsize:  # one value per grid point
        return data.reshape(gshape)[ind].sum()
    if data.ndim == 2 and data.shape[1] == gsize:
        h = np.zeros(data.shape[0])
        for i in range(data.shape[0]):
            h[i] = data[i].reshape(gshape)[ind].sum()
        return h
    return None
