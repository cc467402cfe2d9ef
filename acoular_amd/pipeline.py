"""Fused CSM estimation + beamforming on device-resident time data, replayable as a hipGraph.

``FusedCsmBeamform`` chains the two halves of Acoular's frequency-domain path --
``PowerSpectra.calc_csm`` (``spectra.py:366-416``: Welch blocks, windowed FFT, ``calcCSM``
rank-1 updates, Hermitian completion, scaling) and ``BeamformerBase._calc``
(``fbeamform.py:475-510``: ``beamformerFreq`` for every requested bin) -- as one stream-ordered
sequence of C-ABI launches over buffers allocated once at construction:

    abf_welch_csm   time [N, M] float32 -> CSM [F, M, M] complex64 (HBM resident)
    abf_beamform    CSM slices of the band, grid, microphones -> maps [F_band, G] float32

Nothing in ``run`` allocates, copies to the host or synchronises, so ``capture`` records the
whole pipeline into one ``torch.cuda.CUDAGraph`` (hipGraph) and ``replay`` re-runs every kernel
of it with a single launch call; new time data copied into ``time`` in place is picked up by the
next replay.  The library clears its accumulators with fill kernels (no memset nodes).
"""

from __future__ import annotations

import numpy as np

from . import _lib
from .fbeamform import make_job
from .spectra import _PASS_BYTES, OVERLAPS, WINDOWS


class FusedCsmBeamform:
    """Welch CSM of ``time`` followed by the beamformer over the bins ``[ind_low, ind_high)``.

    ``time``: float32 device tensor ``[N, M]`` (kept by reference: refill it in place between
    replays).  The CSM scaling, block layout and frequency axis follow ``PowerSpectra``; the
    beamformer follows ``BeamformerBase`` with ``steer_type`` and ``r_diag`` (normalisation
    M / (M - 1) when the diagonal is removed), reference point ``ref`` (default origin).
    """

    def __init__(self, time, sample_freq: float, grid_pos, mic_pos, block_size: int = 2048, window: str = 'Hanning',
                 overlap: str = '50%', steer_type: str = 'true location', r_diag: bool = True, ind_low: int = 1,
                 ind_high: int | None = None, form: str = 'mfma_bf16x3', phase_mode: str = 'reference', c: float = 343.0,
                 ref=(0.0, 0.0, 0.0)):
        torch = _lib.torch()
        if window not in WINDOWS:
            raise ValueError(f'window must be one of {list(WINDOWS)}')
        if overlap not in OVERLAPS:
            raise ValueError(f'overlap must be one of {list(OVERLAPS)}')
        if time.dtype != torch.float32 or time.device.type != 'cuda' or not time.is_contiguous():
            raise TypeError('time must be a contiguous float32 device tensor [N, M]')
        dev = _lib.device()
        self.time = time
        n, m = time.shape
        mic_pos = np.asarray(mic_pos, dtype=np.float64)
        grid_pos = np.asarray(grid_pos, dtype=np.float64)
        if mic_pos.shape != (3, m):
            raise ValueError(f'{m:d} channels do not fit {mic_pos.shape[1]:d} mics')
        bs = int(block_size)
        ov = OVERLAPS[overlap]
        self.block_size, self.hop = bs, bs // ov
        self.num_blocks = max(0, (n - bs) // self.hop + 1)
        # PowerSpectra._get_num_blocks (spectra.py:281-283): a float
        nb_float = ov * n / bs - ov + 1
        wind = WINDOWS[window](bs)
        self.scale = 2.0 / bs / float(np.dot(wind, wind)) / nb_float
        self.freqs = np.abs(np.fft.fftfreq(bs, 1.0 / sample_freq)[: bs // 2 + 1])
        nf = bs // 2 + 1
        ind_high = nf if ind_high is None else (ind_high if ind_high >= 0 else nf + ind_high)
        if not 0 <= ind_low < ind_high <= nf:
            raise ValueError('empty or out-of-range frequency band')
        self.ind = np.arange(ind_low, ind_high)
        lib = _lib.lib()
        self.window = torch.from_numpy(wind.astype(np.float32)).to(dev)
        self.csm = torch.empty((nf, m, m), dtype=torch.complex64, device=dev)
        self.bpp = int(max(1, min(max(self.num_blocks, 1), _PASS_BYTES // max(1, nf * m * 8))))
        self.ws_csm = _lib.workspace(lib.abf_welch_csm_workspace_size(m, bs, self.bpp))
        # beamformer job over the CSM slices of the band (a view: no copy)
        self.gpos = torch.from_numpy(np.ascontiguousarray(grid_pos)).to(dev)
        self.mpos = torch.from_numpy(np.ascontiguousarray(mic_pos)).to(dev)
        self.k = torch.from_numpy(2 * np.pi * self.freqs[self.ind] / c).to(dev)
        self.out = torch.empty((len(self.ind), grid_pos.shape[1]), dtype=torch.float32, device=dev)
        band = self.csm[ind_low:ind_high]
        norm = m / (m - 1.0) if r_diag else 1.0
        self.job = make_job(band, self.k, self.gpos, self.mpos, np.asarray(ref, dtype=float), steer_type, r_diag, norm,
                            form, phase_mode, self.out, None, clip=r_diag)
        self.ws_bf = _lib.workspace(lib.abf_beamform_workspace_size(self.job))
        self._band = band
        self.graph = None

    @property
    def map_bins(self) -> int:
        return int(self.out.numel())

    def run_csm(self):
        lib = _lib.lib()
        n, m = self.time.shape
        _lib.check(lib.abf_welch_csm(_lib.ptr(self.time), n, m, self.block_size, self.hop, self.num_blocks,
                                     _lib.ptr(self.window), float(self.scale), _lib.ptr(self.csm), self.bpp,
                                     _lib.ptr(self.ws_csm), self.ws_csm.numel(), _lib.stream_handle()), 'fused CSM')

    def run_beamform(self):
        lib = _lib.lib()
        _lib.check(lib.abf_beamform(self.job, _lib.ptr(self.ws_bf), self.ws_bf.numel(), _lib.stream_handle()),
                   'fused beamform')

    def run(self):
        """Enqueue the whole pipeline on the current stream (no allocation, no synchronisation)."""
        self.run_csm()
        self.run_beamform()
        return self.out

    def capture(self):
        """Record ``run`` into a hipGraph (after one eager run that sets kernel attributes)."""
        torch = _lib.torch()
        self.run()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.run()
        torch.cuda.synchronize()
        self.graph = g
        return g

    def replay(self):
        if self.graph is None:
            raise RuntimeError('capture() first')
        self.graph.replay()
        return self.out
