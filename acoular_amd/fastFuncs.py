"""Drop-in replacements of the numba kernels in ``acoular/fastFuncs.py``, on the GPU.

Same names, argument meaning, return types (host numpy, float64/complex128) and error types as
the reference:

* ``calcCSM(csm, SpecAllMics)``                                     fastFuncs.py:19-58
* ``beamformerFreq(steerVecType, boolRemovedDiagOfCSM, normFactor,
  inputTupleSteer, inputTupleCsm)``                                 fastFuncs.py:61-214
* ``calcTransfer(distGridToArrayCenter, distGridToAllMics, waveNumber)``  fastFuncs.py:859-897

Arithmetic is the GPU's (float32 accumulation, float64 distances and phase reduction); the
results agree with the reference's float64 numba kernels to the parity tolerance stated in
the tests.
"""

from __future__ import annotations

import numpy as np

from . import _lib

_STEER = {'classic': 1, 'inverse': 2, 'true level': 3, 'true location': 4}


def _require(arr, dtypes, ndim, name):
    """Reject what numba's explicit C-contiguous signatures reject (TypeError)."""
    if not isinstance(arr, np.ndarray) or arr.dtype not in dtypes or arr.ndim != ndim or not arr.flags.c_contiguous:
        msg = f'No matching definition for argument {name}: need a C-contiguous {ndim}-d array of {dtypes}'
        raise TypeError(msg)


def calcCSM(csm, SpecAllMics):  # noqa: N802, N803
    """Add one block spectrum to the upper triangle of the CSM, in place (``fastFuncs.py:30-58``)."""
    _require(csm, (np.complex128, np.complex64), 3, 'csm')
    _require(SpecAllMics, (csm.dtype,), 2, 'SpecAllMics')
    nf, nm = SpecAllMics.shape
    if csm.shape != (nf, nm, nm):
        msg = 'csm and SpecAllMics shapes do not match'
        raise ValueError(msg)
    dc = _lib.to_device(csm)
    ds = _lib.to_device(SpecAllMics)
    _lib.check(_lib.lib().abf_calc_csm(_lib.ptr(dc), _lib.ptr(ds), nf, nm, int(csm.dtype == np.complex128),
                                       _lib.stream_handle()), 'calcCSM')
    csm[...] = dc.cpu().numpy()
    return csm


def _transfer_device(r0, rm, k, steer_mode, c128=True):
    r0 = np.ascontiguousarray(r0, dtype=np.float64)
    rm = np.ascontiguousarray(np.atleast_2d(rm), dtype=np.float64)
    ng, nm = rm.shape
    torch = _lib.torch()
    d0 = _lib.to_device(r0)
    dm = _lib.to_device(rm)
    out = torch.empty((ng, nm), dtype=torch.complex128 if c128 else torch.complex64, device=d0.device)
    kk = float(np.asarray(k).ravel()[0])
    _lib.check(_lib.lib().abf_transfer(_lib.ptr(d0), _lib.ptr(dm), ng, nm, kk, _lib.PHASE_MODES['reference'],
                                       steer_mode, int(c128), _lib.ptr(out), _lib.stream_handle()), 'calcTransfer')
    return out.cpu().numpy()


def calcTransfer(distGridToArrayCenter, distGridToAllMics, waveNumber):  # noqa: N802, N803
    """Transfer functions ``exp(-1j f32(k (rm - r0))) r0 / rm`` as complex128 ``[G, M]``
    (``fastFuncs.py:859-897``)."""
    return _transfer_device(distGridToArrayCenter, distGridToAllMics, waveNumber, 0)


def beamformerFreq(steerVecType, boolRemovedDiagOfCSM, normFactor, inputTupleSteer, inputTupleCsm):  # noqa: N802, N803
    """Conventional beamformer for one frequency (``fastFuncs.py:61-214``).

    Returns ``(beamformOutput[G], steerNormalizeOutput[G])`` as float64 numpy arrays.
    """
    is_eig = isinstance(inputTupleCsm, tuple)
    torch = _lib.torch()
    lib = _lib.lib()
    if steerVecType == 'custom':
        steer = np.asarray(inputTupleSteer)
        ng, nm = steer.shape
        dh = _lib.to_device(steer, torch.complex64)
    else:
        st = _STEER[steerVecType]  # KeyError for unknown types, as the reference's dict lookup
        r0, rm, k = inputTupleSteer
        if not isinstance(k, np.ndarray):
            k = np.array([k])
        _require(np.asarray(r0), (np.float64,), 1, 'distGridToArrayCenter')
        _require(np.asarray(rm), (np.float64,), 2, 'distGridToAllMics')
        ng, nm = rm.shape
    dev = _lib.device()
    out = torch.empty(ng, dtype=torch.float32, device=dev)
    onorm = torch.empty(ng, dtype=torch.float32, device=dev)
    sp = _lib.stream_handle()
    if is_eig:
        eva, eve = inputTupleCsm
        eva = np.ascontiguousarray(eva, dtype=np.float64)
        eve = np.asarray(eve)
        if eve.shape[0] != nm:
            msg = 'eigenvector length does not match the number of microphones'
            raise ValueError(msg)
        de = _lib.to_device(eva)
        dv = _lib.to_device(np.ascontiguousarray(eve.T), torch.complex64)  # [K][M]
        if steerVecType == 'custom':
            rc = lib.abf_beamform_eig_custom(_lib.ptr(de), _lib.ptr(dv), eva.shape[0], _lib.ptr(dh), ng, nm,
                                             int(bool(boolRemovedDiagOfCSM)), float(normFactor), _lib.ptr(out),
                                             _lib.ptr(onorm), None, 0, sp)
        else:
            d0 = _lib.to_device(r0)
            dm = _lib.to_device(rm)
            rc = lib.abf_beamform_eig(_lib.ptr(de), _lib.ptr(dv), eva.shape[0], _lib.ptr(d0), _lib.ptr(dm), ng, nm,
                                      float(k[0]), st, int(bool(boolRemovedDiagOfCSM)), float(normFactor),
                                      _lib.PHASE_MODES['reference'], _lib.ptr(out), _lib.ptr(onorm), None, 0, sp)
    else:
        csm = inputTupleCsm
        if steerVecType != 'custom':
            _require(csm, (np.complex128,), 2, 'csm')
        if csm.shape != (nm, nm):
            msg = 'csm shape does not match the number of microphones'
            raise ValueError(msg)
        dc = _lib.to_device(np.asarray(csm), torch.complex64)
        if steerVecType == 'custom':
            rc = lib.abf_beamform_custom(_lib.ptr(dc), _lib.ptr(dh), ng, nm, int(bool(boolRemovedDiagOfCSM)),
                                         float(normFactor), _lib.ptr(out), _lib.ptr(onorm), None, 0, sp)
        else:
            d0 = _lib.to_device(r0)
            dm = _lib.to_device(rm)
            ws = _lib.workspace(lib.abf_beamform_table_workspace_size(nm))
            rc = lib.abf_beamform_table(_lib.ptr(dc), _lib.ptr(d0), _lib.ptr(dm), ng, nm, float(k[0]), st,
                                        int(bool(boolRemovedDiagOfCSM)), float(normFactor),
                                        _lib.PHASE_MODES['reference'], _lib.ptr(out), _lib.ptr(onorm), _lib.ptr(ws),
                                        ws.numel(), sp)
    _lib.check(rc, 'beamformerFreq')
    return out.cpu().numpy().astype(np.float64), onorm.cpu().numpy().astype(np.float64)
