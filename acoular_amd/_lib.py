"""ctypes binding of ``libabf.so`` (the gfx950 kernels) and the device plumbing around it.

PyTorch provides device memory and the current HIP stream; every compute call goes through
the C ABI declared in ``include/abf.h``.  There is no CPU fallback: without the built library
or without a GPU the calls raise.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# ACOULAR_AMD_LIB names another build of the same library (the sanitizer build of the host code
# used by the CPU tests); by default the in-tree libabf.so
LIB_PATH = os.environ.get('ACOULAR_AMD_LIB') or os.path.join(_HERE, 'libabf.so')

ABF_OK = 0
ABF_ERR_INVALID = -1
ABF_ERR_UNSUPPORTED = -2
ABF_ERR_HIP = -3
ABF_ERR_WORKSPACE = -4

FORMS = {'plain_f32': 0, 'mfma_bf16x3': 1, 'mfma_f32': 2, 'mfma_f16x2': 3}
PHASE_MODES = {'reference': 0, 'fp64': 1}
FAST_TRIG = 0x100
# enum abf_mfma_kernel (include/abf.h) values named by the host code
MFMA_KERNEL_STREAMED = 6
MFMA_KERNEL_STREAMED_PERSISTENT = 8
STEER_CODES = {'classic': 1, 'inverse': 2, 'true level': 3, 'true location': 4}

_c_dp = ctypes.c_void_p


class AbfGeometry(ctypes.Structure):
    _fields_ = [
        ('gpos', _c_dp),
        ('mpos', _c_dp),
        ('num_grid', ctypes.c_int64),
        ('num_mics', ctypes.c_int32),
        ('pad_', ctypes.c_int32),
        ('ref', ctypes.c_double * 3),
        ('ref_dist', ctypes.c_double),
    ]


class AbfGridJob(ctypes.Structure):
    _fields_ = [
        ('geom', AbfGeometry),
        ('csm', _c_dp),
        ('wavenumber', _c_dp),
        ('num_freqs', ctypes.c_int32),
        ('steer_type', ctypes.c_int32),
        ('r_diag', ctypes.c_int32),
        ('clip_negative', ctypes.c_int32),
        ('norm_factor', ctypes.c_double),
        ('form', ctypes.c_int32),
        ('phase_mode', ctypes.c_int32),
        ('out', _c_dp),
        ('out_norm', _c_dp),
        ('tile_cols', ctypes.c_int32),
        ('tile_freqs', ctypes.c_int32),
        ('tile_kernel', ctypes.c_int32),
        ('pad1_', ctypes.c_int32),
    ]


# function name -> (restype, argtypes)
_SIGS = {
    'abf_beamform_workspace_size': (ctypes.c_size_t, [ctypes.POINTER(AbfGridJob)]),
    'abf_beamform_plan': (ctypes.c_int, [ctypes.POINTER(AbfGridJob), ctypes.POINTER(ctypes.c_int32),
                                         ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                         ctypes.POINTER(ctypes.c_int64)]),
    'abf_beamform': (ctypes.c_int, [ctypes.POINTER(AbfGridJob), _c_dp, ctypes.c_size_t, _c_dp]),
    'abf_beamform_prepare': (ctypes.c_int, [ctypes.POINTER(AbfGridJob), _c_dp, ctypes.c_size_t, _c_dp]),
    'abf_beamform_run': (ctypes.c_int, [ctypes.POINTER(AbfGridJob), _c_dp, ctypes.c_size_t, _c_dp]),
    'abf_beamform_noalloc': (ctypes.c_int, [ctypes.POINTER(AbfGridJob), _c_dp]),
    'abf_beamform_batch_workspace_size': (
        ctypes.c_size_t, [ctypes.POINTER(AbfGridJob), ctypes.c_int32, ctypes.POINTER(ctypes.c_size_t)]),
    'abf_beamform_batch': (ctypes.c_int, [ctypes.POINTER(AbfGridJob), ctypes.c_int32, _c_dp, ctypes.c_size_t, _c_dp]),
    'abf_beamform_table': (ctypes.c_int, [_c_dp, _c_dp, _c_dp, ctypes.c_int64, ctypes.c_int32, ctypes.c_double,
                                          ctypes.c_int32, ctypes.c_int32, ctypes.c_double, ctypes.c_int32, _c_dp,
                                          _c_dp, _c_dp, ctypes.c_size_t, _c_dp]),
    'abf_beamform_table_workspace_size': (ctypes.c_size_t, [ctypes.c_int32]),
    'abf_beamform_eig': (ctypes.c_int, [_c_dp, _c_dp, ctypes.c_int32, _c_dp, _c_dp, ctypes.c_int64, ctypes.c_int32,
                                        ctypes.c_double, ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                                        ctypes.c_int32, _c_dp, _c_dp, _c_dp, ctypes.c_size_t, _c_dp]),
    'abf_beamform_eig_workspace_size': (ctypes.c_size_t, [ctypes.c_int32, ctypes.c_int32]),
    'abf_beamform_eig_grid': (ctypes.c_int, [ctypes.POINTER(AbfGeometry), _c_dp, ctypes.c_int32, _c_dp, _c_dp,
                                             ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                             ctypes.c_double, ctypes.c_int32, ctypes.c_int32, _c_dp, _c_dp, _c_dp]),
    'abf_csm_eig_workspace_size': (ctypes.c_size_t, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    'abf_csm_eig': (ctypes.c_int, [_c_dp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _c_dp, _c_dp, _c_dp, _c_dp,
                                   ctypes.c_size_t, _c_dp]),
    'abf_beamform_f64_workspace_size': (ctypes.c_size_t, [ctypes.c_int32, ctypes.c_int32]),
    'abf_beamform_f64': (ctypes.c_int, [ctypes.POINTER(AbfGeometry), _c_dp, _c_dp, ctypes.c_int32, ctypes.c_int32,
                                        ctypes.c_int32, ctypes.c_double, ctypes.c_int32, ctypes.c_int32, _c_dp, _c_dp,
                                        _c_dp, ctypes.c_size_t, _c_dp]),
    'abf_point_source_workspace_size': (ctypes.c_size_t, [ctypes.c_int32, ctypes.c_int64]),
    'abf_point_source': (ctypes.c_int, [_c_dp, ctypes.c_int64, _c_dp, _c_dp, ctypes.c_int32, ctypes.c_int32,
                                        ctypes.c_int32, ctypes.c_int64, _c_dp, _c_dp, ctypes.c_size_t, _c_dp]),
    'abf_white_noise': (ctypes.c_int, [_c_dp, ctypes.c_int32, ctypes.c_int64, ctypes.c_double, _c_dp, _c_dp]),
    'abf_mix': (ctypes.c_int, [_c_dp, ctypes.c_double, ctypes.c_int64, ctypes.c_int32, _c_dp, _c_dp]),
    'abf_cleansc_workspace_size': (ctypes.c_size_t, [ctypes.c_int32, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32]),
    'abf_cleansc': (ctypes.c_int, [ctypes.POINTER(AbfGeometry), _c_dp, ctypes.c_int32, _c_dp, _c_dp, ctypes.c_int32,
                                   ctypes.c_int32, ctypes.c_double, ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                                   ctypes.c_int32, _c_dp, _c_dp, ctypes.c_size_t, _c_dp]),
    'abf_resample_workspace_size': (ctypes.c_size_t, [ctypes.c_int64, ctypes.c_int64]),
    'abf_resample': (ctypes.c_int, [_c_dp, ctypes.c_int64, ctypes.c_int64, _c_dp, _c_dp, ctypes.c_size_t, _c_dp]),
    'abf_rfft_workspace_size': (ctypes.c_size_t, [ctypes.c_int32]),
    'abf_rfft_blocks': (ctypes.c_int, [_c_dp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                       ctypes.c_int32, _c_dp, _c_dp, _c_dp, _c_dp, ctypes.c_size_t, _c_dp]),
    'abf_irfft_workspace_size': (ctypes.c_size_t, [ctypes.c_int32]),
    'abf_irfft_blocks': (ctypes.c_int, [_c_dp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _c_dp, _c_dp,
                                        ctypes.c_size_t, _c_dp]),
    'abf_auto_power': (ctypes.c_int, [_c_dp, ctypes.c_int64, ctypes.c_double, _c_dp, _c_dp]),
    'abf_cross_power': (ctypes.c_int, [_c_dp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                                       _c_dp, _c_dp]),
    'abf_clean_workspace_size': (ctypes.c_size_t, [ctypes.c_int32, ctypes.c_int64, ctypes.c_int32]),
    'abf_clean': (ctypes.c_int, [ctypes.POINTER(AbfGeometry), _c_dp, ctypes.c_int32, _c_dp, ctypes.c_int32,
                                 ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_double, _c_dp, _c_dp,
                                 ctypes.c_size_t, _c_dp]),
    'abf_psf': (ctypes.c_int, [ctypes.POINTER(AbfGeometry), _c_dp, ctypes.c_int32, _c_dp, ctypes.c_int64,
                               ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _c_dp, _c_dp]),
    'abf_damas': (ctypes.c_int, [_c_dp, ctypes.c_int32, _c_dp, _c_dp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                 ctypes.c_double, _c_dp]),
    'abf_beamform_custom': (ctypes.c_int, [_c_dp, _c_dp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                           ctypes.c_double, _c_dp, _c_dp, _c_dp, ctypes.c_size_t, _c_dp]),
    'abf_beamform_eig_custom': (ctypes.c_int, [_c_dp, _c_dp, ctypes.c_int32, _c_dp, ctypes.c_int64, ctypes.c_int32,
                                               ctypes.c_int32, ctypes.c_double, _c_dp, _c_dp, _c_dp,
                                               ctypes.c_size_t, _c_dp]),
    'abf_dist_mat': (ctypes.c_int, [_c_dp, _c_dp, ctypes.c_int64, ctypes.c_int32, _c_dp, _c_dp]),
    'abf_transfer': (ctypes.c_int, [_c_dp, _c_dp, ctypes.c_int64, ctypes.c_int32, ctypes.c_double, ctypes.c_int32,
                                    ctypes.c_int32, ctypes.c_int32, _c_dp, _c_dp]),
    'abf_calc_csm': (ctypes.c_int, [_c_dp, _c_dp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _c_dp]),
    'abf_welch_csm_workspace_size': (ctypes.c_size_t, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    'abf_welch_csm': (ctypes.c_int, [_c_dp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                     ctypes.c_int32, _c_dp, ctypes.c_double, _c_dp, ctypes.c_int32, _c_dp,
                                     ctypes.c_size_t, _c_dp]),
    'abf_last_error': (ctypes.c_char_p, []),
    'abf_version': (ctypes.c_char_p, []),
}

_lib = None


class AbfError(RuntimeError):
    """A HIP-level failure inside libabf."""


def lib():
    """The loaded ``libabf.so``; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            msg = f'{LIB_PATH} is missing: build it with `python -m acoular_amd._build` (hipcc, gfx950)'
            raise ImportError(msg)
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(rc: int, what: str = 'libabf') -> None:
    """Map an ``abf_status`` to the exception type the reference raises for that condition."""
    if rc == ABF_OK:
        return
    msg = f'{what}: {lib().abf_last_error().decode()}'
    if rc == ABF_ERR_INVALID:
        raise ValueError(msg)
    if rc == ABF_ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    raise AbfError(msg)


# ---------------------------------------------------------------------------------------------
# device plumbing
# ---------------------------------------------------------------------------------------------
def torch():
    import torch as _torch

    return _torch


def device():
    """The GPU the kernels run on; raises when there is none (no CPU fallback)."""
    t = torch()
    if not t.cuda.is_available():
        msg = 'acoular_amd needs a ROCm GPU (torch.cuda.is_available() is False)'
        raise RuntimeError(msg)
    lib()
    return t.device('cuda', t.cuda.current_device())


def stream_handle():
    return ctypes.c_void_p(torch().cuda.current_stream().cuda_stream)


def ptr(t) -> ctypes.c_void_p | None:
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def to_device(a, dtype=None):
    """numpy array or tensor -> contiguous device tensor (optionally cast).  Host data is staged
    in pinned memory and copied without blocking: the copy is ordered on the current stream like
    a kernel, so the host does not wait for the work queued before it (a pageable copy does)."""
    t = torch()
    dev = device()
    if isinstance(a, t.Tensor):
        x = a.to(dev) if a.is_cuda else a.contiguous().pin_memory().to(dev, non_blocking=True)
    else:
        x = t.from_numpy(np.ascontiguousarray(a)).pin_memory().to(dev, non_blocking=True)
    if dtype is not None and x.dtype != dtype:
        x = x.to(dtype)
    return x.contiguous()


def workspace(nbytes: int):
    """Caller-owned scratch for a libabf call (a uint8 device tensor)."""
    t = torch()
    return t.empty(max(int(nbytes), 1), dtype=t.uint8, device=device())


def header_functions() -> list[str]:
    """Names of the functions ``include/abf.h`` declares (used by the export test)."""
    import re

    path = os.path.join(_HERE, '..', 'include', 'abf.h')
    with open(path) as f:
        text = f.read()
    return sorted(set(re.findall(r'\b(abf_[a-z0-9_]+)\s*\(', text)))
