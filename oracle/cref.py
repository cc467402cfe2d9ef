"""ctypes binding of ``oracle/cpu_ref.c`` (the plain-C restatement).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_CPU_REF_LIB names another build of cpu_ref.c (the sanitizer build used by the CPU tests)
_LIB = os.environ.get('ORACLE_CPU_REF_LIB') or os.path.join(_HERE, 'libcpu_ref.so')
_lib = None


def build() -> None:
    subprocess.run(['make', '-C', _HERE, '-B', 'libcpu_ref.so'], check=True, capture_output=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = ctypes.CDLL(_LIB)
        dp = ctypes.POINTER(ctypes.c_double)
        _lib.cref_beamformer_full_csm.argtypes = [dp, dp, dp, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                                  ctypes.c_int, ctypes.c_int64, ctypes.c_int64, dp, dp]
        _lib.cref_calc_csm.argtypes = [dp, dp, ctypes.c_int64, ctypes.c_int64]
        _lib.cref_dist_mat.argtypes = [dp, dp, ctypes.c_int64, ctypes.c_int64, dp]
        _lib.cref_num_threads.restype = ctypes.c_int
        _lib.cref_set_threads.argtypes = [ctypes.c_int]
        ip = ctypes.POINTER(ctypes.c_int64)
        _lib.cref_psf.argtypes = [dp, dp, ctypes.c_int64, ctypes.c_int64, ctypes.c_double, ip, ctypes.c_int64,
                                  ctypes.c_int, dp]
        _lib.cref_damas_gauss_seidel.argtypes = [dp, dp, dp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                                 ctypes.c_double]
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def num_threads() -> int:
    return lib().cref_num_threads()


def set_threads(n: int) -> None:
    """Threads of the OpenMP loops that follow."""
    lib().cref_set_threads(int(n))


def dist_mat(gpos, mpos):
    g = np.ascontiguousarray(gpos, dtype=np.float64)
    m = np.ascontiguousarray(mpos, dtype=np.float64)
    out = np.empty((g.shape[1], m.shape[1]))
    lib().cref_dist_mat(_p(g), _p(m), g.shape[1], m.shape[1], _p(out))
    return out


def beamformer_full_csm(csm, r0, rm, k, norm, r_diag, steer_type):
    """``_freqBeamformer_FullCSM`` (``fastFuncs.py:238-311``) via the C loop nest."""
    c = np.ascontiguousarray(csm, dtype=np.complex128)
    r0 = np.ascontiguousarray(r0, dtype=np.float64)
    rm = np.ascontiguousarray(rm, dtype=np.float64)
    ng, nm = rm.shape
    res = np.empty(ng)
    nst = np.empty(ng)
    lib().cref_beamformer_full_csm(c.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double)), _p(r0),
                                   _p(rm), float(k), float(norm), int(bool(r_diag)), int(steer_type), ng, nm,
                                   _p(res), _p(nst))
    return res, nst


def calc_csm(csm, spec):
    """``calcCSM`` (``fastFuncs.py:30-58``) in place on a complex128 ``csm``."""
    assert csm.dtype == np.complex128 and csm.flags.c_contiguous
    s = np.ascontiguousarray(spec, dtype=np.complex128)
    lib().cref_calc_csm(csm.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                        s.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double)), s.shape[0], s.shape[1])
    return csm


def psf(steer_type, r0, rm, k, ind):
    """``calcPointSpreadFunction`` (``fastFuncs.py:563-801``) via the C loop nest: ``[G, len(ind)]``."""
    r0 = np.ascontiguousarray(r0, dtype=np.float64)
    rm = np.ascontiguousarray(rm, dtype=np.float64)
    ind = np.ascontiguousarray(ind, dtype=np.int64).ravel()
    ng, nm = rm.shape
    out = np.empty((ng, len(ind)))
    st = {'classic': 1, 'inverse': 2, 'true level': 3, 'true location': 4}[steer_type]
    lib().cref_psf(_p(r0), _p(rm), ng, nm, float(k), ind.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), len(ind),
                   st, _p(out))
    return out


def damas_gauss_seidel(A, dirty, x, n_iter, relax):
    """``damasSolverGaussSeidel`` (``fastFuncs.py:820-855``) over ``[F, G, G]`` systems, in place on ``x``."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    y = np.ascontiguousarray(dirty, dtype=np.float64)
    assert x.dtype == np.float64 and x.flags.c_contiguous and x.shape == y.shape
    nf, ng = y.shape
    lib().cref_damas_gauss_seidel(_p(A), _p(y), _p(x), nf, ng, int(n_iter), float(relax))
    return x
