"""Float64 numpy restatement of Acoular's point spread function and DAMAS solver.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

* ``psf``: ``calcPointSpreadFunction`` and its four formulations (``fastFuncs.py:563-801``),
  ``psf[g, s] = |sum_m exp(-1j f32(k (rm[g, m] - rm[s, m]))) w(g, s, m)|^2 * norm(g, s)`` with the
  phase difference rounded to float32 before cos/sin are taken in float32 (``:671-673``), the
  per-formulation weights and normalisations (classic ``:657-677``, inverse ``:693-717``, true
  level ``:733-759``, true location ``:775-801``).  Vectorised over grid points; one source
  column at a time.
* ``damas_gauss_seidel``: ``damasSolverGaussSeidel`` (``fastFuncs.py:820-855``), the
  row-by-row relaxed Gauss-Seidel sweep with the positivity clip, one scalar loop over rows
  per iteration (the dot product of a row with the current solution is the only vectorised
  step).
"""

from __future__ import annotations

import numpy as np

from .beamform import STEER_TYPES, _unit_phasor


def psf(steer_type, r0, rm, k, ind, phase_mode='reference', dtype='float64') -> np.ndarray:
    """``calcPointSpreadFunction`` (``fastFuncs.py:563-622``): ``[G, len(ind)]``."""
    st = STEER_TYPES[steer_type] if isinstance(steer_type, str) else int(steer_type)
    r0 = np.asarray(r0, dtype=np.float64)
    rm = np.atleast_2d(np.asarray(rm, dtype=np.float64))
    k = float(np.asarray(k).ravel()[0])
    ind = np.asarray(ind, dtype=np.int64).ravel()
    ng, nm = rm.shape
    out = np.zeros((ng, len(ind)), dtype=dtype)
    help_grid = np.sum(1.0 / (rm * rm), axis=1)
    for c, s in enumerate(ind):
        e = _unit_phasor(k * (rm - rm[s][None, :]), phase_mode) / rm[s][None, :]
        if st == 2:
            e = e * rm
        elif st in (3, 4):
            e = e / rm
        sp = np.sum(e, axis=1)
        sq = (sp * sp.conjugate()).real
        if st == 1:
            nf = r0[s] / nm
            res = sq * (nf * nf)
        elif st == 2:
            nf = r0[s] / r0 / nm
            res = sq * (nf * nf)
        elif st == 3:
            nf = r0[s] / r0 / help_grid
            res = sq * (nf * nf)
        else:
            res = sq * (r0[s] * r0[s]) / nm / help_grid
        out[:, c] = res
    return out


def damas_gauss_seidel(A, dirty, n_iter, relax, x) -> np.ndarray:
    """``damasSolverGaussSeidel`` (``fastFuncs.py:820-855``) for one frequency, in place on ``x``."""
    A = np.asarray(A, dtype=np.float64)
    y = np.asarray(dirty, dtype=np.float64)
    relax = float(np.asarray(relax).ravel()[0])
    n = len(y)
    for _ in range(int(np.asarray(n_iter).ravel()[0])):
        for i in range(n):
            s = A[i] @ x - A[i, i] * x[i]
            s = (1 - relax) * x[i] + relax * (y[i] - s)
            x[i] = s if s > 0.0 else 0.0
    return x
