"""Float64 numpy restatement of Acoular's point spread function and DAMAS solver.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

* ``psf``: ``calcPointSpreadFunction`` and its four formulations (``fastFuncs.py:563-801``),
  ``psf[g, s] = |sum_m exp(-1j f32(k (rm[g, m] - rm[s, m]))) w(g, s, m)|^2 * norm(g, s)`` with the
  phase difference rounded to float32 before cos/sin are taken in float32 (``:671-673``), the
  per-formulation weights and normalisations (classic ``:657-677``, inverse ``:693-717``, true
  level ``:733-759``, true location ``:775-801``).  Vectorised over grid points; one source
  column at a time.
* ``clean_sc``: ``BeamformerCleansc._calc`` (``fbeamform.py:1449-1492``), the reference's
  iteration restated on the oracle's beamformer forms.
* ``clean``: the iteration of ``BeamformerClean._calc`` (``fbeamform.py:1578-1598``) on one
  dirty map, with PSF columns from ``psf``.
* ``damas_plus``: one bin of ``BeamformerDamasPlus._calc`` (``fbeamform.py:1286-1331``), the
  dirty map scaled by ``unit_mult`` and handed to scipy's ``nnls``/``linprog`` or
  scikit-learn's ``LassoLars``/``OrthogonalMatchingPursuitCV`` (normalised columns), the
  third-party solvers the reference calls.  Pinned by exact recovery of a sparse non-negative
  source distribution from its own blurred map (``tests/test_oracle_deconv.py``).
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


def clean_sc(csm, freqs, gpos, mpos, steer_type='true level', r_diag=True, r_diag_norm=0.0, n_iter=0, damp=0.6,
             stopn=3, c=343.0, ref=(0.0, 0.0, 0.0), phase_mode='reference'):
    """``BeamformerCleansc._calc`` (``fbeamform.py:1449-1492``) over every slice of ``csm``:
    float32 result maps ``[F, G]``.  The dirty map and the maps of the removed coherent
    components use the oracle's full-CSM and eigenvalue forms (``fastFuncs.py:238-311,
    335-411``); the steering vector at the peak is the transfer function with the steering
    normalisation of ``SteeringVector.steer_vector`` (``fbeamform.py:145-153, 237-258``)."""
    from .beamform import (calc_transfer, dist_mat, freq_beamformer_eig, freq_beamformer_full_csm, sig_loss_norm,
                           steer_r0, steer_vector)

    csm = np.asarray(csm)
    nm = csm.shape[1]
    r0 = steer_r0(gpos, ref if np.isscalar(ref) else np.asarray(ref, dtype=float))
    rm = dist_mat(gpos, mpos)
    normfactor = sig_loss_norm(r_diag, r_diag_norm, nm)
    J = nm * 2 if not n_iter else n_iter
    out = np.zeros((csm.shape[0], rm.shape[0]), dtype=np.float32)
    for i, fr in enumerate(freqs):
        k = 2 * np.pi * fr / c
        C = np.array(csm[i], dtype='complex128', copy=True)
        h = freq_beamformer_full_csm(C, r0, rm, k, normfactor, r_diag, steer_type, phase_mode)[0]
        result = np.zeros(rm.shape[0], 'f')
        powers = np.zeros(J, 'd')
        for j in range(J):
            xi_max = h.argmax()
            powers[j] = hmax = h[xi_max]
            result[xi_max] += damp * hmax
            if j > stopn and hmax > powers[j - stopn]:
                break
            wmax = steer_vector(calc_transfer(r0[xi_max:xi_max + 1], rm[xi_max:xi_max + 1], k, phase_mode),
                                steer_type) * np.sqrt(normfactor)
            wmax = wmax[0].conj()
            hh = wmax.copy()
            D1 = np.dot(C.T - np.diag(np.diag(C)), wmax) / hmax
            ww = wmax.conj() * wmax
            for _ in range(20):
                H = hh.conj() * hh
                hh = (D1 + H * wmax) / np.sqrt(1 + np.dot(ww, H))
            hh = hh[:, np.newaxis]
            csm1 = hmax * (hh * hh.conj().T)
            h1 = freq_beamformer_eig(np.array((hmax,)), hh.conj(), r0, rm, k, normfactor, r_diag, steer_type,
                                     phase_mode)[0]
            h -= damp * h1
            C -= damp * csm1.T
        out[i] = result
    return out


def clean(dirty, steer_type, r0, rm, k, n_iter=100, damp=0.6, psf_precision='float64', phase_mode='reference'):
    """The Hoegbom CLEAN loop of ``BeamformerClean._calc`` (``fbeamform.py:1578-1598``) on one
    float64 dirty map (already clipped when r_diag): returns the float64 clean map.  At least one
    iteration always runs; the loop stops after an iteration that does not lower the sum of
    absolute values, after ``n_iter`` iterations, or when the map has no positive value left."""
    dirty = np.array(dirty, dtype=np.float64, copy=True)
    gs = dirty.shape[0]
    cl = np.zeros(gs, dtype=dirty.dtype)
    i_iter = 0
    flag = True
    while flag:
        dirty_sum = abs(dirty).sum(0)
        next_max = dirty.argmax(0)
        col = psf(steer_type, r0, rm, k, [next_max], phase_mode, dtype=psf_precision).reshape(gs)
        new_amp = damp * dirty[next_max]
        cl[next_max] += new_amp
        dirty -= col * new_amp
        i_iter += 1
        flag = dirty_sum > abs(dirty).sum(0) and i_iter < n_iter and max(dirty) > 0
    return cl


def damas_plus(A, y, method='NNLS', alpha=0.0, n_iter=500, unit=1e9):
    """``fbeamform.py:1286-1331`` for one bin: x >= 0 with A x ~ y, solved on y * unit."""
    yu = np.asarray(y, dtype=np.float64) * unit
    if method == 'NNLS':
        from scipy.optimize import nnls

        return nnls(A, yu)[0] / unit
    if method == 'LP':
        from scipy.optimize import linprog

        return linprog(c=-A.sum(1), A_ub=A, b_ub=yu).x / unit
    from sklearn.linear_model import LassoLars, OrthogonalMatchingPursuitCV

    model = (LassoLars(alpha=alpha * unit, max_iter=n_iter, positive=True) if method == 'LassoLars'
             else OrthogonalMatchingPursuitCV())
    nrm = np.sqrt((A * A).sum(0))
    model.fit(A / nrm, yu)
    return model.coef_ / nrm / unit
