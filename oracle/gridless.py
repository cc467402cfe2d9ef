"""Float64 numpy restatement of Acoular's gridless orthogonal beamformer.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

* ``gridless_orth``: ``BeamformerGridlessOrth._calc`` (``fbeamform.py:2337-2410``) for one bin:
  the CSM's eigendecomposition (``numpy.linalg.eigh``, ``spectra.py:351-364``), the selected
  eigenvectors from the largest down, and for each one scipy's ``shgo`` (default options
  ``fbeamform.py:2366-2372``) minimising minus the eigenvalue-form map of that eigenvector
  (eigenvalue 1) at one point clipped to the bounds (``freq_beamformer_eig``).  Returns the found
  positions and eigenvalue / M.  Pinned by the closed form of a lone monopole with true-location
  steering: |v^H h(x)|^2 <= 1/M with equality only where a(x) is parallel to a_s, so the search
  ends on the source and the level is p |a_s|^2 / M (``tests/test_oracle_gridless.py``).
"""

from __future__ import annotations

import numpy as np

from .beamform import apparent_r, freq_beamformer_eig, sig_loss_norm


def gridless_orth(csm, freq, mpos, steer_type='true level', r_diag=True, eva_list=(-1,), n=1,
                  bounds=((-1.0, 1.0), (-1.0, 1.0), (0.01, 1.0)), shgo_opts=None, c=343.0,
                  phase_mode='reference'):
    from scipy.optimize import shgo

    nm = csm.shape[0]
    ev, evec = np.linalg.eigh(np.asarray(csm, dtype=np.complex128))
    norm = sig_loss_norm(r_diag, 1.0, nm)
    k = 2 * np.pi * freq / c
    opts = {'n': 256, 'iters': 1, 'sampling_method': 'sobol', 'minimizer_kwargs': {'method': 'Nelder-Mead'}}
    opts.update(shgo_opts or {})
    bmin = np.array([min(b) for b in bounds])
    bmax = np.array([max(b) for b in bounds])
    sel = np.unique(np.asarray(eva_list) % nm)[::-1]
    pos = np.zeros((3, n))
    lev = np.zeros(n)
    for j, e in enumerate(sel):
        def func(xy, e=e):
            p = np.clip(xy, bmin, bmax)[:, None]
            r0 = apparent_r(p)
            rm = apparent_r(p, mpos)
            return -freq_beamformer_eig(np.ones(1), evec[:, e:e + 1], np.atleast_1d(r0), np.atleast_2d(rm), k, norm,
                                        r_diag, steer_type, phase_mode)[0][0]

        res = shgo(func, bounds, **opts)
        pos[:, j] = res['x']
        lev[j] = ev[e] / nm
    return pos, lev
