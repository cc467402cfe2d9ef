"""Pins of the gridless orthogonal beamformer's restatement (``oracle/gridless.py``,
fbeamform.py:2337-2410) by a closed form.

With true-location steering h = a / sqrt(M sum|a|^2) and no diagonal removal, the eigenvalue-form
map of a unit eigenvector v is |v^H h(x)|^2 <= 1/M (Cauchy-Schwarz), with equality only where a(x)
is parallel to v.  For a rank-one CSM p a_s a_s^H the largest eigenvector is a_s / |a_s|, so the
search must end on the source and report p |a_s|^2 / M.
"""

import numpy as np
import pytest

from oracle import beamform as B
from oracle import gridless as GL


def _mics(nm=32, seed=3):
    rng = np.random.default_rng(seed)
    r = 0.5 * np.sqrt(rng.uniform(0, 1, nm))
    t = rng.uniform(0, 2 * np.pi, nm)
    return np.stack([r * np.cos(t), r * np.sin(t), np.zeros(nm)])


def _rank_one(mpos, spos, p, freq):
    rm = B.dist_mat(spos[:, None], mpos)[0]
    a = np.exp(-1j * 2 * np.pi * freq / 343.0 * rm) / rm
    return p * np.outer(a, a.conj()), p * np.sum(np.abs(a) ** 2)


BOUNDS = [(-0.4, 0.4), (-0.4, 0.4), (0.2, 0.8)]


def test_lone_monopole_is_found_with_its_level():
    mpos = _mics()
    spos = np.array([0.12, -0.07, 0.45])
    csm, lam = _rank_one(mpos, spos, 2.0, 6000.0)
    pos, lev = GL.gridless_orth(csm, 6000.0, mpos, 'true location', False, bounds=BOUNDS, shgo_opts={'n': 64})
    assert np.abs(pos[:2, 0] - spos[:2]).max() <= 1e-3, pos[:, 0]
    assert abs(pos[2, 0] - spos[2]) <= 2e-2, pos[:, 0]
    assert lev[0] == pytest.approx(lam / 32, rel=1e-10)


def test_two_incoherent_sources_largest_eigenvector_first():
    """Two incoherent monopoles far apart: the two largest eigenvectors are (nearly) the two
    steering directions, the stronger source first, so n = 2 finds both in that order."""
    mpos = _mics(64, 5)
    s1 = np.array([-0.2, 0.15, 0.5])
    s2 = np.array([0.22, -0.18, 0.5])
    c1, _ = _rank_one(mpos, s1, 4.0, 8000.0)
    c2, _ = _rank_one(mpos, s2, 1.0, 8000.0)
    # default shgo options (256 Sobol points); fewer land on a sidelobe of the stronger source
    pos, lev = GL.gridless_orth(c1 + c2, 8000.0, mpos, 'true location', False, eva_list=[-1, -2], n=2,
                                bounds=[(-0.4, 0.4), (-0.4, 0.4), (0.45, 0.55)])
    assert lev[0] > lev[1] > 0
    assert np.abs(pos[:2, 0] - s1[:2]).max() <= 1e-3, pos
    assert np.abs(pos[:2, 1] - s2[:2]).max() <= 1e-3, pos
