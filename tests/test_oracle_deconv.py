"""Pins of the oracle's point spread function and DAMAS solver (oracle/deconv.py, oracle/cpu_ref.c).

* PSF (fastFuncs.py:563-801): closed forms on the diagonal (a source seen at its own grid point:
  1 for 'inverse' and 'true level', r0^2 sum(1/rm^2) / M for 'true location',
  (r0/M)^2 (sum 1/rm)^2 for 'classic'), and, with the exact float64 phase, equality with the
  already pinned beamformer oracle applied to the CSM a_s a_s^H of a unit monopole at the
  source point (the identity the reference's docstring states, fastFuncs.py:592-601); the C
  loop nest equals the numpy restatement.
* DAMAS Gauss-Seidel (fastFuncs.py:820-855): one sweep on the identity is the clipped dirty
  map; a hand-worked 2 x 2 sweep fixes the row order (updated values are used at once) and the
  relaxation formula; many sweeps on a diagonally dominant system reach numpy.linalg.solve; the
  C loop nest equals the numpy restatement.
"""

import numpy as np
import pytest

from oracle import beamform as B
from oracle import cref
from oracle import deconv as D

STEERS = ['classic', 'inverse', 'true level', 'true location']


def _geometry(nm=16, n=7, seed=3):
    rng = np.random.default_rng(seed)
    r = np.sqrt(rng.uniform(0, 1, nm)) * 0.5
    ph = rng.uniform(0, 2 * np.pi, nm)
    mpos = np.stack([r * np.cos(ph), r * np.sin(ph), np.zeros(nm)])
    x = np.linspace(-0.3, 0.3, n)
    gpos = np.stack([np.repeat(x, n), np.tile(x, n), np.full(n * n, 0.6)])
    return gpos, mpos


@pytest.mark.parametrize('steer', STEERS)
def test_psf_diagonal_closed_forms(steer):
    gpos, mpos = _geometry()
    r0 = B.steer_r0(gpos, np.zeros(3))
    rm = B.dist_mat(gpos, mpos)
    nm = mpos.shape[1]
    ind = np.arange(gpos.shape[1])
    p = D.psf(steer, r0, rm, 2 * np.pi * 3000 / 343.0, ind)
    d = np.diag(p)
    help_g = np.sum(1.0 / rm**2, axis=1)
    want = {'inverse': np.ones_like(d), 'true level': np.ones_like(d),
            'true location': r0**2 * help_g / nm,
            'classic': (r0 / nm) ** 2 * np.sum(1.0 / rm, axis=1) ** 2}[steer]
    np.testing.assert_allclose(d, want, rtol=1e-13)


@pytest.mark.parametrize('steer', STEERS)
def test_psf_equals_beamformer_of_point_source(steer):
    gpos, mpos = _geometry()
    r0 = B.steer_r0(gpos, np.zeros(3))
    rm = B.dist_mat(gpos, mpos)
    k = 2 * np.pi * 5000 / 343.0
    ind = np.array([0, 11, 24, 48])
    p = D.psf(steer, r0, rm, k, ind, phase_mode='fp64')
    for c, s in enumerate(ind):
        a = B.calc_transfer(r0[s:s + 1], rm[s:s + 1], k, phase_mode='fp64')[0]
        csm = np.outer(a, a.conj())
        bf = B.freq_beamformer_full_csm(csm, r0, rm, k, 1.0, False, steer, phase_mode='fp64')[0]
        np.testing.assert_allclose(p[:, c], bf, rtol=1e-10, atol=1e-13 * np.abs(bf).max())


@pytest.mark.parametrize('steer', STEERS)
def test_psf_c_loop_equals_numpy(steer):
    gpos, mpos = _geometry(nm=24, n=9, seed=8)
    r0 = B.steer_r0(gpos, np.zeros(3))
    rm = B.dist_mat(gpos, mpos)
    k = 2 * np.pi * 9000 / 343.0
    ind = np.array([3, 40, 80, 5])
    # libm cosf/sinf and numpy's correctly rounded float32 cos/sin may differ by one float32 ulp
    a, b = cref.psf(steer, r0, rm, k, ind), D.psf(steer, r0, rm, k, ind)
    assert np.abs(a - b).max() <= 1e-6 * np.abs(b).max()


def test_damas_identity_one_sweep_is_clipped_dirty_map():
    y = np.array([0.5, -0.2, 1.5, 0.0])
    x = y.copy()
    D.damas_gauss_seidel(np.eye(4), y, 1, 1.0, x)
    np.testing.assert_array_equal(x, np.maximum(y, 0))


def test_damas_sweep_order_and_relaxation():
    A = np.array([[1.0, 0.5], [0.25, 1.0]])
    y = np.array([1.0, 2.0])
    x0 = np.array([0.4, 0.8])
    w = 0.7
    # row 0 uses the old x1, row 1 already the new x0 (Gauss-Seidel, not Jacobi)
    n0 = max(0.0, (1 - w) * x0[0] + w * (y[0] - 0.5 * x0[1]))
    n1 = max(0.0, (1 - w) * x0[1] + w * (y[1] - 0.25 * n0))
    x = x0.copy()
    D.damas_gauss_seidel(A, y, 1, w, x)
    np.testing.assert_allclose(x, [n0, n1], rtol=1e-15)
    xc = x0.copy()[None]
    cref.damas_gauss_seidel(A[None], y[None], xc, 1, w)
    np.testing.assert_allclose(xc[0], [n0, n1], rtol=1e-15)


def test_damas_converges_to_the_solution():
    rng = np.random.default_rng(4)
    n = 12
    A = rng.uniform(0, 0.05, (n, n))
    np.fill_diagonal(A, 1.0)
    xs = rng.uniform(0.1, 1.0, n)
    y = A @ xs
    x = y.copy()
    D.damas_gauss_seidel(A, y, 200, 1.0, x)
    np.testing.assert_allclose(x, np.linalg.solve(A, y), rtol=1e-12)


def test_damas_c_loop_equals_numpy():
    rng = np.random.default_rng(6)
    nf, n = 3, 30
    A = rng.uniform(0, 0.2, (nf, n, n))
    for f in range(nf):
        np.fill_diagonal(A[f], 1.0)
    y = rng.uniform(-0.2, 1.0, (nf, n))
    xc = y.copy()
    cref.damas_gauss_seidel(A, y, xc, 25, 0.9)
    for f in range(nf):
        xn = y[f].copy()
        D.damas_gauss_seidel(A[f], y[f], 25, 0.9, xn)
        np.testing.assert_allclose(xc[f], xn, rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize('steer_type', ['classic', 'inverse', 'true level', 'true location'])
def test_clean_sc_single_source_is_removed_exactly(steer_type):
    """One coherent monopole exactly on a grid point, beamformed with the diagonal removed: the
    source component that CLEAN-SC solves for (``fbeamform.py:1474-1480``, Sijtsma 2007) is then
    exactly the source's cross-spectral matrix, so each iteration finds the same peak and removes
    the fraction ``damp`` of it; after n iterations the map holds h0 (1 - (1 - damp)^n) at the
    source (h0 the dirty map's peak) and nothing elsewhere.  A transposed or unconjugated CSM in
    the component estimate, or a missing normalisation, leaves a residue of several per cent."""
    from acoular_amd import synthetic
    from oracle.beamform import dist_mat, freq_beamformer_full_csm, sig_loss_norm, steer_r0

    mpos = synthetic.disk_array(24, 1.0, 3)
    x = np.linspace(-0.3, 0.3, 13)
    gpos = np.stack([np.repeat(x, 13), np.tile(x, 13), np.full(169, 0.5)])
    freqs = np.array([5000.0])
    csm = synthetic.analytic_csm(mpos, gpos[:, 60:61], np.array([2.0]), freqs, dtype=np.complex128)
    h0 = freq_beamformer_full_csm(csm[0], steer_r0(gpos, np.zeros(3)), dist_mat(gpos, mpos), 2 * np.pi * 5000.0 / 343.0,
                                  sig_loss_norm(True, 0.0, 24), True, steer_type, 'fp64')[0]
    assert np.argmax(h0) == 60
    for n in (1, 2, 7, 30):
        out = D.clean_sc(csm, freqs, gpos, mpos, steer_type, r_diag=True, n_iter=n, damp=0.6, phase_mode='fp64')[0]
        np.testing.assert_allclose(out[60], h0[60] * (1 - 0.4**n), rtol=2e-6)
        assert np.count_nonzero(out) == 1


@pytest.mark.parametrize('n_iter', [1, 3, 40])
def test_clean_on_a_psf_shaped_map(n_iter):
    """Classic steering, a dirty map that is p times the PSF column of grid point s (the map of a
    monopole at s, whose peak is at s by the triangle inequality): every iteration picks s and
    scales the map by (1 - damp c), c = psf(s, s), so after n iterations the clean map holds
    p (1 - (1 - damp c)^n) at s and nothing elsewhere.  A PSF row taken for a column, or an
    amplitude not taken from the current map, breaks the geometric series."""
    from acoular_amd import synthetic

    mpos = synthetic.disk_array(32, 1.0, 8)
    x = np.linspace(-0.3, 0.3, 11)
    gpos = np.stack([np.repeat(x, 11), np.tile(x, 11), np.full(121, 0.6)])
    r0 = B.steer_r0(gpos, np.zeros(3))
    rm = B.dist_mat(gpos, mpos)
    k = 2 * np.pi * 4000.0 / 343.0
    s, p = 47, 3.0
    col = D.psf('classic', r0, rm, k, [s])[:, 0]
    assert np.argmax(col) == s
    c = col[s]
    out = D.clean(p * col, 'classic', r0, rm, k, n_iter=n_iter, damp=0.6)
    np.testing.assert_allclose(out[s], p * (1 - (1 - 0.6 * c) ** n_iter), rtol=1e-12)
    assert np.count_nonzero(out) == 1


def test_clean_stop_rules():
    """A map with no positive value runs exactly one iteration (the loop's first pass is
    unconditional); n_iter 0 behaves as 1."""
    from acoular_amd import synthetic

    mpos = synthetic.disk_array(8, 1.0, 8)
    gpos = np.stack([np.linspace(-0.2, 0.2, 5), np.zeros(5), np.full(5, 0.5)])
    r0 = B.steer_r0(gpos, np.zeros(3))
    rm = B.dist_mat(gpos, mpos)
    dirty = -np.arange(1.0, 6.0)
    out = D.clean(dirty, 'true level', r0, rm, 1.0, n_iter=50, damp=0.5)
    assert np.count_nonzero(out) == 1 and out[0] == 0.5 * dirty[0]
    out = D.clean(np.arange(1.0, 6.0), 'classic', r0, rm, 1.0, n_iter=0, damp=0.5)
    assert np.count_nonzero(out) == 1 and out[4] == 2.5


def _recovery_case(sym):
    """A real PSF (true level, 32 mics, 5 x 5 points, 8 kHz) and a two-source map made from it."""
    mpos = np.random.default_rng(4).uniform(-0.5, 0.5, (3, 32))
    mpos[2] = 0.0
    xs, ys = np.meshgrid(np.linspace(-0.2, 0.2, 5), np.linspace(-0.2, 0.2, 5), indexing='ij')
    gpos = np.stack([xs.ravel(), ys.ravel(), np.full(25, 0.5)])
    rm = B.dist_mat(gpos, mpos)
    A = D.psf('true level', B.steer_r0(gpos, np.zeros(3)), rm, 2 * np.pi * 8000 / 343.0, np.arange(25))
    if sym:
        A = 0.5 * (A + A.T)
    x = np.zeros(25)
    x[[6, 18]] = [2.0e-9, 0.5e-9]
    return A, x, A @ x


@pytest.mark.parametrize('method', ['NNLS', 'LassoLars', 'OMPCV'])
def test_damas_plus_recovers_sparse_sources(method):
    """A non-negative source distribution is recovered exactly from its own blurred map by every
    least-squares method (the PSF is non-singular); the transposed PSF (not symmetric) misses it."""
    A, x, y = _recovery_case(False)
    assert np.abs(A - A.T).max() > 1e-3 * A.max()
    got = D.damas_plus(A, y, method)
    assert np.abs(got - x).max() <= 1e-6 * x.max(), (method, got[[6, 18]])
    if method == 'NNLS':
        assert np.abs(D.damas_plus(A.T, y, method) - x).max() > 1e-3 * x.max()


def test_damas_plus_lp_optimum():
    """With a symmetric non-singular PSF, maximising the summed map under A x <= y, x >= 0 has the
    unique optimum x = A^-1 y (1'A x <= 1'y, equality only at A x = y)."""
    A, x, y = _recovery_case(True)
    got = D.damas_plus(A, y, 'LP')
    assert np.abs(got - x).max() <= 1e-6 * x.max(), got[[6, 18]]
