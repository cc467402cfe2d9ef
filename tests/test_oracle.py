"""Pins of the CPU oracle: known-answer vectors from the reference's own doctests and tests,
closed forms, invariants and brute force.  No GPU needed."""

import numpy as np
import pytest

from oracle import beamform as B
from oracle import spectra as S


# ---------------------------------------------------------------------------------------------
# CSM (spectra.py)
# ---------------------------------------------------------------------------------------------
def test_synthetic_ev_known_answer():
    """Doctest of PowerSpectra.synthetic_ev (spectra.py:570-578): window, blocking, float
    num_blocks (7.8125 for 7 blocks) and scaling, through numpy.linalg.eigh."""
    np.random.seed(0)
    data = np.random.rand(1000, 4)
    csm = S.welch_csm(data, 128, 'Hamming', 'None')
    f = S.fftfreq(128, 51200.0)
    ev = np.linalg.eigh(csm)[0]
    # the doctest prints numpy's 8-decimal repr: agreement to half a unit of the last digit
    np.testing.assert_allclose(S.synthetic_ev(ev, f, 5000, 5), [0.00048803, 0.0010141, 0.00234248, 0.00457097],
                               rtol=0, atol=5e-9)
    np.testing.assert_allclose(S.synthetic_ev(ev, f, 5000), [0.00022468, 0.0004589, 0.00088059, 0.00245989],
                               rtol=0, atol=5e-9)
    assert csm.shape == (65, 4, 4)  # spectra.py:391-394


@pytest.mark.parametrize('window,overlap', [('Hanning', '50%'), ('Hamming', 'None'), ('Blackman', '75%')])
def test_welch_of_a_few_bins_equals_full_welch(window, overlap):
    """The direct-DFT Welch of selected bins equals the FFT-based Welch on those bins (another
    transform, the same blocks, window and scaling)."""
    data = np.random.default_rng(4).standard_normal((3000, 5))
    full = S.welch_csm(data, 256, window, overlap)
    bins = [0, 7, 64, 128]
    part = S.welch_csm_bins(data, 256, bins, window, overlap)
    np.testing.assert_allclose(part, full[bins], rtol=0, atol=1e-12 * np.abs(full).max())


def test_num_blocks_is_float():
    assert S.num_blocks(1000, 128, 'None') == 7.8125
    assert len(S.block_starts(1000, 128, 'None')) == 7
    assert S.num_blocks(51200, 128, '50%') == 799.0
    assert len(S.block_starts(51200, 128, '50%')) == 799


def test_fftfreq_known_answer():
    """Doctest of BaseSpectra.fftfreq (spectra.py:146-164)."""
    np.testing.assert_array_equal(S.fftfreq(128, 51200), np.arange(65) * 400.0)


@pytest.mark.parametrize('overlap', ['None', '50%', '75%', '87.5%'])
@pytest.mark.parametrize('n', [100, 128, 129, 1000, 1024, 2047])
def test_block_starts_match_ring_buffer_generator(overlap, n):
    """The closed-form block starts equal what the literal ring-buffer generator of
    spectra.py:171-183 yields when fed block_size chunks (TimeSamples.result)."""
    rng = np.random.default_rng(3)
    d = rng.standard_normal((n, 2))
    blocks = list(S.block_generator(d, 128, overlap))
    starts = S.block_starts(n, 128, overlap)
    assert len(blocks) == len(starts)
    for b, s in zip(blocks, starts):
        np.testing.assert_array_equal(b, d[s : s + 128])


def test_calc_csm_update_bruteforce():
    """calcCSM (fastFuncs.py:30-58) against explicit loops."""
    rng = np.random.default_rng(5)
    spec = rng.standard_normal((3, 5)) + 1j * rng.standard_normal((3, 5))
    csm = np.zeros((3, 5, 5), complex)
    S.calc_csm_update(csm, spec)
    ref = np.zeros_like(csm)
    for f in range(3):
        for col in range(5):
            for row in range(col + 1):
                ref[f, row, col] += np.conj(spec[f, col]) * spec[f, row]
    np.testing.assert_allclose(csm, ref, rtol=0, atol=1e-14)


@pytest.mark.parametrize('window', ['Rectangular', 'Hanning', 'Hamming', 'Blackman', 'Bartlett'])
def test_csm_equals_mean_of_block_cross_spectra(window):
    """Invariant of tests/unittests/test_fprocess.py:79-89: with num_blocks exact, the CSM is
    the mean of the per-block one-sided cross spectra 2 X X^H / (bs sum w^2)."""
    rng = np.random.RandomState(1)
    data = rng.randn(256, 3)
    csm = S.welch_csm(data, 128, window, 'None')
    w = S.WINDOWS[window](128)
    blocks = [np.fft.rfft(data[i : i + 128] * w[:, None], axis=0) for i in (0, 128)]
    per = [2.0 / 128 / np.dot(w, w) * np.einsum('fr,fc->frc', x, x.conj()) for x in blocks]
    np.testing.assert_allclose(csm, np.mean(per, axis=0), rtol=1e-12, atol=1e-18)


def test_csm_hermitian_psd_parseval():
    rng = np.random.default_rng(11)
    sigma = 0.7
    data = rng.standard_normal((64 * 1024, 3)) * sigma
    csm = S.welch_csm(data, 256, 'Hanning', 'None')
    np.testing.assert_allclose(csm, np.conj(csm.transpose(0, 2, 1)), atol=1e-18)
    assert np.all(np.linalg.eigvalsh(csm) > -1e-15)  # no overlap: a sum of rank-1 PSD terms
    # one-sided scaling: interior bins of white noise hold 2 sigma^2 / bs each -> total ~ sigma^2
    total = np.real(np.trace(csm[1:-1], axis1=1, axis2=2)).sum() / 3
    assert abs(total - sigma**2) / sigma**2 < 0.03


# ---------------------------------------------------------------------------------------------
# geometry (grids.py, environments.py)
# ---------------------------------------------------------------------------------------------
def test_rect_grid3d_doctest():
    """RectGrid3D doctest (grids.py:590-596): increment 2 on the default [-1,1]^3 box."""
    pos = B.rect_grid3d_pos(-1, 1, -1, 1, -1, 1, 2)
    np.testing.assert_array_equal(pos, [[-1, -1, -1, -1, 1, 1, 1, 1], [-1, -1, 1, 1, -1, -1, 1, 1],
                                        [-1, 1, -1, 1, -1, 1, -1, 1]])


def test_rect_grid_order_and_steps():
    assert B.rect_grid_steps(-1.0, 1.0, 0.1) == 21
    assert B.rect_grid_steps(-0.2, 0.2, 0.01) == 41
    pos = B.rect_grid_pos(0, 1, 0, 2, 0.5, 1.0)
    # x-major: index = ix * ny + iy (grids.py:351-359)
    np.testing.assert_array_equal(pos[:2].T, [[0, 0], [0, 1], [0, 2], [1, 0], [1, 1], [1, 2]])
    assert np.all(pos[2] == 0.5)


def test_dist_mat_bruteforce():
    rng = np.random.default_rng(2)
    g, m = rng.standard_normal((3, 7)), rng.standard_normal((3, 4))
    ref = np.array([[np.linalg.norm(g[:, i] - m[:, j]) for j in range(4)] for i in range(7)])
    np.testing.assert_allclose(B.dist_mat(g, m), ref, rtol=1e-15)
    assert B.apparent_r(g).shape == (7,)


# ---------------------------------------------------------------------------------------------
# steering and beamforming (fastFuncs.py, fbeamform.py)
# ---------------------------------------------------------------------------------------------
def _scene(m=12, seed=0):
    rng = np.random.default_rng(seed)
    mpos = np.r_[rng.uniform(-0.5, 0.5, (2, m)), np.zeros((1, m))]
    gpos = B.rect_grid_pos(-0.3, 0.3, -0.3, 0.3, 0.5, 0.15)
    return mpos, gpos


def test_transfer_closed_form_and_steer_invariants():
    mpos, gpos = _scene()
    r0 = B.steer_r0(gpos, np.zeros(3))
    rm = B.dist_mat(gpos, mpos)
    k = 2 * np.pi * 3000 / 343.0
    a = B.calc_transfer(r0, rm, k, 'fp64')
    np.testing.assert_allclose(np.abs(a), r0[:, None] / rm, rtol=1e-14)
    np.testing.assert_allclose(np.angle(a * np.exp(1j * k * (rm - r0[:, None]))), 0, atol=1e-9)
    nm = mpos.shape[1]
    np.testing.assert_allclose(np.abs(B.steer_vector(a, 'classic')), 1.0 / nm, rtol=1e-14)
    h = B.steer_vector(a, 'true level')
    np.testing.assert_allclose(np.sum(np.abs(h) ** 2, 1) * np.sum(np.abs(a) ** 2, 1), 1.0, rtol=1e-13)
    h = B.steer_vector(a, 'true location')
    np.testing.assert_allclose(np.sum(np.abs(h) ** 2, 1), 1.0 / nm, rtol=1e-13)
    h = B.steer_vector(a, 'inverse')
    np.testing.assert_allclose(h * a.conj(), 1.0 / nm, rtol=1e-13)


@pytest.mark.parametrize('steer', ['classic', 'inverse', 'true level', 'true location'])
def test_point_source_closed_forms(steer):
    """C = p a a^H of a source at a grid point, full CSM: true level and inverse give exactly
    p, true location p sum|a|^2/M, classic p (sum|a|)^2/M^2; with the diagonal removed
    p (|h^H a|^2 - sum |h|^2 |a|^2) M/(M-1)."""
    mpos, gpos = _scene()
    gi = 7
    nm = mpos.shape[1]
    r0 = B.steer_r0(gpos, np.zeros(3))
    rm = B.dist_mat(gpos, mpos)
    k = 2 * np.pi * 2500 / 343.0
    a = B.calc_transfer(r0[gi : gi + 1], rm[gi : gi + 1], k, 'fp64')[0]
    p = 0.37
    C = p * np.outer(a, a.conj())
    res = B.freq_beamformer_full_csm(C, r0, rm, k, 1.0, False, steer, 'fp64')[0][gi]
    expect = {'classic': p * np.sum(np.abs(a)) ** 2 / nm**2, 'inverse': p, 'true level': p,
              'true location': p * np.sum(np.abs(a) ** 2) / nm}[steer]
    np.testing.assert_allclose(res, expect, rtol=1e-12)
    h = B.steer_vector(a[None], steer)[0]
    norm = nm / (nm - 1)
    res_rd = B.freq_beamformer_full_csm(C, r0, rm, k, norm, True, steer, 'fp64')[0][gi]
    expect_rd = p * (np.abs(np.vdot(h, a)) ** 2 - np.sum(np.abs(h) ** 2 * np.abs(a) ** 2)) * norm
    np.testing.assert_allclose(res_rd, expect_rd, rtol=1e-11)


@pytest.mark.parametrize('steer', ['classic', 'inverse', 'true level', 'true location'])
@pytest.mark.parametrize('r_diag', [False, True])
def test_full_csm_equals_explicit_quadratic_form(steer, r_diag):
    """The triangle-and-normalisation path equals h^H C h with the explicit steering vector of
    fbeamform.py:145-153 over the FULL Hermitian matrix (brute force, every i, j)."""
    mpos, gpos = _scene(9, 4)
    rng = np.random.default_rng(8)
    X = rng.standard_normal((9, 3)) + 1j * rng.standard_normal((9, 3))
    C = X @ X.conj().T
    r0 = B.steer_r0(gpos, np.zeros(3))
    rm = B.dist_mat(gpos, mpos)
    k = 2 * np.pi * 4000 / 343.0
    h = B.steer_vector(B.calc_transfer(r0, rm, k, 'fp64'), steer)
    Cd = C - np.diag(np.diag(C)) if r_diag else C
    ref = np.array([sum(np.conj(h[g, i]) * Cd[i, j] * h[g, j] for i in range(9) for j in range(9)).real
                    for g in range(h.shape[0])])
    got = B.freq_beamformer_full_csm(C, r0, rm, k, 1.0, r_diag, steer, 'fp64')[0]
    np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())
    got_c = B.freq_beamformer_custom(C, h, 1.0, r_diag)[0]
    np.testing.assert_allclose(got_c, ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())


def test_upper_triangle_only_is_read():
    """Only the upper triangle and the diagonal of C enter (fastFuncs.py:277-291)."""
    mpos, gpos = _scene(6, 1)
    rng = np.random.default_rng(1)
    C = rng.standard_normal((6, 6)) + 1j * rng.standard_normal((6, 6))
    junk = C.copy()
    junk[np.tril_indices(6, -1)] = 99.0 + 7j
    r0 = B.steer_r0(gpos, np.zeros(3))
    rm = B.dist_mat(gpos, mpos)
    a = B.freq_beamformer_full_csm(C, r0, rm, 30.0, 1.0, False, 'classic')[0]
    b = B.freq_beamformer_full_csm(junk, r0, rm, 30.0, 1.0, False, 'classic')[0]
    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize('steer', ['classic', 'inverse', 'true level', 'true location'])
@pytest.mark.parametrize('r_diag', [False, True])
def test_eig_form_equals_full_csm(steer, r_diag):
    """sum_k l_k |v_k^H h|^2 over all eigenpairs equals h^H C h; the r_diag correction equals
    removing the diagonal (fastFuncs.py:335-411)."""
    mpos, gpos = _scene(8, 2)
    rng = np.random.default_rng(9)
    X = rng.standard_normal((8, 4)) + 1j * rng.standard_normal((8, 4))
    C = X @ X.conj().T
    lam, V = np.linalg.eigh(C)
    r0 = B.steer_r0(gpos, np.zeros(3))
    rm = B.dist_mat(gpos, mpos)
    full = B.freq_beamformer_full_csm(C, r0, rm, 40.0, 1.1, r_diag, steer)
    eig = B.freq_beamformer_eig(lam, V, r0, rm, 40.0, 1.1, r_diag, steer)
    np.testing.assert_allclose(eig[0], full[0], rtol=1e-9, atol=1e-12 * np.abs(full[0]).max())
    np.testing.assert_allclose(eig[1], full[1], rtol=1e-14)


@pytest.mark.parametrize('r_diag', [False, True])
def test_eig_custom_equals_explicit_quadratic_form(r_diag):
    """The eigenvalue form with explicit steering vectors (fastFuncs.py:479-559) over all
    eigenpairs equals h^H C h summed by brute force over every (i, j) of the full Hermitian CSM;
    with r_diag its correction sum_k l_k sum_m |v_mk|^2 |h_m|^2 equals removing the diagonal of
    C explicitly.  Over a subset of eigenpairs it equals h^H C_K h for the matrix rebuilt from
    that subset.  A conjugated eigenvector product (v_k^T h instead of v_k^H h) or a missing
    eigenvalue weight fails these."""
    rng = np.random.default_rng(21)
    M, G = 7, 5
    X = rng.standard_normal((M, 3)) + 1j * rng.standard_normal((M, 3))
    C = X @ X.conj().T + np.diag(rng.uniform(0.1, 0.5, M))
    h = rng.standard_normal((G, M)) + 1j * rng.standard_normal((G, M))
    lam, V = np.linalg.eigh(C)

    def brute(Cm):
        Cd = Cm - np.diag(np.diag(Cm)) if r_diag else Cm
        return np.array([sum(np.conj(h[g, i]) * Cd[i, j] * h[g, j] for i in range(M) for j in range(M)).real
                         for g in range(G)])

    got, nst = B.freq_beamformer_eig_custom(lam, V, h, 1.3, r_diag)
    np.testing.assert_allclose(got, 1.3 * brute(C), rtol=1e-10, atol=1e-12 * np.abs(brute(C)).max())
    np.testing.assert_allclose(nst, np.sum(np.abs(h) ** 2, axis=1), rtol=1e-14)
    full, _ = B.freq_beamformer_custom(C, h, 1.3, r_diag)
    np.testing.assert_allclose(got, full, rtol=1e-10, atol=1e-12 * np.abs(full).max())
    sub = [2, 4, 6]
    Ck = (V[:, sub] * lam[sub]) @ V[:, sub].conj().T
    got_k, _ = B.freq_beamformer_eig_custom(lam[sub], V[:, sub], h, 1.0, r_diag)
    np.testing.assert_allclose(got_k, brute(Ck), rtol=1e-10, atol=1e-12 * np.abs(brute(Ck)).max())


def test_reference_phase_rounding_differs_from_fp64_only_slightly():
    """The float32 phase of fastFuncs.py:260 moves the map by far less than the 1e-4 gate at
    k r ~ 1e3 rad, and by more than float64 noise (so the two modes are really distinct)."""
    mpos, gpos = _scene(32, 3)
    rng = np.random.default_rng(3)
    X = rng.standard_normal((32, 2)) + 1j * rng.standard_normal((32, 2))
    C = X @ X.conj().T
    r0 = B.steer_r0(gpos, np.zeros(3))
    rm = B.dist_mat(gpos, mpos) * 1500  # k r ~ 1e3..1e4 rad
    a = B.freq_beamformer_full_csm(C, r0, rm, 1.0, 1.0, False, 'classic', 'reference')[0]
    b = B.freq_beamformer_full_csm(C, r0, rm, 1.0, 1.0, False, 'classic', 'fp64')[0]
    err = np.abs(a - b).max() / np.abs(b).max()
    assert 1e-12 < err < 1e-4


def test_c_restatement_matches_numpy_oracle():
    """oracle/cpu_ref.c (the timed CPU baseline) agrees with the numpy oracle."""
    from oracle import cref

    try:
        cref.lib()
    except Exception as e:  # pragma: no cover - gcc missing
        pytest.skip(f'C oracle not buildable: {e}')
    mpos, gpos = _scene(20, 6)
    rng = np.random.default_rng(6)
    X = rng.standard_normal((20, 3)) + 1j * rng.standard_normal((20, 3))
    C = X @ X.conj().T
    r0 = B.steer_r0(gpos, np.zeros(3))
    rm = B.dist_mat(gpos, mpos)
    for st in (1, 2, 3, 4):
        for rd in (False, True):
            a = B.freq_beamformer_full_csm(C, r0, rm, 60.0, 1.2, rd, st)
            b = cref.beamformer_full_csm(C, r0, rm, 60.0, 1.2, rd, st)
            assert np.abs(a[0] - b[0]).max() / np.abs(a[0]).max() < 1e-6
            np.testing.assert_allclose(a[1], b[1], rtol=1e-13)
    spec = rng.standard_normal((4, 20)) + 1j * rng.standard_normal((4, 20))
    c1 = np.zeros((4, 20, 20), complex)
    c2 = c1.copy()
    S.calc_csm_update(c1, spec)
    cref.calc_csm(c2, spec)
    np.testing.assert_allclose(c1, c2, atol=1e-13)
    np.testing.assert_allclose(cref.dist_mat(gpos, mpos), rm, rtol=1e-15)


def test_l_p():
    x = np.array([4e-10, 4e-8, 0.0, -1.0])
    lp = B.l_p(x)
    np.testing.assert_allclose(lp[:2], [0.0, 20.0])
    assert lp[2] == lp[3] == 10 * np.log10(np.finfo(float).eps)
