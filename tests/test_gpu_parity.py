"""GPU parity: the HIP kernels of libabf.so against the CPU oracle on the same seeded inputs.

Tolerance: the maps agree with the oracle (phase_mode 'reference', float64) to a peak-normalised
error max_f max_g |B - B_ref| / max_g |B_ref| <= 1e-4 for the plain, fp32 and split-bf16 forms;
CSMs to 1e-5 of their largest element; transfer functions to 1e-6 absolute (float32 cos/sin).
The opt-in f16x2 form rounds the CSM once to fp16 (unit roundoff 2^-11): its map error is the map
of that rounding matrix, up to about 2e-4 of the peak where the grid misses the sources' beams
(measured 1.8e-4 on the large-k r scene below), so its gate is 5e-4.
"""

import ctypes

import numpy as np
import pytest

import acoular_amd as aa
from acoular_amd import _lib, synthetic
from acoular_amd.fbeamform import beamform_device, make_job
from oracle import beamform as B
from oracle import spectra as S
from tests.conftest import peak_rel_err

pytestmark = pytest.mark.gpu

STEERS = ['classic', 'inverse', 'true level', 'true location']
FORMS = ['plain_f32', 'mfma_f32', 'mfma_bf16x3', 'mfma_f16x2']
TOL = 1e-4
TOL_F16X2 = 5e-4


def _tol(form):
    return TOL_F16X2 if form == 'mfma_f16x2' else TOL


def _scene(nm, grid, freqs, nsrc=3, seed=1, z=0.5, diameter=1.0):
    mpos = synthetic.disk_array(nm, diameter, seed)
    spos = synthetic.source_positions(nsrc, [-0.2, -0.2, z], [0.2, 0.2, z], seed + 1)
    p = synthetic.level_powers(nsrc, 10.0, seed + 2)
    csm = synthetic.analytic_csm(mpos, spos, p, freqs)
    mics = aa.MicGeom(pos_total=mpos)
    return mics, grid, csm


def _gpu_maps(csm, freqs, mics, grid, steer_type, r_diag, form, phase_mode='reference', fast_trig=None, ref=None):
    sv = aa.SteeringVector(grid=grid, mics=mics, steer_type=steer_type)
    if ref is not None:
        sv.ref = ref
    nm = mics.num_mics
    norm = B.sig_loss_norm(r_diag, 0.0, nm)
    dcsm = _lib.to_device(csm, _lib.torch().complex64)
    out = beamform_device(dcsm, freqs, np.arange(len(freqs)), sv, steer_type, r_diag, norm, form, phase_mode,
                          clip=r_diag, fast_trig=fast_trig)
    return out.cpu().numpy()


def _oracle_maps(csm, freqs, mics, grid, steer_type, r_diag, phase_mode='reference', ref=(0.0, 0.0, 0.0)):
    return B.beamformer_map(csm.astype(np.complex128), freqs, grid.pos, mics.pos, steer_type, r_diag,
                            ind=np.arange(len(freqs)), phase_mode=phase_mode, ref=ref)


@pytest.mark.parametrize('form', FORMS)
@pytest.mark.parametrize('steer', STEERS)
@pytest.mark.parametrize('r_diag', [False, True])
def test_grid_beamformer_matches_oracle(gpu, form, steer, r_diag):
    freqs = np.array([800.0, 4000.0, 11000.0, 19000.0])
    grid = aa.RectGrid(x_min=-0.2, x_max=0.2, y_min=-0.2, y_max=0.2, z=0.3, increment=0.01)
    mics, grid, csm = _scene(64, grid, freqs, z=0.3)
    got = _gpu_maps(csm, freqs, mics, grid, steer, r_diag, form)
    ref = _oracle_maps(csm, freqs, mics, grid, steer, r_diag)
    err = peak_rel_err(got, ref)
    print(f'form={form} steer={steer} r_diag={r_diag} max rel map err={err:.2e}')
    assert err <= _tol(form)


@pytest.mark.parametrize('form', FORMS)
@pytest.mark.parametrize('fast_trig', [False, True])
@pytest.mark.parametrize('phase_mode', ['reference', 'fp64'])
def test_phase_modes_and_trig_paths(gpu, form, fast_trig, phase_mode):
    """Large k r (about 1.3e3 rad at 25.6 kHz and r ~ 2.7 m): float64 range reduction keeps the
    map within the gate in both phase modes, with either sin/cos implementation."""
    freqs = np.array([25600.0, 17000.0])
    grid = aa.RectGrid(x_min=-1.0, x_max=1.0, y_min=-1.0, y_max=1.0, z=2.5, increment=0.1)
    mics, grid, csm = _scene(48, grid, freqs, z=2.5, diameter=2.0)
    got = _gpu_maps(csm, freqs, mics, grid, 'true level', False, form, phase_mode, fast_trig)
    ref = _oracle_maps(csm, freqs, mics, grid, 'true level', False, phase_mode)
    err = peak_rel_err(got, ref)
    print(f'form={form} fast_trig={fast_trig} phase={phase_mode} err={err:.2e}')
    assert err <= _tol(form)


@pytest.mark.parametrize('form', FORMS)
@pytest.mark.parametrize('nm,nx,nf', [(5, 3, 1), (37, 7, 3), (100, 13, 2), (300, 9, 2), (33, 1, 1)])
def test_ragged_shapes(gpu, form, nm, nx, nf):
    """Microphone counts off the 16/32 tile multiples (and across the 256-mic LDS chunk of the
    plain kernel), grids off the tile width, a single grid point, a single frequency."""
    freqs = np.linspace(1500.0, 9000.0, nf)
    grid = aa.RectGrid(x_min=-0.3, x_max=0.3, y_min=-0.2, y_max=0.2, z=0.6, increment=0.6 / max(nx - 1, 1))
    if nx == 1:
        grid = aa.ImportGrid(pos=np.array([[0.05], [-0.02], [0.6]]))
    mics, grid, csm = _scene(nm, grid, freqs, z=0.6)
    for steer, r_diag in (('true location', True), ('classic', False)):
        got = _gpu_maps(csm, freqs, mics, grid, steer, r_diag, form)
        ref = _oracle_maps(csm, freqs, mics, grid, steer, r_diag)
        err = peak_rel_err(got, ref)
        print(f'form={form} M={nm} G={grid.size} F={nf} steer={steer} err={err:.2e}')
        assert err <= _tol(form)


@pytest.mark.parametrize('form', ['mfma_bf16x3', 'mfma_f16x2', 'mfma_f32', 'plain_f32'])
def test_many_mics(gpu, form):
    """1024 microphones (the largest array of the workloads), 3D grid slice, true level."""
    freqs = np.array([3000.0, 9000.0])
    grid = aa.RectGrid3D(x_min=-0.5, x_max=0.5, y_min=-0.5, y_max=0.5, z_min=0.8, z_max=1.2, increment=0.1)
    mpos = synthetic.spiral_array(1024, 3.0, 5)
    spos = synthetic.source_positions(6, [-0.5, -0.5, 0.8], [0.5, 0.5, 1.2], 6)
    csm = synthetic.analytic_csm(mpos, spos, synthetic.level_powers(6, 20.0, 7), freqs)
    mics = aa.MicGeom(pos_total=mpos)
    got = _gpu_maps(csm, freqs, mics, grid, 'true level', True, form)
    ref = _oracle_maps(csm, freqs, mics, grid, 'true level', True)
    err = peak_rel_err(got, ref)
    print(f'form={form} M=1024 G={grid.size} err={err:.2e}')
    assert err <= _tol(form)


@pytest.mark.parametrize('env', [{'ABF_PLAIN_JC': '128'}, {'ABF_PLAIN_JC': '256'}, {'ABF_PLAIN_LDSU': '1'}])
def test_plain_kernel_variants(gpu, monkeypatch, env):
    """The plain kernel's other microphone chunk sizes (partial sums split differently: within
    1e-6 of the peak of the default kernel's maps) and its LDS-staged U form (the same accumulation
    order: bit for bit) against the default kernel, and all within the gate: a ragged microphone
    count (more than one chunk, a partial last chunk), both r_diag settings."""
    freqs = np.array([1500.0, 8000.0, 15000.0])
    grid = aa.RectGrid(x_min=-0.3, x_max=0.3, y_min=-0.3, y_max=0.3, z=0.5, increment=0.02)
    mics, grid, csm = _scene(300, grid, freqs, z=0.5, seed=8)
    for steer, r_diag in (('inverse', True), ('true level', False)):
        want = _gpu_maps(csm, freqs, mics, grid, steer, r_diag, 'plain_f32')
        with monkeypatch.context() as m:
            for k, v in env.items():
                m.setenv(k, v)
            got = _gpu_maps(csm, freqs, mics, grid, steer, r_diag, 'plain_f32')
        if 'ABF_PLAIN_LDSU' in env:
            np.testing.assert_array_equal(got, want)
        else:
            assert peak_rel_err(got, want) <= 1e-6
        assert peak_rel_err(got, _oracle_maps(csm, freqs, mics, grid, steer, r_diag)) <= TOL


def test_reference_point_and_distance(gpu):
    freqs = np.array([3000.0])
    grid = aa.RectGrid(x_min=-0.2, x_max=0.2, y_min=-0.2, y_max=0.2, z=0.4, increment=0.05)
    mics, grid, csm = _scene(24, grid, freqs, z=0.4)
    for ref in ([0.1, -0.2, 0.05], 1.7):
        for st in ('inverse', 'true level'):
            got = _gpu_maps(csm, freqs, mics, grid, st, False, 'mfma_bf16x3', ref=ref)
            want = _oracle_maps(csm, freqs, mics, grid, st, False, ref=np.asarray(ref) if not np.isscalar(ref) else ref)
            assert peak_rel_err(got, want) <= TOL


# ---------------------------------------------------------------------------------------------
# fastFuncs drop-ins
# ---------------------------------------------------------------------------------------------
def _tables(nm=20, seed=3):
    mics, grid, _ = _scene(nm, aa.RectGrid(x_min=-0.3, x_max=0.3, y_min=-0.3, y_max=0.3, z=0.5, increment=0.05),
                           [1000.0])
    r0 = B.steer_r0(grid.pos, np.zeros(3))
    rm = B.dist_mat(grid.pos, mics.pos)
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((nm, 4)) + 1j * rng.standard_normal((nm, 4))
    return r0, rm, X @ X.conj().T


@pytest.mark.parametrize('steer', STEERS)
@pytest.mark.parametrize('r_diag', [False, True])
def test_beamformerfreq_dropin(gpu, steer, r_diag):
    r0, rm, C = _tables()
    k = 2 * np.pi * 6000 / 343.0
    got = aa.beamformerFreq(steer, r_diag, 1.3, (r0, rm, k), C)
    ref = B.freq_beamformer_full_csm(C, r0, rm, k, 1.3, r_diag, steer)
    assert got[0].dtype == np.float64 and got[0].shape == ref[0].shape
    assert peak_rel_err(got[0], ref[0]) <= TOL
    np.testing.assert_allclose(got[1], ref[1], rtol=1e-6)
    # eigen-decomposition form with all eigenpairs
    lam, V = np.linalg.eigh(C)
    got_e = aa.beamformerFreq(steer, r_diag, 1.3, (r0, rm, k), (lam, V))
    ref_e = B.freq_beamformer_eig(lam, V, r0, rm, k, 1.3, r_diag, steer)
    assert peak_rel_err(got_e[0], ref_e[0]) <= TOL


@pytest.mark.parametrize('r_diag', [False, True])
def test_beamformerfreq_custom_dropin(gpu, r_diag):
    r0, rm, C = _tables(16, 4)
    k = 2 * np.pi * 3000 / 343.0
    h = B.steer_vector(B.calc_transfer(r0, rm, k), 'true location')
    got = aa.beamformerFreq('custom', r_diag, 1.0, h, C)
    ref = B.freq_beamformer_custom(C, h, 1.0, r_diag)
    assert peak_rel_err(got[0], ref[0]) <= TOL
    np.testing.assert_allclose(got[1], ref[1], rtol=1e-5)
    lam, V = np.linalg.eigh(C)
    got_e = aa.beamformerFreq('custom', r_diag, 1.0, h, (lam[-2:], V[:, -2:]))
    ref_e = B.freq_beamformer_eig_custom(lam[-2:], V[:, -2:], h, 1.0, r_diag)
    assert peak_rel_err(got_e[0], ref_e[0]) <= TOL


def test_calc_transfer_and_steer_vector(gpu):
    r0, rm, _ = _tables(12)
    k = 2 * np.pi * 12000 / 343.0
    got = aa.calcTransfer(r0, rm, k)
    ref = B.calc_transfer(r0, rm, k)
    assert got.dtype == np.complex128
    np.testing.assert_allclose(got, ref, rtol=0, atol=2e-6 * np.abs(ref).max())
    mics, grid, _ = _scene(12, aa.RectGrid(x_min=-0.3, x_max=0.3, y_min=-0.3, y_max=0.3, z=0.5, increment=0.05),
                           [1.0])
    for st in STEERS:
        sv = aa.SteeringVector(grid=grid, mics=mics, steer_type=st)
        h = sv.steer_vector(12000.0)
        want = B.steer_vector(B.calc_transfer(sv.r0, sv.rm, 2 * np.pi * 12000.0 / 343.0), st)
        np.testing.assert_allclose(h, want, rtol=0, atol=2e-6 * np.abs(want).max())
        # a subset of grid points: every row is normalised on its own (fbeamform.py:145-153), so
        # the rows of the subset equal those of the full matrix for all four steer types
        h1 = sv.steer_vector(12000.0, ind=np.array([3, 5]))
        assert h1.shape == (2, 12)
        np.testing.assert_allclose(h1, want[[3, 5]], rtol=0, atol=2e-6 * np.abs(want).max())
        # a scalar index takes the rm[ind, :][np.newaxis] branch (fbeamform.py:229-230)
        h2 = sv.steer_vector(12000.0, ind=4)
        assert h2.shape == (1, 12)
        np.testing.assert_allclose(h2, want[[4]], rtol=0, atol=2e-6 * np.abs(want).max())
        t2 = sv.transfer(12000.0, ind=4)
        np.testing.assert_allclose(t2, B.calc_transfer(sv.r0, sv.rm, 2 * np.pi * 12000.0 / 343.0)[[4]], rtol=0,
                                   atol=2e-6 * np.abs(t2).max())


@pytest.mark.parametrize('dtype', [np.complex128, np.complex64])
@pytest.mark.parametrize('nm', [13, 16, 70, 530])
def test_calc_csm_dropin(gpu, dtype, nm):
    """Odd and even mic counts (two columns per 16-byte access for an even count), a last column
    block that is partly empty; the lower triangle stays as it was."""
    rng = np.random.default_rng(2)
    spec = (rng.standard_normal((9, nm)) + 1j * rng.standard_normal((9, nm))).astype(dtype)
    csm = (rng.standard_normal((9, nm, nm)) + 1j * rng.standard_normal((9, nm, nm))).astype(dtype)
    want = S.calc_csm_update(csm.astype(np.complex128).copy(), spec.astype(np.complex128))
    got = aa.calcCSM(csm, spec)
    assert got is csm
    tol = 1e-12 if dtype == np.complex128 else 2e-6
    np.testing.assert_allclose(got, want, rtol=0, atol=tol * np.abs(want).max())


def test_even_counts_on_views_that_start_off_16_bytes(gpu):
    """calcCSM (complex64) and dist_mat with an even count on caller views that start 8 bytes
    past a 16-byte boundary: the library takes the one-column path and the results stay exact."""
    torch = _lib.torch()
    lib = _lib.lib()
    rng = np.random.default_rng(12)
    nf, nm = 5, 34
    spec = (rng.standard_normal((nf, nm)) + 1j * rng.standard_normal((nf, nm))).astype(np.complex64)
    csm = (rng.standard_normal((nf, nm, nm)) + 1j * rng.standard_normal((nf, nm, nm))).astype(np.complex64)
    want = S.calc_csm_update(csm.astype(np.complex128).copy(), spec.astype(np.complex128))
    buf = torch.zeros(nf * nm * nm + 1, dtype=torch.complex64, device=gpu)
    sbuf = torch.zeros(nf * nm + 1, dtype=torch.complex64, device=gpu)
    dcsm, dspec = buf[1:], sbuf[1:]
    assert dcsm.data_ptr() % 16 == 8 and dspec.data_ptr() % 16 == 8
    dcsm.copy_(torch.from_numpy(csm.ravel()))
    dspec.copy_(torch.from_numpy(spec.ravel()))
    _lib.check(lib.abf_calc_csm(_lib.ptr(dcsm), _lib.ptr(dspec), nf, nm, 0, _lib.stream_handle()))
    got = dcsm.cpu().numpy().reshape(nf, nm, nm)
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-6 * np.abs(want).max())
    g, m = rng.standard_normal((3, 77)), rng.standard_normal((3, 34))
    rbuf = torch.zeros(77 * 34 + 1, dtype=torch.float64, device=gpu)
    drm = rbuf[1:]
    assert drm.data_ptr() % 16 == 8
    dg, dm = _lib.to_device(g), _lib.to_device(m)
    _lib.check(lib.abf_dist_mat(_lib.ptr(dg), _lib.ptr(dm), 77, 34, _lib.ptr(drm), _lib.stream_handle()))
    np.testing.assert_allclose(drm.cpu().numpy().reshape(77, 34), B.dist_mat(g, m), rtol=1e-15)


@pytest.mark.parametrize('nm', [17, 16, 530, 1024])
def test_dist_mat(gpu, nm):
    """Odd and even microphone counts (an even count takes the two-distances-per-store path),
    and counts above 512 (more than one microphone block per grid point)."""
    rng = np.random.default_rng(1)
    g, m = rng.standard_normal((3, 301)), rng.standard_normal((3, nm))
    np.testing.assert_allclose(aa.dist_mat(g, m), B.dist_mat(g, m), rtol=1e-15)
    np.testing.assert_allclose(aa.Environment().apparent_r(g), B.apparent_r(g), rtol=1e-15)


# ---------------------------------------------------------------------------------------------
# Welch CSM
# ---------------------------------------------------------------------------------------------
def test_power_spectra_known_answer(gpu):
    """spectra.py:570-578 doctest through the GPU Welch path (+ eigh of the resident CSM)."""
    np.random.seed(0)
    data = np.random.rand(1000, 4)
    ps = aa.PowerSpectra(source=aa.TimeSamples(data=data, sample_freq=51200), block_size=128, window='Hamming')
    np.testing.assert_allclose(ps.synthetic_ev(5000, 5), [0.00048803, 0.0010141, 0.00234248, 0.00457097], rtol=0,
                               atol=5e-9)
    np.testing.assert_allclose(ps.synthetic_ev(5000), [0.00022468, 0.0004589, 0.00088059, 0.00245989], rtol=0,
                               atol=5e-9)
    assert ps.csm.shape == (65, 4, 4) and ps.csm.dtype == np.complex128


@pytest.mark.parametrize('bs,window,overlap,nm,n', [
    (128, 'Hanning', '50%', 64, 51200),
    (256, 'Rectangular', 'None', 7, 3000),
    (512, 'Blackman', '75%', 33, 9000),
    (1024, 'Bartlett', '87.5%', 16, 20000),
    (2048, 'Hamming', '50%', 40, 30000),
    (128, 'Hanning', '50%', 3, 100),  # fewer samples than one block: zero blocks
])
def test_welch_csm_matches_oracle(gpu, bs, window, overlap, nm, n):
    rng = np.random.default_rng(bs + nm)
    data = rng.standard_normal((n, nm)) * rng.uniform(0.5, 2.0, nm)
    data[:, 1:] += 0.5 * data[:, :1]  # correlated channels
    ps = aa.PowerSpectra(source=aa.TimeSamples(data=data, sample_freq=51200.0), block_size=bs, window=window,
                         overlap=overlap, precision='complex64')
    got = ps.csm
    want = S.welch_csm(data, bs, window, overlap, 'complex128')
    assert got.dtype == np.complex64 and got.shape == want.shape
    if n < bs:
        assert np.all(got == 0)
        return
    err = np.abs(got - want).max() / np.abs(want).max()
    print(f'bs={bs} M={nm} blocks={len(S.block_starts(n, bs, overlap))} rel err={err:.2e}')
    assert err <= 1e-5
    np.testing.assert_array_equal(got, np.conj(got.transpose(0, 2, 1)))


def test_welch_csm_complex128_bound_on_the_64_mic_case(gpu):
    """precision='complex128' (the reference's default) on the 64-mic, 1 s, block size 128 case:
    the CSM is accumulated in complex64 on the GPU and returned as complex128.  Against the float64
    oracle it stays within 1e-5 of the largest element and within 1e-5 relative on every
    auto-spectrum (the diagonal), and it is exactly Hermitian."""
    mpos = synthetic.disk_array(64, 1.0, 1001)
    spos = synthetic.source_positions(3, [-0.2, -0.2, 0.3], [0.2, 0.2, 0.3], 1002)
    data = synthetic.point_source_signals(mpos, spos, 51200, 51200.0, seed=1003, noise_rel=0.1)
    ps = aa.PowerSpectra(source=aa.TimeSamples(data=data, sample_freq=51200.0), block_size=128, window='Hanning',
                         overlap='50%')
    assert ps.precision == 'complex128'
    got = ps.csm
    want = S.welch_csm(data, 128, 'Hanning', '50%', 'complex128')
    assert got.dtype == np.complex128 and got.shape == want.shape == (65, 64, 64)
    err = np.abs(got - want).max() / np.abs(want).max()
    d_got = np.einsum('fii->fi', got).real
    d_want = np.einsum('fii->fi', want).real
    drel = np.max(np.abs(d_got - d_want) / d_want)
    print(f'complex128 CSM: peak-normalised error {err:.2e}, auto-spectra relative error {drel:.2e}')
    assert err <= 1e-5
    assert drel <= 1e-5
    np.testing.assert_array_equal(got, np.conj(got.transpose(0, 2, 1)))


@pytest.mark.parametrize('tr', ['2', '4'])
@pytest.mark.parametrize('nm', [33, 70, 300])
def test_welch_csm_tile_forms(gpu, monkeypatch, tr, nm):
    """Both forms of the CSM update (32 x 64 tiles, and 64 x 64 tiles with the next blocks' loads
    in flight) on ragged mic counts and a pass count that is not a multiple of 4 blocks."""
    monkeypatch.setenv('ABF_HERK_TR', tr)
    rng = np.random.default_rng(nm)
    data = rng.standard_normal((4500, nm))
    data[:, 1:] += 0.3 * data[:, :1]
    ps = aa.PowerSpectra(source=aa.TimeSamples(data=data, sample_freq=51200.0), block_size=256, window='Hanning',
                         overlap='50%', precision='complex64')
    got = ps.csm
    want = S.welch_csm(data, 256, 'Hanning', '50%', 'complex128')
    assert np.abs(got - want).max() / np.abs(want).max() <= 1e-5
    np.testing.assert_array_equal(got, np.conj(got.transpose(0, 2, 1)))


def test_welch_multi_pass(gpu, monkeypatch):
    """The CSM accumulated over several passes of blocks equals one pass."""
    from acoular_amd import spectra

    rng = np.random.default_rng(4)
    data = rng.standard_normal((20000, 10))
    ps = aa.PowerSpectra(source=aa.TimeSamples(data=data, sample_freq=1.0), block_size=256, window='Hanning',
                         overlap='50%')
    one = ps.csm.copy()
    monkeypatch.setattr(spectra, '_PASS_BYTES', 129 * 10 * 8 * 7)  # 7 blocks per pass
    ps2 = aa.PowerSpectra(source=aa.TimeSamples(data=data.copy(), sample_freq=1.0), block_size=256,
                          window='Hanning', overlap='50%')
    np.testing.assert_allclose(ps2.csm, one, rtol=0, atol=1e-6 * np.abs(one).max())


# ---------------------------------------------------------------------------------------------
# BeamformerBase API
# ---------------------------------------------------------------------------------------------
def test_beamformer_base_api(gpu):
    """The reference's entry sequence: PowerSpectra -> SteeringVector -> BeamformerBase.synthetic
    (the reference's README example), lazily per frequency, checked against the oracle chain."""
    mpos = synthetic.disk_array(32, 1.0, 3)
    spos = synthetic.source_positions(2, [-0.1, -0.1, 0.3], [0.1, 0.1, 0.3], 4)
    data = synthetic.point_source_signals(mpos, spos, 25600, 51200.0, seed=5, noise_rel=0.1)
    ts = aa.TimeSamples(data=data, sample_freq=51200.0)
    ps = aa.PowerSpectra(source=ts, block_size=128, window='Hanning', overlap='50%')
    grid = aa.RectGrid(x_min=-0.2, x_max=0.2, y_min=-0.2, y_max=0.2, z=0.3, increment=0.02)
    mics = aa.MicGeom(pos_total=mpos)
    bf = aa.BeamformerBase(freq_data=ps, steer=aa.SteeringVector(grid=grid, mics=mics))
    m1 = bf.synthetic(8000, 1)
    assert m1.shape == grid.shape
    f = ps.fftfreq()
    i1, i2 = np.searchsorted(f, 8000 * 2**-0.5), np.searchsorted(f, 8000 * 2**0.5)
    assert bf._fr.sum() == i2 - i1  # only the band was computed
    csm_ref = S.welch_csm(data, 128, 'Hanning', '50%')
    ref = B.beamformer_map(csm_ref, f, grid.pos, mpos, 'true level', True, ind=np.arange(i1, i2)).sum(0)
    assert np.abs(m1.ravel() - ref).max() / np.abs(ref).max() <= TOL
    with pytest.warns(Warning):
        bf.synthetic(8123.0, 0)
    lp = aa.L_p(m1)
    assert np.isfinite(lp).all()


@pytest.mark.parametrize('steer', STEERS)
def test_integrate_one_point_csm(gpu, steer):
    """tests/unittests/test_integrate.py:22-38: a CSM built from the transfer function of the
    grid point (0, 0, 1); integrating a sector around it gives the map peak there."""
    rng1 = np.random.RandomState(1)
    mics = aa.MicGeom(pos_total=rng1.normal(size=3 * 5).reshape((3, 5)))
    sv1 = aa.SteeringVector(grid=aa.ImportGrid(pos=np.array([[0], [0], [1]])), mics=mics)
    transf = sv1.transfer(2000).T[None]
    csm = transf @ transf.swapaxes(2, 1).conjugate()
    ps = aa.PowerSpectraImport(csm=csm, frequencies=2000)
    grid = aa.RectGrid(x_min=-1, x_max=1, y_min=-1, y_max=1, z=1, increment=0.1)
    bf = aa.BeamformerBase(freq_data=ps, steer=aa.SteeringVector(grid=grid, mics=mics, steer_type=steer))
    res = bf.synthetic(2000)
    ref = _oracle_maps(csm, np.array([2000.0]), mics, grid, steer, True)[0].reshape(grid.shape)
    assert np.abs(res - ref).max() / np.abs(ref).max() <= TOL
    integ = bf.integrate((0.0, 0.0, 0.0001))  # circle around (0, 0): the single nearest point
    assert integ.shape == (1,)
    np.testing.assert_allclose(integ[0], res[10, 10], rtol=1e-12)


# ---------------------------------------------------------------------------------------------
# batching determinism
# ---------------------------------------------------------------------------------------------
def test_batch_equals_alone_bitwise(gpu):
    torch = _lib.torch()
    jobs_spec = [(32, 0.05, [1000.0, 2000.0, 3000.0]), (96, 0.04, [5000.0]), (64, 0.1, [700.0, 9000.0])]
    tensors, jobs, alone = [], [], []
    for i, (nm, inc, freqs) in enumerate(jobs_spec):
        grid = aa.RectGrid(x_min=-0.4, x_max=0.4, y_min=-0.3, y_max=0.3, z=0.5, increment=inc)
        mics, grid, csm = _scene(nm, grid, np.array(freqs), seed=10 + i)
        sv = aa.SteeringVector(grid=grid, mics=mics, steer_type='true location')
        dcsm = _lib.to_device(csm, torch.complex64)
        alone.append(beamform_device(dcsm, freqs, np.arange(len(freqs)), sv, 'true location', True,
                                     nm / (nm - 1.0), 'mfma_bf16x3', clip=True).cpu().numpy())
        gpos, mpos = sv.device_geometry()
        k = torch.from_numpy(2 * np.pi * np.array(freqs) / 343.0).to(gpu)
        out = torch.full((len(freqs), grid.size), -1.0, dtype=torch.float32, device=gpu)
        jobs.append(make_job(dcsm, k, gpos, mpos, sv.ref, 'true location', True, nm / (nm - 1.0), 'mfma_bf16x3',
                             'reference', out, None, True))
        tensors += [dcsm, k, gpos, mpos, out]
    arr = (_lib.AbfGridJob * len(jobs))(*jobs)
    lib = _lib.lib()
    nbytes = lib.abf_beamform_batch_workspace_size(arr, len(jobs), None)
    ws = _lib.workspace(nbytes)
    _lib.check(lib.abf_beamform_batch(arr, len(jobs), _lib.ptr(ws), ws.numel(), _lib.stream_handle()))
    for a, o in zip(alone, tensors[4::5]):
        np.testing.assert_array_equal(o.cpu().numpy(), a)


def test_persistent_streamed_kernel_equals_one_item_per_workgroup(gpu):
    """The persistent streamed launch (kernel 8: one workgroup per CU walking the (job, tile, band)
    items round robin) gives the maps of the one-item-per-workgroup launch (kernel 6) bit for bit:
    a batch of jobs in two launch groups (64 and 128 microphones, two jobs sharing a group), more
    items than CUs, run eagerly twice in a row and replayed from a captured graph with other GPU
    work between the replays."""
    torch = _lib.torch()
    lib = _lib.lib()
    specs = [(64, 0.01, [900.0, 2000.0, 3100.0, 5000.0, 7700.0]), (128, 0.02, [1500.0, 6000.0]),
             (64, 0.03, [12000.0, 400.0, 8000.0])]
    tensors, outs = [], {6: [], 8: []}
    jobs = {6: [], 8: []}
    for i, (nm, inc, freqs) in enumerate(specs):
        grid = aa.RectGrid(x_min=-0.5, x_max=0.5, y_min=-0.4, y_max=0.4, z=0.5, increment=inc)
        mics, grid, csm = _scene(nm, grid, np.array(freqs), seed=30 + i)
        sv = aa.SteeringVector(grid=grid, mics=mics, steer_type='true level')
        gpos, mpos = sv.device_geometry()
        dcsm = _lib.to_device(csm, torch.complex64)
        k = torch.from_numpy(2 * np.pi * np.array(freqs) / 343.0).to(gpu)
        tensors += [dcsm, k, gpos, mpos]
        for kern in (6, 8):
            out = torch.full((len(freqs), grid.size), -3.0, dtype=torch.float32, device=gpu)
            job = make_job(dcsm, k, gpos, mpos, sv.ref, 'true level', True, nm / (nm - 1.0), 'mfma_bf16x3',
                           'reference', out, None, True, fast_trig=True)
            job.tile_kernel, job.tile_cols, job.tile_freqs = kern, 64, 2
            jobs[kern].append(job)
            outs[kern].append(out)
    plan = [ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()]
    _lib.check(lib.abf_beamform_plan(jobs[8][0], *[ctypes.byref(x) for x in plan]))
    assert plan[0].value == 8 and plan[3].value > 256  # more items than the persistent grid
    arrs = {kern: (_lib.AbfGridJob * len(js))(*js) for kern, js in jobs.items()}
    ws = {kern: _lib.workspace(lib.abf_beamform_batch_workspace_size(a, len(jobs[kern]), None))
          for kern, a in arrs.items()}
    stream = _lib.stream_handle()
    for kern in (6, 8):
        _lib.check(lib.abf_beamform_batch(arrs[kern], len(jobs[kern]), _lib.ptr(ws[kern]), ws[kern].numel(), stream))
    want = [o.cpu().numpy() for o in outs[6]]
    for o, w in zip(outs[8], want):
        np.testing.assert_array_equal(o.cpu().numpy(), w)
    # a second eager call
    for o in outs[8]:
        o.fill_(-5.0)
    _lib.check(lib.abf_beamform_batch(arrs[8], len(jobs[8]), _lib.ptr(ws[8]), ws[8].numel(), stream))
    for o, w in zip(outs[8], want):
        np.testing.assert_array_equal(o.cpu().numpy(), w)
    # the first job's run (descriptors, persistent launch) captured and replayed
    job, out = jobs[8][0], outs[8][0]
    wsj = _lib.workspace(lib.abf_beamform_workspace_size(job))
    _lib.check(lib.abf_beamform_prepare(job, _lib.ptr(wsj), wsj.numel(), stream))
    torch.cuda.synchronize()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(side):
        with torch.cuda.graph(graph, stream=side):
            _lib.check(lib.abf_beamform_run(job, _lib.ptr(wsj), wsj.numel(), _lib.stream_handle()))
    torch.cuda.current_stream().wait_stream(side)
    for rep in range(3):
        out.fill_(float(-rep))
        junk = torch.randn(1 << 20, device=gpu).sum()  # other eager GPU work between replays
        graph.replay()
        torch.cuda.synchronize()
        assert torch.isfinite(junk)
        np.testing.assert_array_equal(out.cpu().numpy(), want[0])


@pytest.mark.parametrize('form', ['mfma_bf16x3', 'mfma_f16x2', 'mfma_f32'])
@pytest.mark.parametrize('kernel,tile_cols,tile_freqs', [(1, 0, 0), (1, 16, 1), (1, 64, 3), (2, 0, 0), (2, 64, 2),
                                                         (2, 32, 5), (3, 0, 0), (3, 16, 2), (3, 64, 4), (4, 0, 0),
                                                         (4, 64, 3), (5, 0, 0), (5, 32, 2), (6, 0, 0), (6, 64, 3),
                                                         (6, 16, 1), (7, 0, 0), (7, 64, 3), (8, 0, 0), (8, 64, 3)])
@pytest.mark.parametrize('nm,steer', [(64, 'true level'), (100, 'classic'), (256, 'inverse'), (33, 'true location')])
def test_mfma_kernel_variants_and_tiles(gpu, form, kernel, tile_cols, tile_freqs, nm, steer):
    """Every MFMA kernel variant (alternating, warp-specialised, row pairs, pipelined) and explicit tile shapes, including
    ragged microphone counts and bands that do not divide into the frequency tile."""
    torch = _lib.torch()
    freqs = np.array([600.0, 2500.0, 7000.0, 12000.0, 15000.0, 21000.0, 3100.0])
    grid = aa.RectGrid(x_min=-0.3, x_max=0.3, y_min=-0.25, y_max=0.25, z=0.4, increment=0.02)
    mics, grid, csm = _scene(nm, grid, freqs, z=0.4, seed=nm)
    sv = aa.SteeringVector(grid=grid, mics=mics, steer_type=steer)
    gpos, mpos = sv.device_geometry()
    dcsm = _lib.to_device(csm, torch.complex64)
    k = torch.from_numpy(2 * np.pi * freqs / 343.0).to(gpu)
    out = torch.full((len(freqs), grid.size), -7.0, dtype=torch.float32, device=gpu)
    norm = nm / (nm - 1.0)
    job = make_job(dcsm, k, gpos, mpos, sv.ref, steer, True, norm, form, 'reference', out, None, True)
    job.tile_kernel, job.tile_cols, job.tile_freqs = kernel, tile_cols, tile_freqs
    lib = _lib.lib()
    ws = _lib.workspace(lib.abf_beamform_workspace_size(job))
    _lib.check(lib.abf_beamform(job, _lib.ptr(ws), ws.numel(), _lib.stream_handle()))
    ref = _oracle_maps(csm, freqs, mics, grid, steer, True)
    err = peak_rel_err(out.cpu().numpy(), ref)
    print(f'form={form} kernel={kernel} tile=({tile_cols},{tile_freqs}) nm={nm} steer={steer} err={err:.2e}')
    assert err <= _tol(form)


@pytest.mark.parametrize('kernel, nm', [(kn, nm) for kn in (6, 7, 8) for nm in (33, 64, 65, 100, 128, 160, 192, 224, 256)]
                         + [(7, nm) for nm in (257, 300, 352, 384, 416, 449, 480, 512)])
@pytest.mark.parametrize('steer', STEERS)
def test_streamed_kernel_every_block_count(gpu, kernel, nm, steer):
    """The streamed kernels (one instantiation per 2..8 blocks of 32 microphones, 2..16 for the
    wide kernel): both r_diag settings, a grid whose last tile is partial, a band that does not
    divide into the frequency tile; the plan reports the kernel asked for."""
    torch = _lib.torch()
    freqs = np.array([700.0, 3300.0, 9000.0, 14000.0, 19500.0])
    grid = aa.RectGrid(x_min=-0.3, x_max=0.3, y_min=-0.2, y_max=0.2, z=0.5, increment=0.03)  # 21 x 14
    mics, grid, csm = _scene(nm, grid, freqs, z=0.5, seed=nm + 1)
    sv = aa.SteeringVector(grid=grid, mics=mics, steer_type=steer)
    gpos, mpos = sv.device_geometry()
    dcsm = _lib.to_device(csm, torch.complex64)
    k = torch.from_numpy(2 * np.pi * freqs / 343.0).to(gpu)
    lib = _lib.lib()
    for r_diag in (True, False):
        out = torch.full((len(freqs), grid.size), -7.0, dtype=torch.float32, device=gpu)
        onorm = torch.empty_like(out)
        norm = nm / (nm - 1.0) if r_diag else 1.0
        job = make_job(dcsm, k, gpos, mpos, sv.ref, steer, r_diag, norm, 'mfma_bf16x3', 'reference', out, onorm,
                       r_diag, fast_trig=True)
        job.tile_kernel, job.tile_cols, job.tile_freqs = kernel, 64, 2
        plan = [ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()]
        _lib.check(lib.abf_beamform_plan(job, *[ctypes.byref(x) for x in plan]))
        assert plan[0].value == kernel
        ws = _lib.workspace(lib.abf_beamform_workspace_size(job))
        _lib.check(lib.abf_beamform(job, _lib.ptr(ws), ws.numel(), _lib.stream_handle()))
        ref = _oracle_maps(csm, freqs, mics, grid, steer, r_diag)
        err = peak_rel_err(out.cpu().numpy(), ref)
        print(f'kernel={kernel} nm={nm} steer={steer} r_diag={r_diag} err={err:.2e}')
        assert err <= TOL
        want_norm = np.stack([B.freq_beamformer_full_csm(csm[0].astype(np.complex128), B.steer_r0(grid.pos, sv.ref),
                                                         B.dist_mat(grid.pos, mics.pos), 1.0, 1.0, r_diag, steer)[1]]
                             * len(freqs))
        np.testing.assert_allclose(onorm.cpu().numpy(), want_norm, rtol=1e-5)


@pytest.mark.parametrize('kernel', [6, 7, 8])
def test_streamed_kernel_falls_back_outside_its_cases(gpu, kernel):
    """The streamed kernels asked for the float64 phase, the library sin/cos or more than 256
    microphones (512 for the wide kernel) run the row-pair kernel (the plan says so) and stay
    within the gate."""
    torch = _lib.torch()
    lib = _lib.lib()
    freqs = np.array([2000.0, 17000.0])
    grid = aa.RectGrid(x_min=-0.2, x_max=0.2, y_min=-0.2, y_max=0.2, z=0.5, increment=0.04)
    big = 520 if kernel == 7 else 300
    for nm, phase, fast in ((64, 'fp64', True), (64, 'reference', False), (big, 'reference', True)):
        mics, grid, csm = _scene(nm, grid, freqs, z=0.5, seed=3)
        sv = aa.SteeringVector(grid=grid, mics=mics, steer_type='true level')
        gpos, mpos = sv.device_geometry()
        out = torch.empty((len(freqs), grid.size), dtype=torch.float32, device=gpu)
        # the job holds raw pointers: the device CSM and wavenumbers must stay referenced until the
        # launch is done, or the workspace allocated below may reuse their memory
        dcsm = _lib.to_device(csm, torch.complex64)
        k = torch.from_numpy(2 * np.pi * freqs / 343.0).to(gpu)
        job = make_job(dcsm, k, gpos, mpos, sv.ref, 'true level', True, nm / (nm - 1.0), 'mfma_bf16x3', phase, out,
                       None, True, fast_trig=fast)
        job.tile_kernel, job.tile_cols, job.tile_freqs = kernel, 64, 2
        plan = [ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()]
        _lib.check(lib.abf_beamform_plan(job, *[ctypes.byref(x) for x in plan]))
        assert plan[0].value in (1, 3)
        ws = _lib.workspace(lib.abf_beamform_workspace_size(job))
        _lib.check(lib.abf_beamform(job, _lib.ptr(ws), ws.numel(), _lib.stream_handle()))
        ref = _oracle_maps(csm, freqs, mics, grid, 'true level', True, phase)
        got = out.cpu().numpy()
        err = peak_rel_err(got, ref)
        bad = np.abs(got - ref) > TOL * np.abs(ref).max(axis=1, keepdims=True)
        print(f'kernel={kernel} nm={nm} phase={phase} fast={fast} plan={[x.value for x in plan]} err={err:.2e} '
              f'bad points per bin {bad.sum(axis=1).tolist()} first {np.argwhere(bad)[:4].tolist()}')
        assert err <= TOL


def test_mfma_kernel_variant_rejected_when_unknown(gpu):
    torch = _lib.torch()
    freqs = np.array([1000.0])
    grid = aa.RectGrid(x_min=-0.1, x_max=0.1, y_min=-0.1, y_max=0.1, z=0.4, increment=0.05)
    mics, grid, csm = _scene(16, grid, freqs)
    sv = aa.SteeringVector(grid=grid, mics=mics, steer_type='classic')
    gpos, mpos = sv.device_geometry()
    dcsm = _lib.to_device(csm, torch.complex64)
    k = torch.from_numpy(2 * np.pi * freqs / 343.0).to(gpu)
    out = torch.empty((1, grid.size), dtype=torch.float32, device=gpu)
    job = make_job(dcsm, k, gpos, mpos, sv.ref, 'classic', False, 1.0, 'mfma_bf16x3', 'reference', out)
    job.tile_kernel = 9
    with pytest.raises(ValueError):
        _lib.check(_lib.lib().abf_beamform(job, None, 0, _lib.stream_handle()))


@pytest.mark.parametrize('scale', [1e-20, 1.0, 1e20])
@pytest.mark.parametrize('steer', STEERS)
def test_f16x2_scaling(gpu, scale, steer):
    """The f16x2 form scales U by a power of two per frequency and the steering weights by one per
    grid point; maps of CSMs far outside fp16's range (x 1e-20, x 1e20) and of a grid 60-140 m
    away (weights r or 1/r far from 1) keep the oracle's accuracy."""
    freqs = np.array([900.0, 5000.0])
    grid = aa.RectGrid(x_min=-0.2, x_max=0.2, y_min=-0.2, y_max=0.2, z=0.5, increment=0.04)
    mics, grid, csm = _scene(48, grid, freqs, z=0.5)
    csm = csm * np.float32(scale)
    got = _gpu_maps(csm, freqs, mics, grid, steer, True, 'mfma_f16x2')
    want = _oracle_maps(csm, freqs, mics, grid, steer, True)
    assert peak_rel_err(got, want) <= TOL_F16X2
    far = aa.RectGrid(x_min=-40.0, x_max=40.0, y_min=-40.0, y_max=40.0, z=100.0, increment=8.0)
    mpos = synthetic.disk_array(48, 1.0, 1)
    spos = np.array([[8.0], [-16.0], [100.0]])
    fcsm = synthetic.analytic_csm(mpos, spos, np.array([1.0]), freqs) * np.float32(scale)
    fm = aa.MicGeom(pos_total=mpos)
    got = _gpu_maps(fcsm, freqs, fm, far, steer, False, 'mfma_f16x2')
    want = _oracle_maps(fcsm, freqs, fm, far, steer, False)
    assert peak_rel_err(got, want) <= TOL_F16X2
