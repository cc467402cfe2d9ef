"""Committed golden vectors (tests/golden/oracle_small.npz, written by tests/golden/make_golden.py
from the pinned oracle): the oracle must keep reproducing them (CPU), and the GPU path must match
them within the map / CSM / transfer tolerances (GPU), without recomputing the oracle."""

import os

import numpy as np
import pytest

import acoular_amd as aa
from acoular_amd import _lib
from acoular_amd.fbeamform import beamform_device
from tests.conftest import peak_rel_err

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'oracle_small.npz')
STEERS = ['classic', 'inverse', 'true level', 'true location']


@pytest.fixture(scope='module')
def golden():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_oracle_reproduces_golden_vectors(golden):
    from tests.golden import make_golden

    fresh = make_golden.build()
    assert set(fresh) == set(golden)
    for key, want in golden.items():
        np.testing.assert_allclose(fresh[key], want, rtol=1e-12, atol=1e-14 * max(1.0, np.abs(want).max()),
                                   err_msg=key)


def test_golden_known_answer_values(golden):
    # PowerSpectra.synthetic_ev doctest (spectra.py:570-578), 8 printed digits
    assert golden['kat_ev5'].shape == (4,) and golden['kat_ev0'][0] == pytest.approx(0.00022468)


@pytest.mark.gpu
@pytest.mark.parametrize('form', ['plain_f32', 'mfma_f32', 'mfma_bf16x3', 'mfma_f16x2'])
def test_gpu_maps_match_golden(gpu, golden, form):
    mics = aa.MicGeom(pos_total=golden['bf_mpos'])
    grid = aa.ImportGrid(pos=golden['bf_gpos'])
    freqs = golden['bf_freqs']
    dcsm = _lib.to_device(golden['bf_csm'], _lib.torch().complex64)
    nm = mics.num_mics
    for i, st in enumerate(STEERS):
        sv = aa.SteeringVector(grid=grid, mics=mics, steer_type=st)
        for rd in (0, 1):
            norm = nm / (nm - 1.0) if rd else 1.0
            out = beamform_device(dcsm, freqs, np.arange(3), sv, st, bool(rd), norm, form, clip=bool(rd))
            err = peak_rel_err(out.cpu().numpy(), golden['bf_maps'][i, rd])
            assert err <= (5e-4 if form == 'mfma_f16x2' else 1e-4), (st, rd, err)
        _, onorm = beamform_device(dcsm, freqs, np.arange(3), sv, st, False, 1.0, form, out_norm=True)
        np.testing.assert_allclose(onorm.cpu().numpy(), golden['bf_norms'][i], rtol=1e-5)


@pytest.mark.gpu
def test_gpu_welch_csm_matches_golden(gpu, golden):
    ps = aa.PowerSpectra(source=aa.TimeSamples(data=golden['welch_time'], sample_freq=51200.0), block_size=256,
                         window='Hanning', overlap='50%', precision='complex64')
    want = golden['welch_csm']
    assert np.abs(ps.csm - want).max() <= 1e-5 * np.abs(want).max()


@pytest.mark.gpu
def test_gpu_transfer_matches_golden(gpu, golden):
    from oracle import beamform as B

    r0 = B.steer_r0(golden['tr_gpos'], np.zeros(3))
    rm = B.dist_mat(golden['tr_gpos'], golden['tr_mpos'])
    got = aa.calcTransfer(r0, rm, float(golden['tr_k'][0]))
    want = golden['tr_transfer']
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-6 * np.abs(want).max())
