"""Fused CSM + beamforming pipeline: parity with the oracle and hipGraph replays equal to eager
runs (bit for bit), with the program's own eager GPU work between the replays."""

import numpy as np
import pytest

from acoular_amd import _lib, synthetic
from acoular_amd.pipeline import FusedCsmBeamform
from oracle import beamform as B
from oracle import spectra as S
from tests.conftest import peak_rel_err


def test_rejects_host_time_data():
    torch = _lib.torch()
    with pytest.raises(TypeError):
        FusedCsmBeamform(torch.zeros((4096, 4), dtype=torch.float32), 51200.0, np.zeros((3, 4)), np.zeros((3, 4)),
                         block_size=256)


def _scene(nm=24, n=6000, fs=51200.0, seed=11):
    mpos = synthetic.disk_array(nm, 0.8, seed)
    spos = synthetic.source_positions(3, [-0.2, -0.2, 0.6], [0.2, 0.2, 0.6], seed + 1)
    data = synthetic.point_source_signals(mpos, spos, n, fs, seed=seed + 2, noise_rel=0.05)
    x = np.linspace(-0.3, 0.3, 9)
    gpos = np.stack([np.repeat(x, 9), np.tile(x, 9), np.full(81, 0.6)])
    return mpos, gpos, data


@pytest.mark.gpu
@pytest.mark.parametrize('steer', ['true location', 'classic'])
def test_fused_pipeline_matches_oracle(gpu, steer):
    torch = _lib.torch()
    fs, bs = 51200.0, 256
    mpos, gpos, data = _scene()
    t = torch.from_numpy(data.astype(np.float32)).to(gpu)
    p = FusedCsmBeamform(t, fs, gpos, mpos, block_size=bs, steer_type=steer, r_diag=True, ind_low=3, ind_high=60)
    out = p.run().cpu().numpy()
    csm_ref = S.welch_csm(data.astype(np.float32), bs, 'Hanning', '50%')
    csm = p.csm.cpu().numpy()
    assert np.max(np.abs(csm - csm_ref)) <= 1e-5 * np.max(np.abs(csm_ref))
    ref = B.beamformer_map(csm_ref, p.freqs, gpos, mpos, steer, True, ind=p.ind, phase_mode='reference')
    err = peak_rel_err(out, ref)
    print(f'steer={steer} max rel map err {err:.2e}')
    assert err <= 1e-4


@pytest.mark.gpu
def test_graph_replays_equal_eager_runs(gpu):
    torch = _lib.torch()
    fs, bs = 51200.0, 256
    mpos, gpos, data_a = _scene(seed=11)
    _, _, data_b = _scene(seed=29)
    a = torch.from_numpy(data_a.astype(np.float32)).to(gpu)
    b = torch.from_numpy(data_b.astype(np.float32)).to(gpu)
    t = a.clone()
    p = FusedCsmBeamform(t, fs, gpos, mpos, block_size=bs, steer_type='true level', r_diag=True)
    res_a = p.run().clone()
    csm_a = p.csm.clone()
    t.copy_(b)
    res_b = p.run().clone()
    torch.cuda.synchronize()
    p.capture()
    for i, (src, want, want_csm) in enumerate([(a, res_a, csm_a), (b, res_b, None), (a, res_a, csm_a),
                                                (b, res_b, None), (a, res_a, csm_a)]):
        # eager work of the program between replays
        junk = torch.randn(1 << 16, device=gpu).sum()
        t.copy_(src)
        p.replay()
        torch.cuda.synchronize()
        assert torch.isfinite(junk)
        assert torch.equal(p.out, want), f'replay {i} differs from the eager run'
        if want_csm is not None:
            assert torch.equal(p.csm, want_csm), f'replay {i}: CSM differs from the eager run'
