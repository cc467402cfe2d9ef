"""Parity at the full sizes of the benchmark configurations, through the C ABI, against the oracle.

* 256 mics, 201 x 201 grid, 513 bins (analytic complex64 CSM of 8 sources): the plain FP32,
  fp32 MFMA and split-bf16 MFMA forms (and the opt-in f16x2 form at its own gate) for
  'classic' and 'inverse' with and without r_diag, every bin computed, the oracle on 4 bins.
* 512 mics, 10 s at 51.2 kHz, Welch block size 2048 (499 blocks, 1025 bins) + 128 x 128
  'true location' maps of every bin, replayed as one hipGraph: CSM and maps on 2 bins.
* 1024 mics, RectGrid3D 64 x 64 x 32, 256 bins ('true level', r_diag), split-bf16 MFMA and plain:
  8,192 sampled points (plus the map maximum) of 2 bins.
* 1024 mics x 257 bins of a Welch CSM (above 2^31 bytes) into the MFMA beamformer: the last bins.
* the serving mix: 200 requests drawn from the mixed-size distribution plus one 512-mic,
  250k-point, 64-bin request; bit-identical alone and in the mix, the oracle on sampled points.

Gate: peak-normalised map error max_f max_g |B - B_ref| / max_g |B_ref| <= 1e-4 (5e-4 for f16x2),
CSMs to 1e-5 of their largest element.
"""

import numpy as np
import pytest

from acoular_amd import _lib, workloads
from acoular_amd.fbeamform import make_job
from oracle import beamform as B
from oracle import spectra as S
from tests.conftest import peak_rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-4
TOL_F16X2 = 5e-4


def _run(csm, k, gpos, mpos, steer, r_diag, form, ref=np.zeros(3)):
    """One abf_beamform job over every bin of ``csm``; returns the float32 device maps."""
    torch = _lib.torch()
    nm = mpos.shape[1]
    out = torch.empty((csm.shape[0], gpos.shape[1]), dtype=torch.float32, device=csm.device)
    job = make_job(csm, k, gpos, mpos, ref, steer, r_diag, nm / (nm - 1.0) if r_diag else 1.0, form, 'reference', out,
                   None, clip=r_diag)
    lib = _lib.lib()
    ws = _lib.workspace(lib.abf_beamform_workspace_size(job))
    print(f'launch {form} {steer} r_diag={r_diag} M={nm} G={gpos.shape[1]} F={csm.shape[0]}', flush=True)
    _lib.check(lib.abf_beamform(job, _lib.ptr(ws), ws.numel(), _lib.stream_handle()), form)
    torch.cuda.synchronize()
    return out


# ---------------------------------------------------------------------------------------------
# 256 mics, 201 x 201, 513 bins
# ---------------------------------------------------------------------------------------------
@pytest.fixture(scope='module')
def plain_vs_mfma(gpu):
    wl = workloads.plain_vs_mfma(device=gpu)
    wl['k'] = _lib.to_device(2 * np.pi * wl['freqs'] / 343.0)
    wl['gd'] = _lib.to_device(wl['gpos'])
    wl['md'] = _lib.to_device(wl['mpos'])
    return wl


@pytest.mark.parametrize('steer', ['classic', 'inverse'])
@pytest.mark.parametrize('r_diag', [True, False])
def test_plain_vs_mfma_full_size(plain_vs_mfma, steer, r_diag):
    wl = plain_vs_mfma
    bins = [3, 100, 257, 511]
    ref = B.beamformer_map(wl['csm'][bins].cpu().numpy().astype(np.complex128), wl['freqs'][bins], wl['gpos'],
                           wl['mpos'], steer, r_diag, ind=np.arange(len(bins)))
    for form in ('plain_f32', 'mfma_f32', 'mfma_bf16x3', 'mfma_f16x2'):
        out = _run(wl['csm'], wl['k'], wl['gd'], wl['md'], steer, r_diag, form)
        assert bool(torch_isfinite(out)), form
        got = out[bins].double().cpu().numpy()
        err = peak_rel_err(got, ref)
        print(f'{form} {steer} r_diag={r_diag}: max relative map error {err:.2e} on bins {bins}')
        assert err <= (TOL_F16X2 if form == 'mfma_f16x2' else TOL), (form, err)
        del out


def torch_isfinite(t):
    return _lib.torch().isfinite(t).all().item()


# ---------------------------------------------------------------------------------------------
# 512 mics, Welch bs 2048 + 128 x 128 maps, one hipGraph
# ---------------------------------------------------------------------------------------------
def test_fused_pipeline_full_size(gpu):
    import torch

    from acoular_amd.pipeline import FusedCsmBeamform

    wl = workloads.fused()
    t = wl['signals']
    assert tuple(t.shape) == (512000, 512)
    p = FusedCsmBeamform(t, wl['fs'], wl['gpos'], wl['mpos'], block_size=2048, steer_type='true location',
                         r_diag=True, ind_low=0, ind_high=1025)
    assert p.num_blocks == 499 and p.map_bins == 1025 * 16384
    print('fused: eager run', flush=True)
    eager = p.run().clone()
    p.capture()
    print('fused: graph replays', flush=True)
    for _ in range(2):
        p.replay()
        torch.cuda.synchronize()
        assert torch.equal(p.out, eager)
    bins = [200, 700]
    data = t.cpu().numpy()
    csm_ref = S.welch_csm_bins(data, 2048, bins)
    got_csm = p.csm[bins].cpu().numpy()
    cerr = np.abs(got_csm - csm_ref).max() / np.abs(csm_ref).max()
    print(f'fused CSM error {cerr:.2e} on bins {bins}')
    assert cerr <= 1e-5
    freqs = np.asarray(bins, dtype=float) * wl['fs'] / 2048
    ref = B.beamformer_map(csm_ref, freqs, wl['gpos'], wl['mpos'], 'true location', True, ind=np.arange(2))
    err = peak_rel_err(eager[bins].double().cpu().numpy(), ref)
    print(f'fused maps error {err:.2e} on bins {bins}')
    assert err <= TOL


# ---------------------------------------------------------------------------------------------
# 1024 mics, 64 x 64 x 32, 256 bins
# ---------------------------------------------------------------------------------------------
def test_large3d_full_size(gpu):
    wl = workloads.large3d(device=gpu)
    gpos, mpos, freqs = wl['gpos'], wl['mpos'], wl['freqs']
    assert gpos.shape[1] == 131072 and wl['csm'].shape == (256, 1024, 1024)
    k = _lib.to_device(2 * np.pi * freqs / 343.0)
    gd, md = _lib.to_device(gpos), _lib.to_device(mpos)
    rng = np.random.default_rng(4007)
    bins = [40, 200]
    refs = {}
    for form in ('mfma_bf16x3', 'plain_f32'):
        out = _run(wl['csm'], k, gd, md, 'true level', True, form)
        for b in bins:
            got = out[b].double().cpu().numpy()
            if b not in refs:
                sub = np.unique(np.concatenate([rng.choice(gpos.shape[1], 8192, replace=False),
                                                [int(np.argmax(got))]]))
                refs[b] = (sub, B.beamformer_map(wl['csm'][b:b + 1].cpu().numpy().astype(np.complex128),
                                                 freqs[b:b + 1], gpos[:, sub], mpos, 'true level', True)[0])
            sub, ref = refs[b]
            err = float(np.max(np.abs(got[sub] - ref)) / np.max(np.abs(ref)))
            print(f'{form} bin {b}: max relative map error {err:.2e} on {len(sub)} points')
            assert err <= TOL
        del out


def test_welch_csm_above_2gib_into_mfma(gpu):
    """A Welch CSM of 1024 mics x 257 bins (2.16e9 bytes, above 2^31) and the MFMA beamformer over
    all its bins: the last bins, whose byte offsets pass 2^31, match the oracle."""
    import torch

    import acoular_amd as aa
    from acoular_amd import synthetic

    nm, bs, n = 1024, 512, 512 * 9
    mpos = synthetic.spiral_array(nm, 2.0, 71)
    data = synthetic.point_source_signals(mpos, np.array([[0.1], [-0.2], [0.8]]), n, 51200.0, seed=72,
                                          noise_rel=0.1).astype(np.float32)
    ps = aa.PowerSpectra(source=aa.TimeSamples(data=data, sample_freq=51200.0), block_size=bs, window='Hanning',
                         overlap='50%', precision='complex64')
    csm = ps.csm_device
    assert csm.shape == (257, nm, nm) and csm.numel() * 8 > 2**31
    bins = [3, 255, 256]
    want = S.welch_csm_bins(data, bs, bins)
    got = csm[bins].cpu().numpy()
    cerr = np.abs(got - want).max() / np.abs(want).max()
    print(f'CSM error {cerr:.2e} on bins {bins}')
    assert cerr <= 1e-5
    x = np.linspace(-0.5, 0.5, 21)
    gpos = np.stack([np.repeat(x, 21), np.tile(x, 21), np.full(441, 0.8)])
    freqs = np.arange(257) * 51200.0 / bs
    out = _run(csm, _lib.to_device(2 * np.pi * freqs / 343.0), _lib.to_device(gpos), _lib.to_device(mpos),
               'true level', True, 'mfma_bf16x3')
    ref = B.beamformer_map(want, freqs[bins], gpos, mpos, 'true level', True, ind=np.arange(3))
    err = peak_rel_err(out[bins].double().cpu().numpy(), ref)
    print(f'maps error {err:.2e} on bins {bins}')
    assert err <= TOL
    del csm, out
    torch.cuda.empty_cache()


# ---------------------------------------------------------------------------------------------
# serving mix
# ---------------------------------------------------------------------------------------------
def test_serving_mix_full_distribution(gpu):
    import torch

    from acoular_amd.serving import BeamformRequest, BeamformServer, RequestFactory

    fac = RequestFactory(seed=7)
    reqs = fac.make(200, qps=None, seed_offset=31)
    big = fac.request(200, reqs[-1].arrival, 512, 250000, False, 300, 64, 'true location', 99)
    reqs.append(big)
    assert big.num_grid == 250000 and big.num_freqs == 64 and big.num_mics == 512
    print(f'serving: {len(reqs)} requests, {sum(r.map_bins for r in reqs):.3e} map-bins', flush=True)
    srv = BeamformServer(hbm_budget=24 << 30, max_batch=64, num_streams=2)
    srv.run(reqs, realtime=False)
    assert srv.failed_launches == 0 and max(srv.batch_sizes) > 1
    assert all(r.error is None and r.done.all() for r in reqs)
    rng = np.random.default_rng(5)
    worst = 0.0
    for r in reqs:
        # alone: a fresh request with the same draws, on its own server
        twin = BeamformRequest(rid=r.rid, arrival=0.0, wavenumber=r.wavenumber, mpos=r.mpos, csm_fn=r.csm_fn,
                               grid_fn=r.grid_fn, steer_type=r.steer_type, r_diag=True, num_grid=r.num_grid)
        alone = BeamformServer(hbm_budget=24 << 30, max_batch=1, num_streams=1)
        alone.run([twin], realtime=False)
        np.testing.assert_array_equal(twin.result, r.result)
        # oracle: one bin, 256 sampled points plus the map maximum
        b = r.num_freqs // 2
        gpos = r.grid_fn().cpu().numpy()
        sub = np.unique(np.concatenate([rng.choice(gpos.shape[1], min(256, gpos.shape[1]), replace=False),
                                        [int(np.argmax(r.result[b]))]]))
        k = r.wavenumber[b:b + 1].cpu().numpy()
        csm_b = r.csm_fn()[b:b + 1].cpu().numpy().astype(np.complex128)
        ref = B.beamformer_map(csm_b, k * 343.0 / (2 * np.pi), gpos[:, sub], r.mpos.cpu().numpy(), r.steer_type,
                               True)[0]
        if np.max(np.abs(ref)) > 0:
            err = float(np.max(np.abs(r.result[b][sub] - ref)) / np.max(np.abs(ref)))
            worst = max(worst, err)
            assert err <= TOL, (r.rid, r.num_mics, r.num_grid, r.num_freqs, err)
    print(f'serving mix: worst sampled error {worst:.2e}')
    torch.cuda.empty_cache()
