"""Autotuner of the MFMA launch shape: shape classes and the JSON cache on the CPU, a tuning run
and the tuned result against the oracle on the GPU."""

import json

import numpy as np
import pytest

import acoular_amd as aa
from acoular_amd import _lib, autotune, synthetic
from acoular_amd.fbeamform import make_job
from oracle import beamform as B
from tests.conftest import peak_rel_err


def test_shape_key_buckets():
    assert autotune.shape_key('mfma_bf16x3', 256, 40401, 513) == 'mfma_bf16x3|mp256|g32768|f64'
    assert autotune.shape_key('mfma_f32', 33, 1000, 3) == 'mfma_f32|mp64|g1024|f4'
    # neighbours of one class share a key, the padded mic count separates classes
    assert autotune.shape_key('mfma_bf16x3', 250, 40000, 500) == autotune.shape_key('mfma_bf16x3', 256, 41000, 600)
    assert autotune.shape_key('mfma_bf16x3', 257, 40000, 8) != autotune.shape_key('mfma_bf16x3', 256, 40000, 8)


def test_cache_round_trip_and_apply(tmp_path):
    path = str(tmp_path / 'tune.json')
    c = autotune.TileCache(path)
    key = autotune.shape_key('mfma_bf16x3', 64, 2000, 16)
    c.put(key, {'tile_kernel': 2, 'tile_cols': 32, 'tile_freqs': 8, 'seconds': 1e-3})
    with open(path) as fh:
        assert json.load(fh)[key]['tile_cols'] == 32
    c2 = autotune.TileCache(path)
    assert c2.get(key)['tile_kernel'] == 2
    job = _lib.AbfGridJob()
    job.form = _lib.FORMS['mfma_bf16x3']
    job.geom.num_mics, job.geom.num_grid, job.num_freqs = 64, 2000, 16
    job.tile_freqs = 4  # set by the caller: kept
    assert autotune.apply(job, c2)
    assert (job.tile_kernel, job.tile_cols, job.tile_freqs) == (2, 32, 4)
    job.form = _lib.FORMS['plain_f32']
    assert not autotune.apply(job, c2)


def test_cache_ignores_corrupt_file(tmp_path):
    path = tmp_path / 'tune.json'
    path.write_text('{not json')
    assert autotune.TileCache(str(path)).entries == {}


@pytest.mark.gpu
def test_tune_job_picks_a_correct_shape(gpu, tmp_path):
    torch = _lib.torch()
    freqs = np.linspace(1000.0, 15000.0, 12)
    nm = 64
    mpos = synthetic.disk_array(nm, 1.0, 3)
    spos = synthetic.source_positions(2, [-0.2, -0.2, 0.5], [0.2, 0.2, 0.5], 4)
    csm = synthetic.analytic_csm(mpos, spos, synthetic.level_powers(2, 6.0, 5), freqs)
    grid = aa.RectGrid(x_min=-0.4, x_max=0.4, y_min=-0.4, y_max=0.4, z=0.5, increment=0.02)
    mics = aa.MicGeom(pos_total=mpos)
    sv = aa.SteeringVector(grid=grid, mics=mics, steer_type='true level')
    gpos, mpos_d = sv.device_geometry()
    dcsm = _lib.to_device(csm, torch.complex64)
    k = torch.from_numpy(2 * np.pi * freqs / 343.0).to(gpu)
    out = torch.empty((len(freqs), grid.size), dtype=torch.float32, device=gpu)
    cache = autotune.TileCache(str(tmp_path / 'tune.json'))
    job = make_job(dcsm, k, gpos, mpos_d, sv.ref, 'true level', True, nm / (nm - 1.0), 'mfma_bf16x3', 'reference',
                   out, None, True)
    job.tile_kernel = job.tile_cols = job.tile_freqs = 0
    res = autotune.tune_job(job, reps=1, tile_cols=(16, 32, 64), tile_freqs=(2, 8), tiles=cache)
    print(res['tried'])
    assert len(res['tried']) >= 3
    assert res['seconds'] == min(t['seconds'] for t in res['tried'])
    assert cache.get(autotune.job_key(job))['tile_cols'] == res['tile_cols']
    assert autotune.apply(job, cache)
    lib = _lib.lib()
    ws = _lib.workspace(lib.abf_beamform_workspace_size(job))
    _lib.check(lib.abf_beamform(job, _lib.ptr(ws), ws.numel(), _lib.stream_handle()))
    ref = B.beamformer_map(csm.astype(np.complex128), freqs, grid.pos, mics.pos, 'true level', True,
                           ind=np.arange(len(freqs)), phase_mode='reference')
    assert peak_rel_err(out.cpu().numpy(), ref) <= 1e-4
