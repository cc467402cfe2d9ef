"""abf_beamform_noalloc (workspace in a static device array, no allocation) against abf_beamform
(caller workspace), through the C ABI: two calls in a row, the second on a smaller job, for the
split-bf16 MFMA form (a job larger than one chunk of the array) and the plain form; the map
values agree to within 1e-6 of each map's peak (the beamformer's own parity gate is 1e-4)."""

import ctypes

import numpy as np
import pytest

from acoular_amd import _lib, synthetic
from acoular_amd.fbeamform import make_job

pytestmark = pytest.mark.gpu


def _job(gpu, nm, g_side, nf, form, steer, seed):
    torch = _lib.torch()
    mpos = synthetic.disk_array(nm, 1.5, seed)
    spos = synthetic.source_positions(4, [-0.4, -0.4, 0.7], [0.4, 0.4, 0.7], seed + 1)
    freqs = np.linspace(300.0, 20000.0, nf)
    csm = synthetic.analytic_csm(mpos, spos, synthetic.level_powers(4, 10.0, seed + 2), freqs, device=gpu)
    x = np.linspace(-0.6, 0.6, g_side)
    gpos = np.stack([np.repeat(x, g_side), np.tile(x, g_side), np.full(g_side * g_side, 0.7)])
    k = _lib.to_device(2 * np.pi * freqs / 343.0)
    gd, md = _lib.to_device(gpos), _lib.to_device(mpos)
    outs = [torch.full((nf, gpos.shape[1]), -7.0, dtype=torch.float32, device=gpu) for _ in range(2)]
    jobs = [make_job(csm, k, gd, md, np.zeros(3), steer, True, nm / (nm - 1.0), form, 'reference', o, None,
                     clip=True) for o in outs]
    return jobs, outs, (csm, k, gd, md)


def _compare(gpu, nm, g_side, nf, form, steer, seed):
    torch = _lib.torch()
    lib = _lib.lib()
    jobs, outs, keep = _job(gpu, nm, g_side, nf, form, steer, seed)
    ws = _lib.workspace(lib.abf_beamform_workspace_size(jobs[0]))
    print(f'abf_beamform {form} M={nm} G={g_side * g_side} F={nf}', flush=True)
    _lib.check(lib.abf_beamform(jobs[0], _lib.ptr(ws), ws.numel(), _lib.stream_handle()))
    print(f'abf_beamform_noalloc {form} M={nm} G={g_side * g_side} F={nf}', flush=True)
    _lib.check(lib.abf_beamform_noalloc(jobs[1], _lib.stream_handle()), 'noalloc')
    torch.cuda.synchronize()
    a, b = outs[0].cpu().numpy(), outs[1].cpu().numpy()
    assert np.isfinite(b).all() and (b != -7.0).all()
    peak = np.abs(a).max(axis=1, keepdims=True)
    peak[peak == 0] = 1.0
    err = float((np.abs(a - b) / peak).max())
    print(f'  max difference {err:.2e} of the peak, identical: {np.array_equal(a, b)}')
    assert err <= 1e-6
    return keep


@pytest.mark.parametrize('form,steer', [('mfma_bf16x3', 'classic'), ('plain_f32', 'true level')])
def test_noalloc_matches_workspace_entry_point(gpu, form, steer):
    lib = _lib.lib()
    # first call: 256 mics x 513 bins, a re-laid-out CSM of 151 MB for the MFMA form (three chunks of
    # the 64 MiB array); second call: a smaller job on the same array
    big = (256, 41, 513) if form == 'mfma_bf16x3' else (160, 31, 200)
    _compare(gpu, *big, form, steer, 11)
    _compare(gpu, 48, 23, 7, form, steer, 12)
    assert lib.abf_beamform_noalloc(None, None) == _lib.ABF_ERR_INVALID


def test_noalloc_rejects_a_bin_larger_than_the_array(gpu):
    lib = _lib.lib()
    j = _lib.AbfGridJob()
    j.geom.num_grid = 10
    j.geom.num_mics = 4096  # one plain bin: 4096^2 complex64 = 128 MiB
    j.num_freqs = 1
    j.steer_type = 1
    j.form = _lib.FORMS['plain_f32']
    for name in ('csm', 'wavenumber', 'out'):
        setattr(j, name, 256)
    j.geom.gpos = j.geom.mpos = 256
    rc = lib.abf_beamform_noalloc(ctypes.byref(j), None)
    assert rc == _lib.ABF_ERR_WORKSPACE
    assert b'scratch' in lib.abf_last_error()
