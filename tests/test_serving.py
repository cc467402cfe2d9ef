"""Serving engine: admission logic (CPU) and alone-vs-in-mix results (GPU)."""

import numpy as np
import pytest

from acoular_amd.serving import admit


class _Req:
    def __init__(self, n):
        self.n = n

    def footprint(self):
        return self.n


def test_admission_respects_budget_order_and_batch_cap():
    q = [_Req(4), _Req(3), _Req(5), _Req(1)]
    b = admit(q, used=0, budget=10, max_batch=8, any_inflight=False)
    assert [r.n for r in b] == [4, 3]  # 4+3 fits, +5 would not; FIFO: the 1 behind waits
    assert [r.n for r in q] == [5, 1]
    b = admit(q, used=7, budget=10, max_batch=8, any_inflight=True)
    assert b == []
    q2 = [_Req(1)] * 5
    assert len(admit(list(q2), 0, 100, 3, False)) == 3


def test_oversized_request_runs_alone_when_idle():
    q = [_Req(50), _Req(1)]
    assert admit(q, 0, 10, 4, any_inflight=True) == []
    b = admit(q, 0, 10, 4, any_inflight=False)
    assert [r.n for r in b] == [50]


@pytest.mark.gpu
def test_results_identical_alone_and_in_mix(gpu):
    import torch

    from acoular_amd import _lib
    from acoular_amd.fbeamform import make_job
    from acoular_amd.serving import BeamformServer, RequestFactory

    fac = RequestFactory(seed=3, pool=1, g_min=200, g_max=6000, f_max=6, mic_counts=(32, 64, 96))
    reqs = fac.make(24, qps=None)
    srv = BeamformServer(hbm_budget=64 << 20, max_batch=8, num_streams=2)
    srv.run(reqs, realtime=False)
    assert srv.launches >= 3 and max(srv.batch_sizes) > 1
    lib = _lib.lib()
    for r in reqs:
        gpos = r.grid_fn()
        out = torch.empty((r.num_freqs, gpos.shape[1]), dtype=torch.float32, device=gpu)
        nm = r.num_mics
        job = make_job(r.csm, r.wavenumber, gpos, r.mpos, r.ref, r.steer_type, True, nm / (nm - 1.0),
                       'mfma_bf16x3', 'reference', out, None, clip=True)
        ws = _lib.workspace(lib.abf_beamform_workspace_size(job))
        _lib.check(lib.abf_beamform(job, _lib.ptr(ws), ws.numel(), _lib.stream_handle()))
        np.testing.assert_array_equal(out.cpu().numpy(), r.result)


@pytest.mark.gpu
def test_serving_results_match_oracle(gpu):
    from acoular_amd.serving import BeamformServer, RequestFactory
    from oracle import beamform as B
    from tests.conftest import peak_rel_err

    fac = RequestFactory(seed=5, pool=1, g_min=100, g_max=900, f_max=3, mic_counts=(32, 64))
    reqs = fac.make(6, qps=2000.0)
    srv = BeamformServer(hbm_budget=1 << 30, max_batch=4, num_streams=2)
    srv.run(reqs, realtime=True)
    for r in reqs:
        gpos = r.grid_fn().cpu().numpy()
        csm = r.csm.cpu().numpy()
        k = r.wavenumber.cpu().numpy()
        freqs = k * 343.0 / (2 * np.pi)
        ref = B.beamformer_map(csm.astype(np.complex128), freqs, gpos, r.mpos.cpu().numpy(), r.steer_type, True)
        assert peak_rel_err(r.result, ref) <= 1e-4
        assert r.t_done >= r.arrival
