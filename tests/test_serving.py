"""Serving engine: admission, accounting and failure recovery (CPU, through a CPU executor and
an injected library error), alone-vs-in-mix results and stream ordering (GPU)."""

from collections import Counter

import numpy as np
import pytest

from acoular_amd import _lib
from acoular_amd.serving import BeamformRequest, BeamformServer, admit


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


class _StubLib:
    """Stands in for libabf where the CPU tests inject a failed launch: ``_lib.check`` maps the
    status and reads the message from here."""

    @staticmethod
    def abf_last_error():
        return b'serve batch: hipErrorLaunchFailure (injected)'


class _CpuExecutor:
    """The executor interface on the CPU: launch n fails with ABF_ERR_HIP when n is in
    ``fail_on`` (through ``_lib.check``, as a failed ``abf_beamform_batch`` does); otherwise the
    maps are a deterministic function of (request, bin, grid point), ready on the second poll."""

    def __init__(self, slots=2, fail_on=(), fail_rid=None):
        self.busy = [False] * slots
        self.fail_on = set(fail_on)
        self.fail_rid = fail_rid
        self.n = 0
        self.computed = Counter()

    @staticmethod
    def expect(rid, b, g):
        return rid * 1000.0 + b + 1e-3 * np.arange(g)

    def ready_event(self):
        return None

    def free_slot(self):
        return next((i for i, b in enumerate(self.busy) if not b), None)

    def launch(self, slot, parts):
        self.n += 1
        fail = self.n in self.fail_on or any(r.rid == self.fail_rid for r, _, _ in parts)
        _lib.check(_lib.ABF_ERR_HIP if fail else _lib.ABF_OK, 'serve batch')
        maps = []
        for r, b0, b1 in parts:
            r.num_grid = r.gpos.shape[1]
            maps.append(np.stack([self.expect(r.rid, b, r.num_grid) for b in range(b0, b1)]))
            for b in range(b0, b1):
                self.computed[(r.rid, b)] += 1
        self.busy[slot] = True
        return {'slot': slot, 'maps': maps, 'polls': 0}

    def poll(self, h):
        h['polls'] += 1
        return h['polls'] > 1

    def fetch(self, h):
        return h['maps']

    def release(self, h):
        self.busy[h['slot']] = False


def _cpu_requests(n=3, nf=10, g=7, nm=8):
    return [BeamformRequest(rid=i, arrival=0.001 * i, wavenumber=np.ones(nf), mpos=np.zeros((3, nm)),
                            csm=np.zeros((nf, nm, nm), np.complex64), gpos=np.zeros((3, g)), num_grid=g)
            for i in range(n)]


def test_failed_launch_requeues_and_keeps_finished_bins(monkeypatch):
    """A launch failing with ABF_ERR_HIP puts its requests back at the head of the queue; bins
    finished by earlier launches are kept (the per-bin bitmap, as BeamformerBase's _fr flags,
    fbeamform.py:279-285) and every bin is computed exactly once; nothing escapes the engine."""
    monkeypatch.setattr(_lib, 'lib', lambda: _StubLib)
    ex = _CpuExecutor(slots=2, fail_on={2, 5})
    reqs = _cpu_requests()
    srv = BeamformServer(hbm_budget=1 << 30, max_batch=2, max_bins_per_launch=4, executor=ex)
    srv.run(reqs, realtime=False)
    assert srv.failed_launches == 2 and all('injected' in e for e in srv.errors)
    assert all(r.error is None and r.done.all() for r in reqs)
    for r in reqs:
        np.testing.assert_array_equal(r.result, np.stack([ex.expect(r.rid, b, 7) for b in range(10)]))
    assert set(ex.computed.values()) == {1} and len(ex.computed) == 30
    assert srv.used == 0 and not srv.inflight and not srv.queue
    # chunks of at most 4 bins: three per request went through successful launches
    assert sum(srv.batch_sizes) == 9 and ex.n == srv.launches + 2


def test_request_gives_up_after_max_retries(monkeypatch):
    """A request whose every launch fails completes with the error recorded after max_retries
    retries; the others finish normally and the HBM accounting returns to zero."""
    monkeypatch.setattr(_lib, 'lib', lambda: _StubLib)
    ex = _CpuExecutor(slots=1, fail_rid=1)
    reqs = _cpu_requests()
    srv = BeamformServer(hbm_budget=1 << 30, max_batch=1, max_retries=2, executor=ex)
    srv.run(reqs, realtime=False)
    assert reqs[1].error is not None and 'injected' in reqs[1].error
    assert reqs[1].attempts == 3 and not reqs[1].done.any()
    assert reqs[0].error is None and reqs[2].error is None and reqs[0].done.all() and reqs[2].done.all()
    assert srv.failed_launches == 3 and srv.used == 0


def test_footprint_counts_the_csm():
    r = _cpu_requests(1, nf=4, g=100, nm=64)[0]
    # CSM 4 x 64 x 64 complex64 + maps + grid + re-laid-out CSM (2 x 2 upper blocks of 8 KiB per bin)
    assert r.footprint() == 8 * 4 * 64 * 64 + 4 * 4 * 100 + 24 * 100 + 4 * 3 * 8192 + 4096
    assert r.footprint(1) == 8 * 4 * 64 * 64 + 4 * 100 + 24 * 100 + 3 * 8192 + 4096
    assert r.pending_bins(3) == (0, 3)
    r.done[1] = True
    assert r.pending_bins(3) == (0, 1) and r.pending_bins() == (0, 1)
    r.done[0] = True
    assert r.pending_bins() == (2, 4)


@pytest.mark.gpu
def test_results_identical_alone_and_in_mix(gpu):
    import torch

    from acoular_amd import _lib
    from acoular_amd.fbeamform import make_job
    from acoular_amd.serving import BeamformServer, RequestFactory

    fac = RequestFactory(seed=3, g_min=200, g_max=6000, f_max=6, mic_counts=(32, 64, 96))
    reqs = fac.make(24, qps=None)
    srv = BeamformServer(hbm_budget=64 << 20, max_batch=8, num_streams=2)
    srv.run(reqs, realtime=False)
    assert srv.launches >= 3 and max(srv.batch_sizes) > 1
    assert all(r.csm is None for r in reqs)  # each request's own CSM was freed with it
    lib = _lib.lib()
    for r in reqs:
        gpos = r.grid_fn()
        csm = r.csm_fn()
        out = torch.empty((r.num_freqs, gpos.shape[1]), dtype=torch.float32, device=gpu)
        nm = r.num_mics
        job = make_job(csm, r.wavenumber, gpos, r.mpos, r.ref, r.steer_type, True, nm / (nm - 1.0),
                       'mfma_bf16x3', 'reference', out, None, clip=True)
        ws = _lib.workspace(lib.abf_beamform_workspace_size(job))
        _lib.check(lib.abf_beamform(job, _lib.ptr(ws), ws.numel(), _lib.stream_handle()))
        np.testing.assert_array_equal(out.cpu().numpy(), r.result)


@pytest.mark.gpu
@pytest.mark.parametrize('persistent', [True, False])
def test_persistent_serving_matches_alone(gpu, persistent):
    """Serving with the streamed jobs as persistent launches (the default) or as one work item per
    workgroup gives each request the maps of its launch shape run alone, bit for bit (the two
    launches of the streamed kernel compute every item alike); the mix holds streamed jobs."""
    import torch

    from acoular_amd import _lib
    from acoular_amd.fbeamform import make_job
    from acoular_amd.serving import BeamformServer, RequestFactory

    fac = RequestFactory(seed=11, g_min=3000, g_max=9000, f_max=40, mic_counts=(64, 128, 256))
    reqs = fac.make(12, qps=None)
    srv = BeamformServer(hbm_budget=1 << 30, max_batch=6, num_streams=2, persistent=persistent)
    srv.run(reqs, realtime=False)
    lib = _lib.lib()
    streamed = 0
    for r in reqs:
        gpos = r.grid_fn()
        csm = r.csm_fn()
        out = torch.empty((r.num_freqs, gpos.shape[1]), dtype=torch.float32, device=gpu)
        nm = r.num_mics
        job = make_job(csm, r.wavenumber, gpos, r.mpos, r.ref, r.steer_type, True, nm / (nm - 1.0),
                       'mfma_bf16x3', 'reference', out, None, clip=True)
        streamed += job.tile_kernel in (_lib.MFMA_KERNEL_STREAMED, _lib.MFMA_KERNEL_STREAMED_PERSISTENT)
        ws = _lib.workspace(lib.abf_beamform_workspace_size(job))
        _lib.check(lib.abf_beamform(job, _lib.ptr(ws), ws.numel(), _lib.stream_handle()))
        np.testing.assert_array_equal(out.cpu().numpy(), r.result)
    assert streamed > 0


@pytest.mark.gpu
def test_serving_results_match_oracle(gpu):
    from acoular_amd.serving import BeamformServer, RequestFactory
    from oracle import beamform as B
    from tests.conftest import peak_rel_err

    fac = RequestFactory(seed=5, g_min=100, g_max=900, f_max=3, mic_counts=(32, 64))
    reqs = fac.make(6, qps=2000.0)
    srv = BeamformServer(hbm_budget=1 << 30, max_batch=4, num_streams=2)
    srv.run(reqs, realtime=True)
    for r in reqs:
        gpos = r.grid_fn().cpu().numpy()
        csm = r.csm_fn().cpu().numpy()
        k = r.wavenumber.cpu().numpy()
        freqs = k * 343.0 / (2 * np.pi)
        ref = B.beamformer_map(csm.astype(np.complex128), freqs, gpos, r.mpos.cpu().numpy(), r.steer_type, True)
        assert peak_rel_err(r.result, ref) <= 1e-4
        assert r.t_done >= r.arrival


@pytest.mark.gpu
def test_fresh_csm_without_synchronize(gpu):
    """A CSM still being computed on the caller's stream when the request is submitted (no
    synchronisation in between) is read only after it is written: the launch waits for an event
    recorded at submission."""
    import torch

    from acoular_amd import synthetic
    from oracle import beamform as B
    from tests.conftest import peak_rel_err

    nm = 48
    mpos = synthetic.disk_array(nm, 1.0, 4)
    spos = synthetic.source_positions(3, [-0.3, -0.3, 0.6], [0.3, 0.3, 0.6], 5)
    freqs = np.array([2000.0, 6000.0, 12000.0])
    x = np.linspace(-0.4, 0.4, 41)
    gpos = np.stack([np.repeat(x, 41), np.tile(x, 41), np.full(41 * 41, 0.6)])
    srv = BeamformServer(hbm_budget=1 << 30, max_batch=4, num_streams=2)
    reqs = []
    for i in range(3):
        # a long chain of device work ahead of the CSM's last write on the current stream
        big = torch.randn(4096, 4096, device=gpu)
        for _ in range(4):
            big = big @ big * 1e-3
        csm = synthetic.analytic_csm(mpos, spos, np.ones(3), freqs, device=gpu) * (1.0 + 0.0 * big[0, 0])
        r = BeamformRequest(rid=i, arrival=0.0, wavenumber=_lib.to_device(2 * np.pi * freqs / 343.0),
                            mpos=_lib.to_device(mpos), csm=csm, gpos=_lib.to_device(gpos), steer_type='true level',
                            num_grid=gpos.shape[1])
        srv.submit(r)
        reqs.append(r)
    t = 0.0
    while any(not r.finished for r in reqs):
        srv.retire(lambda: t)
        srv.admit_and_launch(lambda: t)
    want = B.beamformer_map(synthetic.analytic_csm(mpos, spos, np.ones(3), freqs).astype(np.complex128), freqs,
                            gpos, mpos, 'true level', True)
    for r in reqs:
        assert peak_rel_err(r.result, want) <= 1e-4


@pytest.mark.gpu
def test_chunked_launches_match_oracle_and_mix(gpu):
    """Bands launched in chunks of bins: every chunk filed into its place, results equal the
    oracle and are bit-identical alone and in the mix with the same chunking."""
    from acoular_amd.serving import RequestFactory
    from oracle import beamform as B
    from tests.conftest import peak_rel_err

    fac = RequestFactory(seed=8, g_min=100, g_max=1500, f_max=9, mic_counts=(32, 64))
    reqs = fac.make(8, qps=None)
    srv = BeamformServer(hbm_budget=1 << 30, max_batch=4, num_streams=2, max_bins_per_launch=3)
    srv.run(reqs, realtime=False)
    assert srv.launches > len(reqs) // 4
    for r in reqs:
        alone = BeamformServer(hbm_budget=1 << 30, max_batch=1, num_streams=1, max_bins_per_launch=3)
        twin = fac.make(8, qps=None)[r.rid]
        alone.run([twin], realtime=False)
        np.testing.assert_array_equal(twin.result, r.result)
        gpos = r.grid_fn().cpu().numpy()
        k = r.wavenumber.cpu().numpy()
        ref = B.beamformer_map(r.csm_fn().cpu().numpy().astype(np.complex128), k * 343.0 / (2 * np.pi), gpos,
                               r.mpos.cpu().numpy(), r.steer_type, True)
        assert peak_rel_err(r.result, ref) <= 1e-4
