"""Continuous-batching beamforming server on one GPU.

Requests (a CSM of a band of bins, a grid, microphone positions and the steering parameters)
arrive over time.  The engine loop

1. retires finished launches and files their maps, bin by bin, into the requests' results;
   each request keeps a per-bin completion bitmap, as ``BeamformerBase`` keeps its ``_fr``
   flags (``fbeamform.py:279-285``), and only its missing bins are ever launched again;
2. admits waiting requests in arrival order while the admitted HBM footprint (the request's CSM,
   its grid, the re-laid-out CSM and the maps of the bins launched) stays under the budget and a
   launch slot is free;
3. launches the admitted requests as one ``abf_beamform_batch`` call (one MFMA launch per tile
   shape, work items of every request interleaved) on a free stream, so requests that arrive
   while earlier launches run overlap with them on the device.

A launch that fails (a library error, or the allocator out of memory) does not stop the engine:
its requests go back to the head of the queue with their finished bins kept, and a request
that fails ``max_retries`` times in a row is completed with its error recorded.  With
``max_bins_per_launch`` a request is launched in chunks of contiguous bins, so a failure costs
at most one chunk.

Each map value is computed by one workgroup with a fixed reduction order, so a request's result
is bit-identical alone and in any mix.  The latency of a request is measured from its arrival to
the moment its last map is in host memory.

The device side sits behind a small executor interface (``GpuExecutor``: free slot, launch,
poll, fetch, release), so the engine's scheduling, accounting and recovery logic is plain host
code that a CPU executor can drive in tests.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .fbeamform import make_job

_STEER = ('classic', 'inverse', 'true level', 'true location')


@dataclass
class BeamformRequest:
    rid: int
    arrival: float                 # seconds from the start of the run
    wavenumber: object             # device float64 tensor [F]
    mpos: object                   # device float64 [3, M]
    csm: object | None = None      # device complex64 [F, M, M], or None with csm_fn
    gpos: object | None = None     # device float64 [3, G], or None with grid_fn
    steer_type: str = 'true level'
    r_diag: bool = True
    ref: object = field(default_factory=lambda: np.zeros(3))
    grid_fn: object = None         # callable returning gpos on the device at admission
    csm_fn: object = None          # callable building this request's own CSM on the device at admission
    num_grid: int = 0
    result: np.ndarray | None = None
    done: np.ndarray | None = None  # per-bin completion flags
    attempts: int = 0              # consecutive failed launches
    error: str | None = None
    t_admit: float = 0.0
    t_done: float = 0.0
    ready: object = None           # event recorded on the submitting stream (caller-provided tensors)

    @property
    def num_mics(self) -> int:
        return int(self.mpos.shape[1])

    @property
    def num_freqs(self) -> int:
        return int(self.wavenumber.shape[0])

    @property
    def map_bins(self) -> int:
        return self.num_freqs * self.num_grid

    @property
    def finished(self) -> bool:
        return self.error is not None or (self.done is not None and bool(self.done.all()))

    def pending_bins(self, max_bins: int | None = None) -> tuple[int, int]:
        """The next contiguous run of bins not yet computed, at most ``max_bins`` long."""
        if self.done is None:
            self.done = np.zeros(self.num_freqs, dtype=bool)
        todo = np.flatnonzero(~self.done)
        b0 = int(todo[0])
        b1 = b0 + 1
        while b1 < self.num_freqs and not self.done[b1] and (max_bins is None or b1 - b0 < max_bins):
            b1 += 1
        return b0, b1

    def footprint(self, nbins: int | None = None) -> int:
        """HBM bytes a launch of ``nbins`` of this request's bins (all by default) holds while in
        flight: the request's CSM, its grid, the re-laid-out CSM and the maps of those bins."""
        nb = self.num_freqs if nbins is None else nbins
        nm = self.num_mics
        mp = (nm + 31) // 32 * 32
        nk = mp // 32
        prep = nb * nk * (nk + 1) // 2 * 8192
        csm = 8 * self.num_freqs * nm * nm
        return csm + 4 * nb * self.num_grid + 24 * self.num_grid + prep + 4096


def admit(queue: list, used: int, budget: int, max_batch: int, any_inflight: bool, max_bins: int | None = None) -> list:
    """Pop the requests of the next launch from the head of ``queue`` (arrival order): as many
    as fit the HBM budget next to the ``used`` bytes of launches in flight, at most ``max_batch``.
    A request larger than the whole budget is admitted alone once nothing else is in flight, so
    it cannot block the queue forever."""
    batch = []
    nbytes = 0
    while queue and len(batch) < max_batch:
        r = queue[0]
        if hasattr(r, 'pending_bins'):
            b0, b1 = r.pending_bins(max_bins)
            fp = r.footprint(b1 - b0)
        else:
            fp = r.footprint()
        if used + nbytes + fp > budget and (batch or any_inflight):
            break
        batch.append(queue.pop(0))
        nbytes += fp
    return batch


class _Launch:
    """One batched launch: ``parts`` = [(request, b0, b1)], the executor's handle and the bytes
    it holds."""

    def __init__(self, parts, nbytes, handle=None):
        self.parts = parts
        self.bytes = nbytes
        self.handle = handle


class GpuExecutor:
    """Launch slots on default-priority HIP streams, one ``abf_beamform_batch`` call per launch,
    maps copied to pinned host memory behind it, completion polled through a HIP event."""

    def __init__(self, form: str = 'mfma_bf16x3', phase_mode: str = 'reference', num_streams: int = 2,
                 staging_bytes: int = 0, persistent: bool | None = None):
        """``staging_bytes`` > 0 gives every stream a reusable pinned host buffer of that size for the
        device-to-host copies instead of a pinned allocation per launch.  ``persistent=True`` runs the
        jobs whose launch shape is the streamed kernel as its persistent launch (one workgroup per CU
        walking the work items round robin; same tiles, same maps bit for bit), ``False`` as one
        work item per workgroup, ``None`` (default) as the tile table says."""
        torch = _lib.torch()
        self.dev = _lib.device()
        self.form = form
        self.phase_mode = phase_mode
        self.persistent = persistent
        # default-priority streams only
        self.streams = [torch.cuda.Stream(device=self.dev) for _ in range(num_streams)]
        self.busy = [False] * num_streams
        self.staging = [torch.empty(int(staging_bytes), dtype=torch.uint8, pin_memory=True) if staging_bytes > 0
                        else None for _ in self.streams]

    def ready_event(self):
        """An event on the caller's current stream: work a request's tensors depend on."""
        torch = _lib.torch()
        ev = torch.cuda.Event()
        ev.record(torch.cuda.current_stream())
        return ev

    def free_slot(self):
        for i, b in enumerate(self.busy):
            if not b:
                return i
        return None

    def _batch_call(self, arr, n, ws, stream) -> int:
        lib = _lib.lib()
        return lib.abf_beamform_batch(arr, n, _lib.ptr(ws), ws.numel(), _lib.ctypes.c_void_p(stream.cuda_stream))

    def launch(self, slot: int, parts) -> dict:
        """Queue one launch of ``parts`` on stream ``slot``; raises (nothing left in flight) on a
        library error or when the allocator runs out of memory."""
        torch = _lib.torch()
        stream = self.streams[slot]
        keep = []
        with torch.cuda.stream(stream):
            jobs, outs = [], []
            for r, b0, b1 in parts:
                if r.ready is not None:
                    stream.wait_event(r.ready)
                if r.csm is None:
                    r.csm = r.csm_fn()
                if r.gpos is None:
                    r.gpos = r.grid_fn()
                r.num_grid = int(r.gpos.shape[1])
                for t in (r.csm, r.wavenumber, r.gpos, r.mpos):
                    t.record_stream(stream)  # not reused by the allocator before this launch is done
                out = torch.empty((b1 - b0, r.num_grid), dtype=torch.float32, device=self.dev)
                nm = r.num_mics
                norm = nm / (nm - 1.0) if r.r_diag else 1.0
                job = make_job(r.csm[b0:b1], r.wavenumber[b0:b1], r.gpos, r.mpos, r.ref, r.steer_type, r.r_diag,
                               norm, self.form, self.phase_mode, out, None, clip=r.r_diag)
                if self.persistent and job.tile_kernel == _lib.MFMA_KERNEL_STREAMED:
                    job.tile_kernel = _lib.MFMA_KERNEL_STREAMED_PERSISTENT
                elif self.persistent is False and job.tile_kernel == _lib.MFMA_KERNEL_STREAMED_PERSISTENT:
                    job.tile_kernel = _lib.MFMA_KERNEL_STREAMED
                jobs.append(job)
                outs.append(out)
            arr = (_lib.AbfGridJob * len(jobs))(*jobs)
            nb = _lib.lib().abf_beamform_batch_workspace_size(arr, len(jobs), None)
            ws = torch.empty(max(nb, 1), dtype=torch.uint8, device=self.dev)
            _lib.check(self._batch_call(arr, len(jobs), ws, stream), 'serve batch')
            stage = self.staging[slot]
            host, off = [], 0
            for o in outs:
                nbytes = o.numel() * o.element_size()
                if stage is not None and off + nbytes <= stage.numel():
                    h = stage[off:off + nbytes].view(o.dtype).view(o.shape)
                    off += (nbytes + 255) // 256 * 256
                else:
                    h = torch.empty(o.shape, dtype=o.dtype, pin_memory=True)
                h.copy_(o, non_blocking=True)
                host.append(h)
            ev = torch.cuda.Event()
            ev.record(stream)
        keep += outs + [ws]
        self.busy[slot] = True
        return {'slot': slot, 'event': ev, 'host': host, 'tensors': keep}

    def poll(self, h) -> bool:
        return h['event'].query()

    def fetch(self, h) -> list:
        """Host maps of a finished launch, one [b1 - b0, G] array per part (copies when they sit
        in the reusable staging buffer)."""
        return [x.numpy().copy() if self.staging[h['slot']] is not None else x.numpy() for x in h['host']]

    def release(self, h) -> None:
        self.busy[h['slot']] = False
        h['tensors'] = None


class BeamformServer:
    """Admission by HBM footprint + continuous batching over the executor's launch slots.

    ``batched=False`` with ``form='plain_f32'`` and ``num_streams=1`` gives the unbatched plain
    baseline: one request per launch, one launch at a time.
    """

    def __init__(self, hbm_budget: int = 32 << 30, max_batch: int = 64, num_streams: int = 2,
                 form: str = 'mfma_bf16x3', batched: bool = True, phase_mode: str = 'reference',
                 keep_results: bool = True, staging_bytes: int = 0, max_bins_per_launch: int | None = None,
                 max_retries: int = 3, executor=None, persistent: bool | None = None):
        """``keep_results=False`` drops each map once it has reached host memory (long benchmark
        runs, latency only); ``max_bins_per_launch`` launches long bands in chunks of bins;
        ``persistent`` as in ``GpuExecutor``."""
        self.executor = executor if executor is not None else GpuExecutor(form, phase_mode, num_streams, staging_bytes,
                                                                          persistent)
        self.budget = int(hbm_budget)
        self.max_batch = max_batch if batched else 1
        self.max_bins = max_bins_per_launch
        self.max_retries = int(max_retries)
        self.keep_results = keep_results
        self.form = form
        self.inflight: list[_Launch] = []
        self.queue: list[BeamformRequest] = []
        self.used = 0
        self.launches = 0
        self.failed_launches = 0
        self.batch_sizes: list[int] = []
        self.errors: list[str] = []

    # -- engine ------------------------------------------------------------------------------
    def submit(self, req: BeamformRequest) -> None:
        """Queue a request.  Work queued on the caller's current stream before this call (e.g. the
        kernels that computed a caller-provided CSM) is waited for before the request's launch."""
        req.ready = self.executor.ready_event()
        self.queue.append(req)

    def _requeue(self, reqs) -> None:
        # back to the head of the queue, in arrival order
        self.queue[:0] = sorted(reqs, key=lambda r: (r.arrival, r.rid))

    def _complete(self, r, t) -> None:
        r.t_done = t
        if r.csm_fn is not None:
            r.csm = None  # the request's own CSM is freed with it
        if r.grid_fn is not None:
            r.gpos = None  # regenerated grids are not kept after completion
        if not self.keep_results:
            r.result = None

    def retire(self, now_fn) -> int:
        """File the maps of finished launches; returns the number of requests completed."""
        done = 0
        keep = []
        for la in self.inflight:
            if not self.executor.poll(la.handle):
                keep.append(la)
                continue
            t = now_fn()
            maps = self.executor.fetch(la.handle)
            self.executor.release(la.handle)
            self.used -= la.bytes
            again = []
            for (r, b0, b1), m in zip(la.parts, maps):
                if self.keep_results:
                    if r.result is None:
                        r.result = np.empty((r.num_freqs, m.shape[1]), dtype=m.dtype)
                    r.result[b0:b1] = m
                r.done[b0:b1] = True
                r.attempts = 0
                if r.finished:
                    self._complete(r, t)
                    done += 1
                else:
                    again.append(r)
            self._requeue(again)
        self.inflight = keep
        return done

    def admit_and_launch(self, now_fn) -> int:
        """Launch the next batch if a slot is free; returns the number of requests completed by a
        failure (0 normally)."""
        slot = self.executor.free_slot()
        if slot is None or not self.queue:
            return 0
        batch = admit(self.queue, self.used, self.budget, self.max_batch, bool(self.inflight), self.max_bins)
        if not batch:
            return 0
        t = now_fn()
        parts = []
        for r in batch:
            b0, b1 = r.pending_bins(self.max_bins)
            if r.t_admit == 0.0:
                r.t_admit = t
            parts.append((r, b0, b1))
        nbytes = sum(r.footprint(b1 - b0) for r, b0, b1 in parts)
        try:
            handle = self.executor.launch(slot, parts)
        except (_lib.AbfError, RuntimeError) as e:  # torch's out-of-memory error is a RuntimeError
            self.failed_launches += 1
            self.errors.append(str(e))
            return self._failed(parts, str(e), now_fn)
        la = _Launch(parts, nbytes, handle)
        self.used += la.bytes
        self.inflight.append(la)
        self.launches += 1
        self.batch_sizes.append(len(parts))
        return 0

    def _failed(self, parts, msg, now_fn) -> int:
        """Requeue the requests of a failed launch with their finished bins kept; give up on a
        request after ``max_retries`` consecutive failures."""
        again, gave_up = [], 0
        for r, _, _ in parts:
            if r.csm_fn is not None:
                r.csm = None  # rebuilt at the retry; frees its memory meanwhile
            if r.grid_fn is not None:
                r.gpos = None
            r.attempts += 1
            if r.attempts > self.max_retries:
                r.error = msg
                self._complete(r, now_fn())
                gave_up += 1
            else:
                again.append(r)
        self._requeue(again)
        return gave_up

    def run(self, requests, realtime: bool = True, progress=None):
        """Serve ``requests`` (sorted by arrival).  With ``realtime`` the arrivals follow the
        wall clock from the start of the call; otherwise every request is available at once.
        Returns the wall time of the run."""
        t0 = time.perf_counter()

        def now():
            return time.perf_counter() - t0

        pending = list(requests)
        i = 0
        n = len(pending)
        finished = 0
        last_report = 0.0
        while finished < n:
            t = now()
            while i < n and (not realtime or pending[i].arrival <= t):
                self.submit(pending[i])
                i += 1
            finished += self.retire(now)
            before = self.launches
            finished += self.admit_and_launch(now)
            launched = self.launches > before
            if progress is not None and t - last_report > 10.0:
                progress(f'{t:.1f} s: {finished}/{n} done, {len(self.queue)} queued, {len(self.inflight)} in flight')
                last_report = t
            if not launched:
                if self.inflight:
                    # wait for the oldest launch or the next arrival, whichever is sooner
                    if realtime and i < n:
                        dt = pending[i].arrival - now()
                        if dt > 0:
                            time.sleep(min(dt, 2e-4))
                    else:
                        time.sleep(5e-5)
                elif realtime and i < n:
                    dt = pending[i].arrival - now()
                    if dt > 0:
                        time.sleep(min(dt, 1e-3))
        return now()


def latency_stats(reqs):
    lat = np.array([r.t_done - r.arrival for r in reqs])
    return {'p50_ms': float(np.percentile(lat, 50) * 1e3), 'p99_ms': float(np.percentile(lat, 99) * 1e3),
            'mean_ms': float(lat.mean() * 1e3), 'max_ms': float(lat.max() * 1e3)}


# ---------------------------------------------------------------------------------------------
# the mixed-size serving workload
# ---------------------------------------------------------------------------------------------
MIC_COUNTS = (32, 64, 128, 256, 512)


class RequestFactory:
    """Seeded mixed-size requests, each with its own analytic CSM.

    A request draws a microphone count (a seeded random array per count), F contiguous bins of
    block size 1024 at 51.2 kHz, 1..16 seeded sources at seeded positions and levels (a 20 dB
    spread), a 2D or 3D rectangular grid of log-uniform size and a steer type.  Its CSM
    ``sum_s p_s a_s a_s^H`` over its bins is built on the device when the request is admitted
    (``csm_fn``) and freed when it completes; its grid likewise (``grid_fn``).
    """

    def __init__(self, seed: int = 7, g_min: float = 1e3, g_max: float = 2.5e5, f_max: int = 64,
                 mic_counts=MIC_COUNTS, pool: int | None = None):
        """``pool`` is accepted for compatibility and ignored: every request builds its own CSM."""
        from . import synthetic

        torch = _lib.torch()
        self.dev = _lib.device()
        self.seed = seed
        self.freqs = np.arange(513) * 51200.0 / 1024
        self.k = torch.from_numpy(2 * np.pi * self.freqs / 343.0).to(self.dev)
        self.g_min, self.g_max, self.f_max = g_min, g_max, f_max
        self.mic_counts = tuple(mic_counts)
        self.mpos_host = {nm: synthetic.disk_array(nm, 1.0 + nm / 256.0, seed + 100 * j)
                          for j, nm in enumerate(self.mic_counts)}
        self.mpos = {nm: torch.from_numpy(mp).to(self.dev) for nm, mp in self.mpos_host.items()}
        torch.cuda.synchronize()

    @staticmethod
    def grid_shape(g_target: int, three_d: bool):
        if three_d:
            n = max(2, round(g_target ** (1 / 3)))
            return (n, n, max(1, round(g_target / (n * n))))
        n = max(1, round(g_target**0.5))
        return (n, n)

    def grid(self, g_target: int, three_d: bool):
        torch = _lib.torch()
        shape = self.grid_shape(g_target, three_d)
        if three_d:
            n, _, nz = shape
            x = torch.linspace(-0.6, 0.6, n, dtype=torch.float64, device=self.dev)
            z = torch.linspace(0.5, 1.0, nz, dtype=torch.float64, device=self.dev)
            gx, gy, gz = torch.meshgrid(x, x, z, indexing='ij')
        else:
            n = shape[0]
            x = torch.linspace(-0.6, 0.6, n, dtype=torch.float64, device=self.dev)
            gx, gy = torch.meshgrid(x, x, indexing='ij')
            gz = torch.full_like(gx, 0.75)
        return torch.stack([gx.reshape(-1), gy.reshape(-1), gz.reshape(-1)]).contiguous()

    def sources(self, rid_seed: int):
        """Source count, positions and powers of one request (seeded)."""
        from . import synthetic

        rng = np.random.default_rng(rid_seed)
        s = int(rng.integers(1, 17))
        spos = synthetic.source_positions(s, [-0.5, -0.5, 0.5], [0.5, 0.5, 1.0], rid_seed + 1)
        return spos, synthetic.level_powers(s, 20.0, rid_seed + 2)

    def csm(self, nm: int, b0: int, nf: int, rid_seed: int):
        """The request's analytic CSM ``[nf, nm, nm]`` complex64 on the device."""
        from . import synthetic

        spos, p = self.sources(rid_seed)
        return synthetic.analytic_csm(self.mpos_host[nm], spos, p, self.freqs[b0:b0 + nf], device=self.dev)

    def request(self, rid: int, arrival: float, nm: int, g: int, three_d: bool, b0: int, nf: int, steer: str,
                rs: int) -> BeamformRequest:
        """One request: ``nm`` mics, a grid of about ``g`` points, bins ``b0 .. b0 + nf - 1``, sources
        and levels seeded by ``rs``."""
        shape = self.grid_shape(g, three_d)
        return BeamformRequest(rid=rid, arrival=arrival, wavenumber=self.k[b0:b0 + nf], mpos=self.mpos[nm],
                               csm_fn=(lambda: self.csm(nm, b0, nf, rs)), grid_fn=(lambda: self.grid(g, three_d)),
                               steer_type=steer, r_diag=True, num_grid=int(np.prod(shape)))

    def make(self, n: int, qps: float | None, seed_offset: int = 0):
        rng = np.random.default_rng(12345 + seed_offset)
        t = 0.0
        reqs = []
        for rid in range(n):
            if qps:
                t += rng.exponential(1.0 / qps)
            nm = int(rng.choice(self.mic_counts))
            g = int(np.exp(rng.uniform(np.log(self.g_min), np.log(self.g_max))))
            three_d = bool(rng.integers(0, 2))
            nf = int(rng.integers(1, self.f_max + 1))
            b0 = int(rng.integers(1, 513 - nf + 1))
            steer = _STEER[int(rng.integers(0, 4))]
            rs = int(rng.integers(0, 2**31))
            reqs.append(self.request(rid, t, nm, g, three_d, b0, nf, steer, rs))
        return reqs
