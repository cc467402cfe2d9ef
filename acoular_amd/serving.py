"""Continuous-batching beamforming server on one GPU.

Requests (a resident CSM slice, a grid, microphone positions, a band of bins and the steering
parameters) arrive over time.  The engine loop

1. retires finished launches (HIP events) and copies their maps to pinned host memory,
   releasing their HBM footprint;
2. admits waiting requests in arrival order while the admitted HBM footprint stays under the
   budget and a launch slot (one of ``num_streams`` streams) is free;
3. launches all admitted requests as one ``abf_beamform_batch`` call (one MFMA launch per tile
   shape, work items of every request interleaved) on a free stream, so requests that arrive
   while earlier launches run overlap with them on the device.

Each map value is computed by one workgroup with a fixed reduction order, so a request's result
is bit-identical alone and in any mix.  The latency of a request is measured from its arrival to
the moment its map is in host memory.
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
    csm: object                    # device complex64 tensor [F, M, M] (resident)
    wavenumber: object             # device float64 tensor [F]
    gpos: object | None            # device float64 [3, G] (or None with grid_fn)
    mpos: object                   # device float64 [3, M]
    steer_type: str = 'true level'
    r_diag: bool = True
    ref: object = field(default_factory=lambda: np.zeros(3))
    grid_fn: object = None         # callable returning gpos on the device at admission
    num_grid: int = 0
    result: np.ndarray | None = None
    t_admit: float = 0.0
    t_done: float = 0.0

    @property
    def num_mics(self) -> int:
        return int(self.mpos.shape[1])

    @property
    def num_freqs(self) -> int:
        return int(self.csm.shape[0])

    @property
    def map_bins(self) -> int:
        return self.num_freqs * self.num_grid

    def footprint(self) -> int:
        """HBM bytes a request holds while in flight: map, grid, CSM re-layout workspace."""
        nm = self.num_mics
        mp = (nm + 31) // 32 * 32
        nk = mp // 32
        prep = self.num_freqs * nk * (nk + 1) // 2 * 8192
        return 4 * self.map_bins + 24 * self.num_grid + prep + 4096


def admit(queue: list, used: int, budget: int, max_batch: int, any_inflight: bool) -> list:
    """Pop the requests of the next launch from the head of ``queue`` (arrival order): as many
    as fit the HBM budget next to the ``used`` bytes of launches in flight, at most ``max_batch``.
    A request larger than the whole budget is admitted alone once nothing else is in flight, so
    it cannot block the queue forever."""
    batch = []
    nbytes = 0
    while queue and len(batch) < max_batch:
        fp = queue[0].footprint()
        if used + nbytes + fp > budget and (batch or any_inflight):
            break
        batch.append(queue.pop(0))
        nbytes += fp
    return batch


class _Launch:
    def __init__(self, reqs, stream, event, tensors, host_out):
        self.reqs = reqs
        self.stream = stream
        self.event = event
        self.tensors = tensors
        self.host_out = host_out
        self.bytes = sum(r.footprint() for r in reqs)


class BeamformServer:
    """Admission by HBM footprint + continuous batching over ``num_streams`` launch slots.

    ``batched=False`` with ``form='plain_f32'`` and ``num_streams=1`` gives the unbatched plain
    baseline: one request per launch, one launch at a time.
    """

    def __init__(self, hbm_budget: int = 32 << 30, max_batch: int = 64, num_streams: int = 2,
                 form: str = 'mfma_bf16x3', batched: bool = True, phase_mode: str = 'reference',
                 keep_results: bool = True, staging_bytes: int = 0):
        """``keep_results=False`` drops each map once it has reached host memory (long benchmark
        runs); ``staging_bytes`` > 0 gives every stream a reusable pinned host buffer of that size
        for the device-to-host copies instead of a pinned allocation per request."""
        torch = _lib.torch()
        self.dev = _lib.device()
        self.budget = int(hbm_budget)
        self.max_batch = max_batch if batched else 1
        # default-priority streams only
        self.streams = [torch.cuda.Stream(device=self.dev) for _ in range(num_streams)]
        self.keep_results = keep_results
        self.staging = {id(s): torch.empty(int(staging_bytes), dtype=torch.uint8, pin_memory=True)
                        for s in self.streams} if staging_bytes > 0 else {}
        self.form = form
        self.phase_mode = phase_mode
        self.inflight: list[_Launch] = []
        self.queue: list[BeamformRequest] = []
        self.used = 0
        self.launches = 0
        self.batch_sizes: list[int] = []

    # -- engine ------------------------------------------------------------------------------
    def submit(self, req: BeamformRequest) -> None:
        self.queue.append(req)

    def _free_stream(self):
        busy = {id(la.stream) for la in self.inflight}
        for s in self.streams:
            if id(s) not in busy:
                return s
        return None

    def retire(self, now_fn) -> int:
        done = 0
        keep = []
        for la in self.inflight:
            if la.event.query():
                t = now_fn()
                for r, h in zip(la.reqs, la.host_out):
                    if self.keep_results:
                        r.result = h.numpy().copy() if self.staging else h.numpy()
                    r.t_done = t
                    if r.grid_fn is not None:
                        r.gpos = None  # regenerated grids are not kept after completion
                self.used -= la.bytes
                done += len(la.reqs)
            else:
                keep.append(la)
        self.inflight = keep
        return done

    def admit_and_launch(self, now_fn) -> int:
        stream = self._free_stream()
        if stream is None or not self.queue:
            return 0
        batch = admit(self.queue, self.used, self.budget, self.max_batch, bool(self.inflight))
        if not batch:
            return 0
        self._launch(batch, stream, now_fn)
        return len(batch)

    def _launch(self, batch, stream, now_fn):
        torch = _lib.torch()
        t = now_fn()
        with torch.cuda.stream(stream):
            jobs, tensors, host_out = [], [], []
            for r in batch:
                r.t_admit = t
                if r.gpos is None:
                    r.gpos = r.grid_fn()
                r.num_grid = int(r.gpos.shape[1])
                out = torch.empty((r.num_freqs, r.num_grid), dtype=torch.float32, device=self.dev)
                nm = r.num_mics
                norm = nm / (nm - 1.0) if r.r_diag else 1.0
                jobs.append(make_job(r.csm, r.wavenumber, r.gpos, r.mpos, r.ref, r.steer_type, r.r_diag, norm,
                                     self.form, self.phase_mode, out, None, clip=r.r_diag))
                tensors += [out, r.gpos]
            arr = (_lib.AbfGridJob * len(jobs))(*jobs)
            lib = _lib.lib()
            nb = lib.abf_beamform_batch_workspace_size(arr, len(jobs), None)
            ws = torch.empty(max(nb, 1), dtype=torch.uint8, device=self.dev)
            _lib.check(lib.abf_beamform_batch(arr, len(jobs), _lib.ptr(ws), ws.numel(),
                                              _lib.ctypes.c_void_p(stream.cuda_stream)), 'serve batch')
            stage = self.staging.get(id(stream))
            off = 0
            for o in tensors[0::2]:
                nbytes = o.numel() * o.element_size()
                if stage is not None and off + nbytes <= stage.numel():
                    h = stage[off:off + nbytes].view(o.dtype).view(o.shape)
                    off += (nbytes + 255) // 256 * 256
                else:
                    h = torch.empty(o.shape, dtype=o.dtype, pin_memory=True)
                h.copy_(o, non_blocking=True)
                host_out.append(h)
            ev = torch.cuda.Event()
            ev.record(stream)
        tensors.append(ws)
        la = _Launch(batch, stream, ev, tensors, host_out)
        self.used += la.bytes
        self.inflight.append(la)
        self.launches += 1
        self.batch_sizes.append(len(batch))

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
            launched = self.admit_and_launch(now)
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
    """Seeded mixed-size requests over resident CSM pools.

    Per microphone count, ``pool`` analytic CSMs of all 513 bins (block size 1024 at 51.2 kHz)
    from 1..16 seeded sources stay resident on the device; a request takes ``F`` contiguous bins
    of one of them (a view, no copy), a 2D or 3D rectangular grid of log-uniform size and a steer
    type.
    """

    def __init__(self, seed: int = 7, pool: int = 2, g_min: float = 1e3, g_max: float = 2.5e5, f_max: int = 64,
                 mic_counts=MIC_COUNTS):
        from . import synthetic

        torch = _lib.torch()
        self.dev = _lib.device()
        self.rng = np.random.default_rng(seed)
        self.freqs = np.arange(513) * 51200.0 / 1024
        self.k = torch.from_numpy(2 * np.pi * self.freqs / 343.0).to(self.dev)
        self.g_min, self.g_max, self.f_max = g_min, g_max, f_max
        self.mic_counts = tuple(mic_counts)
        self.pools = {}
        self.mpos = {}
        for j, nm in enumerate(self.mic_counts):
            mp = synthetic.disk_array(nm, 1.0 + nm / 256.0, seed + 100 * j)
            self.mpos[nm] = torch.from_numpy(mp).to(self.dev)
            lst = []
            for p in range(pool):
                s = int(self.rng.integers(1, 17))
                spos = synthetic.source_positions(s, [-0.5, -0.5, 0.5], [0.5, 0.5, 1.0], seed + 1000 * j + p)
                lst.append(synthetic.analytic_csm(mp, spos, synthetic.level_powers(s, 20.0, seed + p), self.freqs,
                                                  device=self.dev))
            self.pools[nm] = lst
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
            pool = self.pools[nm][int(rng.integers(0, len(self.pools[nm])))]
            steer = _STEER[int(rng.integers(0, 4))]
            shape = self.grid_shape(g, three_d)
            reqs.append(BeamformRequest(rid=rid, arrival=t, csm=pool[b0 : b0 + nf], wavenumber=self.k[b0 : b0 + nf],
                                        gpos=None, grid_fn=(lambda g=g, d=three_d: self.grid(g, d)),
                                        mpos=self.mpos[nm], steer_type=steer, r_diag=True,
                                        num_grid=int(np.prod(shape))))
        return reqs
