"""Autotuner of the MFMA beamformer's launch shape.

The MFMA forms of ``abf_beamform`` take three launch parameters per job (``tile_kernel``,
``tile_cols``, ``tile_freqs`` of ``abf_grid_job``): the kernel variant (steering generation
alternating with the MFMA work, or steering waves beside MFMA waves), the grid points per
workgroup and the frequencies per workgroup.  The best choice depends on the microphone count
(LDS footprint and U-fragment reuse), on the grid size (enough workgroups to fill 256 CUs) and on
the number of bins (the frequency tile amortises the float64 distance prologue).

``tune_job`` times every distinct launch shape of one prepared job with HIP events on the launch
stream and returns the fastest; ``TileCache`` keeps the winners per shape class in a JSON file
(``$ABF_TUNE_CACHE`` or ``~/.cache/acoular_amd/tune.json``) so that later jobs of the same class
(``make_job`` consults the cache) start from the measured best without tuning again.  Classes
missing there fall back to the table shipped with the package (``tiles_gfx950.json``), which
``python -m acoular_amd.autotune --out acoular_amd/tiles_gfx950.json`` regenerates on an MI355X
for the standard workload classes (``STANDARD_CLASSES``).
"""

from __future__ import annotations

import ctypes
import json
import math
import os
import threading

from . import _lib

KERNELS = (1, 2, 3, 4, 5, 6, 7, 8)  # enum abf_mfma_kernel (include/abf.h)
SHIPPED_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'tiles_gfx950.json')
TILE_COLS = (16, 32, 64, 128)
TILE_FREQS = (2, 4, 8, 16, 32, 64)


def _bucket(n: int) -> int:
    """Nearest power of two (shape classes of grid size and bin count)."""
    return 1 << max(0, round(math.log2(max(1, n))))


def shape_key(form: str, num_mics: int, num_grid: int, num_freqs: int) -> str:
    """Shape class of a job: form, padded microphone count, grid and bin counts in powers of two."""
    mp = (int(num_mics) + 31) // 32 * 32
    return f'{form}|mp{mp}|g{_bucket(num_grid)}|f{min(_bucket(num_freqs), 64)}'


def default_cache_path() -> str:
    return os.environ.get('ABF_TUNE_CACHE') or os.path.join(os.path.expanduser('~'), '.cache', 'acoular_amd',
                                                            'tune.json')


class TileCache:
    """Measured best launch shapes per shape class, persisted as JSON."""

    def __init__(self, path: str | None = None, fallback: str | None = None):
        self.path = path or default_cache_path()
        self._lock = threading.Lock()
        self.entries = self._load(self.path)
        self.fallback = self._load(fallback) if fallback else {}

    @staticmethod
    def _load(path: str) -> dict[str, dict]:
        try:
            with open(path) as fh:
                data = json.load(fh)
        except (OSError, ValueError):
            return {}
        if not isinstance(data, dict):
            return {}
        return {str(k): dict(v) for k, v in data.items() if isinstance(v, dict)}

    def get(self, key: str) -> dict | None:
        e = self.entries.get(key)
        return e if e is not None else self.fallback.get(key)

    def put(self, key: str, entry: dict, save: bool = True) -> None:
        with self._lock:
            self.entries[key] = dict(entry)
            if save:
                self.save()

    def save(self) -> None:
        d = os.path.dirname(self.path)
        try:
            if d:
                os.makedirs(d, exist_ok=True)
            tmp = self.path + '.tmp'
            with open(tmp, 'w') as fh:
                json.dump(self.entries, fh, indent=1, sort_keys=True)
            os.replace(tmp, self.path)
        except OSError:
            pass  # a read-only home keeps the in-memory entries only


_cache: TileCache | None = None


def cache() -> TileCache:
    global _cache
    if _cache is None:
        _cache = TileCache(fallback=SHIPPED_PATH)
    return _cache


def set_cache(c: TileCache | None) -> None:
    global _cache
    _cache = c


def job_key(job: _lib.AbfGridJob) -> str:
    form = {v: k for k, v in _lib.FORMS.items()}[job.form]
    return shape_key(form, job.geom.num_mics, job.geom.num_grid, job.num_freqs)


def apply(job: _lib.AbfGridJob, tiles: TileCache | None = None) -> bool:
    """Fill the tile fields of ``job`` from the cache (only fields the caller left at 0)."""
    if job.form == _lib.FORMS['plain_f32']:
        return False
    e = (tiles or cache()).get(job_key(job))
    if not e:
        return False
    if job.tile_kernel == 0:
        job.tile_kernel = int(e.get('tile_kernel', 0))
    if job.tile_cols == 0:
        job.tile_cols = int(e.get('tile_cols', 0))
    if job.tile_freqs == 0:
        job.tile_freqs = int(e.get('tile_freqs', 0))
    return True


def plan(job: _lib.AbfGridJob) -> tuple[int, int, int, int]:
    """Effective (kernel, grid points, frequencies, workgroups) of a job after clamping."""
    k, c, f = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    w = ctypes.c_int64()
    _lib.check(_lib.lib().abf_beamform_plan(job, ctypes.byref(k), ctypes.byref(c), ctypes.byref(f),
                                            ctypes.byref(w)), 'plan')
    return k.value, c.value, f.value, w.value


def _time_run(job, ws, reps: int) -> float:
    torch = _lib.torch()
    lib = _lib.lib()
    stream = _lib.stream_handle()
    _lib.check(lib.abf_beamform_run(job, _lib.ptr(ws), ws.numel(), stream), 'tune warm-up')
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(reps):
        _lib.check(lib.abf_beamform_run(job, _lib.ptr(ws), ws.numel(), stream), 'tune run')
    stop.record()
    stop.synchronize()
    return start.elapsed_time(stop) / 1e3 / reps


def tune_job(job: _lib.AbfGridJob, reps: int = 3, kernels=KERNELS, tile_cols=TILE_COLS, tile_freqs=TILE_FREQS,
             tiles: TileCache | None = None, store: bool = True) -> dict:
    """Time every distinct launch shape of ``job`` (its tensors must be alive) and return the
    fastest as ``{'tile_kernel', 'tile_cols', 'tile_freqs', 'seconds', 'tried'}``.  The job's
    own tile fields are restored; the winner is stored in the cache unless ``store`` is false."""
    if job.form == _lib.FORMS['plain_f32']:
        raise ValueError('the plain form has no launch shape to tune')
    lib = _lib.lib()
    saved = (job.tile_kernel, job.tile_cols, job.tile_freqs)
    ws = _lib.workspace(lib.abf_beamform_workspace_size(job))
    _lib.check(lib.abf_beamform_prepare(job, _lib.ptr(ws), ws.numel(), _lib.stream_handle()), 'tune prepare')
    seen = {}
    try:
        for k in kernels:
            for c in tile_cols:
                for f in tile_freqs:
                    job.tile_kernel, job.tile_cols, job.tile_freqs = k, c, f
                    try:
                        eff = plan(job)
                    except (ValueError, NotImplementedError):
                        continue
                    if eff[:3] in seen:
                        continue
                    seen[eff[:3]] = _time_run(job, ws, reps)
    finally:
        job.tile_kernel, job.tile_cols, job.tile_freqs = saved
    if not seen:
        raise NotImplementedError('no MFMA launch shape fits this job')
    best = min(seen, key=seen.get)
    result = {'tile_kernel': best[0], 'tile_cols': best[1], 'tile_freqs': best[2], 'seconds': seen[best],
              'tried': [{'tile_kernel': s[0], 'tile_cols': s[1], 'tile_freqs': s[2], 'seconds': t}
                        for s, t in sorted(seen.items(), key=lambda kv: kv[1])]}
    if store:
        (tiles or cache()).put(job_key(job), {k: result[k] for k in ('tile_kernel', 'tile_cols', 'tile_freqs',
                                                                     'seconds')})
    return result


# (form, mics, grid points, bins) of the classes the shipped table covers: the 256-mic 201 x 201
# plain-vs-MFMA map (all three MFMA forms), the fused 512-mic 128 x 128 map of 1025 bins, the 1024-mic 64 x 64 x 32 map
# and the serving mix (mic counts 32..512, small and large grids, narrow and wide bands)
STANDARD_CLASSES = (
    [('mfma_bf16x3', 256, 40401, 513), ('mfma_f32', 256, 40401, 513), ('mfma_f16x2', 256, 40401, 513),
     ('mfma_bf16x3', 512, 16384, 1025),
     ('mfma_bf16x3', 1024, 131072, 256)]
    + [('mfma_bf16x3', m, g, f) for m in (32, 64, 128, 256, 512) for g in (4096, 65536) for f in (4, 32)]
)


def tune_class(form: str, num_mics: int, num_grid: int, num_freqs: int, reps: int = 2, tiles=None) -> dict:
    """Tune one shape class on a synthetic job of that shape (seeded random Hermitian CSM, grid
    points in a 2 m box at z = 1, microphones in a 2 m disk; timing does not depend on values)."""
    import numpy as np

    from .fbeamform import make_job

    torch = _lib.torch()
    dev = _lib.device()
    rng = np.random.default_rng(num_mics * 7 + num_grid + num_freqs)
    a = rng.standard_normal((num_freqs, num_mics, 4)) + 1j * rng.standard_normal((num_freqs, num_mics, 4))
    csm = torch.from_numpy((a @ np.conj(a.transpose(0, 2, 1))).astype(np.complex64)).to(dev)
    gpos = torch.from_numpy(np.stack([rng.uniform(-1, 1, num_grid), rng.uniform(-1, 1, num_grid),
                                      np.ones(num_grid)])).to(dev)
    r = np.sqrt(rng.uniform(0, 1, num_mics))
    ph = rng.uniform(0, 2 * np.pi, num_mics)
    mpos = torch.from_numpy(np.stack([r * np.cos(ph), r * np.sin(ph), np.zeros(num_mics)])).to(dev)
    k = torch.from_numpy(2 * np.pi * np.linspace(500.0, 20000.0, num_freqs) / 343.0).to(dev)
    out = torch.empty((num_freqs, num_grid), dtype=torch.float32, device=dev)
    job = make_job(csm, k, gpos, mpos, np.zeros(3), 'true level', True, 1.0, form, 'reference', out, None, True)
    job.tile_kernel = job.tile_cols = job.tile_freqs = 0
    return tune_job(job, reps=reps, tiles=tiles)


def main(argv=None):
    import argparse

    ap = argparse.ArgumentParser(description='Tune the MFMA launch shapes of the standard workload classes.')
    ap.add_argument('--out', default=None, help='JSON file to write (default: the user cache)')
    ap.add_argument('--reps', type=int, default=2)
    args = ap.parse_args(argv)
    tiles = TileCache(args.out) if args.out else cache()
    for form, m, g, f in STANDARD_CLASSES:
        res = tune_class(form, m, g, f, reps=args.reps, tiles=tiles)
        print(json.dumps({'class': shape_key(form, m, g, f),
                          **{k: res[k] for k in ('tile_kernel', 'tile_cols', 'tile_freqs', 'seconds')},
                          'tried': len(res['tried'])}), flush=True)
    tiles.save()


if __name__ == '__main__':
    main()
