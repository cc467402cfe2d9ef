#!/usr/bin/env python3
"""Benchmark of the frequency-domain beamforming hot path on one MI355X.

Workload (the plain-vs-MFMA configuration): 256 seeded random microphones in a 2 m disk,
8 seeded monopoles with a 20 dB level spread, analytic complex64 CSM of 513 bins
(block_size 1024 at fs 51200 Hz), RectGrid 201 x 201 (40,401 points) at z = 0.5,
steer_type 'classic', r_diag True.  One step = the whole path over the resident CSM:
re-layout of the CSM (abf_beamform_prepare) + beamforming of all 513 bins x 40,401 points
(abf_beamform_run).  Metric: map-bins/s.

Printed: one JSON line (stdout) with the throughput, the roofline of the dominant kernel
(bf_mfma_kernel) timed with HIP events on the launch stream, the other forms' throughput, the
max relative map error against the CPU oracle on sampled bins and the CPU baseline
(oracle/cpu_ref.c timed on the host cores on a bounded sample of the same workload).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MI355X_PEAK = {'bf16': 2500.0, 'fp16': 2500.0, 'fp32_mfma': 157.3, 'fp32': 157.3}  # TFLOP/s, dense
# FP32 VALU without packed instructions (disabled in this build): 256 CUs x 4 SIMDs x 16 lanes x
# 2 flops x 2.4 GHz
FP32_VALU_NO_PACKED = 78.6


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_workload(dev, seed=2024):
    import torch

    from acoular_amd import synthetic

    nm, fs, bs = 256, 51200.0, 1024
    mpos = synthetic.disk_array(nm, 2.0, seed)
    spos = synthetic.source_positions(8, [-0.8, -0.8, 0.5], [0.8, 0.8, 0.5], seed + 1)
    powers = synthetic.level_powers(8, 20.0, seed + 2)
    freqs = np.arange(bs // 2 + 1) * fs / bs
    x = np.linspace(-1.0, 1.0, 201)
    gpos = np.empty((3, 201 * 201))
    gpos[0] = np.repeat(x, 201)
    gpos[1] = np.tile(x, 201)
    gpos[2] = 0.5
    csm = synthetic.analytic_csm(mpos, spos, powers, freqs, device=dev)
    torch.cuda.synchronize()
    return dict(mpos=mpos, spos=spos, powers=powers, freqs=freqs, gpos=gpos, csm=csm, nm=nm)


class Runner:
    """One prepared job of a given form over the resident workload."""

    def __init__(self, wl, form, steer='classic', r_diag=True, dev=None, tuned=True):
        import torch

        from acoular_amd import _lib
        from acoular_amd.fbeamform import config, make_job

        self.lib = _lib.lib()
        self._lib = _lib
        nm = wl['nm']
        self.gpos = torch.from_numpy(wl['gpos']).to(dev)
        self.mpos = torch.from_numpy(wl['mpos']).to(dev)
        self.k = torch.from_numpy(2 * np.pi * wl['freqs'] / 343.0).to(dev)
        self.out = torch.empty((len(wl['freqs']), wl['gpos'].shape[1]), dtype=torch.float32, device=dev)
        norm = nm / (nm - 1.0) if r_diag else 1.0
        saved, config['autotuned'] = config['autotuned'], tuned
        try:
            self.job = make_job(wl['csm'], self.k, self.gpos, self.mpos, np.zeros(3), steer, r_diag, norm, form,
                                'reference', self.out, None, clip=r_diag)
        finally:
            config['autotuned'] = saved
        self.ws = _lib.workspace(self.lib.abf_beamform_workspace_size(self.job))
        self.csm = wl['csm']
        self.map_bins = self.out.numel()

    def prepare(self):
        s = self._lib.stream_handle()
        self._lib.check(self.lib.abf_beamform_prepare(self.job, self._lib.ptr(self.ws), self.ws.numel(), s))

    def run(self):
        s = self._lib.stream_handle()
        self._lib.check(self.lib.abf_beamform_run(self.job, self._lib.ptr(self.ws), self.ws.numel(), s))

    def step(self):
        self.prepare()
        self.run()


def measured_traffic(form, runner):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3 FETCH_SIZE and
    WRITE_SIZE runs (profiles/hbm_traffic.json, calibrated against a 1 GiB copy) when they were
    taken with this form and launch shape; None otherwise."""
    from acoular_amd import autotune

    try:
        with open(os.path.join(ROOT, 'profiles', 'hbm_traffic.json')) as fh:
            table = json.load(fh)
    except (OSError, ValueError):
        return None
    if form == 'plain_f32':
        key = 'bf_plain_kernel|plain_f32|256mic_201x201_513bins|0,64,1'
    else:
        k, c, f, _ = autotune.plan(runner.job)
        key = f'bf_mfma_kernel|{form}|256mic_201x201_513bins|{k},{c},{f}'
    e = table.get(key)
    return None if e is None else int(e['bytes_per_launch'])


def dominant_kernel_name(form, runner):
    """Name of the beamform launch's kernel (as rocprofv3 lists it) for the job's launch shape."""
    if form == 'plain_f32':
        return 'bf_plain_kernel'
    from acoular_amd import autotune

    k = autotune.plan(runner.job)[0]
    return {6: 'bf_mfma_stream_kernel', 7: 'bf_mfma_wide_kernel',
            8: 'bf_mfma_stream_kernel (persistent)'}.get(k, 'bf_mfma_kernel')


def executed_flops_per_bin(form, nm):
    """Flops the kernels execute per map-bin (the upper block triangle of U only)."""
    if form == 'plain_f32':
        mp = (nm + 15) // 16 * 16
        nrb = mp // 16
        return 8.0 * 16 * (mp * nrb - 16 * nrb * (nrb - 1) / 2)
    nk = (nm + 31) // 32
    # per (row tile, K-block): 12 or 8 16x16x32 MFMAs (split forms), 32 16x16x4 fp32 MFMAs
    per_pair = {'mfma_bf16x3': 12 * 16384 / 16, 'mfma_f16x2': 8 * 16384 / 16}.get(form, 32 * 2048 / 16)
    return per_pair * nk * (nk + 1)


def api_bench(args):
    """The drop-in array functions at their workloads' sizes, all HBM-bound: calcCSM
    (fastFuncs.py:19-58) as the reference calls it once per block, on a 1025-bin, 512-mic
    complex64 CSM (the fused workload's shape: every upper-triangle element read and written
    once per call); calcTransfer (fastFuncs.py:859-897, steer_vector normalisation 'true level')
    of 40,401 points x 256 mics into complex128 (distances read once, transfer written once);
    dist_mat (environments.py:65-103) of a 64 x 64 x 32 grid to 1024 mics (float64 written
    once).  Each against oracle values on a sample; the CPU leg is the oracle's calcCSM."""
    import torch

    from acoular_amd import _lib
    from oracle import beamform as B
    from oracle import spectra as S

    dev = torch.device('cuda', 0)
    lib = _lib.lib()
    sh = _lib.stream_handle
    g = torch.Generator(device=dev).manual_seed(11)
    out = {}
    # calcCSM
    F, M = 1025, 512
    csm = torch.randn((F, M, M, 2), generator=g, device=dev, dtype=torch.float32)
    csm = torch.view_as_complex(csm).contiguous()
    spec = torch.view_as_complex(torch.randn((F, M, 2), generator=g, device=dev, dtype=torch.float32)).contiguous()
    csm0 = csm[[3, 700]].cpu().numpy()

    def step_csm():
        _lib.check(lib.abf_calc_csm(_lib.ptr(csm), _lib.ptr(spec), F, M, 0, sh()), 'calcCSM')

    step_csm()
    want = S.calc_csm_update(csm0.astype(np.complex128), spec[[3, 700]].cpu().numpy().astype(np.complex128))
    got = csm[[3, 700]].cpu().numpy()
    iu = np.triu_indices(M)
    err_csm = float(np.abs(got[:, iu[0], iu[1]] - want[:, iu[0], iu[1]]).max() / np.abs(want).max())
    t_csm = event_time(step_csm, 20, warm=3)
    b_csm = F * (M * (M + 1) // 2) * 16.0 + F * M * 8.0
    log(f'calcCSM {F} x {M} x {M} complex64: {t_csm * 1e3:.3f} ms per call, {b_csm / t_csm / 1e9:.0f} GB/s, '
        f'max error {err_csm:.1e}')
    # calcTransfer + steer_vector 'true level'
    G, Mt = 40401, 256
    rm = torch.rand((G, Mt), generator=g, device=dev, dtype=torch.float64) + 0.5
    r0 = torch.rand((G,), generator=g, device=dev, dtype=torch.float64) + 0.5
    tr = torch.empty((G, Mt), dtype=torch.complex128, device=dev)
    k = 2 * np.pi * 8000.0 / 343.0

    def step_tr():
        _lib.check(lib.abf_transfer(_lib.ptr(r0), _lib.ptr(rm), G, Mt, k, 0, 3, 1, _lib.ptr(tr), sh()), 'transfer')

    step_tr()
    sel = np.arange(0, G, 997)
    a = B.calc_transfer(r0.cpu().numpy()[sel], rm.cpu().numpy()[sel], k)
    want_tr = B.steer_vector(a, 'true level')
    err_tr = float(np.abs(tr.cpu().numpy()[sel] - want_tr).max() / np.abs(want_tr).max())
    t_tr = event_time(step_tr, 20, warm=3)
    b_tr = G * 8.0 + G * Mt * 8.0 + G * Mt * 16.0
    log(f'steer_vector {G} x {Mt} complex128: {t_tr * 1e3:.3f} ms, {b_tr / t_tr / 1e9:.0f} GB/s, max error {err_tr:.1e}')
    # dist_mat
    Gd, Md = 131072, 1024
    gpos = torch.rand((3, Gd), generator=g, device=dev, dtype=torch.float64)
    mpos = torch.rand((3, Md), generator=g, device=dev, dtype=torch.float64)
    rmd = torch.empty((Gd, Md), dtype=torch.float64, device=dev)

    def step_dm():
        _lib.check(lib.abf_dist_mat(_lib.ptr(gpos), _lib.ptr(mpos), Gd, Md, _lib.ptr(rmd), sh()), 'dist_mat')

    step_dm()
    seld = np.arange(0, Gd, 4099)
    want_dm = B.dist_mat(gpos.cpu().numpy()[:, seld], mpos.cpu().numpy())
    err_dm = float(np.abs(rmd.cpu().numpy()[seld] - want_dm).max() / want_dm.max())
    t_dm = event_time(step_dm, 20, warm=3)
    b_dm = Gd * Md * 8.0 + (Gd + Md) * 24.0
    log(f'dist_mat {Gd} x {Md}: {t_dm * 1e3:.3f} ms, {b_dm / t_dm / 1e9:.0f} GB/s, max error {err_dm:.1e}')
    roof = {'bound': 'hbm', 'achieved': round(b_csm / t_csm / 1e9, 1), 'peak': 8000.0, 'unit': 'GB/s',
            'kernel': 'calc_csm_c64_kernel', 'algorithmic_bytes_per_launch': b_csm, 'traffic': None}
    roof['frac'] = round(roof['achieved'] / roof['peak'], 4)
    cpu = None
    if not args.no_cpu:
        c = csm[:64].cpu().numpy().astype(np.complex128)
        sp = spec[:64].cpu().numpy().astype(np.complex128)
        done, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < min(args.cpu_seconds, 10.0):
            S.calc_csm_update(c, sp)
            done += 1
        dt = time.perf_counter() - t0
        cpu = {'value': done * 64 * (M * (M + 1) // 2) * 16.0 / dt / 1e9, 'unit': 'GB/s', 'cores': 1, 'kind': 'port',
               'sample': f'{done} calcCSM updates of 64 of the 1025 bins (oracle, numpy), {dt:.1f} s'}
        log(f"CPU calcCSM: {cpu['value']:.2f} GB/s")
    return {
        'metric': 'GB/s of the drop-in calcCSM update (1025 bins x 512 mics, complex64)',
        'value': b_csm / t_csm / 1e9, 'unit': 'GB/s', 'n_gpus': 1, 'ms_per_step': t_csm * 1e3,
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32 (complex64)',
        'data': 'synthetic (seeded normal values)',
        'config': {'workload': 'drop-in array functions: calcCSM 1025 x 512 x 512 c64, steer_vector 40401 x 256 '
                               'c128, dist_mat 131072 x 1024 f64', 'parallelism': 'single GPU'},
        'roofline': roof, 'max_rel_err': err_csm, 'cpu_baseline': cpu,
        'steer_vector': {'ms': t_tr * 1e3, 'GB/s': b_tr / t_tr / 1e9, 'frac': b_tr / t_tr / 8e12, 'max_rel_err': err_tr},
        'dist_mat': {'ms': t_dm * 1e3, 'GB/s': b_dm / t_dm / 1e9, 'frac': b_dm / t_dm / 8e12, 'max_rel_err': err_dm},
    }


def time_steps(fn, steps, warmup):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def event_time(fn, reps, warm=0):
    """Average device time of ``fn`` from HIP events on the current (launch) stream, after
    ``warm`` untimed calls (the GPU idles while the host checks results before a timing)."""
    import torch

    st = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    for _ in range(warm):
        fn()
    torch.cuda.synchronize()
    e0.record(st)
    for _ in range(reps):
        fn()
    e1.record(st)
    e1.synchronize()
    return e0.elapsed_time(e1) / 1e3 / reps


def parity_sample(wl, runner, bins):
    """Errors of the computed maps against the oracle on sampled bins: the peak-normalised map error
    max_f max_g |B - B_ref| / max_g |B_ref| against the reference phase (float32-rounded k r) and
    against float64 phases, and the largest per-point relative error |B - B_ref| / |B_ref| over the
    points with B_ref >= 1e-2 of the peak."""
    from oracle import beamform as B

    out = runner.out.cpu().numpy()
    csm = wl['csm'][bins].cpu().numpy().astype(np.complex128)
    res = {'max_rel_err': 0.0, 'max_rel_err_vs_fp64_phase': 0.0, 'point_rel_err_above_1pct_peak': 0.0}
    for key, pm in (('max_rel_err', 'reference'), ('max_rel_err_vs_fp64_phase', 'fp64')):
        ref = B.beamformer_map(csm, wl['freqs'][bins], wl['gpos'], wl['mpos'], 'classic', True,
                               ind=np.arange(len(bins)), phase_mode=pm)
        for i, b in enumerate(bins):
            peak = np.abs(ref[i]).max()
            if peak > 0:
                d = np.abs(out[b] - ref[i])
                res[key] = max(res[key], float(d.max() / peak))
                if pm == 'reference':
                    sel = ref[i] >= 1e-2 * peak
                    res['point_rel_err_above_1pct_peak'] = max(res['point_rel_err_above_1pct_peak'],
                                                               float((d[sel] / ref[i][sel]).max()))
    return res


def variants_bench(wl, dev, steps=2):
    """Every form on the 256-mic workload for 'classic' and 'inverse' with and without r_diag:
    map-bins/s (mean of ``steps`` steps after one warm-up, prepare + beamform) and the max relative
    map error against the oracle on two bins."""
    import torch

    from oracle import beamform as B

    bins = [100, 400]
    out = []
    for steer in ('classic', 'inverse'):
        for rd in (True, False):
            ref = B.beamformer_map(wl['csm'][bins].cpu().numpy().astype(np.complex128), wl['freqs'][bins], wl['gpos'],
                                   wl['mpos'], steer, rd, ind=np.arange(len(bins)))
            peak = np.abs(ref).max(axis=1)
            for form in ('mfma_bf16x3', 'mfma_f32', 'plain_f32', 'mfma_f16x2'):
                r = Runner(wl, form, steer=steer, r_diag=rd, dev=dev)
                t = time_steps(r.step, steps, 1)
                got = r.out[bins].double().cpu().numpy()
                err = float(np.max(np.max(np.abs(got - ref), axis=1) / peak))
                out.append({'form': form, 'steer_type': steer, 'r_diag': rd, 'map_bins_per_s': r.map_bins / t,
                            'ms_per_step': t * 1e3, 'max_rel_err': err, 'err_bins': bins,
                            'gate': 5e-4 if form == 'mfma_f16x2' else 1e-4})
                log(f'variant {form:12s} {steer:8s} r_diag={rd!s:5s}: {r.map_bins / t:.4g} map-bins/s, err {err:.2e}')
                del r
                torch.cuda.empty_cache()
    return out


def _cpu_model():
    try:
        with open('/proc/cpuinfo') as fh:
            for line in fh:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(wl, budget_s):
    """oracle/cpu_ref.c (the reference loop nest in C + OpenMP) on the host cores, on a bounded
    sample of the same workload: whole 40,401-point maps, as many bins as fit ~budget_s with all
    threads, then one map on one thread."""
    from oracle import beamform as B
    from oracle import cref

    nm = wl['nm']
    r0 = B.steer_r0(wl['gpos'], np.zeros(3))
    rm = cref.dist_mat(wl['gpos'], wl['mpos'])
    bins = [(7 + 37 * i) % 513 for i in range(513)]  # every bin once, spread over the band
    threads = cref.num_threads()

    def one(b):
        k = 2 * np.pi * wl['freqs'][b] / 343.0
        csm = wl['csm'][b].cpu().numpy().astype(np.complex128)
        cref.beamformer_full_csm(csm, r0, rm, k, nm / (nm - 1.0), True, 1)

    one(bins[0])  # warm-up
    done, t0 = 0, time.perf_counter()
    for b in bins:
        one(b)
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    cref.set_threads(1)
    t1 = time.perf_counter()
    one(bins[0])
    dt1 = time.perf_counter() - t1
    cref.set_threads(threads)
    g = wl['gpos'].shape[1]
    return {'value': g * done / dt, 'unit': 'map-bins/s', 'cores': threads, 'kind': 'port',
            'value_1_thread': g / dt1, 'cpu_model': _cpu_model(), 'nproc': len(os.sched_getaffinity(0)),
            'sample': f'{done} full 201x201 maps (bins 7, 44, 81, ... in steps of 37 mod 513) of the same 256-mic '
                      f'classic r_diag workload, {dt:.1f} s on {threads} threads after one warm-up map; one map on '
                      f'1 thread {dt1:.1f} s'}


def _serve(server_kw, reqs, realtime):
    from acoular_amd import serving

    srv = serving.BeamformServer(**server_kw)
    t = srv.run(reqs, realtime=realtime, progress=log)
    st = serving.latency_stats(reqs)
    bins = sum(r.map_bins for r in reqs)
    st.update({'requests': len(reqs), 'wall_s': t, 'req_per_s': len(reqs) / t, 'map_bins_per_s': bins / t,
               'launches': srv.launches, 'mean_batch': float(np.mean(srv.batch_sizes)) if srv.batch_sizes else 0.0})
    return st


def serving_bench(args):
    """Mixed-size serving: 10,000 requests (mic count in {32..512}, grid 1k..250k points 2D or 3D,
    1..64 contiguous bins, analytic CSMs of 1..16 sources), Poisson arrivals at a fixed fraction of
    the measured saturation rate of the batched MFMA engine; latency = arrival -> map in host
    memory.  Baseline: the unbatched plain kernel (one request per launch, one launch at a time)
    on the same request stream."""
    from acoular_amd import serving

    log('serving: resident CSM pools (analytic, 513 bins, 32..512 mics)')
    fac = serving.RequestFactory(seed=7)
    mfma = dict(hbm_budget=24 << 30, max_batch=64, num_streams=2, keep_results=False, staging_bytes=1 << 30)
    plain = dict(hbm_budget=24 << 30, max_batch=1, num_streams=1, form='plain_f32', batched=False,
                 keep_results=False, staging_bytes=1 << 30)
    _serve(mfma, fac.make(40, None, seed_offset=9), False)  # warm-up: first launches of every shape
    _serve(plain, fac.make(4, None, seed_offset=9), False)
    sat = _serve(mfma, fac.make(args.serve_sat, None, seed_offset=1), False)
    log(f"saturation (all {args.serve_sat} requests queued at once): {sat['req_per_s']:.1f} req/s, "
        f"{sat['map_bins_per_s']:.3g} map-bins/s")
    qps = args.serve_load * sat['req_per_s']
    log(f'timed run: {args.serve_requests} requests, Poisson at {qps:.1f} req/s')
    reqs = fac.make(args.serve_requests, qps, seed_offset=2)
    run = _serve(mfma, reqs, True)
    log(f"p50 {run['p50_ms']:.1f} ms, p99 {run['p99_ms']:.1f} ms, {run['map_bins_per_s']:.3g} map-bins/s")
    # the same request stream with every streamed job as a persistent launch
    prs = _serve({**mfma, 'persistent': True}, fac.make(args.serve_requests, qps, seed_offset=2), True)
    log(f"persistent launches: p50 {prs['p50_ms']:.1f} ms, p99 {prs['p99_ms']:.1f} ms, "
        f"{prs['map_bins_per_s']:.3g} map-bins/s")
    log('baseline: unbatched plain kernel')
    psat = _serve(plain, fac.make(args.serve_plain, None, seed_offset=1), False)
    preqs = fac.make(args.serve_plain, qps, seed_offset=2)  # the first requests of the same stream
    prun = _serve(plain, preqs, True)
    log(f"plain: saturation {psat['req_per_s']:.1f} req/s; at {qps:.1f} req/s offered p50 {prun['p50_ms']:.1f} "
        f"ms, p99 {prun['p99_ms']:.1f} ms")
    return {
        'metric': 'serving p99 latency (mixed-size beamform requests, continuous batching)',
        'value': run['p99_ms'], 'unit': 'ms', 'n_gpus': 1, 'higher_is_better': False, 'vs_baseline': None,
        'dtype': 'bf16x3 (split hi+lo, fp32 accumulate)', 'data': 'synthetic (seeded analytic CSMs)',
        'config': {'workload': 'serving loop: mic count in {32,64,128,256,512}, grid log-uniform 1k..250k points '
                               '(2D/3D), 1..64 contiguous bins, 1..16 sources, r_diag, all four steer types',
                   'requests': args.serve_requests, 'offered_req_per_s': qps, 'load_fraction': args.serve_load,
                   'hbm_budget_bytes': mfma['hbm_budget'], 'max_batch': mfma['max_batch'],
                   'streams': mfma['num_streams'], 'parallelism': 'single GPU'},
        'serving': run, 'serving_persistent': prs, 'saturation': sat,
        'plain_unbatched': {'saturation': psat, 'same_offered_load_first_requests': prun},
    }


def _fused_oracle_bins(data, fs, bs, gpos, mpos, bins):
    """Oracle maps of a few bins of the fused workload: the Welch CSM of those bins by a direct
    float64 DFT of every block (``spectra.py:366-416``), then ``beamformer_map``."""
    from oracle import beamform as B
    from oracle import spectra as S

    n, m = data.shape
    w = np.hanning(bs)
    starts = S.block_starts(n, bs, '50%')
    e = np.exp(-2j * np.pi * np.outer(bins, np.arange(bs)) / bs)  # [nb, bs]
    csm = np.zeros((len(bins), m, m), dtype=np.complex128)
    for st in starts:
        x = e @ (data[st:st + bs].astype(np.float64) * w[:, None])  # [nbins, M]
        csm += x[:, :, None] * np.conj(x[:, None, :])
    csm *= 2.0 / bs / float(np.dot(w, w)) / S.num_blocks(n, bs, '50%')
    freqs = np.asarray(bins, dtype=float) * fs / bs
    return B.beamformer_map(csm, freqs, gpos, mpos, 'true location', True, ind=np.arange(len(bins)))


def fused_bench(args):
    """CSM estimation plus beamforming fused and batched: 512 mics (2 m disk), 10 s at 51.2 kHz
    float32 from 16 seeded correlated white-noise sources (mixing I + 0.3 N(0,1)) at z = 1.0 plus
    5 % uncorrelated noise; Welch bs 2048, Hanning, 50 % (499 blocks, 1025 bins, rank-1 updates on
    fp32 MFMA); RectGrid 128 x 128 on [-1, 1]^2 at z = 1.0, true location, r_diag; the whole
    pipeline captured in one hipGraph and replayed.  One step = one replay."""
    import torch

    from acoular_amd import autotune, synthetic
    from acoular_amd.pipeline import FusedCsmBeamform

    fs, n, nm, bs, seed = 51200.0, 512000, 512, 2048, 3003
    mpos = synthetic.disk_array(nm, 2.0, seed)
    spos = synthetic.source_positions(16, [-0.8, -0.8, 1.0], [0.8, 0.8, 1.0], seed + 1)
    mixing = np.eye(16) + 0.3 * np.random.default_rng(seed + 2).standard_normal((16, 16))
    log('fused: synthesising 10 s x 512 channels on the GPU')
    t = synthetic.point_source_signals_device(mpos, spos, n, fs, seed=seed + 3, noise_rel=0.05, mixing=mixing)
    x = np.linspace(-1.0, 1.0, 128)
    gpos = np.stack([np.repeat(x, 128), np.tile(x, 128), np.full(128 * 128, 1.0)])
    p = FusedCsmBeamform(t, fs, gpos, mpos, block_size=bs, steer_type='true location', r_diag=True, ind_low=0,
                         ind_high=bs // 2 + 1)
    p.run_csm()
    tiles = None
    if args.no_tune:
        p.job.tile_kernel = p.job.tile_cols = p.job.tile_freqs = 0
    else:
        src = 'autotuner table'
        if args.tune or p.job.tile_cols == 0:
            res = autotune.tune_job(p.job, reps=1)
            p.job.tile_kernel, p.job.tile_cols, p.job.tile_freqs = (res['tile_kernel'], res['tile_cols'],
                                                                    res['tile_freqs'])
            src = 'tuned now'
        tiles = {'tile_kernel': p.job.tile_kernel, 'tile_cols': p.job.tile_cols, 'tile_freqs': p.job.tile_freqs,
                 'source': src}
        log(f'launch shape: {tiles}')
    t_csm = event_time(p.run_csm, 2)
    t_bf = event_time(p.run_beamform, 2)
    t_eager = time_steps(p.run, 2, 1)
    p.capture()
    t_graph = time_steps(p.replay, args.steps, args.warmup)
    log(f'CSM {t_csm * 1e3:.1f} ms, beamform {t_bf * 1e3:.1f} ms, eager pipeline {t_eager * 1e3:.1f} ms, '
        f'graph replay {t_graph * 1e3:.1f} ms')
    value = p.map_bins / t_graph
    csm_flops = 8.0 * nm * nm * (bs // 2 + 1) * p.num_blocks
    bins = [200, 700]
    err = None
    if not args.no_compare:
        log(f'oracle on bins {bins}')
        ref = _fused_oracle_bins(t.cpu().numpy(), fs, bs, gpos, mpos, bins)
        got = p.out[bins].double().cpu().numpy()
        peak = np.max(np.abs(ref), axis=1)
        err = float(np.max(np.max(np.abs(got - ref), axis=1) / peak))
        log(f'max relative map error {err:.2e}')
    return {
        'metric': 'map-bins/s (grid pts x freqs) of the fused CSM + beamforming pipeline', 'value': value,
        'unit': 'map-bins/s', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': t_graph * 1e3, 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f32 (CSM, fp32 MFMA) + bf16x3 (beamforming)', 'data': 'synthetic (seeded, generated on the GPU)',
        'config': {'workload': 'CSM + beamforming fused: 512 mics, 10 s at 51.2 kHz float32, 16 correlated sources '
                               '+ 5% noise, Welch bs 2048 Hanning 50% (499 blocks, 1025 bins), RectGrid 128x128 '
                               'z=1.0, true location, r_diag, one hipGraph replay',
                   'num_mics': nm, 'num_grid': int(gpos.shape[1]), 'num_freqs': int(len(p.ind)),
                   'num_blocks': int(p.num_blocks), 'launch_shape': tiles, 'parallelism': 'single GPU'},
        'csm_ms': t_csm * 1e3, 'beamform_ms': t_bf * 1e3, 'eager_ms': t_eager * 1e3,
        'csm_tflops': csm_flops / t_csm / 1e12,
        'beamform_alg_tflops': p.map_bins * 8.0 * nm * (nm + 1) / t_bf / 1e12,
        'max_rel_err': err, 'max_rel_err_bins': bins,
    }


def large3d_bench(args):
    """Large 3D map: 1024 mics on a 3 m spiral + seeded jitter, RectGrid3D 64 x 64 x 32 (131,072
    points; x, y in [-1, 1], z from 0.5, increment 2/63), 256 bins (bins 2..257 of block size 1024
    at 51.2 kHz) of an analytic complex64 CSM of 32 seeded 3D sources (20 dB spread), true level,
    r_diag; autotuned split-bf16 MFMA against the plain FP32 kernel."""
    import torch

    import acoular_amd as aa
    from acoular_amd import _lib, autotune, synthetic
    from acoular_amd.fbeamform import make_job
    from oracle import beamform as B

    dev = torch.device('cuda', 0)
    nm, seed = 1024, 4004
    mpos = synthetic.spiral_array(nm, 3.0, seed)
    inc = 2.0 / 63
    grid = aa.RectGrid3D(x_min=-1.0, x_max=1.0, y_min=-1.0, y_max=1.0, z_min=0.5, z_max=0.5 + 31 * inc,
                         increment=inc)
    gpos = grid.pos
    assert gpos.shape[1] == 64 * 64 * 32
    spos = synthetic.source_positions(32, [-0.9, -0.9, 0.6], [0.9, 0.9, 1.4], seed + 1)
    freqs = np.arange(2, 258) * 51200.0 / 1024
    log('large3d: analytic CSM 256 x 1024 x 1024 on the GPU')
    csm = synthetic.analytic_csm(mpos, spos, synthetic.level_powers(32, 20.0, seed + 2), freqs, device=dev)
    k = torch.from_numpy(2 * np.pi * freqs / 343.0).to(dev)
    gd = torch.from_numpy(gpos).to(dev)
    md = torch.from_numpy(mpos).to(dev)
    lib = _lib.lib()
    res = {}
    outs = {}
    for form in ('mfma_bf16x3', 'plain_f32'):
        out = torch.empty((len(freqs), gpos.shape[1]), dtype=torch.float32, device=dev)
        job = make_job(csm, k, gd, md, np.zeros(3), 'true level', True, nm / (nm - 1.0), form, 'reference', out,
                       None, clip=True)
        tiles = None
        if args.no_tune:
            job.tile_kernel = job.tile_cols = job.tile_freqs = 0
        elif form != 'plain_f32':
            src = 'autotuner table'
            if args.tune or job.tile_cols == 0:
                tr = autotune.tune_job(job, reps=1)
                job.tile_kernel, job.tile_cols, job.tile_freqs = tr['tile_kernel'], tr['tile_cols'], tr['tile_freqs']
                src = 'tuned now'
            tiles = {'tile_kernel': job.tile_kernel, 'tile_cols': job.tile_cols, 'tile_freqs': job.tile_freqs,
                     'source': src}
            log(f'launch shape: {tiles}')
        ws = _lib.workspace(lib.abf_beamform_workspace_size(job))

        def step(job=job, ws=ws):
            _lib.check(lib.abf_beamform(job, _lib.ptr(ws), ws.numel(), _lib.stream_handle()))

        n = max(1, args.steps // 3) if form != 'plain_f32' else 1
        t = time_steps(step, n, 1)
        res[form] = {'ms_per_step': t * 1e3, 'map_bins_per_s': out.numel() / t, 'launch_shape': tiles,
                     'alg_tflops': out.numel() * 8.0 * nm * (nm + 1) / t / 1e12}
        log(f"{form}: {t * 1e3:.1f} ms -> {res[form]['map_bins_per_s']:.4g} map-bins/s")
        outs[form] = out
        del ws
    err = None
    bins = [40, 200]
    if not args.no_compare:
        # oracle on 8192 seeded grid points (plus the map maximum) of two bins
        rng = np.random.default_rng(seed + 3)
        errs = []
        for b in bins:
            got = outs['mfma_bf16x3'][b].double().cpu().numpy()
            sub = np.unique(np.concatenate([rng.choice(gpos.shape[1], 8192, replace=False), [int(np.argmax(got))]]))
            ref = B.beamformer_map(csm[b:b + 1].cpu().numpy(), freqs[b:b + 1], gpos[:, sub], mpos, 'true level',
                                   True)[0]
            errs.append(float(np.max(np.abs(got[sub] - ref)) / np.max(np.abs(ref))))
        err = max(errs)
        log(f'max relative map error on bins {bins} (8192 sampled points): {err:.2e}')
    m = res['mfma_bf16x3']
    return {
        'metric': 'map-bins/s (grid pts x freqs) at <=1e-4 rel err vs FP64', 'value': m['map_bins_per_s'],
        'unit': 'map-bins/s', 'n_gpus': 1, 'ms_per_step': m['ms_per_step'], 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'bf16x3 (split hi+lo, fp32 accumulate)',
        'data': 'synthetic (seeded analytic CSM)',
        'config': {'workload': 'large 3D map: 1024 mics (3 m spiral + jitter), RectGrid3D 64x64x32, 256 bins, '
                               '32 sources, true level, r_diag', 'num_mics': nm, 'num_grid': int(gpos.shape[1]),
                   'num_freqs': len(freqs), 'parallelism': 'single GPU'},
        'forms': res, 'mfma_over_plain': m['map_bins_per_s'] / res['plain_f32']['map_bins_per_s'],
        'max_rel_err': err, 'max_rel_err_bins': bins,
    }


def eig_bench(args):
    """Eigenvalue path on the 256-mic 201 x 201 513-bin workload (classic, r_diag): the batched
    eigendecomposition of the resident CSM (PowerSpectra.eig_device) and the all-bins
    eigenvalue-form beamformer (abf_beamform_eig_grid) for BeamformerEig (the largest eigenpair)
    and BeamformerMusic (the M - 1 smallest)."""
    import torch

    import acoular_amd as aa
    from acoular_amd.fbeamform import beamform_eig_device
    from oracle import beamform as B

    dev = torch.device('cuda', 0)
    wl = build_workload(dev)
    nm, freqs, gpos = wl['nm'], wl['freqs'], wl['gpos']
    ps = aa.PowerSpectraImport(csm=wl['csm'], frequencies=freqs, precision='complex64')
    grid = aa.ImportGrid(pos=gpos)
    sv = aa.SteeringVector(grid=grid, mics=aa.MicGeom(pos_total=wl['mpos']), steer_type='classic')
    norm = nm / (nm - 1.0)

    def eig_step():
        ps._cache.clear()
        return ps.eig_device()

    t_eig = time_steps(eig_step, max(1, args.steps // 5), 1)
    log(f'eigh of {len(freqs)} x {nm} x {nm} complex64 on the device: {t_eig * 1e3:.1f} ms')
    ev, evec = ps.eig_device()
    ind = np.arange(len(freqs))
    res = {}
    outs = {}
    for name, e0, k in (('eig', nm - 1, 1), ('music', 0, nm - 1)):
        box = {}

        def step(e0=e0, k=k, box=box):
            box['out'] = beamform_eig_device(ev, evec, freqs, ind, sv, 'classic', name == 'eig', norm, e0, k,
                                             clip=name == 'eig')

        step()
        t = event_time(step, max(1, args.steps if name == 'eig' else args.steps // 5))
        bins = len(freqs) * gpos.shape[1]
        res[name] = {'eig_first': e0, 'num_eig': k, 'ms_per_step': t * 1e3, 'map_bins_per_s': bins / t,
                     'alg_tflops': bins * 8.0 * nm * k / t / 1e12}
        outs[name] = box['out']
        log(f"{name} (eigenvectors {e0}..{e0 + k - 1}): {t * 1e3:.2f} ms -> {bins / t:.4g} map-bins/s")
    err = None
    music_errs = []
    sample_bins = [40, 300]
    if not args.no_compare:
        r0 = B.steer_r0(gpos, np.zeros(3))
        rm = B.dist_mat(gpos, wl['mpos'])
        errs = []
        for b in sample_bins:
            lam, V = np.linalg.eigh(wl['csm'][b].cpu().numpy().astype(np.complex128))
            ref = B.freq_beamformer_eig(lam[-1:], V[:, -1:], r0, rm, 2 * np.pi * freqs[b] / 343.0, norm, True,
                                        'classic')[0]
            ref = np.where(ref < 0, 0.0, ref)
            got = outs['eig'][b].double().cpu().numpy()
            errs.append(float(np.abs(got - ref).max() / np.abs(ref).max()))
            # the 255 smallest eigenpairs without r_diag (the fp32 MFMA kernel)
            refm = B.freq_beamformer_eig(lam[:-1], V[:, :-1], r0, rm, 2 * np.pi * freqs[b] / 343.0, norm, False,
                                         'classic')[0]
            gotm = outs['music'][b].double().cpu().numpy()
            music_errs.append(float(np.abs(gotm - refm).max() / np.abs(refm).max()))
        err = max(errs)
        log(f'BeamformerEig max relative map error on bins {sample_bins}: {err:.2e}; '
            f'255 smallest eigenpairs: {max(music_errs):.2e}')
    cpu = None
    if not args.no_cpu:
        # the oracle's eigenvalue form (numpy, one thread) on whole maps of the same workload
        r0 = B.steer_r0(gpos, np.zeros(3))
        rm = B.dist_mat(gpos, wl['mpos'])
        evh, evech = ev.cpu().numpy(), evec.cpu().numpy().astype(np.complex128)
        done, t0 = 0, time.perf_counter()
        for b in range(7, len(freqs), 37):
            B.freq_beamformer_eig(evh[b, -1:], evech[b][:, -1:], r0, rm, 2 * np.pi * freqs[b] / 343.0, norm, True,
                                  'classic')
            done += 1
            if time.perf_counter() - t0 > args.cpu_seconds:
                break
        dt = time.perf_counter() - t0
        cpu = {'value': gpos.shape[1] * done / dt, 'unit': 'map-bins/s', 'cores': 1, 'kind': 'port',
               'sample': f'{done} full 201x201 maps of the eigenvalue form (largest eigenpair), oracle numpy, '
                         f'{dt:.1f} s'}
        log(f"CPU oracle: {cpu['value']:.4g} map-bins/s")
    m = res['eig']
    return {
        'metric': 'map-bins/s (grid pts x freqs) at <=1e-4 rel err vs FP64', 'value': m['map_bins_per_s'],
        'unit': 'map-bins/s', 'n_gpus': 1, 'ms_per_step': m['ms_per_step'], 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32', 'data': 'synthetic (seeded analytic CSM)',
        'config': {'workload': 'eigenvalue path: 256 mics (2 m disk), 201x201 grid, 513 bins, 8 sources, classic, '
                               'r_diag; BeamformerEig (largest eigenpair) and BeamformerMusic (M - 1 smallest)',
                   'num_mics': nm, 'num_grid': int(gpos.shape[1]), 'num_freqs': len(freqs),
                   'parallelism': 'single GPU'},
        'eigh_ms': t_eig * 1e3, 'forms': res, 'max_rel_err': err, 'max_rel_err_bins': sample_bins,
        'max_rel_err_music': max(music_errs) if music_errs else None,
        'cpu_baseline': cpu,
    }


def damas_bench(args):
    """DAMAS on the 64-mic workload: 64 mics in a 1 m disk, RectGrid 41 x 41 (increment 0.01,
    z = 0.3), 65 bins of block size 128 at 51.2 kHz (bins 1..64 plus DC) of an analytic complex128
    CSM of 3 seeded sources, true level, r_diag, 100 sweeps: the conventional maps
    (abf_beamform), the full PSFs of every bin (abf_psf, float64) and the Gauss-Seidel sweeps
    (abf_damas), all bins per launch."""
    import torch

    import acoular_amd as aa
    from acoular_amd import synthetic
    from oracle import beamform as B
    from oracle import cref

    dev = torch.device('cuda', 0)
    nm, seed, n_iter = 64, 1001, 100
    mpos = synthetic.disk_array(nm, 1.0, seed)
    grid = aa.RectGrid(x_min=-0.2, x_max=0.2, y_min=-0.2, y_max=0.2, z=0.3, increment=0.01)
    G = grid.size
    assert G == 41 * 41
    spos = synthetic.source_positions(3, [-0.2, -0.2, 0.3], [0.2, 0.2, 0.3], seed + 1)
    freqs = np.arange(65) * 51200.0 / 128
    csm = synthetic.analytic_csm(mpos, spos, synthetic.level_powers(3, 10.0, seed + 2), freqs, dtype=np.complex128)
    mics = aa.MicGeom(pos_total=mpos)
    sv = aa.SteeringVector(grid=grid, mics=mics, steer_type='true level')
    dcsm = torch.from_numpy(csm).to(dev)
    ind = np.arange(len(freqs))
    norm = nm / (nm - 1.0)
    box = {}

    def dirty():
        # float64 dirty maps, as BeamformerDamas forms them
        box['y'] = aa.beamform_f64_device(dcsm, freqs, ind, sv, 'true level', True, norm, clip=True)

    def psf():
        box['A'] = aa.psf_device(sv, freqs, np.arange(G))

    def solve():
        box['x'] = aa.damas_device(box['A'], box['y'], n_iter, 1.0)

    def step():
        dirty()
        psf()
        solve()

    step()
    t_step = time_steps(step, max(1, args.steps), 1)
    parts = {name: event_time(fn, 3) * 1e3 for name, fn in (('dirty_maps', dirty), ('psf', psf), ('damas', solve))}
    bins = len(freqs) * G
    log(f'DAMAS step {t_step * 1e3:.2f} ms ({parts}) -> {bins / t_step:.4g} map-bins/s')
    # DAMAS kernel: each sweep reads every bin's G x G float64 PSF once
    damas_bytes = len(freqs) * n_iter * G * G * 8.0
    roof = {'bound': 'hbm', 'achieved': round(damas_bytes / (parts['damas'] / 1e3) / 1e9, 1), 'peak': 8000.0,
            'unit': 'GB/s', 'kernel': 'damas_kernel<true>'}
    roof['frac'] = round(roof['achieved'] / roof['peak'], 4)
    roof['traffic'] = None
    err = None
    if not args.no_compare:
        r0 = B.steer_r0(grid.pos, np.zeros(3))
        rm = B.dist_mat(grid.pos, mpos)
        sample = [5, 33]
        y = B.beamformer_map(csm, freqs, grid.pos, mpos, 'true level', True, ind=sample)
        A = np.stack([cref.psf('true level', r0, rm, 2 * np.pi * freqs[b] / 343.0, np.arange(G)) for b in sample])
        want = y.copy()
        cref.damas_gauss_seidel(A, y, want, n_iter, 1.0)
        got = box['x'][sample].cpu().numpy()
        err = max(float(np.abs(got[i] - want[i]).max() / np.abs(want[i]).max()) for i in range(len(sample)))
        log(f'DAMAS max relative map error on bins {sample}: {err:.2e}')
    cpu = None
    if not args.no_cpu:
        # the C restatement (OpenMP) of the PSF and the sweeps on whole bins of the same workload
        r0 = B.steer_r0(grid.pos, np.zeros(3))
        rm = B.dist_mat(grid.pos, mpos)
        y = box['y'].cpu().numpy()
        done, t0 = 0, time.perf_counter()
        for b in range(3, len(freqs), 7):
            A = cref.psf('true level', r0, rm, 2 * np.pi * freqs[b] / 343.0, np.arange(G))[None]
            x = y[b:b + 1].copy()
            cref.damas_gauss_seidel(A, y[b:b + 1], x, n_iter, 1.0)
            done += 1
            if time.perf_counter() - t0 > args.cpu_seconds:
                break
        dt = time.perf_counter() - t0
        cpu = {'value': G * done / dt, 'unit': 'map-bins/s', 'cores': cref.num_threads(), 'kind': 'port',
               'sample': f'{done} bins of the same workload (PSF + {n_iter} sweeps, C + OpenMP; the dirty map not '
                         f'timed), {dt:.1f} s'}
        log(f"CPU: {cpu['value']:.4g} map-bins/s")
    return {
        'metric': 'map-bins/s (grid pts x freqs) at <=1e-4 rel err vs FP64', 'value': bins / t_step,
        'unit': 'map-bins/s', 'n_gpus': 1, 'ms_per_step': t_step * 1e3, 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f64',
        'data': 'synthetic (seeded analytic CSM)',
        'config': {'workload': 'DAMAS: 64 mics (1 m disk), RectGrid 41x41 z=0.3, 65 bins (bs 128, 51.2 kHz), '
                               '3 sources, true level, r_diag, 100 sweeps', 'num_mics': nm, 'num_grid': G,
                   'num_freqs': len(freqs), 'parallelism': 'single GPU'},
        'parts_ms': parts, 'roofline': roof, 'max_rel_err': err, 'cpu_baseline': cpu,
    }


def rfft_bench(args):
    """RFFT (block-wise one-sided FFT, fprocess.py:35-180) of the fused workload's time data: 512
    channels, 10 s at 51.2 kHz (float32, 1.05 GB resident), block size 2048, Hanning, 50 %
    overlap (499 blocks), 'energy' scaling: all blocks and channels in one abf_rfft_blocks call,
    output [499][1025][512] complex64.  HBM-bound: algorithmic bytes = every time sample read
    once + every spectrum value written once."""
    import torch

    import acoular_amd as aa
    from acoular_amd import _lib
    from oracle import fprocess as P

    dev = torch.device('cuda', 0)
    nch, fs, n, bs = 512, 51200.0, 512000, 2048
    g = torch.Generator(device=dev).manual_seed(7)
    tdev = torch.randn((n, nch), generator=g, device=dev, dtype=torch.float32)

    class _Src:
        sample_freq = fs
        num_samples = n
        num_channels = nch

        def result_device(self, num):
            return tdev

    fft = aa.RFFT(source=_Src(), block_size=bs, window='Hanning', overlap='50%', scaling='energy')
    hop = bs // 2
    nb = (n - bs) // hop + 1
    nf = bs // 2 + 1
    out = torch.empty((nb, nf, nch), dtype=torch.complex64, device=dev)
    lib = _lib.lib()
    wdev = _lib.to_device(np.hanning(bs).astype(np.float32))
    sdev = _lib.to_device(fft.bin_scale().astype(np.float32))
    ws = _lib.workspace(lib.abf_rfft_workspace_size(bs))

    def step():
        _lib.check(lib.abf_rfft_blocks(_lib.ptr(tdev), n, nch, bs, hop, nb, _lib.ptr(wdev), _lib.ptr(sdev),
                                       _lib.ptr(out), _lib.ptr(ws), ws.numel(), _lib.stream_handle()), 'RFFT')

    step()
    t_step = time_steps(step, max(1, args.steps), args.warmup)
    t_k = event_time(step, 5)
    alg_bytes = n * nch * 4.0 + nb * nf * nch * 8.0
    roof = {'bound': 'hbm', 'achieved': round(alg_bytes / t_k / 1e9, 1), 'peak': 8000.0, 'unit': 'GB/s',
            'kernel': 'welch_fft_radix_kernel (block-major output)', 'algorithmic_bytes_per_launch': alg_bytes,
            'traffic': None}
    roof['frac'] = round(roof['achieved'] / roof['peak'], 4)
    values = nb * nf * nch
    log(f'RFFT step {t_step * 1e3:.2f} ms ({t_k * 1e3:.2f} ms kernel time) -> {values / t_step:.4g} spectrum '
        f'values/s, {roof["achieved"]} GB/s')
    err = None
    if not args.no_compare:
        x = tdev[: 6 * hop + bs].cpu().numpy()
        want = P.rfft_blocks(x, bs, 'Hanning', '50%', 'energy')
        got = out[: want.shape[0]].cpu().numpy().reshape(want.shape[0], -1)
        err = float(np.abs(got - want).max() / np.abs(want).max())
        log(f'RFFT max error relative to the largest magnitude on 7 blocks: {err:.2e}')
    cpu = None
    if not args.no_cpu:
        x = tdev.cpu().numpy()
        done, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < args.cpu_seconds and done < nb:
            b = done
            P.rfft_blocks(x[b * hop : b * hop + bs], bs, 'Hanning', 'None', 'energy')
            done += 1
        dt = time.perf_counter() - t0
        cpu = {'value': done * nf * nch / dt, 'unit': 'spectrum values/s', 'cores': 1, 'kind': 'port',
               'sample': f'{done} blocks of the same workload (oracle: scipy rfft in float64), {dt:.1f} s'}
        log(f"CPU: {cpu['value']:.4g} spectrum values/s")
    return {
        'metric': 'spectrum values/s (blocks x bins x channels)', 'value': values / t_step,
        'unit': 'spectrum values/s', 'n_gpus': 1, 'ms_per_step': t_step * 1e3, 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32', 'data': 'synthetic (seeded normal noise)',
        'config': {'workload': 'RFFT: 512 channels, 10 s at 51.2 kHz, bs 2048, Hanning, 50 %, energy scaling',
                   'blocks': nb, 'num_freqs': nf, 'channels': nch, 'parallelism': 'single GPU'},
        'roofline': roof, 'max_rel_err': err, 'cpu_baseline': cpu,
    }


def clean_bench(args):
    """CLEAN on the 64-mic workload of the DAMAS mode (64 mics, RectGrid 41 x 41 at z = 0.3, 65
    bins of block size 128 at 51.2 kHz, analytic complex128 CSM of 3 seeded sources), true level,
    r_diag, the reference's defaults (up to 100 iterations, damp 0.6, float64 PSF): the float64
    dirty maps (abf_beamform_f64, clipped) and the iterations of all bins together (abf_clean)."""
    import torch

    import acoular_amd as aa
    from acoular_amd import synthetic
    from oracle import beamform as B
    from oracle import deconv as D

    dev = torch.device('cuda', 0)
    nm, seed = 64, 1001
    mpos = synthetic.disk_array(nm, 1.0, seed)
    grid = aa.RectGrid(x_min=-0.2, x_max=0.2, y_min=-0.2, y_max=0.2, z=0.3, increment=0.01)
    G = grid.size
    spos = synthetic.source_positions(3, [-0.2, -0.2, 0.3], [0.2, 0.2, 0.3], seed + 1)
    freqs = np.arange(65) * 51200.0 / 128
    csm = synthetic.analytic_csm(mpos, spos, synthetic.level_powers(3, 10.0, seed + 2), freqs, dtype=np.complex128)
    mics = aa.MicGeom(pos_total=mpos)
    sv = aa.SteeringVector(grid=grid, mics=mics, steer_type='true level')
    dcsm = torch.from_numpy(csm).to(dev)
    ind = np.arange(len(freqs))
    norm = nm / (nm - 1.0)
    box = {}

    def step():
        y = aa.beamform_f64_device(dcsm, freqs, ind, sv, 'true level', True, norm, clip=True)
        box['x'] = aa.clean_device(y, freqs, sv, 100, 0.6)

    step()
    t_step = time_steps(step, max(1, args.steps), 1)
    bins = len(freqs) * G
    log(f'CLEAN step {t_step * 1e3:.2f} ms -> {bins / t_step:.4g} map-bins/s')
    r0 = B.steer_r0(grid.pos, np.zeros(3))
    rm = B.dist_mat(grid.pos, mpos)
    err = None
    if not args.no_compare:
        sample = [5, 33, 64]
        y = B.beamformer_map(csm, freqs, grid.pos, mpos, 'true level', True, ind=sample)
        y[y < 0] = 0.0
        want = np.stack([D.clean(y[i], 'true level', r0, rm, 2 * np.pi * freqs[b] / 343.0) for i, b in
                         enumerate(sample)])
        got = box['x'][sample].cpu().numpy()
        err = max(float(np.abs(got[i] - want[i]).max() / np.abs(want[i]).max()) for i in range(len(sample)))
        same = bool(np.array_equal(got != 0, want != 0))
        log(f'CLEAN max relative map error on bins {sample}: {err:.2e}, same support: {same}')
        if not same:
            err = float('inf')
    cpu = None
    if not args.no_cpu:
        # the oracle (numpy) on whole bins of the same workload: dirty map and iterations
        done, t0 = 0, time.perf_counter()
        for b in range(3, len(freqs), 5):
            y = B.beamformer_map(csm, freqs, grid.pos, mpos, 'true level', True, ind=[b])[0]
            y[y < 0] = 0.0
            D.clean(y, 'true level', r0, rm, 2 * np.pi * freqs[b] / 343.0)
            done += 1
            if time.perf_counter() - t0 > args.cpu_seconds:
                break
        dt = time.perf_counter() - t0
        cpu = {'value': G * done / dt, 'unit': 'map-bins/s', 'cores': 1, 'kind': 'port',
               'sample': f'{done} bins of the same workload (dirty map + iterations, numpy), {dt:.1f} s'}
        log(f"CPU: {cpu['value']:.4g} map-bins/s")
    return {
        'metric': 'map-bins/s (grid pts x freqs) at <=1e-4 rel err vs FP64', 'value': bins / t_step,
        'unit': 'map-bins/s', 'n_gpus': 1, 'ms_per_step': t_step * 1e3, 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f64',
        'data': 'synthetic (seeded analytic CSM)',
        'config': {'workload': 'CLEAN: 64 mics (1 m disk), RectGrid 41x41 z=0.3, 65 bins (bs 128, 51.2 kHz), '
                               '3 sources, true level, r_diag, n_iter 100, damp 0.6, float64 PSF', 'num_mics': nm,
                   'num_grid': G, 'num_freqs': len(freqs), 'parallelism': 'single GPU'},
        'max_rel_err': err, 'cpu_baseline': cpu,
    }


def cleansc_bench(args):
    """CLEAN-SC on the 64-mic workload of the DAMAS mode (64 mics, RectGrid 41 x 41 at z = 0.3,
    65 bins of block size 128 at 51.2 kHz, analytic complex128 CSM of 3 seeded sources), true
    level, r_diag, the reference's defaults (up to 2 M = 128 iterations, damp 0.6, stopn 3): the
    float64 dirty maps (abf_beamform_f64) and the iterations of all bins together (abf_cleansc)."""
    import torch

    import acoular_amd as aa
    from acoular_amd import synthetic
    from oracle import deconv as D

    dev = torch.device('cuda', 0)
    nm, seed = 64, 1001
    mpos = synthetic.disk_array(nm, 1.0, seed)
    grid = aa.RectGrid(x_min=-0.2, x_max=0.2, y_min=-0.2, y_max=0.2, z=0.3, increment=0.01)
    G = grid.size
    spos = synthetic.source_positions(3, [-0.2, -0.2, 0.3], [0.2, 0.2, 0.3], seed + 1)
    freqs = np.arange(65) * 51200.0 / 128
    csm = synthetic.analytic_csm(mpos, spos, synthetic.level_powers(3, 10.0, seed + 2), freqs, dtype=np.complex128)
    csm += 1e-3 * np.eye(nm)[None]
    mics = aa.MicGeom(pos_total=mpos)
    sv = aa.SteeringVector(grid=grid, mics=mics, steer_type='true level')
    dcsm = torch.from_numpy(csm).to(dev)
    ind = np.arange(len(freqs))
    norm = nm / (nm - 1.0)
    box = {}

    def step():
        box['x'] = aa.cleansc_device(dcsm, freqs, ind, sv, 'true level', True, norm, 0, 0.6, 3)

    step()
    t_step = time_steps(step, max(1, args.steps), 1)
    bins = len(freqs) * G
    log(f'CLEAN-SC step {t_step * 1e3:.2f} ms -> {bins / t_step:.4g} map-bins/s')
    err = None
    if not args.no_compare:
        sample = [5, 33, 64]
        want = D.clean_sc(csm[sample], freqs[sample], grid.pos, mpos, 'true level', r_diag=True)
        got = box['x'][sample].cpu().numpy()
        err = max(float(np.abs(got[i] - want[i]).max() / np.abs(want[i]).max()) for i in range(len(sample)))
        same = bool(np.array_equal(got != 0, want != 0))
        log(f'CLEAN-SC max relative map error on bins {sample}: {err:.2e}, same support: {same}')
        if not same:
            err = float('inf')
    cpu = None
    if not args.no_cpu:
        # the oracle (numpy) on whole bins of the same workload
        done, t0 = 0, time.perf_counter()
        for b in range(3, len(freqs), 5):
            D.clean_sc(csm[b:b + 1], freqs[b:b + 1], grid.pos, mpos, 'true level', r_diag=True)
            done += 1
            if time.perf_counter() - t0 > args.cpu_seconds:
                break
        dt = time.perf_counter() - t0
        cpu = {'value': G * done / dt, 'unit': 'map-bins/s', 'cores': 1, 'kind': 'port',
               'sample': f'{done} bins of the same workload (dirty map + iterations, numpy), {dt:.1f} s'}
        log(f"CPU: {cpu['value']:.4g} map-bins/s")
    return {
        'metric': 'map-bins/s (grid pts x freqs) at <=1e-4 rel err vs FP64', 'value': bins / t_step,
        'unit': 'map-bins/s', 'n_gpus': 1, 'ms_per_step': t_step * 1e3, 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f64',
        'data': 'synthetic (seeded analytic CSM)',
        'config': {'workload': 'CLEAN-SC: 64 mics (1 m disk), RectGrid 41x41 z=0.3, 65 bins (bs 128, 51.2 kHz), '
                               '3 sources, true level, r_diag, n_iter 2M, damp 0.6, stopn 3', 'num_mics': nm,
                   'num_grid': G, 'num_freqs': len(freqs), 'parallelism': 'single GPU'},
        'max_rel_err': err, 'cpu_baseline': cpu,
    }


def _c1_scene(seed=1001):
    """The 64-mic configuration's sources: 64 mics in a 1 m disk; 3 unit-variance white-noise
    sources at uniform [-0.2, 0.2]^2, z = 0.3; independent noise of rms 0.1 at every microphone;
    51.2 kHz, 1 s."""
    import acoular_amd as aa
    from acoular_amd import synthetic

    fs, n, nm = 51200.0, 51200, 64
    mics = aa.MicGeom(pos_total=synthetic.disk_array(nm, 1.0, seed))
    spos = synthetic.source_positions(3, [-0.2, -0.2, 0.3], [0.2, 0.2, 0.3], seed + 1)
    srcs = [aa.PointSource(signal=aa.WNoiseGenerator(sample_freq=fs, num_samples=n, seed=seed + 10 + i),
                           loc=spos[:, i], mics=mics) for i in range(3)]
    noise = aa.UncorrelatedNoiseSource(signal=aa.WNoiseGenerator(sample_freq=fs, num_samples=n, seed=seed + 100,
                                                                 rms=0.1), mics=mics)
    mixer = aa.SourceMixer(sources=[*srcs, noise])
    grid = aa.RectGrid(x_min=-0.2, x_max=0.2, y_min=-0.2, y_max=0.2, z=0.3, increment=0.01)
    return mics, srcs, noise, mixer, grid


def c1_bench(args):
    """The reference's own CPU-runnable configuration end to end on the GPU: source synthesis
    (three point sources + microphone noise, mixed), Welch CSM (block size 128, Hanning, 50 %:
    799 blocks, 65 bins) and true-level r_diag maps on a 41 x 41 grid for every bin."""
    import torch

    import acoular_amd as aa
    from oracle import beamform as B
    from oracle import cref
    from oracle import sources as OS
    from oracle import spectra as OSp

    mics, srcs, noise, mixer, grid = _c1_scene()
    box = {}

    def step():
        ps = aa.PowerSpectra(source=mixer, block_size=128, window='Hanning', overlap='50%', precision='complex64')
        bf = aa.BeamformerBase(freq_data=ps, steer=aa.SteeringVector(grid=grid, mics=mics, steer_type='true level'),
                               r_diag=True, precision='float32')
        box['maps'] = bf.result[:]
        box['csm'] = ps.csm_device

    step()
    torch.cuda.synchronize()
    t_step = time_steps(step, max(1, args.steps), 1)
    bins = 65 * grid.size
    log(f'C1 end to end: {t_step * 1e3:.2f} ms -> {bins / t_step:.4g} map-bins/s')
    # parts on the device
    dev_data = mixer.result_device(128)
    t_syn = event_time(lambda: mixer.result_device(128), 3)
    log(f'  synthesis {t_syn * 1e3:.2f} ms')
    err = None
    if not args.no_compare:
        data = dev_data.cpu().numpy()
        csm = OSp.welch_csm(data, 128, 'Hanning', '50%')
        freqs = np.arange(65) * 51200.0 / 128
        sample = [10, 40]
        ref = B.beamformer_map(csm, freqs, grid.pos, mics.pos, 'true level', True, ind=sample)
        got = box['maps'][sample]
        err = max(float(np.abs(got[i] - ref[i]).max() / np.abs(ref[i]).max()) for i in range(len(sample)))
        log(f'max relative map error on bins {sample} (oracle Welch + beamformer on the same samples): {err:.2e}')
    cpu = None
    if not args.no_cpu:
        # the oracle pipeline on the host: sources (numpy, scipy resampling), Welch (numpy), beamformer
        # (C + OpenMP)
        t0 = time.perf_counter()
        parts = []
        for s in srcs:
            rm = s.env.apparent_r(s.loc, s.mics.pos).reshape(1, -1)
            ind0 = (-rm / s.env.c - s.start_t + s.start) * s.sample_freq * s.up
            parts.append(OS.point_source(OS.usignal(s.signal.signal(), s.up), rm, ind0, s.up, s.prepadding,
                                         s.num_samples, 128, float(s.num_samples)))
        parts.append(OS.white_noise(np.arange(64) + noise.signal.seed, noise.num_samples, noise.signal.rms))
        data = OS.mix(parts, [1.0] * 4)
        csm = OSp.welch_csm(data, 128, 'Hanning', '50%')
        freqs = np.arange(65) * 51200.0 / 128
        r0 = B.steer_r0(grid.pos, np.zeros(3))
        rm = cref.dist_mat(grid.pos, mics.pos)
        for b in range(65):
            cref.beamformer_full_csm(csm[b], r0, rm, 2 * np.pi * freqs[b] / 343.0, 64 / 63.0, True, 3)
        dt = time.perf_counter() - t0
        cpu = {'value': bins / dt, 'unit': 'map-bins/s', 'cores': cref.num_threads(), 'kind': 'port',
               'sample': f'the whole configuration once (synthesis and Welch in numpy, beamformer in C + OpenMP), '
                         f'{dt:.1f} s'}
        log(f"CPU pipeline: {dt:.2f} s -> {cpu['value']:.4g} map-bins/s")
    return {
        'metric': 'map-bins/s (grid pts x freqs) at <=1e-4 rel err vs FP64', 'value': bins / t_step,
        'unit': 'map-bins/s', 'n_gpus': 1, 'ms_per_step': t_step * 1e3, 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f64 synthesis, f32 CSM, bf16x3 maps',
        'data': 'synthetic (seeded sources synthesised on the GPU)',
        'config': {'workload': '64 mics (1 m disk), 3 white-noise point sources + mic noise rms 0.1, 51.2 kHz, 1 s; '
                               'Welch bs 128 Hanning 50 % (799 blocks, 65 bins); RectGrid 41x41 z=0.3; true level, '
                               'r_diag; synthesis + CSM + maps', 'num_mics': 64, 'num_grid': grid.size,
                   'num_freqs': 65, 'parallelism': 'single GPU'},
        'synthesis_ms': t_syn * 1e3, 'max_rel_err': err, 'cpu_baseline': cpu,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=10)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--form', default='mfma_bf16x3', choices=['mfma_bf16x3', 'mfma_f16x2', 'mfma_f32', 'plain_f32'])
    ap.add_argument('--no-compare', action='store_true', help='skip the other forms')
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline')
    ap.add_argument('--no-tune', action='store_true', help='default launch shapes (no autotuner table, no tuning)')
    ap.add_argument('--tune', action='store_true', help='tune the launch shape now even if the table has it')
    ap.add_argument('--cpu-seconds', type=float, default=15.0)
    ap.add_argument('--no-extras', action='store_true',
                    help='headline only: skip the per-variant, fused, large-3D and serving figures of the default line')
    ap.add_argument('--mode', default='headline', choices=['headline', 'copy', 'serving', 'fused', 'large3d', 'eig', 'api',
                                                          'damas', 'cleansc', 'clean', 'rfft', 'c1'],
                    help="'copy': a 1 GiB device copy, the calibration kernel for HBM byte counters; "
                         "'serving': the mixed-size request loop (Poisson arrivals) against the unbatched "
                         "plain kernel; 'fused': Welch CSM + beamforming of 512 mics, 10 s at 51.2 kHz, "
                         "1025 bins, 128x128 grid, replayed as one hipGraph; 'large3d': 1024 mics, "
                         "64x64x32 grid, 256 bins, autotuned MFMA vs the plain kernel")
    ap.add_argument('--serve-requests', type=int, default=10000, help='serving: requests of the timed run')
    ap.add_argument('--serve-load', type=float, default=0.65, help='serving: offered load / saturation rate')
    ap.add_argument('--serve-sat', type=int, default=600, help='serving: requests of the saturation run')
    ap.add_argument('--serve-plain', type=int, default=300, help='serving: requests of the plain baseline runs')
    args = ap.parse_args()
    if args.gpus != 1:
        print(f'bench.py runs on one GPU only (asked for {args.gpus}); sharding this path over several GPUs '
              '(bins or grid tiles per GPU, spectra exchanged by one all-to-all) is designed but not built.',
              file=sys.stderr)
        sys.exit(2)

    import torch

    dev = torch.device('cuda', 0)
    if args.mode == 'copy':
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({'mode': 'copy', 'bytes_read_per_dispatch': src.numel() * 4,
                          'bytes_written_per_dispatch': dst.numel() * 4, 'dispatches': 3}))
        return
    if args.mode == 'serving':
        print(json.dumps(serving_bench(args)), flush=True)
        return
    if args.mode == 'fused':
        print(json.dumps(fused_bench(args)), flush=True)
        return
    if args.mode == 'large3d':
        print(json.dumps(large3d_bench(args)), flush=True)
        return
    if args.mode == 'eig':
        print(json.dumps(eig_bench(args)), flush=True)
        return
    if args.mode == 'damas':
        print(json.dumps(damas_bench(args)), flush=True)
        return
    if args.mode == 'rfft':
        print(json.dumps(rfft_bench(args)), flush=True)
        return
    if args.mode == 'clean':
        print(json.dumps(clean_bench(args)), flush=True)
        return
    if args.mode == 'cleansc':
        print(json.dumps(cleansc_bench(args)), flush=True)
        return
    if args.mode == 'c1':
        print(json.dumps(c1_bench(args)), flush=True)
        return
    if args.mode == 'api':
        print(json.dumps(api_bench(args)), flush=True)
        return
    log('building workload (analytic CSM 513 x 256 x 256 on the GPU)')
    wl = build_workload(dev)
    runner = Runner(wl, args.form, dev=dev, tuned=not args.no_tune)
    tiles = None
    if args.form != 'plain_f32' and not args.no_tune:
        # launch shape from the autotuner's table (shipped or user cache); tuned now, outside the
        # timed region, when the table lacks this shape class or --tune asks for it
        from acoular_amd import autotune

        src = 'autotuner table'
        if args.tune or runner.job.tile_cols == 0:
            res = autotune.tune_job(runner.job, reps=2)
            runner.job.tile_kernel, runner.job.tile_cols, runner.job.tile_freqs = (
                res['tile_kernel'], res['tile_cols'], res['tile_freqs'])
            src = f"tuned now ({len(res['tried'])} shapes timed)"
        tiles = {'tile_kernel': runner.job.tile_kernel, 'tile_cols': runner.job.tile_cols,
                 'tile_freqs': runner.job.tile_freqs, 'source': src}
        log(f'launch shape: {tiles}')
    log(f'form {args.form}: warmup {args.warmup}, steps {args.steps}')
    t_step = time_steps(runner.step, args.steps, args.warmup)
    value = runner.map_bins / t_step
    log(f'step {t_step * 1e3:.2f} ms -> {value:.4g} map-bins/s')

    # dominant kernel alone (the beamform launch over the prepared CSM), HIP events on its stream
    runner.prepare()
    t_run = event_time(runner.run, max(3, args.steps))
    nm = wl['nm']
    alg_flops = runner.map_bins * 8.0 * nm * (nm + 1)
    bound_peak = MI355X_PEAK['bf16'] if args.form in ('mfma_bf16x3', 'mfma_f16x2') else MI355X_PEAK['fp32']
    achieved = alg_flops / t_run / 1e12
    executed = executed_flops_per_bin(args.form, nm)
    roofline = {'bound': 'mfma' if args.form != 'plain_f32' else 'valu', 'achieved': round(achieved, 2),
                'peak': bound_peak, 'unit': 'TFLOP/s', 'frac': round(achieved / bound_peak, 4),
                'traffic': measured_traffic(args.form, runner),
                'kernel': dominant_kernel_name(args.form, runner),
                'kernel_ms': round(t_run * 1e3, 3),
                'algorithmic_flops_per_launch': alg_flops,
                'executed_mfma_tflops': round(runner.map_bins * executed / t_run / 1e12, 2)}
    log(f'kernel {t_run * 1e3:.2f} ms: {achieved:.1f} TFLOP/s algorithmic ({achieved / bound_peak:.1%} of '
        f'{bound_peak} peak)')

    bins = [3, 100, 257, 511]
    errs = parity_sample(wl, runner, bins)
    err = errs['max_rel_err']
    log(f'max relative map error vs oracle on bins {bins}: {err:.2e} (vs float64 phases '
        f"{errs['max_rel_err_vs_fp64_phase']:.2e}; per point above 1 % of the peak "
        f"{errs['point_rel_err_above_1pct_peak']:.2e})")

    forms = {args.form: value}
    plain = None
    if not args.no_compare:
        for other in ('mfma_f32', 'plain_f32', 'mfma_bf16x3', 'mfma_f16x2'):
            if other == args.form:
                continue
            r = Runner(wl, other, dev=dev)
            t = time_steps(r.step, 2, 1)
            forms[other] = r.map_bins / t
            log(f'form {other}: {forms[other]:.4g} map-bins/s')
            if other == 'plain_f32':
                r.prepare()
                tk = event_time(r.run, 2)
                ex = r.map_bins * executed_flops_per_bin('plain_f32', nm) / tk / 1e12
                plain = {'kernel': 'bf_plain_kernel', 'kernel_ms': round(tk * 1e3, 2),
                         'executed_tflops': round(ex, 2), 'valu_ceiling_tflops': FP32_VALU_NO_PACKED,
                         'frac_of_valu_ceiling': round(ex / FP32_VALU_NO_PACKED, 4),
                         'algorithmic_tflops': round(r.map_bins * 8.0 * nm * (nm + 1) / tk / 1e12, 2),
                         'mfma_over_plain': round(value / forms[other], 2)}
                log(f'plain kernel {tk * 1e3:.1f} ms: {ex:.1f} TFLOP/s executed '
                    f'({ex / FP32_VALU_NO_PACKED:.1%} of the non-packed FP32 VALU ceiling)')
            del r
    cpu = None
    if not args.no_cpu:
        log('CPU baseline (oracle/cpu_ref.c, OpenMP) on a bounded sample')
        cpu = cpu_baseline(wl, args.cpu_seconds)
        log(f"CPU: {cpu['value']:.4g} map-bins/s on {cpu['cores']} threads")
    num_grid, num_freqs = int(wl['gpos'].shape[1]), len(wl['freqs'])
    variants = configs = serving = None
    if not args.no_extras:
        log('variants: every form x classic/inverse x r_diag on the same workload')
        variants = variants_bench(wl, dev)
        del runner
        wl = None
        import gc

        gc.collect()
        torch.cuda.empty_cache()
        sub = argparse.Namespace(**{**vars(args), 'steps': 3, 'warmup': 1, 'no_compare': False})
        log('configuration: fused CSM + beamforming (512 mics, 10 s, bs 2048, 128x128, hipGraph)')
        fz = fused_bench(sub)
        torch.cuda.empty_cache()
        log('configuration: large 3D map (1024 mics, 64x64x32, 256 bins)')
        l3 = large3d_bench(sub)
        torch.cuda.empty_cache()
        nm3 = l3['config']['num_mics']
        configs = {
            'fused_csm_beamform': {
                'workload': fz['config']['workload'], 'map_bins_per_s': fz['value'], 'ms_per_replay': fz['ms_per_step'],
                'max_rel_err': fz['max_rel_err'], 'csm_ms': fz['csm_ms'], 'beamform_ms': fz['beamform_ms'],
                'roofline': {'kernel': 'bf_mfma (beamforming launch)', 'bound': 'mfma', 'unit': 'TFLOP/s',
                             'achieved': round(fz['beamform_alg_tflops'], 2), 'peak': MI355X_PEAK['bf16'],
                             'frac': round(fz['beamform_alg_tflops'] / MI355X_PEAK['bf16'], 4)},
                'csm_roofline': {'kernel': 'welch_herk_kernel (fp32 MFMA, dense flop count)', 'bound': 'mfma',
                                 'unit': 'TFLOP/s', 'achieved': round(fz['csm_tflops'], 2),
                                 'peak': MI355X_PEAK['fp32_mfma'],
                                 'frac': round(fz['csm_tflops'] / MI355X_PEAK['fp32_mfma'], 4)}},
            'large3d': {
                'workload': l3['config']['workload'], 'map_bins_per_s': l3['value'], 'ms_per_step': l3['ms_per_step'],
                'max_rel_err': l3['max_rel_err'], 'plain_map_bins_per_s': l3['forms']['plain_f32']['map_bins_per_s'],
                'mfma_over_plain': l3['mfma_over_plain'],
                'roofline': {'kernel': 'bf_mfma_kernel', 'bound': 'mfma', 'unit': 'TFLOP/s', 'num_mics': nm3,
                             'achieved': round(l3['forms']['mfma_bf16x3']['alg_tflops'], 2),
                             'peak': MI355X_PEAK['bf16'],
                             'frac': round(l3['forms']['mfma_bf16x3']['alg_tflops'] / MI355X_PEAK['bf16'], 4)}},
        }
        log('serving: mixed-size requests at 65 % of the measured saturation, against the unbatched plain kernel')
        ssub = argparse.Namespace(**{**vars(args), 'serve_requests': 2000, 'serve_sat': 300, 'serve_plain': 100,
                                     'serve_load': 0.65})
        sv = serving_bench(ssub)
        serving = {'p50_ms': sv['serving']['p50_ms'], 'p99_ms': sv['serving']['p99_ms'],
                   'req_per_s': sv['serving']['req_per_s'], 'map_bins_per_s': sv['serving']['map_bins_per_s'],
                   'requests': sv['config']['requests'], 'offered_req_per_s': sv['config']['offered_req_per_s'],
                   'saturation_req_per_s': sv['saturation']['req_per_s'],
                   'mean_batch': sv['serving']['mean_batch'],
                   'persistent_launches': {'p50_ms': sv['serving_persistent']['p50_ms'],
                                           'p99_ms': sv['serving_persistent']['p99_ms'],
                                           'map_bins_per_s': sv['serving_persistent']['map_bins_per_s'],
                                           'launches': 'every streamed job as a persistent launch (one workgroup '
                                                       'per CU walking the work items round robin)'},
                   'plain_unbatched': {'saturation_req_per_s': sv['plain_unbatched']['saturation']['req_per_s'],
                                       'p50_ms_same_load': sv['plain_unbatched']['same_offered_load_first_requests'][
                                           'p50_ms'],
                                       'p99_ms_same_load': sv['plain_unbatched']['same_offered_load_first_requests'][
                                           'p99_ms']},
                   'workload': sv['config']['workload']}

    line = {
        'metric': 'map-bins/s (grid pts\u00d7freqs) at \u22641e-4 rel err vs FP64; serving p50/p99 latency',
        'value': value,
        'unit': 'map-bins/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': t_step * 1e3,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': {'mfma_bf16x3': 'bf16x3 (split hi+lo, fp32 accumulate)',
                  'mfma_f16x2': 'f16x2 (CSM fp16 scaled, steering split hi+lo, fp32 accumulate)', 'mfma_f32': 'f32',
                  'plain_f32': 'f32'}[args.form],
        'data': 'synthetic (seeded analytic CSM)',
        'config': {'workload': 'plain vs MFMA beamforming: 256 mics (2 m disk), 8 sources (20 dB spread), '
                               'analytic c64 CSM 513 bins, RectGrid 201x201 z=0.5, classic, r_diag',
                   'num_mics': nm, 'num_grid': num_grid, 'num_freqs': num_freqs,
                   'form': args.form, 'phase_mode': 'reference', 'launch_shape': tiles,
                   'parallelism': 'single GPU'},
        'max_rel_err': err,
        'max_rel_err_vs_fp64_phase': errs['max_rel_err_vs_fp64_phase'],
        'point_rel_err_above_1pct_peak': errs['point_rel_err_above_1pct_peak'],
        'forms_map_bins_per_s': forms,
        'plain_baseline': plain,
        'roofline': roofline,
        'cpu_baseline': cpu,
        'variants': variants,
        'configs': configs,
        'serving': serving,
    }
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
