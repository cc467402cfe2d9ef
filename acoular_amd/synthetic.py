"""Seeded synthetic inputs for the benchmark and test workloads: microphone arrays, source
layouts, analytic cross-spectral matrices and multichannel time signals.

Analytic CSMs follow the reference's monopole transfer model (``fastFuncs.py:893-897``):
``a_m = exp(-1j k (r_sm - r_s0)) r_s0 / r_sm`` with ``r_s0`` the source distance to the origin,
``C(f) = sum_s p_s a_s a_s^H``.  Time signals are white-noise monopoles propagated with a 1/r
amplitude and a fractional delay applied in the frequency domain.
"""

from __future__ import annotations

import numpy as np


def disk_array(num: int, diameter: float, seed: int) -> np.ndarray:
    """``num`` microphones uniform in a disk of the given diameter at z=0 (sqrt-radius sampling)."""
    rng = np.random.default_rng(seed)
    r = 0.5 * diameter * np.sqrt(rng.uniform(0, 1, num))
    phi = rng.uniform(0, 2 * np.pi, num)
    return np.stack([r * np.cos(phi), r * np.sin(phi), np.zeros(num)])


def spiral_array(num: int, diameter: float, seed: int, jitter: float = 0.01) -> np.ndarray:
    """Logarithmic-spiral layout (golden-angle arms) plus seeded Gaussian jitter, at z=0."""
    rng = np.random.default_rng(seed)
    n = np.arange(num)
    r = 0.5 * diameter * np.sqrt((n + 0.5) / num)
    phi = n * np.pi * (3 - np.sqrt(5))
    pos = np.stack([r * np.cos(phi), r * np.sin(phi), np.zeros(num)])
    pos[:2] += rng.normal(scale=jitter, size=(2, num))
    return pos


def source_positions(num: int, lo, hi, seed: int) -> np.ndarray:
    """``num`` points uniform in the box ``[lo, hi]`` (3-vectors)."""
    rng = np.random.default_rng(seed)
    lo = np.asarray(lo, dtype=float)
    hi = np.asarray(hi, dtype=float)
    return (lo[:, None] + (hi - lo)[:, None] * rng.uniform(0, 1, (3, num)))


def level_powers(num: int, spread_db: float, seed: int) -> np.ndarray:
    """Source powers ``10**(-L/10)`` with ``L`` uniform in ``[0, spread_db]``."""
    rng = np.random.default_rng(seed)
    return 10.0 ** (-rng.uniform(0, spread_db, num) / 10.0)


def analytic_csm(mpos, spos, powers, freqs, c: float = 343.0, dtype=np.complex64, device=None):
    """``C(f) = sum_s p_s a_s a_s^H`` for every frequency (float64 arithmetic, cast to ``dtype``).

    With ``device`` (a torch device) the matrices are built there and a device tensor returned.
    """
    mpos = np.asarray(mpos, dtype=float)
    spos = np.asarray(spos, dtype=float)
    freqs = np.atleast_1d(np.asarray(freqs, dtype=float))
    rsm = np.sqrt(((spos[:, :, None] - mpos[:, None, :]) ** 2).sum(0))  # [S, M]
    rs0 = np.sqrt((spos**2).sum(0))  # [S]
    p = np.asarray(powers, dtype=float)
    if device is None:
        out = np.empty((len(freqs), mpos.shape[1], mpos.shape[1]), dtype=dtype)
        for i, f in enumerate(freqs):
            k = 2 * np.pi * f / c
            a = np.exp(-1j * k * (rsm - rs0[:, None])) * (rs0[:, None] / rsm)  # [S, M]
            out[i] = np.einsum('s,sm,sn->mn', p, a, a.conj()).astype(dtype)
        return out
    import torch

    def up(x):
        # staged in pinned memory and copied without blocking: the host does not wait for the
        # work already queued on the stream (a pageable copy would)
        return torch.from_numpy(np.ascontiguousarray(x)).pin_memory().to(device, non_blocking=True)

    tdt = {np.complex64: torch.complex64, np.complex128: torch.complex128}[np.dtype(dtype).type]
    rsm_t = up(rsm)
    rs0_t = up(rs0)
    amp = (rs0_t[:, None] / rsm_t) * up(np.sqrt(p))[:, None]  # [S, M]
    dr = rsm_t - rs0_t[:, None]
    out = torch.empty((len(freqs), mpos.shape[1], mpos.shape[1]), dtype=tdt, device=device)
    kk = up(2 * np.pi * freqs / c)
    chunk = 64
    for i0 in range(0, len(freqs), chunk):
        k = kk[i0 : i0 + chunk]
        a = torch.polar(amp[None].expand(len(k), -1, -1), -k[:, None, None] * dr[None])  # [Fc, S, M]
        at = a.transpose(1, 2).contiguous()
        out[i0 : i0 + chunk] = torch.bmm(at, a.conj().contiguous()).to(tdt)
    return out


def point_source_signals(mpos, spos, num_samples: int, fs: float, c: float = 343.0, seed: int = 1,
                         noise_rel: float = 0.0, powers=None, dtype=np.float64) -> np.ndarray:
    """Microphone signals ``[N, M]`` of uncorrelated white-noise monopoles.

    Each source signal is unit-variance (times ``sqrt(power)``) white noise; at microphone m it
    arrives delayed by ``r_sm / c`` (fractional delay as a linear phase, circular) and scaled by
    ``1 / r_sm``.  Independent white noise of standard deviation ``noise_rel`` times the rms of
    the summed source signals at each microphone is added.
    """
    rng = np.random.default_rng(seed)
    mpos = np.asarray(mpos, dtype=float)
    spos = np.asarray(spos, dtype=float)
    ns = spos.shape[1]
    nm = mpos.shape[1]
    p = np.ones(ns) if powers is None else np.asarray(powers, dtype=float)
    rsm = np.sqrt(((spos[:, :, None] - mpos[:, None, :]) ** 2).sum(0))  # [S, M]
    omega = 2 * np.pi * np.fft.rfftfreq(num_samples, 1.0 / fs)
    out = np.zeros((num_samples, nm))
    for s in range(ns):
        sig = rng.standard_normal(num_samples) * np.sqrt(p[s])
        spec = np.fft.rfft(sig)
        for m0 in range(0, nm, 64):
            m1 = min(nm, m0 + 64)
            ph = np.exp(-1j * omega[:, None] * (rsm[s, m0:m1] / c)[None, :]) / rsm[s, m0:m1][None, :]
            out[:, m0:m1] += np.fft.irfft(spec[:, None] * ph, n=num_samples, axis=0)
    if noise_rel > 0:
        rms = np.sqrt(np.mean(out**2, axis=0))
        out += rng.standard_normal(out.shape) * (noise_rel * rms)[None, :]
    return out.astype(dtype)


def point_source_signals_device(mpos, spos, num_samples: int, fs: float, c: float = 343.0, seed: int = 1,
                                noise_rel: float = 0.0, mixing=None, chunk: int = 64):
    """Device (float32 ``[N, M]`` torch tensor) counterpart of ``point_source_signals`` for large
    arrays and long signals: the same propagation model (delay ``r_sm / c`` as a circular linear
    phase, amplitude ``1 / r_sm``), computed with the GPU FFT.

    Source signals are seeded unit-variance white noise; with ``mixing`` (``[S, S]``) they are
    mixed as ``s = mixing @ w`` (correlated sources).  Microphone noise: independent white noise of
    ``noise_rel`` times each microphone's source-signal rms (seeded torch generator on the device).
    """
    import torch

    dev = torch.device('cuda', 0)
    rng = np.random.default_rng(seed)
    mpos = np.asarray(mpos, dtype=float)
    spos = np.asarray(spos, dtype=float)
    ns, nm = spos.shape[1], mpos.shape[1]
    w = rng.standard_normal((ns, num_samples))
    if mixing is not None:
        w = np.asarray(mixing, dtype=float) @ w
    rsm = np.sqrt(((spos[:, :, None] - mpos[:, None, :]) ** 2).sum(0))  # [S, M]
    omega = torch.from_numpy(2 * np.pi * np.fft.rfftfreq(num_samples, 1.0 / fs)).to(dev)
    out = torch.zeros((num_samples, nm), dtype=torch.float32, device=dev)
    for s in range(ns):
        spec = torch.fft.rfft(torch.from_numpy(w[s]).to(dev))
        for m0 in range(0, nm, chunk):
            m1 = min(nm, m0 + chunk)
            r = torch.from_numpy(rsm[s, m0:m1]).to(dev)
            ph = torch.exp(-1j * omega[:, None] * (r / c)[None, :]) / r[None, :]
            out[:, m0:m1] += torch.fft.irfft(spec[:, None] * ph, n=num_samples, dim=0).to(torch.float32)
    if noise_rel > 0:
        gen = torch.Generator(device=dev)
        gen.manual_seed(int(seed) + 7919)
        rms = torch.sqrt(torch.mean(out.double() ** 2, dim=0)).float()
        out += torch.randn(out.shape, generator=gen, device=dev, dtype=torch.float32) * (noise_rel * rms)[None, :]
    return out
