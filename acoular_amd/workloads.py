"""The seeded benchmark configurations of the beamforming path, shared by the benchmark and the
full-size parity tests.

* ``plain_vs_mfma``: 256 mics in a 2 m disk, 8 monopoles with a 20 dB level spread at z = 0.5,
  analytic complex64 CSM of 513 bins (block size 1024 at 51.2 kHz), RectGrid 201 x 201 on
  [-1, 1]^2 at z = 0.5.
* ``fused``: 512 mics in a 2 m disk, 10 s at 51.2 kHz float32 from 16 seeded correlated
  white-noise sources (mixing I + 0.3 N(0, 1)) at z = 1.0 plus 5 % uncorrelated noise;
  RectGrid 128 x 128 on [-1, 1]^2 at z = 1.0; Welch block size 2048 (1025 bins).
* ``large3d``: 1024 mics on a 3 m spiral with seeded jitter, RectGrid3D 64 x 64 x 32 (x, y in
  [-1, 1], z from 0.5, increment 2/63), bins 2..257 of block size 1024 at 51.2 kHz, analytic
  complex64 CSM of 32 seeded 3D sources with a 20 dB spread.
* the serving mix: ``serving.RequestFactory``.
"""

from __future__ import annotations

import numpy as np

from . import synthetic

C = 343.0


def plain_vs_mfma(device=None, seed: int = 2024) -> dict:
    """256 mics, 201 x 201 grid, 513 bins; the CSM is built on ``device`` when given."""
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
    csm = synthetic.analytic_csm(mpos, spos, powers, freqs, device=device)
    return dict(mpos=mpos, spos=spos, powers=powers, freqs=freqs, gpos=gpos, csm=csm, nm=nm)


def fused(seed: int = 3003, num_samples: int = 512000) -> dict:
    """512-mic, 10 s signals synthesised on cuda:0 (float32 ``[N, 512]`` tensor) and the grid."""
    fs, nm, bs = 51200.0, 512, 2048
    mpos = synthetic.disk_array(nm, 2.0, seed)
    spos = synthetic.source_positions(16, [-0.8, -0.8, 1.0], [0.8, 0.8, 1.0], seed + 1)
    mixing = np.eye(16) + 0.3 * np.random.default_rng(seed + 2).standard_normal((16, 16))
    t = synthetic.point_source_signals_device(mpos, spos, num_samples, fs, seed=seed + 3, noise_rel=0.05,
                                              mixing=mixing)
    x = np.linspace(-1.0, 1.0, 128)
    gpos = np.stack([np.repeat(x, 128), np.tile(x, 128), np.full(128 * 128, 1.0)])
    return dict(fs=fs, nm=nm, bs=bs, mpos=mpos, spos=spos, signals=t, gpos=gpos)


def large3d(device=None, seed: int = 4004) -> dict:
    """1024 mics, 64 x 64 x 32 grid, 256 bins; the 2.15 GB CSM is built on ``device`` when given."""
    from .grids import RectGrid3D

    nm = 1024
    mpos = synthetic.spiral_array(nm, 3.0, seed)
    inc = 2.0 / 63
    grid = RectGrid3D(x_min=-1.0, x_max=1.0, y_min=-1.0, y_max=1.0, z_min=0.5, z_max=0.5 + 31 * inc, increment=inc)
    gpos = grid.pos
    spos = synthetic.source_positions(32, [-0.9, -0.9, 0.6], [0.9, 0.9, 1.4], seed + 1)
    freqs = np.arange(2, 258) * 51200.0 / 1024
    csm = synthetic.analytic_csm(mpos, spos, synthetic.level_powers(32, 20.0, seed + 2), freqs, device=device)
    return dict(mpos=mpos, spos=spos, freqs=freqs, gpos=gpos, grid=grid, csm=csm, nm=nm)

