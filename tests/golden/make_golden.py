"""Writes tests/golden/oracle_small.npz: small input/output vectors of the CPU oracle (the float64
restatement of Acoular's algorithms, pinned by the reference's doctest known-answer values and
by closed forms in tests/test_oracle.py), for regression of the oracle itself and for GPU parity
without recomputing the oracle.

    python tests/golden/make_golden.py

Contents (seeded, deterministic):
  bf_*      12 microphones (0.6 m disk), 6 x 6 grid at z = 0.4, 3 bins of an analytic CSM of 2
            sources; maps [4 steer types][r_diag 0/1][3][36] and steer normalisation, phase mode
            'reference'
  welch_*   8 channels x 4096 samples of white noise sources at fs 51200, block size 256, Hanning,
            50 % overlap: the CSM [129][8][8]
  tr_*      transfer functions of 10 grid points x 6 microphones at one wavenumber
  kat_*     the reference's own doctest values of PowerSpectra.synthetic_ev (spectra.py:570-578)
"""

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from acoular_amd import synthetic  # noqa: E402
from oracle import beamform as B  # noqa: E402
from oracle import spectra as S  # noqa: E402

STEERS = ['classic', 'inverse', 'true level', 'true location']
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'oracle_small.npz')


def build():
    d = {}
    mpos = synthetic.disk_array(12, 0.6, 71)
    spos = synthetic.source_positions(2, [-0.1, -0.1, 0.4], [0.1, 0.1, 0.4], 72)
    freqs = np.array([1500.0, 6000.0, 14000.0])
    csm = synthetic.analytic_csm(mpos, spos, [1.0, 0.3], freqs, dtype=np.complex128)
    x = np.linspace(-0.15, 0.15, 6)
    gpos = np.stack([np.repeat(x, 6), np.tile(x, 6), np.full(36, 0.4)])
    maps = np.zeros((4, 2, 3, 36))
    norms = np.zeros((4, 3, 36))
    r0 = B.steer_r0(gpos, np.zeros(3))
    rm = B.dist_mat(gpos, mpos)
    for i, st in enumerate(STEERS):
        for rd in (0, 1):
            maps[i, rd] = B.beamformer_map(csm, freqs, gpos, mpos, st, bool(rd))
        for f in range(3):
            norms[i, f] = B.freq_beamformer_full_csm(csm[f], r0, rm, 2 * np.pi * freqs[f] / 343.0, 1.0, False, st)[1]
    d.update(bf_mpos=mpos, bf_gpos=gpos, bf_freqs=freqs, bf_csm=csm, bf_maps=maps, bf_norms=norms)
    sig = synthetic.point_source_signals(synthetic.disk_array(8, 0.5, 73),
                                         synthetic.source_positions(2, [-0.1, -0.1, 0.5], [0.1, 0.1, 0.5], 74),
                                         4096, 51200.0, seed=75, noise_rel=0.1)
    d.update(welch_time=sig.astype(np.float32),
             welch_csm=S.welch_csm(sig.astype(np.float32), 256, 'Hanning', '50%'))
    tg = gpos[:, :10]
    tm = mpos[:, :6]
    k = 2 * np.pi * 9000.0 / 343.0
    d.update(tr_gpos=tg, tr_mpos=tm, tr_k=np.array([k]),
             tr_transfer=B.calc_transfer(B.steer_r0(tg, np.zeros(3)), B.dist_mat(tg, tm), k))
    d.update(kat_ev5=np.array([0.00048803, 0.0010141, 0.00234248, 0.00457097]),
             kat_ev0=np.array([0.00022468, 0.0004589, 0.00088059, 0.00245989]))
    return d


if __name__ == '__main__':
    np.savez_compressed(OUT, **build())
    print('wrote', OUT, os.path.getsize(OUT), 'bytes')
