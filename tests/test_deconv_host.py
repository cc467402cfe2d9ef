"""Host logic of PointSpreadFunction, BeamformerDamas and BeamformerCleansc (no GPU): parameters,
digests, workspace sizes and the index check that runs before any device work (fbeamform.py:948-1492)."""

import warnings

import numpy as np
import pytest

import acoular_amd as aa
from acoular_amd import _lib


def _steer(nm=6):
    pos = np.zeros((3, nm))
    pos[0] = np.arange(nm) * 0.1
    grid = aa.RectGrid(x_min=-0.1, x_max=0.1, y_min=-0.1, y_max=0.1, z=0.5, increment=0.1)
    return aa.SteeringVector(grid=grid, mics=aa.MicGeom(pos_total=pos))


def test_psf_parameters():
    p = aa.PointSpreadFunction(steer=_steer())
    assert p.calcmode == 'single' and p.precision == 'float64' and p.freq == 1.0 and p.grid_indices.size == 0
    p.grid_indices = [1, 4]
    assert p.grid_indices.dtype.kind == 'i'
    with pytest.raises(ValueError):
        p.calcmode = 'all'
    with pytest.raises(ValueError):
        p.precision = 'float16'
    with pytest.raises(TypeError):
        p.steer = None


def test_psf_digest_depends_on_steer_and_precision_only():
    sv = _steer()
    a = aa.PointSpreadFunction(steer=sv, freq=1000.0)
    b = aa.PointSpreadFunction(steer=sv, freq=2000.0, calcmode='full', grid_indices=[2])
    assert a.digest == b.digest
    assert aa.PointSpreadFunction(steer=sv, precision='float32').digest != a.digest
    assert aa.PointSpreadFunction(steer=_steer(7)).digest != a.digest


def test_psf_device_rejects_out_of_range_indices():
    with pytest.raises(IndexError):
        aa.psf_device(_steer(), [1000.0], [0, 9])
    with pytest.raises(IndexError):
        aa.psf_device(_steer(), [1000.0], [-1])


def test_damas_parameters():
    bf = aa.BeamformerDamas(steer=_steer())
    assert bf.n_iter == 100 and bf.damp == 1.0 and bf.psf_precision == 'float64' and bf.calcmode == 'full'
    assert bf.r_diag is True
    with pytest.raises(ValueError):
        bf.psf_precision = 'float16'
    with pytest.raises(ValueError):
        bf.calcmode = 'all'
    d = bf.digest
    bf.calcmode = 'block'
    assert bf.digest == d  # calcmode does not change the result
    bf.n_iter = 50
    assert bf.digest != d


def test_cleansc_parameters():
    bf = aa.BeamformerCleansc(steer=_steer())
    assert bf.n_iter == 0 and bf.damp == 0.6 and bf.stopn == 3 and bf.r_diag is True
    with pytest.raises(ValueError):
        bf.damp = 0.0
    with pytest.raises(ValueError):
        bf.damp = 1.5
    d = bf.digest
    bf.stopn = 4
    assert bf.digest != d
    d = bf.digest
    bf.damp = 0.5
    assert bf.digest != d


def test_cleansc_workspace_and_errors():
    """Workspace: powers [F][J] f64, hh [F][M] c128, hmax [F] f64, stop flags [F] i32, the distance
    and weight table [M][G] (2 x f64) and the per-point sums [G] f64, each on a 256-byte boundary."""
    lib = _lib.lib()

    def up(x):
        return (x + 255) // 256 * 256

    def want(M, G, F, J):
        n = up(F * J * 8)
        n = up(n + F * M * 16)
        n = up(n + F * 8)
        n = up(n + F * 4)
        n = up(n + M * G * 16)
        return up(n + G * 8)

    assert lib.abf_cleansc_workspace_size(64, 1681, 10, 0) == want(64, 1681, 10, 128)
    assert lib.abf_cleansc_workspace_size(64, 1681, 3, 25) == want(64, 1681, 3, 25)
    assert lib.abf_cleansc_workspace_size(0, 10, 10, 0) == 0
    assert lib.abf_cleansc_workspace_size(4, -1, 10, 0) == 0


def test_clean_parameters():
    bf = aa.BeamformerClean(steer=_steer())
    assert bf.n_iter == 100 and bf.damp == 0.6 and bf.psf_precision == 'float64' and bf.calcmode == 'block'
    with pytest.raises(ValueError):
        bf.psf_precision = 'float16'
    with pytest.raises(ValueError):
        bf.calcmode = 'all'
    with pytest.raises(ValueError):
        bf.damp = 0.001
    d = bf.digest
    bf.calcmode = 'single'
    assert bf.digest == d  # calcmode does not change the result
    bf.n_iter = 10
    assert bf.digest != d
    lib = _lib.lib()
    G, M, F = 1681, 64, 10
    n = 4 * 256 + (M * G * 16 + 255) // 256 * 256 + (G * 8 + 255) // 256 * 256
    assert lib.abf_clean_workspace_size(M, G, F) == n
    assert lib.abf_clean_workspace_size(0, G, F) == 0


def test_damas_plus_parameters():
    bf = aa.BeamformerDamasPlus()
    assert (bf.method, bf.alpha, bf.n_iter, bf.unit_mult) == ('NNLS', 0.0, 500, 1e9)
    assert aa.BeamformerDamasPlus(n_iter=7).n_iter == 7
    with pytest.raises(ValueError):
        aa.BeamformerDamasPlus(method='SIMPLEX')
    with pytest.raises(ValueError):
        aa.BeamformerDamasPlus(alpha=1.5)
    d = bf.digest
    bf.method = 'LP'
    assert bf.digest != d
    d = bf.digest
    bf.unit_mult = 1e6
    assert bf.digest != d


@pytest.mark.parametrize('method', ['NNLS', 'LassoLars', 'OMPCV', 'LP'])
def test_damas_plus_solve_recovers_sparse_sources(method):
    """The host solve of every method recovers a sparse non-negative distribution from its own
    blurred map (symmetric PSF for LP, whose optimum is then A^-1 y)."""
    from acoular_amd.fbeamform import damas_plus_solve
    from tests.test_oracle_deconv import _recovery_case

    A, x, y = _recovery_case(method == 'LP')
    with warnings.catch_warnings():
        warnings.simplefilter('ignore')
        got = damas_plus_solve(A, y, method, 0.0, 500, 1e9, r_diag=True)
    assert np.abs(got - x).max() <= 1e-6 * x.max(), (method, got[[6, 18]])


def test_damas_plus_lp_warns_with_r_diag():
    from acoular_amd.fbeamform import damas_plus_solve
    from tests.test_oracle_deconv import _recovery_case

    A, _, y = _recovery_case(True)
    with pytest.warns(Warning, match='Linear programming'):
        damas_plus_solve(A, y, 'LP', 0.0, 500, 1e9, r_diag=True)
