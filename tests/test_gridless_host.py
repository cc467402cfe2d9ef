"""Host logic of the adaptive-grid beamformers (``BeamformerAdaptiveGrid``,
``BeamformerGridlessOrth``, fbeamform.py:2242-2410): parameters, the eigenvector selection,
the result size, the fixed diagonal normalisation and the sector check.  No device calls."""

import numpy as np
import pytest

import acoular_amd as aa


def test_parameters_and_eigenvector_selection():
    bf = aa.BeamformerGridlessOrth()
    assert bf.n == 1 and list(bf.eva_list) == [-1]
    assert bf.bounds == [(-1.0, 1.0), (-1.0, 1.0), (0.01, 1.0)] and bf.shgo == {}
    bf.n = 3
    assert list(bf.eva_list) == [-1, -2, -3]
    bf.eva_list = [0, 5]
    assert list(bf.eva_list) == [0, 5] and bf.n == 3
    assert bf.r_diag_norm == 1.0
    with pytest.raises(ValueError):
        bf.r_diag_norm = 0.0
    with pytest.raises(ValueError):
        bf.bounds = [(-1, 1), (-1, 1)]


def test_digest_follows_search_parameters():
    bf = aa.BeamformerGridlessOrth()
    d = bf.digest
    bf.shgo = {'n': 16}
    d2 = bf.digest
    assert d2 != d
    bf.bounds = [(-0.5, 0.5), (-0.5, 0.5), (0.1, 1.0)]
    assert bf.digest != d2


def test_size_shape_and_sector_check():
    freqs = np.array([1000.0, 2000.0, 3000.0])
    csm = np.zeros((3, 4, 4), dtype=np.complex128)
    csm[:] = np.eye(4)
    bf = aa.BeamformerGridlessOrth(freq_data=aa.PowerSpectraImport(csm=csm, frequencies=freqs), n=2)
    assert bf.size == 6 and bf.shape == (6,)
    assert bf.pos is None
    with pytest.raises(NotImplementedError):
        bf.integrate([0.0, 0.0, 0.1])
    assert issubclass(aa.BeamformerGridlessOrth, aa.Grid)
