"""Host-side logic of acoular_amd and the C-ABI library surface (no GPU compute)."""

import ctypes

import numpy as np
import pytest

import acoular_amd as aa
from acoular_amd import _lib
from oracle import beamform as B
from oracle import spectra as S


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = _lib.header_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_lib._SIGS) == set(names)
    assert b'gfx950' in _lib.lib().abf_version()


def test_struct_layout_matches_header():
    # abf_geometry: 2 pointers, int64, 2 int32, 4 doubles
    assert ctypes.sizeof(_lib.AbfGeometry) == 8 * 2 + 8 + 8 + 8 * 4
    assert _lib.AbfGridJob.out.offset % 8 == 0


def test_workspace_query_without_gpu():
    """Workspace sizing is host arithmetic and runs without a device."""
    j = _lib.AbfGridJob()
    j.geom.num_grid = 40401
    j.geom.num_mics = 256
    j.num_freqs = 513
    j.steer_type = 1
    j.form = _lib.FORMS['mfma_bf16x3']
    n = _lib.lib().abf_beamform_workspace_size(ctypes.byref(j))
    # upper 32x32 blocks of a 256-mic CSM: 36 blocks x 2 tiles x 4 KiB per frequency
    assert n >= 513 * 36 * 8192
    j.form = _lib.FORMS['plain_f32']
    assert _lib.lib().abf_beamform_workspace_size(ctypes.byref(j)) >= 513 * 256 * 256 * 8
    j.steer_type = 9
    assert _lib.lib().abf_beamform_workspace_size(ctypes.byref(j)) == 0
    assert b'steer_type' in _lib.lib().abf_last_error()


def test_invalid_arguments_rejected_before_any_device_work():
    lib = _lib.lib()
    with pytest.raises(ValueError):
        _lib.check(lib.abf_transfer(None, None, 4, 3, 1.0, 0, 7, 1, None, None))
    with pytest.raises(NotImplementedError):
        _lib.check(lib.abf_welch_csm(None, 0, 2, 32768, 100, 0, None, 1.0, ctypes.c_void_p(8), 1,
                                     ctypes.c_void_p(8), 1 << 30, None))


def test_grids_match_oracle():
    g = aa.RectGrid(x_min=-0.2, x_max=0.2, y_min=-0.2, y_max=0.2, z=0.3, increment=0.01)
    assert g.shape == (41, 41) and g.size == 1681
    np.testing.assert_array_equal(g.pos, B.rect_grid_pos(-0.2, 0.2, -0.2, 0.2, 0.3, 0.01))
    g3 = aa.RectGrid3D(increment=2)
    np.testing.assert_array_equal(g3.pos, B.rect_grid3d_pos(-1, 1, -1, 1, -1, 1, 2))
    assert g3.index(0.5, 0.5, 1.0) == (1, 1, 1)  # grids.py doctest
    g.increment = 0.1  # cache follows parameter changes
    assert g.size == 25 and g.pos.shape == (3, 25)
    ig = aa.ImportGrid(pos=np.array([[0.0], [0.0], [1.0]]))
    assert ig.size == 1 and ig.shape == (1,)


def test_micgeom_invalid_channels():
    m = aa.MicGeom(pos_total=np.arange(15.0).reshape(3, 5), invalid_channels=[1, 3])
    assert m.num_mics == 3
    np.testing.assert_array_equal(m.pos, np.arange(15.0).reshape(3, 5)[:, [0, 2, 4]])
    with pytest.raises(ValueError):
        aa.MicGeom(pos_total=np.zeros((2, 4)))


def test_power_spectra_host_quantities():
    ts = aa.TimeSamples(data=np.zeros((1000, 4)), sample_freq=51200)
    ps = aa.PowerSpectra(source=ts, block_size=128, window='Blackman')
    np.testing.assert_array_equal(ps.fftfreq(), S.fftfreq(128, 51200))
    assert ps.num_blocks == 7.8125
    np.testing.assert_array_equal(ps.block_starts(), S.block_starts(1000, 128, 'None'))
    ps.overlap = '75%'
    np.testing.assert_array_equal(ps.block_starts(), S.block_starts(1000, 128, '75%'))
    with pytest.raises(ValueError):
        ps.block_size = 100
    with pytest.raises(ValueError):
        ps.window = 'Kaiser'
    assert list(ps.indices) == list(range(1, 64))


def test_power_spectra_import_checks_shape():
    with pytest.raises(ValueError):
        aa.PowerSpectraImport(csm=np.zeros((2, 3, 4)), frequencies=100.0)
    p = aa.PowerSpectraImport(csm=np.zeros((1, 5, 5), complex), frequencies=2000)
    assert p.num_channels == 5
    np.testing.assert_array_equal(p.fftfreq(), [2000.0])


def test_steering_vector_params():
    sv = aa.SteeringVector()
    assert sv.steer_type == 'true level'
    with pytest.raises(ValueError):
        sv.steer_type = 'foo'
    sv.ref = 2.5
    assert sv.ref == 2.5
    with pytest.raises(TypeError):
        sv.ref = [1, 2]


def test_beamformer_params_and_norm():
    mics = aa.MicGeom(pos_total=np.random.default_rng(0).normal(size=(3, 5)))
    ps = aa.PowerSpectraImport(csm=np.zeros((1, 5, 5), complex), frequencies=1000.0)
    bf = aa.BeamformerBase(freq_data=ps, steer=aa.SteeringVector(mics=mics, grid=aa.RectGrid()))
    assert bf.sig_loss_norm() == 5 / 4
    bf.r_diag_norm = 2.0
    assert bf.sig_loss_norm() == 2.0
    bf.r_diag = False
    assert bf.sig_loss_norm() == 1.0
    assert bf._beamformer_params()[0] == 'true level'

    class MySteer(aa.SteeringVector):
        pass

    bf.steer = MySteer(mics=mics, grid=aa.RectGrid())
    assert bf._beamformer_params()[0] == 'custom'
    with pytest.raises(ValueError):
        bf.form = 'tensor'


def test_device_calls_fail_loudly_without_gpu():
    import torch

    if torch.cuda.is_available():
        pytest.skip('GPU present')
    with pytest.raises(RuntimeError):
        aa.dist_mat(np.zeros((3, 2)), np.zeros((3, 2)))


def test_fastfuncs_type_checks():
    """The drop-in rejects what numba's typed signatures reject, before touching the device."""
    with pytest.raises(TypeError):
        aa.calcCSM(np.zeros((2, 3, 3), dtype=np.float64), np.zeros((2, 3)))
    with pytest.raises(TypeError):
        aa.calcCSM(np.zeros((2, 3, 3), dtype=np.complex128), np.zeros((2, 3), dtype=np.complex64))
    with pytest.raises(KeyError):
        aa.beamformerFreq('nonsense', True, 1.0, (np.zeros(2), np.zeros((2, 3)), 1.0), np.zeros((3, 3), complex))


def test_welch_workspace_holds_the_split_partials_only_for_small_arrays():
    """The CSM update of a small array (few tiles per bin) is split over up to 16 slices of a
    pass's blocks whose partial upper triangles the workspace holds; a large array is not split.
    Host arithmetic, no device."""
    lib = _lib.lib()

    def base(m, bs, bpp):  # twiddles + the spectra of one pass, 256-byte aligned
        def a(n):
            return (n + 255) // 256 * 256

        return a(bs // 2 * 8) + a((bs // 2 + 1) * bpp * m * 8)

    # 64 mics, 65 bins, 799 blocks per pass: 65 waves of 64 x 64 tiles -> 16 slices
    got = lib.abf_welch_csm_workspace_size(64, 128, 799)
    assert got == base(64, 128, 799) + (16 * 65 * 64 * 64 * 8 + 255) // 256 * 256
    # slices keep at least 32 blocks: 100 blocks per pass allow 2
    assert lib.abf_welch_csm_workspace_size(64, 128, 100) == base(64, 128, 100) + (2 * 65 * 64 * 64 * 8 + 255) // 256 * 256
    # 512 mics x 1025 bins fill the GPU with tiles alone: no partials
    assert lib.abf_welch_csm_workspace_size(512, 2048, 125) == base(512, 2048, 125)


def test_calc_csm_rejects_more_mics_than_a_grid_row_holds():
    """calcCSM takes its rows from the grid's second dimension (at most 65535); larger arrays are
    refused before any device work."""
    rc = _lib.lib().abf_calc_csm(ctypes.c_void_p(8), ctypes.c_void_p(8), 1, 70000, 0, None)
    assert rc == _lib.ABF_ERR_UNSUPPORTED


def test_rect_grid_index_and_indices_against_brute_force():
    """RectGrid.index is the nearest grid point (brute force over every point, off the midway
    ties); a circle sub-domain holds exactly the points within the radius; a rectangle's slices
    span the points nearest to its corners; coordinates outside the grid raise ValueError."""
    g = aa.RectGrid(x_min=-0.3, x_max=0.5, y_min=-0.2, y_max=0.2, z=0.4, increment=0.05)
    nx, ny = g.shape
    rng = np.random.default_rng(11)
    for x, y in rng.uniform([-0.3, -0.2], [0.5, 0.2], (200, 2)):
        d2 = (g.pos[0] - x) ** 2 + (g.pos[1] - y) ** 2
        assert g.index(x, y) == divmod(int(np.argmin(d2)), ny)
    for cx, cy, rad in ((0.1, 0.0, 0.12), (-0.3, -0.2, 0.07), (0.49, 0.19, 0.2), (0.12, 0.03, 0.001)):
        got = g.indices(cx, cy, rad)
        flat = np.flatnonzero((g.pos[0] - cx) ** 2 + (g.pos[1] - cy) ** 2 <= rad**2)
        if flat.size == 0:
            assert got == g.index(cx, cy)
            continue
        ix, iy = got
        assert ix.dtype.kind == 'i' and sorted(zip(ix.tolist(), iy.tolist())) == sorted(divmod(int(f), ny) for f in flat)
    sx, sy = g.indices(0.2, -0.1, -0.07, 0.13)
    assert (sx, sy) == (np.s_[g.index(-0.07, 0.0)[0]: g.index(0.2, 0.0)[0] + 1],
                        np.s_[g.index(0.0, -0.1)[1]: g.index(0.0, 0.13)[1] + 1])
    sub = g.pos.reshape(3, nx, ny)[:, sx, sy]
    assert sub[0].min() >= -0.07 - 0.025 and sub[0].max() <= 0.2 + 0.025
    assert sub[1].min() >= -0.1 - 0.025 and sub[1].max() <= 0.13 + 0.025
    for bad in ((0.51, 0.0), (0.0, -0.21)):
        with pytest.raises(ValueError):
            g.index(*bad)
    g3 = aa.RectGrid3D(x_min=0, x_max=1, y_min=0, y_max=2, z_min=-1, z_max=1, increment=0.5)
    assert g3.index(0.25, 1.75, 0.0) == (0, 4, 2)  # half to even on x (0.5 -> 0), y (3.5 -> 4)
    assert g3.indices(1.0, 0.0, 1.0, 0.0, 2.0, -1.0) == (np.s_[0:3], np.s_[0:5], np.s_[0:5])
    with pytest.raises(ValueError):
        g3.index(0.0, 0.0, 1.5)
