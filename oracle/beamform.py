"""Float64 numpy restatement of Acoular's frequency-domain beamforming kernels.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``): imported by tests, the smoke check
and the CPU-baseline leg of the benchmark, never by the product path.

Citations are to the Acoular source tree (``acoular/...``).  The restatement keeps the
reference's arithmetic where it matters for the 1e-4 parity gate:

* the phase argument ``k * r`` is rounded to float32 before cos/sin are taken in float32
  (``fastFuncs.py:259-261``) when ``phase_mode == 'reference'``; ``'fp64'`` keeps the exact
  float64 phase instead;
* only the upper triangle and the diagonal of the CSM are read (``fastFuncs.py:277-291``);
* the four steering normalisations are applied after the quadratic form
  (``fastFuncs.py:293-310``).

The loops over grid points are vectorised (one complex matrix product per frequency), which
changes only the float64 summation order.
"""

from __future__ import annotations

import numpy as np

STEER_TYPES = {'classic': 1, 'inverse': 2, 'true level': 3, 'true location': 4}


# --------------------------------------------------------------------------------------------
# geometry
# --------------------------------------------------------------------------------------------
def dist_mat(gpos: np.ndarray, mpos: np.ndarray) -> np.ndarray:
    """Euclidean distances ``rm[n, m] = |g_n - m_m|`` (``environments.py:65-103``).

    ``gpos`` is ``(3, N)``, ``mpos`` is ``(3, M)``; returns ``(N, M)`` float64.
    """
    g = np.asarray(gpos, dtype=np.float64)
    m = np.asarray(mpos, dtype=np.float64)
    d = g[:, :, None] - m[:, None, :]
    return np.sqrt(d[0] ** 2 + d[1] ** 2 + d[2] ** 2)


def apparent_r(gpos: np.ndarray, mpos=0.0) -> np.ndarray:
    """``Environment.apparent_r`` (``environments.py:202-238``): scalar ``mpos`` means the origin,
    and a single reference point collapses the result to shape ``(N,)``."""
    if np.isscalar(mpos):
        mpos = np.zeros((3, 1))
    rm = dist_mat(gpos, mpos)
    if rm.shape[1] == 1:
        rm = rm[:, 0]
    return rm


def steer_r0(gpos: np.ndarray, ref) -> np.ndarray:
    """``SteeringVector._get_r0`` (``fbeamform.py:187-193``): a positive scalar ``ref`` is a
    constant reference distance, a 3-vector is a reference point (default: the origin)."""
    gpos = np.asarray(gpos, dtype=np.float64)
    if np.isscalar(ref):
        if ref > 0:
            return np.full((gpos.shape[1],), float(ref))
        return apparent_r(gpos)
    return apparent_r(gpos, np.asarray(ref, dtype=np.float64)[:, None])


def rect_grid_steps(vmin: float, vmax: float, increment: float) -> int:
    """``RectGrid._get_nxsteps`` (``grids.py:333-345``)."""
    i = abs(increment)
    if i != 0:
        return round((abs(vmax - vmin) + i) / i)
    return 1


def rect_grid_pos(x_min, x_max, y_min, y_max, z, increment) -> np.ndarray:
    """``RectGrid._get_pos`` (``grids.py:351-359``): x-major order, index = ix * ny + iy."""
    nx = rect_grid_steps(x_min, x_max, increment)
    ny = rect_grid_steps(y_min, y_max, increment)
    x = np.linspace(x_min, x_max, nx)
    y = np.linspace(y_min, y_max, ny)
    pos = np.empty((3, nx * ny))
    pos[0] = np.repeat(x, ny)
    pos[1] = np.tile(y, nx)
    pos[2] = z
    return pos


def rect_grid3d_pos(x_min, x_max, y_min, y_max, z_min, z_max, increment) -> np.ndarray:
    """``RectGrid3D._get_pos`` (``grids.py:553-561``): index = (ix * ny + iy) * nz + iz."""
    inc = np.broadcast_to(np.asarray(increment, dtype=float), (3,))
    nx = rect_grid_steps(x_min, x_max, inc[0])
    ny = rect_grid_steps(y_min, y_max, inc[1])
    nz = rect_grid_steps(z_min, z_max, inc[2])
    x = np.linspace(x_min, x_max, nx)
    y = np.linspace(y_min, y_max, ny)
    z = np.linspace(z_min, z_max, nz)
    pos = np.empty((3, nx * ny * nz))
    pos[0] = np.repeat(x, ny * nz)
    pos[1] = np.tile(np.repeat(y, nz), nx)
    pos[2] = np.tile(z, nx * ny)
    return pos


# --------------------------------------------------------------------------------------------
# steering vectors
# --------------------------------------------------------------------------------------------
def _unit_phasor(phase: np.ndarray, phase_mode: str) -> np.ndarray:
    """``cos(p) - 1j*sin(p)`` with the reference's float32 rounding of ``p`` (``fastFuncs.py:260-261``)."""
    if phase_mode == 'reference':
        # float32 argument; cos/sin of it rounded to float32 (a correctly rounded cosf/sinf,
        # as the libm calls numba emits for float32 np.cos/np.sin)
        p = phase.astype(np.float32).astype(np.float64)
        c = np.cos(p).astype(np.float32).astype(np.float64)
        s = np.sin(p).astype(np.float32).astype(np.float64)
        return c - 1j * s
    if phase_mode == 'fp64':
        return np.cos(phase) - 1j * np.sin(phase)
    msg = f'unknown phase_mode {phase_mode!r}'
    raise ValueError(msg)


def calc_transfer(r0: np.ndarray, rm: np.ndarray, k: float, phase_mode: str = 'reference') -> np.ndarray:
    """``calcTransfer`` / ``_transferCoreFunc`` (``fastFuncs.py:859-897``):
    ``a[g, m] = exp(-1j * f32(k * (rm - r0))) * r0 / rm``."""
    r0 = np.asarray(r0, dtype=np.float64)
    rm = np.atleast_2d(np.asarray(rm, dtype=np.float64))
    k = float(np.asarray(k).ravel()[0])
    return _unit_phasor(k * (rm - r0[:, None]), phase_mode) * (r0[:, None] / rm)


def steer_vector(transfer: np.ndarray, steer_type: str) -> np.ndarray:
    """``SteeringVector._steer_funcs_freq`` (``fbeamform.py:145-153``)."""
    x = transfer
    nm = x.shape[-1]
    if steer_type == 'classic':
        return x / np.abs(x) / nm
    if steer_type == 'inverse':
        return 1.0 / x.conj() / nm
    s = np.einsum('ij,ij->i', x, x.conj())[:, None]
    if steer_type == 'true level':
        return x / s
    if steer_type == 'true location':
        return x / np.sqrt(s * nm)
    raise KeyError(steer_type)


def _steer_weights(rm: np.ndarray, steer_type: int):
    """Per-microphone weight and ``helpNormalize`` sum (``fastFuncs.py:262-273``)."""
    if steer_type == 2:
        return rm, np.sum(rm * rm, axis=1)
    if steer_type in (3, 4):
        return 1.0 / rm, np.sum(1.0 / (rm * rm), axis=1)
    return None, None


def _normalise(q, r0, help_norm, nmics, steer_type, norm):
    """Steer-type specific normalisation (``fastFuncs.py:293-310``)."""
    with np.errstate(divide='ignore', invalid='ignore'):
        if steer_type == 1:
            return q / (nmics * nmics) * norm, np.full_like(q, 1.0 / nmics)
        if steer_type == 2:
            nf2 = (nmics * r0) ** 2
            return q / nf2 * norm, help_norm / nf2
        if steer_type == 3:
            nf = r0 * help_norm
            return q / (nf * nf) * norm, 1.0 / (r0 * r0) / help_norm
        if steer_type == 4:
            return q / (nmics * help_norm) * norm, np.full_like(q, 1.0 / nmics)
    raise KeyError(steer_type)


def steering_phasors(rm: np.ndarray, k: float, steer_type: int, phase_mode: str = 'reference'):
    """The per-grid-point vector ``v`` used inside the beamformer (before normalisation) and
    its ``helpNormalize`` sum (``fastFuncs.py:256-273``)."""
    rm = np.atleast_2d(np.asarray(rm, dtype=np.float64))
    v = _unit_phasor(k * rm, phase_mode)
    w, help_norm = _steer_weights(rm, steer_type)
    if w is not None:
        v = v * w
    return v, help_norm


def quad_form_upper(csm: np.ndarray, v: np.ndarray, r_diag: bool) -> np.ndarray:
    """``q = 2 Re sum_{i<j} C_ij conj(v_i) v_j (+ sum_i Re(C_ii) |v_i|^2)`` reading only the
    upper triangle and diagonal of ``csm`` (``fastFuncs.py:277-291``)."""
    c = np.asarray(csm, dtype=np.complex128)
    cu = np.triu(c, 1)
    w = v @ cu.T  # w[g, i] = sum_j C_ij v_j, j > i
    q = 2.0 * np.real(np.sum(np.conj(v) * w, axis=1))
    if not r_diag:
        q = q + np.abs(v) ** 2 @ np.real(np.diag(c))
    return q


def freq_beamformer_full_csm(csm, r0, rm, k, norm, r_diag, steer_type, phase_mode='reference'):
    """``_freqBeamformer_FullCSM`` (``fastFuncs.py:238-311``) for one frequency.

    Returns ``(result[G], normalizeSteer[G])`` in float64.
    """
    if isinstance(steer_type, str):
        steer_type = STEER_TYPES[steer_type]
    rm = np.atleast_2d(np.asarray(rm, dtype=np.float64))
    r0 = np.asarray(r0, dtype=np.float64)
    k = float(np.asarray(k).ravel()[0])
    nmics = np.asarray(csm).shape[0]
    v, help_norm = steering_phasors(rm, k, steer_type, phase_mode)
    q = quad_form_upper(csm, v, r_diag)
    return _normalise(q, r0, help_norm, nmics, steer_type, norm)


def freq_beamformer_eig(eigval, eigvec, r0, rm, k, norm, r_diag, steer_type, phase_mode='reference'):
    """``_freqBeamformer_EigValues`` (``fastFuncs.py:335-411``)."""
    if isinstance(steer_type, str):
        steer_type = STEER_TYPES[steer_type]
    rm = np.atleast_2d(np.asarray(rm, dtype=np.float64))
    r0 = np.asarray(r0, dtype=np.float64)
    k = float(np.asarray(k).ravel()[0])
    eigvec = np.asarray(eigvec, dtype=np.complex128)
    eigval = np.asarray(eigval, dtype=np.float64)
    nmics = eigvec.shape[0]
    v, help_norm = steering_phasors(rm, k, steer_type, phase_mode)
    t = v @ np.conj(eigvec)  # t[g, e] = sum_m conj(V_me) v_m
    full = np.abs(t) ** 2
    if r_diag:
        diag = np.abs(v) ** 2 @ (np.abs(eigvec) ** 2)
        q = (full - diag) @ eigval
    else:
        q = full @ eigval
    return _normalise(q, r0, help_norm, nmics, steer_type, norm)


def freq_beamformer_custom(csm, steer, norm, r_diag):
    """``_freqBeamformer_SpecificSteerVec_FullCSM`` / ``_CsmRemovedDiag``
    (``fastFuncs.py:414-476``): explicit steering vectors ``steer[G, M]``."""
    h = np.asarray(steer, dtype=np.complex128)
    q = quad_form_upper(csm, h, r_diag)
    return q * norm, np.real(np.sum(h * np.conj(h), axis=1))


def freq_beamformer_eig_custom(eigval, eigvec, steer, norm, r_diag):
    """``_freqBeamformer_EigValProb_SpecificSteerVec_*`` (``fastFuncs.py:479-559``)."""
    h = np.asarray(steer, dtype=np.complex128)
    eigvec = np.asarray(eigvec, dtype=np.complex128)
    t = h @ np.conj(eigvec)
    full = np.abs(t) ** 2
    if r_diag:
        full = full - np.abs(h) ** 2 @ (np.abs(eigvec) ** 2)
    q = full @ np.asarray(eigval, dtype=np.float64)
    return q * norm, np.real(np.sum(h * np.conj(h), axis=1))


def beamformer_freq(steer_type, r_diag, norm, steer_tuple, csm_tuple, phase_mode='reference'):
    """Dispatch of ``beamformerFreq`` (``fastFuncs.py:61-214``)."""
    eig = isinstance(csm_tuple, tuple)
    if steer_type == 'custom':
        if eig:
            return freq_beamformer_eig_custom(csm_tuple[0], csm_tuple[1], steer_tuple, norm, r_diag)
        return freq_beamformer_custom(csm_tuple, steer_tuple, norm, r_diag)
    r0, rm, k = steer_tuple
    st = STEER_TYPES[steer_type]
    if eig:
        return freq_beamformer_eig(csm_tuple[0], csm_tuple[1], r0, rm, k, norm, r_diag, st, phase_mode)
    return freq_beamformer_full_csm(csm_tuple, r0, rm, k, norm, r_diag, st, phase_mode)


def sig_loss_norm(r_diag: bool, r_diag_norm: float, nmics: int) -> float:
    """``BeamformerBase.sig_loss_norm`` (``fbeamform.py:438-451``)."""
    if not r_diag:
        return 1.0
    if r_diag_norm == 0.0:
        return float(nmics) / (float(nmics) - 1.0)
    return r_diag_norm


def beamformer_map(csm, freqs, gpos, mpos, steer_type='true level', r_diag=True, r_diag_norm=0.0,
                   c=343.0, ref=(0.0, 0.0, 0.0), ind=None, phase_mode='reference'):
    """``BeamformerBase._calc`` (``fbeamform.py:475-510``) over the frequency indices ``ind``:
    the complex128 cast of each CSM slice, the signal-loss normalisation and the clipping of
    negative values when ``r_diag``.  Returns ``[len(ind), G]`` float64."""
    csm = np.asarray(csm)
    freqs = np.asarray(freqs, dtype=np.float64)
    if ind is None:
        ind = np.arange(csm.shape[0])
    nmics = csm.shape[1]
    r0 = steer_r0(gpos, ref if np.isscalar(ref) else np.asarray(ref, dtype=float))
    rm = dist_mat(gpos, mpos)
    norm = sig_loss_norm(r_diag, r_diag_norm, nmics)
    out = np.zeros((len(ind), rm.shape[0]))
    for n, i in enumerate(ind):
        k = 2 * np.pi * freqs[i] / c
        res = freq_beamformer_full_csm(np.asarray(csm[i], dtype=np.complex128), r0, rm, k, norm, r_diag,
                                       steer_type, phase_mode)[0]
        if r_diag:
            res[np.sign(res) < 0] = 0.0
        out[n] = res
    return out


def l_p(x):
    """``L_p`` (``fbeamform.py:2412-2432``)."""
    x = np.asarray(x)
    return 10 * np.log10(np.maximum(x / 4e-10, np.finfo(x.dtype).eps))
