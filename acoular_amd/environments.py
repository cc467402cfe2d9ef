"""Sound propagation environment (``acoular/environments.py``): the default resting medium.

``dist_mat`` and ``Environment.apparent_r`` run on the GPU (``abf_dist_mat``) and return
float64 numpy arrays like the reference.
"""

from __future__ import annotations

import numpy as np

from . import _lib
from ._base import HasParams


def dist_mat(gpos, mpos) -> np.ndarray:
    """``dist_mat`` (``environments.py:65-103``): ``(N, M)`` distances of ``gpos (3, N)`` to
    ``mpos (3, M)``, computed in float64 on the GPU."""
    g = np.ascontiguousarray(gpos, dtype=np.float64)
    m = np.ascontiguousarray(mpos, dtype=np.float64)
    if g.ndim != 2 or m.ndim != 2 or g.shape[0] != 3 or m.shape[0] != 3:
        msg = 'gpos and mpos must have shape (3, N) and (3, M)'
        raise ValueError(msg)
    n, mm = g.shape[1], m.shape[1]
    torch = _lib.torch()
    dg = _lib.to_device(g)
    dm = _lib.to_device(m)
    out = torch.empty((n, mm), dtype=torch.float64, device=dg.device)
    _lib.check(_lib.lib().abf_dist_mat(_lib.ptr(dg), _lib.ptr(dm), n, mm, _lib.ptr(out), _lib.stream_handle()),
               'dist_mat')
    return out.cpu().numpy()


class Environment(HasParams):
    """Resting, homogeneous medium with speed of sound ``c`` (``environments.py:172-253``)."""

    _params = ('c', 'roi')

    def __init__(self, **kwargs):
        self.c = 343.0
        self.roi = None
        super().__init__(**kwargs)

    def apparent_r(self, gpos, mpos=0.0):
        """Distances of ``gpos`` to ``mpos``; a scalar ``mpos`` means the origin and a single
        point collapses the result to shape ``(N,)`` (``environments.py:202-238``)."""
        if np.isscalar(mpos):
            mpos = np.zeros((3, 1))
        rm = dist_mat(np.ascontiguousarray(gpos), np.ascontiguousarray(mpos))
        if rm.shape[1] == 1:
            rm = rm[:, 0]
        return rm
