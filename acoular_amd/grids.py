"""Grid geometries of the beamforming map (``acoular/grids.py``).

Only what the frequency-domain path needs: positions, shape and the rectangular
``index``/``indices`` helpers used by ``BeamformerBase.integrate``.
"""

from __future__ import annotations

import numpy as np

from ._base import HasParams


def _steps(vmin: float, vmax: float, inc: float) -> int:
    """Number of points along one axis (``grids.py:333-345``)."""
    i = abs(inc)
    if i != 0:
        return round((abs(vmax - vmin) + i) / i)
    return 1


class Grid(HasParams):
    """Abstract grid (``grids.py:131-215``): ``size``, ``shape`` and ``pos`` (3, size)."""

    @property
    def pos(self) -> np.ndarray:
        return self._cached('pos', self._calc_pos)

    @property
    def size(self) -> int:
        return self.pos.shape[-1]

    def _calc_pos(self):  # pragma: no cover - abstract
        raise NotImplementedError


class RectGrid(Grid):
    """2D Cartesian grid at constant z (``grids.py:284-474``).  Point index = ix * ny + iy."""

    _params = ('x_min', 'x_max', 'y_min', 'y_max', 'z', 'increment')

    def __init__(self, **kwargs):
        self.x_min, self.x_max, self.y_min, self.y_max = -1.0, 1.0, -1.0, 1.0
        self.z, self.increment = 1.0, 0.1
        super().__init__(**kwargs)

    @property
    def nxsteps(self) -> int:
        return _steps(self.x_min, self.x_max, self.increment)

    @property
    def nysteps(self) -> int:
        return _steps(self.y_min, self.y_max, self.increment)

    @property
    def size(self) -> int:
        return self.nxsteps * self.nysteps

    @property
    def shape(self):
        return (self.nxsteps, self.nysteps)

    @property
    def extent(self):
        return (self.x_min, self.x_max, self.y_min, self.y_max)

    def _calc_pos(self):
        # np.mgrid with a complex step count is linspace; x-major order (grids.py:351-359)
        x = np.linspace(self.x_min, self.x_max, self.nxsteps)
        y = np.linspace(self.y_min, self.y_max, self.nysteps)
        pos = np.empty((3, self.size))
        pos[0] = np.repeat(x, self.nysteps)
        pos[1] = np.tile(y, self.nxsteps)
        pos[2] = self.z
        return pos

    def index(self, x, y):
        """Indices of the grid point nearest to (x, y) (``grids.py:388-414``)."""
        if x < self.x_min or x > self.x_max:
            msg = 'x-value out of range'
            raise ValueError(msg)
        if y < self.y_min or y > self.y_max:
            msg = 'y-value out of range'
            raise ValueError(msg)
        xi = int((x - self.x_min) / self.increment + 0.5)
        yi = int((y - self.y_min) / self.increment + 0.5)
        return xi, yi

    def indices(self, *r):
        """Rectangle (4 values) or circle (3 values) sub-domain (``grids.py:416-473``)."""
        if len(r) == 3:
            xpos = self.pos
            dr2 = (xpos[0, :] - r[0]) ** 2 + (xpos[1, :] - r[1]) ** 2
            inds = dr2 <= r[2] ** 2
            xis, yis = [], []
            for n_p in np.arange(self.size)[inds]:
                xi, yi = self.index(xpos[0, n_p], xpos[1, n_p])
                xis.append(xi)
                yis.append(yi)
            if not (xis and yis):
                return self.index(r[0], r[1])
            return np.array(xis), np.array(yis)
        if len(r) == 4:
            xi1, yi1 = self.index(min(r[0], r[2]), min(r[1], r[3]))
            xi2, yi2 = self.index(max(r[0], r[2]), max(r[1], r[3]))
            return np.s_[xi1 : xi2 + 1], np.s_[yi1 : yi2 + 1]
        msg = 'polygon sub-domains are not part of this build'
        raise NotImplementedError(msg)


class RectGrid3D(RectGrid):
    """3D Cartesian grid (``grids.py:475-665``).  Point index = (ix * ny + iy) * nz + iz."""

    _params = ('x_min', 'x_max', 'y_min', 'y_max', 'z_min', 'z_max', 'increment')

    def __init__(self, **kwargs):
        self.z_min, self.z_max = -1.0, 1.0
        super().__init__(**kwargs)

    def _inc(self, axis):
        inc = self.increment
        return abs(inc) if np.isscalar(inc) else abs(np.asarray(inc, dtype=float)[axis])

    @property
    def nxsteps(self) -> int:
        return _steps(self.x_min, self.x_max, self._inc(0))

    @property
    def nysteps(self) -> int:
        return _steps(self.y_min, self.y_max, self._inc(1))

    @property
    def nzsteps(self) -> int:
        return _steps(self.z_min, self.z_max, self._inc(2))

    @property
    def size(self) -> int:
        return self.nxsteps * self.nysteps * self.nzsteps

    @property
    def shape(self):
        return (self.nxsteps, self.nysteps, self.nzsteps)

    def _calc_pos(self):
        nx, ny, nz = self.shape
        x = np.linspace(self.x_min, self.x_max, nx)
        y = np.linspace(self.y_min, self.y_max, ny)
        z = np.linspace(self.z_min, self.z_max, nz)
        pos = np.empty((3, nx * ny * nz))
        pos[0] = np.repeat(x, ny * nz)
        pos[1] = np.tile(np.repeat(y, nz), nx)
        pos[2] = np.tile(z, nx * ny)
        return pos

    def index(self, x, y, z):
        """Indices of the grid point nearest to (x, y, z) (``grids.py:566-606``)."""
        if x < self.x_min or x > self.x_max:
            msg = f'x-value out of range {x:f} ({self.x_min:f}, {self.x_max:f})'
            raise ValueError(msg)
        if y < self.y_min or y > self.y_max:
            msg = f'y-value out of range {y:f} ({self.y_min:f}, {self.y_max:f})'
            raise ValueError(msg)
        if z < self.z_min or z > self.z_max:
            msg = f'z-value out of range {z:f} ({self.z_min:f}, {self.z_max:f})'
            raise ValueError(msg)
        inc = self.increment
        incx, incy, incz = (inc, inc, inc) if np.isscalar(inc) else tuple(inc)
        xi = round((x - self.x_min) / incx)
        yi = round((y - self.y_min) / incy)
        zi = round((z - self.z_min) / incz)
        return xi, yi, zi

    def indices(self, x1, y1, z1, x2, y2, z2):
        """Box sub-domain (``grids.py:608-650``)."""
        xi1, yi1, zi1 = self.index(min(x1, x2), min(y1, y2), min(z1, z2))
        xi2, yi2, zi2 = self.index(max(x1, x2), max(y1, y2), max(z1, z2))
        return np.s_[xi1 : xi2 + 1], np.s_[yi1 : yi2 + 1], np.s_[zi1 : zi2 + 1]


class ImportGrid(Grid):
    """Arbitrary point set (``grids.py:666-833``); positions given directly as ``pos``."""

    _params = ('gpos',)

    def __init__(self, pos=None, **kwargs):
        self.gpos = np.zeros((3, 0)) if pos is None else np.asarray(pos, dtype=float)
        super().__init__(**kwargs)

    def _calc_pos(self):
        return np.atleast_2d(np.asarray(self.gpos, dtype=float))

    @property
    def shape(self):
        return (self.pos.shape[-1],)
