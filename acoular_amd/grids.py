"""Grid geometries of the beamforming map and the sectors integrated over (``acoular/grids.py``).

Positions, shape, the rectangular ``index``/``indices`` helpers and the ``Sector`` classes whose
``contains`` masks select the grid points ``BeamformerBase.integrate`` and ``integrate`` sum
over (``Grid.subdomain``).  Sector membership is host geometry, as in the reference.
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

    def subdomain(self, sector):
        """Indices of the grid points inside ``sector`` in the grid's shape (``grids.py:187-212``)."""
        return np.where(sector.contains(self.pos).reshape(self.shape))


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

    def _axes(self):
        """(name, lower bound, upper bound, step) of each axis."""
        return (('x', self.x_min, self.x_max, self.increment), ('y', self.y_min, self.y_max, self.increment))

    def _snap(self, *coords, half_even=False):
        """Per-axis step numbers of the grid points nearest to the coordinates (scalars or arrays,
        one per axis): (c - lower bound) / step rounded half up (2D, ``grids.py:388-414``) or half
        to even (3D, Python's ``round``, ``grids.py:566-606``).  A coordinate outside its axis
        range raises ValueError."""
        steps = []
        for (name, lo, hi, inc), c in zip(self._axes(), coords):
            c = np.asarray(c, dtype=float)
            outside = (c < lo) | (c > hi)
            if outside.any():
                msg = f'{name} = {np.ravel(c[outside])[0]:g} lies outside the grid range [{lo:g}, {hi:g}]'
                raise ValueError(msg)
            t = (c - lo) / inc
            steps.append((np.round(t) if half_even else np.floor(t + 0.5)).astype(np.int64))
        return steps

    def index(self, x, y):
        """Indices of the grid point nearest to (x, y) (``grids.py:388-414``)."""
        return tuple(int(i) for i in self._snap(x, y))

    def indices(self, *r):
        """Rectangle (x1, y1, x2, y2) or circle (cx, cy, radius) sub-domain (``grids.py:416-473``).

        A rectangle gives one slice per axis between the points nearest to its corners; a circle
        gives the index arrays of every grid point within the radius (the nearest point to the
        centre when none is)."""
        if len(r) == 4:
            lo = self.index(min(r[0], r[2]), min(r[1], r[3]))
            hi = self.index(max(r[0], r[2]), max(r[1], r[3]))
            return tuple(np.s_[a : b + 1] for a, b in zip(lo, hi))
        if len(r) == 3:
            pos = self.pos
            inside = np.flatnonzero((pos[0] - r[0]) ** 2 + (pos[1] - r[1]) ** 2 <= r[2] ** 2)
            if inside.size == 0:
                return self.index(r[0], r[1])
            return tuple(self._snap(pos[0, inside], pos[1, inside]))
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

    def _axes(self):
        return tuple((name, lo, hi, self._inc(a)) for a, (name, lo, hi) in enumerate(
            (('x', self.x_min, self.x_max), ('y', self.y_min, self.y_max), ('z', self.z_min, self.z_max))))

    def index(self, x, y, z):
        """Indices of the grid point nearest to (x, y, z) (``grids.py:566-606``)."""
        return tuple(int(i) for i in self._snap(x, y, z, half_even=True))

    def indices(self, x1, y1, z1, x2, y2, z2):
        """Box sub-domain: one slice per axis between the points nearest to two opposite corners
        (``grids.py:608-650``)."""
        lo = self.index(min(x1, x2), min(y1, y2), min(z1, z2))
        hi = self.index(max(x1, x2), max(y1, y2), max(z1, z2))
        return tuple(np.s_[a : b + 1] for a, b in zip(lo, hi))


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


class LineGrid(Grid):
    """``num_points`` equally spaced points over ``length`` from ``loc`` along ``direction``
    (``grids.py:834-915``)."""

    _params = ('loc', 'direction', 'length', 'num_points')

    def __init__(self, **kwargs):
        object.__setattr__(self, 'loc', (0.0, 0.0, 0.0))
        object.__setattr__(self, 'direction', (1.0, 0.0, 0.0))
        object.__setattr__(self, 'length', 1.0)
        object.__setattr__(self, 'num_points', 1)
        super().__init__(**kwargs)

    def __setattr__(self, name, value):
        if name in ('loc', 'direction'):
            value = tuple(float(v) for v in value)
        if name == 'length':
            value = float(value)
        if name == 'num_points':
            value = int(value)
        object.__setattr__(self, name, value)

    @property
    def shape(self):
        return self.size

    def _calc_pos(self):
        loc = np.array(self.loc, dtype=float).reshape((3, 1))
        if self.num_points <= 1:
            return loc
        dist = self.length / (self.num_points - 1)
        direc = np.array(self.direction, dtype=float)
        direc = direc / np.linalg.norm(direc)
        steps = np.arange(self.num_points, dtype=float)
        return loc + (direc * dist)[:, None] * steps[None, :]


class MergeGrid(Grid):
    """The distinct points of several grids, in ``numpy.unique`` column order (``grids.py:917-989``)."""

    _params = ('grids',)

    def __init__(self, **kwargs):
        object.__setattr__(self, 'grids', [])
        super().__init__(**kwargs)

    def __setattr__(self, name, value):
        if name == 'grids':
            value = list(value)
        object.__setattr__(self, name, value)

    @property
    def shape(self):
        return self.size

    def _digest_items(self):
        # each grid by its own digest, so a change inside a grid changes the merged positions
        return tuple(self.grids)

    @property
    def subgrids(self) -> np.ndarray:
        """Class name and digest of the grid each point came from, before duplicates are merged
        (``grids.py:969-974``), shape ``(1, total points)``."""
        names = [np.full(g.size, g.__class__.__name__ + g.digest) for g in self.grids]
        return (np.concatenate(names) if names else np.zeros(0, dtype=str))[None, :]

    def _calc_pos(self):
        bpos = np.zeros((3, 0))
        for g in self.grids:
            bpos = np.append(bpos, g.pos, axis=1)
        return np.unique(bpos, axis=1)


# ----------------------------------------------------------------------------------------------
# sectors (grids.py:991-1494)
# ----------------------------------------------------------------------------------------------


def polygon_signed_distance(vx, vy, px, py, small=1e-12) -> np.ndarray:
    """Distance of the points (px, py) to the boundary of the polygon with vertices (vx, vy),
    positive inside, negative outside, 0 within ``small`` of the boundary -- the convention of
    the reference's ``Polygon.is_inside`` (``tools/utils.py:25-180``).  Inside/outside by the
    even-odd crossing rule, the distance as the minimum over the edges."""
    vx = np.asarray(vx, dtype=float)
    vy = np.asarray(vy, dtype=float)
    px = np.atleast_1d(np.asarray(px, dtype=float))
    py = np.atleast_1d(np.asarray(py, dtype=float))
    if vx.shape != vy.shape:
        raise IndexError('x and y must be equally sized.')
    if px.shape != py.shape:
        raise IndexError('x and y have different shapes')
    ax, ay = vx, vy
    bx, by = np.roll(vx, -1), np.roll(vy, -1)
    ex, ey = bx - ax, by - ay
    d2 = np.full(px.shape, np.inf)
    inside = np.zeros(px.shape, dtype=bool)
    for i in range(vx.size):
        if ex[i] == 0.0 and ey[i] == 0.0:  # repeated vertex (a closed vertex list)
            t = np.zeros_like(px)
        else:
            t = np.clip(((px - ax[i]) * ex[i] + (py - ay[i]) * ey[i]) / (ex[i] ** 2 + ey[i] ** 2), 0.0, 1.0)
        d2 = np.minimum(d2, (ax[i] + t * ex[i] - px) ** 2 + (ay[i] + t * ey[i] - py) ** 2)
        # even-odd rule: edges that straddle the horizontal line through the point, crossed to its right
        straddle = (ay[i] > py) != (by[i] > py)
        with np.errstate(divide='ignore', invalid='ignore'):
            xcross = ax[i] + (py - ay[i]) * ex[i] / np.where(ey[i] == 0.0, 1.0, ey[i])
        inside ^= straddle & (px < xcross)
    d = np.sqrt(d2)
    out = np.where(inside, d, -d)
    out[np.abs(out) < small] = 0.0
    return out


def in_hull(p, hull, border=True, tol=0):
    """Points of ``p`` inside the convex hull of ``hull`` (``grids.py:84-128``, a Delaunay
    triangulation of the hull points)."""
    from scipy.spatial import Delaunay

    if not isinstance(hull, Delaunay):
        hull = Delaunay(hull)
    if border:
        return hull.find_simplex(p, tol=tol) >= 0
    return hull.find_simplex(p, tol=tol) > 0


class Sector(HasParams):
    """A sector that contains every grid point (``grids.py:991-1043``)."""

    def contains(self, pos) -> np.ndarray:
        return np.ones(np.asarray(pos).shape[1], dtype=bool)


class SingleSector(Sector):
    """Base of the single sectors (``grids.py:1045-1072``): border inclusion, tolerance and the
    fallback to the nearest point when no point is inside."""

    _params = ('include_border', 'abs_tol', 'default_nearest')

    def __init__(self, **kwargs):
        object.__setattr__(self, 'include_border', True)
        object.__setattr__(self, 'abs_tol', 1e-12)
        object.__setattr__(self, 'default_nearest', True)
        super().__init__(**kwargs)


class RectSector(SingleSector):
    """Rectangle in the x-y plane (``grids.py:1074-1158``)."""

    _params = SingleSector._params + ('x_min', 'x_max', 'y_min', 'y_max')

    def __init__(self, **kwargs):
        for k in ('x_min', 'y_min'):
            object.__setattr__(self, k, -1.0)
        for k in ('x_max', 'y_max'):
            object.__setattr__(self, k, 1.0)
        super().__init__(**kwargs)

    def _box(self):
        return (min(self.x_min, self.x_max), max(self.x_min, self.x_max), min(self.y_min, self.y_max),
                max(self.y_min, self.y_max))

    def _inside(self, pos, lo, hi):
        t = self.abs_tol if self.include_border else -self.abs_tol
        inds = np.ones(pos.shape[1], dtype=bool)
        for ax, (a, b) in enumerate(zip(lo, hi)):
            inds &= (pos[ax, :] - a > -t) & (pos[ax, :] - b < t)
        return inds

    def contains(self, pos):
        pos = np.asarray(pos)
        xmin, xmax, ymin, ymax = self._box()
        inds = self._inside(pos, (xmin, ymin), (xmax, ymax))
        return self._nearest(pos, inds, xmin, xmax, ymin, ymax)

    def _nearest(self, pos, inds, xmin, xmax, ymin, ymax):
        if not inds.any() and self.default_nearest:
            x, y = (xmin + xmax) / 2.0, (ymin + ymax) / 2.0
            inds[np.argmin((pos[0, :] - x) ** 2 + (pos[1, :] - y) ** 2)] = True
        return inds


class RectSector3D(RectSector):
    """Box (``grids.py:1160-1246``); the nearest-point fallback measures in x and y only."""

    _params = RectSector._params + ('z_min', 'z_max')

    def __init__(self, **kwargs):
        object.__setattr__(self, 'z_min', -1.0)
        object.__setattr__(self, 'z_max', 1.0)
        super().__init__(**kwargs)

    def contains(self, pos):
        pos = np.asarray(pos)
        xmin, xmax, ymin, ymax = self._box()
        zmin, zmax = min(self.z_min, self.z_max), max(self.z_min, self.z_max)
        inds = self._inside(pos, (xmin, ymin, zmin), (xmax, ymax, zmax))
        return self._nearest(pos, inds, xmin, xmax, ymin, ymax)


class CircSector(SingleSector):
    """Circle in the x-y plane (``grids.py:1248-1313``)."""

    _params = SingleSector._params + ('x', 'y', 'r')

    def __init__(self, **kwargs):
        object.__setattr__(self, 'x', 0.0)
        object.__setattr__(self, 'y', 0.0)
        object.__setattr__(self, 'r', 1.0)
        super().__init__(**kwargs)

    def contains(self, pos):
        pos = np.asarray(pos)
        dr2 = (pos[0, :] - self.x) ** 2 + (pos[1, :] - self.y) ** 2
        t = self.abs_tol if self.include_border else -self.abs_tol
        inds = dr2 - self.r**2 < t
        if not inds.any() and self.default_nearest:
            inds[np.argmin(dr2)] = True
        return inds


def _polygon_fallback(inds, edges):
    # the reference takes argmin over the two coordinates of the vertices' mean and marks that
    # grid index, 0 or 1 (grids.py:1372-1375, 1430-1433); kept as it is
    dr2 = np.array(edges, dtype=float).reshape(-1, 2).mean(0)
    inds[np.argmin(dr2)] = True
    return inds


class PolySector(SingleSector):
    """Polygon in the x-y plane, vertices ``edges = [x0, y0, x1, y1, ...]`` (``grids.py:1316-1377``)."""

    _params = SingleSector._params + ('edges',)

    def __init__(self, **kwargs):
        object.__setattr__(self, 'edges', [])
        super().__init__(**kwargs)

    def contains(self, pos):
        pos = np.asarray(pos)
        v = np.array(self.edges, dtype=float).reshape(-1, 2)
        dists = polygon_signed_distance(v[:, 0], v[:, 1], pos[0, :], pos[1, :])
        inds = dists >= -self.abs_tol if self.include_border else dists > 0
        if not inds.any() and self.default_nearest:
            inds = _polygon_fallback(inds, self.edges)
        return inds


class ConvexSector(SingleSector):
    """Convex hull of the vertices ``edges`` in the x-y plane (``grids.py:1380-1435``)."""

    _params = SingleSector._params + ('edges',)

    def __init__(self, **kwargs):
        object.__setattr__(self, 'edges', [])
        super().__init__(**kwargs)

    def contains(self, pos):
        pos = np.asarray(pos)
        inds = in_hull(pos[:2, :].T, np.array(self.edges, dtype=float).reshape(-1, 2), border=self.include_border,
                       tol=self.abs_tol)
        if not inds.any() and self.default_nearest:
            inds = _polygon_fallback(inds, self.edges)
        return inds


class MultiSector(Sector):
    """Union of sectors (``grids.py:1438-1494``)."""

    _params = ('sectors',)

    def _digest_items(self):
        return tuple(self.sectors)

    def __init__(self, **kwargs):
        object.__setattr__(self, 'sectors', [])
        super().__init__(**kwargs)

    def contains(self, pos):
        pos = np.asarray(pos)
        inds = np.zeros(pos.shape[1], dtype=bool)
        for sec in self.sectors:
            inds |= sec.contains(pos)
        return inds
