"""Rowland-torus geometry and element placement (set-up, host float64).

The interface is that of ``marxs/design/rowland.py`` (class and method
names, arguments, element counts and order); the geometry is written here
from the torus itself:

* ``RowlandTorus`` (``78-366``): in its local frame a core circle of radius
  ``R`` in the x-z plane with a tube of radius ``r`` around it (symmetry
  axis y); ``Q(x) = (|x|^2 + R^2 - r^2)^2 - 4 R^2 (x_x^2 + x_z^2)`` is zero
  on the surface.  Points and normals from (tube angle, azimuth) follow from
  the core-circle construction; the normal at a given point is the
  normalised gradient of ``Q``; ``solve_quartic`` intersects a line with the
  torus: the reference runs ``scipy.optimize.root`` from
  ``k0 = 0.1 max(|r - R|, 0.01 r)``, here the quartic in the line parameter
  ``k`` is expanded into its polynomial coefficients, all roots are found
  with a companion-matrix eigen solve, the real root nearest ``k0`` is taken
  and polished with Newton steps.
* ``design_tilted_torus`` (``369-422``): the circle through the focus and a
  tilted point at distance f.
* ``ElementsOnTorus``/``RectangularGrid``/``CircularMeshGrid``/
  ``GratingArrayStructure`` (``547-888``): element centres from a y-z layout
  pushed along ``optimize_axis`` onto the torus; normals ``-torus.normal``;
  in-plane axis parallel to the torus symmetry axis.  The layouts must give
  the reference's element counts and positions (its tests,
  ``design/tests/test_design.py:209-238,441-456``), so their rules (rows of
  centred elements, ring counts rounded up, arc counts rounded down, equal
  gaps on an arc) are the reference's.
"""
import numpy as np

from .base import MarxsElement
from .geometry import Geometry
from .mathutils import e2h, h2e, axangle2mat, compose
from .simulator import ParallelCalculated

__all__ = ['anglediff', 'RowlandTorus', 'design_tilted_torus',
           'ElementsOnTorus', 'RectangularGrid', 'CircularMeshGrid',
           'GratingArrayStructure']


def anglediff(phi):
    """Angle range covered by ``phi = [phi0, phi1]`` (wraps at 2 pi)."""
    d = phi[1] - phi[0]
    if d < 0. or d > 2. * np.pi:
        d = d % (2. * np.pi)
    return d


class RowlandTorus(MarxsElement, Geometry):
    def __init__(self, R, r, **kwargs):
        self.R = R
        self.r = r
        Geometry.__init__(self, kwargs)
        MarxsElement.__init__(self, **kwargs)

    def _to_local(self, xyz):
        inv = np.linalg.inv(self.pos4d)
        return h2e(np.einsum('...ij,...j', inv, e2h(xyz, 1)))

    def quartic(self, xyz, transform=True):
        xyz = np.asarray(xyz, dtype=float)
        if xyz.shape[-1] != 3:
            raise ValueError('Input coordinates must be defined in Euclidean space.')
        if transform:
            xyz = self._to_local(xyz)
        return (((xyz ** 2).sum(axis=-1) + self.R ** 2 - self.r ** 2) ** 2
                - 4. * self.R ** 2 * (xyz[..., [0, 2]] ** 2).sum(axis=-1))

    def _line_poly(self, o, v):
        """Coefficients (highest power first) of Q(o + k v) in k (local frame)."""
        R2, r2 = self.R ** 2, self.r ** 2
        a2 = np.dot(v, v)
        a1 = 2 * np.dot(o, v)
        a0 = np.dot(o, o) + R2 - r2
        b2 = v[0] ** 2 + v[2] ** 2
        b1 = 2 * (o[0] * v[0] + o[2] * v[2])
        b0 = o[0] ** 2 + o[2] ** 2
        # (a2 k^2 + a1 k + a0)^2 - 4 R^2 (b2 k^2 + b1 k + b0)
        return np.array([a2 * a2,
                         2 * a2 * a1,
                         a1 * a1 + 2 * a2 * a0 - 4 * R2 * b2,
                         2 * a1 * a0 - 4 * R2 * b1,
                         a0 * a0 - 4 * R2 * b0])

    def solve_quartic(self, origin, v, transform=True):
        """Intersection of the line ``origin + k v`` with the torus nearest to
        the reference's starting guess; returns a homogeneous point."""
        origin = h2e(np.asarray(origin, dtype=float))
        v = h2e(np.asarray(v, dtype=float))
        if transform:
            inv = np.linalg.inv(self.pos4d)
            o_l = inv[:3, :3] @ origin + inv[:3, 3]
            v_l = inv[:3, :3] @ v
        else:
            o_l, v_l = origin, v
        k0 = max(np.abs(self.r - self.R), self.r * 1e-2) * 0.1
        coeffs = self._line_poly(o_l, v_l)
        roots = np.roots(coeffs)
        scale = max(np.abs(roots).max(), 1.)
        real = roots[np.abs(roots.imag) <= 1e-6 * scale].real
        if len(real) == 0:
            raise Exception('Intersection with torus not found.')
        k = real[np.argmin(np.abs(real - k0))]
        dcoef = np.polyder(coeffs)
        for _ in range(8):
            dq = np.polyval(dcoef, k)
            if dq == 0:
                break
            step = np.polyval(coeffs, k) / dq
            k = k - step
            if abs(step) <= 1e-15 * max(1., abs(k)):
                break
        return e2h(origin + k * v, 1)

    # -- the torus as a core circle with a tube around it --------------------
    # In the local frame the core circle has radius R in the x-z plane and the
    # symmetry axis is y.  A point is core(phi) + r t(theta, phi) with
    #   u(phi) = (cos phi, 0, sin phi)        (radial unit vector of the core)
    #   t(theta, phi) = cos(theta) u(phi) + sin(theta) y   (unit, outward)
    # so t is also the surface normal there.
    @staticmethod
    def _tube_direction(theta, phi):
        ct = np.cos(theta)
        return np.stack([ct * np.cos(phi), np.sin(theta), ct * np.sin(phi)], axis=-1)

    def _global(self, vec, w):
        return np.einsum('...ij,...j', self.pos4d, e2h(vec, w))

    def parametric(self, theta, phi):
        """Points (homogeneous, global) at tube angle ``theta`` and azimuth
        ``phi``: core point R u(phi) plus r times the tube direction."""
        theta = np.asarray(theta, dtype=float)
        phi = np.asarray(phi, dtype=float)
        core = self.R * np.stack([np.cos(phi), np.zeros_like(phi), np.sin(phi)], axis=-1)
        return self._global(core + self.r * self._tube_direction(theta, phi), 1)

    def normal_parametric(self, theta, phi):
        """Outward unit normals (homogeneous direction vectors, global) at
        the surface points of :meth:`parametric`: the tube direction."""
        theta = np.asarray(theta, dtype=float)
        phi = np.asarray(phi, dtype=float)
        if theta.ndim != 1 or phi.ndim != 1:
            raise ValueError('theta and phi must be 1-d arrays.')
        return self._global(self._tube_direction(theta, phi), 0)

    def _check_on_surface(self, xyz):
        if not np.allclose(self.quartic(xyz, transform=False) / self.R ** 4., 0.):
            raise ValueError('Parametric representation is only defined for '
                             'points on torus surface.')

    def xyzw2parametric(self, xyzw, transform=True, intersectvalid=True):
        """(theta, phi) of surface points.  A point at distance rho from the
        symmetry axis lies on the tube around the core point of its own
        azimuth when |x|^2 + R^2 - r^2 = 2 R rho, else (a spindle torus,
        r > R, near the axis) on the tube around the opposite core point, at
        azimuth + pi.  theta is the angle of the point about that core point,
        measured from u(phi) towards y, in (-pi, pi].  Points on the axis
        have no azimuth: phi = 0 there (NaN with ``intersectvalid=False``)."""
        xyzw = np.asarray(xyzw, dtype=float)
        if transform:
            xyzw = np.einsum('...ij,...j', np.linalg.inv(self.pos4d), xyzw)
        xyz = h2e(xyzw)
        self._check_on_surface(xyz)
        x, y, z = xyz[:, 0], xyz[:, 1], xyz[:, 2]
        rho = np.hypot(x, z)
        same = (xyz * xyz).sum(axis=1) + self.R ** 2 - self.r ** 2 >= 0
        on_axis = rho == 0
        az = np.arctan2(z, x)
        phi = np.where(same, az, az + np.pi)
        theta = np.arctan2(y, np.where(same, rho - self.R, -rho - self.R))
        phi = np.where(on_axis, np.nan if not intersectvalid else 0., phi)
        return theta, phi

    def normal(self, xyzw):
        """Outward unit normals (homogeneous direction vectors, global) at
        surface points ``xyzw`` (N, 4): in the local frame the gradient of
        the quartic, grad Q = 4 a x - 8 R^2 (x, 0, z) with
        a = |x|^2 + R^2 - r^2, normalised and turned by the sign of a (which
        is negative on the inner sheet of a spindle torus, where the gradient
        points into the tube)."""
        xyzw = np.asarray(xyzw, dtype=float)
        if xyzw.ndim != 2:
            raise ValueError('Shape of input array must be (N, 4).')
        xyz = h2e(np.einsum('...ij,...j', np.linalg.inv(self.pos4d), xyzw))
        self._check_on_surface(xyz)
        a = (xyz * xyz).sum(axis=1) + self.R ** 2 - self.r ** 2
        grad = 4 * a[:, None] * xyz
        grad[:, [0, 2]] -= 8 * self.R ** 2 * xyz[:, [0, 2]]
        n = grad / np.linalg.norm(grad, axis=1)[:, None] * np.where(a < 0, -1., 1.)[:, None]
        return self._global(n, 0)


def design_tilted_torus(f, alpha, beta):
    """Torus (R, r, pos4d) whose Rowland circle passes through the origin
    (the focus) and the point at distance ``f`` along x tilted by ``alpha``,
    with the hinge at angle ``beta`` (marxs/design/rowland.py:369-422).

    The circle through both points with its centre on the perpendicular
    bisector of the chord has radius r = f / (2 cos alpha); the core circle
    (radius R = r cos beta) is the circle's centre rotated about the hinge by
    beta - alpha, i.e. the torus frame is turned by beta - alpha about -z and
    placed at the chord midpoint (f/2, f/2 tan alpha) minus R along the
    rotated x axis."""
    r = f / (2. * np.cos(alpha))
    R = r * np.cos(beta)
    turn = beta - alpha
    rot = axangle2mat([0, 0, -1], turn)
    mid = np.array([f / 2, f / 2 * np.tan(alpha), 0.])
    centre = mid + R * np.array([-np.cos(turn), np.sin(turn), 0.])
    return R, r, compose(centre, rot, np.ones(3))


class ElementsOnTorus(ParallelCalculated):
    def __init__(self, **kwargs):
        self.rowland = kwargs.pop('rowland')
        self.guess_distance = kwargs.pop('guess_distance',
                                         self.rowland.r + self.rowland.R)
        self.d_element = kwargs.pop('d_element')
        self.optimize_axis = np.asanyarray(
            kwargs.pop('optimize_axis', self.rowland.pos4d @ np.array([1, 0, 0, 0])))
        self.id_col = kwargs.pop('id_col', 'facet')
        if 'normal_spec' not in kwargs:
            kwargs['normal_spec'] = lambda xyzw: -self.rowland.normal(xyzw)
        if 'parallel_spec' not in kwargs:
            kwargs['parallel_spec'] = self.rowland.pos4d @ np.array([0, 1, 0, 0])
        kwargs['pos_spec'] = self.elempos
        super().__init__(**kwargs)

    def elemposyz(self):
        raise NotImplementedError

    def elempos(self):
        ypos, zpos = self.elemposyz()
        ypos = np.asarray(ypos, dtype=float)
        zpos = np.asarray(zpos, dtype=float)
        posyz = np.vstack([np.zeros_like(ypos), ypos, zpos])
        origin = (np.einsum('...ij,...j', self.rowland.pos4d, e2h(posyz.T, 1))
                  + self.guess_distance * self.optimize_axis)
        return np.vstack([self.rowland.solve_quartic(o, self.optimize_axis)
                          for o in origin])


def _centred_row(lo, hi, pitch):
    """Centres of ceil((hi - lo) / pitch) elements (at least one) of width
    ``pitch`` placed edge to edge and centred on the middle of [lo, hi]."""
    n = max(1, int(np.ceil((hi - lo) / pitch)))
    return 0.5 * (lo + hi) + (np.arange(n) - 0.5 * (n - 1)) * pitch


class RectangularGrid(ElementsOnTorus):
    """Elements on a rectangular y-z grid of pitch ``d_element`` centred on
    ``y_range`` x ``z_range`` (marxs/design/rowland.py:660-700; the default
    z range is a single row), ordered row by row in z."""

    def __init__(self, **kwargs):
        self.y_range = kwargs.pop('y_range')
        self.z_range = kwargs.pop('z_range', [-1e-10, 1e-10])
        super().__init__(**kwargs)

    def elemposyz(self):
        ys = _centred_row(self.y_range[0], self.y_range[1], self.d_element[0])
        zs = _centred_row(self.z_range[0], self.z_range[1], self.d_element[1])
        return np.tile(ys, len(zs)), np.repeat(zs, len(ys))


class CircularMeshGrid(ElementsOnTorus):
    """Elements on rows of pitch d_element[0] across an annulus
    radius[0] .. radius[1] (marxs/design/rowland.py:703-755): the rows start
    at half the row block's width below the centre; on each row the chord
    through the annulus (one chord outside the hole, two mirrored chords
    beside it) is filled with elements of pitch d_element[1] centred on the
    chord, the chord's far end excluded; a row beyond the outer radius holds
    one element at z = 0."""

    def __init__(self, **kwargs):
        self.radius = kwargs.pop('radius')
        super().__init__(**kwargs)

    def _row(self, y):
        r_in, r_out = self.radius[0], self.radius[1]
        dz = self.d_element[1]
        if abs(y) > r_out:
            return np.zeros(1)
        half = np.sqrt(r_out ** 2 - y ** 2)
        if abs(y) > r_in:
            n = int(np.ceil(2 * half / dz))
            return np.arange(-0.5 * n * dz, half, dz)
        hole = np.sqrt(r_in ** 2 - y ** 2)
        n = int(np.ceil((half - hole) / dz))
        side = np.arange(0.5 * (hole + half) - 0.5 * n * dz, half, dz)
        return np.concatenate([-side, side])

    def elemposyz(self):
        dy = self.d_element[0]
        n_rows = int(np.ceil(2 * self.radius[1] / dy))
        rows = np.arange(-0.5 * n_rows * dy, self.radius[1], dy)
        zs = [self._row(y) for y in rows]
        return (np.concatenate([np.full(len(z), y) for y, z in zip(rows, zs)]).tolist(),
                np.concatenate(zs).tolist())


class GratingArrayStructure(ElementsOnTorus):
    """Facets on rings (marxs/design/rowland.py:758-888).

    ``radius`` holds pairs of ring bounds; each pair gets
    ceil(width / d_element[0]) rings of pitch d_element[0] centred in it.  A
    ring of radius rho carries floor((rho - d_element[0] / 2) span /
    d_element[1]) facets over the azimuth range ``phi`` (span = its width),
    spread with equal gaps: with w = d_element[1] / (2 pi rho) the angle
    the reference gives one facet, the gap is (span - n w) / (n + 1) and
    facet k sits at phi[0] + (k + 1) gap + (k + 1/2) w.  Positions
    y = rho sin(angle), z = rho cos(angle), ring by ring."""

    def __init__(self, **kwargs):
        self.phi = kwargs.pop('phi', [0., 2 * np.pi])
        self.radius = kwargs.pop('radius')
        super().__init__(**kwargs)

    def max_elements_on_radius(self, radius):
        """Rings that fit between the two bounds of ``radius``."""
        return int(np.ceil((radius[1] - radius[0]) / self.d_element[0]))

    def distribute_elements_on_radius(self):
        """Ring radii for every pair of bounds in ``self.radius``."""
        bounds = np.asarray(self.radius, dtype=float).reshape(-1, 2)
        rings = []
        for lo, hi in bounds:
            n = self.max_elements_on_radius([lo, hi])
            rings.append(0.5 * (lo + hi) + (np.arange(n) - 0.5 * (n - 1)) * self.d_element[0])
        return np.concatenate(rings) if rings else np.zeros(0)

    def max_elements_on_arc(self, radius):
        """Facets of pitch d_element[1] that fit on an arc of ``radius``
        over the azimuth span (float, rounded down)."""
        return radius * anglediff(self.phi) // self.d_element[1]

    def distribute_elements_on_arc(self, radius):
        """Facet azimuths on the ring of ``radius``, in [0, 2 pi)."""
        if len(self.phi) == 1:
            return self.phi
        span = anglediff(self.phi)
        n = self.max_elements_on_arc(radius - self.d_element[0] / 2)
        w = self.d_element[1] / (2. * np.pi * radius)
        gap = (span - n * w) / (n + 1)
        k = np.arange(n)
        return (self.phi[0] + (gap + 0.5 * w + k * (gap + w))) % (2. * np.pi)

    def elemposyz(self):
        rings = self.distribute_elements_on_radius()
        per_ring = [self.distribute_elements_on_arc(rho) for rho in rings]
        rho = np.concatenate([np.full(len(a), r) for r, a in zip(rings, per_ring)])
        ang = np.concatenate(per_ring)
        return rho * np.sin(ang), rho * np.cos(ang)
