"""Rowland-torus geometry and element placement (set-up, host float64).

Follows ``marxs/design/rowland.py``:

* ``RowlandTorus`` (``78-366``): torus with the local y axis as symmetry
  axis, ``Q(x) = (|x|^2 + R^2 - r^2)^2 - 4 R^2 (x_x^2 + x_z^2)``;
  ``solve_quartic`` intersects a line with the torus.  The reference runs
  ``scipy.optimize.root`` from ``k0 = 0.1 max(|r - R|, 0.01 r)``; here the
  quartic in the line parameter ``k`` is expanded into its polynomial
  coefficients, all roots are found with a companion-matrix eigen solve, the
  real root nearest ``k0`` is taken and polished with Newton steps.
* ``design_tilted_torus`` (``369-422``).
* ``ElementsOnTorus``/``RectangularGrid``/``CircularMeshGrid``/
  ``GratingArrayStructure`` (``547-888``): element centres from a y-z layout
  pushed along ``optimize_axis`` onto the torus; normals ``-torus.normal``;
  in-plane axis parallel to the torus symmetry axis.
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

    def parametric(self, theta, phi):
        theta = np.asarray(theta, dtype=float)
        phi = np.asarray(phi, dtype=float)
        x = (self.R + self.r * np.cos(theta)) * np.cos(phi)
        z = (self.R + self.r * np.cos(theta)) * np.sin(phi)
        y = self.r * np.sin(theta)
        torus = np.array([x, y, z, np.ones_like(z)]).T
        return np.einsum('...ij,...j', self.pos4d, torus)

    def xyzw2parametric(self, xyzw, transform=True, intersectvalid=True):
        xyzw = np.asarray(xyzw, dtype=float)
        if transform:
            xyzw = np.einsum('...ij,...j', np.linalg.inv(self.pos4d), xyzw)
        xyz = h2e(xyzw)
        if not np.allclose(self.quartic(xyz, transform=False) / self.R ** 4., 0.):
            raise ValueError('Parametric representation is only defined for '
                             'points on torus surface.')
        s = np.sign(np.sqrt(xyz[:, 0] ** 2 + xyz[:, 2] ** 2) - self.R)
        theta = np.arcsin(s * xyz[:, 1] / self.r) + (s < 0) * np.pi
        factor = self.R + self.r * np.cos(theta)
        phi = np.zeros_like(factor)
        ind = factor != 0
        phi[ind] = np.arctan2(xyz[ind, 2] / factor[ind], xyz[ind, 0] / factor[ind])
        if not intersectvalid:
            phi[~ind] = np.nan
        return theta, phi

    def normal_parametric(self, theta, phi):
        theta = np.asarray(theta, dtype=float)
        phi = np.asarray(phi, dtype=float)
        if not (theta.ndim == 1 and phi.ndim == 1):
            raise ValueError('theta and phi must be 1-d arrays.')
        normal = np.empty((len(theta), 3))
        normal[:, 0] = np.cos(theta) * np.cos(phi)
        normal[:, 1] = np.sin(theta)
        normal[:, 2] = np.cos(theta) * np.sin(phi)
        return np.einsum('...ij,...j', self.pos4d, e2h(normal, 0))

    def normal(self, xyzw):
        xyzw = np.asarray(xyzw, dtype=float)
        if xyzw.ndim != 2:
            raise ValueError('Shape of input array must be (N, 4).')
        theta, phi = self.xyzw2parametric(xyzw, 1)
        return self.normal_parametric(theta, phi)


def design_tilted_torus(f, alpha, beta):
    """Torus parameters (R, r, pos4d) for focal distance ``f`` with the
    Rowland circle tilted by ``alpha`` and the hinge at ``beta``."""
    r = f / (2. * np.cos(alpha))
    R = r * np.sin(np.pi / 2 - beta)
    orientation = axangle2mat([0, 0, -1], beta - alpha)
    x_ct = f / 2 - R * np.cos(beta - alpha)
    y_ct = f / 2 * np.tan(alpha) + R * np.sin(beta - alpha)
    return R, r, compose([x_ct, y_ct, 0], orientation, np.ones(3))


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


class RectangularGrid(ElementsOnTorus):
    def __init__(self, **kwargs):
        self.y_range = kwargs.pop('y_range')
        self.z_range = kwargs.pop('z_range', [-1e-10, 1e-10])
        super().__init__(**kwargs)

    def elemposyz(self):
        n_y = max(1, int(np.ceil((self.y_range[1] - self.y_range[0]) / self.d_element[0])))
        n_z = max(1, int(np.ceil((self.z_range[1] - self.z_range[0]) / self.d_element[1])))
        width_y = n_y * self.d_element[0]
        width_z = n_z * self.d_element[1]
        ypos = np.arange(0.5 * (self.y_range[0] - width_y + self.y_range[1] + self.d_element[0]),
                         self.y_range[1], self.d_element[0])
        zpos = np.arange(0.5 * (self.z_range[0] - width_z + self.z_range[1] + self.d_element[1]),
                         self.z_range[1], self.d_element[1])
        ypos, zpos = np.meshgrid(ypos, zpos)
        return ypos.flatten(), zpos.flatten()


class CircularMeshGrid(ElementsOnTorus):
    def __init__(self, **kwargs):
        self.radius = kwargs.pop('radius')
        super().__init__(**kwargs)

    def elemposyz(self):
        n_y = int(np.ceil(2 * self.radius[1] / self.d_element[0]))
        y_width = n_y * self.d_element[0]
        ypos, zpos = [], []
        for y in np.arange(-y_width / 2, self.radius[1], self.d_element[0]):
            if np.abs(y) > self.radius[1]:
                z = np.array([0])
            else:
                z_outer = np.sqrt(self.radius[1] ** 2 - y ** 2)
                if np.abs(y) > self.radius[0]:
                    n_z = int(np.ceil(2 * z_outer / self.d_element[1]))
                    z_width = n_z * self.d_element[1]
                    z = np.arange(-z_width / 2, z_outer, self.d_element[1])
                else:
                    z_inner = np.sqrt(self.radius[0] ** 2 - y ** 2)
                    z_mid = 0.5 * (z_inner + z_outer)
                    n_z = int(np.ceil((z_outer - z_inner) / self.d_element[1]))
                    z_width = n_z * self.d_element[1]
                    z = np.arange(z_mid - z_width / 2, z_outer, self.d_element[1])
                    z = np.hstack([-z, z])
            ypos.extend([y] * len(z))
            zpos.extend(z)
        return ypos, zpos


class GratingArrayStructure(ElementsOnTorus):
    def __init__(self, **kwargs):
        self.phi = kwargs.pop('phi', [0., 2 * np.pi])
        self.radius = kwargs.pop('radius')
        super().__init__(**kwargs)

    def max_elements_on_radius(self, radius):
        return int(np.ceil((radius[1] - radius[0]) / self.d_element[0]))

    def distribute_elements_on_radius(self):
        radii = []
        for i in range(len(self.radius) // 2):
            bracket = self.radius[2 * i: 2 * i + 2]
            n = self.max_elements_on_radius(bracket)
            radii.append(np.mean(bracket)
                         + np.arange(-n / 2 + 0.5, n / 2 + 0.5) * self.d_element[0])
        return np.hstack(radii)

    def max_elements_on_arc(self, radius):
        return radius * anglediff(self.phi) // self.d_element[1]

    def distribute_elements_on_arc(self, radius):
        if len(self.phi) == 1:
            return self.phi
        n = self.max_elements_on_arc(radius - self.d_element[0] / 2)
        element_angle = self.d_element[1] / (2. * np.pi * radius)
        d_between = (anglediff(self.phi) - n * element_angle) / (n + 1)
        centerangles = d_between + 0.5 * element_angle + np.arange(n) * (d_between + element_angle)
        return (self.phi[0] + centerangles) % (2. * np.pi)

    def elemposyz(self):
        radii = self.distribute_elements_on_radius()
        angles = [self.distribute_elements_on_arc(r) for r in radii]
        radii = np.concatenate([[radii[i]] * len(a) for i, a in enumerate(angles)])
        angles = np.concatenate(angles)
        return radii * np.sin(angles), radii * np.cos(angles)
