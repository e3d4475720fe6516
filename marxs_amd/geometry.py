"""Element geometries: the pose matrix and the frame derived from it.

``Geometry.__getitem__`` follows ``marxs/math/geometry.py:108-160``: the
``center`` is column 3 of ``pos4d``; ``v_x``/``v_y``/``v_z`` are columns 0-2
and ``e_*`` the same columns divided by their length.  ``FinitePlane`` is the
flat rectangle ``|y| <= |v_y|, |z| <= |v_z|`` in the plane through ``center``
with normal ``e_x`` (``geometry.py:211-261``).  On this path the per-photon
intersection runs on the GPU; the host only derives the frame constants once
per element (:func:`plane_frame`).
"""
from copy import copy

import numpy as np

from .base import parse_position_keywords

__all__ = ['Geometry', 'FinitePlane', 'PlaneWithHole', 'RectangleHole',
           'CircularHole', 'Cylinder', 'plane_frame']


class Geometry:
    shape = 'none'
    _geometry = {}

    def __init__(self, kwargs=None):
        kwargs = {} if kwargs is None else kwargs
        self.pos4d = parse_position_keywords(kwargs)
        self._geometry = copy(self._geometry)

    def __getitem__(self, key):
        if key == 'center':
            return self.pos4d[:, 3]
        if key in ('v_x', 'e_x'):
            val = self.pos4d[:, 0]
        elif key in ('v_y', 'e_y'):
            val = self.pos4d[:, 1]
        elif key in ('v_z', 'e_z'):
            val = self.pos4d[:, 2]
        else:
            val = self._geometry[key]
            if isinstance(val, np.ndarray) and val.shape[-1] == 4:
                val = np.dot(self.pos4d, val)
        if key[:2] == 'e_':
            return val / np.linalg.norm(val)
        return val


class FinitePlane(Geometry):
    shape = 'box'
    loc_coos_name = ['y', 'z']

    def get_local_euklid_bases(self, interpos_local):
        """(e_y, e_z, e_x) of the plane for each of the N local positions,
        homogeneous (N, 4) arrays (``geometry.py:263-281``)."""
        n = np.asarray(interpos_local).shape[0]
        return (np.tile(self['e_y'], (n, 1)), np.tile(self['e_z'], (n, 1)),
                np.tile(self['e_x'], (n, 1)))


class PlaneWithHole(FinitePlane):
    def __init__(self, kwargs):
        self._geometry = dict(self._geometry)
        r_inner = kwargs.pop('r_inner', 0.)
        super().__init__(kwargs)
        self._geometry['r_inner'] = r_inner


class RectangleHole(PlaneWithHole):
    pass


class CircularHole(PlaneWithHole):
    def __init__(self, kwargs):
        phi = kwargs.pop('phi', [0, 2. * np.pi])
        if np.max(np.abs(phi)) > 10:
            raise ValueError('Input angles >> 2 pi. Did you use degrees '
                             '(radian expected)?')
        if phi[0] > phi[1]:
            raise ValueError('phi[1] must be greater than phi[0].')
        if (phi[1] - phi[0]) > (2 * np.pi + 1e-6):
            raise ValueError('phi[1] - phi[0] must be less than 2 pi.')
        self.phi = phi
        super().__init__(kwargs)


class Cylinder(Geometry):
    """A tube: the unit circle in the local x-y plane, |z| <= 1, mapped by
    ``pos4d`` (``geometry.py:383-564``).  ``phi_lim`` restricts the tube to a
    wedge of azimuth (radian, default the full circle).  Local coordinates
    of a hit are the azimuth phi and z in global units (local z times the z
    zoom); of the two intersections of a ray with the tube the nearer forward
    one within the limits counts."""
    shape = 'surface'
    loc_coos_name = ['phi', 'z']
    n_points = 50

    def __init__(self, kwargs=None):
        kwargs = {} if kwargs is None else kwargs
        self.coos_limits = [np.asanyarray(kwargs.pop('phi_lim', [-np.pi, np.pi]), dtype=float),
                            np.array([-1., 1.])]
        super().__init__(kwargs)

    def __getitem__(self, key):
        if key == 'R':
            from .mathutils import decompose44
            return decompose44(self.pos4d)[2][0]
        return super().__getitem__(key)

    def get_local_euklid_bases(self, interpos_local):
        """Unit vectors along phi, along the tube axis and normal to the
        surface at local positions (phi, z), homogeneous (N, 4) arrays
        (``geometry.py:602-627``)."""
        phi = np.asarray(interpos_local, dtype=float)[:, 0]
        zeros = np.zeros_like(phi)
        e_phi = np.einsum('...ij,...j', self.pos4d,
                          np.array([-np.sin(phi), np.cos(phi), zeros, zeros]).T)
        e_n = np.einsum('...ij,...j', self.pos4d,
                        np.array([np.cos(phi), np.sin(phi), zeros, zeros]).T)
        e_z = np.tile(self['e_z'], (len(phi), 1))
        unit = lambda v: v / np.linalg.norm(v, axis=1)[:, None]
        return unit(e_phi), e_z, unit(e_n)

    def parametric_surface(self, phi=None, z=None):
        """Points of the tube in global homogeneous coordinates, shape
        (n_phi, n_z, 4) (``geometry.py:567-600``): azimuths ``phi`` (default
        ``n_points`` across the limits) and axial positions ``z`` in global
        units (default both rims)."""
        from .mathutils import decompose44
        phi = (np.linspace(self.coos_limits[0][0], self.coos_limits[0][1], self.n_points)
               if phi is None else np.asanyarray(phi))
        zoom = decompose44(self.pos4d)[2]
        z = self.coos_limits[1] if z is None else np.asanyarray(z) / zoom[2]
        if phi.ndim != 1 or z.ndim != 1:
            raise ValueError('input parameters have 1-dim shape.')
        phi, z = np.meshgrid(phi, z)
        coos = np.array([np.cos(phi), np.sin(phi), z, np.ones_like(z)]).T
        return np.einsum('...ij,...j', self.pos4d, coos)

    @classmethod
    def from_rowland(cls, rowland, width, rotation=0., kwargs=None):
        """The tube that follows the Rowland circle of ``rowland``
        (``geometry.py:430-462``): radius r around (R, 0, 0) of the torus
        frame, half-width ``width`` along the torus axis, turned by
        ``rotation`` about the tube axis."""
        from .mathutils import axangle2mat, compose
        rot = axangle2mat(np.array([0, 0, 1.]), rotation)
        pos4d = np.dot(rowland.pos4d, compose([rowland.R, 0, 0], rot,
                                              [rowland.r, rowland.r, width]))
        det_kwargs = {'pos4d': pos4d}
        det_kwargs.update(kwargs or {})
        return cls(det_kwargs)


def plane_frame(geometry):
    """Frame constants of a flat element, as the intersection test uses them.

    Returns a dict with 3-vectors ``center``, ``e_x``, ``e_y``, ``e_z``,
    ``v_y``, ``v_z`` and the half sizes ``hy = |v_y|``, ``hz = |v_z|``.
    """
    g = geometry
    return {
        'center': np.array(g['center'][:3], dtype=float),
        'e_x': np.array(g['e_x'][:3], dtype=float),
        'e_y': np.array(g['e_y'][:3], dtype=float),
        'e_z': np.array(g['e_z'][:3], dtype=float),
        'v_y': np.array(g['v_y'][:3], dtype=float),
        'v_z': np.array(g['v_z'][:3], dtype=float),
        'hy': float(np.linalg.norm(g['v_y'])),
        'hz': float(np.linalg.norm(g['v_z'])),
    }
