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
