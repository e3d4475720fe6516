"""Oracle restatement of the tube geometry and the circular detector on the
CPU, against the reference's known answers
(marxs/math/tests/test_geometry.py:56-222) and closed forms."""
import numpy as np
import pytest

import marxs_amd as mx
from marxs_amd.geometry import Cylinder
from marxs_amd.mathutils import axangle2mat
from oracle import trace as OT
from oracle.describe import describe


def _cyl(**kw):
    c = Cylinder(dict(kw))
    return {'pos4d': c.pos4d, 'phi_lim': np.asarray(c.coos_limits[0], dtype=float)}


def _isect(el, dir, pos):
    dir, pos = np.asarray(dir, dtype=float), np.asarray(pos, dtype=float)
    return OT.cylinder_intersect(el, pos[:, :3], dir[:, :3])


def test_tube_misses():
    """test_geometry.py:56-86"""
    hit, q, phi, z = _isect(_cyl(), [[1., 0., .1, 0], [-1., -1., 0., 0.]],
                            [[0, -1.1, 0., 1.], [2., 0., 0., 1.]])
    assert not hit.any()
    hit, q, phi, z = _isect(_cyl(), [[0., 0.1, 1., 0]], [[0.5, 0.5, 0., 1.]])
    assert not hit.any()
    hit, q, phi, z = _isect(_cyl(position=[0.8, 0.8, 1.2]), [[1., 0., .0, 0], [1., 0., 0., 0.]],
                            [[0., 0., 0., 1.], [0., -0.3, 1., 1.]])
    assert not hit.any()


def test_tube_rays_forward():
    """test_geometry.py:89-102"""
    pos = [[20., 0., 0.8, 1.], [0., .0, 0.4, 1.], [-5., .1, -.2, 1]]
    dir = np.array([[-.1, 0., 0., 0], [-1., 0, 0.02, 0.], [-2., -0.02, 0.023, 0.]])
    el = _cyl(zoom=[1, 1, 1])
    assert np.array_equal(_isect(el, dir, pos)[0], [True, True, False])
    assert np.array_equal(_isect(el, -dir, pos)[0], [False, True, True])


def test_tube_hit_miss_zoomed():
    """test_geometry.py:105-118"""
    dir = [[-.1, 0., 0., 0], [-1., 0, 0.2, 0.]]
    pos = [[20., 0., 0.8, 1.], [2., .0, 0.4, 1.]]
    assert _isect(_cyl(zoom=[1, 1, 1]), dir, pos)[0].all()
    assert not _isect(_cyl(zoom=[1, 1, .5]), dir, pos)[0].any()


POS = [[50., 0., 0., 1.], [10., .5, 0., 1.], [2, 0, .3, 1.]]
DIR = [[-.1, 0., 0., 0], [-0.5, 0, 0., 0.], [-1., 0., 0., 0.]]
EXP_POS = np.array([[1., 0., 0.], [np.sqrt(0.75), 0.5, 0.], [1, 0., .3]])
EXP_LOC = np.array([[0., np.arcsin(0.5), 0.], [0., 0., 0.3]])


def test_tube_two_points_and_transforms():
    """test_geometry.py:121-168: nearer root; zoom in z keeps the local z in
    global units; a rotation about the axis shifts phi; a shift moves the
    points."""
    for el in (_cyl(), _cyl(zoom=[1, 1., 5.])):
        hit, q, phi, z = _isect(el, DIR, POS)
        assert hit.all()
        assert np.allclose(q, EXP_POS) and np.allclose(phi, EXP_LOC[0]) and np.allclose(z, EXP_LOC[1])
    hit, q, phi, z = _isect(_cyl(orientation=axangle2mat([0, 0, 1], -.3)), DIR, POS)
    assert hit.all() and np.allclose(q, EXP_POS)
    assert np.allclose(phi, [0.3, 0.3 + np.arcsin(0.5), 0.3])
    hit, q, phi, z = _isect(_cyl(position=[-.3, 0, 0]), DIR, POS)
    assert np.allclose(q, [[.7, 0, 0], [np.sqrt(0.75) - .3, 0.5, 0.], [.7, 0., .3]])
    assert np.allclose(phi, EXP_LOC[0])


def test_tube_phi_limits():
    """test_geometry.py:171-194: limits pick the far side, or no hit."""
    hit, q, phi, z = _isect(_cyl(phi_lim=[.6, np.pi]), DIR, POS)
    assert hit.all()
    assert np.allclose(q, [[-1., 0., 0.], [-np.sqrt(0.75), 0.5, 0.], [-1, 0., .3]])
    assert np.allclose(phi, [np.pi, np.pi - np.arcsin(0.5), np.pi])
    assert not _isect(_cyl(phi_lim=[-.4, -.3]), DIR, POS)[0].any()


def test_tube_translated_zoomed():
    """test_geometry.py:212-223"""
    hit, q, phi, z = _isect(_cyl(position=[0.8, 0.8, 1.2], zoom=[1, 1, 2]),
                            [[1., 0., .0, 0], [1., 0., 0., 0.]], [[0., 0., 0., 1.], [0., 0., 1., 1.]])
    assert hit.all() and np.allclose(q, [[.2, 0., 0.], [.2, 0., 1.]])


def test_angle_between_wraps():
    """utils.py:180-215 docstring example and the cyclic case."""
    assert OT.angle_between(np.array([-0.1]), -0.2, 6)[0]
    assert OT.angle_between(np.deg2rad(np.array([355.])), np.deg2rad(350), np.deg2rad(10))[0]
    assert not OT.angle_between(np.deg2rad(np.array([20.])), np.deg2rad(350), np.deg2rad(10))[0]


def test_random_rays_on_surface():
    """Brute force: rays aimed at random points of a rotated, zoomed tube hit
    them (nearer root from far outside), radius r from the axis."""
    rng = np.random.default_rng(3)
    c = Cylinder({'zoom': [3., 3., 4.], 'orientation': axangle2mat([1, 2, 3], 0.7),
                  'position': [1., -2., 0.5]})
    el = {'pos4d': c.pos4d, 'phi_lim': np.array([-np.pi, np.pi])}
    phi = rng.uniform(-np.pi, np.pi, 200)
    zl = rng.uniform(-0.9, 0.9, 200)
    pl = np.stack([np.cos(phi), np.sin(phi), zl, np.ones(200)], axis=1)
    pts = (c.pos4d @ pl.T).T[:, :3]
    # start outside along the outward normal, aim back
    nrm = (c.pos4d[:3, :3] @ np.stack([np.cos(phi), np.sin(phi), np.zeros(200)])).T
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    start = pts + 50 * nrm
    hit, q, ph, z = OT.cylinder_intersect(el, start, -nrm)
    assert hit.all()
    assert np.allclose(q, pts, atol=1e-9)
    assert np.allclose(ph, phi, atol=1e-12) and np.allclose(z, 4 * zl, atol=1e-12)


def test_circular_detector_pixels_and_description():
    """detector.py:113-116: detpix_x = phi R / pixsize, detpix_y = z /
    pixsize; photons that miss keep NaN."""
    det = mx.CircularDetector(pixsize=0.01, zoom=[2., 2., 3.])
    t = mx.generate_test_photons(3)
    ph = {'pos': np.array([[10., 0.5, 1.], [10., 0., 0.], [10., 0., 5.]]),
          'dir': np.asarray(t['dir'], dtype=float)[:, :3].copy(),
          'pol': np.asarray(t['polarization'], dtype=float)[:, :3].copy(),
          'energy': np.ones(3), 'prob': np.ones(3)}
    out = OT.trace(describe(det), ph, np.zeros((0, 3)))
    phi0 = np.arctan2(0.5, np.sqrt(4 - 0.25))
    assert np.allclose(out['det_phi'][:2], [phi0, 0.])
    assert np.allclose(out['detpix_x'][:2], [phi0 * 2 / 0.01, 0.])
    assert np.allclose(out['detpix_y'][:2], [100., 0.])
    assert np.isnan(out['detpix_x'][2]) and np.isnan(out['det_y'][2])


def test_tube_parametric_points_are_hit():
    """marxs/math/tests/test_geometry.py:226-239: points of the parametric
    surface, approached along +x from their own position, are their own
    intersection points."""
    circ = Cylinder({'zoom': [2., 3., 4.]})
    par = circ.parametric_surface(phi=[-.1, 0., 1.]).reshape((6, 4))
    dirs = np.tile([1., 0, 0], (6, 1))
    hit, q, phi, z = _isect(_cyl(zoom=[2., 3., 4.]), dirs, par[:, :3])
    assert hit.all() and np.allclose(q, par[:, :3])
    with pytest.raises(ValueError, match='1-dim'):
        circ.parametric_surface(phi=np.zeros((2, 2)))


def test_circular_detector_from_rowland_lies_on_torus():
    """marxs/optics/tests/test_detector.py:44-54: a very narrow tube from
    Cylinder.from_rowland lies on the Rowland torus."""
    from scipy.spatial.transform import Rotation
    rowland = mx.RowlandTorus(R=6e4, r=5e4, position=[123., 345., -678.],
                              orientation=Rotation.from_euler('yxz', [1, 2, 3]).as_matrix())
    det = mx.CircularDetector(geometry=Cylinder.from_rowland(rowland, width=1e-6))
    points = det.geometry.parametric_surface(np.mgrid[0:2. * np.pi:360j])
    assert np.max(np.abs(rowland.quartic(points[..., :3].reshape(-1, 3)))) < 1e5


@pytest.mark.parametrize('geom', [mx.FlatDetector, mx.CircularDetector])
def test_local_coordsys(geom):
    """marxs/math/tests/test_geometry.py:242-258: local bases are orthonormal
    and right-handed."""
    from scipy.spatial.transform import Rotation
    rng = np.random.default_rng(9)
    rot = Rotation.from_euler('xyz', 2 * np.pi * rng.random(3)).as_matrix()
    g = geom(orientation=rot, position=rng.random(3)).geometry
    x, y, z = g.get_local_euklid_bases(rng.random((5, 2)))
    assert np.allclose(np.einsum('ij,ij->i', x, y), 0) and np.allclose(np.einsum('ij,ij->i', x, z), 0)
    for v in (x, y, z):
        assert np.allclose(np.linalg.norm(v, axis=1), 1.)
    assert np.allclose(np.cross(x[:, :3], y[:, :3]), z[:, :3])
