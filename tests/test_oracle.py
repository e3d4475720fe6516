"""Pin the CPU oracle against the reference's own known answers and closed
forms (CPU only).

Each test cites the reference test or formula it encodes.  Planted-mistake
checks (a wrong sign, a transposed operand) are in test_oracle_mutants.py.
"""
import math
from io import StringIO

import numpy as np
import pytest

from oracle import trace as OT
from oracle import placement as PL

E2W = 1.2398419292004202e-06


def plane(pos4d=None, **kw):
    from marxs_amd.base import parse_position_keywords
    if pos4d is None:
        pos4d = parse_position_keywords(dict(kw))
    return OT.frame(pos4d)


# ------------------------------------------------------- intersection
def test_intersect_forwards_backwards():
    """marxs/math/tests/test_geometry.py:11-25"""
    pos = np.array([[20., 0., 0.8], [0., .0, 0.4], [-5., .1, -.2]])
    d = np.array([[-.1, 0., 0.], [-1., 0, 0.02], [-2., -0.02, 0.023]])
    fr = plane(np.eye(4))
    hit, q, ly, lz = OT.intersect(fr, pos, d)
    assert list(hit) == [True, True, False]
    hit, q, ly, lz = OT.intersect(fr, pos, -d)
    assert list(hit) == [False, True, True]


def test_intersect_zoomed():
    """test_geometry.py:28-41"""
    pos = np.array([[1., 4., 0.8], [1., -4., 8.], [1., -6., -8.]])
    d = np.array([[-.1, 0., 0.], [-1., 0, 0.02], [-2., -0.02, 0.023]])
    fr = plane(zoom=[1, 5, 10])
    hit, q, ly, lz = OT.intersect(fr, pos, d)
    assert list(hit) == [True, True, False]
    assert np.all(q[0] == [0, 4, .8])
    assert ly[0] == 4 and lz[0] == .8


def test_intersect_moved():
    """test_geometry.py:44-55"""
    pos = np.array([[4.1, 5.5, 6.8], [5., 4.1, 4.99], [4.1, 3., 6.]])
    d = np.array([[-.1, 0., 0.], [-1., 0, 0.02], [-2., -0.02, 0.023]])
    fr = plane(position=[4, 5, 6])
    hit, q, ly, lz = OT.intersect(fr, pos, d)
    assert list(hit) == [True, True, False]
    assert np.all(q[1] == [4, 4.1, 5.01])
    assert np.allclose([ly[1], lz[1]], [-0.9, -.99])


def test_intersect_closed_form_random():
    """Random rotated/zoomed planes against an independent solve of
    p + k d = c + y e_y + z e_z (3x3 linear system)."""
    from marxs_amd.mathutils import euler2mat_sxyz, compose
    rng = np.random.default_rng(3)
    for _ in range(20):
        rot = euler2mat_sxyz(*rng.uniform(-3, 3, 3))
        zoom = rng.uniform(0.5, 5, 3)
        pos4d = compose(rng.uniform(-10, 10, 3), rot, zoom)
        fr = OT.frame(pos4d)
        p = rng.uniform(-20, 20, (50, 3))
        d = rng.normal(size=(50, 3))
        hit, q, ly, lz = OT.intersect(fr, p, d)
        for i in range(50):
            a = np.column_stack([-d[i], fr['e_y'], fr['e_z']])
            k, y, z = np.linalg.solve(a, p[i] - fr['center'])
            exp_hit = (k >= 0) and abs(y) <= zoom[1] and abs(z) <= zoom[2]
            margin = min(abs(abs(y) - zoom[1]), abs(abs(z) - zoom[2]), abs(k))
            if margin > 1e-9:
                assert hit[i] == exp_hit
            assert np.isclose(ly[i], y, atol=1e-9 * (1 + abs(y)))
            assert np.isclose(lz[i], z, atol=1e-9 * (1 + abs(z)))


def test_frame_rotation():
    """marxs/optics/tests/test_optics.py:41-47: 90 deg about [1,1,0]."""
    from marxs_amd.mathutils import axangle2mat
    rot = axangle2mat(np.array([1., 1., 0.]), np.deg2rad(90))
    fr = plane(orientation=rot)
    assert np.allclose(fr['e_y'], [.5, .5, 1. / np.sqrt(2)])
    assert np.allclose(fr['e_z'], np.array([1., -1, 0]) / np.sqrt(2))
    fr = plane(position=np.array([-2, 3, 4.3]))
    assert np.all(fr['center'] == [-2, 3, 4.3])


# -------------------------------------------------- parallel transport
def test_paralleltransport_known():
    """marxs/math/tests/test_polarization.py:36-72"""
    d1 = OT.unit(np.array([[1., -1, 0]] * 5))
    d2 = OT.unit(np.array([[1., -1, 0], [1, 1, 0], [1, 1, 0], [1, 0.2, 0], [1, 0.2, 0]]))
    pin = OT.unit(np.array([[1., 1, 1], [0, 0, 1], [1, 1, 0], [0, 0, 1], [1, 1, 0]]))
    out = OT.parallel_transport(d1, d2, pin)
    assert np.allclose(pin[0], out[0])
    assert np.allclose(pin[:, 2], out[:, 2])
    assert np.allclose(np.linalg.norm(out, axis=1), 1)
    assert np.allclose(OT.dot(out, d2), 0)


def test_paralleltransport_rodrigues():
    """For pol perpendicular to d1 the transport is the rotation taking d1 to
    d2 about d1 x d2 (Rodrigues' formula)."""
    rng = np.random.default_rng(5)
    d1 = OT.unit(rng.normal(size=(200, 3)))
    d2 = OT.unit(rng.normal(size=(200, 3)))
    pol = OT.unit(np.cross(d1, rng.normal(size=(200, 3))))
    out = OT.parallel_transport(d1, d2, pol)
    ax = np.cross(d1, d2)
    s = np.linalg.norm(ax, axis=1)
    c = np.sum(d1 * d2, axis=1)
    k = ax / s[:, None]
    exp = (pol * c[:, None] + np.cross(k, pol) * s[:, None]
           + k * np.sum(k * pol, axis=1)[:, None] * (1 - c)[:, None])
    assert np.allclose(out, exp, atol=1e-12)


# ----------------------------------------------------------- gratings
def grating_el(d=1e-3, sign='flat', zoom=20, groove=0., selector=None, **kw):
    from marxs_amd.base import parse_position_keywords
    el = {'kind': 'grating', 'pos4d': parse_position_keywords({'zoom': zoom}),
          'd': d, 'sign': sign, 'groove_angle': groove,
          'selector': selector or {'type': 'orderselector', 'orders': [1], 'p': [1.]}}
    el.update(kw)
    return el


def run_grating(el, pos, d, energy, orders):
    fr = OT.frame(el['pos4d'])
    hit, q, ly, lz = OT.intersect(fr, pos, d)
    assert hit.all()
    new, blaze = OT.diffract(el, fr, d, energy, ly, np.asarray(orders), None)
    return new, blaze


def test_grating_equation_angles():
    """marxs/optics/tests/test_grating.py:65-92: sin(alpha) = sin(beta) + lambda/d"""
    beta = np.deg2rad([0., 5., 10., 20.])
    d = np.zeros((4, 3))
    d[:, 1] = np.sin(beta)
    d[:, 0] = -np.cos(beta)
    new, _ = run_grating(grating_el(d=0.001), np.ones((4, 3)), d, np.ones(4), np.ones(4, int))
    alpha = np.arctan2(new[:, 1], -new[:, 0])
    assert np.allclose(np.sin(alpha), np.sin(beta) + E2W / 0.001)


def test_grating_order_symmetry():
    """test_grating.py:94-111"""
    d = np.tile([-1., 0, 0], (5, 1))
    new, _ = run_grating(grating_el(d=1. / 500), np.ones((5, 3)), d, np.ones(5),
                         np.array([-2, -1, 0, 1, 2]))
    assert np.allclose(new[:, 2], 0.)
    assert np.allclose(new[[0, 1], 1], -new[[4, 3], 1])
    assert abs(new[4, 1] / new[3, 1] - 2) < 0.00001


def test_grating_energy():
    """test_grating.py:113-128: n lambda = d sin(theta)"""
    d = np.tile([-1., 0, 0], (5, 1))
    e = np.arange(1., 6)
    new, _ = run_grating(grating_el(d=1. / 500), np.ones((5, 3)), d, e, np.ones(5, int))
    theta = np.arctan2(new[:, 1], new[:, 0])
    assert np.allclose(E2W / e, 1. / 500. * np.sin(theta))


def test_zero_order_passes():
    """test_grating.py:16-46: order 0 keeps the direction."""
    rng = np.random.default_rng(0)
    d = OT.unit(np.column_stack([-rng.uniform(1, 2, 10), rng.normal(0, .1, (10, 2))]))
    new, _ = run_grating(grating_el(d=1. / 500, zoom=[1, 50, 50]), np.ones((10, 3)), d,
                         rng.uniform(.5, 2, 10), np.zeros(10, int))
    assert np.allclose(new, d, atol=1e-14)


def test_cat_sign_convention():
    """test_grating.py:200-223: sign pattern and direction of +-5 orders."""
    dirs = np.array([[-1, 0., 0.], [-1, 0.01, -0.01], [-1, 0.01, 0.01],
                     [-1, -0.01, 0.01], [-1, -0.01, -0.01]])
    fr = OT.frame(np.eye(4))
    l, e_perp, n = OT.groove_frame(fr, 0.)
    s = np.sign(OT.dot(dirs, e_perp))
    s[s == 0] = 1
    assert np.all(s == [1, 1, 1, -1, -1])
    p5, _ = run_grating(grating_el(d=1. / 5000, sign='cat', zoom=2), np.ones((5, 3)), dirs,
                        np.ones(5), np.full(5, 5))
    m5, _ = run_grating(grating_el(d=1. / 5000, sign='cat', zoom=2), np.ones((5, 3)), dirs,
                        np.ones(5), np.full(5, -5))
    assert np.all(p5[1:3, 1] > 0) and np.all(p5[3:, 1] < 0)
    assert np.all(m5[1:3, 1] < 0) and np.all(m5[3:, 1] > 0)


def test_blaze_angle():
    """test_grating.py:130-141 passes floor(10 blaze) to a mock selector and
    expects [0, 7]: blaze 0 for normal incidence, 45 deg for dir (-1,1,0)."""
    d = np.array([[-1., 0, 0], [-1, 1, 0]])
    new, blaze = run_grating(grating_el(d=1., zoom=5), np.array([[1., 0, 0], [1, 0, 0]]), d,
                             np.ones(2), np.zeros(2, int))
    assert np.allclose(np.floor(blaze * 10), [0, 7])


def test_groove_angle_equivalence():
    """test_grating.py:143-176: groove_angle=.3 equals rotating the grating
    by .3 about x."""
    from marxs_amd.mathutils import axangle2mat
    d = np.tile([-1., 0, 0], (1, 1))
    e1 = grating_el(d=1. / 500, zoom=1, groove=.3)
    e2 = grating_el(d=1. / 500, zoom=1)
    e2['pos4d'] = np.eye(4)
    e2['pos4d'][:3, :3] = axangle2mat([1, 0, 0], .3)
    n1, _ = run_grating(e1, np.array([[1., 0, 0]]), d, np.ones(1), np.ones(1, int))
    n2, _ = run_grating(e2, np.array([[1., 0, 0]]), d, np.ones(1), np.ones(1, int))
    assert np.allclose(n1, n2, atol=5e-8)


def test_reflection_grating_flips_normal_component():
    d = np.tile([-1., 0, 0], (1, 1))
    t, _ = run_grating(grating_el(d=1. / 500), np.array([[1., 0, 0]]), d, np.ones(1), np.ones(1, int))
    r, _ = run_grating(grating_el(d=1. / 500, transmission=False), np.array([[1., 0, 0]]), d,
                       np.ones(1), np.ones(1, int))
    assert np.allclose(t[:, 1:], r[:, 1:]) and np.allclose(t[:, 0], -r[:, 0])


# --------------------------------------------------------- selectors
def test_orderselector_inverse_cdf_matches_numpy_choice():
    """numpy's legacy RandomState.choice(a, size, p) is the reference draw
    (grating.py:55-57); with the same seed, the first j with cdf[j] > u for
    u = random_sample() must give the same orders."""
    orders = np.array([-1, 0, 1, 3])
    p = np.array([.25, .35, .2, .1])
    n = 20000
    exp = np.random.RandomState(42).choice(orders, size=n, p=p / p.sum())
    u = np.random.RandomState(42).random_sample(n)
    got, prob, ok = OT.select_order(OT.selector_table(
        {'type': 'orderselector', 'orders': orders, 'p': p}), np.ones(n), u)
    assert ok.all()
    assert np.array_equal(got, exp)
    assert np.allclose(prob, 0.9)


def test_orderselector_probability():
    """test_grating.py:235-244"""
    tab = OT.selector_table({'type': 'orderselector', 'orders': [0, 1], 'p': [.3, .6]})
    u = np.random.default_rng(1).random(10000)
    o, prob, ok = OT.select_order(tab, np.ones(10000), u)
    assert set(o) == {0, 1}
    assert np.allclose(prob, 0.9)
    assert (o == 0).sum() < 0.8 * (o == 1).sum()


def test_efficiencyfile_known():
    """test_grating.py:269-282 and the per-photon loop of grating.py:90-96."""
    dat = np.loadtxt(StringIO(".5 .1 .1 .1 .4\n1. .1 .1 .1 .5\n1.5 0. .1 .0 .5"))
    sel = {'type': 'efficiency', 'energy': dat[:, 0], 'prob': dat[:, 1:],
           'orders': [1, 0, -1, -2]}
    tab = OT.selector_table(sel)
    o, prob, ok = OT.select_order(tab, np.array([.7, 1.1]), np.array([.3, .3]))
    assert np.allclose(prob, [.7, .8])
    u = np.random.default_rng(2).random(1000)
    o, prob, ok = OT.select_order(tab, 1.6 * np.ones(1000), u)
    assert np.allclose(prob, .6)
    assert set(o) == {-2, 0}
    assert (o == 0).sum() < (o == -2).sum()
    # per-photon restatement of the reference loop
    energy = np.random.default_rng(3).uniform(0.2, 1.8, 500)
    u = np.random.default_rng(4).random(500)
    o, prob, ok = OT.select_order(tab, energy, u)
    for i, e in enumerate(energy):
        row = np.argmin(np.abs(tab['energy'] - e))
        j = np.min(np.nonzero(tab['cdf'][row] > u[i]))
        assert o[i] == tab['orders'][j] and prob[i] == tab['tot'][row]


# ----------------------------------------------------------- apertures
def test_circle_aperture_radius_and_wedge():
    """marxs/optics/tests/test_apertures.py:46-57"""
    from marxs_amd.base import parse_position_keywords
    el = {'kind': 'aperture_circle', 'pos4d': parse_position_keywords({'zoom': 25}),
          'phi': [0, np.pi / 2], 'r_inner': 15}
    fr = OT.frame(el['pos4d'])
    u = np.random.default_rng(0).random((2, 500))
    x, y, q = OT.aperture_points(el, fr, u[0], u[1])
    assert np.all(q[:, 1] >= 0) and np.all(q[:, 2] >= 0)
    rad = np.linalg.norm(q, axis=1)
    assert np.all(rad >= 15. - 1e-12) and np.all(rad <= 25. + 1e-12)
    # closed form of the sampling map: r^2 uniform in [ri^2, 1]
    r = np.hypot(x, y)
    assert np.allclose(r ** 2, (15 / 25) ** 2 + (1 - (15 / 25) ** 2) * u[1])


def test_projected_area_probability():
    """test_apertures.py:89-112: photons from ra = 0,45,90,135,315 deg."""
    ra = np.deg2rad([0., 45., 90., 135., 315.])
    d = -np.column_stack([np.cos(ra), np.sin(ra), np.zeros(5)])
    ph = {'dir': d, 'pol': np.tile([0., 0, 1], (5, 1)), 'energy': np.ones(5), 'prob': np.ones(5)}
    el = {'kind': 'aperture_rect', 'pos4d': np.eye(4)}
    res = OT.trace([el], ph, np.random.default_rng(0).random((2, 5)))
    assert np.allclose(res['prob'], [1., 1. / np.sqrt(2), 0, 0, 1. / np.sqrt(2)])
    el['pos4d'] = np.eye(4)
    el['pos4d'][:3, :3] = np.array([[0, 0, 1], [1, 0, 0], [0, 1, 0]])
    res = OT.trace([el], ph, np.random.default_rng(0).random((2, 5)))
    assert np.allclose(res['prob'], [0, 1. / np.sqrt(2), 1., 1. / np.sqrt(2), 0.])


# --------------------------------------------------------------- lenses
@pytest.mark.parametrize('ra', [0., 30., -60.])
def test_perfect_lens_focus(ra):
    """marxs/optics/tests/test_mirror.py:14-40"""
    from marxs_amd.base import parse_position_keywords
    a = np.deg2rad(ra)
    n = 11
    d = np.tile([-np.cos(a), 0, -np.sin(a)], (n, 1))
    ph = {'dir': d, 'pol': np.tile([0., 1, 0], (n, 1)), 'energy': np.ones(n), 'prob': np.ones(n)}
    u = np.random.default_rng(1).random((2, n))
    f = 1000.
    els = [{'kind': 'aperture_rect', 'pos4d': parse_position_keywords({'zoom': 2, 'position': [100, 0, 0]})},
           {'kind': 'lens_perfect', 'pos4d': parse_position_keywords({'zoom': 400}), 'focallength': f}]
    res = OT.trace(els, ph, u)
    assert np.allclose(res['pos'][:, 0], 0.)
    dd = f * np.cos(a)
    for dist, small in ((dd, True), (2 * dd, False)):
        det = {'kind': 'detector', 'pos4d': parse_position_keywords({'zoom': 1e5, 'position': [-dist, 0, 0]}),
               'pixsize': 0.01}
        r2 = OT.trace([det], {k: res[k] for k in ('pos', 'dir', 'pol', 'energy', 'prob')}, u)
        if small:
            assert np.std(r2['det_x']) < 1e-4 and np.std(r2['det_y']) < 1e-4
        else:
            assert np.std(r2['det_x']) > 1e-4 and np.std(r2['det_y']) > 1e-4


def test_perfect_lens_offset():
    """test_mirror.py:43-59"""
    from marxs_amd.base import parse_position_keywords
    ph = {'pos': np.array([[1., 0, 0]]), 'dir': np.array([[-1., 0, 0]]),
          'pol': np.array([[0., 1, 0]]), 'energy': np.ones(1), 'prob': np.ones(1)}
    l1 = {'kind': 'lens_perfect', 'pos4d': parse_position_keywords({'zoom': 400, 'position': [0, 0, 10]}),
          'focallength': 100}
    l2 = {'kind': 'lens_perfect', 'pos4d': parse_position_keywords({'zoom': 400}),
          'focallength': 100, 'd_center_optax': -10}
    l3 = {'kind': 'lens_perfect', 'pos4d': parse_position_keywords({'zoom': 400}),
          'focallength': 100, 'd_center_optax': 0}
    r1 = OT.trace([l1], ph, np.zeros((0, 1)))
    r2 = OT.trace([l2], ph, np.zeros((0, 1)))
    r3 = OT.trace([l3], ph, np.zeros((0, 1)))
    assert np.allclose(r1['dir'], r2['dir'])
    assert not np.allclose(r1['dir'], r3['dir'])


def _refl_spline(f, x=None, y=None):
    from scipy.interpolate import RectBivariateSpline
    x = np.linspace(-3, 3, 100) if x is None else x
    y = np.linspace(-3, 3, 100) if y is None else y
    xg, yg = np.meshgrid(x, y, indexing='ij')
    return RectBivariateSpline(x, y, f(xg, yg), kx=1, ky=1)


def test_perfect_lens_reflectivity_known_answer():
    """test_mirror.py:62-80: exp(-sqrt((E/2)^2 + a^2)) as a degree-1 spline
    on [-3, 3]^2, one test photon (1 keV, along -x through the centre, so the
    angle is 0): probability 0.367205."""
    from marxs_amd.base import parse_position_keywords
    interp = _refl_spline(lambda x, y: np.exp(-np.sqrt((x / 2) ** 2 + y ** 2)))
    ph = {'pos': np.array([[1., 0, 0]]), 'dir': np.array([[-1., 0, 0]]),
          'pol': np.array([[0., 1, 0]]), 'energy': np.ones(1), 'prob': np.ones(1)}
    el = {'kind': 'lens_perfect', 'pos4d': parse_position_keywords({}), 'focallength': 123.,
          'reflectivity': interp}
    res = OT.trace([el], ph, np.zeros((0, 1)))
    assert np.allclose(res['energy'], 1)
    assert np.allclose(res['pol'], [0, 1, 0])
    assert np.allclose(res['prob'], 0.367205)


def test_perfect_lens_reflectivity_angle_quarter_squared():
    """Closed form: rays parallel to the axis at height r turn by
    atan(r / f) towards the focus; a reflectivity linear in the angle (which
    a degree-1 spline represents exactly) gives probability
    (1 - 2 atan(r / f) / 4)^2 times the energy factor."""
    from marxs_amd.base import parse_position_keywords
    f = 1000.
    r = np.array([0., 10., 50., 100., 200.])
    n = len(r)
    interp = _refl_spline(lambda e, a: (1 - 0.1 * e) * (1 - 2 * a),
                          x=np.linspace(0., 5., 11), y=np.linspace(0., 0.5, 6))
    e = np.array([0.5, 1., 1.5, 2., 4.])
    ph = {'pos': np.column_stack([np.full(n, 10.), r, np.zeros(n)]),
          'dir': np.tile([-1., 0, 0], (n, 1)), 'pol': np.tile([0., 0, 1], (n, 1)),
          'energy': e, 'prob': np.ones(n)}
    el = {'kind': 'lens_perfect', 'pos4d': parse_position_keywords({'zoom': 300}), 'focallength': f,
          'reflectivity': interp}
    res = OT.trace([el], ph, np.zeros((0, n)))
    ang = np.arctan(r / f)
    assert np.allclose(res['prob'], ((1 - 0.1 * e) * (1 - 2 * ang / 4)) ** 2, rtol=1e-12)
    # a planted mistake: the full angle instead of a quarter of it
    assert not np.allclose(res['prob'][1:], ((1 - 0.1 * e) * (1 - 2 * ang))[1:] ** 2, rtol=1e-3)


def test_thin_lens_paraxial_focus():
    """Thin-lens rotation by rho/f toward the axis: a ray at height rho
    crosses the axis at rho / tan(rho / f) behind the lens."""
    from marxs_amd.base import parse_position_keywords
    f = 1000.
    rho = np.array([0., 1., 5., 20.])
    n = len(rho)
    ph = {'pos': np.column_stack([np.full(n, 10.), rho, np.zeros(n)]),
          'dir': np.tile([-1., 0, 0], (n, 1)), 'pol': np.tile([0., 0, 1], (n, 1)),
          'energy': np.ones(n), 'prob': np.ones(n)}
    el = {'kind': 'lens_thin', 'pos4d': parse_position_keywords({'zoom': 100}), 'focallength': f}
    res = OT.trace([el], ph, np.zeros((0, n)))
    assert np.allclose(res['dir'][0], [-1, 0, 0])
    slope = res['dir'][1:, 1] / -res['dir'][1:, 0]
    assert np.allclose(-slope, np.tan(rho[1:] / f))
    assert np.allclose(np.linalg.norm(res['dir'], axis=1), 1)
    assert np.allclose(OT.dot(res['dir'], res['pol']), 0)


# ------------------------------------------------------------ detector
def test_detector_pixels():
    """marxs/optics/tests/test_detector.py:16-35"""
    from marxs_amd.base import parse_position_keywords
    pos = np.array([[0, 0., -0.25], [0., 9.9, 1.], [0., 10.1, 1.]])
    ph = {'pos': pos, 'dir': np.ones((3, 3)), 'pol': np.tile([0., 1, -1], (3, 1)),
          'energy': np.ones(3), 'prob': np.ones(3)}
    el = {'kind': 'detector', 'pos4d': parse_position_keywords({'zoom': [1., 10., 5.]}),
          'pixsize': 0.5}
    npix, cpix = OT.detector_params(el)
    assert npix == [40, 20] and cpix == [19.5, 9.5]
    res = OT.trace([el], ph, np.zeros((0, 3)))
    nan = np.nan
    for name, exp in (('det_x', [0, 9.9, nan]), ('det_y', [-0.25, 1., nan]),
                      ('detpix_x', [19.5, 39.3, nan]), ('detpix_y', [9, 11.5, nan])):
        assert np.allclose(res[name], exp, equal_nan=True), name
    el = {'kind': 'detector', 'pos4d': np.eye(4), 'pixsize': 0.05}
    assert OT.detector_params(el)[0] == [40, 40]


# ---------------------------------------------------------- placement
def test_placement_compose_equals_product():
    """T . R . Z composition (marxs/math/tests/test_utils.py:7-28)."""
    rng = np.random.default_rng(0)
    for _ in range(5):
        t = rng.random(3)
        z = rng.random(3)
        rot = PL.euler_sxyz(*rng.uniform(-1, 1, 3))
        exp = PL.translation(t) @ PL.linear(rot) @ PL.linear(np.diag(z))
        assert np.allclose(PL.compose(t, rot, z), exp)


# ------------------------------------------- interpolated efficiency table
def _interp_table(k=1):
    import os
    from marxs_amd.optics import InterpolateEfficiencyTable
    path = os.path.join(os.path.dirname(__file__), 'golden', 'grating_efficiency.csv')
    return InterpolateEfficiencyTable(path, k=k)


def _interp_desc(t):
    return OT.selector_table({'type': 'interp', 'wave': t.wave, 'theta': t.theta,
                              'orders': t.orders, 'data': t.data, 'k': t.k})


def test_interp_table_known_answers():
    """marxs/missions/mitsnl/tests/test_cat.py:98-106 (table data file of the
    reference's tests), at the grid nodes where every spline degree agrees:
    totals [.8, .8, 1, 1], order-0 probabilities [.3, .3, .2] (the first
    column there is the first three orders at 0.5 keV, 1 deg)."""
    t = _interp_table()
    assert list(t.orders) == [0, -1, -2]
    p = OT.interp_probabilities(_interp_desc(t), np.array([0.5, 0.5, 1., 1.]),
                                np.deg2rad([1., 2., 1., 2.]))
    assert np.allclose(p.sum(axis=0), [.8, .8, 1., 1.], rtol=1e-4)
    assert np.allclose(p[:, 0], [.3, .3, .2], rtol=1e-4)


def test_interp_matches_scipy_rectbivariatespline():
    """Between and beyond the nodes: the restated degree-1 evaluation against
    scipy's RectBivariateSpline(kx=1, ky=1).ev (FITPACK), the reference's
    interpolator (catgrating.py:102,134)."""
    from scipy.interpolate import RectBivariateSpline
    rng = np.random.default_rng(4)
    wave = np.array([0.6, 0.9, 1.3, 2.0, 2.5, 4.1])
    theta = np.deg2rad([0.5, 1.0, 1.7, 2.2])
    data = rng.uniform(0, 0.3, (5, len(wave), len(theta)))
    tab = OT.selector_table({'type': 'interp', 'wave': wave, 'theta': theta,
                             'orders': np.arange(5), 'data': data, 'k': 1})
    e = OT.KEV_NM / rng.uniform(0.3, 5., 3000)        # includes both edges
    b = rng.uniform(0., np.deg2rad(3.), 3000)
    got = OT.interp_probabilities(tab, OT.KEV_NM / e, b)
    for m in range(5):
        ref = RectBivariateSpline(wave, theta, data[m], kx=1, ky=1).ev(e, b)
        assert np.allclose(got[m], ref, rtol=1e-13, atol=1e-15)


def test_interp_order_rule_and_fallback():
    """catgrating.py:132-144: first order whose normalised cumulative
    probability exceeds u; the first order when none does (u = 1 can only
    happen with rounding), probability factor = total."""
    t = _interp_table()
    tab = _interp_desc(t)
    e = np.full(4, 1.0)
    b = np.deg2rad(np.full(4, 1.))
    o, prob, ok = OT.select_order(tab, e, np.array([0.1, 0.55, 0.85, 1.0]), b)
    # probabilities at (1.2398 nm ~ 1 keV, 1 deg): .5, .3, .2
    assert list(o) == [0, -1, -2, 0]
    assert ok.all() and np.allclose(prob, 1.0, rtol=1e-4)


def test_interp_table_wrong_format():
    """test_cat.py:120-124: a table with missing rows is rejected."""
    import os
    from marxs_amd.optics import InterpolateEfficiencyTable, DataFileFormatException
    path = os.path.join(os.path.dirname(__file__), 'golden', 'grating_efficiency_broken.csv')
    with pytest.raises(DataFileFormatException):
        InterpolateEfficiencyTable(path)
