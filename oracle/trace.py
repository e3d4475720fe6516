"""numpy float64 restatement of the traced path (TEST INFRASTRUCTURE ONLY).

Instrument description: a list of plain dicts, one per leaf element in
execution order (see ``element dicts`` below).  Photons: dict of numpy
arrays ``pos`` (N,3) (optional), ``dir`` (N,3), ``pol`` (N,3), ``energy``
(N,), ``prob`` (N,).  Uniforms: (n_slots, N).

Each function names the reference code it restates.  Arithmetic is written
out component by component in a fixed left-to-right order (the reference's
``np.dot``/``einsum`` leave the summation order to BLAS); numpy never fuses
multiply-adds, so these are plain IEEE double operations.

How each part is pinned (tests/test_oracle.py, tests/test_oracle_driver.py,
tests/test_setup.py; committed vectors in tests/golden):

* ``intersect`` -- known answers of ``marxs/math/tests/test_geometry.py:11-55``
  and a brute-force check against the closed-form ray-plane solution.
* ``parallel_transport`` -- ``marxs/math/tests/test_polarization.py:36-72``
  invariants (identity for equal directions, s component kept, unit length,
  perpendicular to the new direction) and a Rodrigues-rotation closed form.
* ``diffract`` / ``select_order`` -- grating equation known answers of
  ``marxs/optics/tests/test_grating.py`` (``sin a = sin b + lambda/d`` at
  0/5/10/20 deg, order symmetry, CAT sign pattern [1,1,1,-1,-1],
  EfficiencyFile totals [0.7, 0.8] and 0.6), and numpy's own
  ``RandomState.choice`` run with a fixed seed for the inverse-CDF rule.
* ``aperture_*`` -- ``test_apertures.py`` radius/quadrant/projected-area
  answers ``[1, 1/sqrt2, 0, 0, 1/sqrt2]`` and the closed-form sampling map.
* ``perfect_lens`` -- ``test_mirror.py`` focus answers (spread < 1e-4 at
  ``d = f cos(ra)``; off-axis equivalence).
* ``thin_lens`` -- focusing of paraxial rays (closed form) and unit
  rotation invariants.
* ``detector`` -- ``test_detector.py:16-35`` pixel answers.
* ``stack_layers`` (tests/test_oracle_stack.py) -- CAT support answers of
  ``marxs/missions/mitsnl/tests/test_cat.py`` (Si transmission 0.22458 of
  4 um at 1 keV; L2 open fraction 0.81 on axis, smaller off axis; L2
  broadening ~0.5 arcsec at 0.25 keV; NonParallelCATGrating == CATGrating
  bit for bit), the L1 table interpolation against scipy's ``interp1d``,
  the scatter answers of ``marxs/optics/tests/test_scatter.py`` (widths
  arctan(sigma) on a detector 1 mm away; exact zeros for sigma 0 and for the
  coordinate a rotation cannot reach), the transposed axis-angle matrix
  against scipy's ``Rotation``, and invariants (unit directions,
  polarization perpendicular, probabilities in [0, 1]).
* ``catsupportbars`` -- the rule itself (probability 0 where the facet id is
  negative), checked on a hand-made table.
"""
import math

import numpy as np

E2W = 1.2398419292004202e-06   # keV -> mm, marxs/__init__.py:14
ORDER_NONE = -2 ** 31

APERTURES = ('aperture_circle', 'aperture_rect')


# ------------------------------------------------------------------ helpers
def dot(a, b):
    """Row-wise a . b for (N,3) a and (3,) or (N,3) b, fixed order."""
    if b.ndim == 1:
        return a[:, 0] * b[0] + a[:, 1] * b[1] + a[:, 2] * b[2]
    return a[:, 0] * b[:, 0] + a[:, 1] * b[:, 1] + a[:, 2] * b[:, 2]


def unit(v):
    """norm_vector, marxs/math/utils.py:150-164."""
    n = np.sqrt(v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1] + v[:, 2] * v[:, 2])
    return v / n[:, None]


def cross(a, b):
    """np.cross component formulas."""
    return np.stack([a[:, 1] * b[:, 2] - a[:, 2] * b[:, 1],
                     a[:, 2] * b[:, 0] - a[:, 0] * b[:, 2],
                     a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]], axis=1)


def frame(pos4d):
    """Geometry.__getitem__ (marxs/math/geometry.py:108-160) for a flat
    element: centre, unit axes and half sizes."""
    pos4d = np.asarray(pos4d, dtype=float)
    cols = [pos4d[:, i] for i in range(3)]
    e = [c / np.linalg.norm(c) for c in cols]
    return {'center': pos4d[:3, 3].copy(), 'e_x': e[0][:3], 'e_y': e[1][:3], 'e_z': e[2][:3],
            'v_y': cols[1][:3].copy(), 'v_z': cols[2][:3].copy(),
            'hy': float(np.linalg.norm(cols[1])), 'hz': float(np.linalg.norm(cols[2]))}


def zooms(pos4d):
    """Column lengths of the linear part after orthogonalisation (QR)."""
    q, r = np.linalg.qr(np.asarray(pos4d, dtype=float)[:3, :3])
    return np.abs(np.diag(r))


# ----------------------------------------------------------- intersection
def intersect(fr, pos, dir):
    """FinitePlane.intersect, marxs/math/geometry.py:211-261."""
    c = fr['center']
    kn = dot(c[None, :] - pos, fr['e_x'])
    kd = dot(dir, fr['e_x'])
    with np.errstate(divide='ignore', invalid='ignore'):
        k = kn / kd
        q = pos + k[:, None] * dir
        v = q - c[None, :]
        ly = dot(v, fr['e_y'])
        lz = dot(v, fr['e_z'])
        hit = (kd != 0) & (k >= 0) & (np.abs(ly) <= fr['hy']) & (np.abs(lz) <= fr['hz'])
    return hit, q, ly, lz


# ------------------------------------------------------------ polarization
def parallel_transport(d1, d2, pol):
    """paralleltransport_matrix with J = I and parallel_transport,
    marxs/math/polarization.py:90-170.  Rows with |d1 x d2| <= 1e-8
    (np.isclose to 0) keep their polarization."""
    a = unit(d1)
    b = unit(d2)
    s = cross(a, b)
    with np.errstate(invalid='ignore', divide='ignore'):
        sn = np.sqrt(s[:, 0] * s[:, 0] + s[:, 1] * s[:, 1] + s[:, 2] * s[:, 2])
        same = sn <= 1e-8
        s = s / sn[:, None]
        pin = cross(a, s)
        pout = cross(b, s)
        out = np.empty_like(pol)
        # pmat[j][l] = s_j s_l + pout_j pin_l + b_j a_l
        m = np.empty((len(pol), 3, 3))
        for j in range(3):
            for l in range(3):
                m[:, j, l] = s[:, j] * s[:, l] + pout[:, j] * pin[:, l] + b[:, j] * a[:, l]
        for j in range(3):
            out[:, j] = m[:, j, 0] * pol[:, 0] + m[:, j, 1] * pol[:, 1] + m[:, j, 2] * pol[:, 2]
    out[same] = pol[same]
    return out


# -------------------------------------------------------- order selection
def selector_table(sel):
    """CDF table of an order selector description.

    ``{'type': 'orderselector', 'orders', 'p'}`` follows OrderSelector
    (grating.py:39-57) and numpy's legacy ``choice`` (cdf = cumsum(p/sum p),
    renormalised so that the last entry is 1).  ``{'type': 'efficiency',
    'energy', 'prob', 'orders'}`` follows EfficiencyFile (grating.py:78-96).
    """
    if sel['type'] == 'interp':
        if int(sel.get('k', 1)) != 1:
            raise ValueError('only degree-1 interpolation tables are restated')
        return {'type': 'interp', 'wave': np.asarray(sel['wave'], dtype=float),
                'theta': np.asarray(sel['theta'], dtype=float),
                'orders': np.asarray(sel['orders'], dtype=np.int64),
                'data': np.asarray(sel['data'], dtype=float)}
    if sel['type'] == 'orderselector':
        p = np.asarray(sel['p'], dtype=float)
        p_sum = p.sum()
        cdf = np.cumsum(p / p_sum)
        cdf = cdf / cdf[-1]
        return {'energy': None, 'cdf': cdf[None, :], 'tot': np.array([p_sum]),
                'orders': np.asarray(sel['orders'], dtype=np.int64)}
    prob = np.asarray(sel['prob'], dtype=float)
    tot = np.sum(prob, axis=1)
    return {'energy': np.asarray(sel['energy'], dtype=float),
            'cdf': np.cumsum(prob, axis=1) / tot[:, None], 'tot': tot,
            'orders': np.asarray(sel['orders'], dtype=np.int64)}


KEV_NM = 1.2398419843320026   # h c / keV in nm (SI-exact h, c, e)


def _hat(grid, x):
    """Degree-1 B-spline basis on ``grid`` (knots = grid points, doubled at
    the ends) at x clamped to the grid: interval index i and the two basis
    values f (g[i+1] - x), f (x - g[i]) with f = 1 / (g[i+1] - g[i]), as
    FITPACK's fpbspl forms them."""
    g = np.asarray(grid, dtype=float)
    x = np.minimum(np.maximum(x, g[0]), g[-1])
    i = np.clip(np.searchsorted(g, x, side='right') - 1, 0, len(g) - 2)
    f = 1.0 / (g[i + 1] - g[i])
    return i, f * (g[i + 1] - x), f * (x - g[i])


def interp_probabilities(tab, energy, blaze):
    """InterpolateEfficiencyTable.probabilities (catgrating.py:97-130) for
    k = 1: per order the tensor-product degree-1 spline through the grid
    values at (wavelength, blaze), constant beyond the edges; terms added in
    FITPACK's fpbisp order (c00 hx0 hy0, c01 hx0 hy1, c10 hx1 hy0, c11 hx1 hy1).
    Returns (n_orders, n) probabilities."""
    wave = KEV_NM / np.asarray(energy, dtype=float)
    i, hx0, hx1 = _hat(tab['wave'], wave)
    j, hy0, hy1 = _hat(tab['theta'], np.asarray(blaze, dtype=float))
    d = tab['data']
    out = np.empty((d.shape[0], len(wave)))
    for m in range(d.shape[0]):
        sp = np.zeros(len(wave))
        sp = sp + d[m, i, j] * hx0 * hy0
        sp = sp + d[m, i, j + 1] * hx0 * hy1
        sp = sp + d[m, i + 1, j] * hx1 * hy0
        sp = sp + d[m, i + 1, j + 1] * hx1 * hy1
        out[m] = sp
    return out


def select_order(tab, energy, u, blaze=None):
    """Nearest energy row (first on ties), then the first order whose
    cumulative probability exceeds u.  Returns (order, prob, ok).

    For an interpolated table (catgrating.py:132-144) the probabilities come
    from :func:`interp_probabilities`, the cumulative sum is divided by the
    total, and the first order is taken when no entry exceeds u (argmax of
    an all-false row); ``ok`` is then always true."""
    if tab.get('type') == 'interp':
        p = interp_probabilities(tab, energy, blaze)
        total = np.sum(p, axis=0)
        with np.errstate(invalid='ignore', divide='ignore'):
            cum = np.cumsum(p, axis=0) / total
        j = np.argmax(cum > u[None, :], axis=0)
        return tab['orders'][j], total, np.ones(len(u), dtype=bool)
    if tab['energy'] is None or len(tab['energy']) == 1:
        row = np.zeros(len(energy), dtype=int)
    else:
        row = np.argmin(np.abs(tab['energy'][None, :] - energy[:, None]), axis=1)
    cdf = tab['cdf'][row]
    above = cdf > u[:, None]
    ok = above.any(axis=1)
    j = np.argmax(above, axis=1)
    order = np.where(ok, tab['orders'][j], ORDER_NONE)
    return order, tab['tot'][row], ok


def select_order_loop(tab, energy, u):
    """EfficiencyFile.__call__ as the reference writes it (grating.py:90-96):
    one photon at a time, the nearest energy row by argmin, then the smallest
    index whose cumulative probability exceeds the photon's draw (the
    injected uniform in place of np.random.rand()).  Same results as
    :func:`select_order` for a tabulated selector; the CPU baseline times
    both."""
    n = len(energy)
    order = np.empty(n, dtype=np.int64)
    tot = np.empty(n)
    ok = np.ones(n, dtype=bool)
    eg = tab['energy']
    for i, e in enumerate(energy):
        r = 0 if eg is None or len(eg) == 1 else int(np.argmin(np.abs(eg - e)))
        nz = np.nonzero(tab['cdf'][r] > u[i])[0]
        if len(nz) == 0:
            ok[i] = False
            order[i] = ORDER_NONE
        else:
            order[i] = tab['orders'][np.min(nz)]
        tot[i] = tab['tot'][r]
    return order, tot, ok


# ----------------------------------------------------------------- elements
def groove_frame(fr, groove_angle):
    """FlatGrating.e_groove_coos, grating.py:186-207."""
    ex, ey, en = fr['e_y'], fr['e_z'], fr['e_x']
    e_groove = math.sin(-groove_angle) * ex + math.cos(-groove_angle) * ey
    e_perp = math.cos(-groove_angle) * ex - math.sin(-groove_angle) * ey
    return e_groove, e_perp, en


def blaze_angle(el, fr, dir, ly):
    """The blaze angle of grating.py:240-249 (as computed in :func:`diffract`)."""
    p = unit(dir)
    l, e_perp, n = groove_frame(fr, el.get('groove_angle', 0.))
    with np.errstate(invalid='ignore', divide='ignore'):
        pl = dot(p, l)
        pperp = unit(p - pl[:, None] * l[None, :])
        blaze = np.arccos(np.clip(np.abs(dot(pperp, n)), 0, 1))
        return blaze + (el.get('blaze_center', 0.) + ly * el.get('d_blaze_mm', 0.))


def diffract(el, fr, dir, energy, ly, order, ptot):
    """FlatGrating.diffract_photons, grating.py:233-266 (orders given).
    Returns new direction and blaze angle."""
    p = unit(dir)
    l, e_perp, n = groove_frame(fr, el.get('groove_angle', 0.))
    dg = -e_perp
    wave = E2W / energy
    with np.errstate(invalid='ignore', divide='ignore'):
        pl = dot(p, l)
        pperp = unit(p - pl[:, None] * l[None, :])
        blaze = np.arccos(np.clip(np.abs(dot(pperp, n)), 0, 1))
        blaze = blaze + (el.get('blaze_center', 0.) + ly * el.get('d_blaze_mm', 0.))
        pd0 = dot(p, dg)
        if el.get('sign', 'flat') == 'cat':
            sign = np.sign(pd0)
            sign[sign == 0] = 1
        else:
            sign = -np.ones(len(pd0))
        p_d = pd0 + ((sign * order) * wave) / el['d']
        p_l = pl
        p_n = np.sqrt(1. - p_d * p_d - p_l * p_l)
        direction = np.sign(dot(p, n))
        if not el.get('transmission', True):
            direction = direction * -1
        w = direction * p_n
        new = (p_d[:, None] * dg[None, :] + p_l[:, None] * l[None, :]) + w[:, None] * n[None, :]
    return new, blaze


def aperture_points(el, fr, u0, u1):
    """CircleAperture / RectangleAperture generate_local_xy and
    FlatAperture.specific_process_photons (aperture.py:68-78,92-95,138-146)."""
    if el['kind'] == 'aperture_circle':
        phi0, phi1 = el.get('phi', [0, 2 * np.pi])
        phi = phi0 + (phi1 - phi0) * u0
        rin = el.get('r_inner', 0.) / np.linalg.norm(np.asarray(el['pos4d'])[:, 1])
        low = rin ** 2
        r = np.sqrt(low + (1. - low) * u1)
        x = r * np.cos(phi)
        y = r * np.sin(phi)
    else:
        x = u0 * 2. - 1.
        y = u1 * 2. - 1.
    c, vy, vz = fr['center'], fr['v_y'], fr['v_z']
    q = c[None, :] + x[:, None] * vy[None, :] + y[:, None] * vz[None, :]
    return x, y, q


def perfect_lens(el, fr, dir, q, pol, energy=None):
    """PerfectLens.specific_process_photons, mirror.py:53-82.  Returns the
    new direction, polarization and probability factor: the element's
    ``reflectivity`` callable (the reference's ``reflectivity_interpolator``,
    called as the reference calls it, with ``grid=False``) at the energy and
    a quarter of the angle between the new and the old direction, squared;
    1 without one."""
    pos4d = np.asarray(el['pos4d'], dtype=float)
    e_z4 = pos4d[:, 2] / np.linalg.norm(pos4d[:, 2])
    popt = (pos4d[:, 3] - el.get('d_center_optax', 0.) * e_z4)[:3]
    dn = unit(dir)
    fp = popt[None, :] + el['focallength'] * dn
    new = unit(fp - q)
    refl = el.get('reflectivity')
    if refl is None:
        prob = np.ones(len(new))
    else:
        angle = np.arccos(np.abs(new[:, 0] * dn[:, 0] + new[:, 1] * dn[:, 1] + new[:, 2] * dn[:, 2]))
        prob = np.asarray(refl(energy, angle / 4, grid=False), dtype=float) ** 2
    return new, parallel_transport(dir, new, pol), prob


def thin_lens(el, fr, dir, q, pol):
    """ThinLens.process_photon, mirror.py:130-148, with the axis-angle matrix
    of transforms3d.axangles.axangle2mat in its operation order."""
    c = fr['center']
    r = q - c[None, :]
    dist = np.sqrt(r[:, 0] * r[:, 0] + r[:, 1] * r[:, 1] + r[:, 2] * r[:, 2])
    new_d = dir.copy()
    new_p = pol.copy()
    mv = dist != 0
    if mv.any():
        ang = -dist[mv] / el['focallength']
        ax = cross(dir[mv], r[mv])
        an = np.sqrt(ax[:, 0] * ax[:, 0] + ax[:, 1] * ax[:, 1] + ax[:, 2] * ax[:, 2])
        x, y, z = ax[:, 0] / an, ax[:, 1] / an, ax[:, 2] / an
        cc, ss = np.cos(ang), np.sin(ang)
        C = 1 - cc
        xs, ys, zs = x * ss, y * ss, z * ss
        xC, yC, zC = x * C, y * C, z * C
        xyC, yzC, zxC = x * yC, y * zC, z * xC
        m = [[x * xC + cc, xyC - zs, zxC + ys],
             [xyC + zs, y * yC + cc, yzC - xs],
             [zxC - ys, yzC + xs, z * zC + cc]]
        dm, pm = dir[mv], pol[mv]
        for i in range(3):
            new_d[mv, i] = m[i][0] * dm[:, 0] + m[i][1] * dm[:, 1] + m[i][2] * dm[:, 2]
            new_p[mv, i] = m[i][0] * pm[:, 0] + m[i][1] * pm[:, 1] + m[i][2] * pm[:, 2]
    return new_d, new_p


def angle_between(angle, border1, border2):
    """marxs/math/utils.py:180-215: angle in the cyclic interval
    [border1, border2] (all taken modulo 2 pi with numpy's floor modulo)."""
    twopi = 2 * np.pi
    b1 = (twopi + (border1 % twopi)) % twopi
    b2 = (twopi + (border2 % twopi)) % twopi
    ang = (twopi + (angle % twopi)) % twopi
    if b1 < b2:
        return (b1 <= ang) & (ang <= b2)
    return (b1 <= ang) | (ang <= b2)


def zooms_gram_schmidt(pos4d):
    """Zooms of transforms3d.affines.decompose44 (used by the reference's
    Cylinder, geometry.py:428-432,556-562): Gram-Schmidt on the columns of
    the linear part; x zoom negative for a reflection."""
    m = np.array(pos4d, dtype=float)[:3, :3]
    m0, m1, m2 = m[:, 0].copy(), m[:, 1].copy(), m[:, 2].copy()
    sx = math.sqrt(np.sum(m0 ** 2))
    m0 /= sx
    m1 -= np.dot(m0, m1) * m0
    sy = math.sqrt(np.sum(m1 ** 2))
    m1 /= sy
    m2 -= (np.dot(m0, m2) * m0 + np.dot(m1, m2) * m1)
    sz = math.sqrt(np.sum(m2 ** 2))
    m2 /= sz
    if np.linalg.det(np.column_stack([m0, m1, m2])) < 0:
        sx = -sx
    return sx, sy, sz


def _affine_apply(m, v, w):
    """np.dot(m, e2h(v, w).T).T, summed in index order."""
    return np.stack([((m[j, 0] * v[:, 0] + m[j, 1] * v[:, 1]) + m[j, 2] * v[:, 2]) + m[j, 3] * w
                     for j in range(3)], axis=1)


def cylinder_intersect(el, pos, dir):
    """Cylinder.intersect, marxs/math/geometry.py:495-564: in the local
    frame (inverse pose) solve a t^2 + b t + c = 0 for the unit circle; both
    roots that are forward, within |z| <= 1 and within the azimuth limits
    are hits, the nearer wins.  Returns hit, global point, phi, z (times
    the z zoom)."""
    P = np.array(el['pos4d'], dtype=float)
    inv = np.linalg.inv(P)
    xyz = _affine_apply(inv, pos, 1.)
    d = _affine_apply(inv, dir, 0.)
    x, y = xyz[:, 0], xyz[:, 1]
    rx, ry = d[:, 0], d[:, 1]
    with np.errstate(invalid='ignore', divide='ignore'):
        c = (x * x + y * y) - 1.
        b = 2 * (x * rx + y * ry)
        a = rx * rx + ry * ry
        under = b ** 2 - 4 * a * c
        denom = 2 * a
        root = np.sqrt(under)
        a1 = (-b + root) / denom
        a2 = (-b - root) / denom
        phi1 = np.arctan2(y + a1 * ry, x + a1 * rx)
        phi2 = np.arctan2(y + a2 * ry, x + a2 * rx)
        z1 = xyz[:, 2] + a1 * d[:, 2]
        z2 = xyz[:, 2] + a2 * d[:, 2]
        lim = el['phi_lim']
        h1 = (under >= 0) & (a1 >= 0) & (np.abs(z1) <= 1.) & angle_between(phi1, lim[0], lim[1])
        h2 = (under >= 0) & (a2 >= 0) & (np.abs(z2) <= 1.) & angle_between(phi2, lim[0], lim[1])
    h1 = h1 & ~(h2 & (a2 < a1))
    h2 = h2 & ~(h1 & (a2 >= a1))
    hit = h1 | h2
    t = np.where(h1, a1, a2)
    phi = np.where(h1, phi1, phi2)
    z = np.where(h1, z1, z2) * zooms_gram_schmidt(P)[2]
    ql = np.stack([x + t * rx, y + t * ry, xyz[:, 2] + t * d[:, 2]], axis=1)
    q = _affine_apply(P, ql, 1.)
    return hit, q, phi, z


def detector_params(el):
    """FlatDetector.__init__, detector.py:59-70."""
    z = zooms(el['pos4d'])
    npix = [int(np.round(2. * z[i + 1] / el['pixsize'])) for i in (0, 1)]
    centerpix = [(n - 1) / 2 for n in npix]
    return npix, centerpix


def pixel_index(detpix, npix):
    f = np.floor(detpix + 0.5)
    return np.where((f >= 0) & (f < npix), f, -1).astype(np.int64)


# ------------------------------------------------------------------ driver
STACK_DRAWS = 5   # draws per stack interaction: order, L1 order, L1 open, normal, angle


def stack_loc_pairs(el):
    """Distinct local-coordinate column pairs of a stack (its own first,
    then those of its layers in order)."""
    pairs = []
    for loc in [el.get('loc')] + [ly.get('loc') for ly in el['layers']]:
        if loc is not None and tuple(loc) not in pairs:
            pairs.append(tuple(loc))
    return pairs


def columns_of(el):
    kind = el['kind']
    cols = {'loc0': None, 'loc1': None, 'id': el.get('id_col'), 'order': None,
            'blaze': None, 'detpix0': None, 'detpix1': None, 'pix0': None, 'pix1': None}
    if kind == 'catsupportbars':
        return {k: None for k in cols}
    if kind == 'stack':
        cols.update({'loc0b': None, 'loc1b': None, 'order2': None, 'blaze2': None, 'scat': None,
                     'scat2': None})
        pairs = stack_loc_pairs(el)
        if len(pairs) > 2:
            raise ValueError('more than two local-coordinate column pairs in a stack')
        if pairs:
            cols['loc0'], cols['loc1'] = pairs[0]
        if len(pairs) > 1:
            cols['loc0b'], cols['loc1b'] = pairs[1]
        for ly in el['layers']:
            if ly['op'] == 'grating':
                cols['order'], cols['blaze'] = ly['order_name'], ly['blaze_name']
            elif ly['op'] == 'l1':
                cols['order2'], cols['blaze2'] = ly['order_name'], ly['blaze_name']
            elif ly['op'] == 'scatter' and not (ly['mode'] == 'const' and ly['sigma'] == 0):
                cols['scat'] = ly['name']
            elif ly['op'] == 'radial':
                cols['scat'], cols['scat2'] = ly['name_in'], ly['name_perp']
        return cols
    default_loc = {'aperture_circle': ['y', 'z'], 'aperture_rect': ['y', 'z'],
                   'lens_perfect': ['mirror_x', 'mirror_y'], 'lens_thin': ['y', 'z'],
                   'grating': ['grat_y', 'grat_z'], 'detector': ['det_x', 'det_y'],
                   'cylinder': ['det_phi', 'det_y'],
                   }[kind]
    loc = el.get('loc', default_loc)
    if loc is not None:
        cols['loc0'], cols['loc1'] = loc
    if kind == 'grating':
        cols['order'] = el.get('order_name', 'order')
        cols['blaze'] = el.get('blaze_name', 'blaze')
    if kind == 'cylinder':
        cols['detpix0'], cols['detpix1'] = el.get('detpix', ['detpix_x', 'detpix_y'])
    if kind == 'detector':
        cols['detpix0'], cols['detpix1'] = el.get('detpix', ['detpix_x', 'detpix_y'])
        pix = el.get('pix', ['pix_x', 'pix_y'])
        if pix is not None:
            cols['pix0'], cols['pix1'] = pix
    return cols


def uniform_slots(elements, n_draws=2):
    """Slot base per element: consecutive elements of one kind writing the same
    columns (stacks: and with the same layer operations) share a stage; an
    aperture stage takes 2 draws (phi or y, then r or z), a grating stage
    ``n_draws`` (k-th interaction of a photon in the stage uses draw k, draws
    reused cyclically beyond that), a stack stage 5 per interaction."""
    out = []
    slot = 0
    prev = None
    stage_id = -1
    for el in elements:
        key = (el['kind'], tuple(sorted((k, v) for k, v in columns_of(el).items() if v)),
               stack_ops(el))
        new_stage = (prev is None or key != prev or el['kind'] in APERTURES)
        if new_stage:
            n = (2 if el['kind'] in APERTURES else
                 (n_draws if el['kind'] == 'grating' else
                  (STACK_DRAWS * n_draws if el['kind'] == 'stack' else 0)))
            base = slot
            slot += n
            stage_id += 1
        else:
            base, n = out[-1][0], out[-1][1]
        out.append((base, n, stage_id))
        prev = key
    return out, slot


INT_COLS = ('id', 'order', 'pix0', 'pix1', 'order2')


def stack_ops(el):
    """The layer operations of a stack that change photons (a constant
    scatter of sigma 0 does nothing); () for other elements."""
    if el['kind'] != 'stack':
        return ()
    ops = []
    for ly in el['layers']:
        if ly['op'] == 'scatter':
            if ly['mode'] == 'const' and ly['sigma'] == 0:
                continue
            ops.append(('scatter', ly['mode']))
        else:
            ops.append((ly['op'],))
    return tuple(ops)


def axangle_t_apply(axes, angles, v):
    """einsum('...ij,...i->...j', axangle2mat(axes, angles), v) of the
    reference's scatter (rotations.py:54-87, scatter.py:133-139): the
    transposed rotation matrix (axis normalised) applied to v."""
    axes = axes / np.linalg.norm(axes, axis=1)[:, None]
    c = np.cos(angles)
    s = np.sin(angles)
    C = 1 - c
    x, y, z = axes[:, 0], axes[:, 1], axes[:, 2]
    xs, ys, zs = x * s, y * s, z * s
    xC, yC, zC = x * C, y * C, z * C
    xyC, yzC, zxC = x * yC, y * zC, z * xC
    m = [[x * xC + c, xyC - zs, zxC + ys],
         [xyC + zs, y * yC + c, yzC - xs],
         [zxC - ys, yzC + xs, z * zC + c]]
    return np.stack([(m[0][j] * v[:, 0] + m[1][j] * v[:, 1]) + m[2][j] * v[:, 2]
                     for j in range(3)], axis=1)


def transmission(ly, energy):
    """interp1d(energy, trans) of the L1 layer (catgrating.py:189-192),
    scipy's linear arithmetic; raises outside the table like interp1d."""
    x, y = ly['trans_energy'], ly['trans_value']
    if np.any(energy < x[0]) or np.any(energy > x[-1]):
        raise ValueError('A value in x_new is outside the interpolation range.')
    hi = np.clip(np.searchsorted(x, energy), 1, len(x) - 1)
    lo = hi - 1
    slope = (y[hi] - y[lo]) / (x[hi] - x[lo])
    return slope * (energy - x[lo]) + y[lo]


def _check_prob(v):
    if np.any(v < 0) or np.any(v > 1):
        raise ValueError('Found probability outside of the 0..1 arange.')


def stack_layers(el, fr, ph, hit, ly, uniforms, base, nslot, k, cols_out):
    """FlatStack.process_photons (optics/base.py:258-281): every layer in
    order on the photons ``hit``, each a process_photons of its own
    (probability factor checked and multiplied, columns written).  Layer
    draws: slot 5 k + (0 order, 1 L1 order, 2 L1 open, 3 normal, 4 angle).
    Returns {column key: values} for the bookkeeping."""
    from scipy.special import ndtri
    idx = np.nonzero(hit)[0]
    energy = ph['energy'][hit]

    def u(j):
        sl = STACK_DRAWS * k + j
        sl = np.where(sl < nslot, sl, sl % nslot)
        return uniforms[base + sl, idx]
    order_last = None
    for layer in el['layers']:
        op = layer['op']
        d_old, pol_old = ph['dir'][hit].copy(), ph['pol'][hit].copy()
        if op in ('grating', 'l1'):
            lel = dict(layer, kind='grating')
            tab = selector_table(layer['selector'])
            blz = blaze_angle(lel, fr, d_old, ly) if tab.get('type') == 'interp' else None
            order, ptot, ok = select_order(tab, energy, u(0 if op == 'grating' else 1), blz)
            if not ok.all():
                raise ValueError('no order selected')
            new, blaze = diffract(lel, fr, d_old, energy, ly, order, ptot)
            pol = parallel_transport(d_old, new, pol_old)
            if op == 'l1':
                through = u(2) > layer['openfraction']
                new[through] = d_old[through]
                pol[through] = pol_old[through]
                order = np.where(through, 0, order)
                ptot = np.where(through, 0., ptot)
                if through.any():
                    ptot[through] = transmission(layer, energy[through])
                cols_out['order2'] = order
                cols_out['blaze2'] = blaze
            else:
                order_last = order
                cols_out['order'] = order
                cols_out['blaze'] = blaze
            _check_prob(ptot)
            ph['prob'][hit] = ph['prob'][hit] * ptot
            ph['dir'][hit] = new
            ph['pol'][hit] = pol
        elif op == 'qf':
            if order_last is None:
                raise ValueError('QualityFactor needs a grating layer before it')
            f = layer['factor'] ** (order_last.astype(float) ** 2)
            _check_prob(f)
            ph['prob'][hit] = ph['prob'][hit] * f
        elif op == 'l2abs':
            p3 = unit(d_old)
            angle = np.arccos(np.abs(dot(p3, fr['e_x'])))
            f = layer['openfraction'] - layer['coef'] * np.sin(angle)
            _check_prob(f)
            ph['prob'][hit] = ph['prob'][hit] * f
        elif op == 'radial':
            # RadialMirrorScatter.specific_process_photons, scatter.py:49-78:
            # radial from the centre to the last interaction point (the
            # position before this element), rotations applied transposed
            d0 = ph['dir'][hit]
            radial = ph['pos'][hit] - layer['center'][None, :]
            perpplane = np.cross(d0, radial)
            if layer['sigma_in'] != 0:
                inplane = ndtri(u(3)) * layer['sigma_in']
                out = axangle_t_apply(perpplane, inplane, d0)
            else:
                inplane = np.zeros(len(d0))
                out = d0.copy()
            if layer['sigma_perp'] != 0:
                perp = ndtri(u(4)) * layer['sigma_perp']
                out = axangle_t_apply(radial, perp, out)
            else:
                perp = np.zeros(len(d0))
            ph['pol'][hit] = parallel_transport(d0, out, pol_old)
            ph['dir'][hit] = out
            cols_out['scat'] = inplane
            cols_out['scat2'] = perp
        elif op == 'scatter':
            if layer['mode'] == 'const' and layer['sigma'] == 0:
                continue
            if layer['mode'] == 'l2':
                wave = KEV_NM * 1e-6 / energy
                sigma = 1.22 * 0.4 * np.arcsin(wave / layer['innerfree'])
            else:
                sigma = layer['sigma']
            angle = ndtri(u(3)) * sigma
            pdir = unit(d_old)
            guess = np.zeros_like(pdir)
            ind = np.abs(pdir[:, 0]) < 0.99999
            guess[ind, 0] = 1
            guess[~ind, 1] = 1
            perp = np.cross(pdir, guess)
            out = axangle_t_apply(perp, angle, pdir)
            angle2 = u(4) * 2 * np.pi
            out = axangle_t_apply(pdir, angle2, out)
            ph['pol'][hit] = parallel_transport(d_old, out, pol_old)
            ph['dir'][hit] = out
            cols_out['scat'] = angle
        else:
            raise ValueError(op)


def trace(elements, photons, uniforms, n_draws=2, selector_loop=False):
    """Element-major trace (simulator.py:42-49 + optics/base.py:144-211).

    Returns a new dict of columns.  Raises ValueError where the reference
    raises (probability factor outside [0, 1]; EfficiencyFile without a
    valid order).  ``selector_loop``: tabulated order selectors run one
    photon at a time as the reference's EfficiencyFile does
    (:func:`select_order_loop`; identical results, for timing).
    """
    ph = {k: np.array(v, dtype=float) for k, v in photons.items() if k in
          ('pos', 'dir', 'pol', 'energy', 'prob')}
    n = len(ph['dir'])
    if 'pos' not in ph:
        ph['pos'] = np.full((n, 3), np.nan)
    slots, _ = uniform_slots(elements, n_draws)
    cols = {}
    int_names = {}
    count = None
    cur_stage = None
    for el, (base, nslot, stage_id) in zip(elements, slots):
        if stage_id != cur_stage:
            count = np.zeros(n, dtype=np.int64)
            cur_stage = stage_id
        kind = el['kind']
        if kind == 'catsupportbars':
            # catgrating.py:314-324: the facet frame absorbs photons that
            # met no facet (all photons when there is no facet column)
            if 'facet' in cols:
                ph['prob'][cols['facet'] < 0] = 0.
            else:
                ph['prob'][:] = 0.
            continue
        fr = frame(el['pos4d'])
        cnames = columns_of(el)
        if kind in APERTURES:
            hit = np.ones(n, dtype=bool)
            x, y, q = aperture_points(el, fr, uniforms[base], uniforms[base + 1])
            ly, lz = x, y
            f = dot(ph['dir'], -fr['e_x'])
            f = np.clip(f, 0, 1.)
            out = {'probability': f}
        elif kind == 'cylinder':
            # CircularDetector.specific_process_photons, detector.py:113-116
            hit, q, ly, lz = cylinder_intersect(el, ph['pos'], ph['dir'])
            if not hit.any():
                continue
            q, ly, lz = q[hit], ly[hit], lz[hit]
            R = zooms_gram_schmidt(el['pos4d'])[0]
            out = {'detpix0': ly * R / el['pixsize'] + el['centerpix'][0],
                   'detpix1': lz / el['pixsize'] + el['centerpix'][1]}
            count[hit] += 1
        else:
            hit, q, ly, lz = intersect(fr, ph['pos'], ph['dir'])
            if not hit.any():
                continue
            q, ly, lz = q[hit], ly[hit], lz[hit]
            d_old, pol_old = ph['dir'][hit], ph['pol'][hit]
            out = {}
            if kind == 'grating':
                k = count[hit]
                slot = np.where(k < nslot, k, k % nslot)
                u = uniforms[base + slot, np.nonzero(hit)[0]]
                tab = selector_table(el['selector'])
                blz = blaze_angle(el, fr, d_old, ly) if tab.get('type') == 'interp' else None
                if selector_loop and tab.get('type') != 'interp':
                    order, ptot, ok = select_order_loop(tab, ph['energy'][hit], u)
                else:
                    order, ptot, ok = select_order(tab, ph['energy'][hit], u, blz)
                if not ok.all():
                    raise ValueError('no order selected')
                new, blaze = diffract(el, fr, d_old, ph['energy'][hit], ly, order, ptot)
                out['dir'] = new
                out['probability'] = ptot
                out['polarization'] = parallel_transport(d_old, new, pol_old)
                out['order'] = order
                out['blaze'] = blaze
            elif kind == 'lens_perfect':
                new, pol, pf = perfect_lens(el, fr, d_old, q, pol_old, ph['energy'][hit])
                out['dir'] = new
                out['polarization'] = pol
                if el.get('reflectivity') is not None:
                    out['probability'] = pf
            elif kind == 'lens_thin':
                new, pol = thin_lens(el, fr, d_old, q, pol_old)
                out['dir'] = new
                out['polarization'] = pol
            elif kind == 'stack':
                # photons meeting the stack for the k-th time in this stage use
                # draws 5 k ... 5 k + 4: process each multiplicity in turn
                k = count[hit]
                stack_out = {}
                for kk in np.unique(k):
                    sub = hit.copy()
                    sub[hit] = k == kk
                    sub_out = {}
                    stack_layers(el, fr, ph, sub, ly[k == kk], uniforms, base, nslot, int(kk), sub_out)
                    for key, v in sub_out.items():
                        arr = stack_out.setdefault(key, np.zeros(int(hit.sum()),
                                                                 dtype=np.asarray(v).dtype))
                        arr[k == kk] = v
                out.update(stack_out)
            elif kind == 'detector':
                npix, cpix = detector_params(el)
                dp0 = ly / el['pixsize'] + cpix[0]
                dp1 = lz / el['pixsize'] + cpix[1]
                out['detpix0'] = dp0
                out['detpix1'] = dp1
                out['pix0'] = pixel_index(dp0, npix[0])
                out['pix1'] = pixel_index(dp1, npix[1])
            else:
                raise ValueError(kind)
            count[hit] += 1
        # _assign_col_value, optics/base.py:12-51
        if 'probability' in out:
            v = out['probability']
            if np.any(v < 0) or np.any(v > 1):
                raise ValueError('Found probability outside of the 0..1 arange.')
            ph['prob'][hit] = ph['prob'][hit] * v
        if 'dir' in out:
            ph['dir'][hit] = out['dir']
        if 'polarization' in out:
            ph['pol'][hit] = out['polarization']
        extra = {'loc0': ly, 'loc1': lz, 'id': el.get('id_num', -9)}
        if kind == 'stack':
            extra['loc0b'], extra['loc1b'] = ly, lz
        for key in ('order', 'blaze', 'detpix0', 'detpix1', 'pix0', 'pix1', 'order2',
                    'blaze2', 'scat', 'scat2'):
            if key in out:
                extra[key] = out[key]
        for key, val in extra.items():
            name = cnames.get(key)
            if name is None:
                continue
            if name not in cols:
                if key in INT_COLS:
                    cols[name] = np.full(n, ORDER_NONE if key in ('order', 'order2') else -1,
                                         dtype=np.int64)
                    int_names[name] = key
                else:
                    cols[name] = np.full(n, np.nan)
            cols[name][hit] = val
        ph['pos'][hit] = q
    res = dict(ph)
    res.update(cols)
    return res
