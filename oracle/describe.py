"""Plain-data description of an instrument tree for the oracle
(TEST INFRASTRUCTURE ONLY).

Reads only the user-level attributes an element was built with (its pose
``pos4d``, grating period ``_d``, ``groove_angle``, order-selector inputs,
lens focal length, detector pixel size, id columns) by duck typing, so that
the oracle derives every trace constant itself and shares no code with the
product's compile step.
"""
import numpy as np


def _leaves(tree):
    if hasattr(tree, 'elements'):
        out = []
        for e in tree.elements:
            out += _leaves(e)
        return out
    if isinstance(tree, (list, tuple)):
        out = []
        for e in tree:
            out += _leaves(e)
        return out
    return [tree]


def _selector(sel):
    name = type(sel).__name__
    if name == 'OrderSelector':
        return {'type': 'orderselector', 'orders': np.asarray(sel.orderlist),
                'p': np.asarray(sel.p, dtype=float)}
    if name == 'EfficiencyFile':
        return {'type': 'efficiency', 'energy': np.asarray(sel.energy, dtype=float),
                'prob': np.asarray(sel.prob, dtype=float), 'orders': np.asarray(sel.orders)}
    if name == 'InterpolateEfficiencyTable':
        return {'type': 'interp', 'wave': np.asarray(sel.wave, dtype=float),
                'theta': np.asarray(sel.theta, dtype=float), 'orders': np.asarray(sel.orders),
                'data': np.asarray(sel.data, dtype=float), 'k': int(sel.k)}
    raise TypeError(name)


def describe(tree):
    out = []
    for e in _leaves(tree):
        name = type(e).__name__
        g = e.geometry
        d = {'pos4d': np.array(e.pos4d, dtype=float), 'id_col': e.id_col,
             'id_num': e.id_num, 'loc': e.loc_coos_name}
        if name in ('CircleAperture', 'RectangleAperture'):
            d['kind'] = 'aperture_circle' if name == 'CircleAperture' else 'aperture_rect'
            if name == 'CircleAperture':
                d['phi'] = list(g.phi)
                d['r_inner'] = g['r_inner']
        elif name == 'PerfectLens':
            d['kind'] = 'lens_perfect'
            d['focallength'] = e.focallength
            d['d_center_optax'] = e.d_center_optax
        elif name == 'ThinLens':
            d['kind'] = 'lens_thin'
            d['focallength'] = e.focallength
        elif name in ('FlatGrating', 'CATGrating', 'NonParallelCATGrating'):
            d['kind'] = 'grating'
            d['d'] = e._d
            d['groove_angle'] = g['groove_angle']
            d['sign'] = 'flat' if name == 'FlatGrating' else 'cat'
            d['transmission'] = e.transmission
            d['selector'] = _selector(e.order_selector)
            d['order_name'] = e.order_name
            d['blaze_name'] = e.blaze_name
            if name == 'NonParallelCATGrating':
                d['blaze_center'] = e.blaze_center
                d['d_blaze_mm'] = e.d_blaze_mm
        elif name == 'FlatDetector':
            d['kind'] = 'detector'
            d['pixsize'] = e.pixsize
            d['detpix'] = e.detpix_name
            d['pix'] = e.pixel_name
        else:
            raise TypeError(f'oracle has no model of {name}')
        out.append(d)
    return out
