"""Plain-data description of an instrument tree for the oracle
(TEST INFRASTRUCTURE ONLY).

Reads only the user-level attributes an element was built with (its pose
``pos4d``, grating period ``_d``, ``groove_angle``, order-selector inputs,
lens focal length, detector pixel size, id columns) by duck typing, so that
the oracle derives every trace constant itself and shares no code with the
product's compile step.
"""
import numpy as np


def _is_stack(e):
    return any(c.__name__ == 'FlatStack' for c in type(e).__mro__)


def _leaves(tree):
    if hasattr(tree, 'elements') and not _is_stack(tree):
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


LAYER_TYPES = ('QualityFactor', 'L1', 'L2Abs', 'L2Diffraction', 'RandomGaussianScatter',
               'RadialMirrorScatter')


def _layer(e):
    """Plain description of one stack layer (or a stand-alone layer)."""
    name = type(e).__name__
    loc = e.loc_coos_name
    if name in ('FlatGrating', 'CATGrating', 'NonParallelCATGrating', 'L1'):
        g = e.geometry
        d = {'op': 'l1' if name == 'L1' else 'grating', 'd': e._d,
             'groove_angle': g['groove_angle'],
             'sign': 'flat' if name == 'FlatGrating' else 'cat',
             'transmission': e.transmission, 'selector': _selector(e.order_selector),
             'order_name': e.order_name, 'blaze_name': e.blaze_name, 'loc': loc}
        if name == 'NonParallelCATGrating':
            d['blaze_center'] = e.blaze_center
            d['d_blaze_mm'] = e.d_blaze_mm
        if name == 'L1':
            d['openfraction'] = e.openfraction
            d['trans_energy'] = np.asarray(e.trans_energy, dtype=float)
            d['trans_value'] = np.asarray(e.trans_value, dtype=float)
        return d
    if name == 'QualityFactor':
        return {'op': 'qf', 'factor': e.factor, 'loc': loc}
    if name == 'L2Abs':
        openfraction = (e.innerfree / e.period) ** 2
        totalarea = e.period ** 2 / 2 * np.sqrt(3)
        return {'op': 'l2abs', 'openfraction': openfraction,
                'coef': e.bardepth * e.innerfree / totalarea, 'loc': loc}
    if name == 'RadialMirrorScatter':
        return {'op': 'radial', 'sigma_in': float(e.inplanescatter),
                'sigma_perp': float(e.perpplanescatter), 'name_in': e.inplanescattercol,
                'name_perp': e.perpplanescattercol,
                'center': np.array(e.pos4d, dtype=float)[:3, 3], 'loc': loc}
    if name in ('L2Diffraction', 'RandomGaussianScatter'):
        if name == 'L2Diffraction':
            return {'op': 'scatter', 'mode': 'l2', 'innerfree': e.innerfree,
                    'name': e.scattername, 'loc': loc}
        return {'op': 'scatter', 'mode': 'const', 'sigma': float(e.scatter),
                'name': e.scattername, 'loc': loc}
    raise TypeError(f'oracle has no model of stack layer {name}')


def describe(tree):
    out = []
    for e in _leaves(tree):
        if callable(e) and getattr(e, '__name__', '') == 'catsupportbars':
            out.append({'kind': 'catsupportbars'})
            continue
        name = type(e).__name__
        if _is_stack(e) or name in LAYER_TYPES:
            stack = _is_stack(e)
            out.append({'kind': 'stack', 'pos4d': np.array(e.pos4d, dtype=float),
                        'id_col': e.id_col, 'id_num': e.id_num,
                        'loc': e.loc_coos_name if stack else None,
                        'layers': [_layer(x) for x in (e.elements if stack else [e])]})
            continue
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
            # the callable itself: the reference evaluates it (mirror.py:78-80)
            if getattr(e, 'reflectivity_interpolator', None) is not None:
                d['reflectivity'] = e.reflectivity_interpolator
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
        elif name == 'CircularDetector':
            d['kind'] = 'cylinder'
            d['pixsize'] = e.pixsize
            d['centerpix'] = list(e.centerpix)
            d['detpix'] = e.detpix_name
            d['phi_lim'] = np.array(e.geometry.coos_limits[0], dtype=float)
        elif name == 'FlatDetector':
            d['kind'] = 'detector'
            d['pixsize'] = e.pixsize
            d['detpix'] = e.detpix_name
            d['pix'] = e.pixel_name
        else:
            raise TypeError(f'oracle has no model of {name}')
        out.append(d)
    return out
