"""Misalignment helpers and tolerancing runs.

``generate_facet_uncertainty``/``wiggle``/``varyperiod`` follow
``marxs/design/uncertainties.py:8-39`` and ``marxs/design/tolerancing.py:44-58,
175-201`` (Gaussian translations and fixed-axis x-y-z rotation angles turned
into ``compose(t, R, 1)`` matrices in ``elem_uncertainty``; Gaussian periods),
with the random generator passed in instead of numpy's global stream.
``run_tolerances`` keeps the reference's calling convention
(``tolerancing.py:224-283``); ``run_tolerances_batched`` traces all variants
of the same instrument in one batched GPU launch (``marxs_amd.batch``).
"""
import collections.abc
from copy import copy
from functools import wraps

import numpy as np

from .mathutils import compose, euler2mat_sxyz

__all__ = ['oneormoreelements', 'generate_facet_uncertainty', 'wiggle', 'moveglobal',
           'moveindividual', 'moveelem', 'varyperiod', 'varyorderselector', 'varyattribute',
           'run_tolerances', 'run_tolerances_batched', 'generate_6d_wigglelist', 'reset_6d',
           'select_1dof_changed']


def oneormoreelements(func):
    """Let a function written for one element accept a list of elements
    (tolerancing.py:26-40)."""
    @wraps(func)
    def wrapper(elements, *args, **kwargs):
        if isinstance(elements, collections.abc.Iterable):
            for e in elements:
                func(e, *args, **kwargs)
        else:
            func(elements, *args, **kwargs)
    return wrapper


def _move(dx, dy, dz, rx, ry, rz):
    return compose([dx, dy, dz], euler2mat_sxyz(rx, ry, rz), np.ones(3))


def generate_facet_uncertainty(n, xyz, angle_xyz, rng=None, trans_offset=0, rot_offset=0):
    rng = np.random.default_rng() if rng is None else rng
    translation = rng.normal(size=(n, 3), loc=trans_offset, scale=xyz)
    rotation = rng.normal(size=(n, 3), loc=rot_offset, scale=angle_xyz)
    return [compose(t, euler2mat_sxyz(a[0], a[1], a[2]), np.ones(3))
            for t, a in zip(translation, rotation)]


@oneormoreelements
def wiggle(e, dx=0, dy=0, dz=0, rx=0., ry=0., rz=0., rng=None):
    """Random per-element misalignment of a Parallel container
    (Gaussian sigmas in mm and rad; tolerancing.py:44-58)."""
    e.elem_uncertainty = generate_facet_uncertainty(len(e.elements), [dx, dy, dz],
                                                    [rx, ry, rz], rng=rng)
    e.generate_elements()


@oneormoreelements
def moveglobal(e, dx=0, dy=0, dz=0, rx=0., ry=0., rz=0.):
    """Move and rotate the whole Parallel container about its own origin
    (tolerancing.py:62-78)."""
    e.uncertainty = _move(dx, dy, dz, rx, ry, rz)
    e.generate_elements()


@oneormoreelements
def moveindividual(e, dx=0, dy=0, dz=0, rx=0, ry=0, rz=0):
    """Move and rotate every element of a Parallel container by the same
    amount, each about its own centre (tolerancing.py:81-103)."""
    e.elem_uncertainty = [_move(dx, dy, dz, rx, ry, rz)] * len(e.elements)
    e.generate_elements()


@oneormoreelements
def moveelem(e, dx=0, dy=0, dz=0, rx=0., ry=0., rz=0.):
    """Move and rotate one element; its original pose is kept as
    ``geometry.pos4d_orig`` (tolerancing.py:106-132)."""
    if not hasattr(e.geometry, 'pos4d_orig'):
        e.geometry.pos4d_orig = e.geometry.pos4d.copy()
    e.geometry.pos4d = _move(dx, dy, dz, rx, ry, rz) @ e.geometry.pos4d_orig


@oneormoreelements
def varyattribute(element, **kwargs):
    """Set existing attributes of an element (tolerancing.py:135-170)."""
    for key, val in kwargs.items():
        if not hasattr(element, key):
            raise ValueError(f'Object {element} does not have {key} attribute.')
        setattr(element, key, val)


@oneormoreelements
def varyperiod(el, period_mean, period_sigma, rng=None):
    """Gaussian grating period, one draw per element
    (tolerancing.py:175-201)."""
    rng = np.random.default_rng() if rng is None else rng
    if not hasattr(el, '_d'):
        raise ValueError(f'Object {el} does not have grating period `_d` attribute.')
    el._d = rng.normal(period_mean, period_sigma)


@oneormoreelements
def varyorderselector(element, order_selector, *args, **kwargs):
    """Replace the order selector of a grating (tolerancing.py:204-221)."""
    if not hasattr(element, 'order_selector'):
        raise ValueError(f'Object {element} does not have an order_selector attribute.')
    element.order_selector = order_selector(*args, **kwargs)


def generate_6d_wigglelist(trans, rot, names=('dx', 'dy', 'dz', 'rx', 'ry', 'rz')):
    """Parameter lists changing one degree of freedom at a time
    (tolerancing.py:454-516).  ``trans`` in mm and ``rot`` in rad (plain
    numbers; the first entry of each should be 0).  Returns (changes in both
    directions for the movers, one-sided changes for ``wiggle``)."""
    trans = np.asarray(trans, dtype=float)
    rot = np.asarray(rot, dtype=float)
    if trans[0] != 0 or rot[0] != 0:
        import warnings
        warnings.warn('First element of trans and rot should be 0.')
    n_t, n_r = len(trans), len(rot)
    n = 3 * n_t
    g = np.zeros((2 * 3 * n_t + 2 * 3 * n_r, 6))
    for i in range(3):
        g[i * n_t:(i + 1) * n_t, i] = trans
        g[n + i * n_r:n + (i + 1) * n_r, i + 3] = rot
    half = g.shape[0] // 2
    g[half:, :] = -g[:half, :]
    ind = g[:half, :]
    g = np.unique(g, axis=0)
    ind = np.unique(ind, axis=0)
    return ([dict(zip(names, row)) for row in g], [dict(zip(names, row)) for row in ind])


reset_6d = {'dx': 0., 'dy': 0., 'dz': 0., 'rx': 0., 'ry': 0., 'rz': 0.}


def select_1dof_changed(rows, par, parlist=('dx', 'dy', 'dz', 'rx', 'ry', 'rz')):
    """Rows (dicts, e.g. the output of run_tolerances) in which every
    parameter of ``parlist`` except ``par`` is zero (tolerancing.py:523-548)."""
    others = set(parlist) - {par}
    return [r for r in rows if all(r[p] == 0 for p in others)]


def run_tolerances(photons_in, instrum, wigglefunc, wiggleparts, parameters, analyzefunc):
    """Sequential variant loop (tolerancing.py:224-283): modify, trace a copy
    of the input photons on the GPU, analyse."""
    out = []
    for pars in parameters:
        wigglefunc(wiggleparts, **pars)
        photons = instrum(photons_in.copy())
        cpars = copy(pars)
        cpars.update(analyzefunc(photons))
        out.append(cpars)
    return out


def run_tolerances_batched(photons_in, instrum, wigglefunc, wiggleparts, parameters, analyzefunc,
                           precision=64, uniforms=None, generator=None, tune=False,
                           analyze_on_device=False):
    """``run_tolerances`` with every variant traced in one batched launch.

    Each parameter set modifies the instrument as in the sequential loop and
    the modified tree is compiled at once (a snapshot of its poses and
    periods); then all variants are traced together (``mx_trace_batched``)
    and each variant's photon table is analysed in turn.  ``uniforms`` (slots,
    N) fixes the random draws (the same for every variant: common random
    numbers); otherwise every variant draws its own.  ``tune`` runs the tile
    autotuner first.  ``analyze_on_device``: ``analyzefunc`` receives each
    variant's ``DevicePhotons`` instead of a host table (e.g.
    ``marxs_amd.analysis.CaptureResAeff``, whose reductions then run on the
    columns in HBM).  Photons must not hold output columns of the
    instrument yet.
    """
    import torch
    from .tracer import CompiledInstrument, DevicePhotons, write_back
    from .batch import BatchedTrace
    cis = []
    for pars in parameters:
        wigglefunc(wiggleparts, **pars)
        cis.append(CompiledInstrument(instrum))
    if not cis:
        return []
    clash = [c for c in photons_in.colnames
             if c in cis[0].tables.fcols or c in cis[0].tables.icols]
    if clash:
        raise ValueError(f'photons already hold output columns {clash}')
    src = DevicePhotons.from_table(photons_in, device=cis[0].device)
    if uniforms is not None:
        u = torch.as_tensor(np.ascontiguousarray(uniforms), dtype=torch.float64,
                            device=cis[0].device)
        unis = [u] * len(cis)
    else:
        unis = [ci.draw_uniforms(src.n, generator) for ci in cis]
    bt = BatchedTrace(cis, [src] * len(cis), unis, precision=precision)
    if tune:
        from .autotune import tune_batched
        tune_batched(bt)
    bt.run()
    out = []
    for pars, ci, dp in zip(parameters, cis, bt.outputs):
        cpars = copy(pars)
        if analyze_on_device:
            cpars.update(analyzefunc(dp))
        else:
            photons = photons_in.copy()
            write_back(dp, photons, ci.tables)
            cpars.update(analyzefunc(photons))
        out.append(cpars)
    return out
