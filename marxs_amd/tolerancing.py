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
from copy import copy

import numpy as np

from .mathutils import compose, euler2mat_sxyz

__all__ = ['generate_facet_uncertainty', 'wiggle', 'varyperiod', 'run_tolerances',
           'run_tolerances_batched']


def generate_facet_uncertainty(n, xyz, angle_xyz, rng=None, trans_offset=0, rot_offset=0):
    rng = np.random.default_rng() if rng is None else rng
    translation = rng.normal(size=(n, 3), loc=trans_offset, scale=xyz)
    rotation = rng.normal(size=(n, 3), loc=rot_offset, scale=angle_xyz)
    return [compose(t, euler2mat_sxyz(a[0], a[1], a[2]), np.ones(3))
            for t, a in zip(translation, rotation)]


def wiggle(e, dx=0, dy=0, dz=0, rx=0., ry=0., rz=0., rng=None):
    """Random per-element misalignment of a Parallel container."""
    e.elem_uncertainty = generate_facet_uncertainty(len(e.elements), [dx, dy, dz],
                                                    [rx, ry, rz], rng=rng)
    e.generate_elements()


def varyperiod(elements, period_mean, period_sigma, rng=None):
    """Gaussian grating periods, one draw per element."""
    rng = np.random.default_rng() if rng is None else rng
    if not isinstance(elements, (list, tuple)):
        elements = [elements]
    for el in elements:
        if not hasattr(el, '_d'):
            raise ValueError(f'Object {el} does not have grating period `_d` attribute.')
        el._d = rng.normal(period_mean, period_sigma)


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
                           precision=64, uniforms=None, generator=None, tune=False):
    """``run_tolerances`` with every variant traced in one batched launch.

    Each parameter set modifies the instrument as in the sequential loop and
    the modified tree is compiled at once (a snapshot of its poses and
    periods); then all variants are traced together (``mx_trace_batched``)
    and each variant's photon table is analysed in turn.  ``uniforms`` (slots,
    N) fixes the random draws (the same for every variant: common random
    numbers); otherwise every variant draws its own.  ``tune`` runs the tile
    autotuner first.  Photons must not hold output columns of the
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
        photons = photons_in.copy()
        write_back(dp, photons, ci.tables)
        cpars = copy(pars)
        cpars.update(analyzefunc(photons))
        out.append(cpars)
    return out
