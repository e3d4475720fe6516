"""Shared test helpers: run the GPU trace and compare it with the oracle."""
import numpy as np

from oracle import trace as OT
from oracle.describe import describe

ORDER_NONE = -2 ** 31


def gpu_trace(inst, ph, uniforms, precision=64, grid=True, n_draws=2, device=None):
    """Trace host photon arrays (oracle layout) on the GPU; returns a dict in
    the oracle's layout plus the status counters."""
    import torch
    from marxs_amd.tracer import CompiledInstrument, DevicePhotons
    ci = inst if hasattr(inst, 'trace_device') else CompiledInstrument(inst, n_draws=n_draws, grid=grid)
    dp = DevicePhotons(ph['dir'].T, ph['pol'].T, ph['energy'], ph['prob'],
                       pos=ph['pos'].T if 'pos' in ph else None)
    u = torch.as_tensor(np.ascontiguousarray(uniforms), dtype=torch.float64, device=dp.device)
    st = ci.trace_device(dp, uniforms=u, precision=precision, grid=grid)
    out = {'pos': dp.pos.cpu().numpy().T, 'dir': dp.dir.cpu().numpy().T,
           'pol': dp.pol.cpu().numpy().T, 'prob': dp.prob.cpu().numpy(),
           'energy': dp.energy.cpu().numpy()}
    for k, v in dp.fcols.items():
        out[k] = v.cpu().numpy()
    for k, v in dp.icols.items():
        out[k] = v.cpu().numpy().astype(np.int64)
    return out, st


def oracle_trace(inst, ph, uniforms, n_draws=2):
    return OT.trace(describe(inst), ph, uniforms, n_draws=n_draws)


def near_pixel_boundary(ref, tol=1e-6):
    """Photons whose oracle detector pixel coordinate lies within ``tol``
    pixels of a pixel edge (k + 0.5): the integer pixel index there depends on
    the last bits of the coordinate."""
    out = {}
    for a, b in (('pix_x', 'detpix_x'), ('pix_y', 'detpix_y')):
        if b in ref:
            v = np.asarray(ref[b])
            with np.errstate(invalid='ignore'):
                f = (v + 0.5) - np.floor(v + 0.5)
                out[a] = np.nan_to_num(np.minimum(f, 1 - f), nan=1.) < tol
    return out


def compare(ref, got, rtol=1e-9, atol=1e-11, max_int_mismatch=0, near=None, max_near_frac=1e-3,
            verbose=True, skip=()):
    """Return a report dict; raise AssertionError when outside tolerance.

    Floats: |a - b| <= rtol * |a| + atol per value (NaN must match NaN);
    vector columns (N, 3) per component against the vector's length,
    |a_k - b_k| <= rtol * |a| + atol.  The absolute floor atol (1e-11 mm,
    rad, pixel or probability: about five units in the last place of the
    1.2e4 mm global coordinates that local ones are differences of) only
    matters for values that are zero up to rounding, such as the coordinates
    of a photon focused onto the origin (there the relative bound does not
    exist; measured differences there reach 3e-12 mm and 7e-12 pixels); the
    report counts the values that pass only because of it.  Integers: exact, with at most ``max_int_mismatch``
    differing photons outside the near-boundary masks ``near[col]``; the
    differing photons inside them (a pixel index decided by the last bit of
    a coordinate on a pixel edge) are counted and must stay at most
    ``max_near_frac`` of the photons.
    """
    near = near or {}
    report = {}
    names = [k for k in ref if k in got and k not in skip]
    n_ph = max((len(np.asarray(ref[k])) for k in names), default=0)
    for k in names:
        a, b = np.asarray(ref[k]), np.asarray(got[k])
        if a.dtype.kind in 'iu':
            diff = a != b
            nb = int((diff & near[k]).sum()) if k in near else 0
            report[k] = ('int', int(diff.sum()) - nb, nb)
            continue
        if a.ndim == 2:
            scale = np.linalg.norm(a, axis=1)[:, None] * np.ones_like(a)
        else:
            scale = np.abs(a)
        nan_a, nan_b = np.isnan(a), np.isnan(b)
        nan_mism = int((nan_a != nan_b).sum())
        with np.errstate(invalid='ignore'):
            d = np.where(nan_a | nan_b, 0., np.abs(a - b))
            rel_bound = rtol * np.where(nan_a | nan_b, 0., scale)
            fail = d > rel_bound + atol
            floor_only = (d > rel_bound) & ~fail
            rel = np.where(scale > 0, d / np.where(scale > 0, scale, 1.), np.where(d > 0, np.inf, 0.))
        report[k] = ('float', float(rel.max()) if rel.size else 0., nan_mism, int(fail.sum()),
                     int(floor_only.sum()))
        if fail.any():
            # the largest absolute difference among the failing values and
            # its oracle value, for the failure message
            j = np.unravel_index(np.argmax(np.where(fail, d, -1.)), d.shape)
            report[k] = report[k] + (float(d[j]), float(a[j]))
    bad = []
    for k, r in report.items():
        if r[0] == 'int' and (r[1] > max_int_mismatch or r[2] > max_near_frac * n_ph):
            bad.append((k, r))
        if r[0] == 'float' and (r[3] > 0 or r[2] > max_int_mismatch):
            bad.append((k, r))
    if verbose:
        print({k: r for k, r in report.items()})
    assert not bad, f'parity failures: {bad}'
    return report


def make_inputs(n, seed, n_slots, energies='uniform'):
    from marxs_amd import configs
    rng = np.random.default_rng(seed)
    u = rng.random((2 + n_slots + 1, n))
    if energies == 'uniform':
        e = configs.uniform_energies(u[0])
    else:
        e = configs.line_spectrum_energies(u[0], u[-1])
    ph = configs.onaxis_photons(e, u[1])
    return ph, np.ascontiguousarray(u[2:2 + n_slots])


# ------------------------------------------------------------ golden vectors
GOLDEN = __import__('os').path.join(__import__('os').path.dirname(__file__), 'golden')


def encode_description(els):
    """Instrument description (oracle.describe output) as JSON-able data;
    identical selector tables are stored once (also those of stack layers,
    which nest as lists of dicts), and so are identical arrays of 32 or more
    values (e.g. the L1 transmission table of every stack facet)."""
    sels, arrays, seen = [], [], {}

    def arr(a):
        a = np.asarray(a)
        e = {'__array__': a.tolist(), 'dtype': str(a.dtype)}
        if a.size < 32:
            return e
        key = (str(a.dtype), a.shape, a.tobytes())
        if key not in seen:
            seen[key] = len(arrays)
            arrays.append(e)
        return {'__arrayref__': seen[key]}

    def enc(v):
        if isinstance(v, np.ndarray):
            return arr(v)
        if isinstance(v, np.bool_):
            return bool(v)
        if isinstance(v, np.integer):
            return int(v)
        if isinstance(v, np.floating):
            return float(v)
        if isinstance(v, dict):
            out = {}
            for k, vv in v.items():
                if k == 'selector' and isinstance(vv, dict):
                    e = {kk: enc(x) for kk, x in vv.items()}
                    if e not in sels:
                        sels.append(e)
                    out[k] = {'__selector__': sels.index(e)}
                else:
                    out[k] = enc(vv)
            return out
        if isinstance(v, (list, tuple)):
            return [enc(x) for x in v]
        return v

    elements = [enc(el) for el in els]
    out = {'selectors': sels, 'elements': elements}
    if arrays:
        out['arrays'] = arrays
    return out


def decode_description(obj):
    arrays = obj.get('arrays', [])

    def dec(v):
        if isinstance(v, dict):
            if '__array__' in v:
                return np.array(v['__array__'], dtype=v['dtype'])
            if '__arrayref__' in v:
                a = arrays[v['__arrayref__']]
                return np.array(a['__array__'], dtype=a['dtype'])
            if '__selector__' in v:
                return dict(sels[v['__selector__']])
            return {k: dec(x) for k, x in v.items()}
        if isinstance(v, list):
            return [dec(x) for x in v]
        return v
    sels = []
    sels.extend({k: dec(v) for k, v in s.items()} for s in obj['selectors'])
    return [dec(d) for d in obj['elements']]


def load_golden(name):
    """(description, photons, uniforms, outputs) of a committed case."""
    import json
    import os
    with open(os.path.join(GOLDEN, name + '.json')) as f:
        meta = json.load(f)
    z = np.load(os.path.join(GOLDEN, name + '.npz'), allow_pickle=False)
    ph = {k[3:]: z[k] for k in z.files if k.startswith('in_')}
    out = {k[4:]: z[k] for k in z.files if k.startswith('out_')}
    return meta, decode_description(meta['description']), ph, z['uniforms'], out
