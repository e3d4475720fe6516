"""Figures of merit of a traced photon list per grating order
(``marxs/analysis/gratings.py``), computed on the GPU.

``effectivearea_from_photonlist`` (``gratings.py:333-363``),
``resolvingpower_from_photonlist`` (``215-270``),
``resolvingpower_from_photonlist_robust`` (``273-330``),
``resolvingpower_per_order`` (``26-130``), ``weighted_per_order``
(``133-178``), ``average_R_Aeff`` (``187-212``),
``identify_photon_in_subaperture`` (``366-386``), ``CaptureResAeff`` and
``CaptureResAeff_CCDgaps`` (``389-553``), and from
``marxs/analysis/analysis.py`` ``sigma_clipped_std``, ``mean_width_2d``,
``find_best_detector_position`` and ``detected_fraction`` keep their names,
arguments and results.  The photon list may
be a :class:`~marxs_amd.tracer.DevicePhotons` (columns stay in HBM), a
:class:`~marxs_amd.table.PhotonTable` or a dict of arrays (columns are copied
to the GPU).  The per-order sums and the sigma clipping (astropy's
``sigma_clipped_stats`` with the reference's defaults: 3 sigma, at most 5
rounds, median centre) run in ``mx_order_stats``; only the few per-order
numbers come back to the host.  The functions that trace (the per-order
loop of ``resolvingpower_per_order``, the detector positions tried by
``find_best_detector_position``) keep the photons in HBM between traces and
run the FP64 kernels.
"""
import numpy as np
import torch

from . import _lib
from .compiler import ORDER_NONE

__all__ = ['AnalysisError', 'order_stats', 'sigma_clipped_stats',
           'effectivearea_from_photonlist', 'resolvingpower_from_photonlist',
           'resolvingpower_from_photonlist_robust', 'resolvingpower_per_order',
           'weighted_per_order', 'identify_photon_in_subaperture',
           'average_R_Aeff', 'CaptureResAeff', 'CaptureResAeff_CCDgaps',
           'sigma_clipped_std', 'mean_width_2d', 'find_best_detector_position',
           'detected_fraction']

RES_FWHM = 2.3548     # FWHM / sigma (gratings.py:268)


class AnalysisError(Exception):
    pass


def _device():
    if not torch.cuda.is_available():
        raise RuntimeError('the analysis reductions run on the GPU; no GPU is visible')
    return torch.device('cuda', torch.cuda.current_device())


def _column(photons, name, device):
    """Column ``name`` of a photon list as a device tensor (1-d)."""
    if hasattr(photons, 'fcols') and hasattr(photons, 'icols'):
        if name in photons.fcols:
            return photons.fcols[name]
        if name in photons.icols:
            return photons.icols[name]
        if name in ('probability', 'prob'):
            return photons.prob
        if name == 'energy':
            return photons.energy
        raise KeyError(name)
    v = photons[name]
    if isinstance(v, torch.Tensor):
        return v.to(device)
    return torch.as_tensor(np.asarray(v), device=device)


def _orders_i32(col):
    """Order column as int32 with the no-order sentinel (NaN or non-integer
    values never equal an order number)."""
    if col.dtype == torch.int32:
        return col.contiguous()
    if col.dtype.is_floating_point:
        ok = torch.isfinite(col) & (col == torch.round(col)) & (col.abs() < 2 ** 31 - 1)
        return torch.where(ok, col, torch.full_like(col, float(ORDER_NONE))).to(torch.int32).contiguous()
    return col.to(torch.int32).contiguous()


def order_stats(order, prob, orders, value=None, res_filter=False, sigma=3., maxiters=5,
                select=None, stream=None):
    """Per-order reductions on the GPU (``mx_order_stats``).

    order, prob, value: device tensors of one photon list; orders: distinct
    order numbers.  ``res_filter``: count only photons with a
    finite value and probability > 0; ``select``: boolean tensor, photons
    outside it are ignored.  More than 64 orders take several launches.
    Returns numpy arrays ``prob_sum``, ``count`` and,
    with a value column, the sigma-clipped ``mean``, ``median``, ``std`` and
    ``n_kept``."""
    dev = order.device
    orders = np.asarray(orders, dtype=np.int64).ravel()
    if len(orders) == 0:
        empty = np.zeros(0)
        return {'prob_sum': empty, 'count': np.zeros(0, np.int64), 'mean': empty,
                'median': empty, 'std': empty, 'n_kept': np.zeros(0, np.int64)}
    if len(np.unique(orders)) != len(orders):
        raise ValueError('order numbers must be distinct')
    if len(orders) > _lib.MX_MAX_ORDERS:
        # the kernels take at most MX_MAX_ORDERS orders per call
        parts = [order_stats(order, prob, orders[k:k + _lib.MX_MAX_ORDERS], value, res_filter,
                             sigma, maxiters, select, stream)
                 for k in range(0, len(orders), _lib.MX_MAX_ORDERS)]
        return {key: np.concatenate([r[key] for r in parts]) for key in parts[0]}
    o = _orders_i32(order)
    if select is not None:
        o = torch.where(torch.as_tensor(select, device=dev), o,
                        torch.full_like(o, ORDER_NONE)).contiguous()
    p = prob.to(torch.float64).contiguous()
    v = value.to(torch.float64).contiguous() if value is not None else None
    n = o.numel()
    if p.numel() != n or (v is not None and v.numel() != n):
        raise ValueError('columns of different lengths')
    no = len(orders)
    if n == 0:
        res = {'prob_sum': np.zeros(no), 'count': np.zeros(no, np.int64)}
        if v is not None:
            res.update({'mean': np.full(no, np.nan), 'median': np.full(no, np.nan),
                        'std': np.full(no, np.nan), 'n_kept': np.zeros(no, np.int64)})
        return res
    ords = torch.as_tensor(orders.astype(np.int32), device=dev)
    wsb = int(_lib.lib.mx_order_stats_workspace(n, no, int(v is not None)))
    if wsb < 0:
        raise ValueError('order statistics: invalid sizes')
    ws = torch.empty(max(wsb, 1), dtype=torch.uint8, device=dev)
    # the six result arrays in one buffer, so that one copy brings them back
    keys = ('prob_sum', 'count', 'mean', 'median', 'std', 'n_kept')
    ints = ('count', 'n_kept')
    buf = torch.empty(len(keys) * no, dtype=torch.float64, device=dev)
    out = {}
    for j, k in enumerate(keys):
        part = buf[j * no:(j + 1) * no]
        out[k] = part.view(torch.int64) if k in ints else part
    a = _lib.MxOrderStats()
    a.n = n
    a.order = o.data_ptr()
    a.value = v.data_ptr() if v is not None else None
    a.prob = p.data_ptr()
    a.orders = ords.data_ptr()
    a.n_orders = no
    a.flags = _lib.MX_OS_FILTER if res_filter else 0
    a.sigma = float(sigma)
    a.maxiters = int(maxiters)
    a.workspace = ws.data_ptr()
    a.workspace_bytes = wsb
    for k in ('prob_sum', 'count', 'mean', 'median', 'std', 'n_kept'):
        setattr(a, k, out[k].data_ptr())
    if stream is None:
        stream = torch.cuda.current_stream(dev).cuda_stream
    _lib.check(_lib.lib.mx_order_stats(a, stream))
    host = buf.cpu().numpy()
    res = {}
    for j, k in enumerate(keys):
        part = host[j * no:(j + 1) * no]
        res[k] = part.view(np.int64) if k in ints else part
    if v is None:
        for k in ('mean', 'median', 'std', 'n_kept'):
            res.pop(k)
    return res


def sigma_clipped_stats(data, sigma=3., maxiters=5):
    """(mean, median, std) of ``data`` after sigma clipping (astropy's
    ``sigma_clipped_stats`` with median centre and std ddof 0), on the GPU."""
    dev = data.device if isinstance(data, torch.Tensor) else _device()
    x = torch.as_tensor(data, dtype=torch.float64, device=dev).reshape(-1)
    zero = torch.zeros(x.numel(), dtype=torch.int32, device=dev)
    r = order_stats(zero, torch.ones_like(x), [0], value=x, sigma=sigma, maxiters=maxiters)
    return r['mean'][0], r['median'][0], r['std'][0]


def effectivearea_from_photonlist(photons, orders, n_photons, A_geom=1., ordercol='order'):
    """Summed probability per order / ``n_photons`` * ``A_geom``
    (gratings.py:333-363)."""
    dev = _device()
    order = _column(photons, ordercol, dev)
    prob = _column(photons, 'probability', dev)
    r = order_stats(order, prob, orders)
    return r['prob_sum'] / n_photons * A_geom


def _resolving_power(order, value, orders, zeropos, res_filter, prob, select=None):
    """gratings.py:255-270 on device columns; orders with 20 photons or
    fewer give NaN; zeroth order found from >= 20 photons when not given."""
    orders = np.asarray(orders)
    want = list(dict.fromkeys([int(x) for x in orders] + ([0] if zeropos is None else [])))
    r = order_stats(order, prob, want, value=value, res_filter=res_filter, select=select)
    idx = {o: i for i, o in enumerate(want)}
    if zeropos is None:
        i0 = idx[0]
        if r['count'][i0] < 20:
            raise AnalysisError('Too few photons in list to determine position of order 0 '
                                'automatically.')
        zeropos = r['mean'][i0]
    pos = np.array([r['mean'][idx[int(o)]] if r['count'][idx[int(o)]] > 20 else np.nan
                    for o in orders], dtype=float)
    std = np.array([r['std'][idx[int(o)]] if r['count'][idx[int(o)]] > 20 else np.nan
                    for o in orders], dtype=float)
    with np.errstate(invalid='ignore', divide='ignore'):
        res = np.abs(pos - zeropos) / (std * RES_FWHM)
    return res, pos, std


def resolvingpower_from_photonlist(photons, orders, col='proj_x', zeropos=None, ordercol='order'):
    """Resolving power, mean position and width of ``col`` per order
    (gratings.py:215-270): sigma-clipped statistics; NaN for orders with 20
    photons or fewer; the zeroth order position from order 0 when
    ``zeropos`` is None (AnalysisError below 20 photons)."""
    dev = _device()
    order = _column(photons, ordercol, dev)
    value = _column(photons, col, dev)
    prob = torch.ones(order.numel(), dtype=torch.float64, device=dev)
    return _resolving_power(order, value, orders, zeropos, False, prob)


def average_R_Aeff(r, aeff, axis=None):
    """Summed effective area and area-weighted average of the finite
    resolving powers (gratings.py:187-212)."""
    aeff_avg = aeff.sum(axis=axis)
    res_avg = np.ma.average(np.ma.masked_invalid(r), weights=aeff, axis=axis)
    return res_avg, aeff_avg


class CaptureResAeff:
    """Resolving power and effective area per order of a simulation
    (gratings.py:389-503).  ``aeff_filter``/``res_filter`` select photons;
    the defaults use every photon for the area and photons with a finite
    dispersion coordinate and probability > 0 for the resolving power.
    Overrides return boolean arrays or tensors over the photon list, or
    None for "all photons"."""

    def __init__(self, A_geom=1, order_col='order', orders=np.arange(-10, 11),
                 dispersion_coord='det_x', zeropos=None):
        self.A_geom = A_geom
        self.order_col = order_col
        self.orders = np.asanyarray(orders)
        self.dispersion_coord = dispersion_coord
        self.zeropos = zeropos

    def aeff_filter(self, photons):
        return None

    def res_filter(self, photons):
        return None

    def __call__(self, photons, n_photons=None):
        dev = _device()
        order = _column(photons, self.order_col, dev)
        prob = _column(photons, 'probability', dev)
        value = _column(photons, self.dispersion_coord, dev)
        if n_photons is None:
            n_photons = order.numel()
        sel = self.aeff_filter(photons)
        rsel = self.res_filter(photons)
        if sel is None and rsel is None:
            # one pass: the per-order probability sums of the resolving-power
            # pass run over all photons of the order, as the area needs
            res, aeff = self._both(order, prob, value, n_photons)
        else:
            r = order_stats(order, prob, self.orders, select=sel)
            aeff = r['prob_sum'] / n_photons * self.A_geom
            try:
                res, pos, std = _resolving_power(order, value, self.orders, self.zeropos,
                                                 rsel is None, prob, select=rsel)
            except AnalysisError:
                res = np.nan * np.ones(len(self.orders))
        disporders = self.orders != 0
        avggratres, aeffgrat = average_R_Aeff(res[disporders], aeff[disporders])
        aeff = np.asarray(aeff)
        aeff0 = np.sum(aeff[~disporders])
        return {'Aeff0': aeff0, 'Aeffgrat': aeffgrat, 'Aeff': aeff, 'Rgrat': avggratres, 'R': res}


    def _both(self, order, prob, value, n_photons):
        """Areas and resolving powers from one mx_order_stats call (default
        filters)."""
        orders = np.asarray(self.orders)
        want = list(dict.fromkeys([int(x) for x in orders] + ([0] if self.zeropos is None else [])))
        r = order_stats(order, prob, want, value=value, res_filter=True)
        idx = {o: i for i, o in enumerate(want)}
        aeff = np.array([r['prob_sum'][idx[int(o)]] for o in orders]) / n_photons * self.A_geom
        zeropos = self.zeropos
        if zeropos is None:
            if r['count'][idx[0]] < 20:
                return np.nan * np.ones(len(orders)), aeff
            zeropos = r['mean'][idx[0]]
        ok = np.array([r['count'][idx[int(o)]] > 20 for o in orders])
        pos = np.where(ok, [r['mean'][idx[int(o)]] for o in orders], np.nan)
        std = np.where(ok, [r['std'][idx[int(o)]] for o in orders], np.nan)
        with np.errstate(invalid='ignore', divide='ignore'):
            res = np.abs(pos - zeropos) / (std * RES_FWHM)
        return res, aeff


class CaptureResAeff_CCDgaps(CaptureResAeff):
    """As CaptureResAeff, with only photons where ``aeff_filter_col`` >= 0
    counted for the effective area (gratings.py:505-553)."""

    def __init__(self, aeff_filter_col='CCD_ID', **kwargs):
        super().__init__(**kwargs)
        self.aeff_filter_col = aeff_filter_col

    def aeff_filter(self, photons):
        return _column(photons, self.aeff_filter_col, _device()) >= 0


def resolvingpower_from_photonlist_robust(lphotons, orders, cols, zeropositions, ordercol='order'):
    """Resolving power of each order from several photon lists / columns, the
    lowest per order with its position and width (gratings.py:273-330)."""
    if len(cols) != len(zeropositions):
        raise ValueError('Number of elements in cols and zeropositions is not the same.')
    if len(cols) != len(lphotons):
        raise ValueError('Number of elements in cols and photon lists is not the same.')
    results = np.zeros((len(cols), len(orders)))
    positions = np.zeros_like(results)
    stds = np.zeros_like(results)
    for i, (photons, col, zeropos) in enumerate(zip(lphotons, cols, zeropositions)):
        results[i], positions[i], stds[i] = resolvingpower_from_photonlist(
            photons, orders, col=col, zeropos=zeropos, ordercol=ordercol)
    ind = np.argmin(results, axis=0)
    ind_n = np.arange(results.shape[1])
    return results[ind, ind_n], positions[ind, ind_n], stds[ind, ind_n]


def weighted_per_order(data, orders, energy, gratingeff):
    """Mean over the orders of a per-order, per-energy quantity (e.g. the
    resolving power of each order at each energy), weighted by the
    probability that ``gratingeff`` (an ``EfficiencyFile``) diffracts a
    photon of that energy into that order, linearly interpolated in energy
    (marxs/analysis/gratings.py:133-178; masked values do not count).

    data: (n_orders, n_energies); returns (n_energies,) masked array."""
    data = np.asanyarray(data)
    if data.shape[0] != len(orders):
        raise ValueError('First dimension of "data" must match length of "orders".')
    if data.shape[1] != len(energy):
        raise ValueError('Second dimension of "data" must match length of "energy".')
    table_orders = np.asarray(gratingeff.orders)
    # the efficiency column of each requested order (exactly one)
    hits = table_orders[None, :] == np.asarray(orders)[:, None]
    for o, h in zip(orders, hits):
        if h.sum() != 1:
            raise KeyError('No data for order {0} in gratingeff'.format(o))
    cols = np.argmax(hits, axis=1)
    by_energy = np.argsort(gratingeff.energy)
    e_tab = np.asarray(gratingeff.energy)[by_energy]
    p_tab = np.asarray(gratingeff.prob)[by_energy][:, cols]
    weights = np.stack([np.interp(energy, e_tab, p_tab[:, i]) for i in range(len(cols))])
    return np.ma.average(data, axis=0, weights=weights.astype(data.dtype, copy=False))


def identify_photon_in_subaperture(angle, max_ang, ang_0=np.pi / 2):
    """Photons in the two mirrored sectors ang_0 +- max_ang and
    2 pi - ang_0 +- max_ang (gratings.py:366-386; numpy's isclose with its
    default relative tolerance, as the reference)."""
    a = np.mod(angle, 2 * np.pi)
    return np.isclose(a, ang_0, atol=max_ang) | np.isclose(a, 2 * np.pi - ang_0, atol=max_ang)


# ------------------------------------------------------ traces for analysis
def _device_photons(photons):
    from .tracer import DevicePhotons
    if isinstance(photons, DevicePhotons):
        return photons
    return DevicePhotons.from_table(photons, device=_device())


def _trace(tree, src, precision=64):
    """A copy of device photons ``src`` traced through ``tree`` (the
    reference's ``tree(photons.copy())``); columns already on ``src`` keep
    their values where the tree does not write them."""
    from .tracer import CompiledInstrument
    ci = CompiledInstrument(tree, device=src.device)
    out = src.clone()
    t = ci.tables
    existing = [c for c in list(out.fcols) + list(out.icols) if c in t.fcols or c in t.icols]
    ci.trace_device(out, precision=precision, existing=existing)
    return out


def _select(dp, mask):
    """The photons of ``dp`` where ``mask`` (a device bool tensor) holds."""
    from .tracer import DevicePhotons
    idx = torch.nonzero(mask).squeeze(1)
    out = DevicePhotons.__new__(DevicePhotons)
    out.device, out.n, out.has_pos = dp.device, int(idx.numel()), dp.has_pos
    for a in ('pos', 'dir', 'pol'):
        setattr(out, a, getattr(dp, a)[:, idx].contiguous())
    out.energy = dp.energy[idx].contiguous()
    out.prob = dp.prob[idx].contiguous()
    out.fcols = {k: v[idx].contiguous() for k, v in dp.fcols.items()}
    out.icols = {k: v[idx].contiguous() for k, v in dp.icols.items()}
    return out


def sigma_clipped_std(photons, colname='det_x', sigma=3., maxiters=5):
    """Standard deviation of the sigma-clipped column (analysis.py:9-24)."""
    return sigma_clipped_stats(_column(photons, colname, _device()), sigma=sigma,
                               maxiters=maxiters)[2]


def mean_width_2d(photons):
    """Mean distance of (det_x, det_y) from their means (analysis.py:27-39)."""
    dev = _device()
    x = _column(photons, 'det_x', dev).to(torch.float64)
    y = _column(photons, 'det_y', dev).to(torch.float64)
    r = torch.sqrt((x - x.mean()) ** 2 + (y - y.mean()) ** 2)
    return float(r.sum() / r.numel())


def find_best_detector_position(photons, objective_func=sigma_clipped_std,
                                objective_func_args={'colname': 'det_x'},
                                orientation=np.eye(3), **kwargs):
    """Position of best focus (analysis.py:42-81): ``minimize_scalar`` over x
    of ``objective_func`` on the photons traced (on the GPU, photons kept in
    HBM) to a flat detector at ``orientation . (x, 0, 0)``.  The objective
    receives a :class:`~marxs_amd.tracer.DevicePhotons`."""
    import scipy.optimize
    from .optics import FlatDetector
    src = _device_photons(photons)

    def width(x):
        det = FlatDetector(position=np.dot(orientation, np.array([x, 0, 0])),
                           orientation=orientation, zoom=1e5, pixsize=1.)
        return objective_func(_trace(det, src), **objective_func_args)

    return scipy.optimize.minimize_scalar(width, **kwargs)


def detected_fraction(photons, labels, col='order'):
    """Summed probability per label / number of photons (analysis.py:84-107)."""
    dev = _device()
    lab = _column(photons, col, dev)
    r = order_stats(lab, _column(photons, 'probability', dev), labels)
    return r['prob_sum'] / lab.numel()


def resolvingpower_per_order(gratings, photons, orders, detector=None, colname='det_x'):
    """Resolving power and FWHM of every order (gratings.py:26-130): the same
    photons (ready to hit the gratings) are sent into one order at a time
    (the order selectors of ``gratings`` are replaced, as in the reference),
    traced to the detector and measured with sigma-clipped statistics.
    ``detector``: an element, a ``RowlandTorus`` (a tube on its Rowland
    circle, column ``detpix_x``) or None (a flat detector at the x of best
    focus of each order, found by ``find_best_detector_position``).  Returns
    (res, fwhm, info)."""
    from .geometry import Cylinder
    from .optics import CircularDetector, FlatDetector, OrderSelector
    from .rowland import RowlandTorus
    orders = np.asarray(orders)
    res = np.zeros_like(orders, dtype=float)
    fwhm = np.zeros_like(orders, dtype=float)
    info = {}
    src = _device_photons(photons)
    if detector is None:
        info['method'] = 'Detector position numerically optimized'
        info['fit_results'] = []
        col = 'det_x'
        zeropos = 0.
    elif isinstance(detector, RowlandTorus):
        det = CircularDetector(geometry=Cylinder.from_rowland(
            detector, width=1e5, rotation=np.pi, kwargs={'phi_lim': [-np.pi / 2, np.pi / 2]}))
        info['method'] = 'Circular detector on Rowland circle'
        col = 'detpix_x'
        pg = _trace(det, src)
        pg = _select(pg, pg.prob > 0.)
        zeropos = sigma_clipped_stats(_column(pg, col, src.device))[0]
    else:
        det = detector
        info['method'] = 'User defined detector'
        col = colname
        zeropos = sigma_clipped_stats(_column(_trace(det, src), col, src.device))[0]
    for i, order in enumerate(orders):
        gratingeff = OrderSelector([order])
        gratings.elem_args['order_selector'] = gratingeff
        for elem in gratings.elements:
            elem.order_selector = gratingeff
        pg = _trace(gratings, src)
        keep = (_orders_i32(_column(pg, 'order', src.device)) == int(order)) & (pg.prob > 0.)
        pg = _select(pg, keep)
        if detector is None:
            xbest = find_best_detector_position(pg, objective_func=sigma_clipped_std)
            info['fit_results'].append(xbest)
            det = FlatDetector(position=np.array([xbest.x, 0, 0]), zoom=1e5)
        pg = _trace(det, pg)
        meanpos, medianpos, stdpos = sigma_clipped_stats(_column(pg, col, src.device))
        fwhm[i] = RES_FWHM * stdpos
        res[i] = np.abs((meanpos - zeropos) / fwhm[i])
    return res, fwhm, info
