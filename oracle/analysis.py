"""numpy restatement of the per-order figures of merit (TEST INFRASTRUCTURE
ONLY): ``marxs/analysis/gratings.py:215-270`` (resolving power from a photon
list), ``333-363`` (effective area), ``187-212`` (``average_R_Aeff``),
``389-503`` (``CaptureResAeff``).

The reference computes the clipped statistics with astropy's
``sigma_clipped_stats`` (astropy >= 6.0.1, ``pyproject.toml:23``; astropy is
not importable here).  :func:`sigma_clipped_stats` restates its documented
algorithm with the defaults the reference uses (sigma 3, at most 5
iterations, centre = median, spread = standard deviation with ddof 0):
drop non-finite values; repeat {bounds = median -/+ 3 std of the kept
values; keep values inside the bounds, inclusive} until nothing changes or 5
rounds ran; report mean, median and std of what is kept.

How it is pinned (tests/test_oracle_analysis.py): the reference's known
answers (``analysis/tests/test_grating.py``: effective areas 0 / 0.04 / 0.6;
``CaptureResAeff`` areas [2.5, 0, 7.5] with R NaN for an empty order and
R ~ 50 for a unit-width order at 100; zero probabilities and non-finite
coordinates excluded from R only; the resolving power of a ThinLens +
FlatGrating set-up above 10 in orders -1/+1 and below 1 in order 0), a
closed form for the clipping (uniform values plus far outliers give exactly
numpy's mean / median / std of the uniform values after two rounds), and a
brute-force loop over single-value removals.
"""
import numpy as np

RES_FWHM = 2.3548     # FWHM / sigma used by the reference (gratings.py:268)


class AnalysisError(Exception):
    pass


def sigma_clipped_stats(data, sigma=3., maxiters=5):
    """(mean, median, std) after sigma clipping, see the module docstring."""
    x = np.asarray(data, dtype=float).ravel()
    x = x[np.isfinite(x)]
    nchanged, it = 1, 0
    while nchanged != 0 and it < maxiters:
        it += 1
        size = x.size
        with np.errstate(invalid='ignore'):
            cen = np.median(x) if size else np.nan
            std = np.std(x) if size else np.nan
            lo = cen - std * sigma
            hi = cen + std * sigma
            x = x[(x >= lo) & (x <= hi)]
        nchanged = size - x.size
    if x.size == 0:
        return np.nan, np.nan, np.nan
    return np.mean(x), np.median(x), np.std(x)


def effectivearea_from_photonlist(order, prob, orders, n_photons, A_geom=1.):
    """gratings.py:333-363: summed probability per order / n * A_geom."""
    order, prob = np.asarray(order), np.asarray(prob, dtype=float)
    aeff = np.zeros(len(orders))
    for i, o in enumerate(orders):
        aeff[i] = prob[order == o].sum()
    return aeff / n_photons * A_geom


def resolvingpower_from_photonlist(order, value, orders, zeropos=None):
    """gratings.py:215-270: sigma-clipped mean and std of the dispersion
    coordinate per order (NaN for 20 photons or fewer), the zeroth order
    position from order 0 when not given (at least 20 photons), R = |pos -
    zeropos| / (2.3548 std)."""
    order, value = np.asarray(order), np.asarray(value, dtype=float)
    if zeropos is None:
        ind = order == 0
        if ind.sum() < 20:
            raise AnalysisError('Too few photons in list to determine position of order 0 '
                                'automatically.')
        zeropos = sigma_clipped_stats(value[ind])[0]
    pos = np.zeros(len(orders))
    std = np.zeros(len(orders))
    for i, o in enumerate(orders):
        ind = order == o
        if ind.sum() > 20:
            m, med, s = sigma_clipped_stats(value[ind])
        else:
            m, s = np.nan, np.nan
        pos[i] = m
        std[i] = s
    with np.errstate(invalid='ignore', divide='ignore'):
        res = np.abs(pos - zeropos) / (std * RES_FWHM)
    return res, pos, std


def average_R_Aeff(r, aeff, axis=None):
    """gratings.py:187-212: summed area, area-weighted mean of the finite R."""
    aeff_avg = aeff.sum(axis=axis)
    res_avg = np.ma.average(np.ma.masked_invalid(r), weights=aeff, axis=axis)
    return res_avg, aeff_avg


def capture_res_aeff(order, value, prob, orders, A_geom=1., zeropos=None, n_photons=None):
    """CaptureResAeff.__call__ (gratings.py:464-503) with its default
    filters: every photon for the area, finite coordinate and probability >
    0 for the resolving power."""
    order, value, prob = np.asarray(order), np.asarray(value, dtype=float), np.asarray(prob)
    orders = np.asarray(orders)
    n_photons = len(order) if n_photons is None else n_photons
    aeff = effectivearea_from_photonlist(order, prob, orders, n_photons, A_geom)
    ind = np.isfinite(value) & (prob > 0)
    try:
        res, pos, std = resolvingpower_from_photonlist(order[ind], value[ind], orders, zeropos)
    except AnalysisError:
        res = np.nan * np.ones(len(orders))
    disp = orders != 0
    avggratres, aeffgrat = average_R_Aeff(res[disp], aeff[disp])
    aeff0 = np.sum(aeff[~disp])
    return {'Aeff0': aeff0, 'Aeffgrat': aeffgrat, 'Aeff': aeff, 'Rgrat': avggratres, 'R': res}
