"""CAT gratings with their support structure (the ``marxs.missions.mitsnl``
module of the reference, ``marxs/missions/mitsnl/catgrating.py``).

The layers of a fabricated CAT grating facet act at one place and are traced
as one *stack* element (``FlatStack``): the grating membrane
(``NonParallelCATGrating``), a Debye-Waller quality factor on the
diffraction probability (``QualityFactor``), the L1 cross-support as a
second, perpendicular CAT grating whose bars absorb a fraction of the photons
(``L1``), shadowing by the hexagonal L2 support (``L2Abs``) and a small
Gaussian broadening by diffraction on the L2 openings (``L2Diffraction``).
``CATL1L2Stack`` combines them as the reference does
(``catgrating.py:327-374``).

Lengths are plain numbers in mm (the reference uses astropy quantities).
The L1 transmission table (transmission of 1 um of Si against energy, from
the Henke tables) ships with the package as ``data/SiTransmission.csv``,
the same table and layout the reference reads at import
(``catgrating.py:41``); ``l1_transmission`` overrides it with
``(energy_keV, transmission)`` or the path of another table in that layout
(energy in eV).
"""
import math
import os

import numpy as np

from .optics import (CATGrating, FlatOpticalElement, FlatStack, NonParallelCATGrating,
                     OrderSelector, read_efficiency_table)
from .scatter import RandomGaussianScatter

__all__ = ['l1transtab_path', 'l1_order_selector', 'l1_dims', 'l2_dims', 'qualityfactor', 'd',
           'QualityFactor', 'L1', 'L2Abs', 'L2Diffraction', 'CATL1L2Stack',
           'check_lx_dims', 'read_transmission_table', 'catsupportbars']

d = 0.0002
"""Spacing of grating bars (mm)."""

l1transtab_path = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'data',
                               'SiTransmission.csv')
"""Transmission of 1 um of Si against energy (eV), the L1 default
(catgrating.py:41)."""

l1_order_selector = OrderSelector(orderlist=np.array([-4, -3, -2, -1, 0, 1, 2, 3, 4]),
                                  p=np.array([0.006, 0.0135, 0.022, 0.028, 0.861, 0.028,
                                              0.022, 0.0135, 0.006]))
"""Order selector for diffraction on L1 (catgrating.py:45-52)."""

l1_dims = {'bardepth': 0.004, 'period': 0.005, 'barwidth': 0.0009}
"""L1 support bars perpendicular to the grating bars, mm (catgrating.py:54-55)."""

l2_dims = {'bardepth': 0.5, 'period': 0.9622504, 'barwidth': 0.0962250}
"""Hexagonal L2 support, mm (catgrating.py:57-58)."""

qualityfactor = {'d': 0.2, 'sigma': 0.00175}
"""Debye-Waller scaling of grating efficiencies, mm (catgrating.py:60-61)."""

KEV_MM = 1.2398419843320026e-6   # h c / keV in mm (SI-exact h, c, e)


def check_lx_dims(lx_dims):
    """Bar width must be smaller than the period (catgrating.py:163-166)."""
    if not (lx_dims['barwidth'] < lx_dims['period']):
        raise ValueError('Period of grating must be larger than bar width.')


def read_transmission_table(path):
    """(energy in keV, transmission) from a table in the layout of the
    reference's ``data/SiTransmission.csv`` (energy column in eV)."""
    names, units, data = read_efficiency_table(path)
    energy = data[:, 0]
    if units.get(names[0], 'eV') == 'eV':
        energy = energy * 0.001
    return energy, data[:, 1]


class QualityFactor(FlatOpticalElement):
    """Probability factor exp(-(2 pi sigma / d)^2) ** (order^2) of the last
    grating interaction (catgrating.py:147-160)."""

    def __init__(self, qualityfactor=qualityfactor, **kwargs):
        self.factor = math.exp(-(2 * math.pi * qualityfactor['sigma'] / qualityfactor['d']) ** 2)
        super().__init__(**kwargs)


class L1(CATGrating):
    """L1 cross-support as a CAT grating (catgrating.py:169-212): every photon
    is diffracted, then with probability 1 - open fraction it went through a
    support bar instead: direction and polarization unchanged, ``order_L1``
    0, probability = transmission of the bar at its energy."""
    blaze_name = 'blaze_L1'
    order_name = 'order_L1'

    def __init__(self, l1_dims=l1_dims, l1_transmission=None, **kwargs):
        check_lx_dims(l1_dims)
        self.openfraction = 1 - l1_dims['barwidth'] / l1_dims['period']
        if l1_transmission is None:
            l1_transmission = l1transtab_path
        if isinstance(l1_transmission, str):
            l1_transmission = read_transmission_table(l1_transmission)
        energy, trans = (np.asarray(x, dtype=float) for x in l1_transmission)
        if np.any(np.diff(energy) <= 0):
            raise ValueError('transmission table energies must increase')
        with np.errstate(divide='ignore'):
            logtrans = np.log(trans)
        # scale 1 um of Si to the bar depth (in um)
        self.trans_energy = energy
        self.trans_value = np.exp(logtrans * (l1_dims['bardepth'] * 1000.))
        kwargs['d'] = l1_dims['period']
        super().__init__(**kwargs)

    def transfunc(self, energy):
        """Linear interpolation in the table with scipy ``interp1d``'s
        arithmetic (left interval index, slope times offset plus lower
        value); energies outside the table raise like ``interp1d``."""
        x, y = self.trans_energy, self.trans_value
        e = np.asarray(energy, dtype=float)
        if np.any(e < x[0]) or np.any(e > x[-1]):
            raise ValueError('A value in x_new is outside the interpolation range.')
        hi = np.clip(np.searchsorted(x, e), 1, len(x) - 1)
        lo = hi - 1
        slope = (y[hi] - y[lo]) / (x[hi] - x[lo])
        return slope * (e - x[lo]) + y[lo]


class L2Abs(FlatOpticalElement):
    """L2 absorption and shadowing (catgrating.py:215-249): probability
    factor = ((period - barwidth)/period)^2 - bardepth (period - barwidth)
    sin(angle) / (period^2 sqrt(3) / 2), angle between ray and normal."""

    def __init__(self, l2_dims=l2_dims, **kwargs):
        check_lx_dims(l2_dims)
        self.bardepth = l2_dims['bardepth']
        self.period = l2_dims['period']
        self.barwidth = l2_dims['barwidth']
        super().__init__(**kwargs)
        self.innerfree = self.period - self.barwidth

    def coefficients(self):
        openfraction = (self.innerfree / self.period) ** 2
        totalarea = self.period ** 2 / 2 * np.sqrt(3)
        return openfraction, self.bardepth * self.innerfree / totalarea


class L2Diffraction(RandomGaussianScatter):
    """Gaussian broadening by the L2 openings (catgrating.py:252-271):
    sigma = 1.22 * 0.4 * arcsin(lambda / (period - barwidth))."""
    scattername = 'L2Diffraction'

    def __init__(self, l2_dims=l2_dims, **kwargs):
        check_lx_dims(l2_dims)
        self.innerfree = l2_dims['period'] - l2_dims['barwidth']
        kwargs['scatter'] = 'l2'
        super().__init__(**kwargs)


class CATL1L2Stack(FlatStack):
    """A fabricated CAT grating facet (catgrating.py:327-374): membrane,
    quality factor, L1, L2 absorption and L2 diffraction at one place."""

    def __init__(self, l1_dims=l1_dims, l2_dims=l2_dims, l1_order_selector=l1_order_selector,
                 qualityfactor=qualityfactor, l1_transmission=None, **kwargs):
        kwargs['elements'] = [NonParallelCATGrating, QualityFactor, L1, L2Abs, L2Diffraction]
        groove_angle = kwargs.pop('groove_angle', 0.)
        kwargs['keywords'] = [{'order_selector': kwargs.pop('order_selector'),
                               'd': kwargs.pop('d', d), 'groove_angle': groove_angle},
                              {'qualityfactor': qualityfactor},
                              {'l1_dims': l1_dims, 'order_selector': l1_order_selector,
                               'groove_angle': np.pi / 2. + groove_angle,
                               'l1_transmission': l1_transmission},
                              {'l2_dims': l2_dims},
                              {'l2_dims': l2_dims}]
        super().__init__(**kwargs)


def catsupportbars(photons):
    """The metal frame holding the facets absorbs every photon that passed no
    facet (catgrating.py:314-324): probability 0 where ``facet`` < 0 (or
    everywhere without a ``facet`` column).  A host-side hook."""
    if 'facet' in photons.colnames:
        p = np.asarray(photons['probability'], dtype=float).copy()
        p[np.asarray(photons['facet']) < 0] = 0.
        photons['probability'] = p
    else:
        photons['probability'] = np.zeros(len(photons))
    return photons
