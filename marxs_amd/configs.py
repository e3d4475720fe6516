"""Instrument and photon builders for the benchmark workloads.

Geometry is built with the public set-up API (``Sequence``, ``Parallel``,
``GratingArrayStructure``, ``RectangularGrid``), exactly as a user script for
the reference would build it (e.g. ``marxs/tests/test_analysis.py:30-87``).

Workloads
---------
``flat_grid``    1e5 photons, FixedPointing on-axis, 0.3-2.0 keV, annulus
                 300-600 mm at x = 10000 mm, PerfectLens f = 10000 mm, 11 x 11
                 FlatGrating facets 60 x 60 mm at x = 5000 mm (d = 2e-4 mm,
                 orders -1/0/+1 with p = .25/.5/.25), one 8192^2 FlatDetector
                 of 0.024 mm pixels at x = 0.
``cat_rowland``  CAT facets 25 x 25 mm (d = 2e-4 mm, blaze 1.5 deg) placed by
                 ``GratingArrayStructure`` on RowlandTorus(R=5900, r=6000) and
                 16 CCDs of 1024^2 x 0.024 mm placed by ``RectangularGrid`` on
                 the same torus near the focus; orders -12..+2 from a
                 synthetic efficiency table (Gaussian in m around -6, sigma 3,
                 0.1 keV grid).  A Rowland circle of radius 6000 mm through
                 the focus puts the on-axis facet at x = 12000 mm, so aperture
                 and PerfectLens sit at x = 12500 mm with f = 12500 mm (focus
                 at the origin); the torus is shifted by +100 mm in x so that
                 it passes through the focus.
"""
import math

import numpy as np

from .mathutils import axangle2mat
from .optics import (CircleAperture, PerfectLens, FlatGrating, CATGrating,
                     OrderSelector, EfficiencyFile, FlatDetector)
from .rowland import RowlandTorus, GratingArrayStructure, RectangularGrid
from .simulator import Sequence, Parallel

__all__ = ['flat_grid_instrument', 'cat_rowland_instrument', 'cat_stack_instrument', 'blaze_efficiency',
           'design_variants',
           'onaxis_photons', 'line_spectrum_energies', 'uniform_energies',
           'CAT_ORDERS', 'LINES']

CAT_ORDERS = np.arange(-12, 3)
LINES = np.array([0.5, 0.65, 0.9, 1.2])


def flat_grid_instrument(det_shift=0.):
    """11 x 11 FlatGrating grid between a PerfectLens and one large detector.

    The zero order focuses onto the detector centre, which is the corner of
    four pixels (8192 pixels, centre pixel coordinate 4095.5);
    ``det_shift`` (mm) moves the detector along y and z, e.g. by a third of a
    pixel (0.008 mm) so that the focus lies inside a pixel."""
    aper = CircleAperture(position=[10000, 0, 0], zoom=[1, 600, 600], r_inner=300)
    lens = PerfectLens(focallength=10000, position=[10000, 0, 0], zoom=[1, 700, 700])
    yz = np.arange(-300, 301, 60, dtype=float)
    pos = [[5000., y, z] for y in yz for z in yz]
    grat = Parallel(elem_class=FlatGrating,
                    elem_args={'zoom': [1, 30, 30], 'd': 2e-4,
                               'order_selector': OrderSelector([-1, 0, 1], [.25, .5, .25])},
                    elem_pos={'position': pos}, id_col='facet')
    det = FlatDetector(pixsize=0.024, position=[0, det_shift, det_shift], zoom=[1, 98.304, 98.304])
    return Sequence(elements=[aper, lens, grat, det])


def blaze_efficiency(energy_grid=None, orders=CAT_ORDERS, center=-6., sigma=3.):
    """Synthetic smooth order efficiency: Gaussian in order number, total
    efficiency rising smoothly from 0.46 at 0.3 keV to 0.8 at 2 keV."""
    if energy_grid is None:
        energy_grid = np.round(np.arange(0.3, 2.0001, 0.1), 10)
    w = np.exp(-0.5 * ((orders - center) / sigma) ** 2)
    w = w / w.sum()
    total = 0.4 + 0.2 * energy_grid
    prob = total[:, None] * w[None, :]
    return EfficiencyFile.from_arrays(energy_grid, prob, orders)


def blaze_interp_efficiency(orders=CAT_ORDERS, d_nm=200., sigma=2.5):
    """Synthetic blaze-dependent order efficiency as an
    ``InterpolateEfficiencyTable`` on wavelength 0.6-4.2 nm x blaze angle
    0.5-2.5 deg: Gaussian in order number around the blaze condition
    m = -2 d sin(blaze) / lambda (sigma 2.5 orders), total 0.55 + 0.1 blaze/deg."""
    from .optics import InterpolateEfficiencyTable
    wave = np.round(np.arange(0.6, 4.2001, 0.2), 10)
    theta = np.deg2rad(np.arange(0.5, 2.5001, 0.25))
    m = np.asarray(orders, dtype=float)[:, None, None]
    center = -2 * d_nm * np.sin(theta)[None, None, :] / wave[None, :, None]
    w = np.exp(-0.5 * ((m - center) / sigma) ** 2)
    w = w / w.sum(axis=0, keepdims=True)
    total = 0.55 + 0.1 * np.rad2deg(theta)[None, None, :]
    return InterpolateEfficiencyTable.from_arrays(wave, theta, orders, w * total)


def _gas_count(radius, d_element, phi):
    """Number of facets GratingArrayStructure places for these arguments
    (its yz layout only, without solving for positions on the torus)."""
    probe = GratingArrayStructure.__new__(GratingArrayStructure)
    probe.radius, probe.d_element, probe.phi = radius, d_element, phi
    return len(probe.elemposyz()[0])


def _gas_with_n(rowland, n, radius, d_element, elem_args, elem_class=CATGrating):
    """GratingArrayStructure over a wedge around +z, widened until it holds at
    least ``n`` facets, then cut to the first ``n``."""
    lo, hi = 0.01, math.pi
    best = None
    if _gas_count(radius, d_element, [-hi, hi]) < n:
        raise ValueError(f'cannot place {n} facets in radius {radius}')
    for _ in range(60):
        half = 0.5 * (lo + hi)
        c = _gas_count(radius, d_element, [-half, half])
        if c >= n:
            best, hi = half, half
        else:
            lo = half
        if c == n:
            break
    half = best if best is not None else hi
    gas = GratingArrayStructure(rowland=rowland, d_element=d_element, radius=radius,
                                phi=[-half, half], elem_class=elem_class,
                                elem_args=dict(elem_args), id_col='facet')
    if len(gas.elem_pos) > n:
        gas.elem_pos = gas.elem_pos[:n]
        gas.elem_uncertainty = gas.elem_uncertainty[:n]
        gas.generate_elements()
    return gas


def cat_rowland_instrument(n_facets=200, R=5900., r=6000., blaze_deg=1.5,
                           n_ccd=16, ccd_y=(320., 720.), efficiency=None,
                           focal=12500., d=2e-4, facet_class=CATGrating, facet_kwargs=None):
    """Aperture, PerfectLens, a CAT facet array on a Rowland torus and a row of
    CCDs on the Rowland circle.  Returns (Sequence, parts dict).
    ``facet_class``/``facet_kwargs``: e.g. ``catgrating.CATL1L2Stack`` with
    ``{'l1_transmission': ...}`` for facets with their support structure."""
    aper = CircleAperture(position=[focal, 0, 0], zoom=[1, 600, 600], r_inner=300)
    lens = PerfectLens(focallength=focal, position=[focal, 0, 0], zoom=[1, 700, 700])
    # Rowland circle through the focus (origin) and x = 2 r on axis
    rowland = RowlandTorus(R, r, position=[r - R, 0, 0])
    blaze = axangle2mat([0, 0, 1], np.deg2rad(blaze_deg))
    eff = efficiency if efficiency is not None else blaze_efficiency()
    x_on = 2 * r
    r_in = 300. * x_on / focal
    r_out = 600. * x_on / focal
    facet_args = {'zoom': [1, 12.5, 12.5], 'd': d, 'order_selector': eff,
                  'orientation': blaze}
    facet_args.update(facet_kwargs or {})
    gas = _gas_with_n(rowland, n_facets, [r_in + 13.5, r_out - 13.5], [27., 27.], facet_args,
                      elem_class=facet_class)
    pitch = 24.576 + 0.5
    ccd = RectangularGrid(rowland=rowland, elem_class=FlatDetector,
                          elem_args={'pixsize': 0.024, 'zoom': [1, 12.288, 12.288]},
                          d_element=[pitch, pitch],
                          y_range=[ccd_y[0], ccd_y[0] + n_ccd * pitch - 1e-6],
                          guess_distance=-(r - R), id_col='CCD')
    inst = Sequence(elements=[aper, lens, gas, ccd])
    return inst, {'aperture': aper, 'lens': lens, 'gratings': gas, 'ccds': ccd,
                  'rowland': rowland, 'efficiency': eff}


def cat_stack_instrument(n_facets=600, **kw):
    """``cat_rowland_instrument`` with ``CATL1L2Stack`` facets (membrane,
    quality factor, L1 cross-support with the packaged Si transmission table,
    L2 absorption and diffraction).  Returns (Sequence, parts dict)."""
    from . import catgrating
    return cat_rowland_instrument(n_facets, facet_class=catgrating.CATL1L2Stack,
                                  facet_kwargs={'l1_transmission': catgrating.l1transtab_path}, **kw)


def design_variants(n_variants=64, rng=None, R_range=(5000., 7000.), blaze_range=(1.0, 2.0),
                    facet_range=(50, 800), n_ccd=16):
    """Instrument design variants for a batched trade study: Rowland torus
    R ~ U(R_range) mm with r = R + 100 mm, blaze ~ U(blaze_range) deg, facet
    count ~ U{facet_range}; aperture and lens at x = 2 r + 500 mm (focus at
    the origin) and the CCD row scaled with r so that it covers the same
    dispersion angles as the r = 6000 mm layout.  Returns a list of
    (Sequence, parts, parameters) triples."""
    rng = np.random.default_rng() if rng is None else rng
    out = []
    for _ in range(n_variants):
        R = float(rng.uniform(*R_range))
        blaze = float(rng.uniform(*blaze_range))
        n = int(rng.integers(facet_range[0], facet_range[1] + 1))
        r = R + 100.
        scale = r / 6000.
        inst, parts = cat_rowland_instrument(n, R=R, r=r, blaze_deg=blaze, n_ccd=n_ccd,
                                             ccd_y=(320. * scale, 720. * scale),
                                             focal=2 * r + 500.)
        out.append((inst, parts, {'R': R, 'r': r, 'blaze_deg': blaze, 'n_facets': n}))
    return out


def uniform_energies(u, lo=0.3, hi=2.0):
    return lo + (hi - lo) * u


def line_spectrum_energies(u, u_cont, lines=LINES, frac_cont=0.2, lo=0.3, hi=2.0):
    """20 % flat continuum lo..hi, 80 % equally in the lines."""
    u = np.asarray(u)
    cont = u < frac_cont
    idx = np.minimum(((u - frac_cont) / ((1 - frac_cont) / len(lines))).astype(int), len(lines) - 1)
    e = np.where(cont, lo + (hi - lo) * np.asarray(u_cont), np.asarray(lines)[np.maximum(idx, 0)])
    return e


def onaxis_photons(energy, u_pol):
    """FixedPointing on axis with default roll (marxs/source/pointing.py:101-145):
    dir = (-1, 0, 0), pol = cos(phi) (0, 0, 1) + sin(phi) (0, 1, 0) with
    phi = 2 pi u_pol.  Returns a dict of host arrays (N,3)/(N,)."""
    energy = np.asarray(energy, dtype=float)
    n = len(energy)
    phi = 2 * np.pi * np.asarray(u_pol, dtype=float)
    d = np.zeros((n, 3))
    d[:, 0] = -1.
    pol = np.zeros((n, 3))
    pol[:, 1] = np.sin(phi)
    pol[:, 2] = np.cos(phi)
    return {'dir': d, 'pol': pol, 'energy': energy, 'prob': np.ones(n)}
