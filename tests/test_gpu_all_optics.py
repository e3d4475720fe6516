"""The reference's interface checks over every optical element
(marxs/optics/tests/test_all_optics.py:145-231), for the elements this path
traces, on the GPU: probabilities only go down, and the polarization stays a
unit vector perpendicular to the direction.  Element set and arguments follow
test_all_optics.py:48-108 (angles in radian where the reference passes
astropy quantities; the L1 layers with the Si transmission table of the
reference's tests)."""
import os

import numpy as np
import pytest

import marxs_amd as mx
from marxs_amd import catgrating as CG
from marxs_amd.scatter import RadialMirrorScatter, RandomGaussianScatter

pytestmark = pytest.mark.gpu

SI = os.path.join(os.path.dirname(__file__), 'golden', 'SiTransmission.csv')
ARCSEC = np.pi / 180 / 3600


def _elements():
    torus = mx.RowlandTorus(0.5, 0.5)
    return [
        ('ThinLens', mx.ThinLens(focallength=100)),
        ('RectangleAperture', mx.RectangleAperture()),
        ('CircleAperture', mx.CircleAperture()),
        ('FlatDetector', mx.FlatDetector(pixsize=2.0, zoom=100.0)),
        ('CircularDetector', mx.CircularDetector()),
        ('FlatGrating', mx.FlatGrating(d=0.001, order_selector=mx.OrderSelector([0]))),
        ('GratingArrayStructure', mx.GratingArrayStructure(
            rowland=torus, d_element=[0.1, 0.1], radius=[0, 0.5], elem_class=mx.FlatGrating,
            elem_args={'zoom': 0.05, 'd': 0.002, 'order_selector': mx.OrderSelector([1])})),
        ('CircularMeshGrid', mx.CircularMeshGrid(
            rowland=torus, d_element=[0.05, 0.05], radius=[0., 0.5], elem_class=mx.FlatGrating,
            elem_args={'zoom': 0.05, 'd': 0.002, 'order_selector': mx.OrderSelector([2])})),
        ('RectangularGrid', mx.RectangularGrid(
            rowland=torus, elem_class=mx.FlatGrating,
            elem_args={'zoom': 0.04, 'd': 1e-3, 'order_selector': mx.OrderSelector([-1, 0])},
            d_element=[0.1, 0.1], y_range=[0, 0.1], z_range=[0, 0.1], guess_distance=1.0)),
        ('Sequence', mx.Sequence(elements=[])),
        ('RadialMirrorScatter', RadialMirrorScatter(inplanescatter=0.1 * ARCSEC)),
        ('RandomGaussianScatter', RandomGaussianScatter(scatter=6 * ARCSEC)),
        ('QualityFactor', CG.QualityFactor()),
        ('L1', CG.L1(order_selector=mx.OrderSelector([0]), l1_transmission=SI)),
        ('L2Abs', CG.L2Abs()),
        ('L2Diffraction', CG.L2Diffraction()),
        ('CATL1L2Stack', CG.CATL1L2Stack(order_selector=mx.OrderSelector([2]), l1_transmission=SI)),
        ('NonParallelCATGrating', mx.NonParallelCATGrating(order_selector=mx.OrderSelector([2]),
                                                           d=0.0001)),
    ]


ELEMENTS = _elements()


@pytest.fixture()
def photons(gpu):
    """500 photons on a slit at the origin moving along -x, polarizations
    spread over all angles, an existing order column (test_all_optics.py:
    118-137)."""
    rng = np.random.default_rng(4)
    p = mx.generate_test_photons(500)
    ang = rng.uniform(0, 2 * np.pi, 500)
    p['polarization'] = np.stack([np.zeros(500), np.cos(ang), np.sin(ang), np.zeros(500)], 1)
    p = mx.RectangleAperture()(p)
    p['order'] = np.zeros(500)
    return p


def _run(elem, photons):
    from marxs_amd.optics import FlatAperture
    if isinstance(elem, FlatAperture):
        photons = mx.PhotonTable({k: photons[k] for k in photons.colnames if k != 'pos'})
    return elem(photons)


@pytest.mark.parametrize('name, elem', ELEMENTS, ids=[n for n, e in ELEMENTS])
def test_prob(photons, name, elem):
    """test_all_optics.py:191-198"""
    photons['probability'] = np.full(len(photons), 1e-4)
    out = _run(elem, photons)
    assert np.all(np.asarray(out['probability']) <= 1e-4)


@pytest.mark.parametrize('name, elem', ELEMENTS, ids=[n for n, e in ELEMENTS])
def test_polarization_direction(photons, name, elem):
    """test_all_optics.py:200-214"""
    out = _run(elem, photons)
    d = np.asarray(out['dir'], dtype=float)
    pol = np.asarray(out['polarization'], dtype=float)
    assert np.allclose(np.einsum('...i,...i', d, pol), 0.)
    assert np.allclose(np.linalg.norm(pol, axis=1), 1)
