"""Shared fixtures.

Devices: the 'cpu' mode is the CPU checker (oracle/, registered here as the
reference `_cpu` class set); the 'gpu' mode is the product.  GPU parameters
carry the ``gpu`` marker, so ``-m "not gpu"`` runs the checker against the
reference's golden images and ``-m gpu`` runs the product against the same
images and against the checker.
"""

import math
import os
import pathlib
import sys

import numpy
import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

import fresnel_amd as fresnel  # noqa: E402
import oracle.backend  # noqa: E402

fresnel._register_cpu_module(oracle.backend)

GOLDEN = ROOT / 'tests' / 'golden'
REFERENCE_PNG = GOLDEN / 'reference_png'


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (runs the HIP backend)')


DEVICE_PARAMS = [
    pytest.param(('cpu', None), id='cpu(all)'),
    pytest.param(('gpu', 1), id='gpu(1)', marks=pytest.mark.gpu),
]


@pytest.fixture(scope='session', params=DEVICE_PARAMS)
def device_(request):
    """A Device of each mode (the GPU one only under the gpu marker)."""
    mode, n = request.param
    return fresnel.Device(mode=mode, n=n)


@pytest.fixture(scope='session')
def gpu_device():
    """The GPU device (tests using it must be marked gpu).

    Primary rays of pinhole and orthographic cameras take the per-tile candidate lists from the
    first render here (by default a tracer builds them once a few million primary rays have used one
    scene and camera, and takes the BVH search before): the parity tests render small images and
    are meant to compare the list search with the checker.  `test_tile_list_threshold` covers the
    default."""
    dev = fresnel.Device(mode='gpu', n=1)
    dev._device.tune('tile_min_rays', 1)
    return dev


@pytest.fixture(scope='session')
def cpu_device():
    """The CPU checker device."""
    return fresnel.Device(mode='cpu')


def test_lights():
    """Key and fill lights used by the reference test scenes (fresnel/test/conftest.py:25-44)."""
    key_polar, key_az = (90 - 20) * math.pi / 180, 1 * 45 * math.pi / 180
    fill_polar, fill_az = 90 * math.pi / 180, -1 * 45 * math.pi / 180
    return [
        fresnel.light.Light(direction=(math.sin(key_polar) * math.sin(key_az), math.cos(key_polar),
                                       math.sin(key_polar) * math.cos(key_az)),
                            color=(0.75, 0.75, 0.75),
                            theta=math.pi / 8),
        fresnel.light.Light(direction=(math.sin(fill_polar) * math.sin(fill_az), math.cos(fill_polar),
                                       math.sin(fill_polar) * math.cos(fill_az)),
                            color=(0.1, 0.1, 0.1),
                            theta=math.pi / 2),
    ]


test_lights.__test__ = False


def load_png(name):
    import PIL.Image
    im = PIL.Image.open(REFERENCE_PNG / f'{name}.png').convert('RGBA')
    return numpy.asarray(im)


def image_msd(a, ref):
    """(foreground count of a, of ref, mean squared 8-bit difference over shared foreground)."""
    sa = a[:, :, 3] > 0
    sr = ref[:, :, 3] > 0
    sel = sa & sr
    d = a.astype(numpy.float32) - ref.astype(numpy.float32)
    return int(sa.sum()), int(sr.sum()), float(numpy.mean(d[sel]**2)) if sel.any() else 0.0


def assert_image_approx_equal(a, name, tolerance=1.0):
    """The reference's golden-image criterion (fresnel/test/conftest.py:104-131).

    The foreground (alpha > 0) must cover more than 3/4 of the reference's,
    and the mean squared difference of 8-bit RGBA values over pixels that are
    foreground in both must be below ``tolerance``.
    """
    ref = load_png(name)
    assert a.shape == ref.shape
    na, nr, msd = image_msd(a, ref)
    assert na > 3 / 4 * nr, (na, nr)
    assert msd < tolerance, msd
