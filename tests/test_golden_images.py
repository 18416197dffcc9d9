"""Every golden image of the reference's test suite, on each device.

The CPU checker must reproduce them (this pins the checker); the GPU backend
must too.  Criterion and tolerances are the reference's own
(fresnel/test/conftest.py:104-131: MSD < 1 for direct comparisons, < 2 in the
perspective-camera helpers, < 16 for the path-tracer image).
"""

import pytest

import fresnel_amd as fresnel
from conftest import assert_image_approx_equal
from scenes import CASES, hex_sphere, render_case


@pytest.mark.parametrize('case', CASES, ids=[c.name for c in CASES])
def test_golden(case, device_):
    img = render_case(case, device_)
    assert_image_approx_equal(img, case.name, case.tol)


def test_tracer_path_golden(device_):
    """fresnel/test/test_tracer_path.py:16-32: seed 11, 64 samples x 40 light samples."""
    scene = hex_sphere(device_)
    tracer = fresnel.tracer.Path(device=device_, w=100, h=100)
    tracer.seed = 11
    assert tracer.output[:].shape == (100, 100, 4)
    out = tracer.sample(scene, samples=64, light_samples=40)
    assert_image_approx_equal(out[:], 'test_tracer_path.test_render', 16)
