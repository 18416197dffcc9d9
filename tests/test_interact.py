"""Progressive interactive viewing (fresnel_amd.interact).

The camera controller is checked against rotation matrices built
independently (Rodrigues' formula), and against the reference's pan and zoom
rules (fresnel/interact.py:58-99).  The progressive loop must end with the
image that ``Path.sample`` gives for the same number of samples and seed, on
the CPU checker device here and on the GPU in the gpu-marked test.
"""

import math

import numpy
import pytest

import fresnel_amd as fresnel
from fresnel_amd import interact


def _rodrigues(axis, angle):
    k = numpy.asarray(axis, dtype=numpy.float64)
    k = k / numpy.linalg.norm(k)
    K = numpy.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return numpy.eye(3) + math.sin(angle) * K + (1 - math.cos(angle)) * K @ K


def _camera():
    return fresnel.camera.Orthographic(position=(1, 2, 10), look_at=(0.5, -1, 0), up=(0.2, 1, 0), height=6)


def test_orbit_matches_rotation_matrices():
    cam = _camera()
    start = numpy.array(cam.basis, dtype=numpy.float64)
    p0, l0 = numpy.array(cam.position, numpy.float64), numpy.array(cam.look_at, numpy.float64)
    c = interact._CameraController3D(cam)
    c.start()
    yaw, pitch, roll = 70, -35, 120
    c.orbit(yaw=yaw, pitch=pitch, roll=roll)
    f = -0.0025
    # q = q_pitch * q_yaw * q_roll: the roll acts first, then yaw, then pitch
    R = _rodrigues(start[0], f * pitch) @ _rodrigues(start[1], f * yaw) @ _rodrigues(start[2], f * roll)
    numpy.testing.assert_allclose(cam.position, l0 + R @ (p0 - l0), rtol=0, atol=2e-5)
    numpy.testing.assert_allclose(cam.up, R @ start[1], rtol=0, atol=2e-6)
    numpy.testing.assert_allclose(cam.look_at, l0, rtol=0, atol=0)
    # the distance to look_at is kept; moves start from the recorded pose
    c.orbit(yaw=yaw, slight=True)
    R = _rodrigues(start[1], 0.1 * f * yaw)
    numpy.testing.assert_allclose(cam.position, l0 + R @ (p0 - l0), rtol=0, atol=2e-5)


def test_pan_and_zoom():
    cam = _camera()
    b = numpy.array(cam.basis, dtype=numpy.float64)
    p0, l0, h0 = numpy.array(cam.position), numpy.array(cam.look_at), cam.height
    c = interact._CameraController3D(cam)
    c.start()
    c.pan(0.25, -0.5)
    d = h0 * (0.25 * b[0] - 0.5 * b[1])
    numpy.testing.assert_allclose(cam.position, p0 + d, atol=1e-5)
    numpy.testing.assert_allclose(cam.look_at, l0 + d, atol=1e-5)
    c.pan(0.25, -0.5, slight=True)
    numpy.testing.assert_allclose(cam.position, p0 + 0.1 * d, atol=1e-5)
    c.zoom(200)
    assert cam.height == pytest.approx(h0 * (1 - 200 * 0.0015), rel=1e-6)
    c.zoom(200, slight=True)
    assert cam.height == pytest.approx(h0 * (1 - 200 * 0.00015), rel=1e-6)


def _scene(device):
    s = fresnel.Scene(device, lights=fresnel.light.cloudy())
    fresnel.geometry.Sphere(s, position=[[0, 0, 0], [1.2, 0.3, -0.5], [-1, -0.4, 0.4]], radius=[0.6, 0.5, 0.4],
                            material=fresnel.material.Material(color=fresnel.color.linear([0.3, 0.6, 0.9])))
    s.camera = fresnel.camera.Orthographic.fit(s)
    return s


def _progressive_equals_sample(device, w, h, max_samples):
    s = _scene(device)
    view = interact.ProgressiveView(s, max_samples=max_samples, frame_time=10.0)
    view.resize(w, h)
    n = 0
    while view.rendering:
        img = view.repaint()
        n += 1
        assert n <= max_samples
    assert view.samples == max_samples
    got = numpy.array(img[:])
    t = fresnel.tracer.Path(device, w=w, h=h)
    t.seed = view.tracer.seed - 1
    want = numpy.array(t.sample(s, samples=max_samples)[:])
    assert numpy.array_equal(got, want)
    # moving the camera shows quarter-size frames and restarts the refinement
    view.begin_interaction()
    view.camera_controller.orbit(yaw=40)
    view.restart()
    low = view.repaint()
    assert low.shape == (h // 4, w // 4, 4) and view.samples == 0
    view.end_interaction()
    assert view.repaint().shape == (h, w, 4) and view.samples >= 1
    return n


def test_progressive_view_on_the_cpu_checker(cpu_device):
    _progressive_equals_sample(cpu_device, 24, 16, 5)


@pytest.mark.gpu
def test_progressive_view_on_the_gpu(gpu_device):
    # with a generous frame budget the batch grows 1, 4, 16, ... per repaint
    n = _progressive_equals_sample(gpu_device, 160, 96, 300)
    assert n < 300


@pytest.mark.gpu
def test_progressive_view_on_the_gpu_equals_checker(gpu_device, cpu_device):
    """The GPU's refined image equals, bit for bit in linear HDR, the checker's Path.sample with the same
    seed and sample count."""
    s = _scene(gpu_device)
    view = interact.ProgressiveView(s, max_samples=40, frame_time=10.0)
    view.resize(96, 64)
    while view.rendering:
        view.repaint()
    got = numpy.array(view.tracer.linear_output[:])
    cs = _scene(cpu_device)
    t = fresnel.tracer.Path(cpu_device, w=96, h=64)
    t.seed = view.tracer.seed - 1
    t.sample(cs, samples=40)
    assert numpy.array_equal(got.view(numpy.uint32), numpy.array(t.linear_output[:]).view(numpy.uint32))


@pytest.mark.gpu
def test_progressive_view_async_readback(gpu_device):
    """repaint_async returns the image the previous call read back: the sequence equals repaint()'s, one
    call behind."""
    def run(method):
        s = _scene(gpu_device)
        view = interact.ProgressiveView(s, max_samples=30, frame_time=1e-9)  # one sample per repaint
        view.resize(64, 48)
        out = []
        for _ in range(6):
            out.append(numpy.array(getattr(view, method)()[:]))
        return out

    sync = run('repaint')
    asy = run('repaint_async')
    assert numpy.array_equal(asy[0], sync[0])
    for k in range(1, 6):
        assert numpy.array_equal(asy[k], sync[k - 1]), k


def test_scene_view_needs_qt():
    if interact.QtWidgets is not None:
        pytest.skip("a Qt binding is installed")
    with pytest.raises(ImportError):
        interact.SceneView(None)
