"""Progressive interactive viewing (reference: fresnel/interact.py).

`ProgressiveView` is the rendering state machine of the reference's Qt
`SceneView` (interact.py:102-381) without the widget: a full-resolution Path
tracer that refines the image while the camera is still, and a quarter-size
Path tracer that renders one sample per repaint while the camera moves.  Where
the reference adds one sample per repaint, a repaint here adds as many samples
as fit a frame-time budget, measured on the previous repaint: a single sample
of a large scene takes a fraction of a millisecond on the GPU, so one sample
per repaint would leave the display loop bound by launch latency.  The
accumulated image is bit-identical to ``Path.sample`` with the same total
number of samples.

`SceneView` wraps it in a Qt widget when PySide6 or PySide2 is importable
(neither is required by the rest of the package).
"""

import copy
import math
import time

import numpy

from . import tracer

try:  # pragma: no cover - Qt is optional and absent on compute nodes
    from PySide6 import QtCore, QtGui, QtWidgets
except ImportError:  # pragma: no cover
    try:
        from PySide2 import QtCore, QtGui, QtWidgets
    except ImportError:
        QtCore = QtGui = QtWidgets = None


def _quat_axis_angle(axis, angle):
    """Unit quaternion (s, x, y, z) of a rotation by ``angle`` about ``axis``."""
    axis = numpy.asarray(axis, dtype=numpy.float64)
    axis = axis / numpy.linalg.norm(axis)
    return numpy.concatenate([[math.cos(angle / 2)], math.sin(angle / 2) * axis])


def _quat_mul(a, b):
    """Hamilton product a b of quaternions (s, x, y, z)."""
    sa, va = a[0], a[1:]
    sb, vb = b[0], b[1:]
    return numpy.concatenate([[sa * sb - numpy.dot(va, vb)], sa * vb + sb * va + numpy.cross(va, vb)])


def _quat_rotate(q, v):
    """Rotate vector v by unit quaternion q (q v q*)."""
    s, u = q[0], q[1:]
    v = numpy.asarray(v, dtype=numpy.float64)
    return v + 2 * s * numpy.cross(u, v) + 2 * numpy.cross(u, numpy.cross(u, v))


class _CameraController3D:
    """Orbit, pan and zoom a camera from the pose recorded by `start`
    (reference: fresnel/interact.py:40-99, same factors and axis order).

    Args:
        camera (Camera): the camera to move (a view onto the scene's camera).
    """

    def __init__(self, camera):
        self.camera = camera
        self._start_camera = copy.copy(camera)

    def start(self):
        """Record the pose that the next moves start from."""
        self._start_camera = copy.copy(self.camera)

    def orbit(self, yaw=0, pitch=0, roll=0, factor=-0.0025, slight=False):
        """Rotate the camera about look_at: yaw about the up axis, pitch about
        the right axis, roll about the view axis."""
        if slight:
            factor = factor * 0.1
        basis = numpy.array(self._start_camera.basis, dtype=numpy.float64)
        q1 = _quat_axis_angle(basis[1], factor * yaw)
        q2 = _quat_axis_angle(basis[0], factor * pitch)
        q3 = _quat_axis_angle(basis[2], factor * roll)
        q = _quat_mul(q2, _quat_mul(q1, q3))
        look_at = numpy.asarray(self._start_camera.look_at, dtype=numpy.float64)
        v = numpy.asarray(self._start_camera.position, dtype=numpy.float64) - look_at
        self.camera.position = look_at + _quat_rotate(q, v)
        self.camera.up = _quat_rotate(q, basis[1])

    def pan(self, x, y, slight=False):
        """Move position and look_at parallel to the image plane by (x, y)
        image heights."""
        factor = self._start_camera.height
        if slight:
            factor = factor * 0.1
        basis = numpy.array(self._start_camera.basis, dtype=numpy.float64)
        delta = factor * (x * basis[0] + y * basis[1])
        self.camera.position = numpy.asarray(self._start_camera.position, dtype=numpy.float64) + delta
        self.camera.look_at = numpy.asarray(self._start_camera.look_at, dtype=numpy.float64) + delta

    def zoom(self, s, slight=False):
        """Scale the image-plane height by (1 - 0.0015 s)."""
        factor = 0.0015
        if slight:
            factor = factor * 0.1
        self.camera.height = self._start_camera.height * (1 - s * factor)


class ProgressiveView:
    """Progressive rendering of a scene for an interactive display.

    Args:
        scene (Scene): the scene to display; its camera is the view.
        max_samples (int): stop refining after this many samples per pixel.
        frame_time (float): seconds of rendering per repaint at full resolution.
        low_res_factor (int): the moving-camera image is this many times
            smaller in each dimension.

    Call `resize` with the display size, `repaint` from the display loop
    (it returns the image to show), `begin_interaction` / `end_interaction`
    around camera moves, and `restart` after any change to the scene.
    """

    def __init__(self, scene, max_samples=2000, frame_time=1 / 30, low_res_factor=4):
        self._scene = scene
        self.max_samples = int(max_samples)
        self.frame_time = float(frame_time)
        self.low_res_factor = int(low_res_factor)
        self._tracer = tracer.Path(device=scene.device, w=10, h=10)
        self._low_res_tracer = tracer.Path(device=scene.device, w=10, h=10)
        self.camera_controller = _CameraController3D(scene.camera)
        self._high_res = True
        self._rendering = False
        self._samples = 0
        self._batch = 1
        self._pending = None
        self.restart()

    @property
    def scene(self):
        """Scene: the scene shown."""
        return self._scene

    @scene.setter
    def scene(self, scene):
        self._scene = scene
        self.camera_controller = _CameraController3D(scene.camera)
        self.restart()

    @property
    def samples(self):
        """int: samples per pixel accumulated in the full-resolution image."""
        return self._samples

    @property
    def rendering(self):
        """bool: whether a repaint still adds samples."""
        return self._rendering

    @property
    def high_res(self):
        """bool: whether repaints show the full-resolution image."""
        return self._high_res

    @property
    def tracer(self):
        """Path: the full-resolution tracer."""
        return self._tracer

    def resize(self, w, h):
        """Size the images for a w x h display and start over."""
        w, h = max(int(w), 1), max(int(h), 1)
        self._tracer.resize(w=w, h=h)
        f = self.low_res_factor
        self._low_res_tracer.resize(w=max(w // f, 1), h=max(h // f, 1))
        self._pending = None
        self.restart()

    def restart(self):
        """Discard the accumulated samples (the camera or the scene changed)."""
        self._rendering = True
        self._samples = 0
        self._batch = 1
        self._tracer.reset()
        self._low_res_tracer.reset()

    def begin_interaction(self):
        """The camera starts moving: show low-resolution frames."""
        self.camera_controller.start()
        self._high_res = False
        self.restart()

    def end_interaction(self):
        """The camera is still again: refine at full resolution."""
        self._high_res = True

    def repaint(self):
        """Render for one repaint and return the `ImageArray` to display."""
        self._render()
        return self._tracer.output if self._high_res else self._low_res_tracer.output

    def repaint_async(self):
        """Render for one repaint with the image read back asynchronously.

        Renders the next samples, starts reading the new image back, and
        returns (as an (h, w, 4) uint8 array) the image the previous call
        started reading: that copy ran while the display drew and while this
        call traced.  The first call waits for its own image.  GPU devices.
        """
        prev = None
        if self._pending is not None:
            prev = numpy.array(self._pending.wait())
        self._render()
        t = self._tracer if self._high_res else self._low_res_tracer
        self._pending = t.output_async()
        if prev is None:
            prev = numpy.array(self._pending.wait())
        return prev

    def _render(self):
        if self._rendering:
            if self._high_res:
                k = min(self._batch, self.max_samples - self._samples)
                t0 = time.perf_counter()
                self._tracer.sample(self._scene, samples=k, reset=False)
                dt = time.perf_counter() - t0
                self._samples += k
                if self._samples >= self.max_samples:
                    self._rendering = False
                # samples for the next repaint from this one's rate, grown at
                # most 4x per repaint
                per_sample = dt / k
                want = int(self.frame_time / per_sample) if per_sample > 0 else 4 * k
                self._batch = max(1, min(4 * k, want))
            else:
                self._low_res_tracer.render(self._scene)


if QtWidgets is not None:  # pragma: no cover - needs a Qt binding and a display

    class SceneView(QtWidgets.QWidget):
        """Qt widget that shows a scene and refines it progressively
        (reference: fresnel/interact.py:102-381).

        Left drag: pitch and yaw; right drag: roll and zoom; middle drag: pan;
        wheel: zoom; hold Ctrl for small moves.
        """

        rendering = QtCore.Signal(object)
        TIMEOUT = 100

        def __init__(self, scene, max_samples=2000):
            super().__init__()
            self.setWindowTitle("fresnel: scene viewer")
            self.setMinimumSize(10, 10)
            self._view = ProgressiveView(scene, max_samples=max_samples)
            self._repaint_timer = QtCore.QTimer(self)
            self._repaint_timer.timeout.connect(self.update)
            self._resize_timer = QtCore.QTimer(self)
            self._resize_timer.setSingleShot(True)
            self._resize_timer.timeout.connect(self._resize_done)
            self._low_res_timer = QtCore.QTimer(self)
            self._low_res_timer.setSingleShot(True)
            self._low_res_timer.timeout.connect(self._view.end_interaction)
            self._initial_resize = True
            self._mode = None
            self._press = None

        @property
        def scene(self):
            return self._view.scene

        @scene.setter
        def scene(self, scene):
            self._view.scene = scene
            self._start()

        def _start(self):
            self.rendering.emit(self._view.scene.camera)
            self._repaint_timer.start()

        def _resize_done(self):
            self._view.resize(self.width(), self.height())
            self._start()

        def minimumSizeHint(self):  # noqa: N802 - Qt naming
            return QtCore.QSize(1610, 1000)

        def paintEvent(self, event):  # noqa: N802 - Qt naming
            image = self._view.repaint()
            if not self._view.rendering:
                self._repaint_timer.stop()
            a = numpy.ascontiguousarray(image[:])
            img = QtGui.QImage(a.data, a.shape[1], a.shape[0], QtGui.QImage.Format_RGBA8888)
            qp = QtGui.QPainter(self)
            qp.drawImage(QtCore.QRectF(0, 0, self.width(), self.height()), img,
                         QtCore.QRectF(0.0, 0.0, a.shape[1], a.shape[0]))
            qp.end()

        def resizeEvent(self, event):  # noqa: N802 - Qt naming
            if self._initial_resize:
                self._resize_done()
                self._initial_resize = False
            else:
                self._resize_timer.start(self.TIMEOUT)

        def mousePressEvent(self, event):  # noqa: N802 - Qt naming
            self._press = event.pos()
            b = event.button()
            self._mode = {QtCore.Qt.LeftButton: 'pitch/yaw', QtCore.Qt.RightButton: 'roll/zoom',
                          QtCore.Qt.MiddleButton: 'pan'}.get(b)
            self._view.begin_interaction()
            self._start()
            event.accept()

        def mouseMoveEvent(self, event):  # noqa: N802 - Qt naming
            if self._press is None:
                return
            d = event.pos() - self._press
            slight = bool(event.modifiers() & QtCore.Qt.ControlModifier)
            c = self._view.camera_controller
            if self._mode == 'pitch/yaw':
                c.orbit(yaw=d.x(), pitch=d.y(), slight=slight)
            elif self._mode == 'roll/zoom':
                c.orbit(roll=d.x(), slight=slight)
                c.zoom(-d.y(), slight=slight)
            elif self._mode == 'pan':
                h = self.height()
                c.pan(x=-d.x() / h, y=d.y() / h, slight=slight)
            self._view.restart()
            self._start()
            self.update()
            event.accept()

        def mouseReleaseEvent(self, event):  # noqa: N802 - Qt naming
            self._mode = None
            self._press = None
            self._view.end_interaction()
            event.accept()

        def wheelEvent(self, event):  # noqa: N802 - Qt naming
            self._view.begin_interaction()
            self._view.camera_controller.zoom(event.angleDelta().y(),
                                              slight=bool(event.modifiers() & QtCore.Qt.ControlModifier))
            self._low_res_timer.start(self.TIMEOUT)
            self._start()

else:

    class SceneView:
        """Qt widget for interactive viewing; needs PySide6 or PySide2."""

        def __init__(self, scene, max_samples=2000):
            raise ImportError("fresnel_amd.interact.SceneView needs PySide6 or PySide2; "
                              "use ProgressiveView for the rendering loop without Qt")
