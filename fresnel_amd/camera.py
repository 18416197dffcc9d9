"""Cameras: orthographic and thin-lens perspective (reference: fresnel/camera.py).

The Python objects wrap a `_common.UserCamera` record; the device derives its
ray generation constants from it (fresnel/common/Camera.h:141-209).
"""

import math

import numpy

from . import _common


def _vec3(value, name):
    if len(value) != 3:
        raise ValueError(f"{name} must have length 3")
    return _common.vec3f(*value)


def _t(a):
    """Tuple of plain Python floats (for a repr that eval() can read back)."""
    return tuple(float(x) for x in a)


def _np3(v):
    return numpy.array([v.x, v.y, v.z], dtype=numpy.float32)


class Camera(object):
    """Base class: position, look_at, up and image-plane height.

    Camera space is centered on the camera position: +x right, +y up, +z out
    of the screen.  The visible width is ``height * w / h`` of the tracer.
    """

    def __init__(self, _camera):
        self._camera = _camera

    def __copy__(self):
        """Copy of this camera with its own UserCamera record."""
        cam = _common.UserCamera()
        for name in ('model', 'position', 'look_at', 'up', 'h', 'f', 'f_stop', 'focus_distance'):
            setattr(cam, name, getattr(self._camera, name))
        return _from_cpp(cam)

    @property
    def position(self):
        """(3,) float32: camera position."""
        return _np3(self._camera.position)

    @position.setter
    def position(self, value):
        self._camera.position = _vec3(value, "position")

    @property
    def look_at(self):
        """(3,) float32: the point at the center of the image."""
        return _np3(self._camera.look_at)

    @look_at.setter
    def look_at(self, value):
        self._camera.look_at = _vec3(value, "look_at")

    @property
    def up(self):
        """(3,) float32: a vector toward +y in camera space."""
        return _np3(self._camera.up)

    @up.setter
    def up(self, value):
        self._camera.up = _vec3(value, "up")

    @property
    def height(self):
        """float: height of the image plane."""
        return self._camera.h

    @height.setter
    def height(self, value):
        self._camera.h = float(value)

    @property
    def basis(self):
        """(3, 3) float32: camera-space +x, +y, +z axes expressed in scene space."""
        b = _common.CameraBasis(self._camera)
        return numpy.array([(b.u.x, b.u.y, b.u.z), (b.v.x, b.v.y, b.v.z), (b.w.x, b.w.y, b.w.z)],
                           dtype=numpy.float32)


class Orthographic(Camera):
    """Orthographic camera: parallel rays from the image plane.

    Args:
        position ((3,) float32): center of the image plane.
        look_at ((3,) float32): point at the center of the image.
        up ((3,) float32): vector toward +y in camera space.
        height (float): height of the image plane.
    """

    def __init__(self, position, look_at, up, height):
        cam = _common.UserCamera()
        cam.model = _common.CameraModel.orthographic
        super().__init__(cam)
        self.position = position
        self.look_at = look_at
        self.up = up
        self.height = height

    def __repr__(self):
        """str: constructor expression for this camera."""
        return (f"fresnel.camera.Orthographic(position={_t(self.position)}, "
                f"look_at={_t(self.look_at)}, up={_t(self.up)}, height={self.height})")

    @classmethod
    def fit(cls, scene, view='auto', margin=0.05):
        """Orthographic camera that frames the whole scene.

        Args:
            scene (Scene): scene to frame.
            view (str): 'auto', 'isometric' or 'front'.  'auto' picks 'front'
                for flat scenes (z extent < 0.51 of the larger of x, y).
            margin (float): fraction of extra space above and below.

        Same construction as the reference (fresnel/camera.py:196-279): the
        eight corners of the scene box are projected on the view's up and
        view axes.
        """
        views = {
            'front': (numpy.array([0, 0, 1]), numpy.array([0, 1, 0])),
            'isometric': (numpy.array([1, 1, 1]) / math.sqrt(3), numpy.array([-1, 2, -1]) / math.sqrt(6)),
        }
        if len(scene.geometry) == 0:
            raise ValueError('The camera cannot be fit because the scene has no geometries. '
                             'Add geometries to the scene before calling fit.')
        ext = scene.get_extents()
        if view == 'auto':
            w = ext[1] - ext[0]
            view = 'front' if w[2] < 0.51 * max(w[0], w[1]) else 'isometric'
        v, up = views[view]
        corners = numpy.array([[ext[a, 0], ext[b, 1], ext[c, 2]]
                               for a in (0, 1) for b in (0, 1) for c in (0, 1)])
        center = (ext[0, :] + ext[1, :]) / 2
        corners = corners - center
        height = (1 + margin) * numpy.max(numpy.abs(numpy.dot(corners, up))) * 2
        distance = numpy.max(numpy.dot(corners, v)) * 1.10
        return cls(position=center + distance * v, look_at=center, up=up, height=height)


class Perspective(Camera):
    """Thin-lens perspective camera.

    Args:
        position ((3,) float32): center of projection.
        look_at ((3,) float32): point at the center of the image.
        up ((3,) float32): vector toward +y in camera space.
        focal_length (float): distance from the position to the image plane.
        focus_distance (float): distance to the plane in sharp focus.
        f_stop (float): aperture diameter is ``focal_length / f_stop``.
        height (float): height of the image sensor.
    """

    def __init__(self,
                 position,
                 look_at,
                 up,
                 focal_length=.5,
                 focus_distance=10,
                 f_stop=math.inf,
                 height=0.24):
        cam = _common.UserCamera()
        cam.model = _common.CameraModel.perspective
        super().__init__(cam)
        self.position = position
        self.look_at = look_at
        self.up = up
        self.focus_distance = focus_distance
        self.focal_length = focal_length
        self.f_stop = f_stop
        self.height = height

    def __repr__(self):
        """str: constructor expression for this camera."""
        return (f"fresnel.camera.Perspective(position={_t(self.position)}, "
                f"look_at={_t(self.look_at)}, up={_t(self.up)}, "
                f"focus_distance={self.focus_distance}, focal_length={self.focal_length}, "
                f"f_stop={self.f_stop}, height={self.height})")

    @property
    def focal_length(self):
        """float: focal length of the lens."""
        return self._camera.f

    @focal_length.setter
    def focal_length(self, value):
        self._camera.f = value

    @property
    def f_stop(self):
        """float: aperture as an f-number; inf gives a pinhole."""
        return self._camera.f_stop

    @f_stop.setter
    def f_stop(self, value):
        self._camera.f_stop = value

    @property
    def focus_distance(self):
        """float: distance from the position to the focal plane."""
        return self._camera.focus_distance

    @focus_distance.setter
    def focus_distance(self, value):
        self._camera.focus_distance = value

    @property
    def depth_of_field(self):
        """float: extent about the focal plane in acceptable focus.

        Uses a circle of confusion of ``height / 800`` (0.03 mm on a 24 mm
        sensor) and the hyperfocal distance ``H = f^2 / (N c)``.  Setting it
        solves for ``f_stop``.
        """
        f = self.focal_length
        N = self.f_stop
        s = self.focus_distance
        c = self.height / 800
        H = f**2 / (N * c)
        if H - (s - f) <= 0 or H + (s - f) <= 0:
            return math.inf
        return s * H * (1 / (H - (s - f)) - 1 / (H + (s - f)))

    @depth_of_field.setter
    def depth_of_field(self, value):
        f = self.focal_length
        d = value
        s = self.focus_distance
        c = self.height / 800
        a = math.sqrt(c**2 * f**4 * (d**2 + s**2) * (f - s)**2)
        b = c * f**2 * s * (f - s)
        self.f_stop = (a + b) / (c**2 * d * (f - s)**2)

    @property
    def focus_on(self):
        """(3,) float32: the point on the focal plane along the view axis.

        Setting it moves the focal plane through the given point.
        """
        d = -self.basis[2, :]
        return self.position + d * self.focus_distance

    @focus_on.setter
    def focus_on(self, value):
        if len(value) != 3:
            raise ValueError("focus_on must have length 3")
        d = -self.basis[2, :]
        self.focus_distance = numpy.dot(d, value - self.position)

    @property
    def vertical_field_of_view(self):
        """float: vertical field of view in radians; setting it sets focal_length."""
        return 2 * math.atan(self.height / (2 * self.focal_length))

    @vertical_field_of_view.setter
    def vertical_field_of_view(self, value):
        self.focal_length = self.height / (2 * math.tan(value / 2))


def _from_cpp(cam):
    """Python camera object sharing the given UserCamera record."""
    if cam.model == _common.CameraModel.orthographic:
        result = Orthographic(position=(0, 0, 0), look_at=(0, 0, 1), up=(0, 1, 0), height=1)
    elif cam.model == _common.CameraModel.perspective:
        result = Perspective(position=(0, 0, 0), look_at=(0, 0, 1), up=(0, 1, 0), focus_distance=1)
    else:
        raise RuntimeError("Invalid camera model")
    result._camera = cam
    return result
