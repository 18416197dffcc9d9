"""fresnel_amd: the fresnel ray tracing API on an MI355X-native HIP backend.

The public surface follows fresnel 0.13.5 (fresnel/__init__.py:23-315):
`Device`, `Scene`, `preview`, `pathtrace` and the `geometry`, `material`,
`light`, `camera`, `tracer`, `color` and `util` modules.  Rendering runs on
the GPU through the `_gpu` module, which binds the C ABI of
libfresnel_hip.so (include/fresnel_hip.h).
"""

import os

import numpy

from . import _common
from . import geometry  # noqa: F401
from . import tracer
from . import camera
from . import color  # noqa: F401
from . import light
from . import version  # noqa: F401

try:
    from . import _gpu
    _gpu_import_error = None
except ImportError as exc:  # the product needs the HIP module; say why it is missing
    _gpu = None
    _gpu_import_error = exc

# A CPU device module.  None in this package: the only renderer is the GPU.
# A module with the reference `_cpu` class set may be registered with
# `_register_cpu_module` (the test suite registers its CPU checker this way).
_cpu = None


class Device(object):
    """Hardware device to render with.

    Args:
        mode (str): 'auto', 'gpu' or 'cpu'.  'auto' selects the GPU when one is
            available.
        n (int): number of GPUs (or CPU threads) to use; None uses all.  A GPU
            device renders on the first GPU.

    Attributes:
        available_modes (list[str]): modes that can be created.
        available_gpus (list[str]): descriptions of the GPUs.
    """

    available_modes = []
    available_gpus = []

    def __init__(self, mode='auto', n=None):
        num_gpus = _num_gpus()
        selected = ''
        if mode == 'auto':
            if num_gpus > 0:
                selected = 'gpu'
            else:
                if _cpu is None:
                    raise RuntimeError("No GPUs available AND CPU implementation is not compiled"
                                       + _gpu_error_suffix())
                selected = 'cpu'
        if mode == 'gpu':
            if _gpu is None:
                raise RuntimeError("GPU implementation is not compiled" + _gpu_error_suffix())
            if num_gpus == 0:
                raise RuntimeError("No GPUs are available")
            selected = 'gpu'
        if mode == 'cpu':
            if _cpu is None:
                raise RuntimeError("CPU implementation is not compiled")
            selected = 'cpu'
        limit = -1 if n is None else int(n)
        if selected == 'gpu':
            self.module = _gpu
            self._device = _gpu.Device(os.path.dirname(os.path.realpath(__file__)), limit)
            self._mode = 'gpu'
        elif selected == 'cpu':
            self.module = _cpu
            self._device = _cpu.Device(limit)
            self._mode = 'cpu'
        else:
            raise ValueError("Invalid mode")

    @property
    def mode(self):
        """str: the active mode."""
        return self._mode

    def __str__(self):
        """Human readable summary."""
        return '<fresnel.Device: ' + self._device.describe() + '>'


def _gpu_error_suffix():
    return f" ({_gpu_import_error})" if _gpu_import_error is not None else ""


def _num_gpus():
    if _gpu is None:
        return 0
    return _gpu.get_num_available_devices()


def _detect_modes():
    modes = []
    if _num_gpus() > 0:
        modes.append('gpu')
    if _cpu is not None:
        modes.append('cpu')
    if modes:
        modes.append('auto')
    Device.available_modes = modes
    gpus = []
    if _gpu is not None:
        lines = _gpu.Device.getAllGPUs().split('\n')
        if len(lines) >= 2:
            gpus = lines[:-1]
    Device.available_gpus = gpus


def _register_cpu_module(module):
    """Register a module with the reference `_cpu` class set as the 'cpu' mode."""
    global _cpu
    _cpu = module
    _detect_modes()


_detect_modes()


class Scene(object):
    """The content of a rendered image.

    Args:
        device (Device): device to render with; None creates a default Device.
        camera (camera.Camera): None gives an orthographic camera at
            (0, 0, 100) looking at the origin with height 100.
        lights (list[light.Light]): None gives `light.rembrandt()`.

    Attributes:
        geometry (list): the geometries added to this scene.
    """

    def __init__(self, device=None, camera=None, lights=None):
        if device is None:
            device = Device()
        self._device = device
        self._scene = self.device.module.Scene(self.device._device)
        self.geometry = []
        if camera is None:
            self.camera = globals()['camera'].Orthographic(position=(0, 0, 100),
                                                           look_at=(0, 0, 0),
                                                           up=(0, 1, 0),
                                                           height=100)
        else:
            self.camera = camera
        self.lights = light.rembrandt() if lights is None else lights
        self._tracer = None

    def get_extents(self):
        """(2, 3) float32 array: lower and upper corners over all geometries."""
        if len(self.geometry) == 0:
            return numpy.array([[0, 0, 0], [0, 0, 0]], dtype=numpy.float32)
        ext = self.geometry[0].get_extents()
        for geom in self.geometry[1:]:
            e = geom.get_extents()
            ext[0, :] = numpy.min([ext[0, :], e[0, :]], axis=0)
            ext[1, :] = numpy.max([ext[1, :], e[1, :]], axis=0)
        return ext

    @property
    def device(self):
        """Device: the device of this scene."""
        return self._device

    def set_bvh_refit(self, enable=True, limit=1.5):
        """Refit the BVH instead of rebuilding it when primitives only move.

        Not in the reference, which rebuilds on every change.  Images are
        bit-identical either way (the closest hit does not depend on the
        hierarchy); a refit is faster to produce and, for small moves between
        frames of a trajectory, about as fast to traverse.  When a refit raises
        the hierarchy's surface-area cost above ``limit`` times its value after
        the last build, the scene is rebuilt instead.
        """
        if not hasattr(self._scene, 'setRefit'):
            raise RuntimeError("BVH refit needs the GPU backend")
        self._scene.setRefit(bool(enable))
        self._scene.setRefitLimit(float(limit))

    @property
    def camera(self):
        """camera.Camera: the view (a view onto the scene's own camera record)."""
        return camera._from_cpp(self._scene.getCamera())

    @camera.setter
    def camera(self, value):
        if isinstance(value, camera.Camera):
            self._scene.setCamera(value._camera)
        else:
            raise TypeError(f"camera {value} is not a fresnel.camera.Camera")

    @property
    def background_color(self):
        """(3,) float32: linear background color."""
        c = self._scene.getBackgroundColor()
        return numpy.array([c.r, c.g, c.b], dtype=numpy.float32)

    @background_color.setter
    def background_color(self, value):
        self._scene.setBackgroundColor(_common.RGBf(*value))

    @property
    def background_alpha(self):
        """float: background opacity in [0, 1]."""
        return self._scene.getBackgroundAlpha()

    @background_alpha.setter
    def background_alpha(self, value):
        self._scene.setBackgroundAlpha(value)

    @property
    def lights(self):
        """list[light.Light]: up to 4 lights (a view onto the scene's light record)."""
        return light._LightListProxy(self._scene.getLights())

    @lights.setter
    def lights(self, values):
        tmp = light._LightListProxy()
        for v in values:
            tmp.append(v)
        self._scene.setLights(tmp._lights)


def preview(scene, w=600, h=370, anti_alias=True):
    """Render ``scene`` with `tracer.Preview` and return the output image."""
    t = tracer.Preview(scene.device, w=w, h=h, anti_alias=anti_alias)
    return t.render(scene)


def pathtrace(scene, w=600, h=370, samples=64, light_samples=1):
    """Render ``scene`` with `tracer.Path` (``samples`` passes) and return the output image."""
    t = tracer.Path(scene.device, w=w, h=h)
    t.sample(scene, samples=samples, light_samples=light_samples)
    return t.output
