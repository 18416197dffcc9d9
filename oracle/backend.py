"""The reference `_cpu` class set implemented over the oracle library.

TEST INFRASTRUCTURE ONLY.  The test suite registers this module as the 'cpu'
device mode of fresnel_amd (``fresnel_amd._register_cpu_module``) so that
tests written against fresnel's Python API run unchanged on the CPU checker
and on the GPU.  Class and method names follow fresnel/cpu/module-cpu.cc and
the export_* functions it calls; buffer shapes follow fresnel/cpu/Array.h.
"""

import ctypes
import math
import os

import numpy

from fresnel_amd import _common

from . import lib, ptr

_U32 = ctypes.c_uint32
_U8 = ctypes.c_uint8
GEOM_SPHERE, GEOM_CYLINDER, GEOM_POLYGON, GEOM_POLYHEDRON, GEOM_MESH = range(5)


def _mat9(m):
    return numpy.array([m.solid, m.color.r, m.color.g, m.color.b, m.primitive_color_mix,
                        m.roughness, m.specular, m.metal, m.spec_trans], dtype=numpy.float32)


def _copy_material(m):
    c = _common.Material()
    for k in ('solid', 'primitive_color_mix', 'roughness', 'specular', 'spec_trans', 'metal'):
        setattr(c, k, getattr(m, k))
    c.color = _common.RGBf(m.color.r, m.color.g, m.color.b)
    return c


def _new_material(r, g, b, solid=0.0):
    # Material(RGB, solid) of fresnel/common/Material.h:61-65
    m = _common.Material()
    m.solid = solid
    m.color = _common.RGBf(r, g, b)
    m.primitive_color_mix = 0.0
    m.roughness = 0.1
    m.specular = 0.5
    m.metal = 0.0
    m.spec_trans = 0.0
    return m


class _Array(object):
    """Host buffer with map()/unmap() and numpy access."""

    def __init__(self, shape, dtype=numpy.float32):
        self.data = numpy.zeros(shape, dtype=dtype)

    def map(self):
        pass

    def unmap(self):
        pass

    def __array__(self, dtype=None, copy=None):
        return self.data


class Device(object):

    def __init__(self, limit=-1):
        self.limit = limit
        self.nthreads = limit if limit > 0 else (os.cpu_count() or 1)

    def describe(self):
        if self.limit == -1:
            return "All available CPU threads (oracle)"
        return f"{self.limit} CPU threads (oracle)"


class Scene(object):

    def __init__(self, device):
        self.device = device
        self._camera = _common.UserCamera()
        self._bg = (0.0, 0.0, 0.0)
        self._bg_alpha = 0.0
        self._lights = _common.Lights()
        # default lights of fresnel/cpu/Scene.cc:21-25
        self._lights.N = 2
        self._lights.setDirection(0, _common.vec3f(-1, 0.3, 1))
        self._lights.setColor(0, _common.RGBf(1, 1, 1))
        self._lights.setDirection(1, _common.vec3f(1, 0, 1))
        self._lights.setColor(1, _common.RGBf(0.1, 0.1, 0.1))
        self._geoms = {}  # geomID -> geometry
        self.version = 0
        self._handle = None
        self._handle_version = -1
        self.brute = False

    def __del__(self):
        if self._handle is not None:
            lib().oracle_scene_free(self._handle)

    # ids are the smallest free slot, reused after removal
    def _attach(self, geom):
        gid = 0
        while gid in self._geoms:
            gid += 1
        self._geoms[gid] = geom
        self.version += 1
        return gid

    def _detach(self, gid):
        self._geoms.pop(gid, None)
        self.version += 1

    def setCamera(self, cam):
        c = _common.UserCamera()
        for k in ('model', 'position', 'look_at', 'up', 'h', 'f', 'f_stop', 'focus_distance'):
            setattr(c, k, getattr(cam, k))
        self._camera = c

    def getCamera(self):
        return self._camera

    def setBackgroundColor(self, c):
        self._bg = (c.r, c.g, c.b)

    def getBackgroundColor(self):
        return _common.RGBf(*self._bg)

    def setBackgroundAlpha(self, a):
        self._bg_alpha = float(numpy.float32(a))

    def getBackgroundAlpha(self):
        return self._bg_alpha

    def setLights(self, lights):
        c = _common.Lights()
        c.N = lights.N
        for i in range(4):
            c.setDirection(i, lights.getDirection(i))
            c.setColor(i, lights.getColor(i))
            c.setTheta(i, lights.getTheta(i))
        self._lights = c

    def getLights(self):
        return self._lights

    def handle(self):
        """Oracle scene for the current geometry, rebuilt when geometry changed."""
        L = lib()
        if self._handle is None or self._handle_version != self.version:
            if self._handle is not None:
                L.oracle_scene_free(self._handle)
            h = L.oracle_scene_new()
            for gid in sorted(self._geoms):
                self._geoms[gid]._add_to(h, gid)
            L.oracle_scene_set_brute(h, 1 if self.brute else 0)
            L.oracle_scene_commit(h)
            self._handle = h
            self._handle_version = self.version
        h = self._handle
        c = self._camera
        v = lambda p: numpy.array([p.x, p.y, p.z], dtype=numpy.float32)
        L.oracle_scene_set_camera(h, ptr(v(c.position)), ptr(v(c.look_at)), ptr(v(c.up)), c.h, c.f,
                                  c.f_stop, c.focus_distance, int(c.model))
        L.oracle_scene_set_background(h, ptr(numpy.array(self._bg, dtype=numpy.float32)),
                                      self._bg_alpha)
        n = self._lights.N
        dirs = numpy.zeros((4, 3), numpy.float32)
        cols = numpy.zeros((4, 3), numpy.float32)
        th = numpy.zeros(4, numpy.float32)
        for i in range(min(n, 4)):
            d = self._lights.getDirection(i)
            col = self._lights.getColor(i)
            dirs[i] = (d.x, d.y, d.z)
            cols[i] = (col.r, col.g, col.b)
            th[i] = self._lights.getTheta(i)
        L.oracle_scene_set_lights(h, n, ptr(dirs), ptr(cols), ptr(th))
        return h


class Geometry(object):

    def __init__(self, scene):
        self.scene = scene
        self.valid = True
        self.enabled = True
        self.material = _new_material(1, 0, 1)
        self.outline = _new_material(0, 0, 0, 1.0)
        self.outline_width = 0.0
        self.gid = scene._attach(self)

    def _check(self):
        if not self.valid:
            raise RuntimeError("Cannot modify a removed Geometry")

    def getMaterial(self):
        return _copy_material(self.material)

    def setMaterial(self, m):
        self.material = _copy_material(m)
        self.scene.version += 1

    def getOutlineMaterial(self):
        return _copy_material(self.outline)

    def setOutlineMaterial(self, m):
        self.outline = _copy_material(m)
        self.scene.version += 1

    def getOutlineWidth(self):
        return self.outline_width

    def setOutlineWidth(self, w):
        self.outline_width = float(numpy.float32(w))
        self.scene.version += 1

    def enable(self):
        if not self.valid:
            raise RuntimeError("Cannot enable inactive Geometry")
        self.enabled = True
        self.scene.version += 1

    def disable(self):
        if not self.valid:
            raise RuntimeError("Cannot disable inactive Geometry")
        self.enabled = False
        self.scene.version += 1

    def remove(self):
        if self.valid:
            self.scene._detach(self.gid)
            self.valid = False

    def update(self):
        self.scene.version += 1

    def _common_args(self, gid):
        return (gid, 1 if self.enabled else 0, ptr(_mat9(self.material)), ptr(_mat9(self.outline)),
                self.outline_width)


class GeometrySphere(Geometry):

    def __init__(self, scene, N):
        super().__init__(scene)
        self.N = N
        self.position = _Array((N, 3))
        self.radius = _Array((N,))
        self.color = _Array((N, 3))

    def getPositionBuffer(self):
        return self.position

    def getRadiusBuffer(self):
        return self.radius

    def getColorBuffer(self):
        return self.color

    def _add_to(self, h, gid):
        lib().oracle_add_sphere(h, *self._common_args(gid), self.N, ptr(self.position.data),
                                ptr(self.radius.data), ptr(self.color.data))


class GeometryCylinder(Geometry):

    def __init__(self, scene, N):
        super().__init__(scene)
        self.N = N
        self.points = _Array((N, 2, 3))
        self.radius = _Array((N,))
        self.color = _Array((N, 2, 3))

    def getPointsBuffer(self):
        return self.points

    def getRadiusBuffer(self):
        return self.radius

    def getColorBuffer(self):
        return self.color

    def _add_to(self, h, gid):
        lib().oracle_add_cylinder(h, *self._common_args(gid), self.N, ptr(self.points.data),
                                  ptr(self.radius.data), ptr(self.color.data))


class GeometryConvexPolyhedron(Geometry):

    def __init__(self, scene, origins, normals, colors, N, r):
        super().__init__(scene)
        o = numpy.ascontiguousarray(origins, dtype=numpy.float32)
        n = numpy.ascontiguousarray(normals, dtype=numpy.float32)
        c = numpy.ascontiguousarray(colors, dtype=numpy.float32)
        for a, name in ((o, 'plane_origins'), (n, 'plane_normals'), (c, 'plane_colors')):
            if a.ndim != 2:
                raise RuntimeError(f"{name} must be a 2-dimensional array")
            if a.shape[1] != 3:
                raise RuntimeError(f"{name} must be a Nvert by 3 array")
        if n.shape[0] != o.shape[0]:
            raise RuntimeError("Number of vertices must match in origin and normal arrays")
        if c.shape[0] != o.shape[0]:
            raise RuntimeError("Number of vertices must match in origin and color arrays")
        self.planes = (o, n, c)
        self.r = float(numpy.float32(r))
        self.color_by_face = 0.0
        self.N = N
        self.position = _Array((N, 3))
        self.orientation = _Array((N, 4))
        self.color = _Array((N, 3))

    def getPositionBuffer(self):
        return self.position

    def getOrientationBuffer(self):
        return self.orientation

    def getColorBuffer(self):
        return self.color

    def setColorByFace(self, f):
        self.color_by_face = float(numpy.float32(f))
        self.scene.version += 1

    def getColorByFace(self):
        return self.color_by_face

    def _add_to(self, h, gid):
        o, n, c = self.planes
        lib().oracle_add_polyhedron(h, *self._common_args(gid), self.N, ptr(self.position.data),
                                    ptr(self.orientation.data), ptr(self.color.data), o.shape[0],
                                    ptr(o), ptr(n), ptr(c), self.r, self.color_by_face)


class GeometryMesh(Geometry):

    def __init__(self, scene, vertices, N):
        super().__init__(scene)
        v = numpy.ascontiguousarray(vertices, dtype=numpy.float32)
        if v.ndim != 2:
            raise RuntimeError("vertices must be a 2-dimensional array")
        if v.shape[1] != 3:
            raise RuntimeError("vertices must be a Nvert by 3 array")
        if v.shape[0] % 3 != 0:
            raise RuntimeError("the number of triangle vertices must be a multiple of three.")
        self.vertices = v
        self.N = N
        self.position = _Array((N, 3))
        self.orientation = _Array((N, 4))
        self.color = _Array((v.shape[0], 3))

    def getPositionBuffer(self):
        return self.position

    def getOrientationBuffer(self):
        return self.orientation

    def getColorBuffer(self):
        return self.color

    def _add_to(self, h, gid):
        lib().oracle_add_mesh(h, *self._common_args(gid), self.N, ptr(self.position.data),
                              ptr(self.orientation.data), self.vertices.shape[0], ptr(self.vertices),
                              ptr(self.color.data))


class GeometryPolygon(Geometry):

    def __init__(self, scene, vertices, rounding_radius, N):
        super().__init__(scene)
        v = numpy.ascontiguousarray(vertices, dtype=numpy.float32)
        if v.ndim != 2:
            raise RuntimeError("vertices must be a 2-dimensional array")
        if v.shape[1] != 2:
            raise RuntimeError("vertices must be a Nvert by 2 array")
        self.vertices = v
        self.rounding = float(numpy.float32(rounding_radius))
        r = numpy.float32(0)
        for p in v:
            r = max(r, numpy.sqrt(numpy.float32(p[0] * p[0] + p[1] * p[1]), dtype=numpy.float32))
        self.xy_radius = float(numpy.float32(r + numpy.float32(self.rounding)))
        self.N = N
        self.position = _Array((N, 2))
        self.angle = _Array((N,))
        self.color = _Array((N, 3))

    def getPositionBuffer(self):
        return self.position

    def getAngleBuffer(self):
        return self.angle

    def getColorBuffer(self):
        return self.color

    def getRadius(self):
        return self.xy_radius

    def _add_to(self, h, gid):
        lib().oracle_add_polygon(h, *self._common_args(gid), self.N, ptr(self.position.data),
                                 ptr(self.angle.data), ptr(self.color.data), self.vertices.shape[0],
                                 ptr(self.vertices), self.rounding)


class Tracer(object):

    def __init__(self, device, w, h):
        self.device = device
        self.seed = 0
        self.highlight = False
        self.hw_color = numpy.array([1, 0, 1], dtype=numpy.float32)
        self.resize(w, h)

    def resize(self, w, h):
        if w == 0 or h == 0:
            raise RuntimeError("Invalid dimensions")
        self.w, self.h = w, h
        self.linear = _Array((h, w, 4))
        self.srgb = _Array((h, w, 4), dtype=numpy.uint8)

    def getSRGBOutputBuffer(self):
        return self.srgb

    def getLinearOutputBuffer(self):
        return self.linear

    def enableHighlightWarning(self, c):
        self.highlight = True
        self.hw_color = numpy.array([c.r, c.g, c.b], dtype=numpy.float32)

    def disableHighlightWarning(self):
        self.highlight = False

    def setSeed(self, s):
        self.seed = int(s) & 0xffffffff

    def getSeed(self):
        return self.seed

    def _check_scene(self, scene):
        if scene.device is not self.device:
            raise RuntimeError("Scene and Tracer devices do not match")


class TracerDirect(Tracer):

    def __init__(self, device, w, h):
        self.aa_n = 8
        super().__init__(device, w, h)

    def setAntialiasingN(self, n):
        self.aa_n = int(n)

    def getAntialiasingN(self):
        return self.aa_n

    def render(self, scene):
        self._check_scene(scene)
        h = scene.handle()
        lib().oracle_preview(h, self.w, self.h, self.seed, self.aa_n, ptr(self.linear.data),
                             ptr(self.srgb.data, _U8), 1 if self.highlight else 0, ptr(self.hw_color),
                             self.device.nthreads)


class TracerPath(Tracer):

    def __init__(self, device, w, h, light_samples):
        # the base constructor's resize is Tracer::resize (no seed change)
        super().__init__(device, w, h)
        self.light_samples = light_samples
        self.n_samples = 0
        self.reset()

    def resize(self, w, h):
        super().resize(w, h)
        if hasattr(self, 'n_samples'):
            self.n_samples = 0
            self.seed = (self.seed + 1) & 0xffffffff

    def reset(self):
        self.n_samples = 0
        self.seed = (self.seed + 1) & 0xffffffff
        self.linear.data[...] = 0
        self.srgb.data[...] = 0

    def setLightSamples(self, n):
        self.light_samples = int(n)

    def getNumSamples(self):
        return self.n_samples

    def render(self, scene):
        self._check_scene(scene)
        h = scene.handle()
        self.n_samples += 1
        lib().oracle_path_pass(h, self.w, self.h, self.seed, self.n_samples, self.light_samples,
                               ptr(self.linear.data), ptr(self.srgb.data, _U8),
                               1 if self.highlight else 0, ptr(self.hw_color), self.device.nthreads)
