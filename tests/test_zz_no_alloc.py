"""The allocation-free histogram entry point against the allocating one.

fh_tracer_histogram_noalloc accumulates the bins in a file-scope device
array instead of a buffer the tracer allocates.  Called through the C ABI
(ctypes on libfresnel_hip.so) on two images in a row, the second smaller, it
must return exactly what fh_tracer_histogram returns for each.
"""

import ctypes
import os

import numpy
import pytest

import fresnel_amd as fresnel

pytestmark = pytest.mark.gpu


def _lib():
    path = os.path.join(os.path.dirname(fresnel.__file__), 'libfresnel_hip.so')
    L = ctypes.CDLL(path)
    for name in ('fh_tracer_histogram', 'fh_tracer_histogram_noalloc'):
        f = getattr(L, name)
        f.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32)]
        f.restype = ctypes.c_int
    L.fh_last_error.restype = ctypes.c_char_p
    return L


def _scene(device, n, seed):
    rng = numpy.random.default_rng(seed)
    s = fresnel.Scene(device, lights=fresnel.light.cloudy())
    fresnel.geometry.Sphere(s, position=rng.uniform(-8, 8, (n, 3)).astype(numpy.float32),
                            radius=rng.uniform(0.3, 0.9, n).astype(numpy.float32), outline_width=0.05,
                            color=rng.uniform(0, 1, (n, 3)).astype(numpy.float32),
                            material=fresnel.material.Material(primitive_color_mix=1.0, roughness=0.4))
    s.camera = fresnel.camera.Orthographic.fit(s)
    return s


def _hist(L, fn, handle):
    out = numpy.zeros((512, 4), numpy.uint32)
    rc = getattr(L, fn)(handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    assert rc == 0, L.fh_last_error()
    return out


def test_histogram_noalloc_equals_allocating(gpu_device):
    L = _lib()
    for (w, h, n, seed) in ((320, 240, 3000, 1), (96, 64, 200, 2)):
        s = _scene(gpu_device, n, seed)
        t = fresnel.tracer.Path(gpu_device, w=w, h=h)
        t.sample(s, samples=4, light_samples=1)
        handle = t._tracer.handle()
        ref = _hist(L, 'fh_tracer_histogram', handle)
        got = _hist(L, 'fh_tracer_histogram_noalloc', handle)
        assert ref.sum() > 0
        numpy.testing.assert_array_equal(got, ref)
        # each column counts at most the visible pixels of this image (nothing left from the previous call)
        vis = int((numpy.array(t.linear_output[:])[:, :, 3] > 0).sum())
        assert int(got[:, 1].sum()) <= vis
