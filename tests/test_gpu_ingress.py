"""Scene-update ingress and image egress on the device.

* Writing rows of a geometry buffer (``g.position[5:9] = ...``) uploads only
  those rows; the next image must equal, bit for bit, the image of a scene
  built from scratch with the final data (same camera, same seed).  Colour
  rows leave the hierarchy alone (no build or refit counted), other rows
  rebuild it, or refit it when refit is enabled.  Every geometry type and
  index form (integer, slice, negative step, integer array, boolean mask) is
  covered.
* ``Tracer.histogram`` bins on the device.  It must equal the reference's
  numpy formulation (fresnel/tracer.py:79-108) applied to the same linear
  output, up to values whose gamma-corrected level lies within a few float
  ulps of a bin edge, where the C library's and the device's powf may round
  to different sides.
"""

import numpy
import pytest

import fresnel_amd as fresnel
from conftest import test_lights

pytestmark = pytest.mark.gpu

W, H = 80, 60


def _mat():
    return fresnel.material.Material(color=fresnel.color.linear([0.6, 0.5, 0.4]), roughness=0.4,
                                     primitive_color_mix=1.0)


def _cube():
    v = numpy.array([[x, y, z] for x in (-0.5, 0.5) for y in (-0.5, 0.5) for z in (-0.25, 0.25)])
    return fresnel.util.convex_polyhedron_from_vertices(v)


def _tetra_faces():
    p = numpy.array([[1, 1, 1], [-1, -1, 1], [-1, 1, -1], [1, -1, -1]], numpy.float32) * 0.4
    return numpy.array([p[0], p[1], p[2], p[0], p[3], p[1], p[0], p[2], p[3], p[1], p[3], p[2]], numpy.float32)


def _quat(rng, n):
    q = rng.normal(size=(n, 4))
    return (q / numpy.linalg.norm(q, axis=1)[:, None]).astype(numpy.float32)


def _data(kind, rng, n):
    d = {}
    if kind == 'cylinder':
        a = rng.uniform(-6, 6, (n, 3))
        d['points'] = numpy.stack([a, a + rng.normal(0, 0.8, (n, 3))], axis=1).astype(numpy.float32)
        d['radius'] = rng.uniform(0.1, 0.3, n).astype(numpy.float32)
        d['color'] = rng.uniform(0, 1, (n, 2, 3)).astype(numpy.float32)
        return d
    if kind == 'polygon':
        d['position'] = rng.uniform(-6, 6, (n, 2)).astype(numpy.float32)
        d['angle'] = rng.uniform(0, 6.28, n).astype(numpy.float32)
        d['color'] = rng.uniform(0, 1, (n, 3)).astype(numpy.float32)
        return d
    d['position'] = rng.uniform(-6, 6, (n, 3)).astype(numpy.float32)
    if kind == 'sphere':
        d['radius'] = rng.uniform(0.2, 0.6, n).astype(numpy.float32)
        d['color'] = rng.uniform(0, 1, (n, 3)).astype(numpy.float32)
    else:
        d['orientation'] = _quat(rng, n)
        d['color'] = rng.uniform(0, 1, ((12 if kind == 'mesh' else n), 3)).astype(numpy.float32)
    return d


def _build(device, kind, d, n):
    s = fresnel.Scene(device, lights=test_lights())
    m = _mat()
    if kind == 'sphere':
        g = fresnel.geometry.Sphere(s, position=d['position'], radius=d['radius'], color=d['color'], material=m)
    elif kind == 'cylinder':
        g = fresnel.geometry.Cylinder(s, points=d['points'], radius=d['radius'], color=d['color'], material=m)
    elif kind == 'polyhedron':
        g = fresnel.geometry.ConvexPolyhedron(s, _cube(), position=d['position'], orientation=d['orientation'],
                                              color=d['color'], material=m)
    elif kind == 'mesh':
        g = fresnel.geometry.Mesh(s, _tetra_faces(), position=d['position'], orientation=d['orientation'],
                                  color=d['color'], material=m, N=n)
    else:
        g = fresnel.geometry.Polygon(s, [[-0.4, -0.3], [0.5, -0.3], [0.0, 0.45]], position=d['position'],
                                     angle=d['angle'], color=d['color'], material=m)
    return s, g


def _render(device, s, seed=7):
    t = fresnel.tracer.Path(device, w=W, h=H)
    t.seed = seed
    t.sample(s, samples=3, light_samples=2)
    return numpy.array(t.linear_output[:])


# (buffer, index) edits applied in order; each is followed by a render
EDITS = {
    'sphere': [('color', slice(10, 40)), ('position', slice(5, 9)), ('radius', 17), ('position', -1),
               ('color', numpy.array([3, 90, 44])), ('radius', 'mask'), ('position', slice(80, 20, -7))],
    'cylinder': [('color', (slice(4, 9), 1)), ('points', slice(0, 3)), ('radius', numpy.array([7, 2]))],
    'polyhedron': [('color', 5), ('orientation', slice(10, 30)), ('position', 'mask')],
    'mesh': [('color', slice(3, 6)), ('orientation', 4), ('position', numpy.array([0, 11]))],
    'polygon': [('color', slice(0, 5)), ('angle', slice(3, 12, 2)), ('position', 8)],
}


@pytest.mark.parametrize('refit', [False, True])
@pytest.mark.parametrize('kind', sorted(EDITS))
def test_row_updates_equal_fresh_scene(kind, refit, gpu_device):
    n = 120 if kind != 'mesh' else 16
    rng = numpy.random.default_rng(sum(map(ord, kind)))
    d = _data(kind, rng, n)
    s, g = _build(gpu_device, kind, d, n)
    s.set_bvh_refit(refit, limit=100.0)
    s.camera = fresnel.camera.Orthographic.fit(s, margin=0.2)
    _render(gpu_device, s)
    for which, index in EDITS[kind]:
        arr = d[which]
        if isinstance(index, str):
            index = rng.random(arr.shape[0]) < 0.1
        old = arr.copy()
        new = arr.copy()
        if which == 'color':
            new[index] = rng.uniform(0, 1, new[index].shape)
        else:
            new[index] = new[index] + rng.normal(0, 0.3, new[index].shape).astype(numpy.float32)
        d[which] = new
        assert not numpy.array_equal(old, new)
        b0, r0, _ = s._scene.bvhInfo()
        getattr(g, which)[index] = new[index]
        img = _render(gpu_device, s)
        b1, r1, _ = s._scene.bvhInfo()
        if which == 'color':
            assert (b1, r1) == (b0, r0), f"colour rows changed the hierarchy ({kind})"
        elif refit:
            assert b1 + r1 == b0 + r0 + 1
        else:
            assert (b1, r1) == (b0 + 1, r0)
        s2, _ = _build(gpu_device, kind, d, n)
        s2.camera = s.camera
        assert numpy.array_equal(img, _render(gpu_device, s2)), f"{kind} after {which}[{index}]"


@pytest.mark.parametrize('kind', sorted(EDITS))
def test_row_updates_match_checker(kind, gpu_device, cpu_device):
    """After every row edit of EDITS, the edited GPU scene's primary hits (renderer kernel) and linear image
    equal, bit for bit, those of the CPU checker's scene built from the edited arrays."""
    import ctypes
    import oracle
    from oracle import ptr
    n = 120 if kind != 'mesh' else 16
    rng = numpy.random.default_rng(7 + sum(map(ord, kind)))
    d = _data(kind, rng, n)
    s, g = _build(gpu_device, kind, d, n)
    s.set_bvh_refit(True, limit=100.0)
    s.camera = fresnel.camera.Orthographic.fit(s, margin=0.2)
    _render(gpu_device, s)
    for which, index in EDITS[kind]:
        arr = d[which]
        if isinstance(index, str):
            index = rng.random(arr.shape[0]) < 0.1
        new = arr.copy()
        if which == 'color':
            new[index] = rng.uniform(0, 1, new[index].shape)
        else:
            new[index] = new[index] + rng.normal(0, 0.3, new[index].shape).astype(numpy.float32)
        if which == 'orientation':
            # unit quaternions: a non-unit one scales the shape beyond its bounds (position +- circumradius,
            # as in the reference), and the closest hit then depends on the hierarchy
            q = numpy.atleast_2d(new[index])
            new[index] = (q / numpy.linalg.norm(q, axis=1, keepdims=True)).reshape(new[index].shape)
        if which == 'radius':
            # radii >= 0: a negative radius shrinks or inverts the bounds (the reference's A -+ r) while the
            # intersector, which squares it, reports hits outside them, so the closest hit again depends on
            # the hierarchy's culling
            new[index] = numpy.abs(new[index])
        d[which] = new
        getattr(g, which)[index] = new[index]
        img = _render(gpu_device, s)
        cs, _ = _build(cpu_device, kind, d, n)
        cs.camera = s.camera
        assert numpy.array_equal(img.view(numpy.uint32), _render(cpu_device, cs).view(numpy.uint32)), \
            f"{kind} after {which}[{index}] differs from the checker"
        t = fresnel.tracer.Path(gpu_device, w=W, h=H)
        t.seed = 3
        gid, prim, tt = t._tracer.debugPrimaryRender(s._scene, 1)
        rays = numpy.zeros((W * H, 6), numpy.float32)
        oracle.lib().oracle_camera_rays(cs._scene.handle(), W, H, 3, 1, 1, ptr(rays))
        rg = numpy.zeros(W * H, numpy.uint32)
        rp = numpy.zeros(W * H, numpy.uint32)
        rt = numpy.zeros(W * H, numpy.float32)
        oracle.lib().oracle_trace(cs._scene.handle(), ptr(rays), W * H, 1e-3, ptr(rg, ctypes.c_uint32),
                                  ptr(rp, ctypes.c_uint32), ptr(rt), None, 0)
        assert numpy.array_equal(gid, rg) and numpy.array_equal(prim, rp), f"{kind} hits after {which}"
        assert numpy.array_equal(tt.view(numpy.uint32), rt.view(numpy.uint32))


def test_async_output_equals_map_and_survives_a_render(gpu_device):
    """output_async returns the image of the last render, equal to map(); a render started before the
    wait does not change it; both outputs; the size is the one at the copy's start."""
    rng = numpy.random.default_rng(9)
    d = _data('sphere', rng, 200)
    s, _ = _build(gpu_device, 'sphere', d, 200)
    s.camera = fresnel.camera.Orthographic.fit(s, margin=0.2)
    t = fresnel.tracer.Path(gpu_device, w=W, h=H)
    t.sample(s, samples=2)
    first = (numpy.array(t.output[:]), numpy.array(t.linear_output[:]))
    a = t.output_async()
    b = t.output_async(linear=True)
    t.sample(s, samples=3, reset=False)  # renders while the copies may still run
    assert numpy.array_equal(a.wait(), first[0])
    assert numpy.array_equal(b.wait(), first[1])
    assert a.ready() and b.ready()
    second = numpy.array(t.linear_output[:])
    assert not numpy.array_equal(second, first[1])
    c = t.output_async(linear=True)
    t.resize(40, 30)
    got = c.wait()
    assert got.shape == (H, W, 4) and numpy.array_equal(got, second)


def test_update_rows_rejects_rows_outside_the_buffer(gpu_device):
    s, g = _build(gpu_device, 'sphere', _data('sphere', numpy.random.default_rng(0), 10), 10)
    buf = g._geometry.getPositionBuffer()
    buf.update_rows(9, 1)
    buf.update_rows(10, 0)
    for first, count in ((10, 1), (0, 11), (2**63, 2**63)):
        with pytest.raises(RuntimeError):
            buf.update_rows(first, count)


def _numpy_histogram(linear):
    """The reference's formulation (fresnel/tracer.py:90-107), plus the values
    whose level lies within 4 float32 ulps of a bin edge or of 1."""
    img = linear[linear[:, :, 3] > 0]
    r, g, b = img[:, 0], img[:, 1], img[:, 2]
    lightness = 0.21 * r + 0.72 * g + 0.07 * b
    cols, close = [], []
    for ch in (lightness, r, g, b):
        level = ch**(1 / 2.2)
        hist, bins = numpy.histogram(level, bins=512, range=[0, 1])
        cols.append(hist)
        ulp = numpy.spacing(numpy.abs(level).astype(numpy.float32)) * 4
        lo = numpy.floor((level.astype(numpy.float64) - ulp) * 512)
        hi = numpy.floor((level.astype(numpy.float64) + ulp) * 512)
        close.append(int(numpy.count_nonzero((lo != hi) | (numpy.abs(level - 1.0) <= ulp))))
    return numpy.stack(cols, axis=1), bins[1:], close


@pytest.mark.parametrize('tracer', ['path', 'preview'])
def test_histogram_on_device_matches_reference_formula(tracer, gpu_device):
    rng = numpy.random.default_rng(4)
    d = _data('sphere', rng, 300)
    s, _ = _build(gpu_device, 'sphere', d, 300)
    s.camera = fresnel.camera.Orthographic.fit(s, margin=0.3)
    s.background_alpha = 0.0
    if tracer == 'path':
        t = fresnel.tracer.Path(gpu_device, w=200, h=150)
        t.sample(s, samples=8, light_samples=2)
    else:
        t = fresnel.tracer.Preview(gpu_device, w=200, h=150)
        t.render(s)
    linear = numpy.array(t.linear_output[:])
    assert (linear[:, :, 3] == 0).any() and (linear[:, :, 3] > 0).any()
    want, want_bins, close = _numpy_histogram(linear)
    got, bins = t.histogram()
    assert got.shape == (512, 4)
    assert numpy.array_equal(bins, want_bins)
    assert got[:, 0].sum() > 0
    for k in range(4):
        diff = numpy.abs(got[:, k].astype(numpy.int64) - want[:, k])
        assert diff.sum() <= 2 * close[k], (k, diff.sum(), close[k])


def test_histogram_of_a_blank_image_is_empty(gpu_device):
    t = fresnel.tracer.Path(gpu_device, w=16, h=8)
    got, bins = t.histogram()
    assert got.shape == (512, 4) and got.sum() == 0
    assert bins[0] == numpy.float32(1 / 512) and bins[-1] == 1.0
