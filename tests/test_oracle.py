"""Pins of the CPU checker that do not need the reference's images.

* Philox4x32-10 against known-answer vectors that rocRAND, PyTorch and
  Random123's published list agree on (tests/golden/philox_kat.json).
* u01 against its closed form and invariants.
* The sphere intersector against an FP64 solution of the same quadratic.
* BVH closest hit against brute force over all primitives.
* Camera rays against closed forms.
"""

import ctypes
import json
import math

import numpy
import pytest

import fresnel_amd as fresnel
import oracle
from conftest import GOLDEN
from oracle import ptr

U32 = ctypes.c_uint32


def test_philox_known_answers():
    data = json.load(open(GOLDEN / 'philox_kat.json'))['vectors']
    inp = numpy.array([v['counter'] + v['key'] for v in data], dtype=numpy.uint32)
    exp = numpy.array([v['out'] for v in data], dtype=numpy.uint32)
    out = numpy.zeros_like(exp)
    oracle.lib().oracle_philox(ptr(inp, U32), len(inp), ptr(out, U32))
    numpy.testing.assert_array_equal(out, exp)


def test_philox_detects_a_wrong_round_count():
    """A 9-round Philox must not pass the fixture (the test has teeth)."""
    data = json.load(open(GOLDEN / 'philox_kat.json'))['vectors'][0]

    def philox(c, k, rounds):
        c = list(c)
        k0, k1 = k
        for r in range(rounds):
            if r:
                k0 = (k0 + 0x9E3779B9) & 0xffffffff
                k1 = (k1 + 0xBB67AE85) & 0xffffffff
            p0 = 0xD2511F53 * c[0]
            p1 = 0xCD9E8D57 * c[2]
            c = [((p1 >> 32) ^ c[1] ^ k0) & 0xffffffff, p1 & 0xffffffff, ((p0 >> 32) ^ c[3] ^ k1) & 0xffffffff,
                 p0 & 0xffffffff]
        return c

    assert philox(data['counter'], data['key'], 10) == data['out']
    assert philox(data['counter'], data['key'], 9) != data['out']


def test_u01_closed_form():
    u = numpy.array([0, 1, 2, 0x7fffffff, 0x80000000, 0xfffffffe, 0xffffffff, 12345678], dtype=numpy.uint32)
    out = numpy.zeros(len(u), dtype=numpy.float32)
    oracle.lib().oracle_u01(ptr(u, U32), len(u), ptr(out))
    # float(u) * 2^-32 + 2^-33 rounded to float: exact values
    assert out[0] == numpy.float32(2.0**-33)
    assert out[-2] == numpy.float32(1.0)
    expect = (u.astype(numpy.float32) * numpy.float32(2.0**-32) + numpy.float32(2.0**-33)).astype(numpy.float32)
    numpy.testing.assert_array_equal(out, expect)
    # range (0, 1] and monotone
    many = numpy.sort(numpy.random.default_rng(0).integers(0, 2**32, 100000, dtype=numpy.uint64)).astype(
        numpy.uint32)
    v = numpy.zeros(len(many), numpy.float32)
    oracle.lib().oracle_u01(ptr(many, U32), len(many), ptr(v))
    assert numpy.all(v > 0) and numpy.all(v <= 1) and numpy.all(numpy.diff(v) >= 0)


def test_cr_sincosf_is_correctly_rounded():
    """The checker's cosf/sinf of lobe and aperture angles are the correctly rounded values: each lies
    within half an ulp of cos/sin evaluated in 80-bit extended precision, and is the nearest float to it."""
    rng = numpy.random.default_rng(2)
    x = (numpy.float32(2.0) * numpy.float32(math.pi)
         * rng.uniform(0, 1, 200000).astype(numpy.float32)).astype(numpy.float32)
    x[:4] = [0.0, numpy.float32(math.pi / 2), numpy.float32(math.pi), numpy.float32(2 * math.pi)]
    s = numpy.zeros_like(x)
    c = numpy.zeros_like(x)
    oracle.lib().oracle_cr_sincosf(ptr(x), len(x), ptr(s), ptr(c))
    xl = x.astype(numpy.longdouble)
    for got, exact in ((s, numpy.sin(xl)), (c, numpy.cos(xl))):
        err = numpy.abs(got.astype(numpy.longdouble) - exact)
        # the neighbouring floats are no closer than the result
        up = numpy.nextafter(got, numpy.float32(numpy.inf)).astype(numpy.longdouble)
        dn = numpy.nextafter(got, numpy.float32(-numpy.inf)).astype(numpy.longdouble)
        assert numpy.all(err <= numpy.abs(up - exact)) and numpy.all(err <= numpy.abs(dn - exact))
    assert c[0] == 1.0 and s[0] == 0.0


def _sphere_scene(pos, rad, brute=False):
    L = oracle.lib()
    h = L.oracle_scene_new()
    n = len(pos)
    mat = numpy.zeros(9, numpy.float32)
    col = numpy.zeros((n, 3), numpy.float32)
    L.oracle_add_sphere(h, 0, 1, ptr(mat), ptr(mat), 0.0, n, ptr(numpy.ascontiguousarray(pos, numpy.float32)),
                        ptr(numpy.ascontiguousarray(rad, numpy.float32)), ptr(col))
    L.oracle_scene_set_brute(h, 1 if brute else 0)
    L.oracle_scene_commit(h)
    return h


def _trace(h, rays, tnear=1e-3):
    n = len(rays)
    g = numpy.zeros(n, numpy.uint32)
    p = numpy.zeros(n, numpy.uint32)
    t = numpy.zeros(n, numpy.float32)
    attr = numpy.zeros((n, 7), numpy.float32)
    oracle.lib().oracle_trace(h, ptr(numpy.ascontiguousarray(rays, numpy.float32)), n, tnear, ptr(g, U32),
                              ptr(p, U32), ptr(t), ptr(attr), 0)
    return g, p, t, attr


def test_sphere_hand_cases():
    """Axis-aligned hit, origin inside (second root), miss, and the outline distance."""
    h = _sphere_scene([[0, 0, 0]], [1.0])
    rays = numpy.array([[0, 0, 10, 0, 0, -1],   # head on: t = 9
                        [0, 0, 0, 0, 0, -1],    # from the centre: second root t = 1
                        [0, 2, 10, 0, 0, -1],   # miss
                        [0.6, 0, 10, 0, 0, -1]], numpy.float32)
    g, p, t, attr = _trace(h, rays)
    assert t[0] == 9.0 and t[1] == 1.0
    assert g[2] == 0xffffffff
    # d_edge = r - |v x d| = 1 - 0.6
    assert attr[3, 3] == numpy.float32(1.0) - numpy.float32(0.6)
    # Ng = hit - centre (not flipped for the inside hit, cpu/GeometrySphere.cc:108-110)
    numpy.testing.assert_array_equal(attr[1, 0:3], [0, 0, -1])
    oracle.lib().oracle_scene_free(h)


def test_sphere_t_vs_fp64():
    """Float hit distances vs an FP64 solve; the reference formula is kept, so
    the relative error is only bounded where the discriminant does not cancel."""
    rng = numpy.random.default_rng(5)
    pos = numpy.array([[0, 0, 0]], numpy.float32)
    r = numpy.float32(0.5)
    h = _sphere_scene(pos, [r])
    n = 2000
    o = (rng.normal(size=(n, 3)) * 0.1 + [0, 0, 20]).astype(numpy.float32)
    tgt = rng.uniform(-0.35, 0.35, (n, 3)).astype(numpy.float32)
    d = tgt - o
    d = (d / numpy.linalg.norm(d, axis=1, keepdims=True)).astype(numpy.float32)
    rays = numpy.concatenate([o, d], axis=1).astype(numpy.float32)
    g, p, t, _ = _trace(h, rays)
    L = oracle.lib()
    rel = []
    for k in range(n):
        t64 = L.oracle_sphere_t64(ptr(rays[k, 0:3].copy()), ptr(rays[k, 3:6].copy()), ptr(pos[0].copy()), r, 1e-3)
        assert g[k] == 0 and not math.isnan(t64)
        rel.append(abs(float(t[k]) - t64) / t64)
    # rays aimed well inside the silhouette: |v| = 20, cancellation error ~ ulp(400)/sqrt(det)
    assert numpy.max(rel) < 2e-5
    oracle.lib().oracle_scene_free(h)


@pytest.mark.parametrize('seed', [0, 1])
def test_bvh_equals_brute_force(seed):
    rng = numpy.random.default_rng(seed)
    n = 3000
    pos = rng.uniform(-5, 5, (n, 3)).astype(numpy.float32)
    rad = rng.uniform(0.05, 0.5, n).astype(numpy.float32)
    hb = _sphere_scene(pos, rad)
    hf = _sphere_scene(pos, rad, brute=True)
    m = 5000
    o = rng.uniform(-6, 6, (m, 3)).astype(numpy.float32)
    d = rng.normal(size=(m, 3))
    d = (d / numpy.linalg.norm(d, axis=1, keepdims=True)).astype(numpy.float32)
    rays = numpy.concatenate([o, d], axis=1)
    a = _trace(hb, rays)
    b = _trace(hf, rays)
    for x, y in zip(a[:3], b[:3]):
        numpy.testing.assert_array_equal(x, y)
    oracle.lib().oracle_scene_free(hb)
    oracle.lib().oracle_scene_free(hf)


def test_tie_count_bvh_equals_brute_force():
    """Exact-tie counts found through the hierarchy equal brute force over every primitive,
    on a scene where a quarter of the spheres are duplicated (every hit on them is a tie)."""
    rng = numpy.random.default_rng(11)
    n = 2000
    pos = rng.uniform(-5, 5, (n, 3)).astype(numpy.float32)
    rad = rng.uniform(0.1, 0.6, n).astype(numpy.float32)
    pos = numpy.concatenate([pos, pos[:n // 4]])
    rad = numpy.concatenate([rad, rad[:n // 4]])
    hb = _sphere_scene(pos, rad)
    hf = _sphere_scene(pos, rad, brute=True)
    m = 4000
    o = rng.uniform(-6, 6, (m, 3)).astype(numpy.float32)
    d = rng.normal(size=(m, 3))
    d = (d / numpy.linalg.norm(d, axis=1, keepdims=True)).astype(numpy.float32)
    rays = numpy.ascontiguousarray(numpy.concatenate([o, d], axis=1), numpy.float32)
    tb = numpy.zeros(m, numpy.uint32)
    tf = numpy.zeros(m, numpy.uint32)
    oracle.lib().oracle_count_ties(hb, ptr(rays), m, 1e-3, ptr(tb, U32))
    oracle.lib().oracle_count_ties(hf, ptr(rays), m, 1e-3, ptr(tf, U32))
    numpy.testing.assert_array_equal(tb, tf)
    assert (tf >= 2).sum() > 50  # the duplicates do produce ties
    oracle.lib().oracle_scene_free(hb)
    oracle.lib().oracle_scene_free(hf)


def test_orthographic_rays_closed_form(cpu_device):
    """Orthographic rays point along -w exactly and start on the image plane."""
    scene = fresnel.Scene(cpu_device)
    fresnel.geometry.Sphere(scene, position=[[0, 0, 0]], radius=1)
    scene.camera = fresnel.camera.Orthographic(position=(0, 0, 10), look_at=(0, 0, 0), up=(0, 1, 0), height=4)
    W, H = 8, 4
    out = numpy.zeros((W * H, 6), numpy.float32)
    oracle.lib().oracle_camera_rays(scene._scene.handle(), W, H, 1, 1, 0, ptr(out))
    numpy.testing.assert_array_equal(out[:, 3:6], numpy.tile([0, 0, -1], (W * H, 1)))
    assert numpy.all(out[:, 2] == 10)
    # pixel (i, j) centre: x = ((i + 0.5)/H - W/(2H)) * h, y = -((j + 0.5)/H - 0.5) * h
    i, j = numpy.meshgrid(numpy.arange(W), numpy.arange(H))
    x = ((i.ravel() + 0.5) / H - W / (2 * H)) * 4
    y = -((j.ravel() + 0.5) / H - 0.5) * 4
    numpy.testing.assert_allclose(out[:, 0], x, rtol=0, atol=1e-6)
    numpy.testing.assert_allclose(out[:, 1], y, rtol=0, atol=1e-6)


def test_perspective_rays_unit_and_through_focal_plane(cpu_device):
    """Pinhole rays are unit length and pass through the focal-plane point of their pixel."""
    scene = fresnel.Scene(cpu_device)
    fresnel.geometry.Sphere(scene, position=[[0, 0, 0]], radius=1)
    cam = fresnel.camera.Perspective(position=(0, 0, 10), look_at=(0, 0, 0), up=(0, 1, 0), focal_length=0.5,
                                     focus_distance=10, height=0.24)
    scene.camera = cam
    W, H = 6, 4
    out = numpy.zeros((W * H, 6), numpy.float32)
    oracle.lib().oracle_camera_rays(scene._scene.handle(), W, H, 1, 1, 0, ptr(out))
    numpy.testing.assert_allclose(numpy.linalg.norm(out[:, 3:6], axis=1), 1, atol=1e-6)
    numpy.testing.assert_array_equal(out[:, 0:3], numpy.tile([0, 0, 10], (W * H, 1)))
    # the focal plane is z = 0 with height h/f*focus = 4.8
    i, j = numpy.meshgrid(numpy.arange(W), numpy.arange(H))
    x = ((i.ravel() + 0.5) / H - W / (2 * H)) * 4.8
    y = -((j.ravel() + 0.5) / H - 0.5) * 4.8
    s = -10 / out[:, 5]
    numpy.testing.assert_allclose(out[:, 3] * s, x, atol=1e-5)
    numpy.testing.assert_allclose(out[:, 4] * s, y, atol=1e-5)


def _f32_sphere_t(o, d, p, r):
    """The reference's sphere solve (IntersectSphere.h) in float32, first root."""
    f = numpy.float32
    v = (p - o).astype(f)
    b = f(f(f(v[0] * d[0]) + f(v[1] * d[1])) + f(v[2] * d[2]))
    vv = f(f(f(v[0] * v[0]) + f(v[1] * v[1])) + f(v[2] * v[2]))
    det = f(f(f(b * b) - vv) + f(r * r))
    return f(b - f(numpy.sqrt(det)))


def _f32_slab_tmin(o, d, lo, hi):
    f = numpy.float32
    inv = [f(f(1) / d[k]) for k in range(3)]
    t0 = [f(f(lo[k] - o[k]) * inv[k]) for k in range(3)]
    t1 = [f(f(hi[k] - o[k]) * inv[k]) for k in range(3)]
    return max(min(t0[0], t1[0]), min(t0[1], t1[1]), min(t0[2], t1[2]))


def test_cull_keeps_grazing_hits_below_their_box_entry():
    """A distant grazing hit whose computed t lies 5e-4 (relative) below its box's slab entry.

    Sphere B (primitive 0) is tested first and hit at t_B, between sphere A's
    computed distance and A's box entry.  A cull by entry distance with a small
    relative slack would drop A; the error-bounded cull keeps it, so the
    hierarchy's answer equals brute force: A.
    """
    f = numpy.float32
    o = numpy.zeros(3, f)
    d = numpy.array([0.9991304874420166, -0.03203902393579483, 0.026680011302232742], f)
    pA = numpy.array([74.91876220703125, -2.902667999267578, 2.0010008811950684], f)
    r = f(0.5)
    tA = _f32_sphere_t(o, d, pA, r)
    entryA = _f32_slab_tmin(o, d, pA - r, pA + r)
    assert (entryA - tA) / tA > 5e-4
    pB = (o + d * f(74.97 + 0.5)).astype(f)
    tB = _f32_sphere_t(o, d, pB, r)
    assert tA < tB and entryA > tB * (1 + 2 ** -16)  # a relative-slack cull would drop A
    rays = numpy.array([numpy.concatenate([o, d])], numpy.float32)
    pos = numpy.array([pB, pA], numpy.float32)
    rad = numpy.array([r, r], numpy.float32)
    for brute in (False, True):
        h = _sphere_scene(pos, rad, brute=brute)
        g, p, t, _ = _trace(h, rays)
        oracle.lib().oracle_scene_free(h)
        assert p[0] == 1 and t[0] == tA, f"brute={brute}"


def grazing_case(m=30000, seed=21):
    """Adversarial scene for the cull rule: pairs of a distant grazing hit A and a small
    sphere B hit between A's computed distance and A's box entry.

    Returns (positions, radii, rays (k, 6), index of A for each ray).
    """
    f = numpy.float32
    rng = numpy.random.default_rng(seed)
    d = rng.normal(size=(m, 3)) * [1, 0.2, 0.2] + [5, 0, 0]
    d = (d / numpy.linalg.norm(d, axis=1, keepdims=True)).astype(f)
    dist = rng.uniform(40, 120, m)
    r = rng.uniform(0.3, 0.6, m).astype(f)
    perp = numpy.cross(d, rng.normal(size=(m, 3)))
    perp /= numpy.linalg.norm(perp, axis=1, keepdims=True)
    pA = (d * dist[:, None] + perp * (r[:, None] * (1 - rng.uniform(0, 4e-4, (m, 1))))).astype(f)
    o = numpy.zeros(3, f)
    pos, rad, rays, want = [], [], [], []
    for i in range(m):
        v = (pA[i] - o).astype(f)
        b = f(f(f(v[0] * d[i, 0]) + f(v[1] * d[i, 1])) + f(v[2] * d[i, 2]))
        vv = f(f(f(v[0] * v[0]) + f(v[1] * v[1])) + f(v[2] * v[2]))
        det = f(f(f(b * b) - vv) + f(r[i] * r[i]))
        if det < 0:
            continue
        tA = f(b - f(numpy.sqrt(det)))
        entry = _f32_slab_tmin(o, d[i], pA[i] - r[i], pA[i] + r[i])
        if not entry > tA * (1 + 2 ** -14):
            continue
        tB = 0.5 * (float(tA) + float(entry))
        pos.append(d[i] * f(tB + 0.05))
        rad.append(0.05)
        pos.append(pA[i])
        rad.append(r[i])
        rays.append(numpy.concatenate([o, d[i]]))
        want.append(len(pos) - 1)
    return numpy.array(pos, f), numpy.array(rad, f), numpy.array(rays, f), numpy.array(want)


def test_bvh_equals_brute_force_on_distant_grazing_rays():
    """Hierarchy with culling == brute force on the adversarial grazing pairs, hit for hit."""
    pos, rad, rays, want = grazing_case()
    assert len(rays) > 100
    hb = _sphere_scene(pos, rad, brute=False)
    hf = _sphere_scene(pos, rad, brute=True)
    gb, pb, tb, _ = _trace(hb, rays)
    gf, pf, tf, _ = _trace(hf, rays)
    oracle.lib().oracle_scene_free(hb)
    oracle.lib().oracle_scene_free(hf)
    assert numpy.sum(pf == want) > 50  # the adversarial pairs survive in the shared scene
    numpy.testing.assert_array_equal(pb, pf)
    numpy.testing.assert_array_equal(tb.view(numpy.uint32), tf.view(numpy.uint32))
