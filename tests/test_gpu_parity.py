"""GPU backend vs the CPU checker on the same inputs.

Bars (stated per test):
* Philox words, camera rays, primary hit ids and hit distances: bit-exact.
  A primary hit id may differ only where the checker finds an exact tie in t,
  and such pixels are counted.
* Linear HDR image at equal seed and sample count: bit-identical to the
  checker's (the device's shading performs the reference's float operations
  in the reference's order, with correctly rounded division, square root,
  sine and cosine, as the checker does); the RMSE bar of 1e-3 is printed too.
  The 8-bit sRGB image may differ by one level where powf rounds differently.
* The GPU's own schedules (several samples per submission, several scenes in
  one batch) are bit-identical to one-at-a-time rendering.
"""

import ctypes

import numpy
import pytest

import fresnel_amd as fresnel
import oracle
from oracle import ptr
import scenes

pytestmark = pytest.mark.gpu

MISS = 0xffffffff


def c1_scene(device, n=4096, w_seed=0):
    """The small sphere scene: 4096 spheres in a cube of side 20, r=0.5, outline 0.1."""
    rng = numpy.random.default_rng(w_seed)
    pos = rng.uniform(-10, 10, (n, 3)).astype(numpy.float32)
    s = fresnel.Scene(device, lights=fresnel.light.cloudy())
    fresnel.geometry.Sphere(s, position=pos, radius=0.5, outline_width=0.1,
                            material=fresnel.material.Material(color=fresnel.color.linear([0.1, 0.8, 0.1]),
                                                               roughness=0.8, specular=0.2))
    s.camera = fresnel.camera.Orthographic.fit(s)
    return s


PRIMARY_SCENES = {
    'hex_sphere': scenes.hex_sphere,
    'four_spheres': scenes.four_spheres,
    'four_cylinders': scenes.four_cylinders,
    'box': scenes.box,
    'eight_polyhedra': scenes.eight_polyhedra,
    'tetrahedra': scenes.tetrahedra,
    'one_triangle': scenes.one_triangle,
    'polygons': scenes.polygons,
    'rounded_polygons': scenes.rounded_polygons,
    'perspective_grid': scenes.perspective_grid,
    'c1_spheres': c1_scene,
}


def test_philox_words(gpu_device):
    """Device Philox4x32-10 words equal the checker's for random counters and keys."""
    rng = numpy.random.default_rng(7)
    inp = rng.integers(0, 2**32, size=(4096, 6), dtype=numpy.uint64).astype(numpy.uint32)
    inp[:16] = 0
    inp[16:32] = 0xffffffff
    dev = gpu_device._device.debugPhilox(inp)
    ref = numpy.zeros((4096, 4), dtype=numpy.uint32)
    oracle.lib().oracle_philox(ptr(inp, ctypes.c_uint32), 4096, ptr(ref, ctypes.c_uint32))
    numpy.testing.assert_array_equal(dev, ref)


@pytest.mark.parametrize('name', sorted(PRIMARY_SCENES))
def test_primary_rays_and_hits(name, gpu_device, cpu_device):
    """Camera rays bit-exact; primary hit (geometry, primitive, t) bit-exact except counted exact ties."""
    W, H = 96, 64
    gs = PRIMARY_SCENES[name](gpu_device)
    cs = PRIMARY_SCENES[name](cpu_device)
    tr = fresnel.tracer.Path(gpu_device, w=W, h=H)
    tr.seed = 12345
    seed = tr.seed
    for sample in (1, 7):
        rays, gid, prim, t = tr._tracer.debugPrimary(gs._scene, sample)
        h = cs._scene.handle()
        L = oracle.lib()
        ref_rays = numpy.zeros((W * H, 6), dtype=numpy.float32)
        L.oracle_camera_rays(h, W, H, seed, sample, 1, ptr(ref_rays))
        assert rays.view(numpy.uint32).tobytes() == ref_rays.view(numpy.uint32).tobytes(), \
            f"camera rays differ in {numpy.sum(numpy.any(rays != ref_rays, axis=1))} pixels"
        n = W * H
        rg = numpy.zeros(n, numpy.uint32)
        rp = numpy.zeros(n, numpy.uint32)
        rt = numpy.zeros(n, numpy.float32)
        L.oracle_trace(h, ptr(ref_rays), n, 1e-3, ptr(rg, ctypes.c_uint32), ptr(rp, ctypes.c_uint32), ptr(rt), None,
                       0)
        diff = (gid != rg) | (prim != rp)
        if diff.any():
            ties = numpy.zeros(n, numpy.uint32)
            sel = numpy.ascontiguousarray(ref_rays[diff])
            tsel = numpy.zeros(len(sel), numpy.uint32)
            L.oracle_count_ties(h, ptr(sel), len(sel), 1e-3, ptr(tsel, ctypes.c_uint32))
            assert numpy.all(tsel >= 2), f"{int(numpy.sum(tsel < 2))} primary hits differ without an exact tie"
        same = ~diff
        assert numpy.array_equal(t[same].view(numpy.uint32), rt[same].view(numpy.uint32)), "hit distances differ"
        assert (gid != MISS).any(), "scene not visible"


@pytest.mark.parametrize('tiles', [2, 1], ids=['tile_lists', 'bvh_packets'])
@pytest.mark.parametrize('name', sorted(PRIMARY_SCENES))
def test_render_primary_hits_vs_checker(name, tiles, gpu_device, cpu_device):
    """The renderer's own primary-ray kernel (tile lists, or BVH wavefront packets) returns the
    checker's primary hits, bit for bit, except counted exact ties."""
    W, H = 100, 60  # not multiples of 8: the edge strips take the per-ray search
    d = gpu_device._device
    saved = d.tune('primary_tiles', 0)
    try:
        d.tune('primary_tiles', tiles)
        gs = PRIMARY_SCENES[name](gpu_device)
        cs = PRIMARY_SCENES[name](cpu_device)
        tr = fresnel.tracer.Path(gpu_device, w=W, h=H)
        tr.seed = 321
        for sample in (1, 5):
            gid, prim, t = tr._tracer.debugPrimaryRender(gs._scene, sample)
            h = cs._scene.handle()
            L = oracle.lib()
            ref_rays = numpy.zeros((W * H, 6), dtype=numpy.float32)
            L.oracle_camera_rays(h, W, H, tr.seed, sample, 1, ptr(ref_rays))
            n = W * H
            rg = numpy.zeros(n, numpy.uint32)
            rp = numpy.zeros(n, numpy.uint32)
            rt = numpy.zeros(n, numpy.float32)
            L.oracle_trace(h, ptr(ref_rays), n, 1e-3, ptr(rg, ctypes.c_uint32), ptr(rp, ctypes.c_uint32), ptr(rt),
                           None, 0)
            diff = (gid != rg) | (prim != rp)
            if diff.any():
                sel = numpy.ascontiguousarray(ref_rays[diff])
                tsel = numpy.zeros(len(sel), numpy.uint32)
                L.oracle_count_ties(h, ptr(sel), len(sel), 1e-3, ptr(tsel, ctypes.c_uint32))
                assert numpy.all(tsel >= 2), f"{int(numpy.sum(tsel < 2))} primary hits differ without an exact tie"
            same = ~diff
            assert numpy.array_equal(t[same].view(numpy.uint32), rt[same].view(numpy.uint32)), "hit distances differ"
            assert (gid != MISS).any(), "scene not visible"
    finally:
        d.tune('primary_tiles', saved)


def _linear(device, builder, W, H, samples, L, seed=None):
    s = builder(device)
    t = fresnel.tracer.Path(device, w=W, h=H)
    if seed is not None:
        t.seed = seed
    t.sample(s, samples=samples, light_samples=L, reset=seed is None)
    return numpy.array(t.linear_output[:]), numpy.array(t.output[:])


@pytest.mark.parametrize('name,samples,L', [('c1_spheres', 4, 2), ('hex_sphere', 8, 4), ('four_cylinders', 4, 2),
                                            ('eight_polyhedra', 4, 2), ('tetrahedra', 4, 2),
                                            ('perspective_grid', 4, 2)])
def test_path_image_rmse(name, samples, L, gpu_device, cpu_device):
    """Linear HDR image vs the checker at equal seed and samples: bit-identical (and so RMSE <= 1e-3)."""
    W, H = 96, 64
    g, gs = _linear(gpu_device, PRIMARY_SCENES[name], W, H, samples, L)
    c, cs = _linear(cpu_device, PRIMARY_SCENES[name], W, H, samples, L)
    rmse = float(numpy.sqrt(numpy.mean((g - c)**2)))
    exact = float(numpy.mean(numpy.all(g == c, axis=2)))
    print(f"{name}: RMSE {rmse:.3e}, bit-identical pixels {exact:.4f}, "
          f"max sRGB diff {int(numpy.max(numpy.abs(gs.astype(int) - cs.astype(int))))}")
    assert numpy.all(numpy.isfinite(g))
    assert rmse <= 1e-3
    assert numpy.array_equal(g.view(numpy.uint32), c.view(numpy.uint32)), "linear image differs from the checker"


def test_render_samples_equals_render_loop(gpu_device):
    """S passes in one submission equal S render() calls, bit for bit."""
    s = c1_scene(gpu_device)
    a = fresnel.tracer.Path(gpu_device, w=80, h=60)
    a.sample(s, samples=6, light_samples=3)
    b = fresnel.tracer.Path(gpu_device, w=80, h=60)
    b.reset()  # sample(reset=True) advanced a's seed the same way
    b._tracer.setLightSamples(3)
    for _ in range(6):
        b._tracer.render(s._scene)
    assert a._tracer.getNumSamples() == b._tracer.getNumSamples() == 6
    assert numpy.array_equal(numpy.array(a.linear_output[:]), numpy.array(b.linear_output[:]))
    assert numpy.array_equal(numpy.array(a.output[:]), numpy.array(b.output[:]))


def test_batch_equals_plain(gpu_device):
    """Several scenes of uneven size in one batch equal their plain renders, bit for bit."""
    builders = [lambda d: c1_scene(d, n=1000, w_seed=1), scenes.four_cylinders, scenes.eight_polyhedra,
                lambda d: c1_scene(d, n=5000, w_seed=2)]
    sizes = [(64, 48), (80, 40), (50, 50), (96, 64)]
    plain = []
    for b, (w, h) in zip(builders, sizes):
        plain.append(_linear(gpu_device, b, w, h, 5, 2)[0])
    sc = [b(gpu_device) for b in builders]
    trs = [fresnel.tracer.Path(gpu_device, w=w, h=h) for (w, h) in sizes]
    for t in trs:
        t.reset()
        t._tracer.setLightSamples(2)
    fresnel._gpu.batch_render([s._scene for s in sc], [t._tracer for t in trs], 5)
    for k, t in enumerate(trs):
        assert numpy.array_equal(numpy.array(t.linear_output[:]), plain[k]), f"scene {k} differs"


def _spheres_and_cylinders(device):
    s = scenes.four_cylinders(device)
    fresnel.geometry.Sphere(s, position=[[0.5, 0.5, 0.5], [-1, 0.2, 0.3], [0.3, -1, -0.4]], radius=0.35)
    return s


def test_batches_in_flight_equal_one_batch(gpu_device):
    """Many small batches (two in flight, spheres-only and mixed groups) give the one-batch images bit for bit."""
    d = gpu_device._device
    saved = d.tune('batch_paths', 0)
    builders = [lambda dv: c1_scene(dv, n=3000, w_seed=5), _spheres_and_cylinders]
    sizes = [(64, 48), (40, 56)]

    def render():
        sc = [b(gpu_device) for b in builders]
        trs = [fresnel.tracer.Path(gpu_device, w=w, h=h) for (w, h) in sizes]
        for t in trs:
            t.reset()
            t._tracer.setLightSamples(1)
        fresnel._gpu.batch_render([x._scene for x in sc], [t._tracer for t in trs], 9)
        alone = [_linear(gpu_device, b, w, h, 9, 1)[0] for b, (w, h) in zip(builders, sizes)]
        return [numpy.array(t.linear_output[:]) for t in trs], alone

    try:
        d.tune('batch_paths', 1 << 26)
        one_batch, one_alone = render()
        d.tune('batch_paths', 2000)  # one camera sample per batch: 9 batches of two groups
        many, many_alone = render()
    finally:
        d.tune('batch_paths', saved)
    for k in range(len(builders)):
        assert numpy.array_equal(many[k], one_batch[k]), f"scene {k} differs (batched)"
        assert numpy.array_equal(many_alone[k], one_alone[k]), f"scene {k} differs (alone)"
        assert numpy.array_equal(one_alone[k], one_batch[k]), f"scene {k}: batched differs from alone"


def test_batches_of_whole_pairs_equal_alone(gpu_device):
    """Pairs that take several batches go in batches of whole pairs, each batch starting once its own scenes
    are built while the device stream builds the next pairs' scenes: every image equals its render alone,
    bit for bit (spheres-only and mixed scenes, uneven sizes, one scene shared by two tracers)."""
    d = gpu_device._device
    saved = d.tune('batch_paths', 0)
    shared = c1_scene(gpu_device, n=2500, w_seed=11)
    builders = [lambda dv: c1_scene(dv, n=1500, w_seed=9), _spheres_and_cylinders, lambda dv: shared,
                scenes.eight_polyhedra, lambda dv: shared, lambda dv: c1_scene(dv, n=4000, w_seed=10)]
    sizes = [(64, 48), (40, 56), (48, 48), (56, 40), (64, 32), (72, 40)]
    S = 6
    try:
        d.tune('batch_paths', 40000)  # about two pairs' 6 samples per batch: three or four batches
        sc = [b(gpu_device) for b in builders]
        trs = [fresnel.tracer.Path(gpu_device, w=w, h=h) for (w, h) in sizes]
        for t in trs:
            t.reset()
            t._tracer.setLightSamples(1)
        fresnel._gpu.batch_render([x._scene for x in sc], [t._tracer for t in trs], S)
        batched = [numpy.array(t.linear_output[:]) for t in trs]
    finally:
        d.tune('batch_paths', saved)
    for k, (b, (w, h)) in enumerate(zip(builders, sizes)):
        alone = _linear(gpu_device, b, w, h, S, 1)[0]
        assert numpy.array_equal(alone, batched[k]), f"pair {k} differs from its render alone"


def test_graph_replays_equal_eager(gpu_device):
    """Batches captured into HIP graphs and replayed (tuning key 'graph' = 2) give the eager images bit for
    bit on every replay, with eager renders and synchronizations in between; several batches per render on
    both streams, spheres-only and mixed scenes."""
    d = gpu_device._device
    saved = {k: d.tune(k, 0) for k in ('graph', 'batch_paths')}
    try:
        d.tune('batch_paths', 20000)  # several batches per render
        for b in (lambda dv: c1_scene(dv, n=3000, w_seed=7), _spheres_and_cylinders):
            d.tune('graph', 1)
            ref = _linear(gpu_device, b, 64, 48, 6, 2)[0]
            c0, r0 = d.graphStats()
            for k in range(4):
                d.tune('graph', 2)
                img = _linear(gpu_device, b, 64, 48, 6, 2)[0]
                assert numpy.array_equal(img, ref), f"replay {k} differs"
                d.tune('graph', 1)
                assert numpy.array_equal(_linear(gpu_device, b, 64, 48, 6, 2)[0], ref)
            c1, r1 = d.graphStats()
            assert c1 > c0 and r1 > r0, (c0, r0, c1, r1)
    finally:
        for k, v in saved.items():
            d.tune(k, v)


def test_random_rays_vs_checker(gpu_device, cpu_device):
    """Closest hits of random rays (many starting inside spheres) are bit-exact."""
    rng = numpy.random.default_rng(3)

    def build(dev):
        s = fresnel.Scene(dev)
        pos = rng_pos
        fresnel.geometry.Sphere(s, position=pos, radius=rad)
        return s

    rng_pos = rng.uniform(-5, 5, (2000, 3)).astype(numpy.float32)
    rad = rng.uniform(0.05, 0.6, 2000).astype(numpy.float32)
    gs, cs = build(gpu_device), build(cpu_device)
    n = 20000
    o = rng.uniform(-6, 6, (n, 3)).astype(numpy.float32)
    d = rng.normal(size=(n, 3))
    d = (d / numpy.linalg.norm(d, axis=1, keepdims=True)).astype(numpy.float32)
    rays = numpy.ascontiguousarray(numpy.concatenate([o, d], axis=1), dtype=numpy.float32)
    gid, prim, t = gs._scene.debugTrace(rays, 1e-3)
    rg = numpy.zeros(n, numpy.uint32)
    rp = numpy.zeros(n, numpy.uint32)
    rt = numpy.zeros(n, numpy.float32)
    oracle.lib().oracle_trace(cs._scene.handle(), ptr(rays), n, 1e-3, ptr(rg, ctypes.c_uint32),
                              ptr(rp, ctypes.c_uint32), ptr(rt), None, 0)
    assert numpy.array_equal(gid, rg)
    assert numpy.array_equal(prim, rp)
    assert numpy.array_equal(t.view(numpy.uint32), rt.view(numpy.uint32))


@pytest.mark.parametrize('builder', [lambda d: c1_scene(d, n=20000, w_seed=4), scenes.eight_polyhedra],
                         ids=['spheres', 'polyhedra'])
def test_tuning_knobs_do_not_change_image(builder, gpu_device):
    """Launch-shape knobs (LDS stack depth, batch size, tail threshold) leave the image bit-identical.

    A 4096-path batch forces many batches per sample and an 8-entry LDS stack
    forces spills to the global overflow stack on the 20,000-sphere BVH.
    """
    d = gpu_device._device
    keys = ('stack_lds', 'batch_paths', 'tail_rays', 'trav_chunk', 'trav_refill', 'trav_packet', 'primary_tiles',
            'primary_chunk', 'fuse_primary', 'fuse_bounce', 'inflight', 'xcd_bands', 'graph')
    saved = {k: d.tune(k, 0) for k in keys}
    settings = [dict(stack_lds=16, batch_paths=1 << 25, tail_rays=32768, trav_chunk=256, trav_refill=65,
                     trav_packet=1, primary_tiles=1, fuse_primary=1, fuse_bounce=1, xcd_bands=1),
                dict(stack_lds=8, batch_paths=4096, tail_rays=1, trav_chunk=64, trav_refill=1, trav_packet=2,
                     primary_tiles=2, primary_chunk=64),
                dict(stack_lds=8, batch_paths=1 << 20, tail_rays=1 << 20, trav_chunk=4096, trav_refill=65,
                     trav_packet=2, primary_tiles=1, fuse_primary=2, primary_chunk=256, xcd_bands=2),
                dict(stack_lds=8, batch_paths=2000, tail_rays=64, trav_chunk=64, trav_refill=65,
                     trav_packet=2, primary_tiles=2, fuse_primary=2, primary_chunk=256, xcd_bands=1),
                dict(stack_lds=16, batch_paths=10000, tail_rays=100, trav_chunk=192, trav_refill=64,
                     trav_packet=2, primary_tiles=2, fuse_bounce=1, inflight=1),
                dict(stack_lds=16, batch_paths=1 << 26, tail_rays=100, trav_chunk=64, trav_refill=17,
                     trav_packet=1, primary_tiles=2),
                dict(stack_lds=8, batch_paths=3000, tail_rays=200, trav_chunk=256, trav_refill=65, trav_packet=2,
                     primary_tiles=2, fuse_primary=2, fuse_bounce=1, inflight=2, graph=2),
                dict(stack_lds=8, batch_paths=1 << 20, tail_rays=1, trav_chunk=128, trav_refill=65, trav_packet=2,
                     primary_tiles=1, fuse_primary=1, fuse_bounce=2, inflight=1, graph=2)]
    images = []
    try:
        for st in settings:
            for k, v in st.items():
                d.tune(k, v)
            images.append(_linear(gpu_device, builder, 72, 40, 3, 2)[0])
    finally:
        for k, v in saved.items():
            d.tune(k, v)
    for k, img in enumerate(images[1:], 1):
        assert numpy.array_equal(img, images[0]), f"setting {settings[k]} changed the image"


def test_tune_rejects_bad_keys(gpu_device):
    d = gpu_device._device
    with pytest.raises(Exception, match='stack_lds'):
        d.tune('stack_lds', 12)
    with pytest.raises(Exception, match='trav_chunk'):
        d.tune('trav_chunk', 100)
    with pytest.raises(Exception, match='primary_chunk'):
        d.tune('primary_chunk', 65)
    with pytest.raises(Exception, match='fuse_primary'):
        d.tune('fuse_primary', 3)
    with pytest.raises(Exception, match='trav_refill'):
        d.tune('trav_refill', 66)
    with pytest.raises(Exception, match='unknown tuning key'):
        d.tune('warp_size', 32)


def test_served_requests_equal_plain_sample(gpu_device):
    """Requests advanced together by the render server equal Path.sample on each alone, bit for bit."""
    from fresnel_amd import serve
    builders = [lambda d: c1_scene(d, n=3000, w_seed=5), scenes.four_cylinders, scenes.eight_polyhedra,
                lambda d: c1_scene(d, n=800, w_seed=6)]
    specs = [((64, 48), 7, 1), ((80, 40), 3, 2), ((50, 50), 12, 1), ((96, 64), 1, 3)]
    plain = [_linear(gpu_device, b, w, h, S, L)[0] for b, ((w, h), S, L) in zip(builders, specs)]
    sc = [b(gpu_device) for b in builders]
    trs = [fresnel.tracer.Path(gpu_device, w=w, h=h) for ((w, h), _, _) in specs]
    # a small per-step target splits requests over several steps with changing mixes
    srv = serve.Server(gpu_device, hbm_budget=1 << 34, target_paths=9000)
    reqs = [srv.submit(s, t, S, L) for s, t, (_, S, L) in zip(sc, trs, specs)]
    steps = 0
    while not all(r.done for r in reqs):
        srv.step()
        steps += 1
        assert steps < 100
    assert steps > 1
    for k, (r, t) in enumerate(zip(reqs, trs)):
        st, n, _ = r.poll()
        assert st == serve.DONE and n == specs[k][1]
        assert t._tracer.getNumSamples() == specs[k][1]
        assert numpy.array_equal(numpy.array(t.linear_output[:]), plain[k]), f"request {k} differs"
    assert srv.memory()[0] == 0


def test_path_sample_between_server_steps(gpu_device):
    """Path.sample on the same device between server steps (it renders on the streams the server's batches
    occupy) leaves both its own image and the requests' images equal to their renders alone."""
    from fresnel_amd import serve
    builders = [lambda d: c1_scene(d, n=2000, w_seed=12), scenes.four_cylinders, lambda d: c1_scene(d, n=600, w_seed=13)]
    specs = [((64, 48), 9, 1), ((80, 40), 5, 2), ((56, 56), 6, 1)]
    plain = [_linear(gpu_device, b, w, h, S, L)[0] for b, ((w, h), S, L) in zip(builders, specs)]
    extra = lambda d: c1_scene(d, n=1500, w_seed=14)
    extra_alone = _linear(gpu_device, extra, 72, 40, 4, 1)[0]
    sc = [b(gpu_device) for b in builders]
    trs = [fresnel.tracer.Path(gpu_device, w=w, h=h) for ((w, h), _, _) in specs]
    srv = serve.Server(gpu_device, hbm_budget=1 << 34, target_paths=6000)
    reqs = [srv.submit(s, t, S, L) for s, t, (_, S, L) in zip(sc, trs, specs)]
    srv.step()
    srv.step()  # batches in flight on the device's render streams
    between = _linear(gpu_device, extra, 72, 40, 4, 1)[0]
    steps = 2
    while not all(r.done for r in reqs):
        srv.step()
        steps += 1
        assert steps < 100
    assert numpy.array_equal(between, extra_alone), "Path.sample between server steps differs"
    for k, (r, t) in enumerate(zip(reqs, trs)):
        assert numpy.array_equal(numpy.array(t.linear_output[:]), plain[k]), f"request {k} differs"


def test_server_rejects_over_budget(gpu_device):
    from fresnel_amd import serve
    s = c1_scene(gpu_device, n=100)
    t = fresnel.tracer.Path(gpu_device, w=32, h=32)
    srv = serve.Server(gpu_device, hbm_budget=1024, target_paths=1 << 20)
    r = srv.submit(s, t, 4)
    srv.step()
    assert r.state == serve.REJECTED
    srv.release(r)


def test_grazing_rays_vs_brute_force(gpu_device):
    """Distant grazing hits whose computed t lies below their box entry: the GPU's culled
    traversal returns the brute-force closest hit, bit for bit."""
    from test_oracle import grazing_case, _sphere_scene, _trace
    pos, rad, rays, _ = grazing_case()
    s = fresnel.Scene(gpu_device)
    fresnel.geometry.Sphere(s, position=pos, radius=rad)
    gid, prim, t = s._scene.debugTrace(rays, 1e-3)
    h = _sphere_scene(pos, rad, brute=True)
    _, pf, tf, _ = _trace(h, rays)
    oracle.lib().oracle_scene_free(h)
    assert numpy.array_equal(prim, pf)
    assert numpy.array_equal(t.view(numpy.uint32), tf.view(numpy.uint32))


def _sincos_hard_args():
    """The x column of the device's table of hard-to-round sine/cosine arguments."""
    import pathlib
    import re
    src = (pathlib.Path(__file__).resolve().parent.parent / 'fresnel_amd' / 'csrc' / 'rt' / 'sincos_hard.h').read_text()
    rows = re.findall(r'\{(0x[0-9a-fp.+-]+)f, (0x[0-9a-fp.+-]+)f, (0x[0-9a-fp.+-]+)f\}', src)
    return numpy.array([float.fromhex(r[0]) for r in rows], numpy.float32)


def test_shading_sincos_vs_checker(gpu_device):
    """The device's sinf/cosf of lobe and aperture angles equal the checker's (correctly rounded) bit for bit:
    on the tabulated hard-to-round arguments and their neighbours, on angles 2 pi u01(u) of random words u,
    and on both ends of the range."""
    hard = _sincos_hard_args()
    assert len(hard) > 50
    nb = numpy.concatenate([numpy.nextafter(hard, numpy.float32(0)), numpy.nextafter(hard, numpy.float32(9))])
    rng = numpy.random.default_rng(17)
    u = rng.integers(0, 2**32, 2000000, dtype=numpy.uint64).astype(numpy.uint32)
    u01 = numpy.zeros(len(u), numpy.float32)
    oracle.lib().oracle_u01(ptr(u, ctypes.c_uint32), len(u), ptr(u01))
    ang = (numpy.float32(2.0) * numpy.float32(numpy.pi) * u01).astype(numpy.float32)
    ends = numpy.array([0.0, 1e-30, 1e-7, 0.5, 1.0, numpy.pi / 2, numpy.pi, 2 * numpy.pi, 6.2831855, 8.0],
                       numpy.float32)
    x = numpy.ascontiguousarray(numpy.concatenate([hard, nb, ang, ends]), numpy.float32)
    s, c = gpu_device._device.debugSincos(x)
    rs = numpy.zeros_like(x)
    rc = numpy.zeros_like(x)
    oracle.lib().oracle_cr_sincosf(ptr(x), len(x), ptr(rs), ptr(rc))
    assert numpy.array_equal(s.view(numpy.uint32), rs.view(numpy.uint32)), int((s != rs).sum())
    assert numpy.array_equal(c.view(numpy.uint32), rc.view(numpy.uint32)), int((c != rc).sum())
