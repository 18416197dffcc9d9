"""GPU parity on the benchmark workloads themselves, against the CPU checker.

The scenes come from bench.py's own generators (the 1,048,576-sphere FCC
lattice with HSV colours and a perspective fit, the 4096-sphere scene with a
perspective fit, the polyhedron / spherocylinder / mesh mix, the Brownian
trajectory, the mixed request stream), so the code paths the benchmark times
are the ones compared here: tile lists with the sphere-footprint bound, 4-wide
nodes with sphere leaves, XCD bands, deep stacks, 256 Mi-path batches.

Bars:
* primary hits of the renderer's own primary-ray kernel: (geometry, primitive,
  t) bit-exact with the checker; a hit may differ only where the checker finds
  an exact tie in t (counted and printed, as are all exact-tie pixels);
* linear HDR image at equal seed and samples: bit-identical to the checker's
  (and so within the RMSE bar of 1e-3, which is printed): the device's shading
  performs the reference's float operations in order, with correctly rounded
  division, square root, sine and cosine, as the checker does;
* batched, refit and served renders: bit-identical to plain renders.

The hit-distance distribution against an FP64 solve of the same sphere
(oracle_sphere_t64) is printed for the sphere scenes: the fraction of primary
hits within 1e-5 relative, the worst case and the distance at which it
occurs.  The reference's float solve (det = b^2 - |v|^2 + r^2,
fresnel/common/IntersectSphere.h:50-51) is kept for bit-exact hit ids, so
that distribution is measured, not asserted to be within 1e-5 everywhere.
"""

import ctypes
import gc

import numpy
import pytest

import bench
import fresnel_amd as fresnel
import oracle
from oracle import ptr

pytestmark = pytest.mark.gpu

MISS = 0xffffffff
U32 = ctypes.c_uint32


def _checker_rays(cs, W, H, seed, sample):
    rays = numpy.zeros((W * H, 6), dtype=numpy.float32)
    oracle.lib().oracle_camera_rays(cs._scene.handle(), W, H, seed, sample, 1, ptr(rays))
    return rays


def _checker_trace(cs, rays):
    n = len(rays)
    g = numpy.zeros(n, numpy.uint32)
    p = numpy.zeros(n, numpy.uint32)
    t = numpy.zeros(n, numpy.float32)
    oracle.lib().oracle_trace(cs._scene.handle(), ptr(rays), n, 1e-3, ptr(g, U32), ptr(p, U32), ptr(t), None, 0)
    return g, p, t


def _ties(cs, rays):
    n = len(rays)
    out = numpy.zeros(n, numpy.uint32)
    if n:
        oracle.lib().oracle_count_ties(cs._scene.handle(), ptr(numpy.ascontiguousarray(rays)), n, 1e-3,
                                       ptr(out, U32))
    return out


def primary_parity(name, gs, cs, W, H, seed, samples, spheres=None):
    """Renderer primary hits (debugPrimaryRender) vs the checker for camera samples `samples`.

    Asserts bit-exact (geometry, primitive, t) except exact ties; prints the
    tie counts and, for a sphere geometry given as (position, radius) arrays,
    the FP64 relative-error distribution of t.  Returns the stats.
    """
    tr = fresnel.tracer.Path(gs.device, w=W, h=H)
    tr.seed = seed
    stats = dict(rays=0, hits=0, tie_pixels=0, tie_mismatches=0, rel_max=0.0, within_1e5=0, t64_rays=0)
    worst = None
    for sample in samples:
        gid, prim, t = tr._tracer.debugPrimaryRender(gs._scene, sample)
        rays = _checker_rays(cs, W, H, seed, sample)
        rg, rp, rt = _checker_trace(cs, rays)
        hit = rg != MISS
        ties = _ties(cs, numpy.ascontiguousarray(rays[hit]))
        diff = (gid != rg) | (prim != rp)
        stats['rays'] += W * H
        stats['hits'] += int(hit.sum())
        stats['tie_pixels'] += int((ties >= 2).sum())
        if diff.any():
            tie_full = numpy.zeros(W * H, numpy.uint32)
            tie_full[hit] = ties
            bad = diff & (tie_full < 2)
            assert not bad.any(), f"{name}: {int(bad.sum())} primary hits differ without an exact tie " \
                                  f"(first at pixel {int(numpy.argmax(bad))})"
            stats['tie_mismatches'] += int(diff.sum())
        same = ~diff
        assert numpy.array_equal(t[same].view(numpy.uint32), rt[same].view(numpy.uint32)), \
            f"{name}: hit distances differ in {int((t[same].view(numpy.uint32) != rt[same].view(numpy.uint32)).sum())}"
        if spheres is not None:
            pos, rad = spheres
            sel = numpy.nonzero(hit & (rg == 0))[0]
            xyz = numpy.ascontiguousarray(pos[rp[sel]], numpy.float32)
            rr = numpy.ascontiguousarray(rad[rp[sel]], numpy.float32)
            sub = numpy.ascontiguousarray(rays[sel])
            t64 = numpy.zeros(len(sel), numpy.float64)
            oracle.lib().oracle_sphere_t64_batch(ptr(sub), ptr(xyz), ptr(rr), len(sel), 1e-3,
                                                 t64.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
            ok = numpy.isfinite(t64)
            rel = numpy.abs(rt[sel][ok].astype(numpy.float64) - t64[ok]) / t64[ok]
            stats['t64_rays'] += int(ok.sum())
            stats['within_1e5'] += int((rel <= 1e-5).sum())
            if len(rel) and rel.max() > stats['rel_max']:
                k = int(numpy.argmax(rel))
                dist = float(numpy.linalg.norm(xyz[ok][k] - sub[ok][k, :3]))
                stats['rel_max'] = float(rel.max())
                worst = dict(rel=float(rel.max()), t=float(rt[sel][ok][k]), center_distance=round(dist, 3),
                             radius=float(rr[ok][k]))
    msg = (f"PARITY {name}: {stats['rays']} primary rays, {stats['hits']} hits, bit-exact except "
           f"{stats['tie_mismatches']} exact-tie pixels (exact ties among all hits: {stats['tie_pixels']})")
    if spheres is not None and stats['t64_rays']:
        msg += (f"; FP64 t: {stats['within_1e5'] / stats['t64_rays']:.6f} of {stats['t64_rays']} hits within 1e-5 "
                f"relative, worst {stats['rel_max']:.3e} ({worst})")
    print(msg, flush=True)
    assert stats['hits'] > 0, f"{name}: nothing visible"
    return stats


def image_rmse(name, gs, cs, W, H, samples, L, seed=None):
    """Path.sample on both devices at the same seed and samples: linear images bit-identical."""
    out = {}
    for key, s in (('gpu', gs), ('cpu', cs)):
        tr = fresnel.tracer.Path(s.device, w=W, h=H)
        if seed is not None:
            tr.seed = seed
        tr.sample(s, samples=samples, light_samples=L, reset=seed is None)
        out[key] = numpy.array(tr.linear_output[:])
    g, c = out['gpu'], out['cpu']
    rmse = float(numpy.sqrt(numpy.mean((g - c)**2)))
    exact = float(numpy.mean(numpy.all(g == c, axis=2)))
    print(f"PARITY {name}: {W}x{H} S={samples} L={L} linear RMSE {rmse:.3e}, bit-identical pixels {exact:.4f}",
          flush=True)
    assert numpy.all(numpy.isfinite(g))
    assert rmse <= 1e-3
    assert numpy.array_equal(g.view(numpy.uint32), c.view(numpy.uint32)), f"{name}: linear image differs"
    return rmse


def _sphere_arrays(scene):
    g = scene.geometry[0]
    return (numpy.ascontiguousarray(numpy.array(g.position[:]), numpy.float32),
            numpy.ascontiguousarray(numpy.array(g.radius[:]), numpy.float32))


@pytest.fixture(scope='module')
def c2_pair(gpu_device, cpu_device):
    """The benchmark's 1,048,576-sphere scene on both devices (camera: the benchmark's perspective fit)."""
    return bench.c2_scene(gpu_device), bench.c2_scene(cpu_device)


@pytest.mark.parametrize('tiles', [2, 1], ids=['tile_lists', 'bvh_packets'])
def test_c2_primary_hits(c2_pair, tiles, gpu_device):
    """1M FCC spheres at 480x270 (the full-size frame's camera and view): primary hits bit-exact."""
    gs, cs = c2_pair
    d = gpu_device._device
    saved = d.tune('primary_tiles', 0)
    try:
        d.tune('primary_tiles', tiles)
        primary_parity(f"c2 1M spheres 480x270 ({'tile lists' if tiles == 2 else 'BVH packets'})", gs, cs, 480,
                       270, 77, (1, 2), spheres=_sphere_arrays(gs))
    finally:
        d.tune('primary_tiles', saved)


def test_c2_image_rmse(c2_pair):
    gs, cs = c2_pair
    image_rmse("c2 1M spheres", gs, cs, 480, 270, 4, 1, seed=9)


def test_c1_full_size(gpu_device, cpu_device):
    """The 4096-sphere scene at its benchmark size: 256x256, 16 samples, 8 light samples, perspective fit."""
    gs, cs = bench.c1_scene(gpu_device), bench.c1_scene(cpu_device)
    primary_parity("c1 4096 spheres 256x256", gs, cs, 256, 256, 5, (1, 16), spheres=_sphere_arrays(gs))
    image_rmse("c1 4096 spheres", gs, cs, 256, 256, 16, 8)


def _c3_small(device):
    # a tenth of every count in a cube holding a tenth of the volume (same densities)
    return bench.c3_scene(device, n_poly=20000, n_cyl=200000, n_mesh=5000, side=120.0 * 0.1**(1 / 3))


@pytest.fixture(scope='module')
def c3_pair(gpu_device, cpu_device):
    return _c3_small(gpu_device), _c3_small(cpu_device)


@pytest.mark.parametrize('group', [1, 4, 16], ids=['primitive_entries', 'subtrees_4_tiles', 'subtrees_16_tiles'])
def test_c3_scaled(c3_pair, gpu_device, group):
    """The mixed-geometry scene at a tenth of its counts (same densities): primary hits and image, with tile
    lists of primitives (tile_group 1) and of hierarchy subtrees covering at most 4 or 16 tiles."""
    gs, cs = c3_pair
    d = gpu_device._device
    saved = d.tune('tile_group', 0)
    try:
        d.tune('tile_group', group)
        primary_parity(f"c3 20k polyhedra + 200k spherocylinders + 5k meshes 384x216 (tile_group {group})", gs, cs,
                       384, 216, 3, (1, 2))
        if group <= 4:
            image_rmse(f"c3 scaled mix (tile_group {group})", gs, cs, 384, 216, 2, 1, seed=4)
    finally:
        d.tune('tile_group', saved)


def test_c4_scaled_batched_plain_refit(gpu_device):
    """8 trajectory frames of 262,144 Brownian spheres at 320x180, 4 samples: one batched submission,
    plain renders with rebuilds and plain renders with refits are bit-identical."""
    W, H, S, F = 320, 180, 4, 8
    frames = bench.trajectory_frames(F)
    cam = bench.trajectory_scene(gpu_device, frames[0])[0].camera

    def plain(refit):
        scene, g = bench.trajectory_scene(gpu_device, frames[0], cam)
        if refit:
            scene.set_bvh_refit(True)
        out = []
        for f in range(F):
            if f:
                g.position[:] = frames[f]
            t = fresnel.tracer.Path(gpu_device, w=W, h=H)
            t.sample(scene, samples=S, light_samples=1)
            out.append(numpy.array(t.linear_output[:]))
        return out

    rebuild = plain(False)
    refit = plain(True)
    scenes = [bench.trajectory_scene(gpu_device, frames[f], cam)[0] for f in range(F)]
    trs = [fresnel.tracer.Path(gpu_device, w=W, h=H) for _ in range(F)]
    for t in trs:
        t.reset()
        t._tracer.setLightSamples(1)
    fresnel._gpu.batch_render([s._scene for s in scenes], [t._tracer for t in trs], S)
    for f in range(F):
        b = numpy.array(trs[f].linear_output[:])
        assert numpy.array_equal(rebuild[f], b), f"frame {f}: batched differs from plain"
        assert numpy.array_equal(rebuild[f], refit[f]), f"frame {f}: refit differs from rebuild"
    print(f"PARITY c4 {F} frames x 262144 spheres {W}x{H} S={S}: batched == plain rebuild == plain refit",
          flush=True)


def test_c5_mixed_requests_served_equal_alone(gpu_device):
    """Eight requests of the served stream's generator (spheres or spherocylinders, N log-uniform in
    [1e3, 2e6], 256^2 to 4K, 8/64/256 samples): served together == Path.sample alone, bit for bit."""
    from fresnel_amd import serve
    specs = bench.make_requests(8, seed=5)
    arrays = [bench.request_arrays(sp) for sp in specs]
    alone = []
    for sp, ar in zip(specs, arrays):
        s = bench.request_scene(gpu_device, sp, ar)
        t = fresnel.tracer.Path(gpu_device, w=sp['w'], h=sp['h'])
        t.sample(s, samples=sp['samples'], light_samples=1)
        alone.append(numpy.array(t.linear_output[:]))
        del s, t
    scenes = [bench.request_scene(gpu_device, sp, ar) for sp, ar in zip(specs, arrays)]
    trs = [fresnel.tracer.Path(gpu_device, w=sp['w'], h=sp['h']) for sp in specs]
    srv = serve.Server(gpu_device, hbm_budget=240 << 30, target_paths=64 << 20)
    reqs = []
    steps = 0
    # staggered arrivals: later requests join while earlier ones are in flight
    for k in range(len(specs)):
        reqs.append(srv.submit(scenes[k], trs[k], specs[k]['samples'], 1))
        srv.step()
        steps += 1
    while not all(r.done for r in reqs):
        srv.step()
        steps += 1
        assert steps < 10000
    for k, (r, t) in enumerate(zip(reqs, trs)):
        st, n, _ = r.poll()
        assert st == serve.DONE and n == specs[k]['samples']
        assert numpy.array_equal(numpy.array(t.linear_output[:]), alone[k]), f"request {k} ({specs[k]}) differs"
    print(f"PARITY c5 8 mixed requests ({sum(s['w'] * s['h'] * s['samples'] for s in specs)} paths, {steps} server "
          f"steps): served == alone", flush=True)


def test_thin_lens_primary_rays(gpu_device, cpu_device):
    """Perspective camera with f_stop = 0.2 (aperture sample: Philox, cos, sin; fresnel/common/Camera.h:189-205):
    camera rays equal the checker's bit for bit (mismatches are counted and printed); hits are bit-exact
    except exact ties."""
    W, H = 256, 192

    def scene(dev):
        s = bench.c1_scene(dev)
        cam = s.camera
        cam.f_stop = 0.2
        s.camera = cam
        return s

    gs, cs = scene(gpu_device), scene(cpu_device)
    tr = fresnel.tracer.Path(gpu_device, w=W, h=H)
    tr.seed = 31
    mism = 0
    for sample in (1, 2, 3):
        rays, _, _, _ = tr._tracer.debugPrimary(gs._scene, sample)
        ref = _checker_rays(cs, W, H, 31, sample)
        mism += int(numpy.any(rays.view(numpy.uint32) != ref.view(numpy.uint32), axis=1).sum())
    print(f"PARITY thin lens f_stop=0.2 {W}x{H}: {mism} of {3 * W * H} camera rays differ from the checker",
          flush=True)
    # both evaluate the aperture's cosf/sinf correctly rounded
    assert mism == 0
    primary_parity("thin lens f_stop=0.2", gs, cs, W, H, 31, (1, 2, 3), spheres=_sphere_arrays(gs))


def test_tile_list_threshold(gpu_device):
    """The default policy: a tracer searches the BVH for primary rays until 4 Mi primary rays have used one
    scene and camera, then builds the per-tile lists and takes them; the images equal those of a tracer
    that takes the lists from its first render."""
    d = gpu_device._device
    saved = d.tune('tile_min_rays', 0)
    rng = numpy.random.default_rng(5)
    s = fresnel.Scene(gpu_device, lights=fresnel.light.cloudy())
    fresnel.geometry.Sphere(s, position=rng.uniform(-10, 10, (3000, 3)).astype(numpy.float32), radius=0.5,
                            outline_width=0.1)
    s.camera = fresnel.camera.Orthographic(position=(30, 30, 30), look_at=(0, 0, 0), up=(0, 1, 0), height=35)
    W, H, S = 256, 256, 24  # 1.5 Mi primary rays per sample() call

    def renders(tracer, seeds=None):
        out = []
        d.countersEnable(True)
        try:
            for k in range(3):
                if seeds is not None:
                    tracer.seed = seeds[k] - 1  # sample() resets, which advances the seed by one
                tracer.sample(s, samples=S, light_samples=1)
                out.append((tracer.seed, int(d.countersLast()[10]), numpy.array(tracer.linear_output[:])))
        finally:
            d.countersEnable(False)
        return out

    try:
        d.tune('tile_min_rays', 4 << 20)
        default = renders(fresnel.tracer.Path(gpu_device, w=W, h=H))
        d.tune('tile_min_rays', 1)
        lists = renders(fresnel.tracer.Path(gpu_device, w=W, h=H), seeds=[seed for seed, _, _ in default])
    finally:
        d.tune('tile_min_rays', saved)
    print("tile-list entries tested per sample() call: default", [e for _, e, _ in default], "lists from the start",
          [e for _, e, _ in lists], flush=True)
    # 1.5 Mi and 3 Mi primary rays: BVH search; 4.5 Mi: lists
    assert default[0][1] == 0 and default[1][1] == 0 and default[2][1] > 0
    assert all(e > 0 for _, e, _ in lists)
    for (_, _, a), (_, _, b) in zip(default, lists):
        assert numpy.array_equal(a, b)


def test_tile_lists_not_reused_across_scenes(gpu_device):
    """A Path tracer reused after its scene was destroyed renders a new, smaller scene with the same
    camera exactly as a fresh tracer does (per-tile lists are keyed on the scene, not its address)."""

    def scene(n, seed):
        rng = numpy.random.default_rng(seed)
        s = fresnel.Scene(gpu_device, lights=fresnel.light.cloudy())
        fresnel.geometry.Sphere(s, position=rng.uniform(-10, 10, (n, 3)).astype(numpy.float32), radius=0.5,
                                outline_width=0.1)
        return s

    a = scene(5000, 1)
    cam = fresnel.camera.Orthographic(position=(30, 30, 30), look_at=(0, 0, 0), up=(0, 1, 0), height=35)
    a.camera = cam
    tr = fresnel.tracer.Path(gpu_device, w=128, h=96)
    tr.sample(a, samples=2, light_samples=1)
    for n, seed in ((500, 2), (40, 3), (5000, 4)):
        del a
        gc.collect()
        a = scene(n, seed)
        a.camera = cam
        tr.sample(a, samples=2, light_samples=1)
        fresh = fresnel.tracer.Path(gpu_device, w=128, h=96)
        fresh.seed = tr.seed - 1  # sample(reset=True) advanced both seeds by one
        fresh.sample(a, samples=2, light_samples=1)
        assert numpy.array_equal(numpy.array(tr.linear_output[:]), numpy.array(fresh.linear_output[:])), n
