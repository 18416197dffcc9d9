"""Benchmark: path-traced frames per second on one MI355X, as Mpaths/s.

Default workload (the large particle scene, one GPU):
  1,048,576 spheres on a 64^3 FCC lattice (lattice constant 1.5) with Gaussian
  jitter sigma 0.05, radii uniform in [0.4, 0.6], per-particle colors from
  uniform random HSV (S, V in [0.5, 1]) used through primitive_color_mix = 1,
  roughness 0.3, specular 0.5, metal 0, outline width 0.05, light.cloudy(),
  perspective camera fitted to the lattice; 1920x1080, 256 samples,
  1 light sample.

A step renders one frame: Path.sample(scene, samples=256) on an HBM-resident
scene (the BVH is built during warmup).  Mpaths/s = W*H*samples*light_samples
per second.  The JSON line also carries:
  roofline      the kernel with the most time per frame (the fused
                primary-ray kernel or the secondary-ray traversal):
                algorithmic bytes per launch (64 B per BVH node visited +
                prim_bytes per primitive tested + the ray's own bytes, from
                the instrumented traversal counters) over the average launch
                duration (HIP events on the stream it runs on), against 8 TB/s
                HBM; with its VALU busy and active-lane fractions from the
                committed rocprofv3 counters (profiles/traffic_<workload>.json).
  cpu_baseline  the CPU checker (a scalar restatement of the reference's CPU
                path, oracle/) on the same scene, full-resolution passes on
                host threads for about 10 s.

Other workloads (--workload), each printing one JSON line of the same form:
  c1  4096 spheres, 256x256, samples 16, light_samples 8 (plain).
  c3  mixed geometry: 200k cuboid polyhedra, 2M spherocylinders joining
      nearest-neighbour pairs, 50k icosphere mesh instances; 1920x1080,
      samples 128 (plain).
  c4  trajectory: 64 frames of 262,144 Brownian spheres, 1280x720, samples 32:
      64 plain renders (rebuild, then refit) vs one batched submission of all
      frames; images compared bit for bit.
  c5  served stream: Poisson arrivals of sphere/cylinder scenes with
      log-uniform sizes, mixed resolutions and sample counts: one request at a
      time (plain) vs the render server; Mpaths/s and p50/p95/p99 latency per
      arrival rate; images compared bit for bit.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c1|c2|c3|c4|c5]
"""

import argparse
import hashlib
import json
import math
import os
import sys
import time

import numpy

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "Mpaths/s + p50/p99 frame latency, served/batched vs plain, at fixed image RMSE"
HBM_PEAK_GBS = 8000.0


# ---------------------------------------------------------------- workloads
def fcc_lattice(cells=64, a=1.5, jitter=0.05, seed=0):
    """Positions of a cells^3 FCC lattice (4 per cell) centred on the origin, with Gaussian jitter."""
    rng = numpy.random.default_rng(seed)
    g = numpy.arange(cells, dtype=numpy.float64)
    cx, cy, cz = numpy.meshgrid(g, g, g, indexing='ij')
    corners = numpy.stack([cx.ravel(), cy.ravel(), cz.ravel()], axis=1)
    basis = numpy.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    pos = (corners[:, None, :] + basis[None, :, :]).reshape(-1, 3) * a
    pos -= cells * a / 2
    pos += rng.normal(0.0, jitter, pos.shape)
    return pos.astype(numpy.float32)


def hsv_colors(n, seed=0):
    """Linear colors from uniform random HSV with S, V in [0.5, 1]."""
    import fresnel_amd as fresnel
    rng = numpy.random.default_rng(seed)
    h = rng.uniform(0, 1, n)
    s = rng.uniform(0.5, 1, n)
    v = rng.uniform(0.5, 1, n)
    # vectorised HSV -> sRGB
    i = numpy.floor(h * 6).astype(int) % 6
    f = h * 6 - numpy.floor(h * 6)
    p, q, t = v * (1 - s), v * (1 - f * s), v * (1 - (1 - f) * s)
    r = numpy.choose(i, [v, q, p, p, t, v])
    g = numpy.choose(i, [t, v, v, q, p, p])
    b = numpy.choose(i, [p, p, t, v, v, q])
    return fresnel.color.linear(numpy.stack([r, g, b], axis=1).astype(numpy.float32))


def perspective_fit(scene):
    """Perspective camera on the isometric axis framing the scene.

    Centre, view axis, up vector and height come from
    camera.Orthographic.fit(scene, 'isometric', margin=0.05); the camera sits
    at the fitted distance, focuses on the centre and its vertical field of
    view spans the fitted height at the centre.
    """
    import fresnel_amd as fresnel
    ortho = fresnel.camera.Orthographic.fit(scene, view='isometric', margin=0.05)
    pos = numpy.asarray(ortho.position, dtype=numpy.float64)
    look = numpy.asarray(ortho.look_at, dtype=numpy.float64)
    dist = float(numpy.linalg.norm(pos - look))
    cam = fresnel.camera.Perspective(position=pos, look_at=look, up=ortho.up, focus_distance=dist, height=0.24)
    cam.vertical_field_of_view = 2 * math.atan(ortho.height / 2 / dist)
    return cam


def c2_scene(device):
    import fresnel_amd as fresnel
    pos = fcc_lattice()
    n = len(pos)
    rng = numpy.random.default_rng(1)
    radius = rng.uniform(0.4, 0.6, n).astype(numpy.float32)
    scene = fresnel.Scene(device, lights=fresnel.light.cloudy())
    fresnel.geometry.Sphere(scene, position=pos, radius=radius, color=hsv_colors(n, seed=2), outline_width=0.05,
                            material=fresnel.material.Material(primitive_color_mix=1.0, roughness=0.3, specular=0.5,
                                                               metal=0.0))
    scene.camera = perspective_fit(scene)
    return scene


def c1_scene(device):
    import fresnel_amd as fresnel
    pos = numpy.random.default_rng(0).uniform(-10, 10, (4096, 3)).astype(numpy.float32)
    scene = fresnel.Scene(device, lights=fresnel.light.cloudy())
    fresnel.geometry.Sphere(scene, position=pos, radius=0.5, outline_width=0.1,
                            material=fresnel.material.Material(color=fresnel.color.linear([0.1, 0.8, 0.1]),
                                                               roughness=0.8, specular=0.2))
    scene.camera = perspective_fit(scene)
    return scene


def icosphere(radius=0.4, subdivisions=2):
    """Triangles (3T, 3) of an icosahedron subdivided `subdivisions` times, on a sphere of `radius`."""
    t = (1 + 5 ** 0.5) / 2
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t), (t, 0, -1),
         (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    tris = numpy.array([[v[a], v[b], v[c]] for a, b, c in f], dtype=numpy.float64)
    tris /= numpy.linalg.norm(tris, axis=2, keepdims=True)
    for _ in range(subdivisions):
        a, b, c = tris[:, 0], tris[:, 1], tris[:, 2]
        ab, bc, ca = (a + b), (b + c), (c + a)
        ab /= numpy.linalg.norm(ab, axis=1, keepdims=True)
        bc /= numpy.linalg.norm(bc, axis=1, keepdims=True)
        ca /= numpy.linalg.norm(ca, axis=1, keepdims=True)
        tris = numpy.concatenate([numpy.stack(x, axis=1) for x in ((a, ab, ca), (ab, b, bc), (ca, bc, c), (ab, bc, ca))])
    return (tris.reshape(-1, 3) * radius).astype(numpy.float32)


def random_quaternions(n, rng):
    q = rng.normal(size=(n, 4))
    return (q / numpy.linalg.norm(q, axis=1, keepdims=True)).astype(numpy.float32)


def nn_segments(n_points, side, rng):
    """Segments from each of n_points uniform points to its 1st and 2nd nearest neighbours: (2 n_points, 2, 3)."""
    from scipy.spatial import cKDTree
    p = rng.uniform(-side / 2, side / 2, (n_points, 3))
    _, idx = cKDTree(p).query(p, k=3)
    a = numpy.repeat(p, 2, axis=0)
    b = p[idx[:, 1:3].reshape(-1)]
    return numpy.stack([a, b], axis=1).astype(numpy.float32)


def c3_scene(device, n_poly=200000, n_cyl=2000000, n_mesh=50000, side=120.0, seed=3):
    import fresnel_amd as fresnel
    rng = numpy.random.default_rng(seed)
    scene = fresnel.Scene(device, lights=fresnel.light.cloudy())
    cube = numpy.array([[x, y, z] for x in (-0.5, 0.5) for y in (-0.5, 0.5) for z in (-0.25, 0.25)])
    info = fresnel.util.convex_polyhedron_from_vertices(cube)
    fresnel.geometry.ConvexPolyhedron(scene, info, position=rng.uniform(-side / 2, side / 2, (n_poly, 3)),
                                      orientation=random_quaternions(n_poly, rng), color=hsv_colors(n_poly, seed=4),
                                      material=fresnel.material.Material(primitive_color_mix=1.0, roughness=0.1,
                                                                         metal=0.0), outline_width=0.03)
    seg = nn_segments(n_cyl // 2, side, rng)
    fresnel.geometry.Cylinder(scene, points=seg, radius=0.1, color=[0.8, 0.7, 0.3],
                              material=fresnel.material.Material(color=fresnel.color.linear([0.8, 0.7, 0.3]),
                                                                 roughness=0.5, metal=1.0))
    fresnel.geometry.Mesh(scene, icosphere(0.4, 2), position=rng.uniform(-side / 2, side / 2, (n_mesh, 3)),
                          color=[0.3, 0.4, 0.9],
                          material=fresnel.material.Material(color=fresnel.color.linear([0.3, 0.4, 0.9]),
                                                             roughness=0.9, metal=0.0))
    scene.camera = perspective_fit(scene)
    return scene


WORKLOADS = {
    'c2': dict(build=c2_scene, W=1920, H=1080, samples=256, light_samples=1, prim_bytes=16,
               name="1,048,576 spheres, 64^3 FCC lattice a=1.5 jitter 0.05, r~U[0.4,0.6], HSV colors, "
               "roughness 0.3 specular 0.5, outline 0.05, light.cloudy(), perspective fit; "
               "1920x1080, samples=256, light_samples=1"),
    'c1': dict(build=c1_scene, W=256, H=256, samples=16, light_samples=8, prim_bytes=16,
               name="4096 spheres uniform in a cube of side 20 (seed 0), r=0.5, outline 0.1, "
               "color linear(0.1,0.8,0.1) roughness 0.8 specular 0.2, light.cloudy(), perspective fit; "
               "256x256, samples=16, light_samples=8"),
    'c3': dict(build=c3_scene, W=1920, H=1080, samples=128, light_samples=1, prim_bytes=28,
               name="200k cuboid polyhedra 1x1x0.5 random orientations + 2M spherocylinders r=0.1 joining "
               "1st/2nd nearest neighbours of 1M points + 50k icosphere meshes (320 triangles, r=0.4), "
               "cube of side 120, materials rough 0.1/0.5/0.9 metal 0/1/0, light.cloudy(), perspective fit; "
               "1920x1080, samples=128, light_samples=1"),
}


def log(msg):
    print("# " + msg, file=sys.stderr, flush=True)


def base_record(metric_value, unit, steps, warmup, ms_per_step, config):
    return {
        "metric": METRIC,
        "value": metric_value,
        "unit": unit,
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": config,
    }


def image_digest(tracer):
    return hashlib.sha256(numpy.ascontiguousarray(numpy.array(tracer.linear_output[:])).tobytes()).hexdigest()


# ---------------------------------------------------------------- measurement
def cpu_baseline(wl, seconds=10.0, threads=None):
    """Time the CPU checker on the same scene: full-resolution passes for about `seconds`."""
    import fresnel_amd as fresnel
    import oracle.backend
    fresnel._register_cpu_module(oracle.backend)
    if threads is None:
        threads = min(16, os.cpu_count() or 1)
    dev = fresnel.Device(mode='cpu', n=threads)
    scene = wl['build'](dev)
    tr = fresnel.tracer.Path(dev, w=wl['W'], h=wl['H'])
    tr.reset()
    tr._tracer.setLightSamples(wl['light_samples'])
    scene._scene.handle()  # build the checker's BVH outside the timed region
    passes = 0
    t0 = time.perf_counter()
    while True:
        tr._tracer.render(scene._scene)
        passes += 1
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    paths = wl['W'] * wl['H'] * wl['light_samples'] * passes
    return {
        "value": round(paths / dt / 1e6, 4),
        "unit": "Mpaths/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{passes} full-resolution pass(es) ({wl['W']}x{wl['H']}, light_samples="
                  f"{wl['light_samples']}) of the same scene in {dt:.1f} s; CPU checker (scalar restatement "
                  f"of the reference CPU path, own BVH) on {threads} threads",
    }


def run_plain(args):
    import fresnel_amd as fresnel
    wl = WORKLOADS[args.workload]
    W, H, S, L = wl['W'], wl['H'], wl['samples'], wl['light_samples']
    dev = fresnel.Device(mode='gpu', n=1)
    t_build = time.perf_counter()
    scene = wl['build'](dev)
    tracer = fresnel.tracer.Path(dev, w=W, h=H)
    print(f"# scene built on host in {time.perf_counter() - t_build:.1f} s", file=sys.stderr, flush=True)
    d = dev._device
    if args.copy_calibration:
        ms = d.debugCopy(args.copy_calibration << 20)
        log(f"copy calibration: {args.copy_calibration} MiB read + written in {ms:.3f} ms "
            f"({2 * (args.copy_calibration << 20) / ms / 1e6:.0f} GB/s)")
    tuned = None
    if not args.no_tune:
        from fresnel_amd import autotune
        t0 = time.perf_counter()
        tuned = autotune.tune(dev, scene, W, H, samples=max(1, S // 2), light_samples=L, use_cache=False,
                              log=lambda m: print("# " + m, file=sys.stderr, flush=True))
        print(f"# autotune: {tuned} in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    d.statsEnable(True)

    for k in range(args.warmup):
        t0 = time.perf_counter()
        tracer.sample(scene, samples=S, light_samples=L)
        print(f"# warmup {k}: {time.perf_counter() - t0:.3f} s", file=sys.stderr, flush=True)

    times = []
    kstats = []
    for k in range(args.steps):
        t0 = time.perf_counter()
        tracer.sample(scene, samples=S, light_samples=L)
        dt = time.perf_counter() - t0
        stats = d.statsLast()
        times.append(dt)
        kstats.append(stats)
        log(f"step {k}: {dt:.3f} s, " + ", ".join(f"{n} {v[0]:.1f} ms/{v[1]}" for n, v in stats.items()))
    mean_s = sum(times) / len(times)
    paths = W * H * S * L
    value = paths / mean_s / 1e6

    # deterministic work counters from one instrumented frame
    d.countersEnable(True)
    tracer.sample(scene, samples=S, light_samples=L)
    ctr = d.countersLast()
    d.countersEnable(False)
    pb = wl['prim_bytes']
    traffic, valu_busy, valu_lanes = {}, {}, {}
    tf = os.path.join(ROOT, 'profiles', f'traffic_{args.workload}.json')
    if os.path.exists(tf):
        with open(tf) as f:
            prof = json.load(f)
        traffic = prof.get('hbm_bytes_per_launch', {})
        valu_busy = prof.get('valu_busy_pct', {})
        valu_lanes = prof.get('valu_active_lanes_pct', {})

    def kernel_roofline(name, rays, nodes, prims, ray_bytes, what):
        """Algorithmic bytes per launch (64 B per BVH node visited, prim_bytes per
        primitive tested, ray_bytes per ray) over the kernel's mean launch time."""
        ms = sum(st.get(name, (0.0, 0))[0] for st in kstats) / len(kstats)
        launches = max(1, kstats[-1].get(name, (0.0, 1))[1])
        alg = 64 * nodes + pb * prims + ray_bytes * rays
        achieved = alg / (ms / 1e3) / 1e9 if ms > 0 else 0.0
        return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic.get(name),
                "kernel": what, "alg_bytes_per_launch": int(alg / launches), "launches_per_step": int(launches),
                "avg_launch_ms": round(ms / launches, 3), "rays_per_launch": int(rays / launches),
                "nodes_per_ray": round(nodes / max(rays, 1), 2), "prims_per_ray": round(prims / max(rays, 1), 2),
                "ms_per_step": round(ms, 3),
                # the vector ALU side of the same kernel, from the committed
                # rocprofv3 counters (VALUBusy, VALUUtilization): a kernel far
                # below the HBM peak is bound by its FP32/integer instruction
                # issue instead when its VALU is busy most of the time
                "valu_busy_frac": (round(valu_busy[name] / 100, 3) if name in valu_busy else None),
                "valu_active_lanes_frac": (round(valu_lanes[name] / 100, 3) if name in valu_lanes else None)}

    # primary rays are generated in the kernel; the fused primary kernel
    # (k_primary) writes at least a 16 B record or queue entry per ray, the
    # unfused one a 16 B hit.  Secondary rays read 32 B of ray from the queue
    # and write a 16 B hit.
    primary_name = "primary" if "primary" in kstats[-1] else "traverse_primary"
    roof = {primary_name: kernel_roofline(primary_name, ctr[3], ctr[4], ctr[5], 16,
                                          "primary rays: tile lists or BVH packets, fused with depth-0 shading "
                                          "(k_primary)" if primary_name == "primary" else
                                          "primary-ray traversal (k_traverse_packet / k_traverse)"),
            "traverse": kernel_roofline("traverse", ctr[0], ctr[1], ctr[2], 48,
                                        "secondary-ray traversal (k_traverse)")}
    dominant = max(roof, key=lambda k: roof[k]["ms_per_step"])
    # whole path: every traced segment (primary, secondary, tail) plus 36 B of
    # accumulation per pixel-sample, over the frame time
    seg_rays = ctr[0] + ctr[3] + ctr[6]
    path_bytes = (64 * (ctr[1] + ctr[4] + ctr[7]) + pb * (ctr[2] + ctr[5] + ctr[8])
                  + 36 * W * H * S)
    kernel_ms = {k: round(v[0], 3) for k, v in kstats[-1].items()}
    kernel_launches = {k: int(v[1]) for k, v in kstats[-1].items()}

    out = base_record(round(value, 3), "Mpaths/s", args.steps, args.warmup, round(mean_s * 1e3, 2),
                      {"workload": wl['name'], "mode": "plain", "paths_per_frame": paths,
                       "rays_per_frame": int(seg_rays), "mrays_per_s": round(seg_rays / mean_s / 1e6, 2)})
    out.update({
        "latency_ms": {"p50": round(float(numpy.percentile(times, 50)) * 1e3, 2),
                       "p99": round(float(numpy.percentile(times, 99)) * 1e3, 2)},
        "roofline": roof[dominant],
        "roofline_other": {k: v for k, v in roof.items() if k != dominant},
        "path_roofline": {"alg_bytes_per_path": round(path_bytes / paths, 1),
                          "achieved_GBs": round(path_bytes / mean_s / 1e9, 1),
                          "frac": round(path_bytes / mean_s / 1e9 / HBM_PEAK_GBS, 4)},
        "tail": {"rays": int(ctr[6]), "nodes_per_ray": round(ctr[7] / max(ctr[6], 1), 2)},
        "primary_tile_entries_per_ray": round(ctr[10] / max(ctr[3], 1), 2),
        "kernel_ms_last_step": kernel_ms,
        "kernel_launches_last_step": kernel_launches,
        "tuning": tuned,
        "capped_paths": int(ctr[9]),
    })
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(wl, seconds=args.cpu_seconds)
    print(json.dumps(out), flush=True)


# ---------------------------------------------------------------- c4: trajectory
def trajectory_frames(n_frames=64, n=262144, side=80.0, sigma=0.02, seed=4):
    rng = numpy.random.default_rng(seed)
    pos = rng.uniform(-side / 2, side / 2, (n, 3)).astype(numpy.float32)
    frames = [pos]
    for _ in range(n_frames - 1):
        pos = (pos + rng.normal(0.0, sigma, pos.shape)).astype(numpy.float32)
        frames.append(pos)
    return frames


def trajectory_scene(device, pos, camera=None):
    import fresnel_amd as fresnel
    scene = fresnel.Scene(device, lights=fresnel.light.cloudy())
    g = fresnel.geometry.Sphere(scene, position=pos, radius=0.5, outline_width=0.05,
                                material=fresnel.material.Material(color=fresnel.color.linear([0.2, 0.5, 0.8]),
                                                                   roughness=0.4, specular=0.5))
    scene.camera = camera if camera is not None else perspective_fit(scene)
    return scene, g


def run_trajectory(args):
    """64 frames: 64 plain renders (rebuild / refit) vs one batched submission of every frame."""
    import fresnel_amd as fresnel
    W, H, S, F = 1280, 720, 32, args.frames
    dev = fresnel.Device(mode='gpu', n=1)
    frames = trajectory_frames(F)
    cam = trajectory_scene(dev, frames[0])[0].camera
    paths = F * W * H * S
    results = {}

    def plain(refit):
        scene, g = trajectory_scene(dev, frames[0], cam)
        if refit:
            scene.set_bvh_refit(True)
        # fresnel.pathtrace creates a fresh tracer per call; they are allocated
        # before the clock starts, as the batched form's are
        tracers = [fresnel.tracer.Path(dev, w=W, h=H) for _ in range(F)]
        t0 = time.perf_counter()
        for f in range(F):
            if f:
                g.position[:] = frames[f]
            tracers[f].sample(scene, samples=S, light_samples=1)
        dt = time.perf_counter() - t0
        return dt, [image_digest(t) for t in tracers]

    def batched():
        scenes = [trajectory_scene(dev, frames[f], cam)[0] for f in range(F)]
        tracers = [fresnel.tracer.Path(dev, w=W, h=H) for _ in range(F)]
        t0 = time.perf_counter()
        for t in tracers:
            t.reset()  # what Path.sample(reset=True) does before its passes
            t._tracer.setLightSamples(1)
        fresnel._gpu.batch_render([s._scene for s in scenes], [t._tracer for t in tracers], S)
        dt = time.perf_counter() - t0
        return dt, [image_digest(t) for t in tracers]

    # warm up: one small batched render of two frames
    s2 = [trajectory_scene(dev, frames[f], cam)[0] for f in range(2)]
    t2 = [fresnel.tracer.Path(dev, w=W, h=H) for _ in range(2)]
    fresnel._gpu.batch_render([s._scene for s in s2], [t._tracer for t in t2], 2)
    del s2, t2
    for name, fn in (('plain_rebuild', lambda: plain(False)), ('plain_refit', lambda: plain(True)),
                     ('batched', batched)):
        dt, dig = fn()
        results[name] = (dt, dig)
        log(f"c4 {name}: {dt:.3f} s, {paths / dt / 1e6:.1f} Mpaths/s")
    ref = results['plain_rebuild'][1]
    identical = {k: v[1] == ref for k, v in results.items()}
    dt_b = results['batched'][0]
    out = base_record(round(paths / dt_b / 1e6, 3), "Mpaths/s", 1, 1, round(dt_b * 1e3, 2), {
        "workload": f"{F} frames of 262,144 spheres (r=0.5, uniform in a cube of side 80 at frame 0, Brownian "
                    f"steps sigma 0.02 per axis per frame, seed 4), 1280x720, samples=32, light_samples=1",
        "mode": "batched (all frames in one submission)", "paths": paths})
    out["modes"] = {k: {"seconds": round(v[0], 4), "mpaths_per_s": round(paths / v[0] / 1e6, 2),
                        "bit_identical_to_plain_rebuild": identical[k]} for k, v in results.items()}
    out["speedup_batched_vs_plain_rebuild"] = round(results['plain_rebuild'][0] / dt_b, 3)
    out["frame_digest_0"] = ref[0][:16]
    print(json.dumps(out), flush=True)
    if not all(identical.values()):
        sys.exit(1)


# ---------------------------------------------------------------- c5: served stream
def make_requests(n_req, seed=5):
    """Request specs: kind, primitive count, resolution, samples (uniform/log-uniform draws)."""
    rng = numpy.random.default_rng(seed)
    res = [(256, 256), (1280, 720), (1920, 1080), (3840, 2160)]
    reqs = []
    for i in range(n_req):
        n = int(round(math.exp(rng.uniform(math.log(1e3), math.log(2e6)))))
        kind = 'sphere' if rng.uniform() < 0.5 else 'cylinder'
        w, h = res[rng.integers(0, 4)]
        s = int(rng.choice([8, 64, 256]))
        reqs.append(dict(kind=kind, n=n, w=w, h=h, samples=s, seed=1000 + i))
    return reqs


def request_arrays(spec):
    """Host arrays of a request's scene, from the same generators as the plain workloads."""
    rng = numpy.random.default_rng(spec['seed'])
    n = spec['n']
    if spec['kind'] == 'sphere':
        cells = int(math.ceil((n / 4) ** (1 / 3)))
        lat = fcc_lattice(cells=cells, seed=spec['seed'])
        pos = lat[rng.choice(len(lat), n, replace=False)]
        return dict(position=pos, radius=rng.uniform(0.4, 0.6, n).astype(numpy.float32),
                    color=hsv_colors(n, seed=spec['seed']))
    side = 120.0 * (n / 2e6) ** (1 / 3)
    seg = nn_segments(max(2, (n + 1) // 2), side, rng)[:n]
    return dict(points=seg)


def request_scene(device, spec, arrays):
    import fresnel_amd as fresnel
    scene = fresnel.Scene(device, lights=fresnel.light.cloudy())
    if spec['kind'] == 'sphere':
        fresnel.geometry.Sphere(scene, outline_width=0.05,
                                material=fresnel.material.Material(primitive_color_mix=1.0, roughness=0.3,
                                                                   specular=0.5, metal=0.0), **arrays)
    else:
        fresnel.geometry.Cylinder(scene, radius=0.1, color=[0.8, 0.7, 0.3],
                                  material=fresnel.material.Material(color=fresnel.color.linear([0.8, 0.7, 0.3]),
                                                                     roughness=0.5, metal=1.0), **arrays)
    scene.camera = perspective_fit(scene)
    return scene


def run_served(args):
    """Poisson request streams: plain (one request at a time) vs the render server."""
    import fresnel_amd as fresnel
    from fresnel_amd import serve
    dev = fresnel.Device(mode='gpu', n=1)
    specs = make_requests(args.requests)
    t0 = time.perf_counter()
    arrays = [request_arrays(sp) for sp in specs]
    log(f"c5: {len(specs)} requests generated in {time.perf_counter() - t0:.1f} s")
    total_paths = sum(sp['w'] * sp['h'] * sp['samples'] for sp in specs)
    rates = [float(x) for x in args.rates.split(',')]
    gaps = numpy.random.default_rng(6).exponential(1.0, len(specs))

    def fresh():
        scenes = [request_scene(dev, sp, ar) for sp, ar in zip(specs, arrays)]
        tracers = [fresnel.tracer.Path(dev, w=sp['w'], h=sp['h']) for sp in specs]
        return scenes, tracers

    def wait_until(clock0, t):
        while time.perf_counter() - clock0 < t:
            time.sleep(min(0.002, max(0.0, t - (time.perf_counter() - clock0))))

    def run_plain_stream(arrive, loop):
        scenes, tracers = fresh()
        done = numpy.zeros(len(specs))
        c0 = time.perf_counter()
        for i in numpy.argsort(arrive, kind='stable'):
            wait_until(c0, arrive[i])
            t, sp = tracers[i], specs[i]
            if loop:  # the reference's Path.sample: one render() per camera sample
                t.reset()
                t._tracer.setLightSamples(1)
                for _ in range(sp['samples']):
                    t._tracer.render(scenes[i]._scene)
            else:
                t.sample(scenes[i], samples=sp['samples'], light_samples=1)
            done[i] = time.perf_counter() - c0
        return done, [image_digest(t) for t in tracers]

    def run_served_stream(arrive):
        scenes, tracers = fresh()
        srv = serve.Server(dev, hbm_budget=240 << 30, target_paths=args.target_paths)
        order = numpy.argsort(arrive, kind='stable')
        handles = [None] * len(specs)
        done = numpy.full(len(specs), numpy.nan)
        nxt = 0
        c0 = time.perf_counter()
        while numpy.isnan(done).any():
            now = time.perf_counter() - c0
            while nxt < len(order) and arrive[order[nxt]] <= now:
                i = order[nxt]
                handles[i] = srv.submit(scenes[i], tracers[i], specs[i]['samples'], 1)
                nxt += 1
            a, q = srv.step()
            now = time.perf_counter() - c0
            for i, h in enumerate(handles):
                if h is not None and numpy.isnan(done[i]) and h.state >= serve.DONE:
                    done[i] = now
            if a == 0 and q == 0 and nxt < len(order):
                wait_until(c0, arrive[order[nxt]])
        return done, [image_digest(t) for t in tracers]

    table = []
    identical = True
    for lam in rates:
        arrive = numpy.cumsum(gaps) / lam
        arrive -= arrive[0]
        row = {"lambda_per_s": lam}
        modes = [('plain', lambda: run_plain_stream(arrive, False)), ('served', lambda: run_served_stream(arrive))]
        if not args.no_loop:
            modes.insert(0, ('plain_render_loop', lambda: run_plain_stream(arrive, True)))
        digests = {}
        for name, fn in modes:
            done, dig = fn()
            digests[name] = dig
            lat = done - arrive
            span = done.max()
            row[name] = {"mpaths_per_s": round(total_paths / span / 1e6, 2), "makespan_s": round(span, 3),
                         "latency_s": {p: round(float(numpy.percentile(lat, q)), 4)
                                       for p, q in (("p50", 50), ("p95", 95), ("p99", 99))}}
            log(f"c5 lambda={lam} {name}: {row[name]}")
        same = all(d == digests['plain'] for d in digests.values())
        identical = identical and same
        row["served_vs_plain"] = round(row['served']['mpaths_per_s'] / row['plain']['mpaths_per_s'], 3)
        if 'plain_render_loop' in row:
            row["served_vs_plain_render_loop"] = round(
                row['served']['mpaths_per_s'] / row['plain_render_loop']['mpaths_per_s'], 3)
        row["bit_identical"] = same
        table.append(row)
    top = table[-1]
    out = base_record(top['served']['mpaths_per_s'], "Mpaths/s", len(rates), 0,
                      round(top['served']['makespan_s'] * 1e3, 2), {
                          "workload": f"{len(specs)} requests: spheres (FCC generator) or spherocylinders "
                                      "(nearest-neighbour generator), N log-uniform in [1e3, 2e6], resolution "
                                      "uniform over {256^2, 1280x720, 1920x1080, 3840x2160}, samples uniform "
                                      "over {8, 64, 256}, light_samples 1; Poisson arrivals; 240 GB budget",
                          "mode": "served", "total_paths": total_paths,
                          "target_paths_per_step": args.target_paths})
    out["latency_s"] = top['served']['latency_s']
    out["by_rate"] = table
    out["bit_identical"] = identical
    print(json.dumps(out), flush=True)
    if not identical:
        sys.exit(1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--workload', default='c2', choices=sorted(WORKLOADS) + ['c4', 'c5'])
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--cpu-seconds', type=float, default=10.0)
    ap.add_argument('--no-tune', action='store_true', help='skip the launch-shape autotuner')
    ap.add_argument('--copy-calibration', type=int, default=0, metavar='MIB',
                    help='first run a plain device copy of MIB MiB (the reference kernel for HBM byte counters)')
    ap.add_argument('--frames', type=int, default=64, help='c4: trajectory frames')
    ap.add_argument('--requests', type=int, default=24, help='c5: requests per stream')
    ap.add_argument('--rates', default='5,10,20,50,100,200', help='c5: arrival rates (req/s)')
    ap.add_argument('--target-paths', type=int, default=128 << 20, help='c5: paths per server step')
    ap.add_argument('--no-loop', action='store_true', help='c5: skip the render()-per-sample plain form')
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}). "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.workload == 'c4':
        run_trajectory(args)
    elif args.workload == 'c5':
        run_served(args)
    else:
        run_plain(args)


if __name__ == '__main__':
    main()
