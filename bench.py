"""Benchmark: path-traced frames per second on one MI355X, as Mpaths/s.

Default workload (the large particle scene, one GPU):
  1,048,576 spheres on a 64^3 FCC lattice (lattice constant 1.5) with Gaussian
  jitter sigma 0.05, radii uniform in [0.4, 0.6], per-particle colors from
  uniform random HSV (S, V in [0.5, 1]) used through primitive_color_mix = 1,
  roughness 0.3, specular 0.5, metal 0, outline width 0.05, light.cloudy(),
  perspective camera fitted to the lattice; 1920x1080, 256 samples,
  1 light sample.

A step renders one frame: Path.sample(scene, samples=256) on an HBM-resident
scene (the BVH is built during warmup); the image stays on the device (the
reference's ImageArray maps on construction; here a read maps, once, outside
the timed region).  Mpaths/s = W*H*samples*light_samples per second.  The
JSON line also carries:
  roofline         the kernel with the most time per frame (the
                   secondary-ray traversal or the fused primary kernel):
                   algorithmic bytes per launch (128 B per 4-wide node visit
                   of a spheres-only hierarchy, 64 B per binary node visit,
                   prim_bytes per primitive tested, the ray's own bytes, from
                   the instrumented work counters) over the average launch
                   duration (HIP events on the stream it runs on), against
                   8 TB/s HBM; beside it the same launches' model FP32 flops
                   against 157.3 TFLOP/s, the HBM bytes and VALU busy /
                   active-lane fractions the committed rocprofv3 counters
                   measured (profiles/traffic_<workload>.json), and the
                   resource those counters name as binding.
  rmse_vs_checker  the same scene at 480x270, 4 samples, rendered by the GPU
                   and by the CPU checker (oracle/): linear RMSE and the
                   fraction of bit-identical pixels.
  cpu_baseline     the CPU checker (a scalar restatement of the reference's
                   CPU path, oracle/) on the same scene: full-resolution passes
                   on host threads for about 10 s.
  batched          the trajectory workload (below), batched vs plain.
  served           the mixed request stream (below), served vs plain, at
                   several arrival rates and as a burst, with p50/p95/p99.
  served_small     a stream of small requests, served vs plain.

Other workloads (--workload), each printing one JSON line of the same form:
  c1  4096 spheres, 256x256, samples 16, light_samples 8 (plain).
  c3  mixed geometry: 200k cuboid polyhedra, 2M spherocylinders joining
      nearest-neighbour pairs, 50k icosphere mesh instances; 1920x1080,
      samples 128 (plain).
  c4  trajectory: 64 frames of 262,144 Brownian spheres, 1280x720, samples 32:
      64 plain renders (rebuild, then refit) vs one batched submission of all
      frames; images compared bit for bit (--modes plain_rebuild, plain_refit or
      batched runs fewer of the three, for a kernel trace of one).
  c5  served stream: Poisson arrivals of sphere/cylinder scenes with
      log-uniform sizes, mixed resolutions and sample counts: one request at a
      time (plain) vs the render server; Mpaths/s and p50/p95/p99 latency per
      arrival rate; images compared bit for bit.  --small takes the small-request
      stream instead (10^3-10^5 primitives, 256^2 or 512^2, 8 samples; 32 Mi-path
      steps); --modes plain or --modes served runs one mode alone (for a kernel
      trace of it), without the comparison.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c1|c2|c3|c4|c5] [--no-legs]
"""

import argparse
import gc
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
FP32_PEAK_TFLOPS = 157.3
# Model FP32 operations of the traversal: a child box slab test (3 axes x
# (2 subtractions + 2 multiplications) + 4 min/max + the cull bound) ~20, so
# 80 per 4-wide node visit (4 children) and 40 per binary node visit (2);
# a sphere test (centre offset, two dot products, discriminant, root) ~20; a
# ray's setup (3 reciprocals, 3 multiplications) ~6.
FLOPS_PER_NODE4, FLOPS_PER_NODE2, FLOPS_PER_SPHERE, FLOPS_PER_RAY = 80, 40, 20, 6


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


def c2_arrays():
    """Input arrays of the 1,048,576-sphere lattice: positions, radii, colours."""
    pos = fcc_lattice()
    n = len(pos)
    rng = numpy.random.default_rng(1)
    radius = rng.uniform(0.4, 0.6, n).astype(numpy.float32)
    return dict(position=pos, radius=radius, color=hsv_colors(n, seed=2))


def c2_scene(device, a=None):
    import fresnel_amd as fresnel
    a = a if a is not None else c2_arrays()
    scene = fresnel.Scene(device, lights=fresnel.light.cloudy())
    fresnel.geometry.Sphere(scene, position=a['position'], radius=a['radius'], color=a['color'], outline_width=0.05,
                            material=fresnel.material.Material(primitive_color_mix=1.0, roughness=0.3, specular=0.5,
                                                               metal=0.0))
    scene.camera = perspective_fit(scene)
    return scene


def c1_arrays():
    """Input array of the 4096-sphere scene: positions."""
    return dict(position=numpy.random.default_rng(0).uniform(-10, 10, (4096, 3)).astype(numpy.float32))


def c1_scene(device, a=None):
    import fresnel_amd as fresnel
    a = a if a is not None else c1_arrays()
    scene = fresnel.Scene(device, lights=fresnel.light.cloudy())
    fresnel.geometry.Sphere(scene, position=a['position'], radius=0.5, outline_width=0.1,
                            material=fresnel.material.Material(color=fresnel.color.linear([0.1, 0.8, 0.1]),
                                                               roughness=0.8, specular=0.2))
    scene.camera = perspective_fit(scene)
    return scene


def input_digests(arrays):
    """SHA-256 of each generated input array (its float32/uint32 bytes, C order)."""
    return {k: hashlib.sha256(numpy.ascontiguousarray(v).tobytes()).hexdigest() for k, v in sorted(arrays.items())}


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


def c3_arrays(n_poly=200000, n_cyl=2000000, n_mesh=50000, side=120.0, seed=3):
    """Input arrays of the mixed scene, drawn in this order from one generator: polyhedron positions and
    orientations and colours, spherocylinder end points, mesh instance positions; and the mesh itself."""
    rng = numpy.random.default_rng(seed)
    a = dict(poly_position=rng.uniform(-side / 2, side / 2, (n_poly, 3)))
    a['poly_orientation'] = random_quaternions(n_poly, rng)
    a['poly_color'] = hsv_colors(n_poly, seed=4)
    a['cyl_points'] = nn_segments(n_cyl // 2, side, rng)
    a['mesh_position'] = rng.uniform(-side / 2, side / 2, (n_mesh, 3))
    a['mesh_triangles'] = icosphere(0.4, 2)
    return a


def c3_scene(device, a=None, n_poly=200000, n_cyl=2000000, n_mesh=50000, side=120.0, seed=3):
    import fresnel_amd as fresnel
    a = a if a is not None else c3_arrays(n_poly, n_cyl, n_mesh, side, seed)
    scene = fresnel.Scene(device, lights=fresnel.light.cloudy())
    cube = numpy.array([[x, y, z] for x in (-0.5, 0.5) for y in (-0.5, 0.5) for z in (-0.25, 0.25)])
    info = fresnel.util.convex_polyhedron_from_vertices(cube)
    fresnel.geometry.ConvexPolyhedron(scene, info, position=a['poly_position'], orientation=a['poly_orientation'],
                                      color=a['poly_color'],
                                      material=fresnel.material.Material(primitive_color_mix=1.0, roughness=0.1,
                                                                         metal=0.0), outline_width=0.03)
    fresnel.geometry.Cylinder(scene, points=a['cyl_points'], radius=0.1, color=[0.8, 0.7, 0.3],
                              material=fresnel.material.Material(color=fresnel.color.linear([0.8, 0.7, 0.3]),
                                                                 roughness=0.5, metal=1.0))
    fresnel.geometry.Mesh(scene, a['mesh_triangles'], position=a['mesh_position'], color=[0.3, 0.4, 0.9],
                          material=fresnel.material.Material(color=fresnel.color.linear([0.3, 0.4, 0.9]),
                                                             roughness=0.9, metal=0.0))
    scene.camera = perspective_fit(scene)
    return scene


WORKLOADS = {
    'c2': dict(build=c2_scene, arrays=c2_arrays, W=1920, H=1080, samples=256, light_samples=1, prim_bytes=16, node_bytes=128,
               name="1,048,576 spheres, 64^3 FCC lattice a=1.5 jitter 0.05, r~U[0.4,0.6], HSV colors, "
               "roughness 0.3 specular 0.5, outline 0.05, light.cloudy(), perspective fit; "
               "1920x1080, samples=256, light_samples=1"),
    'c1': dict(build=c1_scene, arrays=c1_arrays, W=256, H=256, samples=16, light_samples=8, prim_bytes=16, node_bytes=128,
               name="4096 spheres uniform in a cube of side 20 (seed 0), r=0.5, outline 0.1, "
               "color linear(0.1,0.8,0.1) roughness 0.8 specular 0.2, light.cloudy(), perspective fit; "
               "256x256, samples=16, light_samples=8"),
    'c3': dict(build=c3_scene, arrays=c3_arrays, W=1920, H=1080, samples=128, light_samples=1, prim_bytes=28, node_bytes=64,
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
    host_cores = os.cpu_count() or 1
    if threads is None:
        threads = min(16, host_cores)
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
        "host_cores": host_cores,
        "cpu_model": cpu_model(),
        "kind": "port",
        "sample": f"{passes} full-resolution pass(es) ({wl['W']}x{wl['H']}, light_samples="
                  f"{wl['light_samples']}) of the same scene in {dt:.1f} s; CPU checker (scalar restatement "
                  f"of the reference CPU path, own BVH) on {threads} threads ({threads} of {host_cores} host "
                  f"cores, {cpu_model()})",
    }


def cpu_model():
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def rmse_vs_checker(wl, W=480, H=270, samples=4, seed=9):
    """The workload's scene rendered at W x H, `samples` samples by the GPU and by the CPU checker
    (same camera, seed and light samples): linear RMSE and the fraction of bit-identical pixels."""
    import fresnel_amd as fresnel
    import oracle.backend
    fresnel._register_cpu_module(oracle.backend)
    imgs = {}
    for mode in ('gpu', 'cpu'):
        dev = fresnel.Device(mode=mode)
        tr = fresnel.tracer.Path(dev, w=W, h=H)
        tr.seed = seed
        tr.sample(wl['build'](dev), samples=samples, light_samples=wl['light_samples'], reset=False)
        imgs[mode] = numpy.array(tr.linear_output[:])
    g, c = imgs['gpu'], imgs['cpu']
    return {"size": f"{W}x{H}", "samples": samples, "light_samples": wl['light_samples'],
            "rmse": float(numpy.sqrt(numpy.mean((g - c)**2))),
            "bit_identical_pixels": float(numpy.mean(numpy.all(g.view(numpy.uint32) == c.view(numpy.uint32), axis=2))),
            "bar": "linear RMSE <= 1e-3 at equal samples"}


def run_plain(args):
    import fresnel_amd as fresnel
    wl = WORKLOADS[args.workload]
    W, H, S, L = wl['W'], wl['H'], wl['samples'], wl['light_samples']
    dev = fresnel.Device(mode='gpu', n=1)
    t_build = time.perf_counter()
    arrays = wl['arrays']()
    digests = input_digests(arrays)
    scene = wl['build'](dev, arrays)
    del arrays
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

    node_bytes = wl.get('node_bytes', 64)
    flops_node = FLOPS_PER_NODE4 if node_bytes == 128 else FLOPS_PER_NODE2

    def kernel_roofline(name, rays, nodes, prims, ray_bytes, what):
        """Algorithmic bytes per launch (node_bytes per hierarchy node visited, prim_bytes per
        primitive tested, ray_bytes per ray) over the kernel's mean launch time, against HBM; the
        model FP32 flops of the same launches against the FP32 peak; the binding resource named
        from the committed counters."""
        ms = sum(st.get(name, (0.0, 0))[0] for st in kstats) / len(kstats)
        launches = max(1, kstats[-1].get(name, (0.0, 1))[1])
        alg = node_bytes * nodes + pb * prims + ray_bytes * rays
        achieved = alg / (ms / 1e3) / 1e9 if ms > 0 else 0.0
        flops = flops_node * nodes + FLOPS_PER_SPHERE * prims + FLOPS_PER_RAY * rays
        tflops = flops / (ms / 1e3) / 1e12 if ms > 0 else 0.0
        t = traffic.get(name)
        measured_gbs = (t / (ms / launches / 1e3) / 1e9) if (t and ms > 0) else None
        busy = valu_busy.get(name)
        lanes = valu_lanes.get(name)
        if measured_gbs is not None and measured_gbs > 0.6 * HBM_PEAK_GBS:
            binding = "HBM bandwidth"
        elif busy is not None and busy >= 90:
            binding = (f"VALU instruction issue (VALU busy {busy:.0f}%, {lanes:.0f}% of lanes active; "
                       f"HBM traffic {measured_gbs / HBM_PEAK_GBS:.0%} of peak)" if measured_gbs else
                       f"VALU instruction issue (VALU busy {busy:.0f}%)")
        elif busy is not None:
            binding = (f"latency of dependent loads and divergence (VALU busy {busy:.0f}%, {lanes:.0f}% of lanes "
                       f"active, HBM traffic {measured_gbs / HBM_PEAK_GBS:.0%} of peak: the hierarchy is served "
                       f"from L2 and the Infinity Cache)" if measured_gbs else
                       f"latency of dependent loads and divergence (VALU busy {busy:.0f}%)")
        else:
            binding = None
        return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": t,
                "kernel": what, "alg_bytes_per_launch": int(alg / launches), "node_bytes": node_bytes,
                "launches_per_step": int(launches), "avg_launch_ms": round(ms / launches, 3),
                "rays_per_launch": int(rays / launches), "nodes_per_ray": round(nodes / max(rays, 1), 2),
                "prims_per_ray": round(prims / max(rays, 1), 2), "ms_per_step": round(ms, 3),
                "valu": {"achieved": round(tflops, 2), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": round(tflops / FP32_PEAK_TFLOPS, 4),
                         "flops_per_launch": int(flops / launches),
                         "model": f"{flops_node} per node visit, {FLOPS_PER_SPHERE} per primitive test, "
                                  f"{FLOPS_PER_RAY} per ray"},
                "measured_hbm_GBs": round(measured_gbs, 1) if measured_gbs else None,
                # the vector ALU side from the committed rocprofv3 counters
                # (VALUBusy, VALUUtilization)
                "valu_busy_frac": (round(busy / 100, 3) if busy is not None else None),
                "valu_active_lanes_frac": (round(lanes / 100, 3) if lanes is not None else None),
                "binding": binding}

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
    path_bytes = (node_bytes * (ctr[1] + ctr[4] + ctr[7]) + pb * (ctr[2] + ctr[5] + ctr[8])
                  + 36 * W * H * S)
    path_flops = (flops_node * (ctr[1] + ctr[4] + ctr[7]) + FLOPS_PER_SPHERE * (ctr[2] + ctr[5] + ctr[8])
                  + FLOPS_PER_RAY * seg_rays)
    kernel_ms = {k: round(v[0], 3) for k, v in kstats[-1].items()}
    kernel_launches = {k: int(v[1]) for k, v in kstats[-1].items()}

    out = base_record(round(value, 3), "Mpaths/s", args.steps, args.warmup, round(mean_s * 1e3, 2),
                      {"workload": wl['name'], "mode": "plain", "paths_per_frame": paths,
                       "rays_per_frame": int(seg_rays), "mrays_per_s": round(seg_rays / mean_s / 1e6, 2),
                       "inputs_sha256": digests})
    out.update({
        "latency_ms": {"p50": round(float(numpy.percentile(times, 50)) * 1e3, 2),
                       "p99": round(float(numpy.percentile(times, 99)) * 1e3, 2)},
        "roofline": roof[dominant],
        "roofline_other": {k: v for k, v in roof.items() if k != dominant},
        "path_roofline": {"alg_bytes_per_path": round(path_bytes / paths, 1),
                          "achieved_GBs": round(path_bytes / mean_s / 1e9, 1),
                          "frac": round(path_bytes / mean_s / 1e9 / HBM_PEAK_GBS, 4),
                          "model_flops_per_path": round(path_flops / paths, 1),
                          "achieved_TFLOPs": round(path_flops / mean_s / 1e12, 2),
                          "valu_frac": round(path_flops / mean_s / 1e12 / FP32_PEAK_TFLOPS, 4)},
        "tail": {"rays": int(ctr[6]), "nodes_per_ray": round(ctr[7] / max(ctr[6], 1), 2)},
        "primary_tile_entries_per_ray": round(ctr[10] / max(ctr[3], 1), 2),
        "primary_tile_entries_by_type": {
            "spherocylinder": round(ctr[11] / max(ctr[3], 1), 2), "polyhedron": round(ctr[12] / max(ctr[3], 1), 2),
            "mesh_triangle": round(ctr[13] / max(ctr[3], 1), 2),
            "other": round((ctr[10] - ctr[11] - ctr[12] - ctr[13]) / max(ctr[3], 1), 2)},
        "kernel_ms_last_step": kernel_ms,
        "kernel_launches_last_step": kernel_launches,
        "tuning": tuned,
        "capped_paths": int(ctr[9]),
    })
    if not args.no_cpu_baseline:
        out["rmse_vs_checker"] = rmse_vs_checker(wl)
        log(f"rmse vs checker: {out['rmse_vs_checker']}")
        out["cpu_baseline"] = cpu_baseline(wl, seconds=args.cpu_seconds)
    return out


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


def trajectory_leg(frames_n, modes=('plain_rebuild', 'plain_refit', 'batched')):
    """`frames_n` frames of the trajectory: plain renders (rebuild / refit) vs one batched submission of every
    frame, images compared bit for bit.  Returns (summary, seconds per mode).  `modes` may name fewer of the
    three (a kernel trace of one); ratios and the comparison need plain_rebuild and batched."""
    import fresnel_amd as fresnel
    W, H, S, F = 1280, 720, 32, frames_n
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
        if name not in modes:
            continue
        dt, dig = fn()
        results[name] = (dt, dig)
        log(f"c4 {name}: {dt:.3f} s, {paths / dt / 1e6:.1f} Mpaths/s")
    ref = next(iter(results.values()))[1]
    identical = all(v[1] == ref for v in results.values()) if len(results) > 1 else None
    both = 'plain_rebuild' in results and 'batched' in results
    summary = {
        "workload": f"{F} frames of 262,144 spheres (r=0.5, uniform in a cube of side 80 at frame 0, Brownian "
                    f"steps sigma 0.02 per axis per frame, seed 4), 1280x720, samples=32, light_samples=1",
        "paths": paths,
        "modes": {k: {"seconds": round(v[0], 4), "mpaths_per_s": round(paths / v[0] / 1e6, 2)}
                  for k, v in results.items()},
        "batched_vs_plain": round(results['plain_rebuild'][0] / results['batched'][0], 3) if both else None,
        "batched_vs_plain_refit": (round(results['plain_refit'][0] / results['batched'][0], 3)
                                   if 'plain_refit' in results and 'batched' in results else None),
        "bit_identical": identical,
    }
    return summary, {k: v[0] for k, v in results.items()}


def run_trajectory(args):
    """64 frames: 64 plain renders (rebuild / refit) vs one batched submission of every frame."""
    modes = tuple(m for m in args.modes.split(',') if m in ('plain_rebuild', 'plain_refit', 'batched')) or \
        ('plain_rebuild', 'plain_refit', 'batched')
    summary, secs = trajectory_leg(args.frames, modes)
    mode = 'batched' if 'batched' in secs else next(iter(secs))
    dt_b = secs[mode]
    out = base_record(round(summary['paths'] / dt_b / 1e6, 3), "Mpaths/s", 1, 1, round(dt_b * 1e3, 2), {
        "workload": summary['workload'],
        "mode": "batched (all frames in one submission)" if mode == 'batched' else mode, "paths": summary['paths']})
    out.update({k: v for k, v in summary.items() if k not in ('workload', 'paths')})
    print(json.dumps(out), flush=True)
    if summary['bit_identical'] is False:
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


def small_requests(n_req, seed=7):
    """Small requests: spheres or spherocylinders, N log-uniform in [1e3, 1e5], 256x256 or 512x512,
    8 samples (interactive previews), from the same generators."""
    rng = numpy.random.default_rng(seed)
    reqs = []
    for i in range(n_req):
        n = int(round(math.exp(rng.uniform(math.log(1e3), math.log(1e5)))))
        kind = 'sphere' if rng.uniform() < 0.5 else 'cylinder'
        w = int(rng.choice([256, 512]))
        reqs.append(dict(kind=kind, n=n, w=w, h=w, samples=8, seed=2000 + i))
    return reqs


def served_leg(specs, rates, target_paths, describe, budget=240 << 30, modes=('plain', 'served')):
    """Poisson streams of `specs` at each rate (req/s; 0 = a burst: every request at time 0): one request at a
    time as it arrives (plain, Path.sample) vs the render server.  Returns the summary with Mpaths/s and
    p50/p95/p99 latency (arrival to complete image) per rate, images compared bit for bit.  `modes` may
    name one of the two alone (a kernel trace of one mode); the comparison then is left out."""
    import fresnel_amd as fresnel
    from fresnel_amd import serve
    dev = fresnel.Device(mode='gpu', n=1)
    t0 = time.perf_counter()
    arrays = [request_arrays(sp) for sp in specs]
    log(f"served leg: {len(specs)} requests generated in {time.perf_counter() - t0:.1f} s")
    total_paths = sum(sp['w'] * sp['h'] * sp['samples'] for sp in specs)
    gaps = numpy.random.default_rng(6).exponential(1.0, len(specs))

    def fresh():
        scenes = [request_scene(dev, sp, ar) for sp, ar in zip(specs, arrays)]
        tracers = [fresnel.tracer.Path(dev, w=sp['w'], h=sp['h']) for sp in specs]
        return scenes, tracers

    def wait_until(clock0, t):
        while time.perf_counter() - clock0 < t:
            time.sleep(min(0.002, max(0.0, t - (time.perf_counter() - clock0))))

    def run_plain_stream(arrive):
        scenes, tracers = fresh()
        done = numpy.zeros(len(specs))
        c0 = time.perf_counter()
        for i in numpy.argsort(arrive, kind='stable'):
            wait_until(c0, arrive[i])
            tracers[i].sample(scenes[i], samples=specs[i]['samples'], light_samples=1)
            done[i] = time.perf_counter() - c0
        return done, [image_digest(t) for t in tracers]

    steps = []  # (start, duration, active requests after it) of each server step of the last served stream

    def run_served_stream(arrive):
        scenes, tracers = fresh()
        srv = serve.Server(dev, hbm_budget=budget, target_paths=target_paths)
        order = numpy.argsort(arrive, kind='stable')
        handles = [None] * len(specs)
        done = numpy.full(len(specs), numpy.nan)
        nxt = 0
        steps.clear()
        c0 = time.perf_counter()
        while numpy.isnan(done).any():
            now = time.perf_counter() - c0
            while nxt < len(order) and arrive[order[nxt]] <= now:
                i = order[nxt]
                handles[i] = srv.submit(scenes[i], tracers[i], specs[i]['samples'], 1)
                nxt += 1
            a, q = srv.step()
            steps.append((now, time.perf_counter() - c0 - now, a))
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
        if lam > 0:
            arrive = numpy.cumsum(gaps) / lam
            arrive -= arrive[0]
        else:
            arrive = numpy.zeros(len(specs))
        row = {"lambda_per_s": lam if lam > 0 else "burst"}
        digests = {}
        for name, fn in (('plain', run_plain_stream), ('served', run_served_stream)):
            if name not in modes:
                continue
            # the request loop's latencies should not include Python's collector:
            # collect before the stream, none during it (both modes alike)
            gc.collect()
            gc.disable()
            try:
                done, dig = fn(arrive)
            finally:
                gc.enable()
            digests[name] = dig
            lat = done - arrive
            span = done.max()
            row[name] = {"mpaths_per_s": round(total_paths / span / 1e6, 2), "makespan_s": round(span, 3),
                         "latency_s": {p: round(float(numpy.percentile(lat, q)), 4)
                                       for p, q in (("p50", 50), ("p95", 95), ("p99", 99))}}
            if name == 'served' and steps:
                dur = numpy.array([d for _, d, _ in steps])
                k = int(dur.argmax())
                row[name]["step_ms"] = {"n": len(steps), "p50": round(float(numpy.median(dur)) * 1e3, 2),
                                        "max": round(float(dur[k]) * 1e3, 2), "max_at_s": round(steps[k][0], 3)}
                row[name]["max_active_requests"] = int(max(a for _, _, a in steps))
            log(f"served leg lambda={lam} {name}: {row[name]}")
        if len(digests) == 2:
            same = digests['served'] == digests['plain']
            identical = identical and same
            row["served_vs_plain"] = round(row['served']['mpaths_per_s'] / row['plain']['mpaths_per_s'], 3)
            row["bit_identical"] = same
        table.append(row)
    both = all("served_vs_plain" in r for r in table)
    return {"workload": describe, "requests": len(specs), "total_paths": total_paths,
            "target_paths_per_step": target_paths, "hbm_budget_bytes": budget, "by_rate": table,
            "best_served_vs_plain": max(r["served_vs_plain"] for r in table) if both else None,
            "bit_identical": identical if both else None}


C5_DESCRIPTION = ("spheres (FCC generator) or spherocylinders (nearest-neighbour generator), N log-uniform in "
                  "[1e3, 2e6], resolution uniform over {256^2, 1280x720, 1920x1080, 3840x2160}, samples uniform "
                  "over {8, 64, 256}, light_samples 1; Poisson arrivals; 240 GiB budget")
SMALL_DESCRIPTION = ("small requests: spheres or spherocylinders, N log-uniform in [1e3, 1e5], 256^2 or 512^2, "
                     "8 samples, light_samples 1; Poisson arrivals; 240 GiB budget")


def run_served(args):
    """Poisson request streams: plain (one request at a time) vs the render server."""
    rates = [float(x) for x in args.rates.split(',')]
    modes = tuple(args.modes.split(','))
    if args.small:
        summary = served_leg(small_requests(args.requests), rates, args.target_paths or 32 << 20,
                             f"{args.requests} requests: " + SMALL_DESCRIPTION, modes=modes)
    else:
        summary = served_leg(make_requests(args.requests), rates, args.target_paths or 256 << 20,
                             f"{args.requests} requests: " + C5_DESCRIPTION, modes=modes)
    rows = [r for r in summary['by_rate'] if r['lambda_per_s'] != 'burst'] or summary['by_rate']
    top = rows[-1]
    mode = 'served' if 'served' in modes else 'plain'
    out = base_record(top[mode]['mpaths_per_s'], "Mpaths/s", len(rates), 0,
                      round(top[mode]['makespan_s'] * 1e3, 2), {
                          "workload": summary['workload'], "mode": mode, "total_paths": summary['total_paths'],
                          "target_paths_per_step": summary['target_paths_per_step']})
    out["latency_s"] = top[mode]['latency_s']
    out["by_rate"] = summary['by_rate']
    out["bit_identical"] = summary['bit_identical']
    print(json.dumps(out), flush=True)
    if summary['bit_identical'] is False:
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
    ap.add_argument('--no-legs', action='store_true', help='c2: skip the batched and served legs')
    ap.add_argument('--copy-calibration', type=int, default=0, metavar='MIB',
                    help='first run a plain device copy of MIB MiB (the reference kernel for HBM byte counters)')
    ap.add_argument('--frames', type=int, default=64, help='c4: trajectory frames')
    ap.add_argument('--requests', type=int, default=24, help='c5: requests per stream')
    ap.add_argument('--rates', default='5,10,20,50,100,200,0', help='c5: arrival rates (req/s; 0 = burst)')
    ap.add_argument('--target-paths', type=int, default=0,
                    help='c5: paths per server step (default 256 Mi; 32 Mi with --small)')
    ap.add_argument('--small', action='store_true',
                    help='c5: the small-request stream (10^3-10^5 primitives, 256^2 or 512^2, 8 samples)')
    ap.add_argument('--modes', default='plain,served',
                    help='c5: plain, served or both; c4: any of plain_rebuild, plain_refit, batched (comma-separated)')
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
        out = run_plain(args)
        if args.workload == 'c2' and not args.no_legs:
            # the other two forms of the metric, each against its plain form
            t0 = time.perf_counter()
            out["batched"] = trajectory_leg(64)[0]
            out["served"] = served_leg(make_requests(24), [5.0, 50.0, 200.0, 0.0], 256 << 20,
                                       "24 requests: " + C5_DESCRIPTION)
            out["served_small"] = served_leg(small_requests(96), [200.0, 0.0], 32 << 20,
                                             "96 requests: " + SMALL_DESCRIPTION)
            # a budget that binds: the server's two scratches take 86 of its 98.8 GB, so the
            # requests' 17.7 GB cannot all be active at once, and the device's Path.sample
            # scratch is released
            out["served_budget"] = served_leg(make_requests(24), [0.0], 256 << 20,
                                              "24 requests: " + C5_DESCRIPTION.replace("240 GiB", "92 GiB"),
                                              budget=92 << 30)
            log(f"batched and served legs: {time.perf_counter() - t0:.1f} s")
        print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
