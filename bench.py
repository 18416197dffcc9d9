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
  roofline      traversal kernel: algorithmic bytes per launch (64 B per BVH
                node visited + 16 B per sphere tested + 48 B of ray in / hit
                out per ray, from the instrumented traversal counters) over the
                average launch duration (HIP events on the render stream),
                against 8 TB/s HBM.
  cpu_baseline  the CPU checker (a scalar restatement of the reference's CPU
                path, oracle/) on the same scene, full-resolution passes on
                host threads for about 10 s.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2|c1]
"""

import argparse
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


WORKLOADS = {
    'c2': dict(build=c2_scene, W=1920, H=1080, samples=256, light_samples=1, prim_bytes=16,
               name="1,048,576 spheres, 64^3 FCC lattice a=1.5 jitter 0.05, r~U[0.4,0.6], HSV colors, "
               "roughness 0.3 specular 0.5, outline 0.05, light.cloudy(), perspective fit; "
               "1920x1080, samples=256, light_samples=1"),
    'c1': dict(build=c1_scene, W=256, H=256, samples=16, light_samples=8, prim_bytes=16,
               name="4096 spheres uniform in a cube of side 20 (seed 0), r=0.5, outline 0.1, "
               "color linear(0.1,0.8,0.1) roughness 0.8 specular 0.2, light.cloudy(), perspective fit; "
               "256x256, samples=16, light_samples=8"),
}


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--workload', default='c2', choices=sorted(WORKLOADS))
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--cpu-seconds', type=float, default=10.0)
    ap.add_argument('--no-tune', action='store_true', help='skip the launch-shape autotuner')
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}). "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import fresnel_amd as fresnel
    wl = WORKLOADS[args.workload]
    W, H, S, L = wl['W'], wl['H'], wl['samples'], wl['light_samples']
    dev = fresnel.Device(mode='gpu', n=1)
    t_build = time.perf_counter()
    scene = wl['build'](dev)
    tracer = fresnel.tracer.Path(dev, w=W, h=H)
    print(f"# scene built on host in {time.perf_counter() - t_build:.1f} s", file=sys.stderr, flush=True)
    d = dev._device
    tuned = None
    if not args.no_tune:
        from fresnel_amd import autotune
        t0 = time.perf_counter()
        tuned = autotune.tune(dev, scene, W, H, samples=max(1, S // 16), light_samples=L, use_cache=False,
                              log=lambda m: print("# " + m, file=sys.stderr, flush=True))
        print(f"# autotune: {tuned} in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    d.statsEnable(True)

    for k in range(args.warmup):
        t0 = time.perf_counter()
        tracer.sample(scene, samples=S, light_samples=L)
        print(f"# warmup {k}: {time.perf_counter() - t0:.3f} s", file=sys.stderr, flush=True)

    times = []
    trav_ms = []
    trav_launches = []
    stats = {}
    for k in range(args.steps):
        t0 = time.perf_counter()
        tracer.sample(scene, samples=S, light_samples=L)
        dt = time.perf_counter() - t0
        stats = d.statsLast()
        times.append(dt)
        trav_ms.append(stats.get('traverse', (0.0, 0))[0])
        trav_launches.append(stats.get('traverse', (0.0, 0))[1])
        print(f"# step {k}: {dt:.3f} s, traverse {trav_ms[-1]:.1f} ms in {trav_launches[-1]} launches",
              file=sys.stderr, flush=True)
    mean_s = sum(times) / len(times)
    paths = W * H * S * L
    value = paths / mean_s / 1e6

    # deterministic work counters from one instrumented frame
    d.countersEnable(True)
    tracer.sample(scene, samples=S, light_samples=L)
    ctr = d.countersLast()
    d.countersEnable(False)
    rays, nodes, prims = ctr[0], ctr[1], ctr[2]
    alg_bytes = 64 * nodes + wl['prim_bytes'] * prims + 48 * rays
    t_trav_s = (sum(trav_ms) / len(trav_ms)) / 1e3
    n_launch = max(1, trav_launches[-1])
    achieved = alg_bytes / t_trav_s / 1e9 if t_trav_s > 0 else 0.0
    traffic = None
    tf = os.path.join(ROOT, 'profiles', f'traffic_{args.workload}.json')
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get('hbm_bytes_per_launch')
    kernel_ms = {k: round(v[0] / len(times) if False else v[0], 3) for k, v in stats.items()}

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mean_s * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": wl['name'], "mode": "plain", "paths_per_frame": paths,
                   "rays_per_frame": int(rays), "mrays_per_s": round(rays / mean_s / 1e6, 2)},
        "latency_ms": {"p50": round(float(numpy.percentile(times, 50)) * 1e3, 2),
                       "p99": round(float(numpy.percentile(times, 99)) * 1e3, 2)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "traverse", "alg_bytes_per_launch": int(alg_bytes / n_launch),
                     "launches_per_step": int(n_launch),
                     "avg_launch_ms": round(t_trav_s * 1e3 / n_launch, 3),
                     "nodes_per_ray": round(nodes / max(rays, 1), 2),
                     "prims_per_ray": round(prims / max(rays, 1), 2)},
        "kernel_ms_last_step": kernel_ms,
        "tuning": tuned,
        "capped_paths": int(ctr[4]),
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(wl, seconds=args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
