"""Launch-shape autotuner for the path tracer.

The tunable shapes never change per-path arithmetic, so images are identical
for every setting; only speed changes.  Knobs (see fh_device_tune):

* ``stack_lds``   traversal stack entries kept in LDS per thread (8 or 16):
                  fewer entries allow more wavefronts per CU, more entries
                  spill less to the global overflow stack.
* ``batch_paths`` light-sample paths per wavefront batch: larger batches
                  amortise host synchronisation and the per-batch tail.
* ``tail_rays``   live-ray count below which a batch finishes its remaining
                  paths in one kernel.
* ``trav_chunk``  rays a secondary-ray traversal wavefront claims per
                  work-counter atomic.
* ``primary_chunk`` the most primary rays a workgroup claims at once (a
                  batch with fewer than 16 such claims per CU claims less).
* ``trav_refill`` idle lanes of a traversal wavefront that trigger taking new
                  rays (65: trace 64 rays at a time).
* ``trav_packet`` primary rays traced as wavefront packets (2) or per ray (1).
* ``primary_tiles`` primary rays tested against per-tile candidate lists (2)
                  or searched in the BVH (1).  With lists on, a tracer builds
                  them once ``tile_min_rays`` primary rays (not tuned here) have
                  used one scene and camera.
* ``tile_group`` largest tile count of a hierarchy subtree listed as one
                  tile-list entry (scenes other than spheres-only; 1: one
                  entry per primitive).
* ``fuse_primary`` primary traversal and depth-0 shading in one kernel (2) or
                  two (1).
* ``fuse_bounce`` the same for each bounce of the secondary rays.
* ``inflight`` batches of one render in flight on the device's two streams.
* ``xcd_bands`` queue segments as image bands claimed per XCD (2) or not (1).

`tune` runs a coordinate descent over the candidates on a caller-provided
scene, timing short renders, and stores the winner in a JSON cache keyed by
GPU architecture and scene class (``$FRESNEL_AMD_TUNING`` or
``~/.cache/fresnel_amd/tuning.json``).
"""

import json
import os
import time

CANDIDATES = {
    'stack_lds': [8, 16],
    'batch_paths': [32 << 20, 64 << 20, 128 << 20, 256 << 20],
    'tail_rays': [8192, 32768, 131072, 524288, 1048576],
    'trav_chunk': [64, 256, 1024],
    'primary_chunk': [256, 1024, 4096],
    'trav_refill': [16, 32, 48, 65],
    'trav_packet': [1, 2],
    'primary_tiles': [1, 2],
    'tile_group': [1, 4, 9, 16],
    'fuse_primary': [1, 2],
    'fuse_bounce': [1, 2],
    'inflight': [1, 2],
    'xcd_bands': [1, 2],
}


MIN_GAIN = 0.02  # relative improvement a knob change must show to be kept


def cache_path():
    return os.environ.get('FRESNEL_AMD_TUNING',
                          os.path.join(os.path.expanduser('~'), '.cache', 'fresnel_amd', 'tuning.json'))


def scene_class(scene, w, h):
    """Key for the cache: geometry types, primitive count decade and image size."""
    types = sorted({type(g).__name__ for g in scene.geometry})
    n = 0
    for g in scene.geometry:
        try:
            n += len(g.position[:]) if hasattr(g, 'position') else len(g.points[:])
        except Exception:
            pass
    decade = len(str(max(n, 1))) - 1
    return f"{'+'.join(types)}|1e{decade}|{w}x{h}"


def arch(device):
    """GPU name of the device's first entry in its description ("[0]: <name>, <n> CUs")."""
    for line in device._device.describe().splitlines():
        if line.startswith('[0]:'):
            return line[4:].strip()
    return 'unknown'


def apply(device, params):
    """Set tuning knobs on a GPU device; returns the previous values."""
    return {k: device._device.tune(k, int(v)) for k, v in params.items()}


def current(device):
    return {k: device._device.tune(k, 0) for k in CANDIDATES}


def load(device, key):
    try:
        with open(cache_path()) as f:
            return json.load(f).get(arch(device), {}).get(key)
    except (OSError, ValueError):
        return None


def store(device, key, params):
    path = cache_path()
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        data = {}
    data.setdefault(arch(device), {})[key] = params
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, 'w') as f:
        json.dump(data, f, indent=1)


def tune(device, scene, w, h, samples=16, light_samples=1, repeats=3, use_cache=True, log=None):
    """Pick the fastest knob values for rendering `scene` at w x h; applies and returns them."""
    from . import tracer as _tracer
    key = scene_class(scene, w, h)
    if use_cache:
        cached = load(device, key)
        if cached:
            apply(device, cached)
            return dict(cached)
    best = current(device)
    t = _tracer.Path(device, w=w, h=h)

    def timed(params):
        apply(device, params)
        t.sample(scene, samples=samples, light_samples=light_samples)  # warm queues at this size
        dt = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            t.sample(scene, samples=samples, light_samples=light_samples)
            dt.append(time.perf_counter() - t0)
        return min(dt)

    best_t = timed(best)
    if log:
        log(f"autotune {key}: {best} -> {best_t * 1e3:.2f} ms (start)")
    for k, values in CANDIDATES.items():
        for v in values:
            if v == best[k]:
                continue
            trial = dict(best)
            trial[k] = v
            dt = timed(trial)
            if log:
                log(f"autotune {key}: {trial} -> {dt * 1e3:.2f} ms")
            # a change must win by more than the timing noise of short renders
            if dt < best_t * (1.0 - MIN_GAIN):
                best, best_t = trial, dt
    apply(device, best)
    store(device, key, best)
    return best
