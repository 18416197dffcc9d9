"""Render-server policy (admission and sample allocation) without a GPU.

The policy lives in host code behind fh_policy_footprint / fh_policy_simulate;
these tests pin it with hand-worked schedules and invariants.
"""

import ctypes
import os

import numpy
import pytest

LIB = os.path.join(os.path.dirname(__file__), '..', 'fresnel_amd', 'libfresnel_hip.so')


@pytest.fixture(scope='module')
def lib():
    if not os.path.exists(LIB):
        pytest.skip('libfresnel_hip.so is not built')
    return ctypes.CDLL(LIB)


def footprint(lib, n_prims, prim_bytes, W, H, L):
    out = ctypes.c_uint64()
    assert lib.fh_policy_footprint(ctypes.c_uint64(n_prims), ctypes.c_uint64(prim_bytes), ctypes.c_uint64(W),
                                   ctypes.c_uint64(H), ctypes.c_uint64(L), ctypes.byref(out)) == 0
    return out.value


def simulate(lib, fp, pps, samples, arrive, budget, target, iters=10000):
    n = len(fp)
    u64 = lambda a: numpy.ascontiguousarray(a, dtype=numpy.uint64)
    u32 = lambda a: numpy.ascontiguousarray(a, dtype=numpy.uint32)
    fp, pps, samples, arrive = u64(fp), u64(pps), u32(samples), u32(arrive)
    adm = numpy.zeros(n, numpy.int32)
    done = numpy.zeros(n, numpy.int32)
    peak = ctypes.c_uint64()
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    assert lib.fh_policy_simulate(ctypes.c_uint32(n), P(fp), P(pps), P(samples), P(arrive), ctypes.c_uint64(budget),
                                  ctypes.c_uint64(target), ctypes.c_uint32(iters), P(adm), P(done),
                                  ctypes.byref(peak)) == 0
    return adm, done, peak.value


def test_footprint_formula(lib):
    # scene: 1M spheres x (32 + 16 + 64 + 128 + 4 + 8 + 4 + 16) bytes; outputs
    # 40 B/pixel; per sample: rays/paths 160 B per path + 64 B per pixel
    W, H, L = 1920, 1080, 1
    expect = (1 << 20) * 272 + W * H * 40 + W * H * L * 160 + W * H * 64
    assert footprint(lib, 1 << 20, 16, W, H, L) == expect
    assert footprint(lib, 0, 16, 1, 1, 1) == 40 + 160 + 64
    # monotone in every argument
    assert footprint(lib, 10, 32, 8, 8, 2) > footprint(lib, 10, 16, 8, 8, 2) > footprint(lib, 10, 16, 8, 8, 1)


def test_equal_share_schedule(lib):
    # two requests, 100 paths/sample each, target 1000 -> 5 samples each per step
    adm, done, peak = simulate(lib, [10, 10], [100, 100], [10, 12], [0, 0], budget=100, target=1000)
    assert list(adm) == [0, 0]
    assert list(done) == [1, 2]   # 10 = 5+5; 12 = 5+5+2
    assert peak == 20


def test_leftover_goes_in_admission_order(lib):
    # request 0 needs 1 sample, the rest of the target goes to request 1
    adm, done, _ = simulate(lib, [1, 1], [100, 100], [1, 9], [0, 0], budget=10, target=1000)
    # step 0: shares 5/5, request 0 capped at 1 -> leftover 400 -> request 1 gets 9
    assert list(done) == [0, 0]


def test_minimum_one_sample(lib):
    # paths per sample exceed the per-request share: each still advances 1 sample per step
    adm, done, _ = simulate(lib, [1, 1, 1], [1000, 1000, 1000], [3, 3, 3], [0, 0, 0], budget=10, target=100)
    assert list(done) == [2, 2, 2]


def test_fcfs_admission_under_budget(lib):
    # budget fits two of the three; the third waits for the first completion, no overtaking
    adm, done, peak = simulate(lib, [40, 40, 40, 5], [10, 10, 10, 10], [1, 3, 1, 1], [0, 0, 0, 0],
                               budget=100, target=1000)
    assert adm[0] == adm[1] == 0
    assert adm[2] == 1 and adm[3] == 1  # the small 4th request does not overtake the 3rd
    assert peak <= 100


def test_reject_larger_than_budget(lib):
    adm, done, _ = simulate(lib, [200, 10], [10, 10], [1, 1], [0, 0], budget=100, target=1000)
    assert adm[0] == -2 and done[0] == -2
    assert adm[1] == 0 and done[1] == 0


def test_random_stream_invariants(lib):
    rng = numpy.random.default_rng(5)
    n = 500
    fp = rng.integers(1, 50, n)
    pps = rng.integers(1, 5000, n)
    samples = rng.choice([8, 64, 256], n)
    arrive = numpy.sort(rng.integers(0, 2000, n))
    budget, target = 400, 100000
    adm, done, peak = simulate(lib, fp, pps, samples, arrive, budget, target)
    assert numpy.all(adm >= arrive)
    assert numpy.all(done >= adm)
    assert peak <= budget
    # FCFS: admission iterations are non-decreasing in arrival order
    assert numpy.all(numpy.diff(adm) >= 0)
    # brute-force replay of the memory bound at every iteration
    for it in range(int(done.max()) + 1):
        live = (adm <= it) & (done >= it)
        assert fp[live].sum() <= budget


def scratch(lib, paths, fixed):
    out = ctypes.c_uint64()
    assert lib.fh_policy_scratch_bytes(ctypes.c_uint64(paths), ctypes.c_uint64(fixed), ctypes.byref(out)) == 0
    return out.value


def test_scratch_bytes_formula(lib):
    # two ray queues of 64 B, hits and records of 16 B per path, plus the size-independent part
    assert scratch(lib, 1000, 7) == 1000 * 160 + 7
    assert scratch(lib, 256 << 20, 0) == (256 << 20) * 160


def test_full_budget_with_scratches_stays_in_hbm(lib):
    """A 240 GiB budget holds the server's two render scratches (256 Mi paths per step), the device's two
    Path.sample scratches (256 Mi-path batches) and the admitted requests: admission gets only what is left,
    so the peak of everything together stays within the budget and inside the 288 GiB of one MI355X."""
    GB = 1 << 30
    budget = 240 * GB
    fixed = 300 << 20  # overflow stacks and counters per scratch (upper bound)
    server = 2 * scratch(lib, 256 << 20, fixed)
    device = 2 * scratch(lib, 256 << 20, fixed)
    room = budget - server - device
    assert room > 0
    rng = numpy.random.default_rng(5)
    n = 200
    res = [(256, 256), (1280, 720), (1920, 1080), (3840, 2160)]
    fp, pps, samples = [], [], []
    for _ in range(n):
        prims = int(numpy.exp(rng.uniform(numpy.log(1e3), numpy.log(2e6))))
        w, h = res[rng.integers(0, 4)]
        fp.append(footprint(lib, prims, 32, w, h, 1))
        pps.append(w * h)
        samples.append(int(rng.choice([8, 64, 256])))
    adm, done, peak = simulate(lib, fp, pps, samples, numpy.zeros(n), budget=room, target=256 << 20)
    assert (done >= 0).all()
    assert peak <= room
    assert peak + server + device <= budget <= 288 * GB
    # the budget binds: not everything is admitted at once
    assert sum(fp) > room and adm.max() > 0
