"""Exactness of two integer/float shortcuts the GPU kernels rely on.

* Camera rays divide pixel positions by the image height H.  The kernels form
  x / H as float(double(x) * (1 / H)) for H <= 2^20 and claim that this is the
  correctly rounded float quotient (a float quotient x / H is never within
  2^-46 of a float rounding midpoint).  Checked here against numpy's IEEE
  float32 division on random and adversarial x for many H.
* Primary rays map a ray index to (sample, pixel) with divisions by per-image
  constants, done as a multiply-high and two shifts (Granlund and Montgomery's
  round-up method).  Checked against integer division, including the extremes
  of the 32-bit range.
"""

import numpy


def _quotient_double(x, h):
    return (x.astype(numpy.float64) * (1.0 / numpy.float64(h))).astype(numpy.float32)


def _heights():
    rng = numpy.random.default_rng(5)
    hs = set(range(1, 2200))
    hs.update(int(v) for v in rng.integers(1, 1 << 20, 300))
    hs.update((1 << e) + d for e in range(1, 21) for d in (-1, 0, 1) if 1 <= (1 << e) + d <= (1 << 20))
    return sorted(hs)


def test_double_product_is_the_rounded_float_quotient():
    rng = numpy.random.default_rng(0)
    for h in _heights():
        hf = numpy.float32(h)
        # pixel positions with jitter: (-1, 2^20 + 1), plus exact multiples of H
        # and values on both sides of them
        x = (rng.random(4000) * (h + 2.0) - 1.0).astype(numpy.float32)
        k = rng.integers(0, 3, 64).astype(numpy.float32) * hf
        near = numpy.concatenate([k, numpy.nextafter(k, numpy.float32(-numpy.inf)),
                                  numpy.nextafter(k, numpy.float32(numpy.inf))])
        x = numpy.concatenate([x, near, numpy.float32([0.0, -0.0, 0.5, 0.25, h - 0.5])])
        want = x / hf
        got = _quotient_double(x, h)
        assert numpy.array_equal(want.view(numpy.uint32), got.view(numpy.uint32)), h


def test_double_product_on_quotients_next_to_midpoints():
    # x / H within a few float ulps of a rounding midpoint: x = RN(m * H) for
    # midpoints m, the hardest inputs for a double-rounded quotient
    rng = numpy.random.default_rng(1)
    for h in (3, 7, 135, 1080, 2160, 4321, (1 << 20) - 1, 1 << 20):
        hf = numpy.float32(h)
        q = (rng.random(20000) * 2.0).astype(numpy.float32)
        mid = (q.astype(numpy.float64) + numpy.nextafter(q, numpy.float32(numpy.inf)).astype(numpy.float64)) / 2
        x = (mid * h).astype(numpy.float32)
        x = numpy.concatenate([x, numpy.nextafter(x, numpy.float32(0)), numpy.nextafter(x, numpy.float32(9e9))])
        want = x / hf
        got = _quotient_double(x, h)
        assert numpy.array_equal(want.view(numpy.uint32), got.view(numpy.uint32)), h


def _fastdiv_constants(d):
    ell = 0
    while ell < 32 and (1 << ell) < d:
        ell += 1
    m = ((1 << 32) * ((1 << ell) - d)) // d + 1
    return m, (1 if ell else 0), (ell - 1 if ell else 0)


def _fastdiv(n, d):
    m, s1, s2 = _fastdiv_constants(d)
    n = n.astype(numpy.uint64)
    t = (n * numpy.uint64(m)) >> numpy.uint64(32)
    return ((t + ((n - t) >> numpy.uint64(s1))) >> numpy.uint64(s2)).astype(numpy.uint64)


def test_fastdiv_matches_integer_division():
    rng = numpy.random.default_rng(2)
    ds = list(range(1, 3000)) + [64 * k for k in range(1, 600)] + [int(v) for v in rng.integers(1, 1 << 32, 500)]
    ds += [(1 << e) + o for e in range(1, 32) for o in (-1, 0, 1)] + [(1 << 32) - 1]
    for d in ds:
        m, _, _ = _fastdiv_constants(d)
        assert 0 <= m < (1 << 32)
        n = numpy.concatenate([rng.integers(0, 1 << 32, 2000, dtype=numpy.uint64),
                               numpy.array([0, 1, d - 1, d, d + 1, (1 << 32) - 1, (1 << 32) - 2], dtype=numpy.uint64)
                               % numpy.uint64(1 << 32)])
        k = rng.integers(0, ((1 << 32) - 1) // d + 1, 200, dtype=numpy.uint64)
        n = numpy.concatenate([n, k * numpy.uint64(d), k * numpy.uint64(d) - numpy.uint64(1)]) % numpy.uint64(1 << 32)
        assert numpy.array_equal(_fastdiv(n, d), n // numpy.uint64(d)), d


def test_sincos_hard_table_is_complete_and_correct(tmp_path):
    """The device's table of hard-to-round sine/cosine arguments equals a fresh exhaustive search over every
    float in [0, 8] (gen_sincos_hard.cpp --check), and each tabulated value is the nearest float to the
    sine/cosine evaluated in 80-bit extended precision."""
    import pathlib
    import re
    import subprocess
    rt = pathlib.Path(__file__).resolve().parent.parent / 'fresnel_amd' / 'csrc' / 'rt'
    exe = tmp_path / 'gen_sincos_hard'
    subprocess.run(['g++', '-O2', '-std=c++17', '-ffp-contract=off', '-pthread', str(rt / 'gen_sincos_hard.cpp'),
                    '-lquadmath', '-o', str(exe)], check=True)
    subprocess.run([str(exe), '--check', str(rt / 'sincos_hard.h')], check=True)
    rows = re.findall(r'\{(0x[0-9a-fp.+-]+)f, (0x[0-9a-fp.+-]+)f, (0x[0-9a-fp.+-]+)f\}',
                      (rt / 'sincos_hard.h').read_text())
    assert len(rows) > 50
    for xs, ss, cs in rows:
        x = numpy.longdouble(float.fromhex(xs))
        for got, exact in ((float.fromhex(ss), numpy.sin(x)), (float.fromhex(cs), numpy.cos(x))):
            g = numpy.float32(got)
            err = abs(numpy.longdouble(g) - exact)
            for nb in (numpy.nextafter(g, numpy.float32(2)), numpy.nextafter(g, numpy.float32(-2))):
                assert err <= abs(numpy.longdouble(nb) - exact)
