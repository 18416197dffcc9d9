"""Rows an assignment to a geometry buffer can touch (util._touched_rows).

The range must contain every row that ``a[index] = value`` changes, for every
index form numpy accepts (a row the range misses would never reach the device),
and it is exact for integers and slices of step +-1.
"""

import numpy
import pytest

from fresnel_amd.util import _touched_rows


def _changed_rows(n, index):
    a = numpy.zeros((n, 3), numpy.float32)
    a[index] = 1.0
    return numpy.nonzero(a.any(axis=1))[0]


INDICES = [0, 7, -1, -12, slice(None), slice(2, 5), slice(5, 2), slice(None, None, -1), slice(9, 1, -3),
           slice(-4, None), slice(1, 12, 5), (3, 1), (slice(2, 6), 0), (slice(None), slice(1, 2)),
           numpy.array([4, 9, 2]), [0, 11], numpy.array([-1, 3]), numpy.array([], dtype=numpy.int64),
           Ellipsis, (Ellipsis, 0), numpy.int32(5), numpy.arange(12) % 4 == 1, numpy.zeros((12, 3), bool)]


@pytest.mark.parametrize('index', INDICES, ids=[repr(i)[:24] for i in INDICES])
def test_range_contains_every_changed_row(index):
    n = 12
    lo, hi = _touched_rows(index, n)
    assert 0 <= lo and hi <= n
    rows = _changed_rows(n, index)
    assert all(lo <= r < hi for r in rows), (index, (lo, hi), rows)


@pytest.mark.parametrize('index, want', [(3, (3, 4)), (-1, (11, 12)), (slice(2, 5), (2, 5)),
                                         (slice(5, 2, -1), (3, 6)), (slice(4, 4), (0, 0)),
                                         ((slice(1, 3), 0), (1, 3)), (numpy.array([8, 2]), (2, 9))])
def test_range_is_tight_for_simple_indices(index, want):
    assert _touched_rows(index, 12) == want


def test_random_fancy_indices():
    rng = numpy.random.default_rng(3)
    for _ in range(200):
        n = int(rng.integers(1, 40))
        k = int(rng.integers(0, n + 1))
        idx = rng.integers(-n, n, k)
        lo, hi = _touched_rows(idx, n)
        rows = _changed_rows(n, idx)
        assert all(lo <= r < hi for r in rows)
