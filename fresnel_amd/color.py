"""Color conversion helpers (reference: fresnel/color.py:9-62)."""

import numpy


def linear(color):
    """Convert sRGB color(s) to the linear color space fresnel computes in.

    Args:
        color: a (3,) or (4,) color, or an (N, 3) / (N, 4) array of colors with
            components in [0, 1]. An alpha channel is ignored.

    Returns:
        numpy.ndarray of float32 with shape (3,) or (N, 3).

    The arithmetic runs in the input's floating point type (float64 for
    integer input), as the reference's elementwise expressions do.
    """
    c = numpy.ascontiguousarray(color)
    if not (c.shape in ((3,), (4,)) or (c.ndim == 2 and c.shape[1] in (3, 4))):
        raise TypeError("color must be a length 3, 4, Nx3, or Nx4 array")
    x = c[..., 0:3]
    if not numpy.issubdtype(x.dtype, numpy.floating):
        x = x.astype(numpy.float64)
    with numpy.errstate(invalid='ignore'):
        out = numpy.where(x < 0.04045, x / 12.92, ((x + 0.055) / 1.055)**2.4)
    return numpy.ascontiguousarray(out, dtype=numpy.float32)
