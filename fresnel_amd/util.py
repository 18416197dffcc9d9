"""Buffer access helpers and polyhedron construction (reference: fresnel/util.py)."""

import io
import itertools

import numpy

try:
    import PIL.Image as PIL_Image
except ImportError:  # pragma: no cover - PIL is optional
    PIL_Image = None


class Array(object):
    """numpy-like view of a backend-owned buffer.

    Reads return copies; writes copy into the buffer and, for geometry
    buffers, notify the geometry so the next render rebuilds its acceleration
    structure (reference: fresnel/util.py:15-70).

    Attributes:
        shape (tuple): dimensions of the array.
        dtype: numpy data type.
    """

    def __init__(self, buf, geom):
        self.buf = buf
        self.geom = geom
        layout = getattr(buf, 'layout', None)
        if layout is not None:
            # shape and type without mapping: a tracer output stays on the
            # device until it is read
            shape, dtype = layout
            self.shape = tuple(int(x) for x in shape)
            self.dtype = numpy.dtype(dtype)
            return
        self.buf.map()
        a = numpy.asarray(self.buf)
        self.shape = a.shape
        self.dtype = a.dtype
        self.buf.unmap()

    def __setitem__(self, slice, data):
        """Assign ``data`` to ``self[slice]``.

        For a geometry buffer only the rows the assignment touches are marked
        for upload when the backend supports it (``update_rows``); otherwise the
        whole geometry is, as in the reference (fresnel/util.py:41-48).
        """
        self.buf.map()
        a = numpy.asarray(self.buf)
        a[slice] = data
        self.buf.unmap()
        if self.geom is None:
            return
        mark = getattr(self.buf, 'update_rows', None)
        rows = _touched_rows(slice, self.shape[0]) if mark is not None else None
        if rows is None:
            self.geom._geometry.update()
        elif rows[1] > rows[0]:
            mark(rows[0], rows[1] - rows[0])

    def __getitem__(self, slice):
        """Return a copy of ``self[slice]``."""
        self.buf.map()
        a = numpy.asarray(self.buf)
        data = numpy.array(a[slice], copy=True)
        self.buf.unmap()
        return data


def _touched_rows(index, n):
    """Rows [lo, hi) along the first axis that ``a[index] = ...`` can write.

    Integers, slices and integer or boolean arrays in the first position are
    resolved exactly (as a bounding range); anything else (an Ellipsis, a
    newaxis, a field name) counts as every row.
    """
    first = index[0] if isinstance(index, tuple) and len(index) > 0 else index
    if isinstance(index, tuple) and len(index) == 0:
        return 0, n
    if isinstance(first, (int, numpy.integer)) and not isinstance(first, bool):
        i = int(first)
        i = i + n if i < 0 else i
        return (i, i + 1) if 0 <= i < n else (0, n)
    if isinstance(first, slice):
        start, stop, step = first.indices(n)
        rows = range(start, stop, step)
        if len(rows) == 0:
            return 0, 0
        return min(rows[0], rows[-1]), max(rows[0], rows[-1]) + 1
    if isinstance(first, (list, numpy.ndarray)):
        idx = numpy.asarray(first)
        if idx.dtype == bool and idx.shape == (n,):
            idx = numpy.nonzero(idx)[0]
        elif idx.dtype.kind not in 'iu':
            return 0, n
        if idx.size == 0:
            return 0, 0
        idx = numpy.where(idx < 0, idx + n, idx)
        return int(idx.min()), int(idx.max()) + 1
    return 0, n


class ImageArray(Array):
    """An (H, W, 4) uint8 RGBA image buffer that Jupyter can display."""

    def _repr_png_(self):
        if PIL_Image is None:
            raise RuntimeError("No PIL.Image module to format png")
        self.buf.map()
        f = io.BytesIO()
        a = numpy.asarray(self.buf)
        PIL_Image.fromarray(numpy.ascontiguousarray(a), mode='RGBA').save(f, 'png')
        self.buf.unmap()
        return f.getvalue()


def _face_corners(pts, n, tol):
    """Corners of the convex polygon spanned by coplanar points (monotone chain).

    Points inside the face or on its edges are not corners, as in a convex
    hull's vertex list.
    """
    a = numpy.cross(n, [1.0, 0.0, 0.0])
    if numpy.linalg.norm(a) < 0.5:
        a = numpy.cross(n, [0.0, 1.0, 0.0])
    a /= numpy.linalg.norm(a)
    b = numpy.cross(n, a)
    q = numpy.stack([pts @ a, pts @ b], axis=1)
    order = numpy.lexsort((q[:, 1], q[:, 0]))

    def cross(o, p, r):
        return (p[0] - o[0]) * (r[1] - o[1]) - (p[1] - o[1]) * (r[0] - o[0])

    def half(idx):
        h = []
        for i in idx:
            while len(h) >= 2 and cross(q[h[-2]], q[h[-1]], q[i]) <= tol:
                h.pop()
            h.append(i)
        return h

    lower = half(order)
    upper = half(order[::-1])
    corners = lower[:-1] + upper[:-1]
    return pts[corners]


def convex_polyhedron_from_vertices(vertices):
    """Build the plane description of the convex hull of ``vertices``.

    The reference delegates to qhull (fresnel/util.py:98-126,
    common/ConvexPolyhedronBuilder.cc:25-83).  This version finds hull faces
    directly: every vertex triple spans a candidate plane, a candidate is a
    face when no vertex lies outside it, and coplanar candidates merge into one
    face.  It is exact for the small vertex counts polyhedra use.

    Returns:
        dict with ``face_origin`` (F,3), ``face_normal`` (F,3) outward unit
        normals, ``face_color`` (F,3), ``face_sides`` (F,) vertex counts per
        face and ``radius`` (circumsphere radius about the origin).
    """
    v = numpy.asarray(vertices, dtype=numpy.float64)
    if v.ndim != 2 or v.shape[1] != 3 or v.shape[0] < 4:
        raise ValueError("vertices must be an (N>=4, 3) array")
    scale = numpy.max(numpy.abs(v))
    tol = 1e-9 * max(scale, 1.0)
    faces = []  # (normal, offset, vertex index set)
    for i, j, k in itertools.combinations(range(v.shape[0]), 3):
        n = numpy.cross(v[j] - v[i], v[k] - v[i])
        norm = numpy.linalg.norm(n)
        if norm <= tol * tol:
            continue
        n = n / norm
        off = numpy.dot(n, v[i])
        side = v @ n - off
        if numpy.all(side <= tol):
            pass
        elif numpy.all(side >= -tol):
            n, off, side = -n, -off, -side
        else:
            continue
        on = frozenset(numpy.nonzero(numpy.abs(side) <= tol)[0].tolist())
        if not any(numpy.allclose(n, f[0], atol=1e-9) and abs(off - f[1]) <= tol for f in faces):
            faces.append((n, off, on))
    origins = []
    normals = []
    sides = []
    for n, off, on in faces:
        pts = _face_corners(v[sorted(on)], n, tol)
        origins.append(numpy.mean(pts, axis=0))
        normals.append(n)
        sides.append(len(pts))
    radius = float(numpy.max(numpy.linalg.norm(v, axis=1)))
    return dict(face_origin=numpy.array(origins, dtype=numpy.float32),
                face_normal=numpy.array(normals, dtype=numpy.float32),
                face_color=numpy.full((len(faces), 3), 0.9, dtype=numpy.float32),
                face_sides=numpy.array(sides, dtype=numpy.uint32),
                radius=radius)
