"""Geometric primitives (reference: fresnel/geometry.py:24-873).

Each class allocates its primitives through the scene's device module and
exposes the per-primitive attributes as `util.Array` views; writing one of
them marks the geometry for an acceleration-structure update before the next
render.
"""

import numpy

from . import material as _material_mod
from . import util


class Geometry(object):
    """Operations common to every geometry type.  Use a subclass."""

    def __init__(self):
        raise RuntimeError("Use a specific geometry class")

    def _setup(self, scene, material, outline_material, outline_width):
        self.material = _material_mod.Material() if material is None else material
        if outline_material is None:
            outline_material = _material_mod.Material(solid=1, color=(0, 0, 0))
        self.outline_material = outline_material
        self.outline_width = outline_width

    def _attach(self, scene):
        self.scene = scene
        self.scene.geometry.append(self)

    def enable(self):
        """Make the geometry visible in the scene."""
        self._geometry.enable()

    def disable(self):
        """Hide the geometry (reversible with `enable`)."""
        self._geometry.disable()

    def remove(self):
        """Remove the geometry from its scene permanently."""
        self._geometry.remove()
        self.scene.geometry.remove(self)

    @property
    def material(self):
        """material.Material: surface material."""
        return _material_mod._MaterialProxy(self)

    @material.setter
    def material(self, mat):
        self._geometry.setMaterial(mat._get_cpp_material())

    @property
    def outline_material(self):
        """material.Material: material of the outline band."""
        return _material_mod._OutlineMaterialProxy(self)

    @outline_material.setter
    def outline_material(self, mat):
        self._geometry.setOutlineMaterial(mat._get_cpp_material())

    @property
    def outline_width(self):
        """float: width of the outline band, scene units."""
        return self._geometry.getOutlineWidth()

    @outline_width.setter
    def outline_width(self, width):
        self._geometry.setOutlineWidth(width)


class Cylinder(Geometry):
    """Spherocylinders: a capsule per primitive between two end points.

    Args:
        scene (Scene): scene to add to.
        points ((N, 2, 3) float32): start and end points.
        radius ((N,) float32): radii.
        color ((N, 2, 3) float32): color at each end (each half takes its end's color).
        N (int): number of primitives; defaults to ``len(points)``.
        material, outline_material, outline_width: see `Geometry`.
    """

    def __init__(self,
                 scene,
                 points=((0, 0, 0), (0, 0, 0)),
                 radius=0.5,
                 color=(0, 0, 0),
                 N=None,
                 material=None,
                 outline_material=None,
                 outline_width=0.0):
        if N is None:
            N = len(points)
        self._geometry = scene.device.module.GeometryCylinder(scene._scene, N)
        self._setup(scene, material, outline_material, outline_width)
        self.points[:] = points
        self.radius[:] = radius
        self.color[:] = color
        self._attach(scene)

    def get_extents(self):
        """(2, 3) array: lower and upper corners of the geometry."""
        A = self.points[:, 0]
        B = self.points[:, 1]
        r = self.radius[:].reshape(-1, 1)
        return numpy.array([
            numpy.min([numpy.min(A - r, axis=0), numpy.min(B - r, axis=0)], axis=0),
            numpy.max([numpy.max(A + r, axis=0), numpy.max(B + r, axis=0)], axis=0)
        ])

    @property
    def points(self):
        """(N, 2, 3) `Array`: end points."""
        return util.Array(self._geometry.getPointsBuffer(), geom=self)

    @property
    def radius(self):
        """(N,) `Array`: radii."""
        return util.Array(self._geometry.getRadiusBuffer(), geom=self)

    @property
    def color(self):
        """(N, 2, 3) `Array`: end colors."""
        return util.Array(self._geometry.getColorBuffer(), geom=self)


class Box(Cylinder):
    """Outline of a (triclinic) box drawn with 12 spherocylinders.

    Args:
        scene (Scene): scene to add to.
        box: a number (cube), ``[Lx, Ly, Lz]``, ``[Lx, Ly, Lz, xy, xz, yz]``
            (HOOMD/GSD convention), a mapping or an object with those fields.
        box_radius (float): edge radius.
        box_color ((3,) float32): edge color (the solid material's color).
    """

    def __init__(self, scene, box, box_radius=0.5, box_color=(0, 0, 0)):
        super().__init__(scene=scene, N=12, material=_material_mod.Material(solid=1.0))
        self._box = self._from_box(box)
        self.points[:] = self._generate_points(self._box)
        self.box_radius = box_radius
        self.box_color = box_color

    @staticmethod
    def _from_box(box):
        """(Lx, Ly, Lz, xy, xz, yz) from any supported box description."""
        if hasattr(box, 'Lx') and hasattr(box, 'Ly'):
            return (box.Lx, box.Ly, getattr(box, 'Lz', 0), getattr(box, 'xy', 0),
                    getattr(box, 'xz', 0), getattr(box, 'yz', 0))
        try:
            return (box['Lx'], box['Ly'], box.get('Lz', 0), box.get('xy', 0), box.get('xz', 0),
                    box.get('yz', 0))
        except (IndexError, KeyError, TypeError):
            pass
        if isinstance(box, (int, float)):
            return (box, box, box, 0, 0, 0)
        try:
            n = len(box)
        except TypeError:
            raise TypeError(f"unsupported box type {type(box)}")
        if n not in (1, 3, 6):
            raise ValueError("List-like objects must have length 1, 3, or 6 to be converted to a box.")
        if n == 1:
            return (box[0], box[0], box[0], 0, 0, 0)
        tilts = tuple(box[3:6]) if n == 6 else (0, 0, 0)
        return (box[0], box[1], box[2]) + tilts

    @staticmethod
    def _generate_points(box):
        """The 12 edges of the box centered on the origin, as (12, 2, 3)."""
        Lx, Ly, Lz, xy, xz, yz = box
        # lattice vectors in the HOOMD convention (columns of the box matrix)
        a1 = numpy.array([Lx, 0, 0])
        a2 = numpy.array([xy * Ly, Ly, 0])
        a3 = numpy.array([xz * Lz, yz * Lz, Lz])
        origin = -(a1 + a2 + a3) / 2

        def corner(i, j, k):
            c = origin
            if i:
                c = c + a1
            if j:
                c = c + a2
            if k:
                c = c + a3
            return c

        # each edge joins two corners that differ in one lattice index
        edges = [((0, 0, 0), (1, 0, 0)), ((0, 0, 0), (0, 1, 0)), ((0, 0, 0), (0, 0, 1)),
                 ((1, 0, 0), (1, 0, 1)), ((1, 0, 0), (1, 1, 0)), ((0, 1, 0), (1, 1, 0)),
                 ((0, 1, 0), (0, 1, 1)), ((0, 0, 1), (1, 0, 1)), ((0, 0, 1), (0, 1, 1)),
                 ((1, 0, 1), (1, 1, 1)), ((0, 1, 1), (1, 1, 1)), ((1, 1, 0), (1, 1, 1))]
        return numpy.asarray([[corner(*a), corner(*b)] for a, b in edges])

    @property
    def box(self):
        """Box parameters (Lx, Ly, Lz, xy, xz, yz); setting regenerates the edges."""
        return self._box

    @box.setter
    def box(self, box):
        self._box = self._from_box(box)
        self.points[:] = self._generate_points(self._box)

    @property
    def box_color(self):
        """(3,) float32: edge color."""
        return self.material.color

    @box_color.setter
    def box_color(self, color):
        self.material.color = color

    @property
    def box_radius(self):
        """float: edge radius."""
        return self.radius[:][0]

    @box_radius.setter
    def box_radius(self, radius):
        self.radius[:] = radius


class Polygon(Geometry):
    """Flat (sphero)polygons in the z = 0 plane.

    Args:
        scene (Scene): scene to add to.
        vertices ((V, 2) float32): polygon vertices (shared by all primitives).
        position ((N, 2) float32): polygon positions.
        angle ((N,) float32): rotation angles about +z, radians.
        color ((N, 3) float32): colors.
        rounding_radius (float): rounding radius of a spheropolygon.
        N (int): number of primitives; defaults to ``len(position)``.
        material, outline_material, outline_width: see `Geometry`.
    """

    def __init__(self,
                 scene,
                 vertices,
                 position=(0, 0),
                 angle=0,
                 color=(0, 0, 0),
                 rounding_radius=0,
                 N=None,
                 material=None,
                 outline_material=None,
                 outline_width=0.0):
        if N is None:
            N = len(position)
        self._geometry = scene.device.module.GeometryPolygon(scene._scene, vertices,
                                                             rounding_radius, N)
        self._setup(scene, material, outline_material, outline_width)
        self.position[:] = position
        self.angle[:] = angle
        self.color[:] = color
        self._attach(scene)

    @property
    def position(self):
        """(N, 2) `Array`: positions."""
        return util.Array(self._geometry.getPositionBuffer(), geom=self)

    @property
    def angle(self):
        """(N,) `Array`: rotation angles (radians)."""
        return util.Array(self._geometry.getAngleBuffer(), geom=self)

    @property
    def color(self):
        """(N, 3) `Array`: colors."""
        return util.Array(self._geometry.getColorBuffer(), geom=self)

    def get_extents(self):
        """(2, 3) array: lower and upper corners of the geometry."""
        pos = self.position[:]
        r = self._geometry.getRadius()
        lo = numpy.min(pos - r, axis=0)
        hi = numpy.max(pos + r, axis=0)
        return numpy.array([[lo[0], lo[1], -1e-5], [hi[0], hi[1], 1e-5]])


class Sphere(Geometry):
    """Spheres.

    Args:
        scene (Scene): scene to add to.
        position ((N, 3) float32): centers.
        radius ((N,) float32): radii.
        color ((N, 3) float32): colors.
        N (int): number of primitives; defaults to ``len(position)``.
        material, outline_material, outline_width: see `Geometry`.
    """

    def __init__(self,
                 scene,
                 position=(0, 0, 0),
                 radius=0.5,
                 color=(0, 0, 0),
                 N=None,
                 material=None,
                 outline_material=None,
                 outline_width=0.0):
        if N is None:
            N = len(position)
        self._geometry = scene.device.module.GeometrySphere(scene._scene, N)
        self._setup(scene, material, outline_material, outline_width)
        self.position[:] = position
        self.radius[:] = radius
        self.color[:] = color
        self._attach(scene)

    def get_extents(self):
        """(2, 3) array: lower and upper corners of the geometry."""
        pos = self.position[:]
        r = self.radius[:].reshape(-1, 1)
        return numpy.array([numpy.min(pos - r, axis=0), numpy.max(pos + r, axis=0)])

    @property
    def position(self):
        """(N, 3) `Array`: centers."""
        return util.Array(self._geometry.getPositionBuffer(), geom=self)

    @property
    def radius(self):
        """(N,) `Array`: radii."""
        return util.Array(self._geometry.getRadiusBuffer(), geom=self)

    @property
    def color(self):
        """(N, 3) `Array`: colors."""
        return util.Array(self._geometry.getColorBuffer(), geom=self)


class Mesh(Geometry):
    """Instances of one triangle mesh.

    Args:
        scene (Scene): scene to add to.
        vertices ((3T, 3) float32): triangle vertices, three per triangle.
        position ((N, 3) float32): instance positions.
        orientation ((N, 4) float32): instance orientations (quaternion s, x, y, z).
        color ((3T, 3) float32): per-vertex colors (shared by instances).
        N (int): number of instances; defaults to ``len(position)``.
        material, outline_material, outline_width: see `Geometry`.
    """

    def __init__(self,
                 scene,
                 vertices,
                 position=(0, 0, 0),
                 orientation=(1, 0, 0, 0),
                 color=(0, 0, 0),
                 N=None,
                 material=None,
                 outline_material=None,
                 outline_width=0.0):
        if N is None:
            N = len(position)
        self.vertices = numpy.asarray(vertices, dtype=numpy.float32)
        self._geometry = scene.device.module.GeometryMesh(scene._scene, self.vertices, N)
        self._setup(scene, material, outline_material, outline_width)
        self.position[:] = position
        self.orientation[:] = orientation
        self.color[:] = color
        self._attach(scene)

    @property
    def position(self):
        """(N, 3) `Array`: instance positions."""
        return util.Array(self._geometry.getPositionBuffer(), geom=self)

    @property
    def orientation(self):
        """(N, 4) `Array`: instance orientations."""
        return util.Array(self._geometry.getOrientationBuffer(), geom=self)

    @property
    def color(self):
        """(3T, 3) `Array`: per-vertex colors."""
        return util.Array(self._geometry.getColorBuffer(), geom=self)

    def get_extents(self):
        """(2, 3) array: lower and upper corners of the geometry.

        As in the reference (fresnel/geometry.py:709-738), the offsets are the
        smallest and largest vertex coordinate over all axes, applied to every
        axis of every instance position; orientation is ignored.
        """
        lo = numpy.min(self.vertices)
        hi = numpy.max(self.vertices)
        pos = self.position[:]
        return numpy.array([numpy.min(pos + lo, axis=0), numpy.max(pos + hi, axis=0)])


class ConvexPolyhedron(Geometry):
    """Convex polyhedra described by outward planes.

    Args:
        scene (Scene): scene to add to.
        polyhedron_info (dict): ``face_origin`` (P,3), ``face_normal`` (P,3),
            ``face_color`` (P,3) and ``radius`` (bounding radius), e.g. from
            `util.convex_polyhedron_from_vertices`.
        position ((N, 3) float32): positions.
        orientation ((N, 4) float32): orientations (quaternion s, x, y, z).
        color ((N, 3) float32): per-polyhedron colors.
        N (int): number of primitives; defaults to ``len(position)``.
        material, outline_material, outline_width: see `Geometry`.
    """

    def __init__(self,
                 scene,
                 polyhedron_info,
                 position=(0, 0, 0),
                 orientation=(1, 0, 0, 0),
                 color=(0, 0, 0),
                 N=None,
                 material=None,
                 outline_material=None,
                 outline_width=0.0):
        if N is None:
            N = len(position)
        r = polyhedron_info['radius']
        self._geometry = scene.device.module.GeometryConvexPolyhedron(
            scene._scene, polyhedron_info['face_origin'], polyhedron_info['face_normal'],
            polyhedron_info['face_color'], N, r)
        self._setup(scene, material, outline_material, outline_width)
        self._radius = r
        self.position[:] = position
        self.orientation[:] = orientation
        self.color[:] = color
        self._attach(scene)

    def get_extents(self):
        """(2, 3) array: lower and upper corners of the geometry."""
        pos = self.position[:]
        r = self._radius
        return numpy.array([numpy.min(pos - r, axis=0), numpy.max(pos + r, axis=0)])

    @property
    def position(self):
        """(N, 3) `Array`: positions."""
        return util.Array(self._geometry.getPositionBuffer(), geom=self)

    @property
    def orientation(self):
        """(N, 4) `Array`: orientations."""
        return util.Array(self._geometry.getOrientationBuffer(), geom=self)

    @property
    def color(self):
        """(N, 3) `Array`: colors."""
        return util.Array(self._geometry.getColorBuffer(), geom=self)

    @property
    def color_by_face(self):
        """float: 0 colors by primitive, 1 by face, values between blend."""
        return self._geometry.getColorByFace()

    @color_by_face.setter
    def color_by_face(self, v):
        self._geometry.setColorByFace(v)
