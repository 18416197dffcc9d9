"""Scenes of the reference's golden-image tests, as data.

Each builder reproduces the scene parameters of one fixture in
fresnel/test/ (file cited per builder); each CASES entry names one golden PNG
of fresnel/test/reference/, the builder, the change made before rendering and
how it is rendered.
"""

import math
from collections import namedtuple

import numpy

import fresnel_amd as fresnel
from conftest import test_lights

lin = fresnel.color.linear


def hex_sphere(device):
    """fresnel/test/conftest.py:73-95"""
    scene = fresnel.Scene(device, lights=test_lights())
    pos = [[2 * math.cos(i * 2 * math.pi / 6), 2 * math.sin(i * 2 * math.pi / 6), 0] for i in range(6)]
    g = fresnel.geometry.Sphere(scene, position=pos, radius=1.0)
    g.material = fresnel.material.Material(solid=0.0, color=lin([1, 0.874, 0.169]))
    g.outline_width = 0.12
    scene.camera = fresnel.camera.Orthographic(position=(0, 0, 10), look_at=(0, 0, 0), up=(0, 1, 0),
                                               height=6)
    return scene


def four_spheres(device):
    """fresnel/test/test_geometry_sphere.py:16-36"""
    scene = fresnel.Scene(device, lights=test_lights())
    fresnel.geometry.Sphere(scene,
                            position=[[1, 0, 1], [1, 0, -1], [-1, 0, 1], [-1, 0, -1]],
                            radius=1.0,
                            material=fresnel.material.Material(color=lin([0.42, 0.267, 1])),
                            color=[[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1]])
    scene.camera = fresnel.camera.Orthographic(position=(10, 10, 10), look_at=(0, 0, 0), up=(0, 1, 0),
                                               height=4)
    return scene


def four_cylinders(device):
    """fresnel/test/test_geometry_cylinder.py:16-39"""
    scene = fresnel.Scene(device, lights=test_lights())
    fresnel.geometry.Cylinder(scene,
                              points=[[[-5, -5, 0], [-5, 5, 0]], [[5, -5, -5], [5, 5, 5]],
                                      [[3, 3, -3], [-3, -3, -3]], [[-2, 2, 2], [2, -2, -2]]],
                              radius=1.0,
                              color=[0.9, 0.9, 0.9],
                              material=fresnel.material.Material(color=lin([0.42, 0.267, 1])))
    scene.camera = fresnel.camera.Orthographic(position=(0, 2, 10), look_at=(0, 0, 0), up=(0, 1, 0),
                                               height=15)
    return scene


def box(device):
    """fresnel/test/test_geometry_box.py:18-33"""
    scene = fresnel.Scene(device, lights=test_lights())
    fresnel.geometry.Box(scene, [1, 2, 3, 0.4, 0.5, 0.6], box_radius=0.2, box_color=[1, 0, 1])
    scene.camera = fresnel.camera.Orthographic(position=(10, 10, 10), look_at=(0, 0, 0), up=(0, 1, 0),
                                               height=4)
    return scene


def eight_polyhedra(device):
    """fresnel/test/test_geometry_convex_polyhedron.py:18-76"""
    scene = fresnel.Scene(device, lights=test_lights())
    position = [[2.5 * i, 2.5 * j, 2.5 * k] for k in range(2) for i in range(2) for j in range(2)]
    origins, normals, colors = [], [], []
    for v in [-1, 1]:
        for axis in range(3):
            e = [0, 0, 0]
            e[axis] = v
            origins.append(list(e))
            normals.append(list(e))
            colors.append([178 / 255, 223 / 255, 138 / 255])
    for x in [-1, 1]:
        for y in [-1, 1]:
            for z in [-1, 1]:
                normals.append([x, y, z])
                origins.append([x * 0.75, y * 0.75, z * 0.75])
                colors.append([166 / 255, 206 / 255, 227 / 255])
    info = {'face_normal': normals, 'face_origin': origins, 'radius': math.sqrt(3),
            'face_color': lin(colors)}
    g = fresnel.geometry.ConvexPolyhedron(scene, info, position=position)
    g.material = fresnel.material.Material(color=lin([1.0, 0, 0]), roughness=0.8, specular=0.5,
                                           primitive_color_mix=0.0)
    g.orientation[:] = [1, 0, 0, 0]
    scene.camera = fresnel.camera.Orthographic(position=(20, 20, 20), look_at=(0, 0, 0), up=(0, 1, 0),
                                               height=7)
    return scene


def one_triangle(device):
    """fresnel/test/test_geometry_mesh.py:16-44"""
    scene = fresnel.Scene(device, lights=test_lights())
    g = fresnel.geometry.Mesh(scene, vertices=[[-1, -1, 0], [1, -1, 0], [-1, 1, 0]], N=1)
    g.color[:] = [[1, 0, 0], [0, 1, 0], [0, 0, 1]]
    g.material = fresnel.material.Material(color=lin([1.0, 0, 0]), roughness=0.8, specular=0.5,
                                           primitive_color_mix=0.0, solid=1)
    g.outline_material = fresnel.material.Material(color=lin([0, 1.0, 0]), roughness=0.8,
                                                   specular=0.5, primitive_color_mix=0.0, solid=1)
    g.orientation[:] = [1, 0, 0, 0]
    scene.camera = fresnel.camera.Orthographic(position=(0, 0, 20), look_at=(0, 0, 0), up=(0, 1, 0),
                                               height=2.1)
    return scene


def tetrahedra(device):
    """fresnel/test/test_geometry_mesh.py:53-91"""
    scene = fresnel.Scene(device, lights=test_lights())
    v = [(1, 1, 1), (1, -1, -1), (-1, 1, -1), (-1, -1, 1)]
    tris = [v[0], v[1], v[2], v[2], v[1], v[3], v[2], v[3], v[0], v[1], v[0], v[3]]
    g = fresnel.geometry.Mesh(scene, vertices=tris, N=4)
    g.color[0:2, :] = [0.9, 0, 0]
    g.color[3:5, :] = [0, 0.9, 0]
    g.color[6:8, :] = [0, 0, 0.9]
    g.color[9:11, :] = [0.9, 0, 0.9]
    g.material = fresnel.material.Material(color=lin([1.0, 0, 0]), roughness=0.8, specular=0.5,
                                           primitive_color_mix=1.0, solid=0)
    g.position[:] = [[-2, -2, 0], [2, -2, 0], [2, 2, 0], [-2, 2, 0]]
    g.orientation[:] = [[0.03723867, 0.38927173, -0.73216521, -0.55768711],
                        [-0.32661186, 0.43644863, -0.09899935, 0.83248808],
                        [0.25624845, 0.32632096, -0.11995704, -0.9019211],
                        [-0.78025512, -0.12102377, 0.24947819, 0.56063877]]
    scene.camera = fresnel.camera.Orthographic(position=(0, 0, -20), look_at=(0, 0, 0), up=(0, 1, 0),
                                               height=7.5)
    return scene


def _polygons(device, rounding):
    scene = fresnel.Scene(device, lights=test_lights())
    mat = fresnel.material.Material(color=lin([0.42, 0.267, 1]), solid=1)
    fresnel.geometry.Polygon(scene, N=2, rounding_radius=rounding,
                             vertices=[[-1, -1], [1, -1], [1, 1], [0, 0], [-1, 1]],
                             position=[[-1.5, 0], [1.5, 0]], angle=[0.1, -0.2],
                             color=[[0, 0, 1], [0, 1, 0]], material=mat)
    scene.camera = fresnel.camera.Orthographic(position=(0, 0, -2), look_at=(0, 0, 0), up=(0, 1, 0),
                                               height=5)
    return scene


def polygons(device):
    """fresnel/test/test_geometry_polygon.py:43-66"""
    return _polygons(device, 0)


def rounded_polygons(device):
    """fresnel/test/test_geometry_polygon.py:16-40"""
    return _polygons(device, 0.3)


def perspective_grid(device):
    """fresnel/test/test_camera_perspective.py:13-36"""
    scene = fresnel.Scene(device, lights=test_lights())
    position = [[i, 0, j] for i in range(-8, 8) for j in range(-8, 8)]
    fresnel.geometry.Sphere(scene, position=position, radius=0.7,
                            material=fresnel.material.Material(color=lin([0.42, 0.267, 0.9])))
    scene.camera = fresnel.camera.Perspective(position=(0, 4, -10), look_at=(0, 0, 0), up=(0, 1, 0),
                                              focus_distance=10)
    return scene


def multiple_geometries(device):
    """fresnel/test/test_scene.py:93-115 (first render)"""
    scene = fresnel.Scene(device, lights=test_lights())
    scene.camera = fresnel.camera.Orthographic(position=(0, 0, 10), look_at=(0, 0, 0), up=(0, 1, 0),
                                               height=7)
    g1 = fresnel.geometry.Sphere(scene, position=[[-4, 1, 0], [-4, -1, 0], [-2, 1, 0], [-2, -1, 0]],
                                 radius=1.0)
    g1.material = fresnel.material.Material(solid=1.0, color=lin([0.42, 0.267, 1]))
    g1.outline_width = 0.12
    g2 = fresnel.geometry.Sphere(scene, position=[[4, 1, 0], [4, -1, 0], [2, 1, 0], [2, -1, 0]],
                                 radius=1.0)
    g2.material = fresnel.material.Material(solid=0.0, color=lin([1, 0.874, 0.169]))
    return scene


# ----------------------------------------------------------------- changes
def _g(scene):
    return scene.geometry[0]


def _set(attr, value):

    def f(scene):
        setattr(_g(scene), attr, value)

    return f


def _set_array(attr, value):

    def f(scene):
        getattr(_g(scene), attr)[:] = value

    return f


def _set_mat(which, attr, value):

    def f(scene):
        setattr(getattr(_g(scene), which), attr, value)

    return f


def _six_colors(scene):
    g = _g(scene)
    for i, c in enumerate([[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [0, 1, 1], [0, 0, 0]]):
        g.color[i] = lin(c)


def _chain(*fs):

    def f(scene):
        for fn in fs:
            fn(scene)

    return f


def _cam(attr, value):

    def f(scene):
        setattr(scene.camera, attr, value)

    return f


def _face_color(scene):
    g = _g(scene)
    g.color_by_face = 1.0
    g.material.primitive_color_mix = 1.0


def _nothing(scene):
    pass


Case = namedtuple('Case', 'name builder change render w h tol')
P = 'preview'
PT = 'pathtrace'

CASES = [
    # spheres, fresnel/test/test_geometry_sphere.py
    Case('test_geometry_sphere.test_render', four_spheres, _nothing, P, 150, 100, 1.0),
    Case('test_geometry_sphere.test_radius', four_spheres,
         _set_array('radius', numpy.array([0.5, 0.6, 0.8, 1.0], dtype=numpy.float32)), P, 150, 100, 1.0),
    Case('test_geometry_sphere.test_position', four_spheres,
         _set_array('position', numpy.array([[1.5, 0, 1], [1.5, 0, -1], [-1.5, 0, 1], [-1.5, 0, -1]],
                                            dtype=numpy.float32)), P, 150, 100, 1.0),
    Case('test_geometry_sphere.test_color', four_spheres,
         _chain(_set_mat('material', 'primitive_color_mix', 1.0),
                _set_array('color', lin(numpy.array([[1, 1, 1], [0, 0, 1], [0, 1, 0], [1, 0, 0]],
                                                    dtype=numpy.float32)))), P, 150, 100, 1.0),
    Case('test_geometry_sphere.test_outline', four_spheres, _set('outline_width', 0.1), P, 150, 100, 1.0),
    # spherocylinders, fresnel/test/test_geometry_cylinder.py
    Case('test_geometry_clyinder.test_render', four_cylinders, _nothing, P, 150, 100, 1.0),
    Case('test_geometry_clyinder.test_radius', four_cylinders,
         _set_array('radius', numpy.array([0.5, 0.6, 0.8, 1.0], dtype=numpy.float32)), P, 150, 100, 1.0),
    Case('test_geometry_clyinder.test_position', four_cylinders,
         _set_array('points', numpy.array([[[-5, -5, 0], [-5, 5, 0]], [[-3, 5, 0], [3, 5, 0]],
                                           [[5, 5, 0], [5, -5, -0]], [[3, -5, 0], [-3, -5, 0]]],
                                          dtype=numpy.float32)), P, 150, 100, 1.0),
    Case('test_geometry_clyinder.test_color', four_cylinders,
         _chain(_set_mat('material', 'primitive_color_mix', 1.0),
                _set_array('color', numpy.array([[[0.9, 0, 0], [0, 0.9, 0]], [[0, 0, 0.9], [0, 0.9, 0.9]],
                                                 [[0.9, 0.9, 0], [0.9, 0, 0.9]],
                                                 [[0.1, 0.2, 0.9], [0.9, 0.2, 0.3]]],
                                                dtype=numpy.float32))), P, 150, 100, 1.0),
    Case('test_geometry_clyinder.test_outline', four_cylinders, _set('outline_width', 0.3), P, 150, 100,
         1.0),
    # box, fresnel/test/test_geometry_box.py
    Case('test_geometry_box.test_render', box, _nothing, P, 150, 100, 1.0),
    Case('test_geometry_box.test_radius', box,
         _set_array('radius', numpy.array(0.1, dtype=numpy.float32)), P, 150, 100, 1.0),
    Case('test_geometry_box.test_color', box,
         _chain(_set_mat('material', 'primitive_color_mix', 1.0),
                _set_array('color', numpy.zeros((12, 2, 3), dtype=numpy.float32) + [1, 0, 0])), P, 150,
         100, 1.0),
    Case('test_geometry_box.test_box_radius', box, _set('box_radius', 0.1), P, 150, 100, 1.0),
    Case('test_geometry_box.test_box_color', box, _set('box_color', numpy.array([1, 0, 1],
                                                                                dtype=numpy.float32)),
         P, 150, 100, 1.0),
    Case('test_geometry_box.test_render0', box, _set('box', 2), P, 150, 100, 1.0),
    Case('test_geometry_box.test_render1', box, _set('box', [1.5]), P, 150, 100, 1.0),
    Case('test_geometry_box.test_render2', box, _set('box', [1, 2, 3]), P, 150, 100, 1.0),
    Case('test_geometry_box.test_render3', box, _set('box', [1, 1.2, 2, 0.2, 0, 0.9]), P, 150, 100, 1.0),
    Case('test_geometry_box.test_render4', box,
         _set('box', namedtuple('box_tuple', 'Lx Ly Lz xy xz yz')(2, 2, 3, 0.5, 0.5, 0.7)), P, 150, 100,
         1.0),
    Case('test_geometry_box.test_render5', box,
         _set('box', {'Lx': 1, 'Ly': 2, 'Lz': 1, 'xy': 0.4, 'xz': 0.5, 'yz': 0.3}), P, 150, 100, 1.0),
    # convex polyhedra, fresnel/test/test_geometry_convex_polyhedron.py
    Case('test_geometry_convex_polyhedron.test_render', eight_polyhedra, _nothing, P, 150, 100, 1.0),
    Case('test_geometry_convex_polyhedron.test_outline', eight_polyhedra, _set('outline_width', 0.1), P,
         150, 100, 1.0),
    Case('test_geometry_convex_polyhedron.test_face_color', eight_polyhedra, _face_color, P, 150, 100,
         1.0),
    # meshes, fresnel/test/test_geometry_mesh.py
    Case('test_geometry_mesh.test_render', one_triangle, _nothing, P, 100, 100, 1.0),
    Case('test_geometry_mesh.test_outline', one_triangle, _set('outline_width', 0.1), P, 100, 100, 1.0),
    Case('test_geometry_mesh.test_color_interp', one_triangle,
         _set_mat('material', 'primitive_color_mix', 1.0), P, 100, 100, 1.0),
    Case('test_geometry_mesh.test_multiple', tetrahedra, _nothing, P, 100, 100, 1.0),
    # polygons, fresnel/test/test_geometry_polygon.py
    Case('test_geometry_polygon.test_render', polygons, _nothing, P, 150, 100, 1.0),
    Case('test_geometry_polygon.test_rounded', rounded_polygons, _set('outline_width', 0.1), P, 150, 100,
         1.0),
    Case('test_geometry_polygon.test_angle', polygons,
         _set_array('angle', numpy.array([-0.8, 0.5], dtype=numpy.float32)), P, 150, 100, 1.0),
    Case('test_geometry_polygon.test_position', polygons,
         _set_array('position', numpy.array([[-2, 0], [3, 0]], dtype=numpy.float32)), P, 150, 100, 1.0),
    Case('test_geometry_polygon.test_color', polygons,
         _chain(_set_mat('material', 'primitive_color_mix', 1.0),
                _set_array('color', numpy.array([lin([0, 1, 0]), lin([1, 0, 0])], dtype=numpy.float32))),
         P, 150, 100, 1.0),
    Case('test_geometry_polygon.test_outline', polygons, _set('outline_width', 0.1), P, 150, 100, 1.0),
    # materials, fresnel/test/test_material.py
    Case('test_material.test_set_material', hex_sphere,
         _set('material', fresnel.material.Material(solid=0.0, color=lin([1, 0, 0]),
                                                    primitive_color_mix=0.0)), P, 100, 100, 1.0),
    Case('test_material.test_solid', hex_sphere, _set_mat('material', 'solid', 1.0), P, 100, 100, 1.0),
    Case('test_material.test_color', hex_sphere, _set_mat('material', 'color', lin([0, 0, 1])), P, 100,
         100, 1.0),
    Case('test_material.test_specular', hex_sphere, _set_mat('material', 'specular', 1.0), P, 100, 100,
         1.0),
    Case('test_material.test_roughness', hex_sphere, _set_mat('material', 'roughness', 1.0), P, 100, 100,
         1.0),
    Case('test_material.test_metal', hex_sphere, _set_mat('material', 'metal', 1.0), P, 100, 100, 1.0),
    Case('test_material.test_primitive_color_mix', hex_sphere,
         _chain(_set('material', fresnel.material.Material(solid=1.0, color=lin([1, 0, 0]),
                                                           primitive_color_mix=1.0)), _six_colors), P,
         100, 100, 1.0),
    # outline materials, fresnel/test/test_outline_material.py
    Case('test_outline_material.test_set_material', hex_sphere,
         _chain(_set('outline_width', 0.3),
                _set('outline_material', fresnel.material.Material(solid=0.0, color=lin([1, 0, 0]),
                                                                   primitive_color_mix=0.0))), P, 100,
         100, 1.0),
    Case('test_outline_material.test_solid', hex_sphere,
         _chain(_set('outline_width', 0.3), _set_mat('outline_material', 'solid', 1.0)), P, 100, 100, 1.0),
    Case('test_outline_material.test_color', hex_sphere,
         _chain(_set('outline_width', 0.3), _set_mat('outline_material', 'color', lin([0, 0, 1]))), P, 100,
         100, 1.0),
    Case('test_outline_material.test_primitive_color_mix', hex_sphere,
         _chain(_set('outline_width', 0.3),
                _set('outline_material', fresnel.material.Material(solid=1.0, color=lin([1, 0, 0]),
                                                                   primitive_color_mix=1.0)),
                _six_colors), P, 100, 100, 1.0),
    # scene, fresnel/test/test_scene.py
    Case('test_scene.test_camera', hex_sphere,
         lambda s: setattr(s, 'camera', fresnel.camera.Orthographic(position=(1, 0, 10), look_at=(1, 0, 0),
                                                                    up=(0, 1, 0), height=6)), P, 100, 100,
         1.0),
    Case('test_scene.test_light_dir', hex_sphere,
         _chain(lambda s: setattr(s.lights[0], 'direction', (1, 0, 0)),
                lambda s: setattr(s.lights[0], 'color', (0.5, 0.5, 0.5))), P, 100, 100, 1.0),
    Case('test_scene.test_multiple_geometries1', multiple_geometries, _nothing, P, 200, 100, 1.0),
    Case('test_scene.test_multiple_geometries2', multiple_geometries, lambda s: s.geometry[0].disable(), P,
         200, 100, 1.0),
    Case('test_scene.test_multiple_geometries3', multiple_geometries,
         _chain(lambda s: s.geometry[0].disable(), lambda s: s.geometry[0].enable()), P, 200, 100, 1.0),
    Case('test_scene.test_multiple_geometries4', multiple_geometries, lambda s: s.geometry[1].remove(), P,
         200, 100, 1.0),
    # tracers, fresnel/test/test_tracer_direct.py
    Case('test_tracer_direct.test_render', hex_sphere, _nothing, P, 100, 100, 1.0),
    Case('test_tracer_direct.test_render_aa', hex_sphere, _nothing, 'preview_aa', 100, 100, 1.0),
    # perspective camera, fresnel/test/test_camera_perspective.py (helpers use tolerance 2)
    Case('test_camera_perspective.test_render', perspective_grid, _nothing, P, 150, 100, 2.0),
    Case('test_camera_perspective.test_focal_length', perspective_grid, _cam('focal_length', 0.25), P, 150,
         100, 2.0),
    Case('test_camera_perspective.test_height', perspective_grid, _cam('height', .125), P, 150, 100, 2.0),
    Case('test_camera_perspective.test_pathtrace', perspective_grid, _nothing, PT, 150, 100, 2.0),
    Case('test_camera_perspective.test_f_stop', perspective_grid, _cam('f_stop', 0.2), PT, 150, 100, 2.0),
    Case('test_camera_perspective.test_focal_distance', perspective_grid,
         _chain(_cam('f_stop', 0.2), _cam('focus_distance', 6)), PT, 150, 100, 2.0),
]


def render_case(case, device):
    """Build, change and render one case; returns the (h, w, 4) uint8 image."""
    scene = case.builder(device)
    case.change(scene)
    if case.render == P:
        return fresnel.preview(scene, w=case.w, h=case.h, anti_alias=False)[:]
    if case.render == 'preview_aa':
        return fresnel.preview(scene, w=case.w, h=case.h, anti_alias=True)[:]
    return fresnel.pathtrace(scene, w=case.w, h=case.h, samples=64 * 40)[:]
