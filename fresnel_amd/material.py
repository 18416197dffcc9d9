"""Materials: how light interacts with a geometry's surface.

Reference: fresnel/material.py:9-332.  The BRDF itself is the principled
Disney model evaluated on the device (fresnel/common/Material.h:47-314).
"""

from . import _common

_FIELDS = ('solid', 'primitive_color_mix', 'roughness', 'specular', 'spec_trans', 'metal')


class Material(object):
    """Material properties.

    Args:
        solid (float): 1 renders a flat color regardless of lighting.
        color ((3,) float32): linear material color.
        primitive_color_mix (float): 0 uses ``color``, 1 uses the per-primitive
            geometry color, values between blend the two.
        roughness (float): surface roughness, nominally in [0.1, 1].
        specular (float): strength of specular highlights, nominally [0, 1].
        spec_trans (float): fraction of specular transmission in [0, 1].
        metal (float): 0 for dielectrics, 1 for metals.

    Colors are linear; convert sRGB inputs with `fresnel_amd.color.linear`.
    """

    def __init__(self,
                 solid=0,
                 color=(0.9, 0.9, 0.9),
                 primitive_color_mix=0,
                 roughness=0.3,
                 specular=0.5,
                 spec_trans=0,
                 metal=0):
        self._material = _common.Material()
        self.solid = solid
        self.color = color
        self.roughness = roughness
        self.specular = specular
        self.metal = metal
        self.spec_trans = spec_trans
        self.primitive_color_mix = primitive_color_mix

    def _get_cpp_material(self):
        return self._material


def _scalar_property(name):

    def getter(self):
        return getattr(self._get_cpp_material(), name)

    def setter(self, value):
        m = self._get_cpp_material()
        setattr(m, name, float(value))
        self._set_cpp_material(m)

    return property(getter, setter)


def _color_getter(self):
    c = self._get_cpp_material().color
    return (c.r, c.g, c.b)


def _color_setter(self, value):
    if len(value) != 3:
        raise ValueError("colors must have length 3")
    m = self._get_cpp_material()
    m.color = _common.RGBf(*value)
    self._set_cpp_material(m)


Material._set_cpp_material = lambda self, m: None
for _name in _FIELDS:
    setattr(Material, _name, _scalar_property(_name))
Material.color = property(_color_getter, _color_setter)


class _MaterialProxy(object):
    """Material view bound to a geometry: every write goes to the geometry."""

    def __init__(self, geometry):
        self._geometry = geometry._geometry

    def _get_cpp_material(self):
        return self._geometry.getMaterial()

    def _set_cpp_material(self, m):
        self._geometry.setMaterial(m)


class _OutlineMaterialProxy(object):
    """Outline material view bound to a geometry."""

    def __init__(self, geometry):
        self._geometry = geometry._geometry

    def _get_cpp_material(self):
        return self._geometry.getOutlineMaterial()

    def _set_cpp_material(self, m):
        self._geometry.setOutlineMaterial(m)


for _cls in (_MaterialProxy, _OutlineMaterialProxy):
    for _name in _FIELDS:
        setattr(_cls, _name, _scalar_property(_name))
    _cls.color = property(_color_getter, _color_setter)
