"""Directional area lights and preset lighting setups.

Reference: fresnel/light.py:20-267.  Lights live at infinity; ``direction``
points toward the light in camera space, ``theta`` is the half angle of the
light's cone and ``color`` its linear intensity.
"""

import math

from . import _common


class Light(object):
    """A directional area light.

    Args:
        direction ((3,) float32): direction toward the light, camera space.
        color ((3,) float32): linear color and intensity.
        theta (float): half angle of the light's cone, radians.
    """

    def __init__(self, direction, color=(1, 1, 1), theta=0.375):
        self.direction = tuple(direction)
        self.color = tuple(color)
        self.theta = float(theta)

    def __str__(self):
        """Human readable representation."""
        return f"<Light object with direction {self.direction} and color {self.color}>"


class _LightProxy(object):
    """One entry of a scene's light list; writes go to the scene."""

    def __init__(self, light_list, idx):
        self._light_list = light_list
        self._idx = idx

    @property
    def direction(self):
        v = self._light_list._lights.getDirection(self._idx)
        return (v.x, v.y, v.z)

    @direction.setter
    def direction(self, v):
        self._light_list._lights.setDirection(self._idx, _common.vec3f(*v))

    @property
    def color(self):
        c = self._light_list._lights.getColor(self._idx)
        return (c.r, c.g, c.b)

    @color.setter
    def color(self, v):
        self._light_list._lights.setColor(self._idx, _common.RGBf(*v))

    @property
    def theta(self):
        return self._light_list._lights.getTheta(self._idx)

    @theta.setter
    def theta(self, v):
        self._light_list._lights.setTheta(self._idx, v)

    def __str__(self):
        return f"<Light proxy object with direction {self.direction} and color {self.color}>"


class _LightListProxy(object):
    """Up to 4 lights stored in a `_common.Lights` record."""

    def __init__(self, _lights=None):
        if _lights is None:
            self._lights = _common.Lights()
            self._lights.N = 0
        else:
            self._lights = _lights

    def __len__(self):
        return self._lights.N

    def clear(self):
        self._lights.N = 0

    def append(self, light):
        if len(self) >= 4:
            raise IndexError("Cannot add more than 4 lights")
        i = self._lights.N
        self._lights.setDirection(i, _common.vec3f(*light.direction))
        self._lights.setColor(i, _common.RGBf(*light.color))
        self._lights.setTheta(i, light.theta)
        self._lights.N = i + 1

    def __getitem__(self, idx):
        if idx >= self._lights.N:
            raise IndexError("Indexing past the end of the list")
        return _LightProxy(self, idx)


def _spherical(polar_deg, azimuth_deg):
    """Unit vector from a polar angle off +y and an azimuth about +y from +z."""
    th = polar_deg * math.pi / 180
    ph = azimuth_deg * math.pi / 180
    return (math.sin(th) * math.sin(ph), math.cos(th), math.sin(th) * math.cos(ph))


def butterfly():
    """Key light high above the camera, fill light below it."""
    key = 50 * math.pi / 180
    fill = -30 * math.pi / 180
    return [
        Light(direction=(0, math.sin(key), math.cos(key)), color=(1.0, 1.0, 1.0), theta=math.pi / 4),
        Light(direction=(0, math.sin(fill), math.cos(fill)), color=(0.1, 0.1, 0.1), theta=math.pi / 2),
    ]


def loop(side='right'):
    """Key light slightly to one side and up, fill light level on the other."""
    sign = {'right': 1, 'left': -1}[side]
    return [
        Light(direction=_spherical(90 - 20, sign * 25), color=(1.0, 1.0, 1.0), theta=math.pi / 4),
        Light(direction=_spherical(90, -sign * 40), color=(0.1, 0.1, 0.1), theta=math.pi / 2),
    ]


def rembrandt(side='right'):
    """Key light 45 degrees to one side and up, fill light level on the other."""
    sign = {'right': 1, 'left': -1}[side]
    return [
        Light(direction=_spherical(90 - 20, sign * 45), color=(1.0, 1.0, 1.0), theta=math.pi / 4),
        Light(direction=_spherical(90, -sign * 45), color=(0.1, 0.1, 0.1), theta=math.pi / 2),
    ]


def lightbox():
    """One area light covering every direction in front of the scene."""
    return [Light(direction=(0, 0, 1), color=(0.9, 0.9, 0.9), theta=math.pi)]


def cloudy():
    """Bright light from the whole upper hemisphere, dim light from below."""
    return [
        Light(direction=(0, 1, 0), color=(0.9, 0.9, 0.9), theta=math.pi / 2),
        Light(direction=(0, -1, 0), color=(0.1, 0.1, 0.1), theta=math.pi / 2),
    ]


def ring():
    """A strong frontal area light."""
    return [Light(direction=(0, 0, 1), color=(0.9, 0.9, 0.9), theta=math.pi / 4)]
