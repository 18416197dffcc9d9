"""Tracers render a `Scene` into an output image (reference: fresnel/tracer.py).

* `Preview`: fast direct lighting with anti-aliasing.
* `Path`: Monte Carlo path tracing with soft shadows, reflections,
  transmission and depth of field.
"""

import numpy

from . import _common
from . import util


class Tracer(object):
    """Base class of the tracers: output buffers, seed, highlight warning."""

    def __init__(self):
        raise RuntimeError("Use a specific tracer class")

    def resize(self, w, h):
        """Resize (and clear) the output buffer to ``w`` x ``h`` pixels."""
        self._tracer.resize(w, h)

    def render(self, scene):
        """Render ``scene`` into the output buffer and return `output`."""
        self._tracer.render(scene._scene)
        return self.output

    def enable_highlight_warning(self, color=(1, 0, 1)):
        """Paint pixels brighter than 1 in any channel with ``color``."""
        self._tracer.enableHighlightWarning(_common.RGBf(*color))

    def disable_highlight_warning(self):
        """Stop flagging over-bright pixels."""
        self._tracer.disableHighlightWarning()

    def histogram(self):
        """Lightness and R, G, B histograms of the visible pixels.

        Returns:
            (histogram, bin_positions): histogram is (512, 4) with columns
            lightness (0.21 R + 0.72 G + 0.07 B), R, G, B, each taken to the
            power 1/2.2 and binned over [0, 1].
        """
        if hasattr(self._tracer, 'histogram'):
            # computed on the device from the resident linear output
            counts = numpy.asarray(self._tracer.histogram()).astype(numpy.int64)
            return counts, numpy.linspace(0, 1, 513, dtype=numpy.float32)[1:]
        a = numpy.array(self.linear_output[:])
        img = a[a[:, :, 3] > 0]
        r, g, b = img[:, 0], img[:, 1], img[:, 2]
        lightness = 0.21 * r + 0.72 * g + 0.07 * b
        n = 512
        cols = []
        for channel in (lightness, r, g, b):
            hist, bins = numpy.histogram(channel**(1 / 2.2), bins=n, range=[0, 1])
            cols.append(hist)
        return numpy.stack(cols, axis=1), bins[1:]

    @property
    def output(self):
        """ImageArray: the (h, w, 4) sRGB uint8 output."""
        return util.ImageArray(self._tracer.getSRGBOutputBuffer(), geom=None)

    def output_async(self, linear=False):
        """Start reading the output back without waiting (GPU devices).

        The copy runs on the device's copy engine while the host continues;
        a later render of this tracer traces concurrently with it and writes
        the output only after it.  Returns an `AsyncOutput`.
        """
        which = 1 if linear else 0
        self._tracer.outputAsync(which)
        return AsyncOutput(self, which)

    @property
    def linear_output(self):
        """Array: the (h, w, 4) linear float32 output."""
        return util.Array(self._tracer.getLinearOutputBuffer(), geom=None)

    @property
    def seed(self):
        """int: random number seed."""
        return self._tracer.getSeed()

    @seed.setter
    def seed(self, value):
        self._tracer.setSeed(value)


class AsyncOutput(object):
    """A readback started by `Tracer.output_async`."""

    def __init__(self, tracer, which):
        self._tracer = tracer
        self._which = which

    def ready(self):
        """bool: whether the copy has completed."""
        return self._tracer._tracer.outputReady(self._which)

    def wait(self):
        """The (h, w, 4) image (uint8 sRGB or float32 linear) as rendered when the copy
        started: a view of the tracer's host buffer, valid until its next
        `output_async` of the same output."""
        return self._tracer._tracer.outputWait(self._which)


class Preview(Tracer):
    """Direct-lighting preview tracer.

    Args:
        device (Device): device to render on.
        w (int): output width.
        h (int): output height.
        anti_alias (bool): 8x8 jittered samples per pixel when True.
    """

    def __init__(self, device, w, h, anti_alias=True):
        self.device = device
        self._tracer = device.module.TracerDirect(device._device, w, h)
        self.anti_alias = anti_alias

    @property
    def anti_alias(self):
        """bool: whether anti-aliasing is on."""
        return self._tracer.getAntialiasingN() > 1

    @anti_alias.setter
    def anti_alias(self, value):
        self._tracer.setAntialiasingN(8 if value else 1)


class Path(Tracer):
    """Monte Carlo path tracer.

    Args:
        device (Device): device to render on.
        w (int): output width.
        h (int): output height.

    Each `render` adds one sample per pixel to the running mean in `output`.
    """

    def __init__(self, device, w, h):
        self.device = device
        self._tracer = device.module.TracerPath(device._device, w, h, 1)

    def reset(self):
        """Clear the output and advance the seed so the next image is independent."""
        self._tracer.reset()

    def sample(self, scene, samples, reset=True, light_samples=1):
        """Take ``samples`` camera samples per pixel.

        Args:
            scene (Scene): scene to render.
            samples (int): camera samples per pixel.
            reset (bool): call `reset` first.
            light_samples (int): paths traced from each primary hit.

        Returns:
            ImageArray: the `output` buffer.

        The result equals ``samples`` calls of `render`; a device module that
        offers ``renderSamples`` receives all passes in one submission.
        """
        if reset:
            self.reset()
        self._tracer.setLightSamples(light_samples)
        if samples > 0:
            if hasattr(self._tracer, 'renderSamples'):
                self._tracer.renderSamples(scene._scene, samples)
            else:
                for _ in range(samples):
                    self._tracer.render(scene._scene)
        self._tracer.setLightSamples(1)
        return self.output
