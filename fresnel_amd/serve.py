"""Render server: many Path.sample requests advanced together on one GPU.

No counterpart in the reference, which renders one (scene, tracer) pair per
call.  A request is ``(scene, tracer, samples, light_samples)``; its result,
left in ``tracer.output`` / ``tracer.linear_output``, is bit-identical to
``tracer.sample(scene, samples, light_samples=light_samples)``.

Each `Server.step` admits queued requests first-come first-served while their
HBM footprints fit the budget, plans a wavefront batch in which every active
request advances by some camera samples (an equal share of ``target_paths``
light-sample paths, at least one sample) when none is waiting, keeps up to two
batches in flight on the device (one's kernels run while the host serves the
other), and retires the requests whose samples have all been accumulated.
"""

QUEUED, ACTIVE, DONE, REJECTED = 0, 1, 2, 3


class Request:
    """Handle of a submitted request."""

    def __init__(self, server, rid, tracer):
        self._server = server
        self.id = rid
        self.tracer = tracer

    def poll(self):
        """Return (state, samples_done, (t_arrival, t_admit, t_done))."""
        return self._server._srv.poll(self.id)

    @property
    def state(self):
        return self.poll()[0]

    @property
    def done(self):
        return self.state >= DONE

    def latency(self):
        """Seconds from arrival to completion (None until done)."""
        st, _, (ta, _, td) = self.poll()
        return td - ta if st == DONE else None


class Server:
    """Render server on a GPU `fresnel_amd.Device`.

    Args:
        device: a GPU device.
        hbm_budget (int): bytes of HBM the server may hold: the active
            requests' footprints plus its two render scratches (about
            2 x 160 B x target_paths, 86 GB at the default 256 Mi) plus any
            render scratch the device keeps for Path.sample (released when
            it would keep a queued request out).
        target_paths (int): light-sample paths traced per step (the batch
            shape of Path.sample).
    """

    def __init__(self, device, hbm_budget=240 << 30, target_paths=256 << 20):
        if device.mode != 'gpu':
            raise RuntimeError("the render server needs a GPU device")
        self.device = device
        self._srv = device.module.Server(device._device, int(hbm_budget), int(target_paths))

    def submit(self, scene, tracer, samples, light_samples=1, now=-1.0):
        """Queue a Path.sample request; `tracer` must be a fresnel_amd.tracer.Path."""
        rid = self._srv.submit(scene._scene, tracer._tracer, int(samples), int(light_samples), float(now))
        return Request(self, rid, tracer)

    def step(self, now=-1.0):
        """One iteration; returns (n_active, n_queued) after it."""
        return self._srv.step(float(now))

    def run(self):
        """Step until no request is queued or active."""
        while True:
            a, q = self.step()
            if a == 0 and q == 0:
                return

    def memory(self):
        """(bytes held by active requests, budget)."""
        return self._srv.memory()

    def release(self, request):
        """Forget a finished request."""
        self._srv.release(request.id)
