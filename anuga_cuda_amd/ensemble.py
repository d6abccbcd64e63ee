"""Ensembles of independent shallow-water domains on one device.

An ``Ensemble`` puts the meshes of several ``Domain`` objects into one device
handle (``sw_create_ensemble``): every fused kernel launch covers all of them
and one hipGraph replays the step for the whole ensemble, while each instance
keeps its own CFL timestep, time, yields, first-order fallback and boundary
values.  Each domain stays usable as usual for setting quantities before the
run and reading them after it (``domain.quantities[...]``).

    ens = Ensemble([domain_a, domain_b, ...])
    for t in ens.evolve(yieldstep=10.0, finaltime=60.0):   # common yield times
        ...
    ens.evolve_steps(300)                                  # or a fixed number of steps
"""
from __future__ import annotations

from . import _lib
from .domain import Domain

__all__ = ["Ensemble"]


class Ensemble:
    def __init__(self, domains, device=0):
        self.domains = list(domains)
        if not self.domains:
            raise ValueError("an ensemble needs at least one domain")
        d0 = self.domains[0]
        for d in self.domains:
            if not isinstance(d, Domain):
                raise TypeError("ensemble members must be Domain objects")
            if d._handle is not None:
                raise RuntimeError("ensemble members must not have device storage yet")
            if d._params != d0._params or d.precision != d0.precision:
                raise ValueError("ensemble members must share their numerical parameters and precision")
            if bytes(d._sw_params()) != bytes(d0._sw_params()):
                raise ValueError("ensemble members must share their timestepping method and limiter settings")
            if d.flow_algorithm == "tsunami":
                raise NotImplementedError("the 'tsunami' flow algorithm has no evolve step")
        self.handle = _lib.Handle.ensemble([d.mesh for d in self.domains], d0._sw_params(), device)
        for i, d in enumerate(self.domains):
            d._attach(self.handle.view(i))
        self.number_of_elements = sum(d.number_of_elements for d in self.domains)

    def __len__(self):
        return len(self.domains)

    def _sync(self):
        for d in self.domains:
            d._host_stale = True
            d._sync_scalars()

    def _push_all(self):
        for d in self.domains:
            d._dev()

    def prepare(self, yieldstep=None, finaltime=None):
        """The evolve preamble of every instance (distribute, boundary values).
        The members share one yield schedule, so they must share a start
        time."""
        t0 = self.domains[0].starttime
        if any(d.starttime != t0 for d in self.domains):
            raise ValueError("ensemble members must share their start time (got %s)"
                             % sorted({d.starttime for d in self.domains}))
        self._push_all()
        if yieldstep is None:
            yieldstep = self.domains[0].evolve_max_timestep
        for d in self.domains:
            d.finaltime = finaltime
            s = d._handle.get_state()
            s.starttime = d.starttime
            d._handle.set_state(s)
        self.handle.call("sw_evolve_prepare", float(yieldstep),
                         float(finaltime) if finaltime is not None else 0.0, int(finaltime is not None))
        self._sync()

    def evolve_steps(self, steps):
        """Run up to ``steps`` Euler steps of every instance (each stops at its
        own yield or final time).  Returns (largest step count, largest event)."""
        self._push_all()
        done, event = self.handle.evolve_steps(int(steps))
        self._sync()
        return done, event

    def autotune(self, steps=3):
        """Choose the fused kernels' workgroup sizes by timing ``steps`` real
        steps per candidate (sw_autotune).  Call after ``prepare``; the tuning
        steps advance every member."""
        self._push_all()
        chosen = self.handle.autotune(steps)
        self._sync()
        return chosen

    def evolve(self, yieldstep, finaltime):
        """Generator over the common yield times: every instance reaches each
        yield (with its own timestep sequence) before the ensemble yields."""
        self.prepare(yieldstep, finaltime)
        yield self.domains[0].get_time()
        while True:
            _, event = self.evolve_steps(1 << 40)
            t = self.domains[0].get_time()
            if event == _lib.SW_EVENT_FINAL:
                yield t
                break
            yield t
            self.handle.call("sw_advance_yield")
            self._sync()

    def close(self):
        for d in self.domains:
            if d._handle is not None:
                d._pull()
                d._handle = None
        self.handle.close()
