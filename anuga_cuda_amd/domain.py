"""ANUGA-compatible shallow-water ``Domain`` running its step on the MI355X.

Mirrors the user-facing API of the reference's device-resident domain
(``src/anuga_cuda/gpu_domain_advanced.py``, itself a subclass of ANUGA's
``Domain``): construction from points / triangles / boundary tags,
``set_quantity``, ``set_boundary`` with ``Reflective_boundary``,
``Dirichlet_boundary`` and ``Transmissive_boundary``, the step operators under
their ANUGA names, and ``evolve(yieldstep, finaltime, duration,
skip_initial_step)`` as a generator.  Every operator runs as HIP kernels
through the C ABI (``include/anuga_sw.h``); host arrays are numpy copies that
are synchronised with the device on access.

Default parameters are ANUGA's shallow-water defaults as the reference passes
them to its kernels (g, epsilon, H0 = minimum_allowed_height, beta_*,
alpha_balance, tight_slope_limiters, use_centroid_velocities,
optimise_dry_cells, extrapolate_velocity_second_order, CFL, evolve_max/min
timestep, max_smallsteps, default order 2); ANUGA itself is not vendored by the
reference, so these values are stated here and every one can be overridden.
"""
from __future__ import annotations

import os

import numpy as np

from . import _lib
from ._lib import (SW_BOUNDARY, SW_CENTROID, SW_EDGE, SW_ELEVATION, SW_EXPLICIT_UPDATE,
                   SW_FRICTION, SW_HEIGHT, SW_SEMI_IMPLICIT_UPDATE, SW_STAGE, SW_VERTEX,
                   SW_XMOM, SW_XVELOCITY, SW_YMOM, SW_YVELOCITY)
from .mesh import Mesh

__all__ = [
    "Domain",
    "Quantity",
    "Reflective_boundary",
    "Dirichlet_boundary",
    "Transmissive_boundary",
    "DEFAULT_PARAMETERS",
    "create_domain_from_file",
]

DEFAULT_PARAMETERS = dict(
    g=9.8,
    epsilon=1.0e-12,
    H0=1.0e-3,
    minimum_allowed_height=1.0e-3,
    maximum_allowed_speed=0.0,
    beta_w=0.9,
    beta_w_dry=0.9,
    beta_uh=0.9,
    beta_uh_dry=0.9,
    beta_vh=0.9,
    beta_vh_dry=0.9,
    alpha_balance=2.0,
    CFL=1.0,
    evolve_max_timestep=1.0e100,
    evolve_min_timestep=1.0e-6,
    optimise_dry_cells=1,
    extrapolate_velocity_second_order=1,
    tight_slope_limiters=1,
    use_centroid_velocities=1,
    use_sloped_mannings=0,
    max_smallsteps=50,
    default_order=2,
    # distribute path (gpu_domain_advanced.py:1814-1990)
    use_edge_limiter=0,
    optimised_gradient_limiter=1,
)

# parameters that select the distribute path rather than map to a sw_params field
_LIMITER_PARAMS = ("use_edge_limiter", "optimised_gradient_limiter")

_QID = {
    "stage": SW_STAGE,
    "xmomentum": SW_XMOM,
    "ymomentum": SW_YMOM,
    "elevation": SW_ELEVATION,
    "friction": SW_FRICTION,
    "height": SW_HEIGHT,
    "xvelocity": SW_XVELOCITY,
    "yvelocity": SW_YVELOCITY,
}
CONSERVED = ("stage", "xmomentum", "ymomentum")
# timestepping methods and their code in sw_params.timestepping_method
TIMESTEPPING = {"euler": 1, "rk2": 2, "rk3": 3}
_LOC = {
    "centroid_values": SW_CENTROID,
    "edge_values": SW_EDGE,
    "vertex_values": SW_VERTEX,
    "boundary_values": SW_BOUNDARY,
    "explicit_update": SW_EXPLICIT_UPDATE,
    "semi_implicit_update": SW_SEMI_IMPLICIT_UPDATE,
}


# ---------------------------------------------------------------------------
# Boundary objects (ANUGA generic_boundary_conditions; kernels in the
# reference: src/anuga_cuda/evaluate_segment/evaluate_segment.cu:3-119)
class Reflective_boundary:
    """Solid wall: stage copied, normal momentum negated."""

    kind = _lib.SW_BC_REFLECTIVE

    def __init__(self, domain=None):
        self.domain = domain

    def __repr__(self):
        return "Reflective_boundary"


class Transmissive_boundary:
    """Boundary state equals the inside edge state."""

    kind = _lib.SW_BC_TRANSMISSIVE

    def __init__(self, domain=None):
        self.domain = domain

    def __repr__(self):
        return "Transmissive_boundary"


class Dirichlet_boundary:
    """Fixed (stage, xmomentum, ymomentum) on the boundary."""

    kind = _lib.SW_BC_DIRICHLET

    def __init__(self, dirichlet_values=None):
        if dirichlet_values is None:
            raise ValueError("Dirichlet_boundary needs [stage, xmomentum, ymomentum]")
        v = np.asarray(dirichlet_values, dtype=np.float64).ravel()
        if v.size < 3:
            raise ValueError("Dirichlet_boundary needs at least 3 values")
        self.dirichlet_values = v

    def __repr__(self):
        return "Dirichlet_boundary(%s)" % self.dirichlet_values[:3].tolist()


# ---------------------------------------------------------------------------
class Quantity:
    """Values of one quantity at centroids, edges, vertices and boundary edges.

    Array attributes are numpy arrays synchronised with the device: reading one
    after a device operation fetches the current device values; arrays read or
    written on the host are sent back before the next device operation.
    """

    def __init__(self, domain, name):
        self.domain = domain
        self.name = name
        # limiter coefficient of the per-quantity limiters (use_edge_limiter or
        # optimised_gradient_limiter False); ANUGA's Quantity default
        self._beta = 1.0
        N = domain.number_of_elements
        Nb = domain.number_of_boundary_elements
        self._arrays = {
            "centroid_values": np.zeros(N),
            "edge_values": np.zeros((N, 3)),
            "vertex_values": np.zeros((N, 3)),
            "boundary_values": np.zeros(Nb),
        }
        if name in CONSERVED:
            self._arrays["explicit_update"] = np.zeros(N)
            self._arrays["semi_implicit_update"] = np.zeros(N)
        if name == "friction":
            self._arrays = {"centroid_values": np.zeros(N), "vertex_values": np.zeros((N, 3)),
                            "edge_values": np.zeros((N, 3))}

    @property
    def beta(self):
        return self._beta

    @beta.setter
    def beta(self, value):
        self._beta = float(value)
        self.domain._params_changed()

    def set_beta(self, beta):
        self.beta = beta

    def _get(self, key):
        d = self.domain
        d._pull_one(self.name, key)
        arr = self._arrays[key]
        if (self.name, key) not in d._touched:
            d._touched[(self.name, key)] = arr.copy()  # what it held when first handed out
        return arr

    def _set(self, key, value):
        arr = self._get(key)
        arr[...] = value

    centroid_values = property(lambda s: s._get("centroid_values"), lambda s, v: s._set("centroid_values", v))
    edge_values = property(lambda s: s._get("edge_values"), lambda s, v: s._set("edge_values", v))
    vertex_values = property(lambda s: s._get("vertex_values"), lambda s, v: s._set("vertex_values", v))
    boundary_values = property(lambda s: s._get("boundary_values"), lambda s, v: s._set("boundary_values", v))
    explicit_update = property(lambda s: s._get("explicit_update"), lambda s, v: s._set("explicit_update", v))
    semi_implicit_update = property(lambda s: s._get("semi_implicit_update"),
                                    lambda s, v: s._set("semi_implicit_update", v))

    # quantity.c:662-684 (_interpolate): centroid = (q0+q1+q2)/3, edge i = mean of the other two
    def interpolate(self):
        v = self.vertex_values
        self.centroid_values[:] = (v[:, 0] + v[:, 1] + v[:, 2]) / 3.0
        self.interpolate_from_vertices_to_edges()

    def interpolate_from_vertices_to_edges(self):
        v = self.vertex_values
        e = self.edge_values
        e[:, 0] = 0.5 * (v[:, 1] + v[:, 2])
        e[:, 1] = 0.5 * (v[:, 0] + v[:, 2])
        e[:, 2] = 0.5 * (v[:, 0] + v[:, 1])

    def extrapolate_first_order(self):
        c = self.centroid_values
        self.edge_values[:] = c[:, None]
        self.vertex_values[:] = c[:, None]

    def set_values(self, numeric=None, location="vertices", indices=None):
        """Constant, array or callable f(x, y) at 'vertices' (then interpolated
        to centroids and edges, ANUGA's default) or 'centroids' (then first
        order extrapolated)."""
        d = self.domain
        if location not in ("vertices", "centroids", "edges"):
            raise ValueError("unknown location %r" % location)
        if indices is not None:
            indices = np.asarray(indices, dtype=np.int64)
        if location == "vertices":
            if callable(numeric):
                xy = d.mesh.vertex_coordinates.reshape(-1, 2)
                vals = np.asarray(numeric(xy[:, 0], xy[:, 1]), dtype=np.float64).reshape(-1, 3)
            else:
                vals = np.broadcast_to(np.asarray(numeric, dtype=np.float64),
                                       (d.number_of_elements, 3)) if np.ndim(numeric) < 2 else \
                    np.asarray(numeric, dtype=np.float64).reshape(-1, 3)
            if indices is None:
                self.vertex_values[:] = vals
            else:
                self.vertex_values[indices] = vals[indices] if vals.shape[0] == d.number_of_elements else vals
            self.interpolate()
        elif location == "centroids":
            if callable(numeric):
                c = d.mesh.centroid_coordinates
                vals = np.asarray(numeric(c[:, 0], c[:, 1]), dtype=np.float64)
            else:
                vals = np.broadcast_to(np.asarray(numeric, dtype=np.float64), (d.number_of_elements,))
            if indices is None:
                self.centroid_values[:] = vals
            else:
                self.centroid_values[indices] = vals[indices] if vals.shape[0] == d.number_of_elements else vals
            self.extrapolate_first_order()
        else:
            if callable(numeric):
                xy = d.mesh.edge_midpoint_coordinates.reshape(-1, 2)
                vals = np.asarray(numeric(xy[:, 0], xy[:, 1]), dtype=np.float64).reshape(-1, 3)
            else:
                vals = np.broadcast_to(np.asarray(numeric, dtype=np.float64), (d.number_of_elements, 3))
            self.edge_values[:] = vals

    def get_integral(self):
        return float(np.sum(self.centroid_values * self.domain.mesh.areas))


# ---------------------------------------------------------------------------
class Domain:
    """Shallow-water domain on an unstructured triangle mesh.

    ``Domain(coordinates, vertices, boundary)`` as in ANUGA, or
    ``Domain(mesh=Mesh(...))``.  Keyword parameters override
    ``DEFAULT_PARAMETERS``; ``precision='mixed'`` selects FP32 flux and
    reconstruction with FP64 accumulation, timestep and wet/dry tests.
    ``locality_order='hilbert'`` (default) lets the library renumber the
    triangles on the device along a Hilbert curve of their centroids;
    ``'keep'`` uses the caller's numbering.  Results and the arrays seen here
    do not depend on it.
    """

    def __init__(self, coordinates=None, vertices=None, boundary=None, mesh=None,
                 precision="fp64", device=0, locality_order="hilbert", **parameters):
        if mesh is None:
            if coordinates is None or vertices is None:
                raise ValueError("need coordinates and vertices, or mesh=")
            mesh = Mesh(coordinates, vertices, boundary=boundary)
        self.mesh = mesh
        self.number_of_elements = mesh.number_of_elements
        self.number_of_boundary_elements = mesh.number_of_boundary_elements
        self.boundary = mesh.boundary
        unknown = set(parameters) - set(DEFAULT_PARAMETERS)
        if unknown:
            raise TypeError("unknown parameters: %s" % sorted(unknown))
        self._params = dict(DEFAULT_PARAMETERS)
        self._params.update(parameters)
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.precision = precision
        if locality_order not in ("hilbert", "keep"):
            raise ValueError("locality_order must be 'hilbert' or 'keep'")
        self.locality_order = locality_order
        self.device = device
        self.quantities = {name: Quantity(self, name) for name in _QID}
        self.conserved_quantities = list(CONSERVED)
        self.evolved_quantities = ["stage", "xmomentum", "ymomentum", "elevation",
                                   "height", "xvelocity", "yvelocity"]
        self.other_quantities = ["elevation", "friction", "height", "xvelocity", "yvelocity"]
        self.boundary_map = None
        self._kinds = np.zeros(self.number_of_boundary_elements, dtype=np.int32)
        self._dirichlet = np.zeros((self.number_of_boundary_elements, 3))
        self._handle = None
        self._stale = set()
        self._touched = {}
        self._all_dirty = True
        self._time = 0.0
        self.starttime = 0.0
        self.finaltime = None
        self.yieldtime = None
        self.timestep = 0.0
        self.flux_timestep = 0.0
        self.number_of_steps = 0
        self.number_of_first_order_steps = 0
        self.recorded_max_timestep = self._params["evolve_min_timestep"]
        self.recorded_min_timestep = self._params["evolve_max_timestep"]
        self._order_ = self._params["default_order"]
        self.smallsteps = 0
        # any name but 'tsunami' selects the default flow, as the reference's
        # dispatch does (gpu_domain_advanced.py:1014, utility.py:267-272)
        self.flow_algorithm = "original"
        # SWW output (ANUGA's names; nothing is written unless store is set)
        self.name = "domain"
        self.datadir = "."
        self.store = False
        self.smooth = True
        self.minimum_storable_height = 1.0e-5
        self.writer = None
        self.quantities_to_be_stored = {"elevation": 1, "stage": 2, "xmomentum": 2, "ymomentum": 2}
        # coordinates are relative to this origin (ANUGA's Geo_reference)
        self.geo_reference = {"xllcorner": 0.0, "yllcorner": 0.0, "zone": -1}
        self._max_speed = np.zeros(self.number_of_elements)

    # ------------------------------------------------------------------ params
    def __getattr__(self, name):
        params = self.__dict__.get("_params")
        if params is not None and name in params:
            return params[name]
        raise AttributeError(name)

    def __setattr__(self, name, value):
        params = self.__dict__.get("_params")
        if params is not None and name in params:
            params[name] = value
            if self.__dict__.get("_handle") is not None:
                self._handle.set_params(self._sw_params())
            return
        object.__setattr__(self, name, value)

    def __len__(self):
        return self.number_of_elements

    def set_default_order(self, n):
        if n not in (1, 2):
            raise ValueError("order must be 1 or 2")
        self.default_order = n
        self._order_ = n

    def set_CFL(self, cfl):
        self.CFL = float(cfl)

    def set_minimum_allowed_height(self, h):
        self.minimum_allowed_height = float(h)
        self.H0 = float(h)

    def get_time(self):
        if self._handle is not None:
            return self._handle.get_state().time
        return self._time

    def set_time(self, t):
        t = float(t)
        if self._handle is not None:
            s = self._handle.get_state()
            s.time = t
            self._handle.set_state(s)
        self._time = t

    def get_starttime(self):
        return self.starttime

    def set_starttime(self, t):
        self.starttime = float(t)
        self.set_time(t)

    def get_quantity(self, name):
        return self.quantities[name]

    def get_timestepping_method(self):
        return self.__dict__.get("_timestepping", "euler")

    def set_timestepping_method(self, method):
        """'euler' (default), 'rk2' or 'rk3' (ANUGA's set_timestepping_method;
        the reference drivers dispatch on it, gpu_domain_advanced.py:2566-2573)."""
        if method in (1, 2, 3):  # ANUGA also accepts the method's number
            method = list(TIMESTEPPING)[method - 1]
        if method not in TIMESTEPPING:
            raise ValueError("timestepping method must be one of %s" % ", ".join(TIMESTEPPING))
        self._timestepping = method
        if self._handle is not None:
            self._handle.set_params(self._sw_params())

    def set_quantities_to_be_stored(self, q):
        """ANUGA's set_quantities_to_be_stored: None stops storing; otherwise a
        dict of quantity name -> 1 (static, stored once) or 2 (every yield).
        The SWW path stores elevation (static) and any of stage, xmomentum,
        ymomentum (every yield); other quantities raise."""
        if q is None:
            self.store = False
            return
        allowed = {"elevation": (1,), "stage": (2,), "xmomentum": (2,), "ymomentum": (2,)}
        for name, flag in dict(q).items():
            if name not in allowed or flag not in allowed[name]:
                raise NotImplementedError(
                    "the SWW output stores elevation (1) and stage / xmomentum / ymomentum (2); got %r: %r"
                    % (name, flag))
        self.quantities_to_be_stored = dict(q)

    def set_name(self, name):
        self.name = name[:-4] if name.endswith(".sww") else name

    def get_name(self):
        return self.name

    def set_datadir(self, name):
        self.datadir = name

    def get_datadir(self):
        return self.datadir

    def set_store(self, flag=True):
        """Write the SWW file (datadir/name.sww) at every yield of evolve."""
        self.store = bool(flag)

    def set_store_vertices_uniquely(self, flag=True, reduction=None):
        """Store 3 values per triangle instead of one smoothed value per node."""
        self.smooth = not flag

    def set_minimum_storable_height(self, minimum_storable_height):
        self.minimum_storable_height = float(minimum_storable_height)

    def get_minimum_storable_height(self):
        return self.minimum_storable_height

    def initialise_storage(self):
        """Create the SWW file with the mesh and the elevation."""
        from .sww import SWWWriter, sww_path

        if self.writer is not None:
            self.writer.close()
        if self.datadir and not os.path.isdir(self.datadir):
            os.makedirs(self.datadir, exist_ok=True)
        dynamic = [n for n, f in self.quantities_to_be_stored.items() if f == 2]
        self.writer = SWWWriter(self, self._dev(), sww_path(self.datadir, self.name), smooth=self.smooth,
                                minimum_storable_height=self.minimum_storable_height,
                                georeference=self.geo_reference, quantities=dynamic)

    def store_timestep(self):
        """Append the current state to the SWW file (gpu_domain_advanced.py:2167-2172)."""
        if self.writer is None:
            self.initialise_storage()
        self.writer.store_timestep(self._time)

    # ------------------------------------------------------------ quantities
    def set_quantity(self, name, numeric=None, function=None, expression=None,
                     location="vertices", indices=None, **kwargs):
        if name not in self.quantities:
            raise KeyError("unknown quantity %r" % name)
        value = numeric if numeric is not None else function
        if expression is not None:
            env = {n: self._expr_values(n, location) for n in self.quantities}
            env["np"] = np
            value = eval(expression, {"__builtins__": {}}, env)
            if location == "vertices":
                value = np.asarray(value).reshape(-1, 3)
        if value is None:
            raise ValueError("set_quantity needs numeric, function or expression")
        if name == "elevation":
            self._all_dirty = True
        self.quantities[name].set_values(value, location=location, indices=indices)

    def _expr_values(self, name, location):
        q = self.quantities[name]
        return q.vertex_values.copy() if location == "vertices" else q.centroid_values.copy()

    def set_boundary(self, boundary_map):
        """Bind each boundary tag to a boundary object (ANUGA Generic_Domain.set_boundary)."""
        tags = self.mesh.boundary_tags
        kinds = np.zeros(len(tags), dtype=np.int32)
        dirichlet = np.zeros((len(tags), 3))
        for tag in set(tags.tolist()):
            if tag not in boundary_map or boundary_map[tag] is None:
                raise ValueError("Tag '%s' has not been bound to a boundary object" % tag)
        for m, tag in enumerate(tags.tolist()):
            B = boundary_map[tag]
            kinds[m] = B.kind
            if B.kind == _lib.SW_BC_DIRICHLET:
                dirichlet[m] = B.dirichlet_values[:3]
        self.boundary_map = dict(boundary_map)
        self._kinds = kinds
        self._dirichlet = dirichlet
        if self._handle is not None:
            self._handle.set_boundary(kinds, dirichlet)

    # ---------------------------------------------------------- host/device
    def limiter_path(self):
        """'optimised' (extrapolate_second_order_sw), 'vertex' or 'edge' (the
        per-quantity limiters), as the reference's distribute chooses."""
        if self._params["use_edge_limiter"]:
            return "edge"
        return "optimised" if self._params["optimised_gradient_limiter"] else "vertex"

    def set_beta(self, beta):
        """One limiter coefficient for every limiter (ANUGA's set_beta)."""
        for n in ("beta_w", "beta_w_dry", "beta_uh", "beta_uh_dry", "beta_vh", "beta_vh_dry"):
            self._params[n] = float(beta)
        for q in CONSERVED:
            self.quantities[q]._beta = float(beta)
        self._params_changed()

    def _params_changed(self):
        if self.__dict__.get("_handle") is not None:
            self._handle.set_params(self._sw_params())

    def _sw_params(self):
        p = _lib.SwParams()
        for k, v in self._params.items():
            if k not in _LIMITER_PARAMS:
                setattr(p, k, v)
        p.limiter = {"optimised": _lib.SW_LIMITER_OPTIMISED, "vertex": _lib.SW_LIMITER_VERTEX,
                     "edge": _lib.SW_LIMITER_EDGE}[self.limiter_path()]
        for i, q in enumerate(CONSERVED):
            p.quantity_beta[i] = self.quantities[q].beta if "quantities" in self.__dict__ else 1.0
        p.precision = _lib.SW_PRECISION_MIXED if self.precision == "mixed" else _lib.SW_PRECISION_FP64
        p.locality_order = _lib.SW_ORDER_HILBERT if self.locality_order == "hilbert" else _lib.SW_ORDER_KEEP
        p.timestepping_method = TIMESTEPPING[self.get_timestepping_method()]
        return p

    def _attach(self, handle):
        """Use ``handle`` (an instance view of an ensemble handle) as this
        domain's device storage and send it the host state."""
        if self._handle is not None:
            raise RuntimeError("domain already has device storage")
        self._handle = handle
        self._in_ensemble = True
        self._init_device_state()
        self._push()

    def _ensure_device(self):
        if self._handle is None:
            self._handle = _lib.Handle(self.mesh, self._sw_params(), self.device)
            self._init_device_state()
        return self._handle

    def _init_device_state(self):
        self._handle.set_boundary(self._kinds, self._dirichlet)
        s = self._handle.get_state()
        s.time = self._time
        s.starttime = self.starttime
        s.order = self._order_
        self._handle.set_state(s)
        self._all_dirty = True

    def _each_array(self):
        for name, q in self.quantities.items():
            for key, arr in q._arrays.items():
                if name == "friction" and key != "centroid_values":
                    continue
                yield name, key, arr

    # Host copies of the device's arrays.  _stale holds the (quantity,
    # location) pairs, and "max_speed", whose host copy is out of date: each is
    # downloaded when it is read, alone.  _touched maps the arrays handed out
    # since the last push to a copy of what they held then: a push uploads
    # those that differ from it, so arrays that were only read stay as the
    # device has them.
    @property
    def _host_stale(self):
        return bool(self._stale)

    @_host_stale.setter
    def _host_stale(self, value):
        self._stale = ({(n, k) for n, k, _ in self._each_array()} | {"max_speed"}) if value else set()

    def _pull(self):
        """Every out-of-date host copy from the device."""
        if self._handle is None:
            return
        for name, key, arr in self._each_array():
            if (name, key) in self._stale:
                self._handle.download(_QID[name], _LOC[key], arr.reshape(-1))
        if "max_speed" in self._stale:
            self._max_speed = self._handle.max_speed()
        self._stale = set()

    def _pull_one(self, name, key):
        if (name, key) in self._stale and self._handle is not None:
            self._handle.download(_QID[name], _LOC[key], self.quantities[name]._arrays[key].reshape(-1))
            self._stale.discard((name, key))

    @property
    def max_speed(self):
        """Per-triangle maximum wave speed of the last flux evaluation (brought
        from the device when read)."""
        if "max_speed" in self._stale and self._handle is not None:
            self._max_speed = self._handle.max_speed()
            self._stale.discard("max_speed")
        return self._max_speed

    def _push(self):
        h = self._ensure_device()
        if self._all_dirty:
            items = list(self._each_array())
        else:
            items = []
            for (n, k), before in self._touched.items():
                arr = self.quantities[n]._arrays[k]
                if n == "friction" and k != "centroid_values":
                    continue
                if arr.shape != before.shape or not np.array_equal(arr.view(np.int64), before.view(np.int64)):
                    items.append((n, k, arr))
        # elevation vertex values last: the device derives the mean bed from them
        items.sort(key=lambda t: (t[0] == "elevation" and t[1] == "vertex_values"))
        for name, key, arr in items:
            h.upload(_QID[name], _LOC[key], arr)
        self._touched.clear()
        self._all_dirty = False

    def _dev(self):
        """Device handle with host changes applied; host copies marked stale.
        The host copies are brought up to date only before a push (pushing
        every array needs every copy current): a call with no host changes,
        such as each yield of evolve(), downloads nothing."""
        if self._handle is None or self._all_dirty or self._touched:
            if self._all_dirty:
                self._pull()
            self._push()
        self._host_stale = True
        return self._handle

    def _sync_scalars(self):
        s = self._handle.get_state()
        self._time = s.time
        self.timestep = s.timestep
        self.flux_timestep = s.flux_timestep
        self.number_of_steps = s.number_of_steps
        self.number_of_first_order_steps = s.number_of_first_order_steps
        self.recorded_max_timestep = s.recorded_max_timestep
        self.recorded_min_timestep = s.recorded_min_timestep
        self._order_ = s.order
        self.smallsteps = s.smallsteps
        self.yieldtime = s.yieldtime
        return s

    # --------------------------------------------------------- operators
    def compute_fluxes(self):
        """compute_fluxes_central_structure + gravity_wb (the reference's
        'wb_2' compute_fluxes, gpu_domain_advanced.py:1595-1696)."""
        h = self._dev()
        self.flux_timestep = h.compute_fluxes()
        h.call("sw_gravity_wb")
        return self.flux_timestep

    def compute_fluxes_central_structure(self):
        h = self._dev()
        self.flux_timestep = h.compute_fluxes()
        return self.flux_timestep

    def gravity_wb(self):
        self._dev().call("sw_gravity_wb")

    def manning_friction_implicit(self):
        self._dev().call("sw_manning_friction")

    def compute_forcing_terms(self):
        self.manning_friction_implicit()

    def update_timestep(self, yieldstep, finaltime):
        h = self._dev()
        has_final = finaltime is not None
        h.call("sw_update_timestep", float(yieldstep or 0.0), float(finaltime or 0.0), int(has_final))
        self._sync_scalars()

    def update_conserved_quantities(self):
        self._dev().call("sw_update_conserved_quantities")

    def set_flow_algorithm(self, flow_algorithm):
        """Names the flow algorithm.  'tsunami' switches
        protect_against_infinitesimal_and_negative_heights to that
        algorithm's protect (gpu_domain_advanced.py:1014-1031); its flux and
        extrapolation are not defined by the reference, so evolve() refuses
        it."""
        self.flow_algorithm = str(flow_algorithm)

    def protect_against_infinitesimal_and_negative_heights(self):
        # gpu_domain_advanced.py:1003-1047: _protect_swb2 for 'tsunami'
        self._dev().call("sw_protect_swb2" if self.flow_algorithm == "tsunami" else "sw_protect")

    def extrapolate_second_order_sw(self):
        self._dev().call("sw_extrapolate_second_order_sw")

    def extrapolate_first_order(self):
        self._dev().call("sw_extrapolate_first_order")

    def balance_deep_and_shallow(self):
        self._dev().call("sw_balance_deep_and_shallow")

    def interpolate_from_vertices_to_edges(self):
        self._dev().call("sw_interpolate_from_vertices_to_edges")

    def distribute_to_vertices_and_edges(self):
        h = self._dev()
        s = h.get_state()
        if s.order != self._order_:
            s.order = self._order_
            h.set_state(s)
        h.call("sw_distribute_to_vertices_and_edges")

    def update_boundary(self):
        self._dev().call("sw_update_boundary")

    def update_other_quantities(self):
        self._dev().call("sw_update_other_quantities")

    def evolve_one_euler_step(self, yieldstep, finaltime):
        self.compute_fluxes()
        self.compute_forcing_terms()
        self.update_timestep(yieldstep, finaltime)
        self.update_conserved_quantities()

    def backup_conserved_quantities(self):
        """Copies of the conserved centroid values (gpu_domain_advanced.py:1449-1475)."""
        self._dev().call("sw_backup_conserved_quantities")

    def saxpy_conserved_quantities(self, a, b):
        """centroid = a * centroid + b * backup for the conserved quantities
        (gpu_domain_advanced.py:1481-1519)."""
        self._dev().call("sw_saxpy_conserved_quantities", float(a), float(b))

    def update_ghosts(self):
        return None

    def update_extrema(self):
        return None

    def apply_fractional_steps(self):
        return None

    # ------------------------------------------------------------ evolve
    def evolve(self, yieldstep=None, finaltime=None, duration=None, skip_initial_step=False,
               fused=True):
        """Generator yielding the model time at every yieldstep and at the end
        (gpu_domain_advanced.py:2296-2698).  ``fused=False`` runs the
        one-launch-per-operator step instead of the fused device loop."""
        if self.__dict__.get("_in_ensemble"):
            raise RuntimeError("this domain belongs to an Ensemble: evolve the ensemble")
        if self.flow_algorithm == "tsunami":
            # the reference's 'tsunami' compute_fluxes only prints and its
            # distribute call is malformed (gpu_domain_advanced.py:1677-1678,
            # :1757-1776): there is no step to reproduce
            raise NotImplementedError("the 'tsunami' flow algorithm has no evolve step; only its protect "
                                      "(protect_against_infinitesimal_and_negative_heights) is provided")
        h = self._dev()
        if yieldstep is None:
            yieldstep = self.evolve_max_timestep
        else:
            yieldstep = float(yieldstep)
        if finaltime is not None and duration is not None:
            raise ValueError("give finaltime or duration, not both")
        if duration is not None:
            finaltime = self.starttime + float(duration)
        if finaltime is not None:
            finaltime = float(finaltime)
            if finaltime < self.starttime:
                raise ValueError("finaltime precedes starttime")
        self.finaltime = finaltime
        s = h.get_state()
        s.starttime = self.starttime
        h.set_state(s)
        h.call("sw_evolve_prepare", yieldstep, finaltime if finaltime is not None else 0.0,
               int(finaltime is not None))
        self._host_stale = True
        self._sync_scalars()
        if self.store and self.writer is None:
            # a later evolve() call continues the same file
            self.initialise_storage()
        if not skip_initial_step:
            if self.store:
                self.store_timestep()
            yield self._time
        while True:
            h = self._dev()
            if fused:
                _, event = h.evolve_steps(1 << 40)
            else:
                event = _lib.SW_EVENT_NONE
                while event == _lib.SW_EVENT_NONE:
                    event = h.evolve_one_step_unfused()
            self._sync_scalars()
            if self.store:
                self.store_timestep()
            if event == _lib.SW_EVENT_FINAL:
                yield self._time
                break
            yield self._time
            self._handle.call("sw_advance_yield")
            self._sync_scalars()

    # ------------------------------------------------------------ reports
    def timestepping_statistics(self):
        return ("Time = %.4f, delta t in [%.8f, %.8f], steps=%d (%d)" % (
            self.get_time(), self.recorded_min_timestep, self.recorded_max_timestep,
            self.number_of_steps, self.number_of_first_order_steps))

    def write_time(self):
        print(self.timestepping_statistics())

    def get_water_volume(self):
        h = self.quantities["stage"].centroid_values - self.quantities["elevation"].centroid_values
        return float(np.sum(h * self.mesh.areas))

    def close(self):
        if self.writer is not None:
            self.writer.close()
            self.writer = None
        if self._handle is not None:
            self._pull()
            self._handle.close()
            self._handle = None


def create_domain_from_file(filename, DomainClass=None, build="host", **parameters):
    """ANUGA's create_domain_from_file, which the reference's Merimbula
    scenario calls (tests/CUDA/merimbula/merimbula.py:86): a domain on the
    mesh of a ``.tsh`` or NetCDF ``.msh`` file (optionally gzip-compressed),
    boundary edges tagged by the file's segments ('exterior' where untagged),
    coordinates kept relative to the file's georeference, and each vertex
    attribute column whose title names a quantity set as that quantity's
    vertex values.  ``build`` ('host' or 'device') picks where the mesh
    geometry is built; ``parameters`` go to the domain."""
    from .mesh import read_mesh

    points, triangles, segments, extra = read_mesh(filename, full=True)
    mesh = Mesh(points, triangles, segments=segments, build=build)
    domain = (DomainClass or Domain)(mesh=mesh, **parameters)
    domain.geo_reference.update(extra.get("georeference") or {})
    attributes = extra.get("vertex_attributes")
    for i, title in enumerate(extra.get("vertex_attribute_titles") or []):
        if title in domain.quantities and attributes is not None:
            domain.set_quantity(title, attributes[:, i][mesh.triangles], location="vertices")
    return domain
