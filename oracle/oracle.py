"""ctypes wrapper of the CPU FP64 oracle (``liboracle_sw.so``) -- TEST INFRASTRUCTURE.

Used only by tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py``, always as the checker; the product never imports this module.

``OracleDomain`` holds its own copy of every array of a domain (mesh,
quantities, boundary kinds, parameters) and runs the restated reference C
routines of ``sw_oracle.c`` on it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIB: another build of the same source (the sanitized build of the
# tests' sanitizer run)
LIB_PATH = os.environ.get("ORACLE_LIB") or os.path.join(_HERE, "liboracle_sw.so")

_I64P = C.POINTER(C.c_int64)
_F64P = C.POINTER(C.c_double)

_PARAM_F = ["epsilon", "H0", "g", "minimum_allowed_height", "maximum_allowed_speed",
            "beta_w", "beta_w_dry", "beta_uh", "beta_uh_dry", "beta_vh", "beta_vh_dry",
            "alpha_balance", "CFL", "evolve_max_timestep", "evolve_min_timestep"]
_PARAM_I = ["optimise_dry_cells", "extrapolate_velocity_second_order", "tight_slope_limiters",
            "use_centroid_velocities", "use_sloped_mannings", "max_smallsteps", "default_order"]
_STATE_F = ["time", "starttime", "finaltime", "yieldtime", "yieldstep", "timestep", "flux_timestep",
            "recorded_max_timestep", "recorded_min_timestep"]
_STATE_I = ["smallsteps", "_order_", "number_of_steps", "number_of_first_order_steps", "has_finaltime"]
_MESH_I = ["neighbours", "neighbour_edges", "surrogate_neighbours", "number_of_boundaries",
           "tri_full_flag"]
_MESH_F = ["normals", "edgelengths", "radii", "areas", "centroid_coordinates", "vertex_coordinates"]
_BND = ["boundary_cells", "boundary_edges", "boundary_type"]
_QNAMES = {"stage": "stage", "xmomentum": "xmom", "ymomentum": "ymom", "elevation": "bed",
           "height": "height", "xvelocity": "xvelocity", "yvelocity": "yvelocity"}
_LOCS = ["centroid_values", "edge_values", "vertex_values", "boundary_values"]


class OrDomain(C.Structure):
    _fields_ = (
        [("number_of_elements", C.c_int64), ("number_of_boundary_elements", C.c_int64)]
        + [(n, C.c_double) for n in _PARAM_F]
        + [(n, C.c_int64) for n in _PARAM_I]
        + [(n, C.c_double) for n in _STATE_F]
        + [(n, C.c_int64) for n in _STATE_I]
        + [(n, _I64P) for n in _MESH_I]
        + [(n, _F64P) for n in _MESH_F]
        + [(n, _I64P) for n in _BND]
        + [("boundary_dirichlet", _F64P)]
        + [(p + "_" + s, _F64P) for p in ("stage", "xmom", "ymom")
           for s in ("centroid_values", "edge_values", "vertex_values", "boundary_values",
                     "explicit_update", "semi_implicit_update")]
        + [("bed_" + s, _F64P) for s in _LOCS]
        + [("friction_centroid_values", _F64P)]
        + [(p + "_" + s, _F64P) for p in ("height", "xvelocity", "yvelocity") for s in _LOCS]
        + [("max_speed", _F64P), ("already_computed_flux", _I64P), ("flux_call", C.c_int64)]
        + [("timestepping_method", C.c_int64)]
        + [(p + "_centroid_backup", _F64P) for p in ("stage", "xmom", "ymom")]
        + [("limiter", C.c_int64), ("beta_stage", C.c_double), ("beta_xmom", C.c_double),
           ("beta_ymom", C.c_double)]
    )

_METHODS = {"euler": 1, "rk2": 2, "rk3": 3}


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    P = C.POINTER(OrDomain)
    sigs = {
        "or_compute_fluxes": ([P], C.c_double),
        "or_gravity_wb": ([P], None),
        "or_manning_friction_flat": ([P], None),
        "or_manning_friction_sloped": ([P], None),
        "or_compute_forcing_terms": ([P], None),
        "or_update": ([C.c_int64, C.c_double, _F64P, _F64P, _F64P], C.c_int),
        "or_update_conserved_quantities": ([P], C.c_int),
        "or_protect": ([P], None),
        "or_protect_swb2": ([P], None),
        "or_extrapolate_second_order_sw": ([P], C.c_int),
        "or_extrapolate_first_order": ([C.c_int64, _F64P, _F64P, _F64P], None),
        "or_balance_deep_and_shallow": ([P], None),
        "or_interpolate_from_vertices_to_edges": ([C.c_int64, _F64P, _F64P], None),
        "or_distribute_to_vertices_and_edges": ([P], C.c_int),
        "or_update_boundary": ([P], None),
        "or_update_other_quantities": ([P], None),
        "or_update_timestep": ([P], C.c_int),
        "or_evolve_one_euler_step": ([P], C.c_int),
        "or_backup_conserved_quantities": ([P], None),
        "or_extrapolate_second_order_and_limit": ([P, _F64P, _F64P, _F64P, C.c_double, C.c_int], None),
        "or_saxpy_conserved_quantities": ([P, C.c_double, C.c_double], None),
        "or_evolve_one_rk_step": ([P, C.c_int], C.c_int),
        "or_evolve_prepare": ([P, C.c_double, C.c_double, C.c_int64], C.c_int),
        "or_evolve_step": ([P, _I64P], C.c_int),
        "or_evolve_steps": ([P, C.c_int64, _I64P, _I64P], C.c_int),
        "or_advance_yield": ([P], None),
        "or_sizeof_domain": ([], C.c_int64),
        "or_flux_function_central": ([_F64P, _F64P, C.c_double, C.c_double, C.c_double,
                                      C.c_double, C.c_double, C.c_double, C.c_double,
                                      C.c_double, _F64P, _F64P], None),
    }
    for name, (a, r) in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = a
        fn.restype = r
    if lib.or_sizeof_domain() != C.sizeof(OrDomain):
        raise RuntimeError("or_domain layout mismatch: C %d vs ctypes %d"
                           % (lib.or_sizeof_domain(), C.sizeof(OrDomain)))
    _lib = lib
    return lib


class OracleError(RuntimeError):
    def __init__(self, code, what):
        super().__init__("oracle %s returned %d" % (what, code))
        self.code = code


def average_vertex_values(vertex_value_indices, number_of_triangles_per_node, vertex_values):
    """ANUGA's quantity_ext.c ``_average_vertex_values`` (the smoothing of
    Quantity.get_vertex_values(smooth=True)), restated as a plain loop: walk
    the indices in order, summing the flat vertex values of the current node,
    and store total / count when the node's count is reached.  (The anuga
    package is not part of the reference.)  A node with no triangles gets 0."""
    vv = np.asarray(vertex_values, dtype=np.float64).ravel()
    counts = [int(c) for c in number_of_triangles_per_node]
    out = np.zeros(len(counts))
    node = 0
    while node < len(counts) and counts[node] == 0:
        node += 1
    k = 0
    total = 0.0
    for index in vertex_value_indices:
        k += 1
        total += float(vv[index])
        if k == counts[node]:
            out[node] = total / k
            total = 0.0
            k = 0
            node += 1
            while node < len(counts) and counts[node] == 0:
                node += 1
    return out


def sww_point_values(elevation, stage, xmom, ymom, minimum_storable_height):
    """The SWW writer's stored values from FP64 point values: rounded to FP32;
    where stage - elevation < minimum_storable_height (FP32 arithmetic, as the
    writer's numpy test on FP32 arrays) stage becomes the elevation and the
    momenta 0.  Returns float32 [4][P]."""
    z, w, u, v = (np.asarray(a, dtype=np.float64).astype(np.float32) for a in (elevation, stage, xmom, ymom))
    storable = (w - z) >= np.float32(minimum_storable_height)
    return np.stack([z, np.where(storable, w, z), np.where(storable, u, np.float32(0)),
                     np.where(storable, v, np.float32(0))]).astype(np.float32)


def flux_function_central(ql, qr, zl, zr, n1, n2, epsilon=1e-12, H0=1e-3, g=9.8):
    """One edge flux, as _flux_function_central (h0 = H0^2, threshold 10 H0)."""
    lib = load()
    ql = np.ascontiguousarray(ql, dtype=np.float64)
    qr = np.ascontiguousarray(qr, dtype=np.float64)
    ef = np.zeros(3)
    ms = np.zeros(1)
    lib.or_flux_function_central(ql.ctypes.data_as(_F64P), qr.ctypes.data_as(_F64P), zl, zr, n1, n2,
                                 epsilon, H0 * H0, 10 * H0, g, ef.ctypes.data_as(_F64P),
                                 ms.ctypes.data_as(_F64P))
    return ef, float(ms[0])


class OracleDomain:
    """CPU copy of a domain stepped by the oracle routines."""

    def __init__(self, mesh, params, quantities, kinds, dirichlet, state=None):
        self.lib = load()
        N, Nb = mesh.number_of_elements, mesh.number_of_boundary_elements
        self.N, self.Nb = N, Nb
        self.mesh = mesh
        d = OrDomain()
        self.d = d
        self.arrays = {}

        def keep(name, arr, dtype):
            a = np.ascontiguousarray(np.array(arr, dtype=dtype, copy=True))
            if a.size == 0:
                a = np.zeros(1, dtype=dtype)
            self.arrays[name] = a
            setattr(d, name, a.ctypes.data_as(_I64P if dtype == np.int64 else _F64P))
            return a

        d.number_of_elements = N
        d.number_of_boundary_elements = Nb
        for n in _PARAM_F:
            setattr(d, n, float(params[n]))
        for n in _PARAM_I:
            setattr(d, n, int(params[n]))
        for n in _MESH_I:
            keep(n, getattr(mesh, n), np.int64)
        for n in _MESH_F:
            keep(n, getattr(mesh, n), np.float64)
        keep("boundary_cells", mesh.boundary_cells, np.int64)
        keep("boundary_edges", mesh.boundary_edges, np.int64)
        keep("boundary_type", kinds, np.int64)
        keep("boundary_dirichlet", dirichlet, np.float64)
        for qname, cname in _QNAMES.items():
            for loc in _LOCS:
                keep(cname + "_" + loc, quantities[qname][loc], np.float64)
        for cname, qname in (("stage", "stage"), ("xmom", "xmomentum"), ("ymom", "ymomentum")):
            for loc in ("explicit_update", "semi_implicit_update"):
                keep(cname + "_" + loc, quantities[qname][loc], np.float64)
        keep("friction_centroid_values", quantities["friction"]["centroid_values"], np.float64)
        keep("max_speed", np.zeros(N), np.float64)
        keep("already_computed_flux", np.zeros(3 * N, dtype=np.int64), np.int64)
        d.flux_call = 0
        d.timestepping_method = _METHODS[params.get("timestepping_method", "euler")]
        for cname in ("stage", "xmom", "ymom"):
            keep(cname + "_centroid_backup", np.zeros(N), np.float64)
        d.limiter = {"optimised": 0, "vertex": 1, "edge": 2}[params.get("limiter", "optimised")]
        betas = params.get("quantity_beta", (1.0, 1.0, 1.0))
        d.beta_stage, d.beta_xmom, d.beta_ymom = (float(b) for b in betas)
        d._order_ = int(params["default_order"])
        d.evolve_max_timestep = float(params["evolve_max_timestep"])
        d.recorded_min_timestep = d.evolve_max_timestep
        d.recorded_max_timestep = float(params["evolve_min_timestep"])
        if state:
            for k, v in state.items():
                setattr(d, k, v)

    @classmethod
    def from_domain(cls, domain):
        """Snapshot of a Domain's current host arrays (pulls device state first)."""
        domain._pull()
        q = {}
        for name, Q in domain.quantities.items():
            q[name] = {k: v.copy() for k, v in Q._arrays.items()}
        state = {"time": domain._time, "starttime": domain.starttime, "_order_": domain._order_}
        params = dict(domain._params, timestepping_method=domain.get_timestepping_method(),
                      limiter=domain.limiter_path(),
                      quantity_beta=tuple(domain.quantities[q].beta for q in ("stage", "xmomentum", "ymomentum")))
        return cls(domain.mesh, params, q, domain._kinds, domain._dirichlet, state)

    # array access by ANUGA quantity name
    def q(self, name, loc):
        cname = _QNAMES.get(name, name)
        a = self.arrays[cname + "_" + loc]
        if loc in ("edge_values", "vertex_values"):
            return a.reshape(-1, 3)
        return a[: self.Nb] if loc == "boundary_values" else a

    @property
    def max_speed(self):
        return self.arrays["max_speed"]

    def _p(self):
        return C.byref(self.d)

    def _rc(self, rc, what):
        if rc:
            raise OracleError(rc, what)

    # operators
    def compute_fluxes(self):
        return self.lib.or_compute_fluxes(self._p())

    def gravity_wb(self):
        self.lib.or_gravity_wb(self._p())

    def manning_friction_implicit(self):
        self.lib.or_compute_forcing_terms(self._p())

    def update_timestep(self, finaltime=None):
        self.d.has_finaltime = int(finaltime is not None)
        if finaltime is not None:
            self.d.finaltime = float(finaltime)
        self._rc(self.lib.or_update_timestep(self._p()), "update_timestep")

    def update_conserved_quantities(self):
        self._rc(self.lib.or_update_conserved_quantities(self._p()), "update")

    def protect(self):
        self.lib.or_protect(self._p())

    def protect_swb2(self):
        self.lib.or_protect_swb2(self._p())

    def extrapolate_second_order_sw(self):
        self._rc(self.lib.or_extrapolate_second_order_sw(self._p()), "extrapolate")

    def balance_deep_and_shallow(self):
        self.lib.or_balance_deep_and_shallow(self._p())

    def interpolate_from_vertices_to_edges(self):
        for c in ("stage", "xmom", "ymom"):
            self.lib.or_interpolate_from_vertices_to_edges(
                self.N, self.arrays[c + "_vertex_values"].ctypes.data_as(_F64P),
                self.arrays[c + "_edge_values"].ctypes.data_as(_F64P))

    def distribute_to_vertices_and_edges(self):
        self._rc(self.lib.or_distribute_to_vertices_and_edges(self._p()), "distribute")

    def update_boundary(self):
        self.lib.or_update_boundary(self._p())

    def update_other_quantities(self):
        self.lib.or_update_other_quantities(self._p())

    def extrapolate_second_order_and_limit(self, name, by_edge, beta):
        """The per-quantity reconstruction of one conserved quantity into its
        vertex and edge arrays."""
        cv, vv, ev = (self.q(name, loc) for loc in ("centroid_values", "vertex_values", "edge_values"))
        self.lib.or_extrapolate_second_order_and_limit(self._p(), cv.ctypes.data_as(_F64P),
                                                        vv.ctypes.data_as(_F64P), ev.ctypes.data_as(_F64P),
                                                        float(beta), int(by_edge))

    def backup_conserved_quantities(self):
        self.lib.or_backup_conserved_quantities(self._p())

    def saxpy_conserved_quantities(self, a, b):
        self.lib.or_saxpy_conserved_quantities(self._p(), float(a), float(b))

    def evolve_prepare(self, yieldstep, finaltime):
        self._rc(self.lib.or_evolve_prepare(self._p(), float(yieldstep),
                                            float(finaltime if finaltime is not None else 0.0),
                                            int(finaltime is not None)), "evolve_prepare")

    def evolve_step(self):
        ev = C.c_int64()
        self._rc(self.lib.or_evolve_step(self._p(), C.byref(ev)), "evolve_step")
        return ev.value

    def evolve_steps(self, max_steps):
        done = C.c_int64()
        ev = C.c_int64()
        self._rc(self.lib.or_evolve_steps(self._p(), int(max_steps), C.byref(done), C.byref(ev)),
                 "evolve_steps")
        return done.value, ev.value

    def advance_yield(self):
        self.lib.or_advance_yield(self._p())

    def evolve(self, yieldstep, finaltime, skip_initial_step=False):
        """Generator mirroring Domain.evolve on the CPU."""
        self.evolve_prepare(yieldstep, finaltime)
        if not skip_initial_step:
            yield self.d.time
        while True:
            _, ev = self.evolve_steps(1 << 40)
            yield self.d.time
            if ev == 2:
                break
            self.advance_yield()
