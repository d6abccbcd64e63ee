"""ctypes binding of ``libanuga_sw.so`` (declarations in ``include/anuga_sw.h``).

The library is built in-tree by ``anuga_cuda_amd/csrc/build.sh``.  There is no
CPU fallback: if the shared library is missing or a GPU is not present, the
first use raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libanuga_sw.so")

SW_STAGE, SW_XMOM, SW_YMOM, SW_ELEVATION, SW_FRICTION, SW_HEIGHT, SW_XVELOCITY, SW_YVELOCITY = range(8)
SW_CENTROID, SW_EDGE, SW_VERTEX, SW_BOUNDARY, SW_EXPLICIT_UPDATE, SW_SEMI_IMPLICIT_UPDATE = range(6)
SW_BC_REFLECTIVE, SW_BC_DIRICHLET, SW_BC_TRANSMISSIVE = range(3)
SW_PRECISION_FP64, SW_PRECISION_MIXED = 0, 1
SW_ORDER_KEEP, SW_ORDER_HILBERT = 0, 1
SW_LIMITER_OPTIMISED, SW_LIMITER_VERTEX, SW_LIMITER_EDGE = 0, 1, 2
SW_EVENT_NONE, SW_EVENT_YIELD, SW_EVENT_FINAL = 0, 1, 2
SW_STEP_AUTO, SW_STEP_EDGE, SW_STEP_VERTEX = 0, 1, 2

SW_ERR_UPDATE = -1
SW_ERR_SMALLSTEPS = -2
SW_ERR_NEIGHBOUR = -3
SW_ERR_ARG, SW_ERR_HIP, SW_ERR_STATE = -10, -11, -12


class SwParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "g", "epsilon", "H0", "minimum_allowed_height", "maximum_allowed_speed",
        "beta_w", "beta_w_dry", "beta_uh", "beta_uh_dry", "beta_vh", "beta_vh_dry",
        "alpha_balance", "CFL", "evolve_max_timestep", "evolve_min_timestep")] + [
        (n, C.c_int32) for n in (
            "optimise_dry_cells", "extrapolate_velocity_second_order",
            "tight_slope_limiters", "use_centroid_velocities", "use_sloped_mannings",
            "max_smallsteps", "default_order", "precision", "locality_order", "timestepping_method",
            "limiter")] + [("quantity_beta", C.c_double * 3)]


_I64P = C.POINTER(C.c_int64)
_F64P = C.POINTER(C.c_double)


class SwMesh(C.Structure):
    _fields_ = [
        ("number_of_elements", C.c_int64),
        ("number_of_boundary_elements", C.c_int64),
        ("neighbours", _I64P),
        ("neighbour_edges", _I64P),
        ("surrogate_neighbours", _I64P),
        ("number_of_boundaries", _I64P),
        ("tri_full_flag", _I64P),
        ("normals", _F64P),
        ("edgelengths", _F64P),
        ("radii", _F64P),
        ("areas", _F64P),
        ("centroid_coordinates", _F64P),
        ("vertex_coordinates", _F64P),
        ("boundary_cells", _I64P),
        ("boundary_edges", _I64P),
    ]


class SwState(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "time", "starttime", "finaltime", "yieldtime", "yieldstep", "timestep", "flux_timestep",
        "recorded_max_timestep", "recorded_min_timestep")] + [
        (n, C.c_int64) for n in (
            "smallsteps", "order", "number_of_steps", "number_of_first_order_steps",
            "has_finaltime", "event", "error")]


class SwError(RuntimeError):
    """Error reported by the HIP library (message from sw_last_error)."""

    def __init__(self, code, message):
        super().__init__("anuga_sw error %d: %s" % (code, message))
        self.code = code


_lib = None

# (name, argtypes) of every exported entry point
_SIGNATURES = {
    "sw_create": ([C.POINTER(SwMesh), C.POINTER(SwParams), C.c_int], C.c_void_p),
    "sw_create_ensemble": ([C.POINTER(SwMesh), C.c_int32, C.POINTER(SwParams), C.c_int], C.c_void_p),
    "sw_select_instance": ([C.c_void_p, C.c_int32], C.c_int),
    "sw_destroy": ([C.c_void_p], C.c_int),
    "sw_last_error": ([C.c_void_p], C.c_char_p),
    "sw_set_params": ([C.c_void_p, C.POINTER(SwParams)], C.c_int),
    "sw_get_params": ([C.c_void_p, C.POINTER(SwParams)], C.c_int),
    "sw_set_boundary": ([C.c_void_p, C.POINTER(C.c_int32), _F64P], C.c_int),
    "sw_upload": ([C.c_void_p, C.c_int, C.c_int, _F64P, C.c_int64], C.c_int),
    "sw_download": ([C.c_void_p, C.c_int, C.c_int, _F64P, C.c_int64], C.c_int),
    "sw_upload_device": ([C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int64], C.c_int),
    "sw_download_device": ([C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int64], C.c_int),
    "sw_restart_device": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double], C.c_int),
    "sw_download_conserved_device": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "sw_get_state": ([C.c_void_p, C.POINTER(SwState)], C.c_int),
    "sw_set_state": ([C.c_void_p, C.POINTER(SwState)], C.c_int),
    "sw_download_max_speed": ([C.c_void_p, _F64P, C.c_int64], C.c_int),
    "sw_compute_fluxes": ([C.c_void_p, _F64P], C.c_int),
    "sw_gravity_wb": ([C.c_void_p], C.c_int),
    "sw_manning_friction": ([C.c_void_p], C.c_int),
    "sw_update_timestep": ([C.c_void_p, C.c_double, C.c_double, C.c_int32], C.c_int),
    "sw_update_conserved_quantities": ([C.c_void_p], C.c_int),
    "sw_protect": ([C.c_void_p], C.c_int),
    "sw_protect_swb2": ([C.c_void_p], C.c_int),
    "sw_extrapolate_second_order_sw": ([C.c_void_p], C.c_int),
    "sw_extrapolate_first_order": ([C.c_void_p], C.c_int),
    "sw_balance_deep_and_shallow": ([C.c_void_p], C.c_int),
    "sw_interpolate_from_vertices_to_edges": ([C.c_void_p], C.c_int),
    "sw_distribute_to_vertices_and_edges": ([C.c_void_p], C.c_int),
    "sw_update_boundary": ([C.c_void_p], C.c_int),
    "sw_update_other_quantities": ([C.c_void_p], C.c_int),
    "sw_evolve_prepare": ([C.c_void_p, C.c_double, C.c_double, C.c_int32], C.c_int),
    "sw_evolve_steps": ([C.c_void_p, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int32)], C.c_int),
    "sw_advance_yield": ([C.c_void_p], C.c_int),
    "sw_evolve_one_step_unfused": ([C.c_void_p, C.POINTER(C.c_int32)], C.c_int),
    "sw_backup_conserved_quantities": ([C.c_void_p], C.c_int),
    "sw_extrapolate_second_order_and_limit_by_vertex": ([C.c_void_p], C.c_int),
    "sw_extrapolate_second_order_and_limit_by_edge": ([C.c_void_p], C.c_int),
    "sw_set_sww_points": ([C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p], C.c_int),
    "sw_sww_values": ([C.c_void_p, C.c_double, C.c_void_p, C.c_int64], C.c_int),
    "sw_saxpy_conserved_quantities": ([C.c_void_p, C.c_double, C.c_double], C.c_int),
    "sw_synchronize": ([C.c_void_p], C.c_int),
    "sw_autotune": ([C.c_void_p, C.c_int32, C.POINTER(C.c_int32)], C.c_int),
    "sw_enable_timing": ([C.c_void_p, C.c_int32], C.c_int),
    "sw_build_mesh": ([C.c_int, C.c_int64, _F64P, C.c_int64, _I64P, C.c_void_p], C.c_int),
    "sw_mesh_last_error": ([], C.c_char_p),
    "sw_build_mesh_noalloc": ([C.c_int, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p], C.c_int),
    "sw_kernel_times": ([C.c_void_p, _F64P, _I64P, C.c_int32], C.c_int),
    "sw_set_step_path": ([C.c_void_p, C.c_int32], C.c_int),
    "sw_get_step_path": ([C.c_void_p, C.POINTER(C.c_int32)], C.c_int),
}


class SwMeshGeometry(C.Structure):
    _fields_ = [("triangles", _I64P), ("vertex_coordinates", _F64P), ("centroid_coordinates", _F64P),
                ("areas", _F64P), ("normals", _F64P), ("edgelengths", _F64P),
                ("edge_midpoint_coordinates", _F64P), ("radii", _F64P), ("neighbours", _I64P),
                ("neighbour_edges", _I64P)]


def build_mesh_geometry(points, triangles, device=0):
    """sw_build_mesh: the oriented triangles, derived geometry and neighbours
    of a mesh, built on GPU ``device``.  Returns a dict of host arrays named
    as the Mesh attributes."""
    lib = load()
    pts = np.ascontiguousarray(points, dtype=np.float64)
    tris = np.ascontiguousarray(triangles, dtype=np.int64)
    N = len(tris)
    out = {"triangles": np.empty((N, 3), dtype=np.int64), "vertex_coordinates": np.empty((N, 6)),
           "centroid_coordinates": np.empty((N, 2)), "areas": np.empty(N), "normals": np.empty((N, 6)),
           "edgelengths": np.empty((N, 3)), "edge_midpoint_coordinates": np.empty((N, 6)),
           "radii": np.empty(N), "neighbours": np.empty((N, 3), dtype=np.int64),
           "neighbour_edges": np.empty((N, 3), dtype=np.int64)}
    g = SwMeshGeometry()
    for name, _ in SwMeshGeometry._fields_:
        a = out[name]
        setattr(g, name, i64p(a) if a.dtype == np.int64 else f64p(a))
    rc = lib.sw_build_mesh(int(device), len(pts), f64p(pts), N, i64p(tris), C.byref(g))
    if rc != 0:
        msg = (lib.sw_mesh_last_error() or b"").decode()
        raise (ValueError if rc == SW_ERR_ARG else SwError)(msg or "sw_build_mesh failed (%d)" % rc)
    return out


def exported_symbols():
    return sorted(_SIGNATURES)


def share_hip_runtime():
    """Makes the HIP runtime that torch ships the one this process uses.

    torch's ROCm build carries its own copy of the HIP and HSA runtimes under
    the same soname (libamdhip64.so.7) as /opt/rocm's.  Loaded after torch,
    the HIP libraries of this package bind to torch's copy and device
    pointers and streams pass freely between the two.  Loaded before torch,
    they map /opt/rocm's copy, torch then maps its own as a second runtime,
    and that second runtime does not see the GPU (torch.cuda.is_available()
    is False).  Importing torch first, when it is installed, keeps one runtime
    per process.  C callers of the library are not affected."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "libanuga_sw.so not built (run anuga_cuda_amd/csrc/build.sh or "
            "__graft_entry__.build()); there is no CPU fallback")
    share_hip_runtime()
    lib = C.CDLL(LIB_PATH)
    for name, (argtypes, restype) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = restype
    _lib = lib
    return lib


def f64p(a):
    return a.ctypes.data_as(_F64P)


def i64p(a):
    return a.ctypes.data_as(_I64P)


def _sw_mesh(mesh, keep):
    """ctypes sw_mesh of a Mesh; the int64 / float64 copies stay alive in ``keep``."""
    m = SwMesh()
    m.number_of_elements = mesh.number_of_elements
    m.number_of_boundary_elements = mesh.number_of_boundary_elements

    def ia(name, arr):
        a = np.ascontiguousarray(arr, dtype=np.int64)
        if a.size == 0:
            a = np.zeros(1, dtype=np.int64)
        keep.append(a)
        setattr(m, name, i64p(a))

    def fa(name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        keep.append(a)
        setattr(m, name, f64p(a))

    ia("neighbours", mesh.neighbours)
    ia("neighbour_edges", mesh.neighbour_edges)
    ia("surrogate_neighbours", mesh.surrogate_neighbours)
    ia("number_of_boundaries", mesh.number_of_boundaries)
    ia("tri_full_flag", mesh.tri_full_flag)
    fa("normals", mesh.normals)
    fa("edgelengths", mesh.edgelengths)
    fa("radii", mesh.radii)
    fa("areas", mesh.areas)
    fa("centroid_coordinates", mesh.centroid_coordinates)
    fa("vertex_coordinates", mesh.vertex_coordinates)
    ia("boundary_cells", mesh.boundary_cells)
    ia("boundary_edges", mesh.boundary_edges)
    return m


class Handle:
    """Owner of one ``sw_handle``: device buffers of one domain, or of an
    ensemble of domains (``Handle.ensemble``).  ``view(i)`` gives a non-owning
    handle whose calls address instance ``i`` of an ensemble."""

    def __init__(self, mesh, params: SwParams, device=0, _meshes=None):
        self.lib = load()
        keep = []
        self.params = params
        self.inst = None
        self._owner = True
        if _meshes is None:
            m = _sw_mesh(mesh, keep)
            h = self.lib.sw_create(C.byref(m), C.byref(params), int(device))
            self.N = mesh.number_of_elements
            self.Nb = mesh.number_of_boundary_elements
            self.n_instances = 1
        else:
            arr = (SwMesh * len(_meshes))(*[_sw_mesh(x, keep) for x in _meshes])
            h = self.lib.sw_create_ensemble(arr, len(_meshes), C.byref(params), int(device))
            self.N = sum(x.number_of_elements for x in _meshes)
            self.Nb = sum(x.number_of_boundary_elements for x in _meshes)
            self.n_instances = len(_meshes)
            self._sizes = [(x.number_of_elements, x.number_of_boundary_elements) for x in _meshes]
        del keep
        if not h:
            raise SwError(-10, self.lib.sw_last_error(None).decode())
        self.h = C.c_void_p(h)

    @classmethod
    def ensemble(cls, meshes, params, device=0):
        return cls(None, params, device, _meshes=list(meshes))

    def view(self, i):
        """Non-owning handle on instance i (its per-instance calls select it first)."""
        v = Handle.__new__(Handle)
        v.__dict__.update(self.__dict__)
        v._owner = False
        v.inst = int(i)
        v.N, v.Nb = self._sizes[i] if self.n_instances > 1 else (self.N, self.Nb)
        return v

    # ------------------------------------------------------------------
    def check(self, rc):
        if rc != 0:
            raise SwError(rc, self.lib.sw_last_error(self.h).decode())

    def select(self):
        """Selects this view's instance (or all, for an ensemble's owner)."""
        if self.inst is not None:
            self.check(self.lib.sw_select_instance(self.h, self.inst))
        elif self.n_instances > 1:
            self.check(self.lib.sw_select_instance(self.h, -1))

    def call(self, name, *args):
        self.select()
        self.check(getattr(self.lib, name)(self.h, *args))

    def close(self):
        if getattr(self, "h", None):
            if self._owner:
                self.lib.sw_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params):
        self.params = params
        self.call("sw_set_params", C.byref(params))

    def set_boundary(self, kinds, dirichlet):
        k = np.ascontiguousarray(kinds, dtype=np.int32)
        d = np.ascontiguousarray(dirichlet, dtype=np.float64)
        if k.size == 0:
            return
        self.call("sw_set_boundary", k.ctypes.data_as(C.POINTER(C.c_int32)), f64p(d))

    def upload(self, q, loc, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        if a.size == 0:
            return
        self.call("sw_upload", int(q), int(loc), f64p(a), a.size)

    def download(self, q, loc, out):
        if out.size == 0:
            return out
        assert out.flags.c_contiguous and out.dtype == np.float64
        self.call("sw_download", int(q), int(loc), f64p(out), out.size)
        return out

    def upload_device(self, q, loc, ptr, count):
        """Upload from a device buffer (address ``ptr``, ``count`` doubles)."""
        self.call("sw_upload_device", int(q), int(loc), C.c_void_p(int(ptr)), int(count))

    def download_device(self, q, loc, ptr, count):
        """Download into a device buffer (address ``ptr``, ``count`` doubles)."""
        self.call("sw_download_device", int(q), int(loc), C.c_void_p(int(ptr)), int(count))

    def restart_device(self, w, uh, vh, time, yieldstep):
        """Restart from device buffers of the conserved centroid values at
        ``time``, yields every ``yieldstep`` (sw_restart_device)."""
        self.call("sw_restart_device", C.c_void_p(int(w)), C.c_void_p(int(uh)), C.c_void_p(int(vh)), float(time),
                  float(yieldstep))

    def download_conserved_device(self, w, uh, vh):
        """Conserved centroid values into three device buffers."""
        self.call("sw_download_conserved_device", C.c_void_p(int(w)), C.c_void_p(int(uh)), C.c_void_p(int(vh)))

    def set_sww_points(self, number_of_triangles_per_node=None, vertex_value_indices=None):
        """Output points for sw_sww_values: mesh nodes (both arrays given) or
        unique vertices (neither)."""
        if number_of_triangles_per_node is None:
            self.call("sw_set_sww_points", 0, None, None)
            return
        counts = np.ascontiguousarray(number_of_triangles_per_node, dtype=np.int64)
        idx = np.ascontiguousarray(vertex_value_indices, dtype=np.int64)
        self.call("sw_set_sww_points", len(counts), counts.ctypes.data, idx.ctypes.data)

    def sww_values(self, npoints, minimum_storable_height, out=None):
        """FP32 [4][npoints] elevation, stage, xmomentum, ymomentum at the points."""
        if out is None:
            out = np.empty((4, npoints), dtype=np.float32)
        assert out.dtype == np.float32 and out.flags.c_contiguous and out.shape == (4, npoints)
        self.call("sw_sww_values", float(minimum_storable_height), out.ctypes.data, out.size)
        return out

    def get_state(self, check=True):
        """The evolve state of the selected instance.  With ``check`` a latched
        device error raises SwError; without, the state (its ``error`` field
        holding the latched error bits) is returned regardless."""
        s = SwState()
        if check:
            self.call("sw_get_state", C.byref(s))
        else:
            self.select()
            self.lib.sw_get_state(self.h, C.byref(s))  # rc ignored: the bits are in s.error
        return s

    def set_state(self, s):
        self.call("sw_set_state", C.byref(s))

    def max_speed(self):
        out = np.empty(self.N)
        self.call("sw_download_max_speed", f64p(out), out.size)
        return out

    def compute_fluxes(self):
        ts = C.c_double()
        self.call("sw_compute_fluxes", C.byref(ts))
        return ts.value

    def evolve_steps(self, max_steps):
        done = C.c_int64()
        ev = C.c_int32()
        self.call("sw_evolve_steps", int(max_steps), C.byref(done), C.byref(ev))
        return done.value, ev.value

    def evolve_one_step_unfused(self):
        ev = C.c_int32()
        self.call("sw_evolve_one_step_unfused", C.byref(ev))
        return ev.value

    def autotune(self, steps=3):
        """Workgroup sizes (flux, update, distribute) chosen by sw_autotune
        after ``steps`` real steps per candidate."""
        chosen = (C.c_int32 * 3)()
        self.call("sw_autotune", int(steps), chosen)
        return tuple(chosen)

    def set_step_path(self, path):
        """SW_STEP_AUTO (vertex-record step where it applies) or SW_STEP_EDGE."""
        self.call("sw_set_step_path", int(path))

    def step_path(self):
        """The step sw_evolve_steps would run now: SW_STEP_VERTEX or SW_STEP_EDGE."""
        p = C.c_int32()
        self.call("sw_get_step_path", C.byref(p))
        return p.value

    def kernel_times(self, n=4):
        ms = np.zeros(n)
        cnt = np.zeros(n, dtype=np.int64)
        self.call("sw_kernel_times", f64p(ms), i64p(cnt), n)
        return ms, cnt
