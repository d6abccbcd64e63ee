"""sw_build_mesh_noalloc: the mesh build on device memory with the library's
static scratch (no allocation by the library), against sw_build_mesh.

Two calls in a row, the second with a smaller mesh, give the arrays of
sw_build_mesh bit for bit (the scratch keeps the first call's keys and flags,
so this also checks that what must start at zero is cleared); a mesh with a
vertex index out of range is rejected and the next call succeeds; a mesh
above the scratch limit is rejected before any work."""
import ctypes as C

import numpy as np
import pytest

from anuga_cuda_amd import _lib
from anuga_cuda_amd.mesh import delaunay_random, rectangular_cross

pytestmark = pytest.mark.gpu

FIELDS = {"triangles": (3, np.int64), "vertex_coordinates": (6, np.float64), "centroid_coordinates": (2, np.float64),
          "areas": (1, np.float64), "normals": (6, np.float64), "edgelengths": (3, np.float64),
          "edge_midpoint_coordinates": (6, np.float64), "radii": (1, np.float64), "neighbours": (3, np.int64),
          "neighbour_edges": (3, np.int64)}


def _noalloc(points, tris):
    import torch

    lib = _lib.load()
    dev = torch.device("cuda:0")
    N = len(tris)
    p = torch.tensor(np.ascontiguousarray(points, dtype=np.float64), device=dev)
    t = torch.tensor(np.ascontiguousarray(tris, dtype=np.int64), device=dev)
    out = {name: torch.empty((N, w) if w > 1 else (N,), dtype=torch.float64 if dt == np.float64 else torch.int64,
                             device=dev) for name, (w, dt) in FIELDS.items()}
    g = _lib.SwMeshGeometry()
    for name, ten in out.items():
        ptype = dict(_lib.SwMeshGeometry._fields_)[name]
        setattr(g, name, C.cast(C.c_void_p(ten.data_ptr()), ptype))
    torch.cuda.synchronize()
    rc = lib.sw_build_mesh_noalloc(0, len(points), p.data_ptr(), N, t.data_ptr(), C.byref(g))
    msg = (lib.sw_mesh_last_error() or b"").decode()
    return rc, msg, {name: ten.cpu().numpy() for name, ten in out.items()}


def _check_equal(points, tris):
    rc, msg, got = _noalloc(points, tris)
    assert rc == 0, msg
    exp = _lib.build_mesh_geometry(points, tris)
    for name in FIELDS:
        a, b = got[name], np.asarray(exp[name]).reshape(got[name].shape)
        assert a.dtype == b.dtype and np.array_equal(a, b), name
    return got


def test_noalloc_mesh_build_matches_the_allocating_one_on_two_calls():
    pts, tris = delaunay_random(20000, extent=(3000.0, 2000.0), seed=5)
    big = _check_equal(pts, tris)
    assert len(tris) > 30000 and np.count_nonzero(big["neighbours"] < 0) > 100
    pts2, tris2, _ = rectangular_cross(12, 9, 120.0, 90.0)
    small = _check_equal(pts2, tris2)
    assert len(tris2) == 432 and np.count_nonzero(small["neighbours"] >= 0) > 1000


def test_noalloc_mesh_build_errors_and_limit():
    pts, tris, _ = rectangular_cross(4, 4, 40.0, 40.0)
    bad = tris.copy()
    bad[5, 1] = len(pts) + 7
    rc, msg, _ = _noalloc(pts, bad)
    assert rc == _lib.SW_ERR_ARG and "out of range" in msg
    _check_equal(pts, tris)  # the flags of the failed call are cleared
    lib = _lib.load()
    g = _lib.SwMeshGeometry()
    for name, _t in _lib.SwMeshGeometry._fields_:
        setattr(g, name, C.cast(C.c_void_p(16), _t))  # never touched: the size check comes first
    rc = lib.sw_build_mesh_noalloc(0, 100, 16, 1048577, 16, C.byref(g))
    assert rc == _lib.SW_ERR_ARG
    assert "scratch limit" in (lib.sw_mesh_last_error() or b"").decode()
