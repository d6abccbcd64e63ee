"""Driver of the sanitized host-code run (tests/test_sanitized.py): loads the
ASan + UBSan build of the library given on the command line and runs its
host-side mesh code -- validation, device numbering (Hilbert order), packing
of the edge, flag, surrogate and neighbour-word arrays, the vertex-record step's
geometry check -- on meshes of several kinds, single and as an ensemble, plus
malformed meshes for the error paths.  Without a GPU, sw_create stops at
hipSetDevice after all of that; with one it creates and destroys the handle.
Prints 'ok' at the end."""
import ctypes as C
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from anuga_cuda_amd import _lib, scenarios  # noqa: E402
from anuga_cuda_amd.domain import Domain  # noqa: E402
from anuga_cuda_amd.mesh import Mesh, delaunay_random, rectangular_cross  # noqa: E402

lib = C.CDLL(sys.argv[1])
lib.sw_create.restype = C.c_void_p
lib.sw_create.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
lib.sw_create_ensemble.restype = C.c_void_p
lib.sw_create_ensemble.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int]
lib.sw_last_error.restype = C.c_char_p
lib.sw_last_error.argtypes = [C.c_void_p]
lib.sw_destroy.argtypes = [C.c_void_p]


def create(meshes, params):
    keep = []
    arr = (_lib.SwMesh * len(meshes))(*[_lib._sw_mesh(m, keep) for m in meshes])
    if len(meshes) == 1:
        h = lib.sw_create(C.byref(arr[0]), C.byref(params), 0)
    else:
        h = lib.sw_create_ensemble(arr, len(meshes), C.byref(params), 0)
    msg = lib.sw_last_error(None).decode()
    if h:
        lib.sw_destroy(h)
        return "created"
    return msg


meshes = [
    Mesh(*rectangular_cross(7, 5, 70.0, 50.0)),
    scenarios.merimbula().mesh,
    Mesh(*delaunay_random(3000, extent=(2000.0, 2000.0), seed=3)).random_renumbered(seed=4),
]
for order in (_lib.SW_ORDER_HILBERT, _lib.SW_ORDER_KEEP):
    p = Domain(mesh=meshes[0])._sw_params()
    p.locality_order = order
    for m in meshes:
        r = create([m], p)
        assert r == "created" or "hipSetDevice" in r, r
    r = create(meshes, p)
    assert r == "created" or "hipSetDevice" in r, r
p = Domain(mesh=meshes[0])._sw_params()
# malformed meshes stop in the host validation, before any device call
bad = Mesh(*rectangular_cross(3, 3, 30.0, 30.0))
bad.neighbours = bad.neighbours.copy()
k = int(np.nonzero((bad.neighbours >= 0).all(axis=1))[0][0])
bad.neighbours[k, 0] = (bad.neighbours[k, 0] + 1) % bad.number_of_elements
assert "not symmetric" in create([bad], p)
bad2 = Mesh(*rectangular_cross(3, 3, 30.0, 30.0))
bad2.surrogate_neighbours = bad2.surrogate_neighbours.copy()
bad2.surrogate_neighbours[2, 1] = 10 ** 6
assert "out of range" in create([bad2], p)
print("ok")
