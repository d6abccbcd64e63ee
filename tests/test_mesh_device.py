"""Mesh ingestion on the device (sw_build_mesh) against the host build.

The host build is pinned by the neighbour columns of the reference's own
mesh files (test_capi_and_mesh.py) and by closed-form geometry; the device
build must give the same arrays bit for bit (same FP64 operations in the same
order, correctly rounded), including on clockwise input and after the
boundary tagging.  Bad meshes raise the host build's errors.
"""
import os

import numpy as np
import pytest

from anuga_cuda_amd.domain import Domain, Reflective_boundary
from anuga_cuda_amd.mesh import Mesh, delaunay_random, read_mesh, rectangular_cross

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ARRAYS = ("triangles", "vertex_coordinates", "centroid_coordinates", "areas", "normals", "edgelengths",
          "edge_midpoint_coordinates", "radii", "neighbours", "neighbour_edges", "surrogate_neighbours",
          "number_of_boundaries", "boundary_cells", "boundary_edges", "boundary_tags")


def _same(a, b):
    for n in ARRAYS:
        x, y = getattr(a, n), getattr(b, n)
        assert x.shape == y.shape and np.array_equal(x, y), n
    assert a.boundary == b.boundary


def _meshes():
    p, t, b = rectangular_cross(13, 9, 130.0, 90.0)
    t = np.array(t)
    rng = np.random.default_rng(4)
    flip = rng.random(len(t)) < 0.5  # clockwise input on half the triangles
    t[flip, 1], t[flip, 2] = t[flip, 2].copy(), t[flip, 1].copy()
    b = {(k, 3 - e if flip[k] and e in (1, 2) else e): tag for (k, e), tag in b.items()}  # same edges, new labels
    yield "rect_cross_mixed_orientation", (p, t), {"boundary": b}
    p, t, segs = read_mesh(os.path.join(ROOT, "tests", "golden", "cairns.msh.gz"))
    yield "cairns", (p, t), {"segments": segs}
    p, t = delaunay_random(3000, extent=(2000.0, 2000.0), seed=9)
    perm = np.random.default_rng(5).permutation(len(t))
    yield "delaunay_random_order", (p, t[perm]), {}


@pytest.mark.parametrize("case", list(_meshes()), ids=lambda c: c[0])
def test_device_build_is_bit_identical_to_host_build(case):
    _, args, kw = case
    _same(Mesh(*args, **kw), Mesh(*args, build="device", **kw))


def test_device_built_mesh_renumbers_and_runs():
    """A Morton renumbering of a device-built mesh is built on the device
    again, and a domain on it evolves exactly as on the host-built mesh."""
    p, t = delaunay_random(2000, extent=(1000.0, 1000.0), seed=3)
    mh = Mesh(p, t).morton_renumbered()
    md = Mesh(p, t, build="device").morton_renumbered()
    _same(mh, md)
    out = []
    for m in (mh, md):
        d = Domain(mesh=m)
        d.set_quantity("elevation", lambda x, y: x / 1000.0)
        d.set_quantity("stage", lambda x, y: np.maximum(x / 1000.0, np.where(x < 500, 1.0, 0.2)))
        d.set_boundary({tag: Reflective_boundary(d) for tag in m.get_tag_names()})
        for _ in d.evolve(yieldstep=1.0, finaltime=2.0):
            pass
        out.append(d.quantities["stage"].centroid_values.copy())
        d.close()
    assert np.array_equal(out[0], out[1])


def test_device_build_rejects_bad_meshes():
    p = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0], [1.0, 1.0], [2.0, 2.0]])
    with pytest.raises(ValueError, match="out of range"):
        Mesh(p, [[0, 1, 7]], build="device")
    with pytest.raises(ValueError, match="degenerate"):
        Mesh(p, [[0, 3, 4]], build="device")  # collinear
    with pytest.raises(ValueError, match="non-manifold"):
        Mesh(p, [[0, 1, 2], [1, 3, 2], [2, 1, 4]], build="device")
