"""Unstructured triangle meshes and their derived geometry.

This is the host-side mesh layer under ``Domain``.  It follows the array
conventions of ANUGA's ``General_mesh`` / ``Mesh`` classes that the reference
driver consumes (``src/anuga_cuda/gpu_domain_advanced.py:582-733`` uploads
exactly these arrays; their layout is spelled out in
``src/anuga_HMPP/sw_domain.h:60-180``):

* triangle ``k`` has vertices ``v0, v1, v2`` in counter-clockwise order;
* edge ``i`` is the edge opposite vertex ``i`` (edge 0 = v1->v2, edge 1 = v2->v0,
  edge 2 = v0->v1), as the reference's vertex->edge interpolation assumes
  (``src/anuga_cuda/interpolate_from_vertices_to_edges/*.cu:27-29``);
* ``normals[k, 2i:2i+2]`` is the outward unit normal of edge ``i``;
* ``neighbours[k, i]`` is the triangle across edge ``i``, or ``-(m+1)`` where
  ``m`` is the boundary index of that edge; boundary indices enumerate the
  boundary edges in ascending ``(vol_id, edge_id)`` order;
* ``surrogate_neighbours`` replaces boundary entries by ``k`` itself and
  ``number_of_boundaries[k]`` counts them (used by the extrapolation,
  ``src/anuga_HMPP/shallow_water_ext.c:2861-2880``);
* ``radii[k]`` is the smallest centroid-to-edge-midpoint distance (ANUGA's
  default ``use_inscribed_circle=False``).

Everything is computed with vectorised float64 numpy in a fixed operation
order so that the same mesh always yields bit-identical geometry.

Mesh generators: ``rectangular_cross`` (ANUGA's mesh factory of the same name,
used by the reference's channel scenarios ``tests/CUDA/channel1.py:25``),
``read_tsh`` (ANUGA's ``.tsh`` text format, the Merimbula fixture
``tests/CUDA/merimbula/merimbula_43200.tsh``) and ``delaunay_random`` (seeded
uniform points).  ``renumber`` reorders triangles (Morton / random order).
"""
from __future__ import annotations

import gzip
import io
import numpy as np

__all__ = [
    "Mesh",
    "rectangular_cross",
    "read_tsh",
    "read_msh",
    "read_mesh",
    "delaunay_random",
    "morton_order",
]


def rectangular_cross(m, n, len1=1.0, len2=1.0, origin=(0.0, 0.0)):
    """Return ``(points, triangles, boundary)`` of an m x n grid of squares, each
    split into 4 triangles around its centre point.

    Same construction as ANUGA's ``mesh_factory.rectangular_cross`` that the
    reference scenarios call (``tests/CUDA/channel1.py:25``,
    ``tests/CUDA/merimbula/merimbula.py:34``): grid points first (column
    major in i, then j), a centre point per cell, triangles left / bottom /
    right / top with edge 1 on the outer side.  ``boundary`` maps
    ``(vol_id, edge_id)`` to one of ``left, right, top, bottom``.
    """
    m = int(m)
    n = int(n)
    delta1 = float(len1) / m
    delta2 = float(len2) / n
    ii, jj = np.meshgrid(np.arange(m + 1), np.arange(n + 1), indexing="ij")
    grid = np.empty(((m + 1) * (n + 1), 2))
    grid[:, 0] = delta1 * ii.ravel() + origin[0]
    grid[:, 1] = delta2 * jj.ravel() + origin[1]

    def vid(i, j):
        return i * (n + 1) + j

    ci, cj = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
    ci = ci.ravel()
    cj = cj.ravel()
    v1 = vid(ci, cj + 1)
    v2 = vid(ci, cj)
    v3 = vid(ci + 1, cj + 1)
    v4 = vid(ci + 1, cj)
    # centre = (p1 + p2 + p3 + p4) * 0.25, summed left to right as in the factory
    cx = (((grid[v1, 0] + grid[v2, 0]) + grid[v3, 0]) + grid[v4, 0]) * 0.25
    cy = (((grid[v1, 1] + grid[v2, 1]) + grid[v3, 1]) + grid[v4, 1]) * 0.25
    ncell = m * n
    v5 = (m + 1) * (n + 1) + np.arange(ncell)
    points = np.concatenate([grid, np.stack([cx, cy], axis=1)], axis=0)

    tris = np.empty((ncell, 4, 3), dtype=np.int64)
    tris[:, 0] = np.stack([v2, v5, v1], axis=1)  # left
    tris[:, 1] = np.stack([v4, v5, v2], axis=1)  # bottom
    tris[:, 2] = np.stack([v3, v5, v4], axis=1)  # right
    tris[:, 3] = np.stack([v1, v5, v3], axis=1)  # top
    tris = tris.reshape(-1, 3)

    boundary = {}
    base = 4 * np.arange(ncell)
    for sel, off, tag in (
        (ci == 0, 0, "left"),
        (cj == 0, 1, "bottom"),
        (ci == m - 1, 2, "right"),
        (cj == n - 1, 3, "top"),
    ):
        for k in base[sel] + off:
            boundary[(int(k), 1)] = tag
    return points, tris, boundary


def read_tsh(path, full=False):
    """Read ANUGA's ``.tsh`` text mesh format (optionally gzip-compressed).

    Sections (see the header lines of
    ``tests/CUDA/merimbula/merimbula_43200.tsh:1,22139,22140,65341``): vertices
    ``<id> <x> <y> [attrs]``, vertex-attribute titles, triangles
    ``<id> <v0> <v1> <v2> <n0> <n1> <n2> [region attr]``, segments
    ``<id> <va> <vb> [tag]``, then holes / regions / geo reference.

    Returns ``(points, triangles, segments)`` where ``segments`` maps the
    unordered vertex pair of each tagged segment to its tag; with ``full``
    also a dict with the file's ``triangle_neighbors`` (the Mesh rebuilds
    neighbours from the triangles; the file's columns serve as a check).
    """
    opener = gzip.open if str(path).endswith(".gz") else open
    with opener(path, "rt") as fh:
        lines = fh.read().splitlines()
    pos = 0

    def header():
        nonlocal pos
        while lines[pos].startswith("#") or not lines[pos].strip():
            pos += 1
        fields = lines[pos].split("#")[0].split()
        pos += 1
        return fields

    nv, nattr = (int(x) for x in header()[:2])
    vtx = np.loadtxt(io.StringIO("\n".join(lines[pos:pos + nv])), ndmin=2)
    pos += nv
    points = vtx[:, 1:3].astype(np.float64)
    attributes = vtx[:, 3:3 + nattr].astype(np.float64) if nattr else None
    # vertex attribute titles
    while lines[pos].startswith("#"):
        pos += 1
    titles = [lines[pos + i].strip() for i in range(nattr)]
    pos += nattr
    fields = header()
    ntri = int(fields[0])
    tri_rows = [lines[pos + t].split() for t in range(ntri)]
    pos += ntri
    triangles = np.array([[int(r[1]), int(r[2]), int(r[3])] for r in tri_rows], dtype=np.int64)
    neighbors = np.array([[int(r[4]), int(r[5]), int(r[6])] for r in tri_rows], dtype=np.int64)
    fields = header()
    nseg = int(fields[0])
    segments = {}
    for s in range(nseg):
        r = lines[pos + s].split()
        tag = r[3] if len(r) > 3 else ""
        a, b = int(r[1]), int(r[2])
        segments[(min(a, b), max(a, b))] = tag
    if full:
        geo = {}
        for i, line in enumerate(lines):
            if line.strip().lower() == "#geo reference":
                vals = [x.strip() for x in lines[i + 1:i + 4]]
                try:
                    geo = {"zone": int(vals[0]), "xllcorner": float(vals[1]), "yllcorner": float(vals[2])}
                except (ValueError, IndexError):
                    geo = {}
                break
        return points, triangles, segments, {"triangle_neighbors": neighbors, "georeference": geo,
                                             "vertex_attributes": attributes, "vertex_attribute_titles": titles}
    return points, triangles, segments


def read_msh(path, full=False):
    """Read ANUGA's NetCDF ``.msh`` mesh format (optionally gzip-compressed),
    as the reference's Cairns scenario ships it (``tests/CUDA/cairns/cairns.msh``):
    variables ``vertices`` [P][2], ``triangles`` [N][3], ``triangle_neighbors``
    [N][3], ``segments`` [S][2] with ``segment_tags`` [S][chars] (blank-padded),
    and the georeference as global attributes.

    Returns ``(points, triangles, segments)`` as ``read_tsh``; with ``full``
    also a dict with ``triangle_neighbors`` and ``georeference``.  Vertex
    coordinates are relative to the georeference's ``xllcorner`` /
    ``yllcorner``, as stored."""
    from scipy.io import netcdf_file

    if str(path).endswith(".gz"):
        with gzip.open(path, "rb") as fh:
            src = io.BytesIO(fh.read())
    else:
        src = str(path)
    f = netcdf_file(src, "r", mmap=False)
    try:
        v = f.variables
        points = np.array(v["vertices"].data, dtype=np.float64)
        triangles = np.array(v["triangles"].data, dtype=np.int64)
        segments = {}
        if "segments" in v:
            segs = np.array(v["segments"].data, dtype=np.int64)
            tags = v["segment_tags"].data if "segment_tags" in v else None
            for i, (a, b) in enumerate(segs.tolist()):
                tag = b"".join(tags[i]).decode("utf-8", "replace").strip("\0 ") if tags is not None else ""
                segments[(min(a, b), max(a, b))] = tag
        attributes, titles = None, []
        if "vertex_attributes" in v:
            attributes = np.array(v["vertex_attributes"].data, dtype=np.float64).reshape(len(points), -1)
            if "vertex_attribute_titles" in v:
                titles = [b"".join(r).decode("utf-8", "replace").strip("\0 ")
                          for r in v["vertex_attribute_titles"].data]
        extras = {"vertex_attributes": attributes, "vertex_attribute_titles": titles,
                  "triangle_neighbors": np.array(v["triangle_neighbors"].data, dtype=np.int64)
                  if "triangle_neighbors" in v else None,
                  "georeference": {k: (a.decode() if isinstance(a, bytes) else a.item() if hasattr(a, "item") else a)
                                   for k, a in f._attributes.items()
                                   if k in ("xllcorner", "yllcorner", "zone", "false_easting", "false_northing",
                                            "datum", "projection", "units")}}
    finally:
        f.close()
    return (points, triangles, segments, extras) if full else (points, triangles, segments)


def read_mesh(path, full=False):
    """``read_tsh`` or ``read_msh`` by the file name (``.tsh`` / ``.msh``, optionally ``.gz``)."""
    name = str(path)[:-3] if str(path).endswith(".gz") else str(path)
    if name.endswith(".msh"):
        return read_msh(path, full=full)
    if name.endswith(".tsh"):
        return read_tsh(path, full=full)
    raise ValueError("unknown mesh format: %s" % path)


def delaunay_random(npoints, extent=(10000.0, 10000.0), seed=12345):
    """Delaunay triangulation of ``npoints`` seeded uniform points in
    ``[0, Lx] x [0, Ly]`` plus evenly spaced points on its sides (spacing equal
    to the mean point spacing, corners included), so the hull is the square
    and its triangles are well shaped.  Returns ``(points, triangles)``; all
    hull edges are boundary ('exterior')."""
    from scipy.spatial import Delaunay  # setup-time only

    rng = np.random.default_rng(seed)
    lx, ly = extent
    pts = rng.random((int(npoints), 2)) * np.array([lx, ly])
    spacing = np.sqrt(lx * ly / max(int(npoints), 1))
    nx = max(int(round(lx / spacing)), 1)
    ny = max(int(round(ly / spacing)), 1)
    sx = np.linspace(0.0, lx, nx + 1)
    sy = np.linspace(0.0, ly, ny + 1)[1:-1]
    side = np.concatenate([
        np.stack([sx, np.zeros_like(sx)], 1), np.stack([sx, np.full_like(sx, ly)], 1),
        np.stack([np.zeros_like(sy), sy], 1), np.stack([np.full_like(sy, lx), sy], 1)])
    # keep interior points off the sides
    pts = np.clip(pts, 1e-6 * lx, (1 - 1e-6) * np.array([lx, ly]))
    pts = np.concatenate([side, pts], axis=0)
    tri = Delaunay(pts).simplices.astype(np.int64)
    return pts, tri


def delaunay_jittered(npoints, extent=(1000.0, 1000.0), seed=12345, jitter=0.35):
    """Delaunay triangulation of about ``npoints`` seeded points: a square grid
    of spacing ``s`` (``npoints`` cells) whose interior points are moved by
    up to ``jitter * s`` in x and y, plus the sides sampled at spacing ``s``.
    Unstructured, and without the slivers of uniform random points.  Returns
    ``(points, triangles)``."""
    from scipy.spatial import Delaunay  # setup-time only

    rng = np.random.default_rng(seed)
    lx, ly = extent
    s = np.sqrt(lx * ly / max(int(npoints), 1))
    nx = max(int(round(lx / s)), 2)
    ny = max(int(round(ly / s)), 2)
    gx, gy = np.meshgrid(np.linspace(0.0, lx, nx + 1), np.linspace(0.0, ly, ny + 1), indexing="ij")
    hx, hy = lx / nx, ly / ny
    px = gx.copy()
    py = gy.copy()
    inner = (slice(1, -1), slice(1, -1))
    px[inner] += rng.uniform(-jitter, jitter, px[inner].shape) * hx
    py[inner] += rng.uniform(-jitter, jitter, py[inner].shape) * hy
    pts = np.stack([px.ravel(), py.ravel()], axis=1)
    tri = Delaunay(pts).simplices.astype(np.int64)
    return pts, tri


def morton_order(cx, cy, bits=16):
    """Permutation that sorts points by the Z-order (Morton) code of their
    coordinates quantised to ``2**bits`` cells per axis."""
    x0, x1 = cx.min(), cx.max()
    y0, y1 = cy.min(), cy.max()
    scale = (1 << bits) - 1
    qx = np.floor((cx - x0) / max(x1 - x0, 1e-300) * scale).astype(np.uint64)
    qy = np.floor((cy - y0) / max(y1 - y0, 1e-300) * scale).astype(np.uint64)

    def spread(v):
        v = v & np.uint64(0xFFFF)
        v = (v | (v << np.uint64(8))) & np.uint64(0x00FF00FF)
        v = (v | (v << np.uint64(4))) & np.uint64(0x0F0F0F0F)
        v = (v | (v << np.uint64(2))) & np.uint64(0x33333333)
        v = (v | (v << np.uint64(1))) & np.uint64(0x55555555)
        return v

    code = spread(qx) | (spread(qy) << np.uint64(1))
    return np.argsort(code, kind="stable")


class Mesh:
    """Triangle mesh plus every derived array the shallow-water step reads.

    ``boundary`` maps ``(vol_id, edge_id)`` to a tag for tagged boundary edges;
    untagged boundary edges get the tag ``'exterior'`` (ANUGA's default).
    ``segments`` (as returned by ``read_tsh``) can tag boundary edges by their
    vertex pair instead.  ``build="device"`` computes the oriented triangles,
    geometry and neighbours on GPU ``device`` (sw_build_mesh; the same bits as
    the host build); the boundary tagging stays on the host.
    """

    def __init__(self, points, triangles, boundary=None, segments=None, build="host", device=0):
        points = np.ascontiguousarray(points, dtype=np.float64)
        tris = np.ascontiguousarray(triangles, dtype=np.int64).copy()
        if tris.ndim != 2 or tris.shape[1] != 3:
            raise ValueError("triangles must be an (N, 3) index array")
        if points.ndim != 2 or points.shape[1] != 2:
            raise ValueError("points must be an (P, 2) array")
        if build not in ("host", "device"):
            raise ValueError("build must be 'host' or 'device'")
        self._build = build
        self._device = device
        self.points = points
        self.number_of_elements = len(tris)
        if build == "device":
            from . import _lib

            g = _lib.build_mesh_geometry(points, tris, device)
            flip = g["triangles"][:, 1] != tris[:, 1]
            for name, a in g.items():
                setattr(self, name, a)
            nb, ne = g["neighbours"].ravel(), g["neighbour_edges"].ravel()
            self._edge_partner = np.where(nb >= 0, 3 * nb + ne, -1)
        else:
            if tris.min() < 0 or tris.max() >= len(points):
                raise ValueError("triangle vertex index out of range")
            # Orientation: make every triangle counter-clockwise.  Swapping
            # v1<->v2 keeps v0 and relabels edges 1<->2.
            p = points[tris]
            cross = (p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1]) - (
                p[:, 2, 0] - p[:, 0, 0]) * (p[:, 1, 1] - p[:, 0, 1])
            if np.any(cross == 0.0):
                raise ValueError("degenerate triangle (zero area)")
            flip = cross < 0
            if flip.any():
                tris[flip, 1], tris[flip, 2] = tris[flip, 2].copy(), tris[flip, 1].copy()
            self.triangles = tris
            self._build_geometry()
            self._build_neighbours()
        if boundary and flip.any():
            # boundary keys follow the relabelled edges
            swapped = {}
            for (k, e), tag in boundary.items():
                if flip[k] and e in (1, 2):
                    e = 3 - e
                swapped[(k, e)] = tag
            boundary = swapped
        self._build_boundary(boundary or {}, segments or {})

    # ------------------------------------------------------------------
    def __len__(self):
        return self.number_of_elements

    def _build_geometry(self):
        p = self.points[self.triangles]  # (N, 3, 2)
        x0, y0 = p[:, 0, 0], p[:, 0, 1]
        x1, y1 = p[:, 1, 0], p[:, 1, 1]
        x2, y2 = p[:, 2, 0], p[:, 2, 1]
        N = self.number_of_elements
        self.vertex_coordinates = np.ascontiguousarray(p.reshape(N, 6))
        # centroid = ((x0 + x1 + x2)/3, (y0 + y1 + y2)/3)
        cx = ((x0 + x1) + x2) / 3.0
        cy = ((y0 + y1) + y2) / 3.0
        self.centroid_coordinates = np.ascontiguousarray(np.stack([cx, cy], axis=1))
        # area = |(x1*y0 - x0*y1) + (x2*y1 - x1*y2) + (x0*y2 - x2*y0)| / 2
        self.areas = np.abs(((x1 * y0 - x0 * y1) + (x2 * y1 - x1 * y2)) + (x0 * y2 - x2 * y0)) / 2.0
        # edge i runs from vertex i+1 to vertex i+2; outward normal is
        # (dy/l, -dx/l) for a counter-clockwise triangle.
        ex = np.stack([x2 - x1, x0 - x2, x1 - x0], axis=1)
        ey = np.stack([y2 - y1, y0 - y2, y1 - y0], axis=1)
        length = np.sqrt(ex * ex + ey * ey)
        if np.any(length == 0.0):
            raise ValueError("zero-length edge")
        nx = ey / length
        ny = -(ex / length)
        normals = np.empty((N, 6))
        normals[:, 0::2] = nx
        normals[:, 1::2] = ny
        self.normals = normals
        self.edgelengths = np.ascontiguousarray(length)
        # edge midpoints: edge 0 = (v1+v2)/2, edge 1 = (v0+v2)/2, edge 2 = (v1+v0)/2
        mx = np.stack([(x1 + x2) / 2, (x0 + x2) / 2, (x1 + x0) / 2], axis=1)
        my = np.stack([(y1 + y2) / 2, (y0 + y2) / 2, (y1 + y0) / 2], axis=1)
        em = np.empty((N, 6))
        em[:, 0::2] = mx
        em[:, 1::2] = my
        self.edge_midpoint_coordinates = em
        dx = cx[:, None] - mx
        dy = cy[:, None] - my
        d = np.sqrt(dx * dx + dy * dy)
        self.radii = np.minimum(np.minimum(d[:, 0], d[:, 1]), d[:, 2])

    def _build_neighbours(self):
        N = self.number_of_elements
        t = self.triangles
        a = np.stack([t[:, 1], t[:, 2], t[:, 0]], axis=1).ravel()
        b = np.stack([t[:, 2], t[:, 0], t[:, 1]], axis=1).ravel()
        lo = np.minimum(a, b)
        hi = np.maximum(a, b)
        nv = np.int64(len(self.points))
        key = lo * nv + hi
        order = np.argsort(key, kind="stable")
        ks = key[order]
        same = ks[1:] == ks[:-1]
        if np.any(same[1:] & same[:-1]):
            raise ValueError("non-manifold mesh: an edge is shared by more than two triangles")
        nb = np.full(3 * N, -1, dtype=np.int64)
        i0 = order[:-1][same]
        i1 = order[1:][same]
        nb[i0] = i1
        nb[i1] = i0
        self._edge_partner = nb  # flat edge index of the partner edge, or -1
        self.neighbours = np.where(nb >= 0, nb // 3, -1).reshape(N, 3)
        self.neighbour_edges = np.where(nb >= 0, nb % 3, -1).reshape(N, 3)

    def _build_boundary(self, boundary, segments):
        N = self.number_of_elements
        flat = np.nonzero(self._edge_partner < 0)[0]  # ascending (vol_id, edge_id)
        vol = flat // 3
        edge = flat % 3
        tags = []
        t = self.triangles
        for k, e in zip(vol.tolist(), edge.tolist()):
            tag = boundary.get((k, e))
            if tag is None and segments:
                va = t[k, (e + 1) % 3]
                vb = t[k, (e + 2) % 3]
                tag = segments.get((min(va, vb), max(va, vb)))
            if not tag:
                tag = "exterior"
            tags.append(tag)
        for key in boundary:
            k, e = key
            if not (0 <= k < N and 0 <= e < 3) or self._edge_partner[3 * k + e] >= 0:
                raise ValueError("boundary key %r is not a boundary edge" % (key,))
        self.boundary_cells = vol.astype(np.int64)
        self.boundary_edges = edge.astype(np.int64)
        self.boundary_tags = np.array(tags, dtype=object)
        self.number_of_boundary_elements = Nb = len(flat)
        nbr = self.neighbours.copy().ravel()
        nbr[flat] = -(np.arange(Nb) + 1)
        self.neighbours = nbr.reshape(N, 3)
        self.boundary = {(int(k), int(e)): tag for k, e, tag in zip(vol, edge, tags)}
        self.surrogate_neighbours = np.where(self.neighbours < 0, np.arange(N)[:, None], self.neighbours)
        self.number_of_boundaries = (self.neighbours < 0).sum(axis=1).astype(np.int64)
        self.tri_full_flag = np.ones(N, dtype=np.int64)

    # ------------------------------------------------------------------
    def get_tag_names(self):
        return sorted(set(self.boundary_tags.tolist()))

    def renumber(self, perm):
        """Return a new Mesh whose triangle ``j`` is triangle ``perm[j]`` of this
        one (local vertex order unchanged, so every per-triangle geometry value
        is bit-identical, only reindexed).  Boundary tags travel with their
        edges."""
        perm = np.asarray(perm, dtype=np.int64)
        N = self.number_of_elements
        if perm.shape != (N,) or len(np.unique(perm)) != N or perm.min() < 0 or perm.max() >= N:
            raise ValueError("perm is not a permutation")
        inv = np.empty_like(perm)
        inv[perm] = np.arange(len(perm))
        boundary = {(int(inv[k]), e): tag for (k, e), tag in self.boundary.items()}
        return Mesh(self.points, self.triangles[perm], boundary=boundary, build=self._build, device=self._device)

    def morton_renumbered(self):
        c = self.centroid_coordinates
        return self.renumber(morton_order(c[:, 0], c[:, 1]))

    def random_renumbered(self, seed=0):
        rng = np.random.default_rng(seed)
        return self.renumber(rng.permutation(self.number_of_elements))


def main(argv=None):
    """Times the host and the device build of the seeded Delaunay meshes'
    derived arrays (geometry, neighbours, boundary) and checks that both give
    the same arrays.  Prints one JSON line per mesh."""
    import argparse
    import json
    import time

    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--points", type=int, nargs="+", default=[500000, 2000000])
    ap.add_argument("--repeats", type=int, default=3)
    args = ap.parse_args(argv)
    for npts in args.points:
        pts, tris = delaunay_random(npts, extent=(10000.0, 10000.0), seed=12345)
        Mesh(*delaunay_random(100, seed=1), build="device")  # device and runtime warm-up
        t = {}
        meshes = {}
        for build in ("host", "device"):
            best = 1e30
            for _ in range(args.repeats):
                t0 = time.perf_counter()
                meshes[build] = Mesh(pts, tris, build=build)
                best = min(best, time.perf_counter() - t0)
            t[build] = best
        h, d = meshes["host"], meshes["device"]
        same = all(np.array_equal(getattr(h, n), getattr(d, n)) for n in (
            "triangles", "vertex_coordinates", "centroid_coordinates", "areas", "normals", "edgelengths",
            "edge_midpoint_coordinates", "radii", "neighbours", "neighbour_edges", "surrogate_neighbours",
            "boundary_cells", "boundary_edges"))
        print(json.dumps({"points": npts, "triangles": h.number_of_elements, "host_s": t["host"],
                          "device_s": t["device"], "speedup": t["host"] / t["device"], "identical": same}))


if __name__ == "__main__":
    main()
