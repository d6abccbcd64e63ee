"""SWW output: ANUGA's NetCDF result file, written at yields.

The reference's GPU domain stores a timestep by handing over to ANUGA's SWW
writer (``store_timestep``, ``src/anuga_cuda/gpu_domain_advanced.py:2167-2172``)
after ``copy_back_necessary_data`` (``:734-737``), which copies nothing back.
Here the device reduces the state to exactly what the file holds -- FP32 point
values of elevation, stage and the momenta, smoothed over the triangles of
each mesh node and masked by the minimum storable height (``sw_sww_values``)
-- so 16 bytes per point cross PCIe instead of the FP64 vertex arrays
(96 bytes per triangle), and the host only appends them to the file.

File layout (ANUGA's ``Write_sww``, restated; the anuga package is not part of
the reference): NetCDF classic with 64-bit offsets; dimensions
``number_of_volumes``, ``number_of_vertices`` (3), ``numbers_in_range`` (2),
``number_of_points``, ``number_of_timesteps`` (unlimited); variables ``x``,
``y``, ``elevation`` (FP32 per point), ``volumes`` (int32, [volumes][3]),
``time`` (FP64 per timestep), ``stage``, ``xmomentum``, ``ymomentum`` (FP32,
[timesteps][points]) and the ``*_range`` pairs (min, max over what was
stored); global attributes ``smoothing``, ``vertices_are_stored_uniquely``,
``order``, ``starttime`` and the georeference (``xllcorner``, ``yllcorner``,
``zone``, ``false_easting``, ``false_northing``, ``datum``, ``projection``,
``units``).

The NetCDF writer is this module's own (the classic format specification:
big-endian header of dimensions, attributes and variables with their file
offsets, then fixed-size variables, then one record per timestep holding every
record variable); it streams records to disk, so a long run never holds more
than one timestep in memory.
"""
from __future__ import annotations

import os
import struct

import numpy as np

__all__ = ["NetCDF3Writer", "SWWWriter", "inverted_triangle_structure"]

_NC_DIMENSION, _NC_VARIABLE, _NC_ATTRIBUTE = 10, 11, 12
_NC_TYPE = {np.dtype("int8"): 1, np.dtype("S1"): 2, np.dtype("int16"): 3, np.dtype("int32"): 4,
            np.dtype("float32"): 5, np.dtype("float64"): 6}


def _pad4(n):
    return (n + 3) & ~3


def _name(s):
    b = s.encode("utf-8")
    return struct.pack(">i", len(b)) + b + b"\0" * (_pad4(len(b)) - len(b))


def _attr_value(v):
    """(nc_type, nelems, big-endian bytes padded to 4) of an attribute value."""
    if isinstance(v, str):
        b = v.encode("utf-8")
        return 2, len(b), b + b"\0" * (_pad4(len(b)) - len(b))
    a = np.atleast_1d(np.asarray(v))
    if a.dtype.kind in "iub":
        a = a.astype(np.int32)
    elif a.dtype.kind == "f":
        a = a.astype(np.float64)
    b = a.astype(a.dtype.newbyteorder(">")).tobytes()
    return _NC_TYPE[a.dtype], a.size, b + b"\0" * (_pad4(len(b)) - len(b))


def _att_list(attrs):
    if not attrs:
        return struct.pack(">ii", 0, 0)
    out = [struct.pack(">ii", _NC_ATTRIBUTE, len(attrs))]
    for k, v in attrs.items():
        t, n, b = _attr_value(v)
        out.append(_name(k) + struct.pack(">ii", t, n) + b)
    return b"".join(out)


class NetCDF3Writer:
    """Streaming writer of a NetCDF classic file with 64-bit offsets.

    ``dims``: list of (name, length), length None for the record (unlimited)
    dimension.  ``variables``: list of (name, dimension names, dtype, attrs).
    Fixed-size variables are written with ``write``; ``append_record`` writes
    one record of every record variable and updates the record count, so the
    file on disk is complete after every call.
    """

    def __init__(self, path, dims, variables, attrs=None):
        self.path = path
        self.dims = list(dims)
        dim_index = {name: i for i, (name, _) in enumerate(self.dims)}
        dim_len = dict(self.dims)
        self.vars = {}
        rec_vars = []
        fixed_vars = []
        for name, vdims, dtype, vattrs in variables:
            dtype = np.dtype(dtype)
            is_rec = bool(vdims) and dim_len[vdims[0]] is None
            shape = tuple(dim_len[d] for d in (vdims[1:] if is_rec else vdims))
            nbytes = int(np.prod(shape, dtype=np.int64)) * dtype.itemsize
            v = {"name": name, "dims": [dim_index[d] for d in vdims], "dtype": dtype, "attrs": vattrs or {},
                 "shape": shape, "record": is_rec, "vsize": _pad4(nbytes), "nbytes": nbytes, "begin": 0}
            self.vars[name] = v
            (rec_vars if is_rec else fixed_vars).append(v)
        # With a single record variable its records are not padded.
        if len(rec_vars) == 1:
            rec_vars[0]["vsize"] = rec_vars[0]["nbytes"]
        self.recsize = sum(v["vsize"] for v in rec_vars)
        self.numrecs = 0
        order = fixed_vars + rec_vars
        header = self._header(order, attrs or {})  # sizes are fixed: offsets are 8 bytes each
        off = len(header)
        for v in fixed_vars:
            v["begin"] = off
            off += v["vsize"]
        self.rec_begin = off
        for v in rec_vars:
            v["begin"] = off
            off += v["vsize"]
        header = self._header(order, attrs or {})
        self.f = open(path, "wb+")
        self.f.write(header)
        self.f.write(b"\0" * (self.rec_begin - len(header)))
        self.f.flush()

    def _header(self, order, attrs):
        out = [b"CDF\x02", struct.pack(">i", 0)]
        out.append(struct.pack(">ii", _NC_DIMENSION, len(self.dims)))
        for name, n in self.dims:
            out.append(_name(name) + struct.pack(">i", 0 if n is None else n))
        out.append(_att_list(attrs))
        out.append(struct.pack(">ii", _NC_VARIABLE, len(order)))
        for v in order:
            out.append(_name(v["name"]) + struct.pack(">i", len(v["dims"])))
            out.append(b"".join(struct.pack(">i", d) for d in v["dims"]))
            out.append(_att_list(v["attrs"]))
            out.append(struct.pack(">iiq", _NC_TYPE[v["dtype"]], v["vsize"], v["begin"]))
        return b"".join(out)

    def _write_var(self, v, array):
        """Writes one variable's values (big-endian, padded to its vsize) at
        the current file position, without an intermediate bytes copy."""
        a = np.ascontiguousarray(array, dtype=v["dtype"])
        if a.size * a.itemsize != v["nbytes"]:
            raise ValueError("%s: %d values for shape %s" % (v["name"], a.size, v["shape"]))
        self.f.write(memoryview(a.astype(v["dtype"].newbyteorder(">"), copy=False)).cast("B"))
        if v["vsize"] > v["nbytes"]:
            self.f.write(b"\0" * (v["vsize"] - v["nbytes"]))

    def write(self, name, array):
        v = self.vars[name]
        if v["record"]:
            raise ValueError("%s is a record variable (append_record)" % name)
        self.f.seek(v["begin"])
        self._write_var(v, array)

    def append_record(self, values):
        rec = [v for v in self.vars.values() if v["record"]]
        missing = [v["name"] for v in rec if v["name"] not in values]
        if missing:
            raise ValueError("record without %s" % ", ".join(missing))
        self.f.seek(self.rec_begin + self.numrecs * self.recsize)
        for v in rec:
            self._write_var(v, values[v["name"]])
        self.numrecs += 1
        self.f.seek(4)
        self.f.write(struct.pack(">i", self.numrecs))
        self.f.flush()

    def close(self):
        if self.f is not None:
            self.f.close()
            self.f = None


def inverted_triangle_structure(triangles, number_of_points):
    """ANUGA's General_mesh.build_inverted_triangle_structure:
    number_of_triangles_per_node[p] and vertex_value_indices, the flat indices
    3k+j of the triangle vertices grouped by node, ascending within a node."""
    flat = np.asarray(triangles, dtype=np.int64).ravel()
    counts = np.bincount(flat, minlength=number_of_points).astype(np.int64)
    return counts, np.argsort(flat, kind="stable").astype(np.int64)


class SWWWriter:
    """The SWW file of one domain: created with the mesh and the static
    elevation, then one record per ``store_timestep``.  The point values come
    from the device (``sw_sww_values``)."""

    QUANTITIES = ("stage", "xmomentum", "ymomentum")

    def __init__(self, domain, handle, path, smooth=True, minimum_storable_height=1.0e-5,
                 georeference=None, quantities=QUANTITIES):
        mesh = domain.mesh
        N = mesh.number_of_elements
        unknown = set(quantities) - set(self.QUANTITIES)
        if unknown:
            raise ValueError("SWW dynamic quantities are stage, xmomentum, ymomentum; got %s" % sorted(unknown))
        # stored every yield, in the file's order
        self.dynamic = [q for q in self.QUANTITIES if q in quantities]
        self.handle = handle
        self.path = path
        self.smooth = bool(smooth)
        self.minimum_storable_height = float(minimum_storable_height)
        geo = {"xllcorner": 0.0, "yllcorner": 0.0, "zone": -1, "false_easting": 500000,
               "false_northing": 10000000, "datum": "wgs84", "projection": "UTM", "units": "m"}
        geo.update(georeference or {})
        if self.smooth:
            P = len(mesh.points)
            counts, vvi = inverted_triangle_structure(mesh.triangles, P)
            handle.set_sww_points(counts, vvi)
            xy = mesh.points
            volumes = mesh.triangles
        else:
            P = 3 * N
            handle.set_sww_points()
            xy = mesh.vertex_coordinates.reshape(P, 2)
            volumes = np.arange(P, dtype=np.int64).reshape(N, 3)
        self.P = P
        dims = [("number_of_volumes", N), ("number_of_vertices", 3), ("numbers_in_range", 2),
                ("number_of_points", P), ("number_of_timesteps", None)]
        variables = [("x", ("number_of_points",), np.float32, None),
                     ("y", ("number_of_points",), np.float32, None),
                     ("elevation", ("number_of_points",), np.float32, None),
                     ("elevation_range", ("numbers_in_range",), np.float32, None),
                     ("volumes", ("number_of_volumes", "number_of_vertices"), np.int32, None)]
        for q in self.dynamic:
            variables.append((q + "_range", ("numbers_in_range",), np.float32, None))
        variables.append(("time", ("number_of_timesteps",), np.float64, None))
        for q in self.dynamic:
            variables.append((q, ("number_of_timesteps", "number_of_points"), np.float32, None))
        attrs = {"institution": "Geoscience Australia",
                 "description": "Output from anuga.abstract_2d_finite_volumes suitable for plotting",
                 "smoothing": "Yes" if self.smooth else "No",
                 "vertices_are_stored_uniquely": "False" if self.smooth else "True",
                 "order": int(domain.default_order), "starttime": float(domain.starttime)}
        attrs.update(geo)
        self.nc = NetCDF3Writer(path, dims, variables, attrs)
        # the domain's coordinates are already relative to its georeference
        self.nc.write("x", xy[:, 0])
        self.nc.write("y", xy[:, 1])
        self.nc.write("volumes", volumes)
        self.values = np.empty((4, P), dtype=np.float32)
        handle.sww_values(P, self.minimum_storable_height, self.values)
        z = self.values[0]
        self.nc.write("elevation", z)
        self.nc.write("elevation_range", [z.min(), z.max()] if P else [0.0, 0.0])
        big = np.finfo(np.float32).max
        self.ranges = {q: [big, -big] for q in self.dynamic}

    def store_timestep(self, time):
        v = self.handle.sww_values(self.P, self.minimum_storable_height, self.values)
        rec = {"time": float(time)}
        for i, q in enumerate(self.QUANTITIES):
            if q not in self.dynamic:
                continue
            a = v[i + 1]
            rec[q] = a
            if a.size:
                r = self.ranges[q]
                r[0] = min(r[0], float(a.min()))
                r[1] = max(r[1], float(a.max()))
                self.nc.write(q + "_range", r)
        self.nc.append_record(rec)

    def close(self):
        self.nc.close()


def sww_path(datadir, name):
    base = name if name.endswith(".sww") else name + ".sww"
    return os.path.join(datadir or ".", base)


def main(argv=None):
    """Times one SWW store on the Delaunay terrain workload (about 2 triangles
    per point): the device reduction and copy of the point values
    (sw_sww_values), the file append, and the alternative of downloading the
    FP64 vertex arrays and smoothing them on the host.  Prints one JSON line."""
    import argparse
    import json
    import tempfile
    import time

    from . import _lib, scenarios

    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--points", type=int, default=500000)
    ap.add_argument("--stores", type=int, default=20)
    ap.add_argument("--steps", type=int, default=20, help="solver steps between stores")
    args = ap.parse_args(argv)
    d = scenarios.delaunay_terrain(npoints=args.points)
    m = d.mesh
    N, P = m.number_of_elements, len(m.points)
    h = d._dev()
    h.call("sw_evolve_prepare", 1e9, 0.0, 0)
    h.evolve_steps(args.steps)
    with tempfile.TemporaryDirectory() as tmp:
        w = SWWWriter(d, h, os.path.join(tmp, "t.sww"))
        t_dev = t_file = 0.0
        for _ in range(args.stores):
            h.evolve_steps(args.steps)
            h.call("sw_synchronize")
            t0 = time.perf_counter()
            v = h.sww_values(P, w.minimum_storable_height, w.values)
            t1 = time.perf_counter()
            w.nc.append_record({"time": 0.0, "stage": v[1], "xmomentum": v[2], "ymomentum": v[3]})
            t2 = time.perf_counter()
            t_dev += t1 - t0
            t_file += t2 - t1
        size = os.path.getsize(w.path)
        w.close()
    # host smoothing of downloaded FP64 vertex arrays
    counts, vvi = inverted_triangle_structure(m.triangles, P)
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
    vv = np.empty(3 * N)
    t_host = 0.0
    reps = max(1, args.stores // 4)
    for _ in range(reps):
        t0 = time.perf_counter()
        for q in (_lib.SW_ELEVATION, _lib.SW_STAGE, _lib.SW_XMOM, _lib.SW_YMOM):
            h.download(q, _lib.SW_VERTEX, vv)
            np.add.reduceat(vv[vvi], starts) / counts
        t_host += time.perf_counter() - t0
    print(json.dumps({"triangles": N, "points": P, "stores": args.stores,
                      "device_ms_per_store": 1e3 * t_dev / args.stores,
                      "file_append_ms_per_store": 1e3 * t_file / args.stores,
                      "host_smoothing_ms_per_store": 1e3 * t_host / reps,
                      "bytes_to_host_per_store": 16 * P, "fp64_vertex_bytes_per_store": 4 * 24 * N,
                      "file_bytes": size}))
    d.close()


if __name__ == "__main__":
    main()
