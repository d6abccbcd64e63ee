"""SWW output: the NetCDF writer (read back by scipy's independent NetCDF
reader), ANUGA's inverted triangle structure and node smoothing, and (GPU) the
stored records against the oracle's state at the same yields.

Pins: the NetCDF layout is checked by an independent reader
(scipy.io.netcdf_file); the node structure by ANUGA's documented example on
the reference's 4-triangle mesh (src/anuga_cuda/anuga_gpu/mytest.py:12-31);
the smoothing by a linear field (every triangle holds the node's exact value)
and by brute force.  ANUGA's own writer is not in this image, so attribute
names and values beyond those read back here are parity unpinned.
"""
import numpy as np
import pytest
from scipy.io import netcdf_file

from anuga_cuda_amd import scenarios
from anuga_cuda_amd.mesh import Mesh, rectangular_cross
from anuga_cuda_amd.sww import NetCDF3Writer, inverted_triangle_structure
from oracle import oracle as O


def _read(path):
    f = netcdf_file(str(path), "r", mmap=False)
    out = {"dims": dict(f.dimensions), "attrs": dict(f._attributes),
           "vars": {k: v.data.copy() for k, v in f.variables.items()}}
    f.close()
    return out


def test_netcdf_writer_reads_back_with_scipy(tmp_path):
    path = tmp_path / "t.nc"
    dims = [("n", 5), ("two", 2), ("t", None)]
    variables = [("a", ("n",), np.float32, {"units": "m"}), ("ij", ("n", "two"), np.int32, None),
                 ("time", ("t",), np.float64, None), ("q", ("t", "n"), np.float32, None),
                 ("r", ("t", "two"), np.float64, None)]
    w = NetCDF3Writer(str(path), dims, variables, {"title": "x", "order": 2, "start": 1.5})
    a = np.arange(5, dtype=np.float32) * 0.5
    ij = np.arange(10, dtype=np.int32).reshape(5, 2)
    w.write("a", a)
    w.write("ij", ij)
    rng = np.random.default_rng(1)
    recs = []
    for i in range(3):
        rec = {"time": 0.25 * i, "q": rng.normal(size=5).astype(np.float32), "r": rng.normal(size=2)}
        w.append_record(rec)
        recs.append(rec)
        got = _read(path)  # complete after every record
        assert got["vars"]["q"].shape == (i + 1, 5)
    w.close()
    got = _read(path)
    assert got["dims"] == {"n": 5, "two": 2, "t": None}
    assert got["attrs"]["title"] == b"x" and got["attrs"]["order"] == 2 and got["attrs"]["start"] == 1.5
    np.testing.assert_array_equal(got["vars"]["a"], a)
    np.testing.assert_array_equal(got["vars"]["ij"], ij)
    np.testing.assert_array_equal(got["vars"]["time"], [r["time"] for r in recs])
    np.testing.assert_array_equal(got["vars"]["q"], np.stack([r["q"] for r in recs]))
    np.testing.assert_array_equal(got["vars"]["r"], np.stack([r["r"] for r in recs]))


def test_netcdf_writer_single_record_variable_and_odd_sizes(tmp_path):
    path = tmp_path / "s.nc"
    w = NetCDF3Writer(str(path), [("n", 3), ("t", None)],
                      [("b", ("n",), np.int16, None), ("q", ("t", "n"), np.int16, None)])
    w.write("b", np.array([1, -2, 3], dtype=np.int16))
    for i in range(4):
        w.append_record({"q": np.array([i, i + 1, i + 2], dtype=np.int16)})
    w.close()
    got = _read(path)
    np.testing.assert_array_equal(got["vars"]["b"], [1, -2, 3])
    np.testing.assert_array_equal(got["vars"]["q"], [[i, i + 1, i + 2] for i in range(4)])


def test_inverted_triangle_structure_known_answer():
    """ANUGA's General_mesh.build_inverted_triangle_structure example, on the
    mesh of the reference's mytest.py: 6 nodes, triangles
    [1,0,2], [1,2,4], [4,2,5], [3,1,4]."""
    tris = np.array([[1, 0, 2], [1, 2, 4], [4, 2, 5], [3, 1, 4]])
    counts, vvi = inverted_triangle_structure(tris, 6)
    assert counts.tolist() == [1, 3, 3, 1, 3, 1]
    assert vvi.tolist() == [1, 0, 3, 10, 2, 4, 7, 9, 5, 6, 11, 8]


def test_node_smoothing_linear_field_and_brute_force():
    points, tris, boundary = rectangular_cross(7, 5, 70.0, 50.0)
    m = Mesh(points, tris, boundary=boundary)
    counts, vvi = inverted_triangle_structure(m.triangles, len(m.points))
    f = lambda x, y: 0.75 * x - 0.5 * y + 3.0
    vv = f(m.vertex_coordinates[:, 0::2], m.vertex_coordinates[:, 1::2])  # [N][3]
    avg = O.average_vertex_values(vvi, counts, vv)
    exact = f(m.points[:, 0], m.points[:, 1])
    np.testing.assert_allclose(avg, exact, rtol=4e-16, atol=0)
    rng = np.random.default_rng(2)
    vv = rng.normal(size=vv.shape)
    avg = O.average_vertex_values(vvi, counts, vv)
    for p in range(len(m.points)):
        k, j = np.nonzero(m.triangles == p)
        vals = vv[k, j][np.argsort(3 * k + j)]
        total = 0.0
        for x in vals:
            total += x
        assert avg[p] == total / len(vals)


def test_storable_rule():
    z = np.array([0.0, 0.0, 1.0, 2.0])
    w = np.array([0.0, 2e-5, 1.0 + 5e-6, 3.0])
    u = np.array([1.0, 2.0, 3.0, 4.0])
    v = -u
    out = O.sww_point_values(z, w, u, v, 1e-5)
    assert out.dtype == np.float32
    np.testing.assert_array_equal(out[1], np.float32([0.0, 2e-5, 1.0, 3.0]))
    np.testing.assert_array_equal(out[2], np.float32([0.0, 2.0, 0.0, 4.0]))
    np.testing.assert_array_equal(out[3], np.float32([0.0, -2.0, 0.0, -4.0]))


# ---------------------------------------------------------------- GPU
def _oracle_records(d, yieldstep, finaltime, smooth):
    od = O.OracleDomain.from_domain(d)
    m = d.mesh
    counts, vvi = inverted_triangle_structure(m.triangles, len(m.points))
    out = []
    for t in od.evolve(yieldstep, finaltime):
        vals = []
        for q in ("elevation", "stage", "xmomentum", "ymomentum"):
            vv = od.q(q, "vertex_values")
            vals.append(O.average_vertex_values(vvi, counts, vv) if smooth else vv.ravel())
        out.append((t, O.sww_point_values(*vals, d.minimum_storable_height)))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("smooth", [True, False], ids=["nodes", "unique_vertices"])
def test_sww_records_match_oracle_at_yields(tmp_path, smooth):
    d = scenarios.dam_break_channel(24, 24, high=2.0, low=0.0, bump=0.5)
    d.set_name("dam")
    d.set_datadir(str(tmp_path))
    d.set_store(True)
    d.set_store_vertices_uniquely(not smooth)
    expected = _oracle_records(d, 0.5, 2.0, smooth)
    times = [t for t in d.evolve(yieldstep=0.5, finaltime=2.0)]
    d.close()
    got = _read(tmp_path / "dam.sww")
    m = d.mesh
    P = len(m.points) if smooth else 3 * m.number_of_elements
    assert got["dims"]["number_of_points"] == P
    assert got["dims"]["number_of_volumes"] == m.number_of_elements
    assert got["attrs"]["smoothing"] == (b"Yes" if smooth else b"No")
    np.testing.assert_array_equal(got["vars"]["time"], times)
    assert len(expected) == len(times) == got["vars"]["stage"].shape[0]
    if smooth:
        np.testing.assert_array_equal(got["vars"]["volumes"], m.triangles)
        np.testing.assert_array_equal(got["vars"]["x"], m.points[:, 0].astype(np.float32))
    else:
        np.testing.assert_array_equal(got["vars"]["y"], m.vertex_coordinates[:, 1::2].ravel().astype(np.float32))
    np.testing.assert_array_equal(got["vars"]["elevation"], expected[0][1][0])
    exact = 0
    for i, (t, e) in enumerate(expected):
        assert abs(t - times[i]) < 1e-12
        for j, q in enumerate(("stage", "xmomentum", "ymomentum")):
            g = got["vars"][q][i]
            # FP32 of FP64 values equal to 1e-12 (Manning's pow may differ in the last FP64 bit)
            np.testing.assert_allclose(g, e[j + 1], rtol=1e-6, atol=1e-6 * max(1.0, float(np.abs(e[j + 1]).max())))
            exact += int(np.array_equal(g, e[j + 1]))
        assert got["vars"]["stage_range"][0] <= got["vars"]["stage"][: i + 1].min()
    print("bit-identical records: %d of %d" % (exact, 3 * len(expected)))
    wet = got["vars"]["stage"][-1] > got["vars"]["elevation"]
    assert wet.any() and (~wet).any()  # a dry region stays, stored as the bed


@pytest.mark.gpu
def test_sww_file_continues_across_evolve_calls(tmp_path):
    """A second evolve() continues the simulation and appends to the same
    file (its first yield, at the time the first call ended, is stored
    again, as ANUGA stores every yield)."""
    d = scenarios.dam_break_channel(8, 8)
    d.set_name("cont")
    d.set_datadir(str(tmp_path))
    d.set_store(True)
    t1 = list(d.evolve(yieldstep=0.5, finaltime=1.0))
    t2 = list(d.evolve(yieldstep=0.5, finaltime=2.0))
    d.close()
    got = _read(tmp_path / "cont.sww")
    np.testing.assert_array_equal(got["vars"]["time"], t1 + t2)
    assert got["vars"]["stage"].shape[0] == len(t1) + len(t2)


def test_quantities_to_be_stored_validation():
    d = scenarios.dam_break_channel(4, 4)
    d.set_store(True)
    d.set_quantities_to_be_stored(None)
    assert d.store is False
    with pytest.raises(NotImplementedError):
        d.set_quantities_to_be_stored({"friction": 1})
    d.set_quantities_to_be_stored({"elevation": 1, "stage": 2})
    assert d.quantities_to_be_stored == {"elevation": 1, "stage": 2}


@pytest.mark.gpu
def test_sww_stores_only_the_requested_quantities(tmp_path):
    d = scenarios.dam_break_channel(8, 8)
    d.set_name("stage_only")
    d.set_datadir(str(tmp_path))
    d.set_store(True)
    d.set_quantities_to_be_stored({"elevation": 1, "stage": 2})
    times = list(d.evolve(yieldstep=0.5, finaltime=1.0))
    d.close()
    got = _read(tmp_path / "stage_only.sww")
    assert "stage" in got["vars"] and "xmomentum" not in got["vars"] and "ymomentum_range" not in got["vars"]
    assert got["vars"]["stage"].shape[0] == len(times)


@pytest.mark.gpu
def test_reference_merimbula_script_end_to_end(tmp_path):
    """The reference's Merimbula script (tests/CUDA/merimbula/merimbula.py:86-173)
    on this package: create_domain_from_file, the stage bump, reflective
    boundaries, SWW storage, yields every 50 s.  The steps to 50 s are the
    reference's published 153, and the stored stage equals the oracle's."""
    import os

    from anuga_cuda_amd import GPU_domain, Reflective_boundary, create_domain_from_file

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    d = create_domain_from_file(os.path.join(root, "tests", "golden", "merimbula_43200.tsh.gz"), GPU_domain)
    d.set_quantity("stage", lambda x, y: 2.0 * ((x > 756000.0) & (x < 756500.0)) + 1.0)
    d.set_boundary({tag: Reflective_boundary(d) for tag in d.mesh.get_tag_names()})
    d.set_datadir(str(tmp_path))
    d.set_name("merimbula_new")
    d.set_store(True)
    expected = _oracle_records(d, 50.0, 100.0, True)
    steps = []
    for t in d.evolve(yieldstep=50, finaltime=100):
        steps.append(d.number_of_steps)
    d.close()
    assert steps[1] == 153
    got = _read(tmp_path / "merimbula_new.sww")
    assert got["vars"]["stage"].shape == (3, len(d.mesh.points))
    for i, (t, e) in enumerate(expected):
        np.testing.assert_allclose(got["vars"]["stage"][i], e[1], rtol=1e-6, atol=1e-6)
