"""CPU checks of the C ABI library (loads, exports every declared symbol;
no compute without a GPU), of the host mesh geometry and of the scenario
builders."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from anuga_cuda_amd import _lib, scenarios
from anuga_cuda_amd.mesh import Mesh, delaunay_random, morton_order, read_tsh, rectangular_cross

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "anuga_sw.h")).read()
    return sorted(set(re.findall(r"\b(sw_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libanuga_sw.so not built")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    declared = _declared()
    assert len(declared) >= 30
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(_lib.exported_symbols())


def test_struct_layouts_match_header(tmp_path):
    """Field offsets and sizes of the ctypes structs equal those the C compiler
    gives the structs of include/anuga_sw.h."""
    structs = {"sw_params": _lib.SwParams, "sw_mesh": _lib.SwMesh, "sw_state": _lib.SwState}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "anuga_sw.h"', "int main(void) {"]
    for cname, cls in structs.items():
        lines.append('printf("%s size %%zu\\n", sizeof(%s));' % (cname, cname))
        for fname, _ in cls._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, fname, cname, fname))
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = dict(line.rsplit(" ", 1) for line in out if line)
    for cname, cls in structs.items():
        assert int(got["%s size" % cname]) == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert int(got["%s.%s" % (cname, fname)]) == getattr(cls, fname).offset, (cname, fname)


def test_rectangular_cross_geometry():
    p, t, b = rectangular_cross(4, 3, 8.0, 6.0)
    m = Mesh(p, t, boundary=b)
    assert m.number_of_elements == 48
    assert m.number_of_boundary_elements == 2 * (4 + 3)
    np.testing.assert_allclose(m.areas.sum(), 48.0)
    # sum of l_i n_i over a closed triangle vanishes
    ln = m.normals.reshape(-1, 3, 2) * m.edgelengths[:, :, None]
    np.testing.assert_allclose(ln.sum(1), 0.0, atol=1e-12)
    # neighbour relation is symmetric and shared edges have opposite normals
    N = m.number_of_elements
    for k in range(N):
        for i in range(3):
            n = m.neighbours[k, i]
            if n >= 0:
                e = m.neighbour_edges[k, i]
                assert m.neighbours[n, e] == k and m.neighbour_edges[n, e] == i
                np.testing.assert_array_equal(m.normals[n, 2 * e:2 * e + 2], -m.normals[k, 2 * i:2 * i + 2])
                assert m.edgelengths[n, e] == m.edgelengths[k, i]
            else:
                b_ = -n - 1
                assert m.boundary_cells[b_] == k and m.boundary_edges[b_] == i
    assert set(m.get_tag_names()) == {"left", "right", "top", "bottom"}
    # radii: smallest centroid-to-midpoint distance
    assert np.all(m.radii > 0) and np.all(m.radii <= np.sqrt(m.areas))


def test_orientation_is_fixed_and_renumbering_is_a_relabelling():
    p, t = delaunay_random(400, extent=(100.0, 100.0), seed=3)
    t2 = t.copy()
    t2[::2, 1], t2[::2, 2] = t[::2, 2], t[::2, 1]  # half clockwise
    m = Mesh(p, t2)
    v = m.vertex_coordinates
    cross = (v[:, 2] - v[:, 0]) * (v[:, 5] - v[:, 1]) - (v[:, 4] - v[:, 0]) * (v[:, 3] - v[:, 1])
    assert np.all(cross > 0)
    mm = m.morton_renumbered()
    perm = morton_order(*m.centroid_coordinates.T)
    np.testing.assert_array_equal(mm.areas, m.areas[perm])
    np.testing.assert_array_equal(mm.normals, m.normals[perm])
    assert mm.number_of_boundary_elements == m.number_of_boundary_elements


def test_merimbula_fixture_reads():
    p, t, s = read_tsh(scenarios.MERIMBULA_TSH)
    assert p.shape == (22137, 2) and t.shape == (43200, 3) and len(s) == 1072
    m = Mesh(p, t, segments=s)
    assert m.number_of_boundary_elements == 1072
    assert set(m.get_tag_names()) == {"exterior", "open"}


def test_dam_break_scenario_has_wet_and_dry_cells():
    d = scenarios.dam_break_channel(20, 20)
    h = d.quantities["stage"].centroid_values - d.quantities["elevation"].centroid_values
    assert (h <= 0).any() and (h > 0.1).any()
    assert d.number_of_elements == 1600


def test_tsunami_flow_algorithm_has_no_evolve_step():
    """'tsunami' selects only its protect: evolve() and ensembles refuse it
    before any device work (the reference defines no such step)."""
    from anuga_cuda_amd import Ensemble
    d = scenarios.dam_break_channel(4, 4)
    assert d.flow_algorithm != "tsunami"
    d.set_flow_algorithm("tsunami")
    with pytest.raises(NotImplementedError):
        next(d.evolve(yieldstep=0.1, finaltime=0.2))
    assert d._handle is None
    with pytest.raises(NotImplementedError):
        Ensemble([d])


def test_product_path_has_no_cpu_fallback(monkeypatch):
    monkeypatch.setattr(_lib, "LIB_PATH", "/nonexistent/libanuga_sw.so")
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(ImportError):
        _lib.load()


def test_one_hip_runtime_per_process():
    """Loading the package's HIP libraries before torch must not leave two HIP
    runtimes mapped (the second would not see the GPU).  Runs in a fresh
    interpreter so that nothing else has imported torch yet."""
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libanuga_sw.so not built")
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from anuga_cuda_amd import _lib, learned\n"
            "_lib.load(); learned._load()\n"
            "import torch\n"
            "maps = {l.split()[-1] for l in open('/proc/self/maps') if 'libamdhip64' in l}\n"
            "print(len(maps))\n" % ROOT)
    out = subprocess.run(["python", "-c", code], capture_output=True, text=True, check=True)
    assert out.stdout.strip().splitlines()[-1] == "1", out.stdout + out.stderr


def test_neighbours_match_the_reference_mesh_files():
    """The host neighbour build against the neighbour columns the reference's
    own mesh files hold: tests/CUDA/cairns/cairns.msh (NetCDF) and
    tests/CUDA/merimbula/merimbula_43200.tsh (text)."""
    from anuga_cuda_amd.mesh import read_mesh

    for name, ntri in (("cairns.msh.gz", 17831), ("merimbula_43200.tsh.gz", 43200)):
        p, t, segs, extra = read_mesh(os.path.join(ROOT, "tests", "golden", name), full=True)
        assert len(t) == ntri
        m = Mesh(p, t, segments=segs)
        assert np.array_equal(m.triangles, t)  # both files are counter-clockwise
        nb = np.where(m.neighbours < 0, -1, m.neighbours)
        assert np.array_equal(nb, extra["triangle_neighbors"]), name


def test_read_msh_cairns_fixture():
    from anuga_cuda_amd.mesh import read_msh

    p, t, segs, extra = read_msh(os.path.join(ROOT, "tests", "golden", "cairns.msh.gz"), full=True)
    assert p.shape == (8974, 2) and t.shape == (17831, 3) and len(segs) == 790
    geo = extra["georeference"]
    assert geo["zone"] == -1 and geo["false_easting"] == 500000 and geo["datum"] == "wgs84"
    assert abs(geo["xllcorner"] - 301278.51) < 1e-9
    m = Mesh(p, t, segments=segs)
    tags, counts = np.unique(m.boundary_tags, return_counts=True)
    assert dict(zip(tags.tolist(), counts.tolist())) == {"bottom": 25, "ocean_east": 35, "onshore": 20, "top": 35}


def test_create_domain_from_file_on_the_reference_meshes(tmp_path):
    """create_domain_from_file (the reference's Merimbula scenario builds its
    domain with it): the mesh, boundary tags and georeference of the file;
    vertex attribute columns become quantities."""
    from anuga_cuda_amd import create_domain_from_file, scenarios

    d = create_domain_from_file(os.path.join(ROOT, "tests", "golden", "merimbula_43200.tsh.gz"))
    ref = scenarios.merimbula()
    assert np.array_equal(d.mesh.triangles, ref.mesh.triangles)
    assert d.mesh.get_tag_names() == ref.mesh.get_tag_names()
    assert d.geo_reference == {"xllcorner": 0.0, "yllcorner": 0.0, "zone": -1}
    c = create_domain_from_file(os.path.join(ROOT, "tests", "golden", "cairns.msh.gz"))
    assert c.number_of_elements == 17831 and c.mesh.get_tag_names() == ["bottom", "ocean_east", "onshore", "top"]
    assert abs(c.geo_reference["xllcorner"] - 301278.51) < 1e-9 and c.geo_reference["zone"] == -1
    # a small .tsh with an elevation attribute column
    lines = ["4 1 # verts, attributes", "0 0.0 0.0 1.5", "1 1.0 0.0 2.5", "2 0.0 1.0 3.5", "3 1.0 1.0 4.5",
             "# attribute column titles", "elevation", "2 # triangles", "0 0 1 2 -1 1 -1", "1 1 3 2 -1 0 -1",
             "1 # segments", "0 0 1 bottom", "#geo reference", "56", "100.0", "200.0"]
    f = tmp_path / "tiny.tsh"
    f.write_text("\n".join(lines) + "\n")
    t = create_domain_from_file(str(f))
    np.testing.assert_array_equal(t.quantities["elevation"].vertex_values, [[1.5, 2.5, 3.5], [2.5, 4.5, 3.5]])
    assert t.geo_reference == {"xllcorner": 100.0, "yllcorner": 200.0, "zone": 56}
    assert sorted(t.mesh.get_tag_names()) == ["bottom", "exterior"]


def test_rocprof_summary_base_names():
    from anuga_cuda_amd.rocprof_summary import base_name

    assert base_name("_ZN3gnn11k_mp_directEllPKiPKDF16bS3_") == "k_mp_direct"
    assert base_name("gnn::k_apply") == "k_apply"
    assert base_name("gnn::k_apply_groups") == "k_apply_groups"
    assert base_name("void sw::k_flux<double, true, 256>(sw::DevMesh, sw::DevFields)") == "k_flux"
    assert base_name("(anonymous namespace)::k_mesh_pairs(long long, int const*)") == "k_mesh_pairs"
