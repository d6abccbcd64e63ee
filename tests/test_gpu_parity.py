"""HIP path vs the CPU FP64 oracle, through the C ABI, on the same inputs.

Bar (stated per comparison): stage / xmomentum / ymomentum within 1e-12
normwise relative error per step (bit mismatches are counted and reported),
identical wet/dry masks, identical step counts and timesteps within 1e-12.
Manning friction calls pow(), whose last bit may differ between libm and the
device math library; everything else is expected bit-identical.
"""
import numpy as np
import pytest

from anuga_cuda_amd import Ensemble, _lib, scenarios
from anuga_cuda_amd.domain import Dirichlet_boundary, Domain, Reflective_boundary, Transmissive_boundary
from anuga_cuda_amd.mesh import Mesh, rectangular_cross
from oracle.oracle import OracleDomain
from tests.helpers import FIELDS, compare, compare_state, device_array, wet_mask

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _dam(**kw):
    return scenarios.dam_break_channel(24, 24, **kw)


def _delaunay(order, **kw):
    return scenarios.delaunay_terrain(npoints=3000, order=order, extent=2000.0, mound_radius=200.0,
                                      seed=7, terrain_seed=11, **kw)


SCENARIOS = {
    "dam_rectcross": lambda: _dam(),
    "delaunay_morton": lambda: _delaunay("morton"),
    "delaunay_random": lambda: _delaunay("random"),
}


def _diagnose(domain, od, step):
    m = domain.mesh
    print("mismatch at step %d" % step)
    for q in ("stage", "xmomentum", "ymomentum"):
        for loc in ("explicit_update", "centroid_values", "vertex_values", "edge_values", "boundary_values"):
            g = device_array(domain, q, loc)
            e = od.q(q, loc).ravel()
            bad = np.nonzero(np.abs(g - e) > 1e-12 * max(np.max(np.abs(e)), 1e-300))[0]
            if bad.size:
                k = bad[:6] if loc in ("centroid_values", "explicit_update") else bad[:6]
                print("  %s.%s: %d bad, first %s" % (q, loc, bad.size, k.tolist()))
                for i in k[:4]:
                    tri = i if loc in ("centroid_values", "explicit_update") else (
                        m.boundary_cells[i] if loc == "boundary_values" else i // 3)
                    print("    idx %d tri %d nob %d nbr %s got %.17g exp %.17g" % (
                        i, tri, m.number_of_boundaries[tri], m.neighbours[tri].tolist(), g[i], e[i]))


def _lockstep(domain, nsteps, fused, stats, yieldstep=1e9, finaltime=None):
    od = OracleDomain.from_domain(domain)
    h = domain._dev()
    h.call("sw_evolve_prepare", yieldstep, finaltime or 0.0, int(finaltime is not None))
    od.evolve_prepare(yieldstep, finaltime)
    compare_state(domain, od, TOL, stats=stats, tag="prepare.")
    eps = domain.minimum_allowed_height
    bed = od.q("elevation", "centroid_values")
    dts = []
    for step in range(nsteps):
        if fused:
            done, ev = h.evolve_steps(1)
            assert done == 1
        else:
            ev = h.evolve_one_step_unfused()
        ev_o = od.evolve_step()
        s = h.get_state()
        assert ev == ev_o, "event mismatch at step %d" % step
        compare([s.timestep], [od.d.timestep], "timestep", TOL, stats)
        compare([s.time], [od.d.time], "time", TOL, stats)
        dts.append((s.timestep, od.d.timestep))
        try:
            compare_state(domain, od, TOL, stats=stats, tag="step.")
            for q in ("stage", "xmomentum", "ymomentum"):
                compare(device_array(domain, q, "explicit_update"), od.q(q, "explicit_update"),
                        "step.%s.explicit_update" % q, TOL, stats)
                if not fused:
                    compare(device_array(domain, q, "boundary_values"), od.q(q, "boundary_values"),
                            "step.%s.boundary_values" % q, TOL, stats)
        except AssertionError:
            _diagnose(domain, od, step)
            raise
        w = device_array(domain, "stage", "centroid_values")
        assert np.array_equal(wet_mask(w, bed, eps), wet_mask(od.q("stage", "centroid_values"), bed, eps)), \
            "wet/dry mask differs at step %d" % step
        if ev != _lib.SW_EVENT_NONE:
            break
    return od, dts


@pytest.mark.parametrize("name", sorted(SCENARIOS))
def test_fused_evolve_matches_oracle(name):
    d = SCENARIOS[name]()
    stats = {}
    od, dts = _lockstep(d, 25, fused=True, stats=stats)
    bits = {k: v[0] for k, v in stats.items() if v[0]}
    print("%s: N=%d steps=%d bit-mismatching fields %s" % (name, d.number_of_elements, len(dts), bits))
    ms = d._handle.max_speed()
    compare(ms, od.max_speed, "max_speed", TOL)


@pytest.mark.parametrize("name", ["dam_rectcross", "delaunay_random"])
def test_unfused_evolve_matches_oracle(name):
    d = SCENARIOS[name]()
    stats = {}
    _lockstep(d, 8, fused=False, stats=stats)
    for name_q in ("stage", "xmomentum", "ymomentum"):
        compare(device_array(d, name_q, "boundary_values"), OracleDomain.from_domain(d).q(name_q, "boundary_values"),
                name_q + ".boundary", TOL)


@pytest.mark.parametrize("method", ["rk2", "rk3"])
@pytest.mark.parametrize("fused", [True, False], ids=["fused", "unfused"])
def test_runge_kutta_steps_match_oracle(method, fused):
    """rk2 / rk3 steps (backup, stages at the first stage's timestep, saxpy
    combinations) in lockstep with the oracle's."""
    d = SCENARIOS["delaunay_random"]() if fused else _dam()
    d.set_timestepping_method(method)
    stats = {}
    od, dts = _lockstep(d, 16 if fused else 6, fused=fused, stats=stats)
    assert len(dts) == (16 if fused else 6)
    compare(d._handle.max_speed(), od.max_speed, "max_speed", TOL)


@pytest.mark.parametrize("limiter", ["vertex", "edge"])
@pytest.mark.parametrize("fused", [True, False], ids=["fused", "unfused"])
def test_per_quantity_limiters_match_oracle(limiter, fused):
    """The distribute paths of optimised_gradient_limiter=False (limit by
    vertex) and use_edge_limiter=True (limit by edge), each quantity with its
    own beta, in lockstep with the oracle."""
    kw = {"use_edge_limiter": 1} if limiter == "edge" else {"optimised_gradient_limiter": 0}
    if fused:
        d = scenarios.delaunay_terrain(npoints=3000, order="random", extent=2000.0, mound_radius=200.0,
                                       seed=7, terrain_seed=11, **kw)
    else:
        d = _dam(**kw)
    d.quantities["xmomentum"].beta = 0.7
    d.quantities["ymomentum"].beta = 1.3
    assert d.limiter_path() == limiter
    if fused:
        assert d._dev().step_path() == _lib.SW_STEP_VERTEX
    stats = {}
    _, dts = _lockstep(d, 16 if fused else 6, fused=fused, stats=stats)
    assert len(dts) == (16 if fused else 6)


def test_backup_and_saxpy_operators_match_oracle():
    d = _dam()
    for _ in d.evolve(yieldstep=0.5, finaltime=1.0):
        pass
    od = OracleDomain.from_domain(d)
    h = d._dev()
    s = h.get_state()
    s.yieldtime = 1e9
    s.event = 0
    h.set_state(s)
    od.d.yieldtime = 1e9
    h.call("sw_backup_conserved_quantities")
    od.backup_conserved_quantities()
    assert h.compute_fluxes() == od.compute_fluxes()
    h.call("sw_update_timestep", 1e9, 0.0, 0)
    od.update_timestep()
    assert h.get_state().timestep == od.d.timestep > 0
    h.call("sw_update_conserved_quantities")
    od.update_conserved_quantities()
    h.call("sw_saxpy_conserved_quantities", 2.0 / 3.0, 1.0 / 3.0)
    od.saxpy_conserved_quantities(2.0 / 3.0, 1.0 / 3.0)
    for q in ("stage", "xmomentum", "ymomentum"):
        compare(device_array(d, q, "centroid_values"), od.q(q, "centroid_values"), q + ".saxpy", TOL)


def test_operators_match_oracle_one_by_one():
    d = _dam()
    # advance a few steps so momenta and wet/dry fronts are non-trivial
    for t in d.evolve(yieldstep=0.5, finaltime=2.0):
        pass
    od = OracleDomain.from_domain(d)
    h = d._dev()
    s = h.get_state()
    s.yieldtime = 1e9
    s.event = 0
    h.set_state(s)
    od.d.yieldtime = 1e9
    st = {}
    ft = h.compute_fluxes()
    ft_o = od.compute_fluxes()
    assert ft == ft_o, (ft, ft_o)
    for q in ("stage", "xmomentum", "ymomentum"):
        compare(device_array(d, q, "explicit_update"), od.q(q, "explicit_update"), q + ".flux", TOL, st)
    compare(h.max_speed(), od.max_speed, "max_speed", 0.0)
    h.call("sw_gravity_wb")
    od.gravity_wb()
    for q in ("xmomentum", "ymomentum"):
        compare(device_array(d, q, "explicit_update"), od.q(q, "explicit_update"), q + ".gravity", 0.0, st)
    h.call("sw_manning_friction")
    od.manning_friction_implicit()
    for q in ("xmomentum", "ymomentum"):
        compare(device_array(d, q, "semi_implicit_update"), od.q(q, "semi_implicit_update"), q + ".friction",
                TOL, st)
    h.call("sw_update_timestep", 1e9, 0.0, 0)
    od.update_timestep(None)
    assert h.get_state().timestep == od.d.timestep
    h.call("sw_update_conserved_quantities")
    od.update_conserved_quantities()
    for q in ("stage", "xmomentum", "ymomentum"):
        compare(device_array(d, q, "centroid_values"), od.q(q, "centroid_values"), q + ".update", TOL, st)
    for op, oop in (("sw_protect", od.protect),
                    ("sw_extrapolate_second_order_sw", od.extrapolate_second_order_sw),
                    ("sw_balance_deep_and_shallow", od.balance_deep_and_shallow),
                    ("sw_interpolate_from_vertices_to_edges", od.interpolate_from_vertices_to_edges),
                    ("sw_update_boundary", od.update_boundary),
                    ("sw_update_other_quantities", od.update_other_quantities)):
        h.call(op)
        oop()
        compare_state(d, od, TOL, locs=("centroid_values", "vertex_values", "edge_values", "boundary_values"),
                      stats=st, tag=op + ".")
    for q in ("height", "xvelocity", "yvelocity"):
        for loc in ("centroid_values", "edge_values", "vertex_values", "boundary_values"):
            compare(device_array(d, q, loc), od.q(q, loc), "%s.%s" % (q, loc), TOL, st)
    print({k: v for k, v in st.items() if v[0]})


def test_protect_swb2_matches_oracle():
    """The 'tsunami' flow algorithm's protect (_protect_swb2,
    protect.cu:64-138), run by protect_against_infinitesimal_and_negative_heights
    when flow_algorithm is 'tsunami', bit for bit against the oracle on the
    3,000-point Delaunay terrain, with depths in every branch: deep, near the
    threshold, dry above the raise level, dry below it."""
    d = _delaunay("morton")
    d.set_flow_algorithm("tsunami")
    h = d._dev()
    od = OracleDomain.from_domain(d)
    N = d.number_of_elements
    rng = np.random.default_rng(12)
    zc = od.q("elevation", "centroid_values")
    low = od.q("elevation", "vertex_values").min(axis=1) - d.minimum_allowed_height
    pick = rng.integers(0, 5, size=N)
    depth = np.choose(pick, [np.full(N, 0.5), np.full(N, 2e-3), np.full(N, 0.0), (low - zc) * 0.5,
                             (low - zc) - 0.2])
    state = {("stage", _lib.SW_CENTROID, "centroid_values"): zc + depth,
             ("stage", _lib.SW_VERTEX, "vertex_values"): zc[:, None] + rng.uniform(-0.3, 0.3, size=(N, 3)),
             ("xmomentum", _lib.SW_CENTROID, "centroid_values"): rng.normal(scale=0.2, size=N),
             ("ymomentum", _lib.SW_CENTROID, "centroid_values"): rng.normal(scale=0.2, size=N)}
    code = dict(FIELDS)
    for (q, loc, name), a in state.items():
        h.upload(code[q], loc, a)
        od.q(q, name)[:] = a
    d.protect_against_infinitesimal_and_negative_heights()
    od.protect_swb2()
    for (q, _, name), a in state.items():
        got = device_array(d, q, name)
        assert np.array_equal(got, od.q(q, name).ravel()), (q, name)
    w = od.q("stage", "centroid_values")
    uh = od.q("xmomentum", "centroid_values")
    raised = w != state[("stage", _lib.SW_CENTROID, "centroid_values")]
    zeroed = uh == 0
    assert raised.any() and (zeroed & ~raised).any() and (~zeroed).any()
    # the default flow algorithm runs the default protect
    d.set_flow_algorithm("original")
    before = device_array(d, "stage", "centroid_values")
    d.protect_against_infinitesimal_and_negative_heights()
    od.protect()
    assert np.array_equal(device_array(d, "stage", "centroid_values"), od.q("stage", "centroid_values"))
    assert not np.array_equal(before, od.q("stage", "centroid_values"))


def test_restart_from_device_buffers_matches_oracle_restart():
    """sw_restart_device / sw_download_conserved_device: after 6 lockstep
    steps, the conserved centroid state goes to torch device buffers, is
    perturbed (a 1 cm rise of stage where wet), and the solver restarts from
    it at the current time; the oracle restarts from the same values
    (starttime = time, evolve_prepare).  6 more steps match the oracle to the
    FP64 bar, timesteps included."""
    import torch

    d = _delaunay("morton")
    od, _ = _lockstep(d, 6, True, {})
    h = d._handle
    N = d.number_of_elements
    w, uh, vh = (torch.empty(N, dtype=torch.float64, device="cuda") for _ in range(3))
    h.download_conserved_device(w.data_ptr(), uh.data_ptr(), vh.data_ptr())
    for name, t in (("stage", w), ("xmomentum", uh), ("ymomentum", vh)):
        compare(t.cpu().numpy(), od.q(name, "centroid_values"), "download." + name, TOL)
    bed = torch.tensor(od.q("elevation", "centroid_values"), device="cuda")
    w += torch.where(w - bed > 1e-3, 0.01, 0.0)
    t_now = od.d.time
    h.restart_device(w.data_ptr(), uh.data_ptr(), vh.data_ptr(), t_now, 1e9)
    od.q("stage", "centroid_values")[:] = w.cpu().numpy()
    od.d.starttime = t_now
    od.evolve_prepare(1e9, None)
    compare_state(d, od, TOL, tag="restart.")
    for step in range(6):
        done, ev = h.evolve_steps(1)
        od.evolve_step()
        s = h.get_state()
        compare([s.timestep, s.time], [od.d.timestep, od.d.time], "timestep/time", TOL)
        compare_state(d, od, TOL, tag="after restart.")


def test_evolve_generator_yields_match_oracle():
    d = _dam()
    od = OracleDomain.from_domain(d)
    times = []
    steps = []
    for t in d.evolve(yieldstep=0.7, finaltime=3.0):
        times.append(t)
        steps.append(d.number_of_steps)
    times_o = []
    steps_o = []
    for t in od.evolve(0.7, 3.0):
        times_o.append(t)
        steps_o.append(od.d.number_of_steps)
    assert steps == steps_o
    compare(times, times_o, "yield times", TOL)
    for q in ("stage", "xmomentum", "ymomentum"):
        compare(d.quantities[q].centroid_values, od.q(q, "centroid_values"), q, TOL)
        compare(d.quantities[q].vertex_values, od.q(q, "vertex_values"), q + ".vertex", TOL)
    compare(d.quantities["height"].edge_values, od.q("height", "edge_values"), "height.edge", TOL)


def test_merimbula_reference_step_count_on_gpu():
    """The reference's published Merimbula run takes 153 steps to t = 50 s
    (docs/profilling/CUDA/profile_on_GTX480_prepared_50_0.256246s)."""
    d = scenarios.merimbula()
    od = OracleDomain.from_domain(d)
    total = 0
    for t in d.evolve(yieldstep=50, finaltime=50):
        total += d.number_of_steps
    total_o = 0
    for t in od.evolve(50, 50):
        total_o += od.d.number_of_steps
    assert total == total_o == 153
    for q in ("stage", "xmomentum", "ymomentum"):
        compare(d.quantities[q].centroid_values, od.q(q, "centroid_values"), q, TOL)


def test_mixed_boundary_kinds():
    points, tris, boundary = rectangular_cross(20, 10, 200.0, 100.0)
    mesh = Mesh(points, tris, boundary=boundary)
    d = Domain(mesh=mesh)
    d.set_quantity("elevation", lambda x, y: -x / 100.0)
    d.set_quantity("stage", lambda x, y: np.where(x < 50, 1.0, 0.2))
    d.set_quantity("friction", 0.01, location="centroids")
    d.set_boundary({"left": Dirichlet_boundary([1.5, 0.2, 0.0]), "right": Transmissive_boundary(d),
                    "top": Reflective_boundary(d), "bottom": Reflective_boundary(d)})
    stats = {}
    _lockstep(d, 20, fused=True, stats=stats)


def test_first_order_fallback_and_smallsteps_error():
    """evolve_min_timestep above every CFL step forces ANUGA's fallback to first
    order after max_smallsteps, then the too-small-timestep error."""
    d = _dam(evolve_min_timestep=100.0, max_smallsteps=3)
    od = OracleDomain.from_domain(d)
    h = d._dev()
    h.call("sw_evolve_prepare", 1e9, 0.0, 0)
    od.evolve_prepare(1e9, None)
    orders = []
    err_dev = err_or = None
    for step in range(12):
        try:
            h.evolve_steps(1)
        except _lib.SwError as e:
            err_dev = (step, e.code)
        try:
            od.evolve_step()
        except Exception as e:  # OracleError
            err_or = (step, e.code)
        if err_dev or err_or:
            break
        orders.append(h.get_state().order)
        assert h.get_state().order == od.d._order_
        compare_state(d, od, TOL)
    assert 1 in orders
    assert err_dev is not None and err_or is not None
    assert err_dev == (err_or[0], _lib.SW_ERR_SMALLSTEPS)


def test_update_error_path_forced_from_host():
    """A semi-implicit update with 1 - dt*si <= 0 is reported as an error, as
    the C update returns -1 (quantity.c:803-805)."""
    d = _dam()
    h = d._dev()
    s = h.get_state()
    s.timestep = 1.0
    h.set_state(s)
    si = np.full(d.number_of_elements, 5.0)
    stage = device_array(d, "stage", "centroid_values")
    h.upload(_lib.SW_STAGE, _lib.SW_SEMI_IMPLICIT_UPDATE, si * stage)  # si/q = 5 -> 1 - 5 < 0
    with pytest.raises(_lib.SwError) as exc:
        h.call("sw_update_conserved_quantities")
    assert exc.value.code == _lib.SW_ERR_UPDATE


def test_uploaded_semi_implicit_terms_apply_once():
    """Semi-implicit updates set by the caller before a fused run: the first
    step's forcing adds Manning's term to them, the update applies them and
    resets them, as the C update does (quantity.c:769-820); the following
    steps match the oracle's."""
    d = _dam()
    for q, f in (("stage", -0.05), ("xmomentum", -0.2), ("ymomentum", 0.1)):
        c = d.quantities[q].centroid_values.copy()
        d.quantities[q].semi_implicit_update[:] = f * c
    stats = {}
    _lockstep(d, 4, fused=True, stats=stats)
    assert np.all(device_array(d, "stage", "semi_implicit_update") == 0.0)


def test_device_renumbering_is_a_relabelling():
    """The library's Hilbert renumbering of the triangles changes only where they
    live on the device: every array seen through the ABI, and the timestep
    sequence, are bitwise those of a handle that keeps the caller's numbering."""
    runs = {}
    for order in ("keep", "hilbert"):
        d = scenarios.delaunay_terrain(npoints=3000, order="random", extent=2000.0, mound_radius=200.0,
                                       seed=7, terrain_seed=11, locality_order=order)
        dts = []
        for _ in d.evolve(yieldstep=0.5, finaltime=2.0):
            dts.append(d.timestep)
        runs[order] = (dts, {q: (d.quantities[q].centroid_values.copy(), d.quantities[q].edge_values.copy(),
                                 d.quantities[q].vertex_values.copy())
                             for q in ("stage", "xmomentum", "ymomentum", "height", "xvelocity")},
                       d.max_speed.copy(), d.number_of_steps)
        d.close()
    a, b = runs["keep"], runs["hilbert"]
    assert a[3] == b[3] and a[3] > 10
    assert a[0] == b[0]
    for q in a[1]:
        for x, y in zip(a[1][q], b[1][q]):
            assert np.array_equal(x, y), q
    assert np.array_equal(a[2], b[2])


@pytest.mark.parametrize("method,limiter", [("euler", "optimised"), ("rk2", "optimised"), ("euler", "edge")])
def test_mixed_precision_error_against_oracle(method, limiter):
    """Mixed precision (FP32 flux and reconstruction arithmetic, FP32 edge
    state; FP64 accumulation, timestep and depth tests) on the 24 x 24 cell dam
    break, against the FP64 oracle at the same simulated time (t = 40 s, 28
    Euler steps).  Bounds: relative L2 error of the centroid values below 1e-6
    for stage and 1e-4 for the momenta (measured with Euler: 2.9e-8, 3.2e-7,
    3.8e-6); wet/dry masks differ in at most 1% of the cells (measured: none).
    The same bounds hold for rk2 steps against the oracle's rk2, and for the
    edge limiter (computed in FP64 inside the mixed loop) against the
    oracle's."""
    d = _dam(precision="mixed", **({"use_edge_limiter": 1} if limiter == "edge" else {}))
    d.set_timestepping_method(method)
    od = OracleDomain.from_domain(d)
    T = 40.0
    for _ in d.evolve(yieldstep=T, finaltime=T):
        pass
    od.evolve_prepare(T, T)
    while not od.evolve_step():
        pass
    assert abs(d.get_time() - T) < 1e-12 and abs(od.d.time - T) < 1e-12
    bed = od.q("elevation", "centroid_values")
    errs = {}
    for q in ("stage", "xmomentum", "ymomentum"):
        a = d.quantities[q].centroid_values
        b = od.q(q, "centroid_values").ravel()
        errs[q] = np.linalg.norm(a - b) / np.linalg.norm(b)
    w = d.quantities["stage"].centroid_values
    mism = np.count_nonzero(wet_mask(w, bed, d.minimum_allowed_height)
                            != wet_mask(od.q("stage", "centroid_values"), bed, d.minimum_allowed_height))
    print("mixed vs oracle at t=%g: steps gpu %d oracle %d, rel L2 %s, wet/dry mismatches %d of %d" % (
        T, d.number_of_steps, od.d.number_of_steps, errs, mism, d.number_of_elements))
    assert errs["stage"] < 1e-6
    assert errs["xmomentum"] < 1e-4 and errs["ymomentum"] < 1e-4
    assert mism <= 0.01 * d.number_of_elements


def test_config1_dam_break_to_10s_matches_oracle():
    """The CPU-runnable configuration in full: 100 x 100 cell rectangular_cross
    (40,000 triangles), bed x/1000 + 0.5 m Gaussian bump (sigma 100 m), stage
    2 m for x < 500 m else 0.5 m, Manning n 0.03, reflective walls, second
    order, evolve to t = 10 s with yieldstep 1 s.  Every yield matches the FP64
    oracle: same times and step counts, stage / momenta within 1e-12
    (normwise relative), identical wet/dry masks."""
    d = scenarios.dam_break_channel(100, 100)
    od = OracleDomain.from_domain(d)
    bed = od.q("elevation", "centroid_values")
    gen_o = od.evolve(1.0, 10.0)
    yields = 0
    total = 0
    for t in d.evolve(yieldstep=1.0, finaltime=10.0):
        t_o = next(gen_o)
        yields += 1
        total += d.number_of_steps
        compare([t], [t_o], "cfg1.time", TOL)
        assert d.number_of_steps == od.d.number_of_steps
        for q in ("stage", "xmomentum", "ymomentum"):
            compare(d.quantities[q].centroid_values, od.q(q, "centroid_values"), "cfg1.%s" % q, TOL)
        assert np.array_equal(wet_mask(d.quantities["stage"].centroid_values, bed, d.minimum_allowed_height),
                              wet_mask(od.q("stage", "centroid_values"), bed, d.minimum_allowed_height))
    assert next(gen_o, None) is None
    assert abs(d.get_time() - 10.0) < 1e-12 and yields == 11
    print("config 1: %d steps to t = 10 s over %d yields" % (total, yields))


@pytest.mark.parametrize("method", ["euler", "rk2", "rk3"])
def test_ensemble_instances_match_separate_runs_and_oracle(method):
    """Three members of different size, Manning n, dam height and inflow in one
    ensemble handle: each instance's timestep sequence and state equal those
    of the same domain evolved alone, bit for bit, and its step counts equal
    the FP64 oracle's, with the state within 1e-12 per step taken (the last
    bit of Manning's pow() may differ each step and is not re-synchronised
    over the run).  The members reach their yields at different steps, so the
    Runge-Kutta runs also check that a halted member skips every stage."""
    specs = [dict(triangles=900, seed=11, manning=0.012, dam_height=1.5, inflow=0.2),
             dict(triangles=2500, seed=12, manning=0.05, dam_height=3.7, inflow=0.9),
             dict(triangles=1400, seed=13, manning=0.03, dam_height=2.2, inflow=0.0)]
    members = [scenarios.ensemble_member(extent=300.0, **sp) for sp in specs]
    alone = [scenarios.ensemble_member(extent=300.0, **sp) for sp in specs]
    for d in members + alone:
        d.set_timestepping_method(method)
    oracles = [OracleDomain.from_domain(d) for d in alone]
    ens = Ensemble(members)
    ens_steps = []
    for t in ens.evolve(yieldstep=2.0, finaltime=6.0):
        ens_steps.append([d.number_of_steps for d in members])
    assert len(ens_steps) == 4
    for i, (d, a, od) in enumerate(zip(members, alone, oracles)):
        steps = [s[i] for s in ens_steps]
        steps_alone = []
        for t in a.evolve(yieldstep=2.0, finaltime=6.0):
            steps_alone.append(a.number_of_steps)
        steps_o = []
        for t in od.evolve(2.0, 6.0):
            steps_o.append(od.d.number_of_steps)
        assert steps == steps_alone == steps_o, (i, steps, steps_alone, steps_o)
        print("member %d: %d triangles, steps per yield %s" % (i, d.number_of_elements, steps))
        assert abs(d.get_time() - 6.0) < 1e-12
        for q in ("stage", "xmomentum", "ymomentum"):
            for loc in ("centroid_values", "vertex_values", "edge_values"):
                x = getattr(d.quantities[q], loc)
                y = getattr(a.quantities[q], loc)
                assert np.array_equal(x, y), (i, q, loc)
            compare(d.quantities[q].centroid_values, od.q(q, "centroid_values"), "ens%d.%s" % (i, q),
                    TOL * sum(steps))
        assert np.array_equal(d.max_speed, a.max_speed)
        a.close()
    ens.close()


def test_autotuned_launch_shapes_do_not_change_results():
    """sw_autotune runs real steps with workgroups of 128, 256 and 512 lanes
    and keeps the fastest per kernel; the state after it and after further
    steps is bitwise that of a run that was never tuned."""
    a = _delaunay("random")
    b = _delaunay("random")
    ha, hb = a._dev(), b._dev()
    for h in (ha, hb):
        h.call("sw_evolve_prepare", 1e9, 0.0, 0)
    chosen = ha.autotune(2)  # 3 candidates x 2 steps
    assert all(c in (128, 256, 512) for c in chosen)
    hb.evolve_steps(6)
    ha.evolve_steps(10)
    hb.evolve_steps(10)
    a._host_stale = b._host_stale = True
    for q in ("stage", "xmomentum", "ymomentum"):
        for loc in ("centroid_values", "vertex_values", "edge_values"):
            assert np.array_equal(getattr(a.quantities[q], loc), getattr(b.quantities[q], loc)), (q, loc)
    assert a._handle.get_state().time == b._handle.get_state().time


NON_DEFAULT = {
    "sloped_mannings": dict(use_sloped_mannings=1),
    "speed_cap": dict(maximum_allowed_speed=2.0),
    "alpha_balance": dict(tight_slope_limiters=0),
    "momentum_balance": dict(use_centroid_velocities=0),
    "momentum_extrapolation": dict(extrapolate_velocity_second_order=0),
    "no_dry_skip": dict(optimise_dry_cells=0),
    "beta_dry": dict(beta_w_dry=0.2, beta_uh_dry=0.1, beta_vh_dry=0.3, beta_w=1.1, beta_uh=0.8, beta_vh=0.95),
}


@pytest.mark.parametrize("fused", [True, False], ids=["fused", "unfused"])
@pytest.mark.parametrize("case", sorted(NON_DEFAULT))
def test_non_default_parameter_branches_match_oracle(case, fused):
    """The HIP branches behind ANUGA's non-default parameters, in lockstep
    with the oracle for 16 steps on the 3,000-point wet/dry Delaunay workload
    (the bar of the other lockstep tests: 1e-12 per step on centroid, vertex
    and edge values, timesteps, wet/dry masks): sloped Manning friction
    (shallow_water_ext.c:1272-1317), the speed cap of protect (:1638-1705),
    the alpha_balance balance and the momentum blend without centroid
    velocities (:1462-1636), momentum extrapolated directly, no dry-cell skip
    (:2717-3324) and beta_*_dry different from beta_* (:2966-2974).  The
    speed-cap case starts from momenta of 5 m/s where the water is shallow,
    so the cap acts."""
    d = _delaunay("random")
    for k, v in NON_DEFAULT[case].items():
        setattr(d, k, v)
    if case == "speed_cap":
        w, z = d.quantities["stage"].centroid_values, d.quantities["elevation"].centroid_values
        h = w - z
        shallow = (h > 0) & (h < 0.1)
        uh = d.quantities["xmomentum"].centroid_values.copy()
        uh[shallow] = 5.0 * h[shallow]
        d.set_quantity("xmomentum", uh, location="centroids")
        assert shallow.sum() > 20
    stats = {}
    od, dts = _lockstep(d, 16, fused=fused, stats=stats)
    assert len(dts) == 16
    assert all(getattr(od.d, k) == v for k, v in NON_DEFAULT[case].items() if hasattr(od.d, k))
    compare(d._handle.max_speed(), od.max_speed, "max_speed", TOL)
    print("%s (%s): bit-mismatching fields %s" % (case, "fused" if fused else "unfused",
                                                  {k: v[0] for k, v in stats.items() if v[0]}))


@pytest.mark.parametrize("method,limiter", [("euler", "optimised"), ("rk2", "optimised"), ("rk3", "optimised"),
                                            ("euler", "vertex"), ("rk2", "edge")])
def test_vertex_record_step_equals_edge_record_step(method, limiter):
    """The vertex-record step (flux from the vertex records; timestep, update
    and friction; extrapolation and balance), with Euler steps or the stages
    of rk2 / rk3 and with the optimised or a per-quantity limiter (betas that
    differ per quantity), and the edge-record step give the same bits:
    timesteps, yields and every array of the operator API, over yields that
    stop a captured 16-step graph part way."""
    kw = {"vertex": {"optimised_gradient_limiter": 0}, "edge": {"use_edge_limiter": 1}}.get(limiter, {})
    runs = {}
    for path in (_lib.SW_STEP_AUTO, _lib.SW_STEP_EDGE):
        d = _delaunay("random", **kw)
        if limiter != "optimised":
            d.quantities["xmomentum"].beta = 0.7
            d.quantities["ymomentum"].beta = 1.3
        assert d.limiter_path() == limiter
        d.set_timestepping_method(method)
        d._dev().set_step_path(path)
        assert d._handle.step_path() == (_lib.SW_STEP_VERTEX if path == _lib.SW_STEP_AUTO else _lib.SW_STEP_EDGE)
        dts, steps = [], []
        for _ in d.evolve(yieldstep=0.37, finaltime=3.0):
            dts.append(d._handle.get_state().timestep)
            steps.append(d.number_of_steps)
        arrays = {(q, loc): getattr(d.quantities[q], loc).copy()
                  for q in ("stage", "xmomentum", "ymomentum", "height", "xvelocity", "yvelocity")
                  for loc in ("centroid_values", "vertex_values", "edge_values", "boundary_values")}
        for q in ("stage", "xmomentum", "ymomentum"):
            arrays[(q, "explicit_update")] = device_array(d, q, "explicit_update")
        runs[path] = (dts, steps, arrays, d.max_speed.copy(), d.get_time())
        d.close()
    a, b = runs[_lib.SW_STEP_AUTO], runs[_lib.SW_STEP_EDGE]
    assert a[0] == b[0] and a[1] == b[1] and sum(a[1]) > 40 and a[4] == b[4]
    for key in a[2]:
        assert np.array_equal(a[2][key], b[2][key]), key
    assert np.array_equal(a[3], b[3])


def _api_arrays(d):
    out = {}
    for q in ("stage", "xmomentum", "ymomentum", "height", "xvelocity", "yvelocity"):
        for loc in ("centroid_values", "vertex_values", "edge_values", "boundary_values"):
            out[(q, loc)] = device_array(d, q, loc)
    for q in ("stage", "xmomentum", "ymomentum"):
        for loc in ("explicit_update", "semi_implicit_update"):
            out[(q, loc)] = device_array(d, q, loc)
    out["max_speed"] = d._handle.max_speed()
    return out


def test_chained_calls_continue_from_the_step_records():
    """Consecutive vertex-record evolve calls continue from the step's records
    at whatever parity the previous call ended on, and the operator API's
    arrays are brought up to date only when read: 40 steps as one call, as
    calls of 1, 2, 17, 3, 16 and 1 steps with nothing read in between, and as
    the same calls with every array read (and the stage re-uploaded
    unchanged, which rebuilds the records) after each, give the same bits."""
    chunks = [1, 2, 17, 3, 16, 1]
    runs = []
    for mode in ("one", "chained", "read"):
        d = _delaunay("morton")
        h = d._dev()
        assert h.step_path() == _lib.SW_STEP_VERTEX
        h.call("sw_evolve_prepare", 1e9, 0.0, 0)
        states = []
        for n in ([sum(chunks)] if mode == "one" else chunks):
            done, ev = h.evolve_steps(n)
            assert (done, ev) == (n, _lib.SW_EVENT_NONE)
            s = h.get_state()
            states.append((s.time, s.timestep, s.number_of_steps))
            if mode == "read":
                _api_arrays(d)
                w = device_array(d, "stage", "centroid_values")
                h.upload(_lib.SW_STAGE, _lib.SW_CENTROID, w)
        runs.append((states[-1], _api_arrays(d)))
        d.close()
    (s0, a0) = runs[0]
    for s, a in runs[1:]:
        assert s == s0
        for key in a0:
            assert np.array_equal(a[key], a0[key]), key


def test_upload_between_chained_calls_matches_oracle():
    """A caller's upload between two evolve calls reaches the next step (the
    step's records are rebuilt from the operator API's arrays), as in the
    oracle given the same values."""
    d = _delaunay("morton")
    od = OracleDomain.from_domain(d)
    h = d._dev()
    h.call("sw_evolve_prepare", 1e9, 0.0, 0)
    od.evolve_prepare(1e9, None)
    assert h.evolve_steps(5)[0] == 5
    for _ in range(5):
        od.evolve_step()
    w = device_array(d, "stage", "centroid_values")
    bed = od.q("elevation", "centroid_values")
    w = np.where(w - bed > 0.05, w + 0.01, w)
    h.upload(_lib.SW_STAGE, _lib.SW_CENTROID, w)
    od.q("stage", "centroid_values")[:] = w
    stats = {}
    for n in (2, 3):
        assert h.evolve_steps(n)[0] == n
        for _ in range(n):
            od.evolve_step()
    s = h.get_state()
    compare([s.time], [od.d.time], "time", TOL, stats)
    compare_state(d, od, TOL, stats=stats, tag="after upload.")


def test_yield_reads_upload_nothing_and_writes_reach_the_next_step():
    """Domain.evolve at yields: reading quantities downloads only what is read
    and uploads nothing (the device keeps its step records); an in-place
    change to a quantity's array is uploaded, that array only, and the run
    continues from it as the oracle does given the same values."""
    d = _delaunay("morton")
    od = OracleDomain.from_domain(d)
    h = d._dev()
    uploads, downloads = [], []
    up, down = h.upload, h.download
    h.upload = lambda q, loc, arr: (uploads.append((q, loc)), up(q, loc, arr))[1]
    h.download = lambda q, loc, out: (downloads.append((q, loc)), down(q, loc, out))[1]
    bed = od.q("elevation", "centroid_values")
    gen, ogen = d.evolve(yieldstep=0.01, finaltime=0.05), od.evolve(0.01, 0.05)
    stats = {}
    for y, (t, t_o) in enumerate(zip(gen, ogen)):
        compare([t], [t_o], "yield time", TOL, stats)
        n_up, n_down = len(uploads), len(downloads)
        w = d.quantities["stage"].centroid_values
        assert downloads[n_down:] == [(_lib.SW_STAGE, _lib.SW_CENTROID)]
        compare(w, od.q("stage", "centroid_values"), "yield %d stage" % y, TOL, stats)
        uh = d.quantities["xmomentum"].centroid_values
        compare(uh, od.q("xmomentum", "centroid_values"), "yield %d xmomentum" % y, TOL, stats)
        if y == 2:
            # a 1 cm rise of the wet cells, on both sides
            w[...] = np.where(w - bed > 0.05, w + 0.01, w)
            od.q("stage", "centroid_values")[:] = w
        n_dev = len(uploads)
        d._dev()
        assert uploads[n_dev:] == ([(_lib.SW_STAGE, _lib.SW_CENTROID)] if y == 2 else [])
        assert len(uploads) == n_up + (1 if y == 2 else 0)
    assert y == 5
    compare_state(d, od, TOL, stats=stats, tag="after the yields.")


def test_vertex_step_with_varying_friction_and_centroid_bed_matches_oracle():
    """The vertex-record step reads Manning n and the mean bed per triangle
    when they are not uniform / not the centroid bed (its update skips those
    loads only where k_step_begin found them so): spatially varying n and a
    centroid bed raised by 1 mm on every third triangle, in lockstep with the
    oracle for 16 steps."""
    d = _delaunay("morton")
    x = d.mesh.centroid_coordinates[:, 0]
    d.set_quantity("friction", 0.015 + 0.02 * (0.5 + 0.5 * np.sin(x / 150.0)), location="centroids")
    # only the centroid values (set_quantity would extrapolate them to the
    # vertices, making the bed discontinuous)
    d.quantities["elevation"].centroid_values[::3] += 1e-3
    d._all_dirty = True
    assert d._dev().step_path() == _lib.SW_STEP_VERTEX
    stats = {}
    od, dts = _lockstep(d, 16, fused=True, stats=stats)
    assert len(dts) == 16


def test_two_kernel_step_needs_recomputable_geometry():
    """A mesh whose geometry arrays are not what the vertex coordinates give
    (here a radius changed in its last bit) keeps the edge-record step, and
    so does the mixed-precision loop; rk2 / rk3 run on the vertex-record
    step."""
    d = _delaunay("morton")
    d.mesh.radii = d.mesh.radii.copy()
    d.mesh.radii[10] = np.nextafter(d.mesh.radii[10], 0.0)
    assert d._dev().step_path() == _lib.SW_STEP_EDGE
    d2 = _delaunay("morton")
    assert d2._dev().step_path() == _lib.SW_STEP_VERTEX
    d2.set_timestepping_method("rk2")
    assert d2._handle.step_path() == _lib.SW_STEP_VERTEX
    d3 = scenarios.delaunay_terrain(npoints=3000, order="morton", extent=2000.0, mound_radius=200.0, seed=7,
                                    terrain_seed=11, precision="mixed")
    assert d3._dev().step_path() == _lib.SW_STEP_EDGE


def test_ensemble_members_must_share_their_start_time():
    specs = [dict(triangles=900, seed=11, manning=0.012, dam_height=1.5, inflow=0.2),
             dict(triangles=900, seed=12, manning=0.05, dam_height=3.7, inflow=0.9)]
    members = [scenarios.ensemble_member(extent=300.0, **sp) for sp in specs]
    members[1].set_starttime(2.0)
    ens = Ensemble(members)
    with pytest.raises(ValueError):
        ens.prepare(1.0, 4.0)
    ens.close()
