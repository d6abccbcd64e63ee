"""Pins of the oracle's parameter branches and special cases (CPU only).

Each test fixes one branch of the reference's FP64 C by a closed form, a
special case that reduces to a hand-computed value, or a brute-force
recomputation, none of which reuses the oracle's own code:

  * _compute_speed's velocity desingularisation for eps <= h < 10 H0
    (shallow_water_ext.c:214-240) against the ANUGA manual's
    u = uh h / (h^2 + H0^2) inside an independently written KNP flux;
  * max_speed[k] and the CFL timestep of _compute_fluxes_central_structure
    (:4519-4696) by brute force over edges, with the dry-edge skip and the
    tri_full_flag gate;
  * the hfactor / beta_dry blend of the limiter (:2966-2974) and the dry-cell
    skip that leaves vertex values stale (:2976-2985), then rescaled by the
    vertex depth (:3292-3315);
  * the auxiliary triangle's area2 <= 0 fallback (:2941-2958) and the
    two-boundary-edge reconstruction (:3130-3215), whose limiter always gives
    phi = 0 (see test_two_boundary_triangles_reconstruct_constant);
  * sloped Manning friction on a planar bed (:1272-1317);
  * the balance's 1e-6 centroid-depth cutoff, the clamp of alpha at 0 and the
    momentum blend without centroid velocities (:1462-1636);
  * the vertex -> edge pairing (quantity.c:687-707) by a linear field;
  * the stage-gradient gravity term (:1942-2037) on a tilted plane;
  * update_timestep's CFL scaling, clips and first-order fallback rule
    (evolve.c:115-160).
"""
import math

import numpy as np
import pytest

from anuga_cuda_amd.domain import Domain, Reflective_boundary
from anuga_cuda_amd.mesh import Mesh, delaunay_jittered, rectangular_cross
from oracle import oracle as O

G = 9.8
EPS = 1e-12
H0 = 1e-3


# ------------------------------------------------------------ helpers
def _domain(m=12, n=10, lx=120.0, ly=100.0, bed=None, **params):
    points, tris, boundary = rectangular_cross(m, n, lx, ly)
    d = Domain(mesh=Mesh(points, tris, boundary=boundary), **params)
    d.set_quantity("elevation", bed if bed is not None else (lambda x, y: 0.01 * x - 0.02 * y))
    d.set_quantity("stage", 0.0)
    d.set_boundary({t: Reflective_boundary(d) for t in d.mesh.get_tag_names()})
    return d


def _set_centroids(od, h, uh=None, vh=None):
    z = od.q("elevation", "centroid_values")
    od.q("stage", "centroid_values")[:] = z + h
    if uh is not None:
        od.q("xmomentum", "centroid_values")[:] = uh
    if vh is not None:
        od.q("ymomentum", "centroid_values")[:] = vh


def _jumps(od, q):
    """Vertex values minus the centroid value (after an extrapolation)."""
    return od.q(q, "vertex_values") - od.q(q, "centroid_values")[:, None]


# ----------------------------------------------------- flux, shallow depths
def _knp_manual(ql, qr, zl, zr, n1, n2, g=G, eps=EPS, H0=H0):
    """Central-upwind flux (Kurganov, Noelle & Petrova 2001, eq. 3.15) in the
    face frame, with ANUGA's desingularised velocities (ANUGA user manual:
    u = uh h / (h^2 + H0^2) below 10 H0, depth and velocity 0 below eps).
    Vectorised over the leading axis; returns (flux in x/y frame, max speed)."""
    ql, qr = np.asarray(ql, float), np.asarray(qr, float)
    z = 0.5 * (zl + zr)

    def side(q):
        w, un, ut = q[:, 0], n1 * q[:, 1] + n2 * q[:, 2], -n2 * q[:, 1] + n1 * q[:, 2]
        h = w - z
        dry = h < eps
        shallow = (h < 10 * H0) & ~dry
        hs = np.where(dry, 0.0, h)
        safe = np.where(dry, 1.0, h)
        u = np.where(shallow, un * safe / (safe * safe + H0 * H0), un / safe)
        v = np.where(shallow, ut * safe / (safe * safe + H0 * H0), ut / safe)
        u = np.where(dry, 0.0, u)
        v = np.where(dry, 0.0, v)
        flux = np.stack([u * hs, u * u * hs + 0.5 * g * hs * hs, u * v * hs], axis=1)
        return np.stack([w, un, ut], axis=1), hs, u, flux

    qL, hL, uL, FL = side(ql)
    qR, hR, uR, FR = side(qr)
    cL, cR = np.sqrt(g * hL), np.sqrt(g * hR)
    ap = np.maximum(np.maximum(uL + cL, uR + cR), 0.0)
    am = np.minimum(np.minimum(uL - cL, uR - cR), 0.0)
    den = ap - am
    ok = den >= eps
    sd = np.where(ok, den, 1.0)
    F = (ap[:, None] * FL - am[:, None] * FR + (ap * am)[:, None] * (qR - qL)) / sd[:, None]
    F = np.where(ok[:, None], F, 0.0)
    out = np.stack([F[:, 0], n1 * F[:, 1] - n2 * F[:, 2], n2 * F[:, 1] + n1 * F[:, 2]], axis=1)
    return out, np.where(ok, np.maximum(np.abs(ap), np.abs(am)), 0.0)


def test_flux_velocity_desingularisation_at_shallow_depths():
    """Depths across every branch of _compute_speed: negative, below eps,
    eps <= h < 10 H0 (the desingularised velocity) and deeper; momenta of
    either sign.  The oracle's flux equals the manual's formula to 1e-12
    relative (the two differ only in rounding: uh / (h + H0^2/h) against
    uh h / (h^2 + H0^2))."""
    rng = np.random.default_rng(21)
    n = 4000
    depths = np.concatenate([
        -rng.uniform(0, 1e-3, n // 8), rng.uniform(0, 1e-12, n // 8),
        10 ** rng.uniform(-11.9, -2.0, n // 2), rng.uniform(0.01, 3.0, n // 4)])
    rng.shuffle(depths)
    hl, hr = depths, rng.permutation(depths)
    z = rng.uniform(-2, 2, n)
    ql = np.stack([z + hl, rng.normal(size=n) * 0.01, rng.normal(size=n) * 0.01], 1)
    qr = np.stack([z + hr, rng.normal(size=n) * 0.01, rng.normal(size=n) * 0.01], 1)
    ang = rng.uniform(0, 2 * np.pi, n)
    shallow = 0
    for i in range(n):
        n1, n2 = math.cos(ang[i]), math.sin(ang[i])
        ef, ms = O.flux_function_central(ql[i], qr[i], z[i], z[i], n1, n2)
        ex, ems = _knp_manual(ql[i:i + 1], qr[i:i + 1], z[i], z[i], n1, n2)
        scale = max(np.max(np.abs(ex)), 1e-300)
        assert np.max(np.abs(ef - ex[0])) <= 1e-12 * scale, (i, hl[i], hr[i], ef, ex[0])
        assert ms == pytest.approx(ems[0], rel=1e-12, abs=1e-300)
        shallow += (EPS <= hl[i] < 10 * H0) or (EPS <= hr[i] < 10 * H0)
    assert shallow > n // 3


def test_flux_of_a_shallow_layer_differs_from_the_raw_velocity():
    """A 2 mm layer moving at 0.5 m/s: the desingularised velocity is
    u h^2 / (h^2 + H0^2) = 0.4 m/s, so the mass flux of a uniform state is
    0.4 * h, not 0.5 * h."""
    h, u = 2e-3, 0.5
    q = np.array([h, u * h, 0.0])
    ef, ms = O.flux_function_central(q, q, 0.0, 0.0, 1.0, 0.0)
    ud = u * h * h / (h * h + H0 * H0)
    assert ef[0] == pytest.approx(ud * h, rel=1e-14)
    assert ms == pytest.approx(ud + math.sqrt(G * h), rel=1e-14)


# ---------------------------------------------- max_speed and CFL timestep
def _wet_dry_delaunay(seed=4):
    points, tris = delaunay_jittered(500, extent=(400.0, 300.0), seed=seed)
    d = Domain(mesh=Mesh(points, tris))
    d.set_quantity("elevation", lambda x, y: 0.004 * x + 0.3 * np.sin(y / 40.0))
    d.set_quantity("stage", 0.0)
    d.set_boundary({t: Reflective_boundary(d) for t in d.mesh.get_tag_names()})
    return d


def test_max_speed_and_cfl_by_brute_force_over_edges():
    """_compute_fluxes_central_structure (shallow_water_ext.c:4519-4696):
    max_speed[k] is the largest wave speed over the edges that triangle k
    computes (those shared with a higher-numbered neighbour, and its boundary
    edges), an edge skipped as dry (both sides |w - z| < eps) counting 0; the
    CFL timestep is the minimum of radius / speed over computed edges whose
    computing triangle is full, on both of its sides.  Checked edge by edge on
    a wet/dry mesh with some triangles not full."""
    d = _wet_dry_delaunay()
    m = d.mesh
    N = m.number_of_elements
    od = O.OracleDomain.from_domain(d)
    rng = np.random.default_rng(8)
    h = np.where(rng.random(N) < 0.3, 0.0, rng.uniform(0.05, 1.0, N))
    _set_centroids(od, h, rng.normal(size=N) * h, rng.normal(size=N) * h)
    od.distribute_to_vertices_and_edges()
    od.update_boundary()
    full = (rng.random(N) < 0.8).astype(np.int64)
    od.arrays["tri_full_flag"][:] = full
    ts = od.compute_fluxes()
    E = {q: od.q(q, "edge_values") for q in ("stage", "xmomentum", "ymomentum")}
    zb = od.q("elevation", "edge_values")
    B = {q: od.q(q, "boundary_values") for q in ("stage", "xmomentum", "ymomentum")}
    ms_exp = np.zeros(N)
    dt = d.evolve_max_timestep
    dry_edges = mixed_edges = 0
    for k in range(N):
        for i in range(3):
            n = m.neighbours[k, i]
            if 0 <= n < k:
                continue  # computed by the lower-numbered neighbour
            ql = np.array([E[q][k, i] for q in E])
            if n < 0:
                qr = np.array([B[q][-n - 1] for q in B])
                zr = zb[k, i]
            else:
                j = m.neighbour_edges[k, i]
                qr = np.array([E[q][n, j] for q in E])
                zr = zb[n, j]
            dry_l, dry_r = abs(ql[0] - zb[k, i]) < EPS, abs(qr[0] - zr) < EPS
            if dry_l and dry_r:
                dry_edges += 1
                continue
            mixed_edges += dry_l != dry_r
            _, s = O.flux_function_central(ql, qr, zb[k, i], zr, *m.normals[k, 2 * i:2 * i + 2])
            ms_exp[k] = max(ms_exp[k], s)
            if full[k] and s > EPS:
                dt = min(dt, m.radii[k] / s)
                if n >= 0:
                    dt = min(dt, m.radii[n] / s)
    assert dry_edges > 20 and mixed_edges > 20
    assert np.array_equal(od.max_speed, ms_exp)
    assert ts == dt
    # a triangle with several computed edges of different speeds: the maximum, not the last
    assert np.count_nonzero(ms_exp > 0) > N // 2


# ---------------------------------------------------- hfactor and beta_dry
def _limiter_state(seed, **params):
    """Interior triangles of a rectangular_cross mesh, depths from dry to deep
    (many near the 1 mm hfactor threshold), random momenta."""
    d = _domain(**params)
    od = O.OracleDomain.from_domain(d)
    N = d.number_of_elements
    rng = np.random.default_rng(seed)
    h = rng.choice([0.0008, 0.001, 0.0011, 0.0015, 0.002, 0.004, 0.02, 0.3], size=N)
    h = h * rng.uniform(1.0, 1.02, size=N)
    h[rng.random(N) < 0.05] = 0.001  # exactly at the threshold
    uh = rng.normal(size=N) * h
    vh = rng.normal(size=N) * h
    _set_centroids(od, h, uh, vh)
    return d, od


def _hmin(d, od):
    sn = d.mesh.surrogate_neighbours
    hc = od.q("stage", "centroid_values") - od.q("elevation", "centroid_values")
    return np.minimum(hc, hc[sn].min(axis=1)), hc


def test_beta_dry_gives_first_order_at_or_below_one_millimetre():
    """beta_*_dry = 0, beta_* = 1: where the shallowest of the triangle and
    its three neighbours is at most 1 mm deep, hfactor = 0 and every quantity
    is reconstructed constant (vertex = centroid, bitwise); at 1.1 mm and
    deeper the reconstruction is second order again."""
    dry = dict(beta_w_dry=0.0, beta_uh_dry=0.0, beta_vh_dry=0.0, beta_w=1.0, beta_uh=1.0, beta_vh=1.0,
               extrapolate_velocity_second_order=0)
    d, od = _limiter_state(31, **dry)
    od.extrapolate_second_order_sw()
    hmin, hc = _hmin(d, od)
    inner = d.mesh.number_of_boundaries == 0
    flat = inner & (hmin <= 0.001)
    steep = inner & (hmin >= 0.0011)
    assert flat.sum() > 30 and steep.sum() > 30
    # the centroid itself is the shallowest somewhere (hmin includes hc)
    assert np.count_nonzero(flat & (hc <= 0.001) & (hc < hmin + 1e-15)) > 5
    for q in ("stage", "xmomentum", "ymomentum"):
        assert np.all(_jumps(od, q)[flat] == 0.0), q
        # (a local extremum among the neighbours is flattened by the limiter itself)
        assert np.count_nonzero(np.any(_jumps(od, q)[steep] != 0.0, axis=1)) > 0.3 * steep.sum(), q


def test_beta_blend_is_linear_in_the_depth_factor():
    """beta = beta_dry + (beta - beta_dry) hfactor, hfactor = (hmin - 1 mm) /
    (hmin + 4 mm) for hmin > 1 mm (shallow_water_ext.c:2966-2974), each
    quantity with its own pair.  With small betas the limiter is active
    (phi = r beta < 1), so the limited jumps scale with beta: against a run
    with beta_dry = beta = 0.01 the jumps of stage (beta_dry 0), xmomentum
    (0.003) and ymomentum (0.006) shrink by exactly the blended factor."""
    base = dict(beta_w=0.01, beta_uh=0.01, beta_vh=0.01, extrapolate_velocity_second_order=0)
    d, ref = _limiter_state(32, beta_w_dry=0.01, beta_uh_dry=0.01, beta_vh_dry=0.01, **base)
    dry = {"stage": 0.0, "xmomentum": 0.003, "ymomentum": 0.006}
    d2, od = _limiter_state(32, beta_w_dry=dry["stage"], beta_uh_dry=dry["xmomentum"],
                            beta_vh_dry=dry["ymomentum"], **base)
    ref.extrapolate_second_order_sw()
    od.extrapolate_second_order_sw()
    hmin, hc = _hmin(d, ref)
    hf = np.where(hmin > 0.001, (hmin - 0.001) / (hmin + 0.004), 0.0)
    inner = d.mesh.number_of_boundaries == 0
    checked = 0
    for q, bd in dry.items():
        jr, j = _jumps(ref, q), _jumps(od, q)
        qmax = np.abs(ref.q(q, "centroid_values")).max()
        big = inner[:, None] & (np.abs(jr) > 1e-9 * qmax)
        factor = (bd + (0.01 - bd) * hf) / 0.01
        # a jump is read back as (q_c + dq) - q_c: a few ulps of q_c
        np.testing.assert_allclose(j[big], (factor[:, None] * jr)[big], rtol=1e-9, atol=1e-15 * qmax)
        # threshold cases present: hmin exactly 1 mm and just above it
        checked += big.sum()
    assert checked > 500
    assert np.count_nonzero(inner & (np.abs(hmin - 0.001) < 1e-12)) > 3  # at the threshold


def test_dry_cells_keep_stale_vertex_values_rescaled_by_vertex_depth():
    """optimise_dry_cells: a triangle whose three neighbours are all shallower
    than eps is skipped (hmax excludes the triangle itself), so its vertex
    stage keeps its previous values bitwise; the velocity extrapolation then
    still multiplies its previous vertex momenta by max(w_v - z_v, 0)
    (shallow_water_ext.c:2976-2985, 3292-3315).  Other triangles are
    recomputed, wet or not."""
    d = _domain(14, 12)
    od = O.OracleDomain.from_domain(d)
    N = d.number_of_elements
    rng = np.random.default_rng(33)
    h = np.where(rng.random(N) < 0.7, 0.0, rng.uniform(0.01, 0.5, N))
    h[rng.random(N) < 0.1] = 0.3  # wet triangles among dry neighbours
    _set_centroids(od, h, rng.normal(size=N) * h, rng.normal(size=N) * h)
    zv = od.q("elevation", "vertex_values")
    wv0 = zv + rng.uniform(0.1, 0.4, size=(N, 3))
    uv0, vv0 = rng.normal(size=(N, 3)), rng.normal(size=(N, 3))
    od.q("stage", "vertex_values")[:] = wv0
    od.q("xmomentum", "vertex_values")[:] = uv0
    od.q("ymomentum", "vertex_values")[:] = vv0
    od.extrapolate_second_order_sw()
    sn = d.mesh.surrogate_neighbours
    hmax = h[sn].max(axis=1)
    inner = d.mesh.number_of_boundaries == 0
    skipped = inner & (hmax < EPS)
    redone = inner & (hmax >= EPS)
    assert np.count_nonzero(skipped & (h > 0)) > 3 and np.count_nonzero(skipped & (h == 0)) > 10
    assert np.count_nonzero(redone & (h == 0)) > 10
    wv, uv, vv = (od.q(q, "vertex_values") for q in ("stage", "xmomentum", "ymomentum"))
    assert np.array_equal(wv[skipped], wv0[skipped])
    dv = np.maximum(wv0 - zv, 0.0)
    assert np.array_equal(uv[skipped], (uv0 * dv)[skipped])
    assert np.array_equal(vv[skipped], (vv0 * dv)[skipped])
    assert not np.any(wv[redone] == wv0[redone])
    # with optimise_dry_cells off nothing is skipped
    d2 = _domain(14, 12, optimise_dry_cells=0)
    od2 = O.OracleDomain.from_domain(d2)
    _set_centroids(od2, h, od.q("xmomentum", "centroid_values"), od.q("ymomentum", "centroid_values"))
    od2.q("stage", "vertex_values")[:] = wv0
    od2.extrapolate_second_order_sw()
    assert not np.any(od2.q("stage", "vertex_values")[skipped] == wv0[skipped])


def test_clockwise_auxiliary_triangle_falls_back_to_constant():
    """The auxiliary triangle of the surrogate neighbours' centroids with
    area2 <= 0 (listed clockwise) gives the constant reconstruction
    (shallow_water_ext.c:2941-2958); listed counter-clockwise the same
    neighbours give a second-order one."""
    d = _domain(10, 10, extrapolate_velocity_second_order=0)
    od = O.OracleDomain.from_domain(d)
    N = d.number_of_elements
    rng = np.random.default_rng(34)
    cx, cy = d.mesh.centroid_coordinates.T
    # smooth fields (few local extrema, which the limiter would flatten anyway)
    _set_centroids(od, 0.5 + 0.004 * cx + 0.001 * rng.random(N), 0.01 * cy + 0.01 * rng.random(N),
                   0.02 * cx - 0.01 * cy + 0.01 * rng.random(N))
    inner = np.nonzero(d.mesh.number_of_boundaries == 0)[0]
    swap = inner[::3]
    sn = od.arrays["surrogate_neighbours"].reshape(N, 3)
    sn[swap, 1], sn[swap, 2] = sn[swap, 2].copy(), sn[swap, 1].copy()
    od.extrapolate_second_order_sw()
    keep = np.setdiff1d(inner, swap)
    for q in ("stage", "xmomentum", "ymomentum"):
        j = _jumps(od, q)
        assert np.all(j[swap] == 0.0), q
        assert np.count_nonzero(np.any(j[keep] != 0.0, axis=1)) > 0.9 * len(keep), q


def _two_triangle_square(**params):
    points = [[0.0, 0.0], [10.0, 0.0], [10.0, 7.0], [0.0, 7.0]]
    d = Domain(points, [[0, 1, 2], [0, 2, 3]], **params)
    d.set_quantity("elevation", lambda x, y: 0.05 * x)
    d.set_quantity("stage", 0.0)
    d.set_boundary({"exterior": Reflective_boundary(d)})
    return d


@pytest.mark.parametrize("beta", [0.3, 0.9, 2.0])
def test_two_boundary_triangles_reconstruct_constant(beta):
    """A triangle with two boundary edges takes the gradient toward its one
    interior neighbour (shallow_water_ext.c:3130-3215): jumps dq1 t_i with
    t_i the projections of the vertex offsets on the centroid-to-centroid
    direction and bounds [min(0, dq1), max(0, dq1)].  The offsets sum to zero,
    so one jump has the sign opposite to dq1 and meets a zero bound: r = 0 and
    phi = 0 for every beta.  The reconstruction is therefore constant (stage,
    and the momenta either as momenta or as velocities times the vertex
    depth), and which beta the branch uses cannot change any result."""
    for vel in (0, 1):
        d = _two_triangle_square(beta_w=beta, beta_uh=beta, beta_vh=beta, extrapolate_velocity_second_order=vel)
        assert list(d.mesh.number_of_boundaries) == [2, 2]
        od = O.OracleDomain.from_domain(d)
        _set_centroids(od, np.array([0.7, 1.9]), np.array([0.4, -1.3]), np.array([-0.2, 0.8]))
        od.extrapolate_second_order_sw()
        wc = od.q("stage", "centroid_values")
        assert np.all(od.q("stage", "vertex_values") == wc[:, None])
        zc, zv = od.q("elevation", "centroid_values"), od.q("elevation", "vertex_values")
        for q in ("xmomentum", "ymomentum"):
            c = od.q(q, "centroid_values")
            if vel:
                u = c / np.maximum(wc - zc, d.minimum_allowed_height)
                exp = u[:, None] * np.maximum(wc[:, None] - zv, 0.0)
            else:
                exp = np.repeat(c[:, None], 3, axis=1)
            assert np.array_equal(od.q(q, "vertex_values"), exp), (q, vel)


# --------------------------------------------------------- sloped Manning
def test_sloped_manning_on_a_planar_bed():
    """_manning_friction_sloped (shallow_water_ext.c:1272-1317) on the plane
    z = a x + b y: S = -g n^2 sqrt(1 + a^2 + b^2) |q| / h^(7/3), h the depth
    over the mean vertex bed; the flat form lacks the slope factor."""
    a, b = 0.3, -0.45
    for sloped in (1, 0):
        d = _domain(4, 4, 40.0, 40.0, bed=lambda x, y: a * x + b * y + 5.0, use_sloped_mannings=sloped)
        d.set_quantity("friction", 0.04, location="centroids")
        od = O.OracleDomain.from_domain(d)
        N = d.number_of_elements
        rng = np.random.default_rng(35)
        h = rng.uniform(0.2, 2.0, N)
        uh, vh = rng.normal(size=N), rng.normal(size=N)
        _set_centroids(od, h, uh, vh)
        od.manning_friction_implicit()
        slope = math.sqrt(1 + a * a + b * b) if sloped else 1.0
        S = -G * 0.04 ** 2 * slope * np.hypot(uh, vh) / h ** (7.0 / 3.0)
        np.testing.assert_allclose(od.q("xmomentum", "semi_implicit_update"), S * uh, rtol=1e-11)
        np.testing.assert_allclose(od.q("ymomentum", "semi_implicit_update"), S * vh, rtol=1e-11)


# ------------------------------------------------------------ balance
def _balance_case(seed, **params):
    d = _domain(10, 8, **params)
    od = O.OracleDomain.from_domain(d)
    N = d.number_of_elements
    rng = np.random.default_rng(seed)
    zc, zv = od.q("elevation", "centroid_values"), od.q("elevation", "vertex_values")
    return d, od, N, rng, zc, zv


def test_balance_centroid_velocity_cutoff_and_alpha_clamp():
    """Tight slope limiters with centroid velocities: a centroid shallower
    than H0 gives alpha < 0, clamped to 0, so every vertex depth becomes the
    centroid depth (w_v = z_v + h_c, bitwise); the vertex momenta carry the
    centroid velocity uh_c / h_c, which is 0 for h_c <= 1e-6
    (shallow_water_ext.c:1462-1636)."""
    d, od, N, rng, zc, zv = _balance_case(36)
    h = rng.choice([3e-7, 1e-6, 4e-6, 2e-5, 3e-4, 8e-4], size=N)
    uh, vh = rng.normal(size=N) * 1e-3, rng.normal(size=N) * 1e-3
    _set_centroids(od, h, uh, vh)
    hc = od.q("stage", "centroid_values") - zc  # the depth the balance sees
    hv0 = hc[:, None] * rng.uniform(0.0, 3.0, size=(N, 3))
    hv0[:, 0] = 0.5 * hc  # one vertex shallower than the centroid: alpha < 0
    od.q("stage", "vertex_values")[:] = zv + hv0
    od.q("xmomentum", "vertex_values")[:] = rng.normal(size=(N, 3))
    od.q("ymomentum", "vertex_values")[:] = rng.normal(size=(N, 3))
    od.balance_deep_and_shallow()
    wv = od.q("stage", "vertex_values")
    assert np.array_equal(wv, zv + hc[:, None])
    cut = hc <= 1e-6
    assert cut.sum() > 10 and (~cut).sum() > 10
    for q, c in (("xmomentum", uh), ("ymomentum", vh)):
        mv = od.q(q, "vertex_values")
        assert np.all(mv[cut] == 0.0), q
        np.testing.assert_allclose(mv[~cut], np.repeat(c[~cut, None], 3, axis=1), rtol=1e-9)


def test_balance_momentum_blend_without_centroid_velocities():
    """use_centroid_velocities = 0: the vertex momenta are blended toward the
    centroid momentum with the same alpha as the depths,
    (m_v' - m_c) / (m_v - m_c) = (h_v' - h_c) / (h_v - h_c)."""
    d, od, N, rng, zc, zv = _balance_case(37, use_centroid_velocities=0)
    hc = rng.uniform(0.002, 0.05, N)
    _set_centroids(od, hc, rng.normal(size=N), rng.normal(size=N))
    hv0 = hc[:, None] * rng.uniform(-0.2, 2.0, size=(N, 3))
    od.q("stage", "vertex_values")[:] = zv + hv0
    m0 = {q: rng.normal(size=(N, 3)) for q in ("xmomentum", "ymomentum")}
    for q in m0:
        od.q(q, "vertex_values")[:] = m0[q]
    od.balance_deep_and_shallow()
    hv = od.q("stage", "vertex_values") - zv
    moved = (hv0.min(axis=1) < d.H0)[:, None] & (np.abs(hv0 - hc[:, None]) > 1e-6)
    assert moved.sum() > 30
    alpha = (hv - hc[:, None]) / (hv0 - hc[:, None])
    assert np.all(alpha[moved] < 1.0 - 1e-9)
    for q in m0:
        mc = od.q(q, "centroid_values")[:, None]
        frac = (od.q(q, "vertex_values") - mc) / (m0[q] - mc)
        np.testing.assert_allclose(frac[moved], alpha[moved], rtol=1e-8, atol=1e-12)


# -------------------------------------------- interpolation and gravity
def test_vertex_to_edge_values_of_a_linear_field():
    """quantity.c:687-707: edge i (opposite vertex i) gets the mean of the
    other two vertices, i.e. a linear field's value at the edge midpoint,
    which is v_i + 3/2 (c - v_i) for the centroid c."""
    d = _domain(6, 5)
    od = O.OracleDomain.from_domain(d)
    m = d.mesh
    f = lambda x, y: 2.5 * x - 1.25 * y + 3.0
    vx, vy = m.vertex_coordinates[:, 0::2], m.vertex_coordinates[:, 1::2]
    cx, cy = m.centroid_coordinates[:, 0:1], m.centroid_coordinates[:, 1:2]
    od.q("stage", "vertex_values")[:] = f(vx, vy)
    od.interpolate_from_vertices_to_edges()
    mid = f(vx + 1.5 * (cx - vx), vy + 1.5 * (cy - vy))
    np.testing.assert_allclose(od.q("stage", "edge_values"), mid, rtol=1e-12, atol=1e-12)


def test_vertex_momenta_carry_the_vertex_depth_factor():
    """extrapolate_velocity_second_order: velocities are reconstructed, then
    multiplied by the vertex depth max(w_v - z_v, 0) (shallow_water_ext.c:
    3292-3315).  A uniform velocity over a varying depth gives constant
    velocity gradients of zero, so the vertex momenta are u0 times the vertex
    depths."""
    d = _domain(10, 8)
    od = O.OracleDomain.from_domain(d)
    N = d.number_of_elements
    rng = np.random.default_rng(38)
    h = rng.uniform(0.5, 2.0, N)
    u0, v0 = 0.7, -0.35
    _set_centroids(od, h, u0 * h, v0 * h)
    od.extrapolate_second_order_sw()
    depth = np.maximum(od.q("stage", "vertex_values") - od.q("elevation", "vertex_values"), 0.0)
    assert np.ptp(depth) > 0.5
    np.testing.assert_allclose(od.q("xmomentum", "vertex_values"), u0 * depth, rtol=1e-12)
    np.testing.assert_allclose(od.q("ymomentum", "vertex_values"), v0 * depth, rtol=1e-12)


def test_gravity_on_a_tilted_plane_of_uniform_depth():
    """gravity_wb (shallow_water_ext.c:1942-2037) for a stage plane w = s x +
    c at uniform depth h: the stage-gradient term is -g s h and the edge
    pressure terms cancel around each triangle (sum of l_e n_e = 0)."""
    s, h = 0.002, 1.7
    d = _domain(8, 6, bed=lambda x, y: 0.002 * x + 3.0 - 1.7)
    od = O.OracleDomain.from_domain(d)
    m = d.mesh
    vx = m.vertex_coordinates[:, 0::2]
    w = lambda x: s * x + 3.0
    od.q("stage", "vertex_values")[:] = w(vx)
    od.interpolate_from_vertices_to_edges()
    od.q("stage", "centroid_values")[:] = od.q("stage", "vertex_values").mean(axis=1)
    od.q("elevation", "edge_values")[:] = od.q("stage", "edge_values") - h
    od.q("elevation", "centroid_values")[:] = od.q("stage", "centroid_values") - h
    for q in ("xmomentum", "ymomentum"):
        od.q(q, "explicit_update")[:] = 0.0
    od.gravity_wb()
    np.testing.assert_allclose(od.q("xmomentum", "explicit_update"), -G * s * h, rtol=1e-9)
    assert np.max(np.abs(od.q("ymomentum", "explicit_update"))) < 1e-9 * G * s * h


# ------------------------------------------------------- update_timestep
def test_update_timestep_cfl_and_clips():
    """update_timestep (evolve.c:115-160): dt = min(CFL flux_timestep,
    evolve_max_timestep), then clipped to the final time and the yield time;
    the recorded extremes follow dt before the clips."""
    d = _domain(2, 2, CFL=0.7, evolve_max_timestep=5.0)
    od = O.OracleDomain.from_domain(d)
    od.d.time, od.d.yieldtime = 10.0, 100.0
    od.d.flux_timestep = 0.3
    od.update_timestep(None)
    assert od.d.timestep == 0.7 * 0.3
    od.d.flux_timestep = 50.0
    od.update_timestep(None)
    assert od.d.timestep == 5.0
    assert od.d.recorded_max_timestep == 5.0 and od.d.recorded_min_timestep == 0.7 * 0.3
    od.d.yieldtime = 11.5
    od.update_timestep(None)
    assert od.d.timestep == 1.5
    od.update_timestep(finaltime=10.25)
    assert od.d.timestep == 0.25


def test_first_order_fallback_rule():
    """More than max_smallsteps consecutive steps below evolve_min_timestep
    switch a second-order run to first order (the counter restarts); a normal
    step switches back to the default order; more small steps at first order
    are an error (evolve.c:115-160, ANUGA's update_timestep)."""
    d = _domain(2, 2, evolve_min_timestep=1.0, max_smallsteps=4)
    od = O.OracleDomain.from_domain(d)
    od.d.yieldtime = 1e9
    od.d.flux_timestep = 0.5
    for i in range(4):
        od.update_timestep(None)
        assert od.d._order_ == 2 and od.d.smallsteps == i + 1
    od.update_timestep(None)
    assert od.d._order_ == 1 and od.d.smallsteps == 0
    od.d.flux_timestep = 2.0
    od.update_timestep(None)
    assert od.d._order_ == 2
    od.d.flux_timestep = 0.5
    for _ in range(5):
        od.update_timestep(None)
    assert od.d._order_ == 1
    for _ in range(4):
        od.update_timestep(None)
    with pytest.raises(O.OracleError) as e:
        od.update_timestep(None)
    assert e.value.code == -2
