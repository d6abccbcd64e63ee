"""CPU checks of the numpy restatement of the learned surrogate's physical
edge flux (tests/gnn_reference.base_flux), which the GPU tests compare the HIP
kernel against: on a flat bed it is the solver's own central-upwind flux
(the oracle's _flux_function_central for deep states), and it is exactly zero
for a lake at rest over any bed, dry cells included."""
import math

import numpy as np

from oracle import oracle as O
from tests import gnn_reference as ref

G = 9.8


def _edge_graph(n1, n2, closed=False, boundary=False):
    return dict(ea=np.array([0]), eb=np.array([-1 if boundary else 1]), nx=np.array([n1]), ny=np.array([n2]),
                closed=np.array([1 if closed else 0]))


def test_flat_bed_base_flux_is_the_solvers_central_upwind_flux():
    """Deep states (h above the solver's velocity-damping threshold 10 H0) on
    a flat bed: the hydrostatic reconstruction is the identity and both
    fluxes are the KNP flux of the same states.  The restatement's side-a
    normal momentum plus its pressure, and side b's, both equal the oracle's
    normal momentum flux."""
    rng = np.random.default_rng(3)
    for _ in range(200):
        zb = rng.uniform(-3.0, 3.0)
        hl, hr = rng.uniform(0.05, 6.0, 2)
        ql = np.array([zb + hl, *rng.uniform(-4, 4, 2) * hl])
        qr = np.array([zb + hr, *rng.uniform(-4, 4, 2) * hr])
        ang = rng.uniform(0, 2 * math.pi)
        n1, n2 = math.cos(ang), math.sin(ang)
        ef, _ = O.flux_function_central(ql, qr, zb, zb, n1, n2)
        kn = np.array([ef[0], n1 * ef[1] + n2 * ef[2], -n2 * ef[1] + n1 * ef[2]])  # normal frame
        w, uh, vh, z = (np.array([ql[0], qr[0]]), np.array([ql[1], qr[1]]), np.array([ql[2], qr[2]]),
                        np.array([zb, zb]))
        P = ref.base_flux(_edge_graph(n1, n2), w, uh, vh, z, 1.0)[0]
        scale = max(np.max(np.abs(kn)), 1.0)
        np.testing.assert_allclose(P[0], kn[0], atol=1e-12 * scale)
        np.testing.assert_allclose(-P[1] + 0.5 * G * hl * hl, kn[1], atol=1e-12 * scale)
        np.testing.assert_allclose(P[3] + 0.5 * G * hr * hr, kn[1], atol=1e-12 * scale)
        np.testing.assert_allclose([-P[2], P[4]], [kn[2], kn[2]], atol=1e-12 * scale)


def test_lake_at_rest_gives_zero_base_flux():
    """Equal stage, zero momentum: every output is zero (to the rounding of
    this restatement's direct form; the HIP kernel's arrangement gives
    exactly 0, checked on the GPU) on interior, wall and open boundary edges,
    over bed steps and with one side dry."""
    rng = np.random.default_rng(5)
    for _ in range(300):
        stage = rng.uniform(-1, 3)
        z = rng.uniform(-4, 4, 2)
        w = np.maximum(z, stage)  # a cell above the lake is dry (w = z)
        ang = rng.uniform(0, 2 * math.pi)
        for closed, boundary in ((False, False), (True, True), (False, True)):
            P = ref.base_flux(_edge_graph(math.cos(ang), math.sin(ang), closed, boundary), w, np.zeros(2),
                              np.zeros(2), z, 0.7)
            assert np.all(np.abs(P) <= 1e-15 * 0.5 * G * 50.0), (P, w, z)


def test_open_boundary_base_flux_is_the_physical_outflow():
    """Open (transmissive) boundary edge: the outside state is the inside
    state, so the flux is the physical flux of the inside state: mass flux
    s u, side-a normal momentum -(s u^2) once its pressure is taken out (a
    boundary edge has no side b)."""
    h, u, v = 2.0, 1.5, -0.4
    P = ref.base_flux(_edge_graph(1.0, 0.0, boundary=True), np.array([h, h]), np.array([u * h, 0]),
                      np.array([v * h, 0]), np.zeros(2), 1.0)[0]
    np.testing.assert_allclose(P[:3], [h * u, -h * u * u, -h * u * v], rtol=1e-14, atol=1e-14)
