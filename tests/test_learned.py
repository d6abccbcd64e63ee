"""Learned edge-flux surrogate: HIP kernels (BF16 MFMA forward and backward)
against the FP64 numpy restatement, mass conservation by construction, and
what training on FV trajectories gains over the untrained network."""
import numpy as np
import pytest

from anuga_cuda_amd.mesh import Mesh, rectangular_cross
from tests import gnn_reference as ref

pytestmark = pytest.mark.gpu


def _mesh():
    pts, tris, bnd = rectangular_cross(4, 3, 40.0, 30.0)
    return Mesh(pts, tris, boundary=bnd)


def _state(mesh, seed):
    rng = np.random.default_rng(seed)
    N = mesh.number_of_elements
    cx = mesh.centroid_coordinates[:, 0]
    z = 0.02 * cx + rng.uniform(-0.2, 0.2, N)
    h = rng.uniform(0.0, 2.0, N)
    h[rng.random(N) < 0.15] = 0.0  # some dry cells
    return z + h, rng.normal(0, 0.5, N) * (h > 0), rng.normal(0, 0.5, N) * (h > 0), z


def _graph_np(graph):
    cp = lambda t: t.detach().cpu().numpy()
    return dict(ea=cp(graph.ea)[:graph.E], eb=cp(graph.eb)[:graph.E], nx=cp(graph.nx), ny=cp(graph.ny),
                len=cp(graph.len), closed=cp(graph.closed), tri_edges=cp(graph.tri_edges), area=cp(graph.area))


def test_gnn_forward_and_gradients_match_cpu_reference():
    """Edge outputs, state change and every layer's weight / bias gradient of
    the HIP kernels against the FP64 numpy restatement (same BF16-rounded
    weights and features).  Bound: relative Frobenius error below 3e-2 (BF16
    activations, FP32 accumulation)."""
    import torch
    from anuga_cuda_amd import learned as L

    mesh = _mesh()
    graph = L.EdgeGraph(mesh, closed_tags=("left", "right"))
    net = L.EdgeGNN(seed=5, depth_scale=2.0, momentum_scale=1.0, length_scale=10.0, flux_scale=0.5,
                    sigma=(0.05, 0.05, 0.05))
    rng = np.random.default_rng(9)
    for name, K, Nn in L.LAYERS:
        net.W[name] = torch.tensor(rng.normal(0, 1.0 / np.sqrt(K), (K, Nn)).astype(np.float32), device="cuda")
        net.b[name] = torch.tensor(rng.normal(0, 0.1, Nn).astype(np.float32), device="cuda")
    net.cast()
    w, uh, vh, z = _state(mesh, 1)
    dev = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")
    tw, tuh, tvh, tz = dev(w), dev(uh), dev(vh), dev(z)
    delta = dev([0.7])
    target = rng.normal(0, 0.05, (graph.N, 3))
    tt = [dev(target[:, c]) for c in range(3)]
    ws = L.Workspace(graph)
    L.forward(net, graph, ws, tw, tuh, tvh, tz)
    w1, uh1, vh1 = torch.empty_like(tw), torch.empty_like(tw), torch.empty_like(tw)
    L.apply(net, graph, ws, delta, tw, tuh, tvh, w1, uh1, vh1, targets=tt)
    L.backward(net, graph, ws, delta)
    torch.cuda.synchronize()

    g = _graph_np(graph)
    params = dict(W={n: net.W[n].cpu().numpy().astype(np.float64) for n, _, _ in L.LAYERS},
                  b={n: net.b[n].cpu().numpy().astype(np.float64) for n, _, _ in L.LAYERS})
    X0 = ref.features(g, w, uh, vh, z, 2.0, 1.0, 10.0)
    X0_gpu = ws.X0[:graph.E].float().cpu().numpy()
    assert np.array_equal(X0_gpu, X0), "features differ"
    loss, F, dq, gW, gb = ref.loss_and_grads(params, g, X0, 0.7, 0.5, target, np.array([0.05, 0.05, 0.05]))

    def rel(a, b):
        return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)

    errs = {"F": rel(ws.F[:graph.E, :5].cpu().numpy(), F[:, :5]),
            "dq": rel(np.stack([(w1 - tw).cpu().numpy(), (uh1 - tuh).cpu().numpy(), (vh1 - tvh).cpu().numpy()], 1),
                      dq),
            "loss": abs(float(ws.loss_terms.sum()) / graph.N - loss) / loss}
    for n, _, _ in L.LAYERS:
        errs["gW." + n] = rel(ws.gW[n].cpu().numpy(), gW[n])
        errs["gb." + n] = rel(ws.gb[n].cpu().numpy(), gb[n])
    print({k: "%.2e" % v for k, v in errs.items()})
    for k, v in errs.items():
        assert v < 3e-2, (k, v)


def test_gnn_rollout_conserves_mass_in_a_closed_basin():
    """With every boundary a wall, the learned step moves mass only between
    triangles: total mass sum(A * h) stays constant to 1e-12 (relative) over 30
    steps of an untrained, randomly weighted network."""
    import torch
    from anuga_cuda_amd import learned as L

    mesh = _mesh()
    graph = L.EdgeGraph(mesh, closed_tags=("left", "right", "top", "bottom"))
    net = L.EdgeGNN(seed=2, flux_scale=0.05)
    rng = np.random.default_rng(4)
    net.W["head"] = torch.tensor(rng.normal(0, 0.3, (L.HID, L.OUT)).astype(np.float32), device="cuda")
    net.cast()
    w, uh, vh, z = _state(mesh, 3)
    dev = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")
    tw, tuh, tvh, tz = dev(w), dev(uh), dev(vh), dev(z)
    area = graph.area
    m0 = float(torch.sum(area * (tw - tz)))
    ws = L.Workspace(graph)
    delta = dev([0.5])
    for _ in range(30):
        L.forward(net, graph, ws, tw, tuh, tvh, tz)
        L.apply(net, graph, ws, delta, tw, tuh, tvh, tw, tuh, tvh)
    m1 = float(torch.sum(area * (tw - tz)))
    moved = float(torch.sum(area * torch.abs(tw - dev(w))))
    print("mass %.17g -> %.17g, relative change %.2e; mass moved %.3e" % (m0, m1, abs(m1 - m0) / m0, moved))
    assert moved > 1e-3 * m0  # the network does move water around
    assert abs(m1 - m0) <= 1e-12 * m0
