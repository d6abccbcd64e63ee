"""Learned edge-flux surrogate of the shallow-water step.

An edge graph network advances a shallow-water state by ``stride`` CFL steps
of the finite-volume (FV) solver at once.  Per edge it predicts a mass flux,
added to one side of the edge and subtracted from the other (so mass changes
only through open boundary edges, by construction), and momentum
contributions to each side.  Architecture (per edge, widths as in the
configuration of the learned workload):

    X0  = features(state)                  32   (depths, normal / tangential
                                                 momenta of both sides, stage
                                                 and bed jumps, flags, length)
    H1  = SiLU(X0 W0 + b0)                 64
    for l = 1, 2, 3 (message passing):
        S_i  = sum of H_l over the edges of triangle i
        C_l  = [H_l, S_a + S_b]            128
        H_l+1 = H_l + SiLU(C_l W_l + b_l)  64
    F   = H4 W4 + b4                        32   (5 used)
    q_i(t + delta) = q_i(t) + delta / A_i * sum of side-signed, rotated F * len

Every layer, its backward pass, the message passing, the state update (FP64)
and Adam run in the hand-written HIP kernels of ``csrc/gnn_kernels.hip``
(BF16 MFMA with FP32 accumulation); torch only allocates device memory and
provides the stream.  Training data are trajectories of the FV kernels.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
GNN_LIB_PATH = os.path.join(_HERE, "libanuga_gnn.so")

FEAT, HID, CAT, OUT = 32, 64, 128, 32
LAYERS = [("enc", FEAT, HID), ("mp1", CAT, HID), ("mp2", CAT, HID), ("mp3", CAT, HID), ("head", HID, OUT)]

_gnn = None


def _load():
    global _gnn
    if _gnn is None:
        if not os.path.exists(GNN_LIB_PATH):
            raise ImportError("libanuga_gnn.so not built (run anuga_cuda_amd/csrc/build.sh); there is no CPU fallback")
        _gnn = C.CDLL(GNN_LIB_PATH)
        for name in ("gnn_features", "gnn_linear_fwd", "gnn_linear_bwd_x", "gnn_linear_bwd_w", "gnn_silu_bwd",
                     "gnn_node_sum", "gnn_concat", "gnn_node_grad", "gnn_edge_grad", "gnn_apply", "gnn_apply_bwd",
                     "gnn_adam", "gnn_cast_weights"):
            getattr(_gnn, name).restype = C.c_int
        f, i64, i32, p = C.c_float, C.c_int64, C.c_int, C.c_void_p
        _gnn.gnn_features.argtypes = [p, i64, i64] + [p] * 10 + [f, f, f, p]
        _gnn.gnn_linear_fwd.argtypes = [p, i64, i32, i32, p, p, p, p, p, p, i32, i32]
        _gnn.gnn_linear_bwd_x.argtypes = [p, i64, i32, i32, p, p, p]
        _gnn.gnn_linear_bwd_w.argtypes = [p, i64, i32, i32, p, p, p, p, p]
        _gnn.gnn_silu_bwd.argtypes = [p, i64, p, p, p]
        _gnn.gnn_node_sum.argtypes = [p, i64, p, p, p]
        _gnn.gnn_concat.argtypes = [p, i64, i64, p, p, p, p, p]
        _gnn.gnn_node_grad.argtypes = [p, i64, p, p, p]
        _gnn.gnn_edge_grad.argtypes = [p, i64, i64, p, p, p, p, p, p]
        _gnn.gnn_apply.argtypes = [p, i64] + [p] * 9 + [p, f] + [p] * 12
        _gnn.gnn_apply_bwd.argtypes = [p, i64, i64] + [p] * 9 + [f, p, p]
        _gnn.gnn_adam.argtypes = [p, i64, p, p, p, p, f, f, f, f, i32]
        _gnn.gnn_cast_weights.argtypes = [p, i32, i32, p, p, p]
    return _gnn


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("the learned surrogate runs on the GPU only")
    return torch


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _ck(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed" % what)


# ---------------------------------------------------------------------------
class EdgeGraph:
    """Edges (interior once, boundary edges too) of ``copies`` copies of a mesh,
    and each triangle's edges, on the device."""

    def __init__(self, mesh, copies=1, closed_tags=("exterior",)):
        torch = _torch()
        N = mesh.number_of_elements
        nbr = mesh.neighbours
        k = np.repeat(np.arange(N), 3)
        i = np.tile(np.arange(3), N)
        n = nbr.ravel()
        keep = (n < 0) | (k < n)
        ka, ia, nb = k[keep], i[keep], n[keep]
        E = len(ka)
        eb = np.where(nb >= 0, nb, -1)
        nrm = mesh.normals.reshape(N, 3, 2)[ka, ia]
        ln = mesh.edgelengths[ka, ia]
        tags = np.array(["" for _ in range(E)], dtype=object)
        bidx = np.nonzero(nb < 0)[0]
        tags[bidx] = mesh.boundary_tags[-nb[bidx] - 1]
        closed = np.isin(tags, list(closed_tags)).astype(np.uint8)
        # triangle -> its edges (edge index of slot i of triangle k)
        slot_edge = -np.ones((N, 3), dtype=np.int64)
        slot_edge[ka, ia] = np.arange(E)
        part = mesh.neighbour_edges.ravel()[keep]
        inter = nb >= 0
        slot_edge[nb[inter], part[inter]] = np.arange(E)[inter]
        B = int(copies)
        self.N1, self.E1, self.copies = N, E, B
        self.N = N * B
        self.E = E * B
        self.Epad = (self.E + 63) // 64 * 64
        offs_n = (np.arange(B) * N)[:, None]
        offs_e = (np.arange(B) * E)[:, None]
        ea_all = (ka[None, :] + offs_n).ravel()
        eb_all = np.where(eb[None, :] >= 0, eb[None, :] + offs_n, -1).ravel()
        te = slot_edge[None, :, :] + offs_e[:, :, None]
        te = np.where(slot_edge[None, :, :] >= 0, te, -1).reshape(-1, 3)
        dev = "cuda"
        self.ea = torch.tensor(ea_all, dtype=torch.int32, device=dev)
        self.eb = torch.tensor(eb_all, dtype=torch.int32, device=dev)
        self.nx = torch.tensor(np.tile(nrm[:, 0], B), dtype=torch.float32, device=dev)
        self.ny = torch.tensor(np.tile(nrm[:, 1], B), dtype=torch.float32, device=dev)
        self.len = torch.tensor(np.tile(ln, B), dtype=torch.float32, device=dev)
        self.closed = torch.tensor(np.tile(closed, B), dtype=torch.uint8, device=dev)
        self.tri_edges = torch.tensor(te, dtype=torch.int32, device=dev)
        self.area = torch.tensor(np.tile(mesh.areas, B), dtype=torch.float64, device=dev)
        self.graph_of = torch.tensor(np.repeat(np.arange(B), N), dtype=torch.int32, device=dev)
        self.mean_len = float(np.mean(ln))


# ---------------------------------------------------------------------------
class EdgeGNN:
    """The network (FP32 master weights, BF16 copies) and its scales."""

    def __init__(self, seed=0, depth_scale=5.0, momentum_scale=5.0, length_scale=10.0, flux_scale=1.0,
                 sigma=(1e-2, 1e-2, 1e-2)):
        torch = _torch()
        self.lib = _load()
        g = np.random.default_rng(seed)
        self.scales = dict(depth=float(depth_scale), momentum=float(momentum_scale), length=float(length_scale),
                           flux=float(flux_scale))
        self.sigma = torch.tensor(np.asarray(sigma, dtype=np.float32), device="cuda")
        self.W, self.b = {}, {}
        for name, K, Nn in LAYERS:
            std = np.sqrt(2.0 / (K + Nn))
            w = g.normal(0.0, std, (K, Nn)).astype(np.float32)
            if name == "head":
                w[:] = 0.0  # the untrained network leaves the state unchanged
            self.W[name] = torch.tensor(w, device="cuda")
            self.b[name] = torch.zeros(Nn, dtype=torch.float32, device="cuda")
        self.Wb = {n: torch.empty(K, Nn, dtype=torch.bfloat16, device="cuda") for n, K, Nn in LAYERS}
        self.WTb = {n: torch.empty(Nn, K, dtype=torch.bfloat16, device="cuda") for n, K, Nn in LAYERS}
        self.m = {}
        self.v = {}
        self.step_count = 0
        self.cast()

    def stream(self):
        import torch

        return C.c_void_p(torch.cuda.current_stream().cuda_stream)

    def cast(self):
        s = self.stream()
        for name, K, Nn in LAYERS:
            _ck(self.lib.gnn_cast_weights(s, K, Nn, _ptr(self.W[name]), _ptr(self.Wb[name]), _ptr(self.WTb[name])),
                "cast")

    def parameters(self):
        for name, _, _ in LAYERS:
            yield name + ".W", self.W[name]
            yield name + ".b", self.b[name]

    def state_dict(self):
        d = {k: v.detach().cpu().numpy() for k, v in self.parameters()}
        d["sigma"] = self.sigma.cpu().numpy()
        for k, v in self.scales.items():
            d["scale." + k] = np.float64(v)
        return d

    def load_state_dict(self, d):
        torch = _torch()
        for name, _, _ in LAYERS:
            self.W[name] = torch.tensor(np.asarray(d[name + ".W"]), dtype=torch.float32, device="cuda")
            self.b[name] = torch.tensor(np.asarray(d[name + ".b"]), dtype=torch.float32, device="cuda")
        self.sigma = torch.tensor(np.asarray(d["sigma"], dtype=np.float32), device="cuda")
        for k in self.scales:
            self.scales[k] = float(d["scale." + k])
        self.cast()

    def save(self, path):
        np.savez(path, **self.state_dict())

    @classmethod
    def load(cls, path):
        net = cls()
        with np.load(path, allow_pickle=False) as d:
            net.load_state_dict({k: d[k] for k in d.files})
        return net


class Workspace:
    """Activations and gradients of one graph (BF16 unless stated)."""

    def __init__(self, graph):
        torch = _torch()
        E, N = graph.Epad, graph.N
        bf = dict(dtype=torch.bfloat16, device="cuda")
        self.X0 = torch.empty(E, FEAT, **bf)
        self.Z = [torch.empty(E, HID, **bf) for _ in range(4)]
        self.H = [torch.empty(E, HID, **bf) for _ in range(5)]  # H[1..4]
        self.Cc = [torch.empty(E, CAT, **bf) for _ in range(4)]  # C[1..3]
        self.S = torch.empty(N, HID, **bf)
        self.F = torch.empty(E, OUT, dtype=torch.float32, device="cuda")
        # backward
        self.dF = torch.empty(E, OUT, **bf)
        self.dH = torch.empty(E, HID, **bf)
        self.dHn = torch.empty(E, HID, **bf)
        self.dZ = torch.empty(E, HID, **bf)
        self.dC = torch.empty(E, CAT, **bf)
        self.dS = torch.empty(N, HID, **bf)
        self.partial = torch.empty(256 * (CAT + 1) * HID, dtype=torch.float32, device="cuda")
        self.gq = torch.empty(N, 3, dtype=torch.float32, device="cuda")
        self.loss_terms = torch.empty(N, dtype=torch.float32, device="cuda")
        self.gW = {n: torch.empty(K, Nn, dtype=torch.float32, device="cuda") for n, K, Nn in LAYERS}
        self.gb = {n: torch.empty(Nn, dtype=torch.float32, device="cuda") for n, K, Nn in LAYERS}


def forward(net, graph, ws, w, uh, vh, z):
    """Edge outputs F (FP32 [Epad][32]) of the state (FP64 device tensors)."""
    L, s = net.lib, net.stream()
    g = graph
    _ck(L.gnn_features(s, g.E, g.Epad, _ptr(g.ea), _ptr(g.eb), _ptr(g.nx), _ptr(g.ny), _ptr(g.len), _ptr(g.closed),
                       _ptr(w), _ptr(uh), _ptr(vh), _ptr(z), net.scales["depth"], net.scales["momentum"],
                       net.scales["length"], _ptr(ws.X0)), "features")
    _ck(L.gnn_linear_fwd(s, g.Epad, FEAT, HID, _ptr(ws.X0), _ptr(net.WTb["enc"]), _ptr(net.b["enc"]), None,
                         _ptr(ws.Z[0]), _ptr(ws.H[1]), 1, 0), "encoder")
    for l in (1, 2, 3):
        name = "mp%d" % l
        _ck(L.gnn_node_sum(s, g.N, _ptr(g.tri_edges), _ptr(ws.H[l]), _ptr(ws.S)), "node sum")
        _ck(L.gnn_concat(s, g.E, g.Epad, _ptr(g.ea), _ptr(g.eb), _ptr(ws.H[l]), _ptr(ws.S), _ptr(ws.Cc[l])),
            "concat")
        _ck(L.gnn_linear_fwd(s, g.Epad, CAT, HID, _ptr(ws.Cc[l]), _ptr(net.WTb[name]), _ptr(net.b[name]),
                             _ptr(ws.H[l]), _ptr(ws.Z[l]), _ptr(ws.H[l + 1]), 1, 0), name)
    _ck(L.gnn_linear_fwd(s, g.Epad, HID, OUT, _ptr(ws.H[4]), _ptr(net.WTb["head"]), _ptr(net.b["head"]), None, None,
                         _ptr(ws.F), 0, 1), "head")
    return ws.F


def apply(net, graph, ws, delta, w0, uh0, vh0, w1, uh1, vh1, targets=None):
    """State after one learned step (FP64); with targets (true changes), also
    the loss gradient in ws.gq and loss terms in ws.loss_terms."""
    L, s, g = net.lib, net.stream(), graph
    dw, duh, dvh = targets if targets is not None else (None, None, None)
    _ck(L.gnn_apply(s, g.N, _ptr(g.tri_edges), _ptr(g.ea), _ptr(g.nx), _ptr(g.ny), _ptr(g.len), _ptr(g.closed),
                    _ptr(g.area), _ptr(delta), _ptr(g.graph_of), _ptr(ws.F), net.scales["flux"], _ptr(w0), _ptr(uh0),
                    _ptr(vh0), _ptr(w1), _ptr(uh1), _ptr(vh1), _ptr(dw), _ptr(duh), _ptr(dvh), _ptr(net.sigma),
                    _ptr(ws.gq), _ptr(ws.loss_terms)), "apply")


def backward(net, graph, ws, delta):
    """Weight gradients (ws.gW, ws.gb) from ws.gq."""
    L, s, g = net.lib, net.stream(), graph
    _ck(L.gnn_apply_bwd(s, g.E, g.Epad, _ptr(g.ea), _ptr(g.eb), _ptr(g.nx), _ptr(g.ny), _ptr(g.len), _ptr(g.closed),
                        _ptr(g.area), _ptr(delta), _ptr(g.graph_of), net.scales["flux"], _ptr(ws.gq), _ptr(ws.dF)),
        "apply bwd")
    _ck(L.gnn_linear_bwd_w(s, g.Epad, HID, OUT, _ptr(ws.H[4]), _ptr(ws.dF), _ptr(ws.partial), _ptr(ws.gW["head"]),
                           _ptr(ws.gb["head"])), "head bwd w")
    _ck(L.gnn_linear_bwd_x(s, g.Epad, HID, OUT, _ptr(ws.dF), _ptr(net.Wb["head"]), _ptr(ws.dH)), "head bwd x")
    for l in (3, 2, 1):
        name = "mp%d" % l
        # H[l+1] = H[l] + SiLU(Z[l]), Z[l] = C[l] W + b; dH holds dL/dH[l+1]
        _ck(L.gnn_silu_bwd(s, g.Epad * HID, _ptr(ws.dH), _ptr(ws.Z[l]), _ptr(ws.dZ)), "silu bwd")
        _ck(L.gnn_linear_bwd_w(s, g.Epad, CAT, HID, _ptr(ws.Cc[l]), _ptr(ws.dZ), _ptr(ws.partial),
                               _ptr(ws.gW[name]), _ptr(ws.gb[name])), name + " bwd w")
        _ck(L.gnn_linear_bwd_x(s, g.Epad, CAT, HID, _ptr(ws.dZ), _ptr(net.Wb[name]), _ptr(ws.dC)), name + " bwd x")
        _ck(L.gnn_node_grad(s, g.N, _ptr(g.tri_edges), _ptr(ws.dC), _ptr(ws.dS)), "node grad")
        _ck(L.gnn_edge_grad(s, g.E, g.Epad, _ptr(g.ea), _ptr(g.eb), _ptr(ws.dH), _ptr(ws.dC), _ptr(ws.dS),
                            _ptr(ws.dHn)), "edge grad")
        ws.dH, ws.dHn = ws.dHn, ws.dH
    _ck(L.gnn_silu_bwd(s, g.Epad * HID, _ptr(ws.dH), _ptr(ws.Z[0]), _ptr(ws.dZ)), "silu bwd")
    _ck(L.gnn_linear_bwd_w(s, g.Epad, FEAT, HID, _ptr(ws.X0), _ptr(ws.dZ), _ptr(ws.partial), _ptr(ws.gW["enc"]),
                           _ptr(ws.gb["enc"])), "enc bwd w")


def adam_step(net, ws, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
    torch = _torch()
    L, s = net.lib, net.stream()
    net.step_count += 1
    for name, K, Nn in LAYERS:
        for key, p, gr in ((name + ".W", net.W[name], ws.gW[name]), (name + ".b", net.b[name], ws.gb[name])):
            if key not in net.m:
                net.m[key] = torch.zeros_like(p)
                net.v[key] = torch.zeros_like(p)
            _ck(L.gnn_adam(s, p.numel(), _ptr(p), _ptr(gr), _ptr(net.m[key]), _ptr(net.v[key]), lr, beta1, beta2,
                           eps, net.step_count), "adam")
    net.cast()
