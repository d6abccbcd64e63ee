"""Learned edge-flux surrogate of the shallow-water step.

An edge graph network advances a shallow-water state by a time step delta
that spans many CFL steps of the finite-volume (FV) solver (whose global
timestep is set by the mesh's worst triangle).  Per edge it predicts a mass flux,
added to one side of the edge and subtracted from the other (so mass changes
only through open boundary edges, by construction), and momentum
contributions to each side.  Architecture (per edge, widths as in the
configuration of the learned workload):

    X0  = features(state)                  32   (depths, normal / tangential
                                                 momenta of both sides, stage
                                                 and bed jumps, flags, length,
                                                 first-order central-upwind
                                                 flux after hydrostatic
                                                 reconstruction)
    H1  = SiLU(X0 W0 + b0)                 64
    for l = 1, 2, 3 (message passing):
        S_i  = sum of H_l over the edges of triangle i
        C_l  = [H_l, S_a + S_b]            128
        H_l+1 = H_l + SiLU(C_l W_l + b_l)  64
    F   = H4 W4 + b4 + P                    32   (5 used)
    q_i(t + delta) = q_i(t) + delta / A_i * sum of side-signed, rotated F * len

P is the physical part of the edge outputs (gnn_base_flux): the first-order
central-upwind flux of the centroid states after hydrostatic reconstruction,
in FP64, exactly zero for a lake at rest.  The network learns the correction
that turns one long first-order step into the FV solver's many short
second-order steps; the untrained network (zero head) is the first-order
scheme with the long step.

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
FOUT = 8  # row of the FP32 edge outputs (5 used)
GRAVITY = 9.8
H_DRY = 1.0e-3  # depth below which a cell's velocity is taken as zero (the solver's minimum_allowed_height)
# Largest Froude number a learned step may leave in a cell (protect caps the
# speed at FROUDE_MAX sqrt(g h)); 0 switches the cap off.
FROUDE_MAX = 5.0
LAYERS = [("enc", FEAT, HID), ("mp1", CAT, HID), ("mp2", CAT, HID), ("mp3", CAT, HID), ("head", HID, OUT)]

_gnn = None


def _load():
    global _gnn
    if _gnn is None:
        if not os.path.exists(GNN_LIB_PATH):
            raise ImportError("libanuga_gnn.so not built (run anuga_cuda_amd/csrc/build.sh); there is no CPU fallback")
        from ._lib import share_hip_runtime

        share_hip_runtime()
        _gnn = C.CDLL(GNN_LIB_PATH)
        for name in ("gnn_features", "gnn_linear_fwd", "gnn_linear_bwd_x", "gnn_linear_bwd_w", "gnn_silu_bwd",
                     "gnn_node_sum", "gnn_concat", "gnn_node_grad", "gnn_edge_grad", "gnn_apply", "gnn_apply_bwd",
                     "gnn_limit", "gnn_protect", "gnn_adam", "gnn_cast_weights", "gnn_base_flux",
                     "gnn_feat_enc", "gnn_head", "gnn_mp_direct"):
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
        _gnn.gnn_apply.argtypes = [p, i64] + [p] * 9 + [f] + [p] * 11 + [C.c_double, p, p, p]
        _gnn.gnn_limit.argtypes = [p, i64] + [p] * 7 + [f, p, p, p, p]
        _gnn.gnn_protect.argtypes = [p, i64, p, p, p, p, C.c_double, C.c_double, C.c_double]
        _gnn.gnn_apply_bwd.argtypes = [p, i64, i64] + [p] * 6 + [f, p, p]
        _gnn.gnn_adam.argtypes = [p, i64, p, p, p, p, f, f, f, f, i32]
        _gnn.gnn_cast_weights.argtypes = [p, i32, i32, p, p, p]
        _gnn.gnn_base_flux.argtypes = [p, i64] + [p] * 9 + [C.c_double, C.c_double, f, p]
        _gnn.gnn_feat_enc.argtypes = [p, i64, i64] + [p] * 10 + [f, f, f, p, p, C.c_double, C.c_double, f] + [p] * 4
        _gnn.gnn_head.argtypes = [p, i64] + [p] * 5
        _gnn.gnn_mp_direct.argtypes = [p, i64, i64] + [p] * 10
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
def small_cell_groups(mesh, merge_radius, max_size=4):
    """Agglomerates small cells for the learned update: a triangle whose
    inscribed-circle radius 2A/P is below ``merge_radius`` joins the group
    across its longest edge, and a group whose radius 2A/P (outer perimeter)
    is still below it joins the group across its longest outer edge, up to
    ``max_size`` members.  Sliver triangles of unstructured meshes would
    otherwise see a change many times their content over one long step.
    Returns (members [N, 4] int32, sorted, -1 padded; group area [N]) or
    (None, None) if no triangle is merged."""
    N = mesh.number_of_elements
    A = np.asarray(mesh.areas, dtype=np.float64)
    Lg = np.asarray(mesh.edgelengths, dtype=np.float64)
    nbr = np.asarray(mesh.neighbours)
    P = Lg.sum(1)
    small = np.nonzero(2.0 * A / P < merge_radius)[0]
    if merge_radius <= 0.0 or small.size == 0:
        return None, None
    small = small[np.argsort(2.0 * A[small] / P[small])]
    root = np.arange(N)
    members = {}
    gA, gP = A.copy(), P.copy()

    def find(i):
        while root[i] != i:
            root[i] = root[root[i]]
            i = root[i]
        return i

    for _ in range(3):
        changed = False
        for i in small:
            gi = find(i)
            if 2.0 * gA[gi] / gP[gi] >= merge_radius:
                continue
            mi = members.get(gi, [gi])
            best, best_len = -1, -1.0
            for m in mi:
                for q in range(3):
                    j = nbr[m, q]
                    if j < 0:
                        continue
                    gj = find(j)
                    if gj == gi or len(mi) + len(members.get(gj, [gj])) > max_size:
                        continue
                    if Lg[m, q] > best_len:
                        best, best_len = gj, Lg[m, q]
            if best < 0:
                continue
            shared = sum(Lg[m, q] for m in mi for q in range(3) if nbr[m, q] >= 0 and find(nbr[m, q]) == best)
            mj = members.pop(best, [best])
            members.pop(gi, None)
            r, o = (gi, best) if gi < best else (best, gi)
            root[o] = r
            gA[r] = gA[gi] + gA[best]
            gP[r] = gP[gi] + gP[best] - 2.0 * shared
            members[r] = sorted(mi + mj)
            changed = True
        if not changed:
            break
    grp = np.tile(np.arange(N, dtype=np.int32)[:, None], (1, 4))
    grp[:, 1:] = -1
    garea = A.copy()
    for r, mem in members.items():
        row = np.full(4, -1, dtype=np.int32)
        row[:len(mem)] = mem
        grp[mem] = row
        garea[mem] = gA[r]
    return grp, garea


class EdgeGraph:
    """Edges (interior once, boundary edges too) of ``copies`` copies of a mesh,
    and each triangle's edges, on the device.  ``merge_radius`` > 0 merges
    small cells for the update (small_cell_groups)."""

    def __init__(self, mesh, copies=1, closed_tags=("exterior",), merge_radius=0.0):
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
        # packed edge records of the per-triangle kernels: ea, eb, nx, ny, len
        # (FP32 bits), closed, 0, 0
        rec = np.zeros((B * E, 8), dtype=np.int32)
        rec[:, 0], rec[:, 1] = ea_all, eb_all
        rec[:, 2] = np.tile(nrm[:, 0], B).astype(np.float32).view(np.int32)
        rec[:, 3] = np.tile(nrm[:, 1], B).astype(np.float32).view(np.int32)
        rec[:, 4] = np.tile(ln, B).astype(np.float32).view(np.int32)
        rec[:, 5] = np.tile(closed, B)
        self.rec = torch.tensor(rec, device=dev)
        # per triangle slot (tri_edges order) the edge's data in the order the
        # limiter and the update stream it: {edge, triangle across (-1:
        # boundary), FP32 bits of len, nx}, ny apart, flags per triangle
        # (bits 3q: an edge, 3q+1: closed, 3q+2: the triangle is side a)
        NB = B * N
        tri = np.repeat(np.arange(NB), 3)
        e_s = te.ravel()
        has = e_s >= 0
        es = np.where(has, e_s, 0)
        side_a = has & (ea_all[es] == tri)
        other = np.where(side_a, eb_all[es], ea_all[es])
        slots = np.zeros((NB * 3, 4), dtype=np.int32)
        slots[:, 0] = np.where(has, e_s, -1)
        slots[:, 1] = np.where(has, other, -1)
        ln_all = np.tile(ln, B).astype(np.float32)
        nx_all = np.tile(nrm[:, 0], B).astype(np.float32)
        ny_all = np.tile(nrm[:, 1], B).astype(np.float32)
        slots[:, 2] = np.where(has, ln_all[es].view(np.int32), 0)
        slots[:, 3] = np.where(has, nx_all[es].view(np.int32), 0)
        sny = np.where(has, ny_all[es], np.float32(0.0)).astype(np.float32)
        closed_all = np.tile(closed, B)
        bits = (has.astype(np.int32) | ((has & (closed_all[es] != 0)).astype(np.int32) << 1)
                | (side_a.astype(np.int32) << 2)).reshape(NB, 3)
        sflags = (bits[:, 0] | (bits[:, 1] << 3) | (bits[:, 2] << 6)).astype(np.int32)
        self.slots = torch.tensor(slots, device=dev)
        self.sny = torch.tensor(sny, device=dev)
        self.sflags = torch.tensor(sflags, device=dev)
        # per edge, the edges of its two triangles in tri_edges order (ea's,
        # then eb's; -1 for a missing triangle and the padding rows): the
        # inference layers form the node sums from these rows
        nb6 = -np.ones((self.Epad, 8), dtype=np.int32)
        nb6[:self.E, 0:3] = te[ea_all]
        has_b = eb_all >= 0
        nb6[:self.E, 3:6][has_b] = te[eb_all[has_b]]
        self.nbr6 = torch.tensor(nb6, device=dev)
        self.area = torch.tensor(np.tile(mesh.areas, B), dtype=torch.float64, device=dev)
        self.graph_of = torch.tensor(np.repeat(np.arange(B), N), dtype=torch.int32, device=dev)
        self.mean_len = float(np.mean(ln))
        self.mean_area = float(np.mean(mesh.areas))
        grp, garea = small_cell_groups(mesh, merge_radius)
        self.group = self.group_area = None
        self.merged = 0
        if grp is not None:
            self.merged = int(np.sum(grp[:, 1] >= 0))
            g_all = np.where(grp[None, :, :] >= 0, grp[None, :, :] + offs_n[:, :, None], -1).reshape(-1, 4)
            self.group = torch.tensor(g_all, dtype=torch.int32, device=dev)
            self.group_area = torch.tensor(np.tile(garea, B), dtype=torch.float64, device=dev)


# ---------------------------------------------------------------------------
class EdgeGNN:
    """The network (FP32 master weights, BF16 copies) and its scales."""

    def __init__(self, seed=0, depth_scale=5.0, momentum_scale=5.0, length_scale=10.0, flux_scale=1.0,
                 sigma=(1e-2, 1e-2, 1e-2), merge_radius=0.0, physics_base=True, delta=0.0):
        torch = _torch()
        self.lib = _load()
        g = np.random.default_rng(seed)
        # merge_radius: the small-cell agglomeration radius of the graphs the
        # network is trained and run on (small_cell_groups).  physics_base:
        # the edge outputs are the first-order central-upwind flux
        # (gnn_base_flux) plus the network's correction; without it the
        # network predicts the whole flux.  delta: the learned step (seconds)
        # the network was trained for (0: not recorded).
        self.scales = dict(depth=float(depth_scale), momentum=float(momentum_scale), length=float(length_scale),
                           flux=float(flux_scale), merge_radius=float(merge_radius),
                           physics_base=1.0 if physics_base else 0.0, delta=float(delta))
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
            if "scale." + k in d:  # files written before a scale existed keep its default
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
        self.F = torch.empty(E, FOUT, dtype=torch.float32, device="cuda")
        self.P = torch.empty(E, FOUT, dtype=torch.float32, device="cuda")  # physical part of F
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
        self.theta = torch.empty(N, dtype=torch.float64, device="cuda")
        self.cells = torch.empty(N, 4, dtype=torch.float64, device="cuda")  # per-triangle sums of merged groups
        self.gW = {n: torch.empty(K, Nn, dtype=torch.float32, device="cuda") for n, K, Nn in LAYERS}
        self.gb = {n: torch.empty(Nn, dtype=torch.float32, device="cuda") for n, K, Nn in LAYERS}


def forward(net, graph, ws, w, uh, vh, z, save=True):
    """Edge outputs F (FP32 [Epad][8]) of the state (FP64 device tensors).
    save=True (training) keeps the features, concatenated inputs and
    pre-activations the backward pass needs; save=False (inference) runs the
    persistent fused kernels (features + physical part + encoder, then
    message-passing layers with the concatenation fused in), with bitwise the
    same result."""
    L, s = net.lib, net.stream()
    g = graph
    base = net.scales["physics_base"] != 0.0
    P = _ptr(ws.P) if base else None
    sc = (net.scales["depth"], net.scales["momentum"], net.scales["length"])
    if save:
        _ck(L.gnn_features(s, g.E, g.Epad, _ptr(g.ea), _ptr(g.eb), _ptr(g.nx), _ptr(g.ny), _ptr(g.len),
                           _ptr(g.closed), _ptr(w), _ptr(uh), _ptr(vh), _ptr(z), *sc, _ptr(ws.X0)), "features")
        if base:
            _ck(L.gnn_base_flux(s, g.E, _ptr(g.ea), _ptr(g.eb), _ptr(g.nx), _ptr(g.ny), _ptr(g.closed), _ptr(w),
                                _ptr(uh), _ptr(vh), _ptr(z), GRAVITY, H_DRY, net.scales["flux"], P), "base flux")
        _ck(L.gnn_linear_fwd(s, g.Epad, FEAT, HID, _ptr(ws.X0), _ptr(net.WTb["enc"]), _ptr(net.b["enc"]), None,
                             _ptr(ws.Z[0]), _ptr(ws.H[1]), 1, 0), "encoder")
    else:
        _ck(L.gnn_feat_enc(s, g.E, g.Epad, _ptr(g.ea), _ptr(g.eb), _ptr(g.nx), _ptr(g.ny), _ptr(g.len),
                           _ptr(g.closed), _ptr(w), _ptr(uh), _ptr(vh), _ptr(z), *sc, _ptr(net.WTb["enc"]),
                           _ptr(net.b["enc"]), GRAVITY, H_DRY, net.scales["flux"], P, None, None, _ptr(ws.H[1])),
            "features + encoder")
    for l in (1, 2, 3):
        name = "mp%d" % l
        if save:
            # the concatenated inputs [H, S_ea + S_eb] (node sums rounded to
            # BF16) are the weight-gradient operand of the backward pass
            _ck(L.gnn_node_sum(s, g.N, _ptr(g.tri_edges), _ptr(ws.H[l]), _ptr(ws.S)), "node sum")
            _ck(L.gnn_concat(s, g.E, g.Epad, _ptr(g.ea), _ptr(g.eb), _ptr(ws.H[l]), _ptr(ws.S), _ptr(ws.Cc[l])),
                "concat")
        last = l == 3 and not save  # inference: the head in the last layer's epilogue, H4 not stored
        _ck(L.gnn_mp_direct(s, g.E, g.Epad, _ptr(g.nbr6), _ptr(ws.H[l]), _ptr(net.WTb[name]), _ptr(net.b[name]),
                            _ptr(ws.Z[l]) if save else None, None if last else _ptr(ws.H[l + 1]),
                            _ptr(net.WTb["head"]) if last else None, _ptr(net.b["head"]) if last else None,
                            P if last else None, _ptr(ws.F) if last else None), name + (" + head" if last else ""))
    if save:
        _ck(L.gnn_head(s, g.Epad, _ptr(ws.H[4]), _ptr(net.WTb["head"]), _ptr(net.b["head"]), P, _ptr(ws.F)),
            "head")
    return ws.F


def apply(net, graph, ws, delta, w0, uh0, vh0, w1, uh1, vh1, targets=None, theta=None):
    """State after one learned step (FP64); with targets (true changes), also
    the loss gradient in ws.gq and loss terms in ws.loss_terms.  ``theta``:
    the positivity limiter's factors (limit())."""
    L, s, g = net.lib, net.stream(), graph
    dw, duh, dvh = targets if targets is not None else (None, None, None)
    _ck(L.gnn_apply(s, g.N, _ptr(g.slots), _ptr(g.sny), _ptr(g.sflags), _ptr(g.area), _ptr(g.group),
                    _ptr(g.group_area), _ptr(delta), _ptr(g.graph_of), _ptr(ws.F), net.scales["flux"], _ptr(theta), _ptr(w0), _ptr(uh0),
                    _ptr(vh0), _ptr(w1), _ptr(uh1), _ptr(vh1), _ptr(dw), _ptr(duh), _ptr(dvh), _ptr(net.sigma),
                    graph.mean_area, _ptr(ws.gq), _ptr(ws.loss_terms), _ptr(ws.cells)), "apply")


def backward(net, graph, ws, delta):
    """Weight gradients (ws.gW, ws.gb) from ws.gq."""
    L, s, g = net.lib, net.stream(), graph
    _ck(L.gnn_apply_bwd(s, g.E, g.Epad, _ptr(g.rec), _ptr(g.area), _ptr(g.group), _ptr(g.group_area), _ptr(delta),
                        _ptr(g.graph_of), net.scales["flux"], _ptr(ws.gq), _ptr(ws.dF)),
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


# ---------------------------------------------------------------------------
# FV trajectories, training, rollout
def training_domain(mesh, seed, extent, base_stage=5.0, friction=0.025):
    """A configuration-2-style scenario on ``mesh`` (a window of side
    ``extent`` with the workload's point density): seeded fBm terrain of
    amplitude 20 m, stage max(z, base_stage) plus a seeded circular mound
    (height 1..4 m, radius 0.1..0.35 extent), Manning n, transmissive
    boundaries."""
    from .domain import Domain, Transmissive_boundary
    from .scenarios import fbm

    rng = np.random.default_rng(seed)
    cx, cy = rng.uniform(0.3, 0.7, 2) * extent
    rad = rng.uniform(0.1, 0.35) * extent
    height = rng.uniform(1.0, 4.0)
    d = Domain(mesh=mesh)

    def bed(x, y):
        return fbm(x, y, seed=int(seed), base_wavelength=5000.0)

    def stage(x, y):
        r2 = (x - cx) ** 2 + (y - cy) ** 2
        return np.maximum(bed(x, y), base_stage) + np.where(r2 < rad * rad, height, 0.0)

    d.set_quantity("elevation", bed)
    d.set_quantity("stage", stage)
    d.set_quantity("friction", friction, location="centroids")
    d.set_boundary({tag: Transmissive_boundary(d) for tag in mesh.get_tag_names()})
    return d


def fv_trajectories(meshes, per_mesh, samples, delta, horizon=5.0, warmup=1, extent=1400.0, seed=0, log=None,
                    first_index=0):
    """FV training pairs: for each mesh, ``per_mesh`` scenarios run together as
    one ensemble, yielding every ``delta`` seconds up to ``horizon`` seconds.
    ``samples`` pairs (the state at a yield and at the next one) are recorded
    per scenario, their first yields spread evenly over [warmup, horizon -
    delta], so the pairs cover the whole time span a rollout of that length
    sees.  Trajectories with non-finite or implausible values (depth > 100 m,
    |momentum| > 200 m^2/s) are dropped.  Returns one dict per mesh: z [T, N],
    states [T, 2 * samples, 3, N] (pair j at rows 2j and 2j + 1; stage,
    xmomentum, ymomentum; FP64, on the device), delta [T]."""
    torch = _torch()
    from . import _lib
    from .ensemble import Ensemble

    nyields = max(int(round(horizon / delta)), warmup + 1)
    starts = np.unique(np.linspace(warmup, nyields - 1, samples).round().astype(int))
    record = {}
    for j, y in enumerate(starts):
        record.setdefault(int(y), []).append(2 * j)
        record.setdefault(int(y) + 1, []).append(2 * j + 1)
    S1 = 2 * len(starts)
    data = []
    for mi, mesh in enumerate(meshes, start=first_index):
        N = mesh.number_of_elements
        doms = [training_domain(mesh, seed * 100000 + mi * 1000 + t, extent) for t in range(per_mesh)]
        ens = Ensemble(doms)
        z = torch.empty(per_mesh, N, dtype=torch.float64, device="cuda")
        states = torch.empty(per_mesh, S1, 3, N, dtype=torch.float64, device="cuda")
        ok = True
        try:
            ens.prepare(delta, None)
            torch.cuda.synchronize()
            for t, d in enumerate(doms):
                d._handle.download_device(_lib.SW_ELEVATION, _lib.SW_CENTROID, z[t].data_ptr(), N)
            for y in range(max(record) + 1):
                for row in record.get(y, ()):
                    for t, d in enumerate(doms):
                        for c, q in enumerate((_lib.SW_STAGE, _lib.SW_XMOM, _lib.SW_YMOM)):
                            d._handle.download_device(q, _lib.SW_CENTROID, states[t, row, c].data_ptr(), N)
                if y == max(record):
                    break
                ens.evolve_steps(1 << 30)
                ens.handle.call("sw_advance_yield")
        except _lib.SwError:
            ok = False  # an unstable member stops its ensemble: this mesh's batch is dropped
        ens.close()
        if not ok:
            if log:
                log("mesh %d: dropped (a member became unstable)" % mi)
            continue
        h = states[:, :, 0, :] - z[:, None, :]
        good = (torch.isfinite(states).flatten(1).all(1) & (h.flatten(1).amax(1) < 100.0)
                & (states[:, :, 1:, :].abs().flatten(1).amax(1) < 200.0))
        keep = torch.nonzero(good).flatten()
        if keep.numel() == 0:
            continue
        data.append(dict(z=z[keep].contiguous(), states=states[keep].contiguous(),
                         delta=torch.full((keep.numel(),), float(delta), dtype=torch.float64, device="cuda"),
                         mesh=mesh, pair_start_times=starts * float(delta)))
        if log:
            log("mesh %d: %d triangles, %d of %d trajectories kept x %d pairs (t = %.2f .. %.2f s)" % (
                mi, N, keep.numel(), per_mesh, len(starts), starts[0] * delta, starts[-1] * delta))
    return data


def fit_scales(net, data):
    """Feature scales, flux scale and loss normalisation from the data."""
    torch = _torch()
    h, q, dq = [], [], []
    for d in data:
        st = d["states"]
        h.append((st[:, :, 0, :] - d["z"][:, None, :]).clamp(min=0).flatten())
        q.append(st[:, :, 1:, :].flatten())
        dq.append((st[:, 1::2] - st[:, 0::2]).transpose(0, 2).reshape(3, -1))
    h = torch.cat(h)
    q = torch.cat(q)
    dq = torch.cat(dq, dim=1)
    hs = float(torch.sqrt(torch.mean(h * h)))
    qs = float(torch.sqrt(torch.mean(q * q)))
    net.scales.update(depth=max(hs, 1e-3), momentum=max(qs, 1e-3), length=float(np.mean(
        [float(np.mean(d["mesh"].edgelengths)) for d in data])), flux=max(qs, 1e-3))
    sig = torch.sqrt(torch.mean(dq * dq, dim=1)).float().clamp(min=1e-12)
    net.sigma = sig.contiguous()


class Trainer:
    """Adam on minibatches of B samples of one mesh (its graph replicated B times)."""

    def __init__(self, net, data, batch=16, lr=1e-3, seed=0):
        torch = _torch()
        self.net, self.data, self.B, self.lr = net, data, batch, lr
        self.rng = np.random.default_rng(seed)
        self.graphs = [EdgeGraph(d["mesh"], copies=batch, closed_tags=(), merge_radius=net.scales["merge_radius"])
                       for d in data]
        self.ws = [Workspace(g) for g in self.graphs]
        self.buf = [dict((k, torch.empty(g.N, dtype=torch.float64, device="cuda"))
                         for k in ("w", "uh", "vh", "z", "dw", "duh", "dvh", "w1", "uh1", "vh1"))
                    for g in self.graphs]
        self.delta = [torch.empty(batch, dtype=torch.float64, device="cuda") for _ in self.graphs]

    def _fill(self, m, pick):
        d, b = self.data[m], self.buf[m]
        T, S1 = d["states"].shape[:2]
        tr = torch_index(pick[:, 0])
        sm = torch_index(2 * pick[:, 1])  # pair j: rows 2j (start) and 2j + 1 (one step later)
        s0 = d["states"][tr, sm]          # [B, 3, N]
        s1 = d["states"][tr, sm + 1]
        b["w"].copy_(s0[:, 0].reshape(-1))
        b["uh"].copy_(s0[:, 1].reshape(-1))
        b["vh"].copy_(s0[:, 2].reshape(-1))
        b["z"].copy_(d["z"][tr].reshape(-1))
        b["dw"].copy_((s1[:, 0] - s0[:, 0]).reshape(-1))
        b["duh"].copy_((s1[:, 1] - s0[:, 1]).reshape(-1))
        b["dvh"].copy_((s1[:, 2] - s0[:, 2]).reshape(-1))
        self.delta[m].copy_(d["delta"][tr])

    def loss(self, m, pick, train):
        g, ws, b = self.graphs[m], self.ws[m], self.buf[m]
        self._fill(m, pick)
        forward(self.net, g, ws, b["w"], b["uh"], b["vh"], b["z"])
        apply(self.net, g, ws, self.delta[m], b["w"], b["uh"], b["vh"], b["w1"], b["uh1"], b["vh1"],
              targets=(b["dw"], b["duh"], b["dvh"]))
        if train:
            backward(self.net, g, ws, self.delta[m])
            adam_step(self.net, ws, lr=self.lr)
        return float(ws.loss_terms.sum()) / g.N

    def sample(self, m, holdout=False):
        """B (trajectory, pair) picks.  The scenarios of every fourth
        trajectory index are held out for evaluation (all their pairs)."""
        d = self.data[m]
        T, S1 = d["states"].shape[:2]
        P = S1 // 2
        held = np.arange(T) % 4 == 3
        pool = np.nonzero(held if (holdout and held.any()) else ~held if (~held).any() else np.ones(T, bool))[0]
        tr = pool[self.rng.integers(0, len(pool), self.B)]
        sm = self.rng.integers(0, P, self.B)
        return np.stack([tr, sm], 1)

    def train(self, iters, log=None, every=100, cosine=True):
        """``iters`` Adam steps; the learning rate decays from lr to lr / 10
        along a half cosine."""
        hist = []
        lr0 = self.lr
        for it in range(iters):
            if cosine:
                self.lr = lr0 * (0.1 + 0.9 * 0.5 * (1.0 + np.cos(np.pi * it / max(iters, 1))))
            m = int(self.rng.integers(0, len(self.data)))
            l = self.loss(m, self.sample(m), train=True)
            hist.append(l)
            if log and (it % every == 0 or it == iters - 1):
                log("iter %d loss %.4g (mean of last %d: %.4g)" % (it, l, min(it + 1, every),
                                                                   float(np.mean(hist[-every:]))))
        self.lr = lr0
        return hist

    def evaluate(self, batches=8, seed=12345, holdout=True):
        """Mean loss (1.0 = no better than leaving the state unchanged for
        unit-variance changes) over ``batches`` batches of held-out pairs (or
        of training pairs), the same pairs for the same seed."""
        rng, self.rng = self.rng, np.random.default_rng(seed)
        ls = []
        for _ in range(batches):
            m = int(self.rng.integers(0, len(self.data)))
            ls.append(self.loss(m, self.sample(m, holdout=holdout), train=False))
        self.rng = rng
        return float(np.mean(ls))


def torch_index(a):
    import torch

    return torch.as_tensor(np.asarray(a, dtype=np.int64), device="cuda")


def limit(net, graph, ws, delta, w, z):
    """Positivity limiter factors of the current edge outputs (ws.theta)."""
    g = graph
    _ck(net.lib.gnn_limit(net.stream(), g.N, _ptr(g.slots), _ptr(g.sflags), _ptr(g.area), _ptr(g.group),
                          _ptr(delta), _ptr(g.graph_of), _ptr(ws.F), net.scales["flux"], _ptr(w), _ptr(z),
                          _ptr(ws.theta), _ptr(ws.cells)), "limit")
    return ws.theta


def protect(net, graph, w, uh, vh, z, h_dry=1e-3, froude_max=None):
    """Zero momentum where the depth is below h_dry; elsewhere speeds capped
    at froude_max sqrt(g h) (FROUDE_MAX unless given; 0: no cap)."""
    fr = FROUDE_MAX if froude_max is None else froude_max
    _ck(net.lib.gnn_protect(net.stream(), graph.N, _ptr(w), _ptr(z), _ptr(uh), _ptr(vh), float(h_dry), GRAVITY,
                            float(fr)), "protect")


def learned_step(net, graph, ws, delta, w, uh, vh, z, h_dry=1e-3):
    """One learned step in place (inference): edge outputs, positivity
    limiter, update, then zero momentum where the depth is below h_dry and
    speeds capped (protect)."""
    forward(net, graph, ws, w, uh, vh, z, save=False)
    theta = limit(net, graph, ws, delta, w, z)
    apply(net, graph, ws, delta, w, uh, vh, w, uh, vh, theta=theta)
    if h_dry > 0.0:
        protect(net, graph, w, uh, vh, z, h_dry)


def mass_balance(net, graph, ws, delta, w, uh, vh, z, steps, h_dry=H_DRY):
    """Mass balance of ``steps`` learned inference steps from copies of (w,
    uh, vh): each step's change of the total volume sum(A (w - z)) against
    the volume its open boundary edges let out (delta times their limited
    mass fluxes).  Returns the largest |change + outflow| / volume."""
    torch = _torch()
    g = graph
    rec = g.rec[:g.E]
    open_b = (rec[:, 1] < 0) & (rec[:, 5] == 0)
    ea_b = rec[open_b, 0].long()
    len_b = rec[open_b, 4].contiguous().view(torch.float32).double()
    fs = net.scales["flux"]
    d = float(delta)
    dl = torch.tensor([d], dtype=torch.float64, device="cuda")
    w, uh, vh = w.clone(), uh.clone(), vh.clone()
    worst = 0.0
    for _ in range(steps):
        m0 = float(torch.sum(g.area * (w - z)))
        forward(net, g, ws, w, uh, vh, z, save=False)
        theta = limit(net, g, ws, dl, w, z)
        c = ws.F[:g.E][open_b, 0].double() * len_b * fs
        c = torch.where(c > 0, c * theta[ea_b], c)  # outflow is limited by the cell it drains
        out = d * float(torch.sum(c))
        apply(net, g, ws, dl, w, uh, vh, w, uh, vh, theta=theta)
        if h_dry > 0.0:
            protect(net, g, w, uh, vh, z, h_dry)
        m1 = float(torch.sum(g.area * (w - z)))
        worst = max(worst, abs(m1 - m0 + out) / abs(m0))
    return worst


def evaluate_rollout(net, domain, steps, delta, judge_every=8, threshold=0.01, judge_precision="fp64", log=None):
    """Learned rollout against the FP64 FV path on ``domain`` (an FV Domain,
    initial state set), in steps of ``delta`` seconds.

    1. Reference: the FV loop with yields every delta, ``steps`` yields; the
       stage at every yield is kept (device copies) and the FV loop is timed.
    2. Learned: from the same initial state, ``steps`` learned steps.  At the
       end of every window of ``judge_every`` steps the FV loop advances the
       state of the window's last step from the same starting point (the
       judge) and its result is kept.  If the learned step's stage L2 error
       against it (area-weighted, relative to the depth norm) exceeds
       ``threshold``, the judge fails or needs more than twice the reference's
       mean FV steps per step (a damaged state: spurious speeds), the whole
       window is recomputed with the FV loop from the window's start and the
       FV loop also runs the next window (fallback).  Should the FV loop fail
       from a window start, it restarts from the window start before.
    ``judge_precision`` 'mixed' runs the judge's FV step with the
    mixed-precision loop (FP32 flux and extrapolation, FP64 accumulation,
    timestep and depth tests: about 1e-7 relative of the FP64 loop, five
    orders below a 1% threshold); the reference trajectory and every
    fallback recomputation run in FP64.
    Returns timings (with a breakdown into learned steps, judges,
    recomputation and the rest), per-step stage errors against the
    reference, the mass change against the reference and the judge's
    record."""
    import time

    torch = _torch()
    from . import _lib

    N = domain.number_of_elements
    h = domain._dev()
    if judge_precision not in ("fp64", "mixed"):
        raise ValueError("judge_precision must be 'fp64' or 'mixed'")
    p64 = _lib.SwParams.from_buffer_copy(h.params)
    p64.precision = _lib.SW_PRECISION_FP64
    pjudge = _lib.SwParams.from_buffer_copy(p64)
    if judge_precision == "mixed":
        pjudge.precision = _lib.SW_PRECISION_MIXED

    def use(p):
        if h.params.precision != p.precision:
            h.set_params(_lib.SwParams.from_buffer_copy(p))

    use(p64)
    h.call("sw_evolve_prepare", 1e9, 0.0, 0)
    dev = lambda: torch.empty(N, dtype=torch.float64, device="cuda")
    w0, uh0, vh0, z = dev(), dev(), dev(), dev()
    for q, buf in ((_lib.SW_STAGE, w0), (_lib.SW_XMOM, uh0), (_lib.SW_YMOM, vh0), (_lib.SW_ELEVATION, z)):
        h.download_device(q, _lib.SW_CENTROID, buf.data_ptr(), N)
    t0 = h.get_state().time
    area = torch.tensor(domain.mesh.areas, dtype=torch.float64, device="cuda")
    # Vertex values of the initial state.  The reconstruction keeps the vertex
    # values of cells in dry neighbourhoods from the previous step, so every
    # restart puts these back first: a failed FV run must not leave its
    # values there, and cells that stay dry then match the reference run.
    vert0 = {}
    for q in (_lib.SW_STAGE, _lib.SW_XMOM, _lib.SW_YMOM):
        vert0[q] = torch.empty(3 * N, dtype=torch.float64, device="cuda")
        h.download_device(q, _lib.SW_VERTEX, vert0[q].data_ptr(), 3 * N)

    def wl2(x):
        """Area-weighted L2 norm over the domain (sqrt of the integral of x^2)."""
        return float(torch.sqrt(torch.sum(area * x * x)))

    def set_state(w, uh, vh, t):
        torch.cuda.synchronize()
        for q, buf in vert0.items():
            h.upload_device(q, _lib.SW_VERTEX, buf.data_ptr(), 3 * N)
        for q, buf in ((_lib.SW_STAGE, w), (_lib.SW_XMOM, uh), (_lib.SW_YMOM, vh)):
            h.upload_device(q, _lib.SW_CENTROID, buf.data_ptr(), N)
        s = h.get_state(check=False)  # a failed judge leaves its error latched
        s.time = t
        s.starttime = t
        s.error = 0
        s.event = 0
        h.set_state(s)
        h.call("sw_evolve_prepare", delta, 0.0, 0)

    def fv_advance(w, uh, vh, t, nyields, judge=False):
        """FV over nyields yields of delta from (w, uh, vh) at time t; result
        in place (the judge's precision if ``judge``, else FP64)."""
        use(pjudge if judge else p64)
        set_state(w, uh, vh, t)
        total = 0
        for k in range(nyields):
            done, ev = h.evolve_steps(1 << 30)
            total += done
            if k + 1 < nyields:
                h.call("sw_advance_yield")
        for q, buf in ((_lib.SW_STAGE, w), (_lib.SW_XMOM, uh), (_lib.SW_YMOM, vh)):
            h.download_device(q, _lib.SW_CENTROID, buf.data_ptr(), N)
        return total

    # 1. reference
    ref_w = torch.empty(steps, N, dtype=torch.float64, device="cuda")
    set_state(w0, uh0, vh0, t0)
    torch.cuda.synchronize()
    tic = time.perf_counter()
    fv_steps = 0
    for k in range(steps):
        done, ev = h.evolve_steps(1 << 30)
        fv_steps += done
        h.download_device(_lib.SW_STAGE, _lib.SW_CENTROID, ref_w[k].data_ptr(), N)
        if k + 1 < steps:
            h.call("sw_advance_yield")
    h.call("sw_synchronize")
    fv_time = time.perf_counter() - tic
    ref_mass = float(torch.sum(area * (ref_w[-1] - z)))
    mass0 = float(torch.sum(area * (w0 - z)))
    if log:
        log("FV reference: %d yields of %.4g s = %d FV steps in %.3f s (%.3e triangle-steps/s)" % (
            steps, delta, fv_steps, fv_time, N * fv_steps / fv_time))

    # 2. learned rollout with the FV judge
    graph = EdgeGraph(domain.mesh, copies=1, closed_tags=(), merge_radius=net.scales["merge_radius"])
    ws = Workspace(graph)
    dl = torch.tensor([delta], dtype=torch.float64, device="cuda")
    # cost of one learned step on this domain (events on the stream the
    # kernels run on), from a scratch copy of the initial state
    sw, suh, svh = w0.clone(), uh0.clone(), vh0.clone()
    learned_step(net, graph, ws, dl, sw, suh, svh, z)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(10):
        learned_step(net, graph, ws, dl, sw, suh, svh, z)
    ev1.record()
    torch.cuda.synchronize()
    step_ms = ev0.elapsed_time(ev1) / 10.0
    del sw, suh, svh
    balance = mass_balance(net, graph, ws, delta, w0, uh0, vh0, z, judge_every)
    if log:
        log("learned step: %.3f ms (%d edges, %d of %d triangles in merged groups); FV: %.3f ms per %g s; mass "
            "balance over %d steps (volume change + open-boundary outflow) / volume <= %.2e" % (
                step_ms, graph.E, graph.merged, N, 1000.0 * fv_time / steps, delta, judge_every, balance))
    errors, judged = [], []
    fv_windows = 0
    fv_in_rollout = 0
    fallback = False
    health_steps = 2.0 * fv_steps / steps  # a judge needing more FV steps than this saw a damaged state
    history = [(0, (w0, uh0, vh0))]  # window starts, oldest first
    spent = dict(judge=0.0, recompute=0.0)
    windows = [(k0, min(judge_every, steps - k0)) for k0 in range(0, steps, judge_every)]

    def now():
        torch.cuda.synchronize()
        return time.perf_counter()

    def recompute(k_end):
        """FV (FP64) from the latest window start that the FV loop can
        advance to k_end (older starts are tried if a newer one fails);
        returns its state at k_end and the FV steps taken."""
        nonlocal history
        while history:
            k0, st = history[-1]
            cur = (st[0].clone(), st[1].clone(), st[2].clone())
            try:
                return cur, fv_advance(cur[0], cur[1], cur[2], t0 + k0 * delta, k_end - k0)
            except _lib.SwError:
                if len(history) == 1:
                    raise
                history = history[:-1]
        raise RuntimeError("no window start")

    def enqueue(state, n):
        """n learned steps from ``state`` (advanced in place), queued on the
        stream without waiting; returns the state, a copy of it before the
        last step and an event after the last step."""
        w, uh, vh = state
        for _ in range(n - 1):
            learned_step(net, graph, ws, dl, w, uh, vh, z)
        last = (w.clone(), uh.clone(), vh.clone())
        learned_step(net, graph, ws, dl, w, uh, vh, z)
        done = torch.cuda.Event()
        done.record()
        return state, last, done

    def judge(last, t_last):
        """One FV yield of delta from the window's last start (in place, the
        judge's precision): restart from device buffers, no vertex restore."""
        use(pjudge)
        h.restart_device(last[0].data_ptr(), last[1].data_ptr(), last[2].data_ptr(), t_last, delta)
        done, _ = h.evolve_steps(1 << 30)
        h.download_conserved_device(last[0].data_ptr(), last[1].data_ptr(), last[2].data_ptr())
        return done

    tic = now()
    queued = enqueue(tuple(x.clone() for x in (w0, uh0, vh0)), windows[0][1])
    for i, (k0, n) in enumerate(windows):
        k_end = k0 + n
        if fallback:
            t1 = time.perf_counter()
            state, done = recompute(k_end)
            spent["recompute"] += time.perf_counter() - t1
            fv_in_rollout += done
            fv_windows += 1
            fallback = False
            if i + 1 < len(windows):
                queued = enqueue(tuple(x.clone() for x in state), windows[i + 1][1])
        else:
            state, last, done_ev = queued
            # the next window's learned steps go on the stream first: the GPU
            # runs them while this window's judge is set up and run
            nxt = enqueue(tuple(x.clone() for x in state), windows[i + 1][1]) if i + 1 < len(windows) else None
            done_ev.synchronize()
            t1 = time.perf_counter()
            try:
                done = judge(last, t0 + (k_end - 1) * delta)
                fv_in_rollout += done
                d, b = state[0] - last[0], last[0] - z
                num, den = torch.stack([torch.sum(area * d * d), torch.sum(area * b * b)]).tolist()
                err = float(np.sqrt(num) / np.sqrt(den))
                # a damaged state: spurious speeds (too many FV steps) or a
                # non-finite value anywhere (NaN compares false with the bar)
                if done > health_steps or not np.isfinite(err):
                    err = float("inf")
            except _lib.SwError:
                done, err = -1, float("inf")
            spent["judge"] += time.perf_counter() - t1
            judged.append((k_end - 1, err))
            if log:
                log("window ending at step %d: judge error %.3g (%d FV steps) -> %s" % (
                    k_end - 1, err, done, "recomputed by FV" if err > threshold else "accepted"))
            if err > threshold:
                torch.cuda.synchronize()  # the queued window is discarded
                t1 = time.perf_counter()
                state, done = recompute(k_end)
                spent["recompute"] += time.perf_counter() - t1
                fv_in_rollout += done
                fv_windows += 1
                fallback = True
            else:
                queued = nxt
        history.append((k_end, state))
        errors.append((k_end - 1, state[0] - ref_w[k_end - 1]))
    w = history[-1][1][0]
    learned_time = now() - tic
    use(p64)
    err_l2 = [(i, wl2(e) / wl2(ref_w[i] - z)) for i, e in errors]
    mass_final = float(torch.sum(area * (w - z)))
    out = dict(triangles=N, steps=steps, delta=delta, fv_steps=fv_steps, fv_seconds=fv_time,
               fv_rate=N * fv_steps / fv_time, learned_seconds=learned_time,
               learned_rate=N * fv_steps / learned_time, speedup=fv_time / learned_time,
               stage_rel_l2=err_l2, max_stage_rel_l2=max(e for _, e in err_l2), final_stage_rel_l2=err_l2[-1][1],
               mass_change_vs_fv=abs(mass_final - ref_mass) / abs(mass0), judged=judged,
               fallback_windows=fv_windows, fv_steps_in_rollout=fv_in_rollout, learned_step_ms=step_ms,
               fv_ms_per_learned_step=1000.0 * fv_time / steps, merged_triangles=graph.merged,
               judge_precision=judge_precision, mass_balance_rel=balance,
               seconds_breakdown=dict(spent, learned_and_other=learned_time - sum(spent.values())))
    if log:
        log("learned (%s judge): %d steps in %.3f s (%.3e FV-equivalent triangle-steps/s, %.2fx FV); stage rel "
            "L2 at window ends max %.3g final %.3g; %d windows judged, %d recomputed by FV; seconds: %s" % (
                judge_precision, steps, learned_time, out["learned_rate"], out["speedup"], out["max_stage_rel_l2"],
                out["final_stage_rel_l2"], len(judged), fv_windows,
                {k: round(v, 4) for k, v in out["seconds_breakdown"].items()}))
    # leave the domain in its initial state
    set_state(w0, uh0, vh0, t0)
    return out


def training_meshes(count, extent=1400.0, density=5000.0, seed=7, first=0):
    """Seeded windows of the configuration-2 mesh family: Delaunay meshes of
    uniform random points (``density`` points per km^2, as 500k points in
    10 km x 10 km) on squares of side ``extent``; window i has seed
    seed * 1000 + i, i = first .. first + count - 1."""
    from .mesh import Mesh, delaunay_random

    npts = int(density * (extent / 1000.0) ** 2)
    return [Mesh(*delaunay_random(npts, extent=(extent, extent), seed=seed * 1000 + i))
            for i in range(first, first + count)]


def run_learned_workload(train_meshes=40, per_mesh=50, samples=8, delta=0.1, iters=3000, batch=16, lr=1e-3,
                         eval_points=500000, eval_steps=100, judge_every=8, threshold=0.01, extent=1400.0,
                         merge_radius=2.5, horizon=None, save_net=None, load_net=None, judge_precision="fp64",
                         log=None, train_trajectories=0, max_train_meshes=400, both_judges=True):
    """Train on FV trajectories of configuration-2-style windows, then roll the
    network out on the configuration-2 workload (1M triangles) against the
    FP64 FV path.  Training windows are taken in order, ``train_meshes`` at a
    time, until at least ``train_trajectories`` trajectories are kept (a
    window whose batch of ``per_mesh`` scenarios has a member the FV scheme
    cannot advance -- sliver triangles of random points -- is dropped whole).
    Returns a dict of the results."""
    import time

    from . import scenarios

    if load_net:
        # a network trained earlier: its scales, merge radius and learned step
        # come with it
        net = EdgeGNN.load(load_net)
        if net.scales["delta"] > 0.0:
            delta = net.scales["delta"]
        if log:
            log("network loaded from %s (no training in this run; learned step %g s)" % (load_net, delta))
    horizon = eval_steps * delta if horizon is None else horizon
    train = dict(train_pairs=samples, train_horizon=horizon, train_iters=iters)
    if load_net:
        train.update(loaded_network=os.path.basename(load_net))
    else:
        t0 = time.time()
        data, first = [], 0
        while first == 0 or (sum(d["states"].shape[0] for d in data) < train_trajectories
                             and first < max_train_meshes):
            meshes = training_meshes(train_meshes, extent=extent, first=first)
            data += fv_trajectories(meshes, per_mesh=per_mesh, samples=samples, delta=delta, horizon=horizon,
                                    extent=extent, log=log, first_index=first)
            first += train_meshes
        n_traj = sum(d["states"].shape[0] for d in data)
        t_data = time.time() - t0
        if log:
            log("training data: %d trajectories x %d pairs on %d of %d windows of %d triangles (%.1f s)" % (
                n_traj, samples, len(data), first, meshes[0].number_of_elements, t_data))
        net = EdgeGNN(seed=1, merge_radius=merge_radius, delta=delta)
        fit_scales(net, data)
        tr = Trainer(net, data, batch=batch, lr=lr)
        untrained = tr.evaluate(16)
        t0 = time.time()
        hist = tr.train(iters, log=log, every=max(iters // 6, 1))
        _torch().cuda.synchronize()
        t_train = time.time() - t0
        trained = tr.evaluate(16)
        if save_net:
            net.save(save_net)
        if log:
            log("held-out one-step loss: untrained (first-order scheme) %.4f, trained %.4f (%d iterations, %.1f s)"
                % (untrained, trained, iters, t_train))
        train.update(train_trajectories=n_traj, train_windows=first, train_seconds=t_train, data_seconds=t_data,
                     holdout_loss_untrained=untrained, holdout_loss_trained=trained,
                     train_loss_last=float(np.mean(hist[-50:])))
        del data, tr
    dom = scenarios.delaunay_terrain(npoints=eval_points)
    res = evaluate_rollout(net, dom, eval_steps, delta, judge_every=judge_every, threshold=threshold,
                           judge_precision=judge_precision, log=log)
    if both_judges:
        other = "fp64" if judge_precision == "mixed" else "mixed"
        alt = evaluate_rollout(net, dom, eval_steps, delta, judge_every=judge_every, threshold=threshold,
                               judge_precision=other, log=log)
        res["with_%s_judge" % other] = {k: alt[k] for k in (
            "speedup", "learned_rate", "fv_rate", "final_stage_rel_l2", "max_stage_rel_l2", "fallback_windows",
            "seconds_breakdown")}
    dom.close()
    res.update(train)
    return res


def learned_kernel_bytes(N, E):
    """Algorithmic HBM bytes per launch of the inference kernels on N
    triangles and E edges: every array a kernel reads or writes, counted once
    (rows gathered from neighbours as cache hits; of a packed record or an
    output row only the fields used).  The three message-passing launches
    are averaged (the third reads the physical part and writes F instead of
    the hidden rows)."""
    return {
        # ea, eb, nx, ny, len, closed; state (w, uh, vh, z) once per triangle;
        # writes H1 (64 BF16) and the physical part (8 FP32)
        "k_feat_enc": E * (8 + 12 + 1 + 128 + 32) + N * 32,
        # own row 128 + neighbour list 32, writes the new row 128; third layer:
        # + physical part 32, writes F 32 instead of the row
        "k_mp_direct": E * (2 * (128 + 32 + 128) + (128 + 32 + 32 + 32)) // 3,
        # slot flags 4, per slot the edge, the other triangle and the length
        # 3x12, group row 16, area, w, z 24, theta 8, graph_of 4; per edge the
        # mass flux 4
        "k_limit": N * (4 + 36 + 16 + 24 + 8 + 4) + E * 4,
        # slot flags 4, slot records 3x16 and ny 3x4, group 16, areas 16,
        # graph_of 4, theta 8, state in 24 and out 24; per edge five outputs 20
        "k_apply": N * (4 + 48 + 12 + 16 + 16 + 4 + 8 + 24 + 24) + E * 20,
        # w and z read, both momenta written (at most: only where dry)
        "k_protect_dry": N * 32,
    }


def time_learned_step(points=500000, steps=20, delta=0.1, merge_radius=2.5):
    """Milliseconds per learned inference step on the seeded Delaunay workload
    of ``points`` points (an untrained network: the same kernels and cost as a
    trained one)."""
    torch = _torch()
    from . import _lib, scenarios

    dom = scenarios.delaunay_terrain(npoints=points)
    N = dom.number_of_elements
    h = dom._dev()
    buf = {q: torch.empty(N, dtype=torch.float64, device="cuda")
           for q in (_lib.SW_STAGE, _lib.SW_XMOM, _lib.SW_YMOM, _lib.SW_ELEVATION)}
    for q, b in buf.items():
        h.download_device(q, _lib.SW_CENTROID, b.data_ptr(), N)
    net = EdgeGNN(seed=1, merge_radius=merge_radius)
    graph = EdgeGraph(dom.mesh, closed_tags=(), merge_radius=merge_radius)
    ws = Workspace(graph)
    dl = torch.tensor([delta], dtype=torch.float64, device="cuda")
    w, uh, vh, z = (buf[q] for q in (_lib.SW_STAGE, _lib.SW_XMOM, _lib.SW_YMOM, _lib.SW_ELEVATION))
    learned_step(net, graph, ws, dl, w, uh, vh, z)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(steps):
        learned_step(net, graph, ws, dl, w, uh, vh, z)
    ev1.record()
    torch.cuda.synchronize()
    dom.close()
    return ev0.elapsed_time(ev1) / steps, N, graph.E


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser(description="time the learned inference step")
    ap.add_argument("--points", type=int, default=500000)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--merge-radius", type=float, default=2.5)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also run a plain device copy of MIB MiB (scale for rocprof FETCH/WRITE_SIZE)")
    a = ap.parse_args()
    if a.copy_calibration:
        torch = _torch()
        src = torch.ones(a.copy_calibration * (1 << 20) // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    ms, n, e = time_learned_step(a.points, a.steps, merge_radius=a.merge_radius)
    print("learned step: %.3f ms on %d triangles, %d edges" % (ms, n, e), flush=True)
    for k, v in learned_kernel_bytes(n, e).items():
        print("algorithmic bytes per launch: %s=%d" % (k, v))
