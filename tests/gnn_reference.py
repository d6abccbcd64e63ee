"""CPU (numpy, FP64) restatement of the learned edge-flux surrogate: forward
pass, state update, loss and gradients, written independently of the HIP
kernels so their BF16 MFMA results can be checked against it.  Test
infrastructure only."""
import numpy as np

FEAT, HID, CAT, OUT = 32, 64, 128, 32


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    b = f.view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


def _sigmoid(z):
    return 0.5 * (1.0 + np.tanh(0.5 * z))


def silu(z):
    return z * _sigmoid(z)


def silu_grad(z):
    s = _sigmoid(z)
    return s * (1.0 + z * (1.0 - s))


def features(g, w, uh, vh, z, hs, qs, ls):
    """Edge features (as k_features), BF16-rounded."""
    E = len(g["ea"])
    X = np.zeros((E, FEAT))
    for e in range(E):
        a, b = g["ea"][e], g["eb"][e]
        n1, n2 = float(g["nx"][e]), float(g["ny"][e])
        wall = g["closed"][e] != 0
        ha = w[a] - z[a]
        qna, qta = uh[a] * n1 + vh[a] * n2, -uh[a] * n2 + vh[a] * n1
        hb, qnb, qtb, dw, dz = ha, (-qna if wall else qna), qta, 0.0, 0.0
        if b >= 0:
            hb = w[b] - z[b]
            qnb, qtb = uh[b] * n1 + vh[b] * n2, -uh[b] * n2 + vh[b] * n1
            dw, dz = w[b] - w[a], z[b] - z[a]
        X[e, :15] = [ha / hs, hb / hs, qna / qs, qnb / qs, qta / qs, qtb / qs, dw / hs, dz / hs,
                     1.0 if b < 0 else 0.0, 1.0 if wall else 0.0, g["len"][e] / ls,
                     np.sqrt(max(ha, 0.0) / hs), np.sqrt(max(hb, 0.0) / hs),
                     1.0 if ha > 1e-3 else 0.0, 1.0 if hb > 1e-3 else 0.0]
        # physics prior: first-order central-upwind flux after hydrostatic reconstruction
        zm = max(z[a], z[b]) if b >= 0 else z[a]
        wb = w[b] if b >= 0 else w[a]
        sa, sb = max(0.0, w[a] - zm), max(0.0, wb - zm)
        G, h0 = 9.8, 1e-6
        ua, ub = qna * sa / (sa * sa + h0), qnb * sb / (sb * sb + h0)
        va, vb = qta * sa / (sa * sa + h0), qtb * sb / (sb * sb + h0)
        ca, cb = np.sqrt(G * sa), np.sqrt(G * sb)
        smax, smin = max(ua + ca, ub + cb, 0.0), min(ua - ca, ub - cb, 0.0)
        k = [0.0, 0.0, 0.0]
        if smax - smin > 1e-6:
            fa = [sa * ua, sa * ua * ua + 0.5 * G * sa * sa, sa * ua * va]
            fb = [sb * ub, sb * ub * ub + 0.5 * G * sb * sb, sb * ub * vb]
            du = [sb - sa, sb * ub - sa * ua, sb * vb - sa * va]
            k = [(smax * fa[i] - smin * fb[i] + smax * smin * du[i]) / (smax - smin) for i in range(3)]
        ps = 0.5 * G * hs * hs
        X[e, 15:20] = [k[0] / qs, k[1] / ps, k[2] / qs, 0.5 * G * (ha * ha - sa * sa) / ps,
                       0.5 * G * (hb * hb - sb * sb) / ps]
    return bf16(X)


def base_flux(g, w, uh, vh, z, fs, G=9.8, h_dry=1e-3):
    """Physical part of the edge outputs (as k_base_flux), [E][5] in the
    outputs' units: first-order central-upwind flux of the centroid states
    after hydrostatic reconstruction; each side's normal momentum without its
    own reconstructed pressure."""
    E = len(g["ea"])
    P = np.zeros((E, 5))
    for e in range(E):
        a, b = int(g["ea"][e]), int(g["eb"][e])
        n1, n2 = float(g["nx"][e]), float(g["ny"][e])
        ha = w[a] - z[a]
        qna, qta = uh[a] * n1 + vh[a] * n2, -uh[a] * n2 + vh[a] * n1
        wb, zb, hb, qnb, qtb = w[a], z[a], ha, (-qna if g["closed"][e] else qna), qta
        if b >= 0:
            wb, zb, hb = w[b], z[b], w[b] - z[b]
            qnb, qtb = uh[b] * n1 + vh[b] * n2, -uh[b] * n2 + vh[b] * n1
        ua, va = (qna / ha, qta / ha) if ha > h_dry else (0.0, 0.0)
        ub, vb = (qnb / hb, qtb / hb) if hb > h_dry else (0.0, 0.0)
        zm = max(z[a], zb)
        sa, sb = max(0.0, w[a] - zm), max(0.0, wb - zm)
        ca, cb = np.sqrt(G * sa), np.sqrt(G * sb)
        smax, smin = max(ua + ca, ub + cb, 0.0), min(ua - ca, ub - cb, 0.0)
        if smax - smin <= 1e-12:
            continue
        # KNP flux of U = (s, s u, s v) with F(U) = (s u, s u^2 + g s^2 / 2, s u v)
        Ua, Ub = np.array([sa, sa * ua, sa * va]), np.array([sb, sb * ub, sb * vb])
        Fa = np.array([sa * ua, sa * ua * ua + 0.5 * G * sa * sa, sa * ua * va])
        Fb = np.array([sb * ub, sb * ub * ub + 0.5 * G * sb * sb, sb * ub * vb])
        k = (smax * Fa - smin * Fb + smax * smin * (Ub - Ua)) / (smax - smin)
        P[e] = [k[0], -(k[1] - 0.5 * G * sa * sa), -k[2], k[1] - 0.5 * G * sb * sb, k[2]]
    return P / fs


def forward(params, g, X0):
    """Returns F and the saved activations."""
    W = {k: bf16(v) for k, v in params["W"].items()}
    b = params["b"]
    Z0 = X0 @ W["enc"] + b["enc"]
    H = {1: silu(Z0)}
    Z = {0: Z0}
    Cs = {}
    for l in (1, 2, 3):
        S = node_sum(g, H[l])
        Cs[l] = np.concatenate([H[l], concat_nodes(g, S)], axis=1)
        Z[l] = Cs[l] @ W["mp%d" % l] + b["mp%d" % l]
        H[l + 1] = H[l] + silu(Z[l])
    F = H[4] @ W["head"] + b["head"]
    return F, dict(W=W, Z=Z, H=H, C=Cs, X0=X0)


def node_sum(g, H):
    S = np.zeros((len(g["tri_edges"]), H.shape[1]))
    for i, es in enumerate(g["tri_edges"]):
        for e in es:
            if e >= 0:
                S[i] += H[e]
    return S


def concat_nodes(g, S):
    out = np.zeros((len(g["ea"]), S.shape[1]))
    for e, (a, b) in enumerate(zip(g["ea"], g["eb"])):
        out[e] = S[a] + (S[b] if b >= 0 else 0.0)
    return out


def apply_update(g, F, delta, fs):
    """State change dq[i] = (dh, duh, dvh) of the learned step.  With
    g["group"] (merged small cells), each triangle's integrated change is
    formed first and every member of a group takes the group's total divided
    by the group's area; the mass flux between two members of one group is
    left out."""
    N = len(g["tri_edges"])
    group = g.get("group")
    members = [set(int(m) for m in group[i] if m >= 0) for i in range(N)] if group is not None else None
    M = np.zeros((N, 3))
    for i, es in enumerate(g["tri_edges"]):
        for e in es:
            if e < 0:
                continue
            side_a = g["ea"][e] == i
            other = g["eb"][e] if side_a else g["ea"][e]
            inside = members is not None and other >= 0 and other in members[i] and len(members[i]) > 1
            L = float(g["len"][e]) * fs
            n1, n2 = float(g["nx"][e]), float(g["ny"][e])
            c = 0.0 if (g["closed"][e] or inside) else F[e, 0] * L
            M[i, 0] += -c if side_a else c
            mn, mt = (F[e, 1], F[e, 2]) if side_a else (F[e, 3], F[e, 4])
            M[i, 1] += (mn * n1 - mt * n2) * L
            M[i, 2] += (mn * n2 + mt * n1) * L
    dq = np.zeros((N, 3))
    for i in range(N):
        if members is not None and len(members[i]) > 1:
            dq[i] = delta / g["group_area"][i] * sum(M[m] for m in sorted(members[i]))
        else:
            dq[i] = delta / g["area"][i] * M[i]
    return dq


def apply_transpose(g, E, delta, fs, gq):
    """dF = J^T gq for the linear map F -> dq of apply_update, J taken column
    by column from apply_update itself (a unit output of one edge at a time)."""
    dF = np.zeros((E, OUT))
    for e in range(E):
        for c in range(5):
            unit = np.zeros((E, OUT))
            unit[e, c] = 1.0
            dF[e, c] = np.sum(apply_update(g, unit, delta, fs) * gq)
    return dF


def loss_and_grads(params, g, X0, delta, fs, target, sigma, base=None):
    """Loss = sum_i (A_i / mean A) sum_c huber((dq - target)/sigma) / N (Huber
    with threshold 3) and its weight gradients."""
    F, saved = forward(params, g, X0)
    if base is not None:  # the physical part (base_flux) does not depend on the weights
        F = F.copy()
        F[:, :5] += base
    N = len(g["tri_edges"])
    dq = apply_update(g, F, delta, fs)
    wgt = (g["area"] / np.mean(g["area"]))[:, None]
    err = (dq - target) / sigma
    k = 3.0
    ae = np.abs(err)
    loss = np.sum(wgt * np.where(ae <= k, err ** 2, k * (2 * ae - k))) / N
    gq = wgt * np.where(ae <= k, 2.0 * err, 2.0 * k * np.sign(err)) / sigma / N
    # d loss / d F through the (linear) update
    dF = apply_transpose(g, len(g["ea"]), delta, fs, gq)
    W, Z, H, Cs = saved["W"], saved["Z"], saved["H"], saved["C"]
    gW, gb = {}, {}
    gW["head"] = H[4].T @ dF
    gb["head"] = dF.sum(0)
    dH = dF @ W["head"].T
    for l in (3, 2, 1):
        name = "mp%d" % l
        dZ = dH * silu_grad(Z[l])
        gW[name] = Cs[l].T @ dZ
        gb[name] = dZ.sum(0)
        dC = dZ @ W[name].T
        dS = node_sum(g, dC[:, HID:])
        dH = dH + dC[:, :HID] + concat_nodes(g, dS)
    dZ = dH * silu_grad(Z[0])
    gW["enc"] = X0.T @ dZ
    gb["enc"] = dZ.sum(0)
    return loss, F, dq, gW, gb


def limiter(g, F, delta, fs, w, z):
    """Positivity limiter factors: per cell (lone cells only), volume over the
    mass it would send out through its open edges over the step, capped at 1."""
    N = len(g["tri_edges"])
    theta = np.ones(N)
    for i, es in enumerate(g["tri_edges"]):
        vol = g["area"][i] * max(w[i] - z[i], 0.0)
        out = 0.0
        for e in es:
            if e < 0 or g["closed"][e]:
                continue
            flux = F[e, 0] * float(g["len"][e]) * fs  # positive: from ea to eb
            out += max(flux, 0.0) if g["ea"][e] == i else max(-flux, 0.0)
        out *= delta
        if out > vol:
            theta[i] = vol / out
    return theta
