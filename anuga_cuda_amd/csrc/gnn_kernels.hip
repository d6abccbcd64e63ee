// Kernels of the learned edge-flux surrogate (declarations: include/anuga_gnn.h).
//
// Dense layers run on the gfx950 matrix cores: v_mfma_f32_16x16x32_bf16, one
// 16 x 16 output block per instruction with K = 32, BF16 operands and FP32
// accumulation.  Operand layout of that instruction (lane l, r = l % 16,
// g = l / 16): A[m = r][k = 8g + j], B[k = 8g + j][n = r], D[m = 4g + j][n = r],
// j = 0..7 for A and B, 0..3 for D.  One wavefront computes 16 rows of a layer.
// Activations and gradients are BF16 in HBM; weight gradients are reduced in
// FP32 (per-workgroup partials, then a deterministic sum); the state update is
// FP64.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <utility>

#include <cmath>

#include "../../include/anuga_gnn.h"

namespace gnn {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kB = 256;      // threads per workgroup
constexpr int kParts = 256;  // workgroups of the weight-gradient reduction
constexpr int kFeat = 32;    // edge input features (padded)
constexpr int kOut = 32;     // head weight columns and output gradient row (padded; 5 used)
constexpr int kFOut = 8;     // row of the FP32 edge outputs F and of the physical part (5 used)

#define GNN_DEV __device__ __forceinline__

// Logistic function with the hardware reciprocal (v_rcp_f32, 1 ulp) rather
// than an IEEE division: the layers' epilogues are bound by these
// instructions, and BF16 activations do not need the last bit.
GNN_DEV float sigmoid(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
GNN_DEV float silu(float z) { return z * sigmoid(z); }
GNN_DEV float silu_grad(float z) {
    const float s = sigmoid(z);
    return s * (1.0f + z * (1.0f - s));
}
// Workgroup b of nb renumbered so that each of the 8 XCDs (workgroups are
// dispatched to them round-robin) takes one contiguous range of the index
// space: the neighbour gathers of a range then hit that XCD's own L2.
GNN_DEV int64_t xcd_block(int64_t b, int64_t nb) {
    const int64_t per = nb >> 3, rem = nb & 7, x = b & 7, idx = b >> 3;
    return x * per + (x < rem ? x : rem) + idx;
}
// Global thread index of a one-thread-per-item kernel, in XCD-contiguous order.
GNN_DEV int64_t gidx() { return xcd_block(blockIdx.x, gridDim.x) * kB + threadIdx.x; }
// Tiles of a persistent kernel: wavefront w of the launch takes tiles w,
// w + nwaves, ...  All wavefronts sweep the array together, so the rows they
// gather from nearby edges are shared in the last-level cache while they are
// hot (measured faster than giving each workgroup a contiguous range).
GNN_DEV void wave_tiles(int64_t &t0, int64_t &step) {
    t0 = ((int64_t)blockIdx.x * kB + threadIdx.x) >> 6;
    step = (int64_t)gridDim.x * (kB / 64);
}

GNN_DEV f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
GNN_DEV bf16x8 ld8(const bf16 *p) { return *reinterpret_cast<const bf16x8 *>(p); }
GNN_DEV void st8(bf16 *p, bf16x8 v) { *reinterpret_cast<bf16x8 *>(p) = v; }
// The same, element idx of a row array as its base plus a 32-bit byte offset:
// with a uniform base the compiler uses the global instructions' scalar-base
// form instead of a 64-bit address per lane (the arrays are below 4 GiB).
GNN_DEV bf16x8 ld8_at(const bf16 *base, uint32_t idx) {
    return *reinterpret_cast<const bf16x8 *>(reinterpret_cast<const char *>(base) + (uint64_t)(idx * 2u));
}
GNN_DEV void st8_at(bf16 *base, uint32_t idx, bf16x8 v) {
    *reinterpret_cast<bf16x8 *>(reinterpret_cast<char *>(base) + (uint64_t)(idx * 2u)) = v;
}

// ---------------------------------------------------------------------------
// Edge features of the current state, in the edge's normal frame (zero for
// the padding rows e >= E).
GNN_DEV void edge_features(int64_t e, int64_t E, const int32_t *ea, const int32_t *eb, const float *nx,
                           const float *ny, const float *len, const uint8_t *closed, const double *w,
                           const double *uh, const double *vh, const double *z, float hs, float qs, float ls,
                           float f[kFeat]) {
#pragma unroll
    for (int i = 0; i < kFeat; i++) f[i] = 0.0f;
    if (e < E) {
        const int a = ea[e], b = eb[e];
        const double n1 = nx[e], n2 = ny[e];
        const bool wall = closed[e] != 0;
        const double ha = w[a] - z[a];
        const double qna = uh[a] * n1 + vh[a] * n2, qta = -uh[a] * n2 + vh[a] * n1;
        double hb = ha, qnb = wall ? -qna : qna, qtb = qta, dw = 0.0, dz = 0.0;
        if (b >= 0) {
            hb = w[b] - z[b];
            qnb = uh[b] * n1 + vh[b] * n2;
            qtb = -uh[b] * n2 + vh[b] * n1;
            dw = w[b] - w[a];
            dz = z[b] - z[a];
        }
        f[0] = float(ha / hs);
        f[1] = float(hb / hs);
        f[2] = float(qna / qs);
        f[3] = float(qnb / qs);
        f[4] = float(qta / qs);
        f[5] = float(qtb / qs);
        f[6] = float(dw / hs);
        f[7] = float(dz / hs);
        f[8] = b < 0 ? 1.0f : 0.0f;
        f[9] = wall ? 1.0f : 0.0f;
        f[10] = len[e] / ls;
        f[11] = sqrtf(float(fmax(ha, 0.0) / hs));
        f[12] = sqrtf(float(fmax(hb, 0.0) / hs));
        f[13] = ha > 1.0e-3 ? 1.0f : 0.0f;
        f[14] = hb > 1.0e-3 ? 1.0f : 0.0f;
        // First-order physics prior: the central-upwind flux of the two
        // centroid states after hydrostatic reconstruction (depths above the
        // higher bed), and the reconstruction's pressure corrections.
        const double zm = b >= 0 ? fmax(z[a], z[b]) : z[a];
        const double wa = w[a], wb = b >= 0 ? w[b] : w[a];
        const float sa = float(fmax(0.0, wa - zm)), sb = float(fmax(0.0, wb - zm));
        const float g = 9.8f, h0 = 1.0e-6f;
        const float ua = float(qna) * sa / (sa * sa + h0), ub = float(qnb) * sb / (sb * sb + h0);
        const float va = float(qta) * sa / (sa * sa + h0), vb = float(qtb) * sb / (sb * sb + h0);
        const float ca = sqrtf(g * sa), cb = sqrtf(g * sb);
        const float smax = fmaxf(fmaxf(ua + ca, ub + cb), 0.0f), smin = fminf(fminf(ua - ca, ub - cb), 0.0f);
        float k0 = 0.0f, k1 = 0.0f, k2 = 0.0f;
        if (smax - smin > 1.0e-6f) {
            const float inv = 1.0f / (smax - smin), sm = smax * smin;
            const float fa0 = sa * ua, fa1 = sa * ua * ua + 0.5f * g * sa * sa, fa2 = sa * ua * va;
            const float fb0 = sb * ub, fb1 = sb * ub * ub + 0.5f * g * sb * sb, fb2 = sb * ub * vb;
            k0 = (smax * fa0 - smin * fb0 + sm * (sb - sa)) * inv;
            k1 = (smax * fa1 - smin * fb1 + sm * (sb * ub - sa * ua)) * inv;
            k2 = (smax * fa2 - smin * fb2 + sm * (sb * vb - sa * va)) * inv;
        }
        const float ps = 0.5f * g * hs * hs;
        f[15] = k0 / qs;
        f[16] = k1 / ps;
        f[17] = k2 / qs;
        f[18] = 0.5f * g * (float(ha * ha) - sa * sa) / ps;
        f[19] = 0.5f * g * (float(hb * hb) - sb * sb) / ps;
    }
}

__global__ __launch_bounds__(kB) void k_features(int64_t E, int64_t Epad, const int32_t *ea, const int32_t *eb,
                                                 const float *nx, const float *ny, const float *len,
                                                 const uint8_t *closed, const double *w, const double *uh,
                                                 const double *vh, const double *z, float hs, float qs, float ls,
                                                 bf16 *X) {
    const int64_t e = gidx();
    if (e >= Epad) return;
    float f[kFeat];
    edge_features(e, E, ea, eb, nx, ny, len, closed, w, uh, vh, z, hs, qs, ls, f);
#pragma unroll
    for (int c = 0; c < kFeat / 8; c++) {
        bf16x8 v;
#pragma unroll
        for (int j = 0; j < 8; j++) v[j] = (bf16)f[c * 8 + j];
        st8(X + e * kFeat + c * 8, v);
    }
}

// ---------------------------------------------------------------------------
// Physical part of the edge outputs: the first-order central-upwind flux
// (Kurganov, Noelle and Petrova, SIAM J. Sci. Comput. 23, 2001) of the two
// centroid states after hydrostatic reconstruction (Audusse, Bouchut,
// Bristeau, Klein and Perthame, SIAM J. Sci. Comput. 25, 2004), added to
// the network's outputs in their units (divided by the flux scale):
//   0: mass flux from a to b,
//   1, 2: normal / tangential momentum change of side a per unit length,
//   3, 4: the same for side b.
// Each side's normal momentum takes the flux minus its own reconstructed
// pressure 0.5 g s^2.  That drops the term 0.5 g h^2 of the cell's centroid
// depth, whose sum over a cell's edges is 0.5 g h^2 times the sum of
// length * normal, zero for a closed polygon.  Written as below, a lake at
// rest (equal stage, zero momentum, any bed, dry cells included) gives
// exactly zero on every edge, so the physical part is well balanced without
// relying on cancellation between edges.
GNN_DEV void base_flux_edge(int64_t e, const int32_t *ea, const int32_t *eb, const float *nx, const float *ny,
                            const uint8_t *closed, const double *w, const double *uh, const double *vh,
                            const double *z, double g, double h_dry, double inv_fs, float *out) {
    // no fused multiply-adds here: pa and pb must be rounded the same way for
    // pb - pa to vanish when sa == sb
#pragma clang fp contract(off)
    const int a = ea[e], b = eb[e];
    const double n1 = nx[e], n2 = ny[e];
    // depths and normal-frame momenta from the FP64 state (no cancellation
    // in FP32), then the flux in FP32: equal inputs still give equal
    // pressures, so a lake at rest stays exactly zero
    const double wa = w[a], za = z[a], ha = wa - za;
    const double qna = uh[a] * n1 + vh[a] * n2, qta = -uh[a] * n2 + vh[a] * n1;
    double wb = wa, zb = za, hb = ha, qnb = closed[e] ? -qna : qna, qtb = qta;
    if (b >= 0) {
        wb = w[b];
        zb = z[b];
        hb = wb - zb;
        qnb = uh[b] * n1 + vh[b] * n2;
        qtb = -uh[b] * n2 + vh[b] * n1;
    }
    const double zm = fmax(za, zb);
    const float sa = (float)fmax(0.0, wa - zm), sb = (float)fmax(0.0, wb - zm);
    const float ua = ha > h_dry ? (float)qna / (float)ha : 0.0f, va = ha > h_dry ? (float)qta / (float)ha : 0.0f;
    const float ub = hb > h_dry ? (float)qnb / (float)hb : 0.0f, vb = hb > h_dry ? (float)qtb / (float)hb : 0.0f;
    const float gf = (float)g;
    const float ca = sqrtf(gf * sa), cb = sqrtf(gf * sb);
    const float smax = fmaxf(fmaxf(ua + ca, ub + cb), 0.0f), smin = fminf(fminf(ua - ca, ub - cb), 0.0f);
    float k0 = 0.0f, mna = 0.0f, k2 = 0.0f, mnb = 0.0f;
    if (smax - smin > 1.0e-6f) {
        const float inv = 1.0f / (smax - smin), sm = smax * smin;
        const float qa = sa * ua, qb = sb * ub;
        const float pa = 0.5f * gf * sa * sa, pb = 0.5f * gf * sb * sb;
        const float core = smax * qa * ua - smin * qb * ub + sm * (qb - qa);
        k0 = (smax * qa - smin * qb + sm * (sb - sa)) * inv;
        mna = -(core - smin * (pb - pa)) * inv;
        mnb = (core + smax * (pa - pb)) * inv;
        k2 = (smax * qa * va - smin * qb * vb + sm * (sb * vb - sa * va)) * inv;
    }
    const float s = (float)inv_fs;
    out[0] = k0 * s;
    out[1] = mna * s;
    out[2] = -k2 * s;
    out[3] = mnb * s;
    out[4] = k2 * s;
    out[5] = out[6] = out[7] = 0.0f;
}

// ---------------------------------------------------------------------------
// Y = act(X W + b) (+ R) of the 16 rows m0 .. m0+15 whose A fragments a[]
// (K/32 of them) are in registers: one wavefront, Nout/16 output blocks.
template <int K, int NOUT, bool OUTF32>
GNN_DEV void linear_tile(int64_t m0, int r, int g, const bf16x8 *a, const bf16 *WT, const float *bias, const bf16 *R,
                         bf16 *Zs, void *Yv, int act) {
#pragma unroll
    for (int nb = 0; nb < NOUT / 16; nb++) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kk = 0; kk < K / 32; kk++) acc = mfma(a[kk], ld8(WT + (nb * 16 + r) * K + kk * 32 + g * 8), acc);
        const int n = nb * 16 + r;
        const float bn = bias[n];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int64_t idx = (m0 + 4 * g + j) * NOUT + n;
            const float zz = acc[j] + bn;
            if (Zs) Zs[idx] = (bf16)zz;
            float y = act ? silu(zz) : zz;
            if (R) y += (float)R[idx];
            if (OUTF32)
                static_cast<float *>(Yv)[idx] = y;
            else
                static_cast<bf16 *>(Yv)[idx] = (bf16)y;
        }
    }
}

// Y = act(X W + b) (+ R): one wavefront per 16 rows; the 16 rows' A
// fragments stay in registers across the Nout/16 output blocks.
template <int K, int NOUT, bool OUTF32>
__global__ __launch_bounds__(kB) void k_linear_fwd(int64_t M, const bf16 *X, const bf16 *WT, const float *bias,
                                                   const bf16 *R, bf16 *Zs, void *Yv, int act) {
    const int64_t wave = (gidx()) >> 6;
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    const int64_t m0 = wave * 16;
    if (m0 >= M) return;
    bf16x8 a[K / 32];
#pragma unroll
    for (int kk = 0; kk < K / 32; kk++) a[kk] = ld8(X + (m0 + r) * K + kk * 32 + g * 8);
    linear_tile<K, NOUT, OUTF32>(m0, r, g, a, WT, bias, R, Zs, Yv, act);
}

// ---------------------------------------------------------------------------
// Persistent dense layers of the inference step.  Each wavefront keeps the
// layer's weight fragments (the B operands of all its output blocks) in
// registers and walks tiles of 16 rows, so the weights are read once per
// wavefront rather than once per tile.  Results leave the accumulators (lane
// (r, g) holds rows 4g .. 4g+3 of column r of each 16-column block) through a
// per-wavefront LDS tile and are stored as 16-byte row segments: lane (r, g)
// writes columns 8g .. 8g+7 and 32+8g .. 32+8g+7 of row r, the layout of its
// own A fragments, so a residual input is already in its registers.  The
// arithmetic is that of k_linear_fwd (same MFMA order, same roundings).
constexpr int kTileLd = 68;  // LDS row of a 16 x 64 FP32 tile (+4 words: rows 4 apart hit other banks)
constexpr int kPersistMax = 1024;  // workgroups of a persistent launch

// Orders one wavefront's LDS writes before its reads of other lanes' data.
GNN_DEV void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// SiLU(acc + bias) of a 16 x 64 tile (acc[nb]: columns nb*16 .. nb*16+15),
// optionally saving the pre-activation (BF16, rows of 64) to Zs; returns lane
// (r, g)'s row segments y0 (columns 8g..) and y1 (columns 32+8g..) of row r.
GNN_DEV void silu_rows(const f32x4 acc[4], const float br[4], int r, int g, float (*T)[kTileLd], bf16 *Zs,
                       int64_t m0, float y0[8], float y1[8]) {
#pragma unroll
    for (int nb = 0; nb < 4; nb++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const float zz = acc[nb][j] + br[nb];
            if (Zs) Zs[(m0 + 4 * g + j) * 64 + nb * 16 + r] = (bf16)zz;
            T[4 * g + j][nb * 16 + r] = silu(zz);
        }
    wave_lds_sync();
#pragma unroll
    for (int i = 0; i < 8; i++) {
        y0[i] = T[r][8 * g + i];
        y1[i] = T[r][32 + 8 * g + i];
    }
    wave_lds_sync();
}

GNN_DEV void store_row_segments(bf16 *Y, int64_t row, int g, const float y0[8], const float y1[8]) {
    bf16x8 o0, o1;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        o0[i] = (bf16)y0[i];
        o1[i] = (bf16)y1[i];
    }
    st8(Y + row * 64 + 8 * g, o0);
    st8(Y + row * 64 + 32 + 8 * g, o1);
}

// Features of the state, the physical part of the edge outputs (P, may be
// NULL) and the encoder layer H1 = SiLU(X0 W + b) in one pass; X0s and Zs
// (may be NULL) receive the features and the pre-activation (training).  A
// wavefront takes 64 edges at a time: each lane forms the features and the
// physical part of one edge, the feature rows go through LDS to the A
// fragments of four 16-row MFMA tiles.  Epad must be a multiple of 64.
constexpr int kXLd = kFeat + 8;  // LDS row of the feature tile (BF16; +8: conflict-free 16-byte reads)

__global__ __launch_bounds__(kB) void k_feat_enc(int64_t E, int64_t Epad, const int32_t *ea, const int32_t *eb,
                                                 const float *nx, const float *ny, const float *len,
                                                 const uint8_t *closed, const double *w, const double *uh,
                                                 const double *vh, const double *z, float hs, float qs, float ls,
                                                 const bf16 *WT, const float *bias, double g_acc, double h_dry,
                                                 double inv_fs, float *P, bf16 *X0s, bf16 *Zs, bf16 *H1) {
    __shared__ float tiles[kB / 64][16][kTileLd];
    __shared__ bf16 xtile[kB / 64][64][kXLd];
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    float(*T)[kTileLd] = tiles[threadIdx.x >> 6];
    bf16(*X)[kXLd] = xtile[threadIdx.x >> 6];
    bf16x8 bw[4];
    float br[4];
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
        bw[nb] = ld8(WT + (nb * 16 + r) * kFeat + g * 8);
        br[nb] = bias[nb * 16 + r];
    }
    int64_t t0, step;
    wave_tiles(t0, step);
    const int64_t t1 = Epad / 64;
    for (int64_t t = t0; t < t1; t += step) {
        const int64_t m64 = t * 64, e = m64 + lane;
        float f[kFeat];
        edge_features(e, E, ea, eb, nx, ny, len, closed, w, uh, vh, z, hs, qs, ls, f);
#pragma unroll
        for (int c = 0; c < kFeat / 8; c++) {
            bf16x8 v;
#pragma unroll
            for (int i = 0; i < 8; i++) v[i] = (bf16)f[c * 8 + i];
            *reinterpret_cast<bf16x8 *>(&X[lane][c * 8]) = v;
            if (X0s) st8(X0s + e * kFeat + c * 8, v);
        }
        if (P) {
            float o[kFOut] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
            if (e < E) base_flux_edge(e, ea, eb, nx, ny, closed, w, uh, vh, z, g_acc, h_dry, inv_fs, o);
            float4 *p = reinterpret_cast<float4 *>(P + e * kFOut);
            p[0] = make_float4(o[0], o[1], o[2], o[3]);
            p[1] = make_float4(o[4], o[5], o[6], o[7]);
        }
        wave_lds_sync();
#pragma unroll
        for (int sub = 0; sub < 4; sub++) {
            const int64_t m0 = m64 + 16 * sub;
            const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&X[16 * sub + r][8 * g]);
            f32x4 acc[4];
#pragma unroll
            for (int nb = 0; nb < 4; nb++) acc[nb] = mfma(a, bw[nb], f32x4{0.0f, 0.0f, 0.0f, 0.0f});
            float y0[8], y1[8];
            silu_rows(acc, br, r, g, T, Zs, m0, y0, y1);
            store_row_segments(H1, m0 + r, g, y0, y1);
        }
        wave_lds_sync();
    }
}

constexpr int kWLd = 128 + 8;  // LDS row of a message-passing layer's weights (BF16)

// One message-passing layer, node sums and concatenation folded into the
// matrix product: [H_e, S_ea + S_eb] W = H_e W1 + sum over the edges t of
// both triangles of H_t W2 (W1, W2 the first and second 64 rows of W), so
// the gathered rows go to the matrix cores as they are (K = 7 x 64: the
// edge's own row, then the three edges of ea and the three of eb in
// tri_edges order from nbr6[e], -1 for a missing triangle), with the sums
// exact in FP32 rather than rounded to BF16 first.  The workgroup's four
// wavefronts share the weights in LDS (rows padded by 16 bytes: the 16 rows a
// fragment load reads fall in different banks); each wavefront walks tiles of
// 16 edges, the next tile's rows and neighbour lists requested before this
// tile's gathers.  Zs (may be NULL) receives the pre-activation; Hout (may be
// NULL when the head follows) the new rows H + SiLU(Z).  With WTh (the head's
// weights, [32][64]) the head follows in registers: F[Epad][kFOut] = the
// first kFOut columns of Hout W_h + b_h (+ P), bitwise what k_head computes
// from the stored rows.
__global__ __launch_bounds__(kB) void k_mp_direct(int64_t E, int64_t Epad, const int32_t *nbr6, const bf16 *H,
                                                  const bf16 *WT, const float *bias, bf16 *Zs, bf16 *Hout,
                                                  const bf16 *WTh, const float *bias_h, const float *P, float *F) {
    __shared__ float tiles[kB / 64][16][kTileLd];
    __shared__ bf16 wts[64][kWLd];
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    float(*T)[kTileLd] = tiles[threadIdx.x >> 6];
    for (int i = threadIdx.x; i < 64 * 16; i += kB) {
        const int row = i >> 4, col = (i & 15) * 8;
        *reinterpret_cast<bf16x8 *>(&wts[row][col]) = ld8(WT + row * 128 + col);
    }
    __syncthreads();
    float br[4];
#pragma unroll
    for (int nb = 0; nb < 4; nb++) br[nb] = bias[nb * 16 + r];
    bf16x8 bh0, bh1;
    float bhn = 0.0f;
    if (WTh) {
        bh0 = ld8(WTh + r * 64 + g * 8);
        bh1 = ld8(WTh + r * 64 + 32 + g * 8);
        bhn = bias_h[r];
    }
    const bf16x8 zero8 = {};
    int64_t t0, step;
    wave_tiles(t0, step);
    const int64_t t1 = Epad / 16;
    bf16x8 h0 = {}, h1 = {};
    int4 nl0 = make_int4(-1, -1, -1, -1), nl1 = make_int4(-1, -1, -1, -1);
    if (t0 < t1) {
        const int64_t e = t0 * 16 + r;
        h0 = ld8_at(H, (uint32_t)(e * 64 + g * 8));
        h1 = ld8_at(H, (uint32_t)(e * 64 + 32 + g * 8));
        nl0 = reinterpret_cast<const int4 *>(nbr6)[2 * e];
        nl1 = reinterpret_cast<const int4 *>(nbr6)[2 * e + 1];
    }
    for (int64_t t = t0; t < t1; t += step) {
        const int64_t m0 = t * 16, e = m0 + r;
        const bf16x8 c0 = h0, c1 = h1;
        const bool real = e < E;
        const int nbr[6] = {real ? nl0.x : -1, real ? nl0.y : -1, real ? nl0.z : -1,
                            real ? nl0.w : -1, real ? nl1.x : -1, real ? nl1.y : -1};
        bf16x8 n0[6], n1[6];
#pragma unroll
        for (int q = 0; q < 6; q++) {
            n0[q] = nbr[q] >= 0 ? ld8_at(H, (uint32_t)nbr[q] * 64u + g * 8) : zero8;
            n1[q] = nbr[q] >= 0 ? ld8_at(H, (uint32_t)nbr[q] * 64u + 32 + g * 8) : zero8;
        }
        const int64_t tn = t + step;
        if (tn < t1) {
            const int64_t en = tn * 16 + r;
            h0 = ld8_at(H, (uint32_t)(en * 64 + g * 8));
            h1 = ld8_at(H, (uint32_t)(en * 64 + 32 + g * 8));
            nl0 = reinterpret_cast<const int4 *>(nbr6)[2 * en];
            nl1 = reinterpret_cast<const int4 *>(nbr6)[2 * en + 1];
        }
        const bf16x8 a0 = real ? c0 : zero8, a1 = real ? c1 : zero8;
        f32x4 acc[4];
#pragma unroll
        for (int nb = 0; nb < 4; nb++) {
            const bf16x8 *wrow = reinterpret_cast<const bf16x8 *>(&wts[nb * 16 + r][g * 8]);
            const bf16x8 w1a = wrow[0], w1b = wrow[4], w2a = wrow[8], w2b = wrow[12];  // k-blocks 0..3
            f32x4 c = mfma(a0, w1a, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
            c = mfma(a1, w1b, c);
#pragma unroll
            for (int q = 0; q < 6; q++) {
                c = mfma(n0[q], w2a, c);
                c = mfma(n1[q], w2b, c);
            }
            acc[nb] = c;
        }
        float y0[8], y1[8];
        silu_rows(acc, br, r, g, T, Zs, m0, y0, y1);
        bf16x8 o0, o1;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            o0[i] = (bf16)(y0[i] + (float)c0[i]);
            o1[i] = (bf16)(y1[i] + (float)c1[i]);
        }
        if (Hout) {
            st8_at(Hout, (uint32_t)(e * 64 + 8 * g), o0);
            st8_at(Hout, (uint32_t)(e * 64 + 32 + 8 * g), o1);
        }
        if (WTh) {
            f32x4 ah = mfma(o0, bh0, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
            ah = mfma(o1, bh1, ah);
            if (r < kFOut) {
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int64_t idx = (m0 + 4 * g + j) * kFOut + r;
                    float v = ah[j] + bhn;
                    if (P) v += P[idx];
                    F[idx] = v;
                }
            }
        }
    }
}

// Head: F[Epad][kFOut] (FP32) = the first kFOut columns of H4 W + b, plus the
// physical part P (may be NULL).
__global__ __launch_bounds__(kB) void k_head(int64_t Epad, const bf16 *H4, const bf16 *WT, const float *bias,
                                             const float *P, float *F) {
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    const bf16x8 b0 = ld8(WT + r * 64 + g * 8), b1 = ld8(WT + r * 64 + 32 + g * 8);
    const float bn = bias[r];
    int64_t t0, step;
    wave_tiles(t0, step);
    const int64_t t1 = Epad / 16;
    for (int64_t t = t0; t < t1; t += step) {
        const int64_t m0 = t * 16, e = m0 + r;
        f32x4 acc = mfma(ld8(H4 + e * 64 + g * 8), b0, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
        acc = mfma(ld8(H4 + e * 64 + 32 + g * 8), b1, acc);
        if (r < kFOut) {
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int64_t idx = (m0 + 4 * g + j) * kFOut + r;
                float v = acc[j] + bn;
                if (P) v += P[idx];
                F[idx] = v;
            }
        }
    }
}

// dX = dZ W^T (W is [K][NOUT]).
template <int K, int NOUT>
__global__ __launch_bounds__(kB) void k_linear_bwd_x(int64_t M, const bf16 *dZ, const bf16 *W, bf16 *dX) {
    const int64_t wave = (gidx()) >> 6;
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    const int64_t m0 = wave * 16;
    if (m0 >= M) return;
    bf16x8 a[NOUT / 32];
#pragma unroll
    for (int nn = 0; nn < NOUT / 32; nn++) a[nn] = ld8(dZ + (m0 + r) * NOUT + nn * 32 + g * 8);
#pragma unroll
    for (int kb = 0; kb < K / 16; kb++) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int nn = 0; nn < NOUT / 32; nn++) acc = mfma(a[nn], ld8(W + (kb * 16 + r) * NOUT + nn * 32 + g * 8), acc);
#pragma unroll
        for (int j = 0; j < 4; j++) dX[(m0 + 4 * g + j) * K + kb * 16 + r] = (bf16)acc[j];
    }
}

// Partial dW = X^T dZ and db = colsum(dZ) of one workgroup's chunks of 32 rows
// (grid-strided), staged through LDS so both operands are read transposed.
template <int K, int NOUT>
__global__ __launch_bounds__(kB) void k_linear_bwd_w(int64_t M, const bf16 *X, const bf16 *dZ, float *partial) {
    constexpr int NBLK = (K / 16) * (NOUT / 16);
    constexpr int PER = (NBLK + 3) / 4;  // output blocks per wavefront
    __shared__ bf16 xs[32][K + 8];
    __shared__ bf16 ds[32][NOUT + 8];
    const int t = threadIdx.x, wv = t >> 6, lane = t & 63, r = lane & 15, g = lane >> 4;
    f32x4 acc[PER];
#pragma unroll
    for (int p = 0; p < PER; p++) acc[p] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    float bsum = 0.0f;
    const int64_t nchunks = M / 32;
    for (int64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const int64_t m0 = c * 32;
        for (int i = t; i < 32 * K / 8; i += kB) {
            const int row = i / (K / 8), col = (i % (K / 8)) * 8;
            const bf16x8 v = ld8(X + (m0 + row) * K + col);
#pragma unroll
            for (int j = 0; j < 8; j++) xs[row][col + j] = v[j];
        }
        for (int i = t; i < 32 * NOUT / 8; i += kB) {
            const int row = i / (NOUT / 8), col = (i % (NOUT / 8)) * 8;
            const bf16x8 v = ld8(dZ + (m0 + row) * NOUT + col);
#pragma unroll
            for (int j = 0; j < 8; j++) ds[row][col + j] = v[j];
        }
        __syncthreads();
        if (t < NOUT) {
            for (int row = 0; row < 32; row++) bsum += (float)ds[row][t];
        }
#pragma unroll
        for (int p = 0; p < PER; p++) {
            const int blk = wv + 4 * p;
            if (blk < NBLK) {
                const int kb = blk / (NOUT / 16), nb = blk % (NOUT / 16);
                bf16x8 av, bv;
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    av[j] = xs[g * 8 + j][kb * 16 + r];  // A[k][m] = X[m][k]
                    bv[j] = ds[g * 8 + j][nb * 16 + r];  // B[m][n] = dZ[m][n]
                }
                acc[p] = mfma(av, bv, acc[p]);
            }
        }
        __syncthreads();
    }
    float *out = partial + (int64_t)blockIdx.x * (K + 1) * NOUT;
#pragma unroll
    for (int p = 0; p < PER; p++) {
        const int blk = wv + 4 * p;
        if (blk < NBLK) {
            const int kb = blk / (NOUT / 16), nb = blk % (NOUT / 16);
#pragma unroll
            for (int j = 0; j < 4; j++) out[(kb * 16 + 4 * g + j) * NOUT + nb * 16 + r] = acc[p][j];
        }
    }
    if (t < NOUT) out[K * NOUT + t] = bsum;
}

// Sum of the per-workgroup partials (fixed order: deterministic).
__global__ __launch_bounds__(kB) void k_reduce_partials(int n, int nparts, const float *partial, float *dW, float *db,
                                                        int nW) {
    const int i = (int)gidx();
    if (i >= n) return;
    float s = 0.0f;
    for (int p = 0; p < nparts; p++) s += partial[(int64_t)p * n + i];
    if (i < nW)
        dW[i] = s;
    else
        db[i - nW] = s;
}

__global__ __launch_bounds__(kB) void k_silu_bwd(int64_t n, const bf16 *dY, const bf16 *Z, bf16 *dZ) {
    const int64_t i = gidx();
    if (i >= n) return;
    dZ[i] = (bf16)((float)dY[i] * silu_grad((float)Z[i]));
}

// ---------------------------------------------------------------------------
// Message passing: node sums over a triangle's edges, edge concatenation.
__global__ __launch_bounds__(kB) void k_node_sum(int64_t N, const int32_t *tri_edges, const bf16 *H, bf16 *S) {
    const int64_t t = gidx();
    if (t >= N * 8) return;
    const int64_t i = t >> 3;
    const int c = (int)(t & 7) * 8;
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const int e = tri_edges[i * 3 + q];
        if (e < 0) continue;
        const bf16x8 v = ld8(H + (int64_t)e * 64 + c);
#pragma unroll
        for (int j = 0; j < 8; j++) s[j] += (float)v[j];
    }
    bf16x8 o;
#pragma unroll
    for (int j = 0; j < 8; j++) o[j] = (bf16)s[j];
    st8(S + i * 64 + c, o);
}

__global__ __launch_bounds__(kB) void k_concat(int64_t E, int64_t Epad, const int32_t *ea, const int32_t *eb,
                                               const bf16 *H, const bf16 *S, bf16 *C) {
    const int64_t t = gidx();
    if (t >= Epad * 16) return;
    const int64_t e = t >> 4;
    const int c = (int)(t & 15) * 8;
    bf16x8 o;
    if (e >= E) {
#pragma unroll
        for (int j = 0; j < 8; j++) o[j] = (bf16)0.0f;
    } else if (c < 64) {
        o = ld8(H + e * 64 + c);
    } else {
        const int a = ea[e], b = eb[e];
        const bf16x8 sa = ld8(S + (int64_t)a * 64 + c - 64);
        if (b >= 0) {
            const bf16x8 sb = ld8(S + (int64_t)b * 64 + c - 64);
#pragma unroll
            for (int j = 0; j < 8; j++) o[j] = (bf16)((float)sa[j] + (float)sb[j]);
        } else {
            o = sa;
        }
    }
    st8(C + e * 128 + c, o);
}

__global__ __launch_bounds__(kB) void k_node_grad(int64_t N, const int32_t *tri_edges, const bf16 *dC, bf16 *dS) {
    const int64_t t = gidx();
    if (t >= N * 8) return;
    const int64_t i = t >> 3;
    const int c = (int)(t & 7) * 8;
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const int e = tri_edges[i * 3 + q];
        if (e < 0) continue;
        const bf16x8 v = ld8(dC + (int64_t)e * 128 + 64 + c);
#pragma unroll
        for (int j = 0; j < 8; j++) s[j] += (float)v[j];
    }
    bf16x8 o;
#pragma unroll
    for (int j = 0; j < 8; j++) o[j] = (bf16)s[j];
    st8(dS + i * 64 + c, o);
}

__global__ __launch_bounds__(kB) void k_edge_grad(int64_t E, int64_t Epad, const int32_t *ea, const int32_t *eb,
                                                  const bf16 *dHn, const bf16 *dC, const bf16 *dS, bf16 *dH) {
    const int64_t t = gidx();
    if (t >= Epad * 8) return;
    const int64_t e = t >> 3;
    const int c = (int)(t & 7) * 8;
    bf16x8 o;
    if (e >= E) {
#pragma unroll
        for (int j = 0; j < 8; j++) o[j] = (bf16)0.0f;
    } else {
        const int a = ea[e], b = eb[e];
        const bf16x8 x0 = ld8(dHn + e * 64 + c), x1 = ld8(dC + e * 128 + c), x2 = ld8(dS + (int64_t)a * 64 + c);
        float s[8];
#pragma unroll
        for (int j = 0; j < 8; j++) s[j] = (float)x0[j] + (float)x1[j] + (float)x2[j];
        if (b >= 0) {
            const bf16x8 x3 = ld8(dS + (int64_t)b * 64 + c);
#pragma unroll
            for (int j = 0; j < 8; j++) s[j] += (float)x3[j];
        }
#pragma unroll
        for (int j = 0; j < 8; j++) o[j] = (bf16)s[j];
    }
    st8(dH + e * 64 + c, o);
}

// The physical part of rows 0 .. E-1 into P[E][kFOut].
__global__ __launch_bounds__(kB) void k_base_flux(int64_t E, const int32_t *ea, const int32_t *eb, const float *nx,
                                                  const float *ny, const uint8_t *closed, const double *w,
                                                  const double *uh, const double *vh, const double *z, double g,
                                                  double h_dry, double inv_fs, float *P) {
    const int64_t e = gidx();
    if (e >= E) return;
    float o[kFOut];
    base_flux_edge(e, ea, eb, nx, ny, closed, w, uh, vh, z, g, h_dry, inv_fs, o);
    float4 *p = reinterpret_cast<float4 *>(P + e * kFOut);
    p[0] = make_float4(o[0], o[1], o[2], o[3]);
    p[1] = make_float4(o[4], o[5], o[6], o[7]);
}

// ---------------------------------------------------------------------------
// State update from the head's edge outputs (FP64 accumulation), loss and its
// gradient with respect to the state change.
// Packed edge record (32 bytes, int32[8]) read by the per-edge backward of
// the update (k_apply_bwd): ea, eb, bits of nx, ny and len (FP32), closed
// flag, two zero words.  One edge is two 16-byte loads instead of six
// scattered ones.
struct EdgeR {
    int a, b;
    float nx, ny, len;
    bool closed;
};
GNN_DEV EdgeR load_edge(const int32_t *rec, int64_t e) {
    const int4 r0 = reinterpret_cast<const int4 *>(rec)[2 * e];
    const int4 r1 = reinterpret_cast<const int4 *>(rec)[2 * e + 1];
    return EdgeR{r0.x, r0.y, __int_as_float(r0.z), __int_as_float(r0.w), __int_as_float(r1.x), r1.y != 0};
}

// Slot records of the per-triangle kernels, streamed in the order those
// kernels consume them (instead of a gather of the edge's record): slot q of
// triangle m (its tri_edges order) at 3m + q holds {the edge, the triangle
// across it (-1: boundary), FP32 bits of the edge's length and nx}, sny[3m +
// q] its ny; sflags[m] holds per slot bits 3q (an edge), 3q + 1 (closed),
// 3q + 2 (m is the edge's side a).  The head's outputs stay per edge (F).
struct SlotR {
    int e, other;
    float len, nx;
};
GNN_DEV SlotR load_slot(const int32_t *slots, int64_t k) {
    const int4 r = reinterpret_cast<const int4 *>(slots)[k];
    return SlotR{r.x, r.y, __int_as_float(r.z), __int_as_float(r.w)};
}

// Contributions of the three edges of triangle m to its integrated change:
// mass (side-signed edge flux), x and y momentum (rotated side outputs).
// With a group, the mass flux of an edge whose other side is in the group is
// left out (it would cancel in the group's sum).
// theta (may be NULL) scales the mass flux of each edge by the limiter factor
// of the cell it drains (k_limit).
__device__ __forceinline__ void edge_sums(int m, const int32_t *slots, const float *sny, int sf, const float *F,
                                          float fs, const double *theta, const int *mem, double &sh, double &sx,
                                          double &sy) {
#pragma unroll
    for (int q = 0; q < 3; q++) {
        if (!((sf >> (3 * q)) & 1)) continue;
        const int64_t k = (int64_t)m * 3 + q;
        const SlotR sr = load_slot(slots, k);
        const bool closed = (sf >> (3 * q + 1)) & 1, side_a = (sf >> (3 * q + 2)) & 1;
        const float4 f03 = *reinterpret_cast<const float4 *>(F + (int64_t)sr.e * kFOut);
        const float f0 = f03.x, f1 = f03.y, f2 = f03.z, f3 = f03.w, f4 = F[(int64_t)sr.e * kFOut + 4];
        const double L = (double)sr.len * (double)fs;
        const double n1 = sr.nx, n2 = sny[k];
        // the mass flux of an edge: one value, added to one side, subtracted from the other
        const int other = sr.other;
        const bool internal = mem && other >= 0 &&
                              (other == mem[0] || other == mem[1] || other == mem[2] || other == mem[3]);
        double c = (closed || internal) ? 0.0 : (double)f0 * L;
        if (theta) {
            const int donor = (c > 0.0) == side_a ? m : other;  // side a drains into b when c > 0
            if (donor >= 0) c *= theta[donor];
        }
        sh += side_a ? -c : c;
        const double mn = side_a ? f1 : f3, mt = side_a ? f2 : f4;
        sx += (mn * n1 - mt * n2) * L;
        sy += (mn * n2 + mt * n1) * L;
    }
}

// New state and (with targets) loss terms of triangle i from its change.
__device__ __forceinline__ void finish_cell(int64_t N, int64_t i, double dh, double dx, double dy,
                                            const double *area, const double *w0, const double *uh0,
                                            const double *vh0, double *w1, double *uh1, double *vh1,
                                            const double *dw_t, const double *duh_t, const double *dvh_t,
                                            const float *sigma, double area_ref, float *gq, float *loss_terms) {
    w1[i] = w0[i] + dh;
    uh1[i] = uh0[i] + dx;
    vh1[i] = vh0[i] + dy;
    if (dw_t) {
        // Area-weighted (the error of the volume and momentum a triangle
        // gains) Huber loss of the normalised error: quadratic up to kHuber
        // standard deviations, linear beyond, so the few violently changing
        // triangles of a trajectory do not dominate.
        constexpr double kHuber = 3.0;
        const double wgt = area[i] / area_ref;
        const double e[3] = {(dh - dw_t[i]) / sigma[0], (dx - duh_t[i]) / sigma[1], (dy - dvh_t[i]) / sigma[2]};
        double l = 0.0;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double ae = fabs(e[c]);
            l += ae <= kHuber ? e[c] * e[c] : kHuber * (2.0 * ae - kHuber);
            const double d = ae <= kHuber ? 2.0 * e[c] : 2.0 * kHuber * (e[c] > 0 ? 1.0 : -1.0);
            gq[i * 3 + c] = (float)(wgt * d / ((double)N * sigma[c]));
        }
        loss_terms[i] = (float)(wgt * l);
    }
}

// The update in two passes, so that every lane does the same work: first
// each triangle sums its own three edges (the mass flux of an edge whose
// other side is in its group left out); a lone triangle takes its update at
// once, a grouped one leaves its sums in M.  Then every member of a group
// adds up its members' sums in member order and takes the group's change.
__global__ __launch_bounds__(kB) void k_apply(int64_t N, const int32_t *slots, const float *sny,
                                              const int32_t *sflags, const double *area, const int32_t *group,
                                              const double *group_area, const double *delta, const int32_t *graph_of,
                                              const float *F, float fs, const double *theta, const double *w0,
                                              const double *uh0,
                                              const double *vh0, double *w1, double *uh1, double *vh1,
                                              const double *dw_t, const double *duh_t, const double *dvh_t,
                                              const float *sigma, double area_ref, float *gq, float *loss_terms,
                                              double4 *M) {
    (void)group_area;
    const int64_t i = gidx();
    if (i >= N) return;
    double sh = 0.0, sx = 0.0, sy = 0.0;
    const int4 grow = group ? reinterpret_cast<const int4 *>(group)[i] : make_int4((int)i, -1, -1, -1);
    const int sf = sflags[i];
    if (grow.y >= 0) {
        const int mem[4] = {grow.x, grow.y, grow.z, grow.w};
        edge_sums((int)i, slots, sny, sf, F, fs, theta, mem, sh, sx, sy);
        M[i] = make_double4(sh, sx, sy, 0.0);
        return;
    }
    edge_sums((int)i, slots, sny, sf, F, fs, theta, nullptr, sh, sx, sy);
    const double s = delta[graph_of[i]] / area[i];
    finish_cell(N, i, s * sh, s * sx, s * sy, area, w0, uh0, vh0, w1, uh1, vh1, dw_t, duh_t, dvh_t, sigma, area_ref,
                gq, loss_terms);
}

__global__ __launch_bounds__(kB) void k_apply_groups(int64_t N, const double *area, const int32_t *group,
                                                     const double *group_area, const double *delta,
                                                     const int32_t *graph_of, const double4 *M, const double *w0,
                                                     const double *uh0, const double *vh0, double *w1, double *uh1,
                                                     double *vh1, const double *dw_t, const double *duh_t,
                                                     const double *dvh_t, const float *sigma, double area_ref,
                                                     float *gq, float *loss_terms) {
    const int64_t i = gidx();
    if (i >= N) return;
    const int4 grow = reinterpret_cast<const int4 *>(group)[i];
    if (grow.y < 0) return;
    const int mem[4] = {grow.x, grow.y, grow.z, grow.w};
    double4 t = M[mem[0]];
#pragma unroll
    for (int j = 1; j < 4; j++) {
        if (mem[j] < 0) continue;
        const double4 m = M[mem[j]];
        t.x += m.x;
        t.y += m.y;
        t.z += m.z;
    }
    const double s = delta[graph_of[i]] / group_area[i];
    finish_cell(N, i, s * t.x, s * t.y, s * t.z, area, w0, uh0, vh0, w1, uh1, vh1, dw_t, duh_t, dvh_t, sigma,
                area_ref, gq, loss_terms);
}

// Positivity limiter of the mass fluxes (the draining-time idea of
// Bollermann, Chen, Kurganov and Noelle, J. Sci. Comput. 56, 2013): the water
// a cell (or merged group) can give over the step is its volume; if the
// outgoing mass fluxes would take more, all of them are scaled by
// theta = volume / outflow.  Each edge takes the factor of the cell it drains,
// on both sides, so mass stays conserved and no cell's volume goes below 0.
__global__ __launch_bounds__(kB) void k_limit(int64_t N, const int32_t *slots, const int32_t *sflags,
                                              const double *area, const int32_t *group, const double *delta,
                                              const int32_t *graph_of, const float *F, float fs, const double *w,
                                              const double *z, double *theta, double4 *M) {
    const int64_t i = gidx();
    if (i >= N) return;
    const int4 grow = group ? reinterpret_cast<const int4 *>(group)[i] : make_int4((int)i, -1, -1, -1);
    const bool grouped = grow.y >= 0;
    const double vol = area[i] * fmax(w[i] - z[i], 0.0);
    double out = 0.0;
    const int sf = sflags[i];
#pragma unroll
    for (int q = 0; q < 3; q++) {
        if (!((sf >> (3 * q)) & 1) || ((sf >> (3 * q + 1)) & 1)) continue;  // no edge, or a closed one
        const int64_t k = i * 3 + q;
        const SlotR sr = load_slot(slots, k);
        const bool side_a = (sf >> (3 * q + 2)) & 1;
        const int other = sr.other;
        if (grouped && other >= 0 && (other == grow.x || other == grow.y || other == grow.z || other == grow.w))
            continue;
        const double c = (double)F[(int64_t)sr.e * kFOut] * (double)sr.len * (double)fs;
        out += side_a ? fmax(c, 0.0) : fmax(-c, 0.0);
    }
    if (grouped) {  // the group's factor is formed by k_limit_groups
        M[i] = make_double4(vol, out, 0.0, 0.0);
        return;
    }
    out *= delta[graph_of[i]];
    theta[i] = out > vol ? vol / out : 1.0;
}

// The factor of a merged group: its members' volumes and outflows summed in
// member order (each member forms it).
__global__ __launch_bounds__(kB) void k_limit_groups(int64_t N, const int32_t *group, const double *delta,
                                                     const int32_t *graph_of, const double4 *M, double *theta) {
    const int64_t i = gidx();
    if (i >= N) return;
    const int4 grow = reinterpret_cast<const int4 *>(group)[i];
    if (grow.y < 0) return;
    const int mem[4] = {grow.x, grow.y, grow.z, grow.w};
    double vol = 0.0, out = 0.0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        if (mem[j] < 0) continue;
        const double4 m = M[mem[j]];
        vol += m.x;
        out += m.y;
    }
    out *= delta[graph_of[i]];
    theta[i] = out > vol ? vol / out : 1.0;
}

// Momentum of cells shallower than h_dry set to zero (as the solver's
// protect); elsewhere, with froude_max > 0, the speed capped at
// froude_max sqrt(g h) with its direction kept (as the solver's
// maximum_allowed_speed caps speeds): a learned step must not leave a cell
// with a speed the flow cannot have, which the network, far outside the
// states it was trained on, would amplify in the next step.  Mass is not
// touched.
__global__ __launch_bounds__(kB) void k_protect_dry(int64_t N, const double *w, const double *z, double *uh,
                                                    double *vh, double h_dry, double g, double froude_max) {
    const int64_t i = gidx();
    if (i >= N) return;
    const double h = w[i] - z[i];
    if (h < h_dry) {
        uh[i] = 0.0;
        vh[i] = 0.0;
    } else if (froude_max > 0.0) {
        const double qu = uh[i], qv = vh[i];
        const double q2 = qu * qu + qv * qv;
        const double qmax = froude_max * sqrt(g * h) * h;
        if (q2 > qmax * qmax) {
            const double s = qmax / sqrt(q2);
            uh[i] = qu * s;
            vh[i] = qv * s;
        }
    }
}

// Gradient with respect to the integrated change of triangle a (scaled by
// 1 / area of its group): the sum of the group members' gq.
__device__ __forceinline__ void group_grad(int a, const int32_t *group, const double *group_area, const double *area,
                                           const float *gq, double &g0, double &g1, double &g2, double &A) {
    if (group && group[(int64_t)a * 4 + 1] >= 0) {
        g0 = g1 = g2 = 0.0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int m = group[(int64_t)a * 4 + j];
            if (m < 0) continue;
            g0 += gq[(int64_t)m * 3];
            g1 += gq[(int64_t)m * 3 + 1];
            g2 += gq[(int64_t)m * 3 + 2];
        }
        A = group_area[a];
    } else {
        g0 = gq[(int64_t)a * 3];
        g1 = gq[(int64_t)a * 3 + 1];
        g2 = gq[(int64_t)a * 3 + 2];
        A = area[a];
    }
}

__global__ __launch_bounds__(kB) void k_apply_bwd(int64_t E, int64_t Epad, const int32_t *rec,
                                                  const double *area, const int32_t *group,
                                                  const double *group_area, const double *delta,
                                                  const int32_t *graph_of, float fs, const float *gq, bf16 *dF) {
    const int64_t e = gidx();
    if (e >= Epad) return;
    float d[kOut];
#pragma unroll
    for (int i = 0; i < kOut; i++) d[i] = 0.0f;
    if (e < E) {
        const EdgeR ed = load_edge(rec, e);
        const int a = ed.a, b = ed.b;
        const double L = (double)ed.len * (double)fs;
        const double n1 = ed.nx, n2 = ed.ny;
        double ga0, ga1, ga2, Aa;
        group_grad(a, group, group_area, area, gq, ga0, ga1, ga2, Aa);
        const double sa = delta[graph_of[a]] / Aa * L;
        double d0 = -ga0 * sa;
        d[1] = (float)(sa * (ga1 * n1 + ga2 * n2));
        d[2] = (float)(sa * (-ga1 * n2 + ga2 * n1));
        bool internal = false;
        if (b >= 0) {
            double gb0, gb1, gb2, Ab;
            group_grad(b, group, group_area, area, gq, gb0, gb1, gb2, Ab);
            internal = group && group[(int64_t)a * 4] == group[(int64_t)b * 4];
            const double sb = delta[graph_of[b]] / Ab * L;
            d0 += gb0 * sb;
            d[3] = (float)(sb * (gb1 * n1 + gb2 * n2));
            d[4] = (float)(sb * (-gb1 * n2 + gb2 * n1));
        }
        if (internal) d0 = 0.0;
        d[0] = ed.closed ? 0.0f : (float)d0;
    }
#pragma unroll
    for (int c = 0; c < kOut / 8; c++) {
        bf16x8 v;
#pragma unroll
        for (int j = 0; j < 8; j++) v[j] = (bf16)d[c * 8 + j];
        st8(dF + e * kOut + c * 8, v);
    }
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kB) void k_adam(int64_t n, float *p, const float *g, float *m, float *v, float lr,
                                             float b1, float b2, float eps, float c1, float c2) {
    const int64_t i = gidx();
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ __launch_bounds__(kB) void k_cast_weights(int K, int NOUT, const float *W, bf16 *Wb, bf16 *WTb) {
    const int i = (int)gidx();
    if (i >= K * NOUT) return;
    const int k = i / NOUT, n = i % NOUT;
    const bf16 x = (bf16)W[i];
    Wb[i] = x;
    if (WTb) WTb[n * K + k] = x;
}

inline unsigned blocks(int64_t n) { return (unsigned)((n + kB - 1) / kB); }
inline hipStream_t S(void *s) { return static_cast<hipStream_t>(s); }

}  // namespace gnn

using namespace gnn;

extern "C" {

int gnn_features(void *stream, int64_t E, int64_t Epad, const int32_t *ea, const int32_t *eb, const float *nx,
                 const float *ny, const float *len, const uint8_t *closed, const double *w, const double *uh,
                 const double *vh, const double *z, float depth_scale, float momentum_scale, float length_scale,
                 void *X0) {
    hipLaunchKernelGGL(k_features, dim3(blocks(Epad)), dim3(kB), 0, S(stream), E, Epad, ea, eb, nx, ny, len, closed,
                       w, uh, vh, z, depth_scale, momentum_scale, length_scale, static_cast<bf16 *>(X0));
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int gnn_linear_fwd(void *stream, int64_t M, int K, int Nout, const void *X, const void *WT, const float *b,
                   const void *R, void *Zsave, void *Y, int act, int out_f32) {
    if (M % 16) return -1;
    const dim3 grid(blocks(M / 16 * 64)), blk(kB);
    const bf16 *x = static_cast<const bf16 *>(X), *wt = static_cast<const bf16 *>(WT), *r = static_cast<const bf16 *>(R);
    bf16 *zs = static_cast<bf16 *>(Zsave);
#define GNN_FWD(KK, NN, F32)                                                                                    \
    if (K == KK && Nout == NN && (out_f32 != 0) == F32) {                                                      \
        hipLaunchKernelGGL((k_linear_fwd<KK, NN, F32>), grid, blk, 0, S(stream), M, x, wt, b, r, zs, Y, act); \
        return hipGetLastError() == hipSuccess ? 0 : -1;                                                       \
    }
    GNN_FWD(32, 64, false)
    GNN_FWD(128, 64, false)
    GNN_FWD(64, 32, true)
    GNN_FWD(64, 64, false)
#undef GNN_FWD
    return -1;
}

// Workgroups of a persistent kernel that are resident on the whole device at
// once (its occupancy per CU, from its registers and LDS, times the CUs).  A
// persistent grid larger than this leaves its last workgroups waiting for a
// second round while every wavefront owns the same number of tiles.
extern "C++" template <typename Kernel>
unsigned resident_blocks(Kernel kernel) {
    static std::mutex lock;
    static std::map<std::pair<const void *, int>, unsigned> cached;  // (kernel, device)
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    const auto key = std::make_pair(reinterpret_cast<const void *>(kernel), dev);
    std::lock_guard<std::mutex> guard(lock);
    auto it = cached.find(key);
    if (it != cached.end()) return it->second;
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kB, 0) != hipSuccess || per_cu < 1) per_cu = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 1;
    return cached[key] = (unsigned)(per_cu * cus);
}

inline unsigned persist_blocks(int64_t rows, unsigned resident = kPersistMax) {
    const int64_t tiles = rows / 16, b = (tiles + kB / 64 - 1) / (kB / 64);
    const int64_t cap = resident < (unsigned)kPersistMax ? (int64_t)resident : (int64_t)kPersistMax;
    return (unsigned)(b < cap ? (b > 0 ? b : 1) : cap);
}

int gnn_feat_enc(void *stream, int64_t E, int64_t Epad, const int32_t *ea, const int32_t *eb, const float *nx,
                 const float *ny, const float *len, const uint8_t *closed, const double *w, const double *uh,
                 const double *vh, const double *z, float depth_scale, float momentum_scale, float length_scale,
                 const void *WT, const float *b, double g, double h_dry, float flux_scale, float *P, void *X0save,
                 void *Zsave, void *H1) {
    if (Epad % 64 || E > Epad || !(flux_scale > 0.0f)) return -1;
    hipLaunchKernelGGL(k_feat_enc, dim3(persist_blocks(Epad / 4, resident_blocks(k_feat_enc))), dim3(kB), 0, S(stream),
                       E, Epad, ea, eb, nx, ny, len,
                       closed, w, uh, vh, z, depth_scale, momentum_scale, length_scale, static_cast<const bf16 *>(WT),
                       b, g, h_dry, 1.0 / (double)flux_scale, P, static_cast<bf16 *>(X0save),
                       static_cast<bf16 *>(Zsave), static_cast<bf16 *>(H1));
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int gnn_mp_direct(void *stream, int64_t E, int64_t Epad, const int32_t *nbr6, const void *H, const void *WT,
                  const float *b, void *Zsave, void *Hout, const void *WTh, const float *bh, const float *P,
                  float *F) {
    if (Epad % 16 || E > Epad || H == Hout || (!Hout && !WTh) || (WTh && (!bh || !F))) return -1;
    hipLaunchKernelGGL(k_mp_direct, dim3(persist_blocks(Epad, resident_blocks(k_mp_direct))), dim3(kB), 0, S(stream), E,
                       Epad, nbr6,
                       static_cast<const bf16 *>(H), static_cast<const bf16 *>(WT), b, static_cast<bf16 *>(Zsave),
                       static_cast<bf16 *>(Hout), static_cast<const bf16 *>(WTh), bh, P, F);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int gnn_head(void *stream, int64_t Epad, const void *H4, const void *WT, const float *b, const float *P, float *F) {
    if (Epad % 16) return -1;
    hipLaunchKernelGGL(k_head, dim3(persist_blocks(Epad, resident_blocks(k_head))), dim3(kB), 0, S(stream), Epad,
                       static_cast<const bf16 *>(H4), static_cast<const bf16 *>(WT), b, P, F);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int gnn_linear_bwd_x(void *stream, int64_t M, int K, int Nout, const void *dZ, const void *W, void *dX) {
    if (M % 16) return -1;
    const dim3 grid(blocks(M / 16 * 64)), blk(kB);
    const bf16 *dz = static_cast<const bf16 *>(dZ), *w = static_cast<const bf16 *>(W);
    bf16 *dx = static_cast<bf16 *>(dX);
    if (K == 128 && Nout == 64) hipLaunchKernelGGL((k_linear_bwd_x<128, 64>), grid, blk, 0, S(stream), M, dz, w, dx);
    else if (K == 64 && Nout == 32) hipLaunchKernelGGL((k_linear_bwd_x<64, 32>), grid, blk, 0, S(stream), M, dz, w, dx);
    else if (K == 32 && Nout == 64) hipLaunchKernelGGL((k_linear_bwd_x<32, 64>), grid, blk, 0, S(stream), M, dz, w, dx);
    else return -1;
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int gnn_linear_bwd_w(void *stream, int64_t M, int K, int Nout, const void *X, const void *dZ, float *partial,
                     float *dW, float *db) {
    if (M % 32) return -1;
    const bf16 *x = static_cast<const bf16 *>(X), *dz = static_cast<const bf16 *>(dZ);
    if (K == 128 && Nout == 64) hipLaunchKernelGGL((k_linear_bwd_w<128, 64>), dim3(kParts), dim3(kB), 0, S(stream), M, x, dz, partial);
    else if (K == 64 && Nout == 32) hipLaunchKernelGGL((k_linear_bwd_w<64, 32>), dim3(kParts), dim3(kB), 0, S(stream), M, x, dz, partial);
    else if (K == 32 && Nout == 64) hipLaunchKernelGGL((k_linear_bwd_w<32, 64>), dim3(kParts), dim3(kB), 0, S(stream), M, x, dz, partial);
    else return -1;
    const int n = (K + 1) * Nout;
    hipLaunchKernelGGL(k_reduce_partials, dim3(blocks(n)), dim3(kB), 0, S(stream), n, kParts, partial, dW, db, K * Nout);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int gnn_silu_bwd(void *stream, int64_t n, const void *dY, const void *Z, void *dZ) {
    hipLaunchKernelGGL(k_silu_bwd, dim3(blocks(n)), dim3(kB), 0, S(stream), n, static_cast<const bf16 *>(dY),
                       static_cast<const bf16 *>(Z), static_cast<bf16 *>(dZ));
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int gnn_node_sum(void *stream, int64_t N, const int32_t *tri_edges, const void *H, void *Sv) {
    hipLaunchKernelGGL(k_node_sum, dim3(blocks(N * 8)), dim3(kB), 0, S(stream), N, tri_edges,
                       static_cast<const bf16 *>(H), static_cast<bf16 *>(Sv));
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int gnn_concat(void *stream, int64_t E, int64_t Epad, const int32_t *ea, const int32_t *eb, const void *H,
               const void *Sv, void *C) {
    hipLaunchKernelGGL(k_concat, dim3(blocks(Epad * 16)), dim3(kB), 0, S(stream), E, Epad, ea, eb,
                       static_cast<const bf16 *>(H), static_cast<const bf16 *>(Sv), static_cast<bf16 *>(C));
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int gnn_node_grad(void *stream, int64_t N, const int32_t *tri_edges, const void *dC, void *dS) {
    hipLaunchKernelGGL(k_node_grad, dim3(blocks(N * 8)), dim3(kB), 0, S(stream), N, tri_edges,
                       static_cast<const bf16 *>(dC), static_cast<bf16 *>(dS));
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int gnn_edge_grad(void *stream, int64_t E, int64_t Epad, const int32_t *ea, const int32_t *eb, const void *dHnext,
                  const void *dC, const void *dS, void *dH) {
    hipLaunchKernelGGL(k_edge_grad, dim3(blocks(Epad * 8)), dim3(kB), 0, S(stream), E, Epad, ea, eb,
                       static_cast<const bf16 *>(dHnext), static_cast<const bf16 *>(dC), static_cast<const bf16 *>(dS),
                       static_cast<bf16 *>(dH));
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int gnn_apply(void *stream, int64_t N, const int32_t *slots, const float *sny, const int32_t *sflags,
              const double *area, const int32_t *group, const double *group_area, const double *delta,
              const int32_t *graph_of, const float *F, float flux_scale, const double *theta, const double *w0, const double *uh0,
              const double *vh0, double *w1, double *uh1, double *vh1, const double *dw_t, const double *duh_t,
              const double *dvh_t, const float *sigma, double area_ref, float *gq, float *loss_terms,
              double *cell_sums) {
    if (group && (!group_area || !cell_sums)) return -1;
    double4 *M = reinterpret_cast<double4 *>(cell_sums);
    hipLaunchKernelGGL(k_apply, dim3(blocks(N)), dim3(kB), 0, S(stream), N, slots, sny, sflags, area, group,
                       group_area, delta, graph_of, F, flux_scale, theta, w0, uh0, vh0, w1, uh1, vh1, dw_t, duh_t, dvh_t, sigma,
                       area_ref, gq, loss_terms, M);
    if (group)
        hipLaunchKernelGGL(k_apply_groups, dim3(blocks(N)), dim3(kB), 0, S(stream), N, area, group, group_area, delta,
                           graph_of, M, w0, uh0, vh0, w1, uh1, vh1, dw_t, duh_t, dvh_t, sigma, area_ref, gq,
                           loss_terms);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int gnn_base_flux(void *stream, int64_t E, const int32_t *ea, const int32_t *eb, const float *nx, const float *ny,
                  const uint8_t *closed, const double *w, const double *uh, const double *vh, const double *z,
                  double g, double h_dry, float flux_scale, float *P) {
    if (!(flux_scale > 0.0f)) return -1;
    hipLaunchKernelGGL(k_base_flux, dim3(blocks(E)), dim3(kB), 0, S(stream), E, ea, eb, nx, ny, closed, w, uh, vh, z,
                       g, h_dry, 1.0 / (double)flux_scale, P);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int gnn_limit(void *stream, int64_t N, const int32_t *slots, const int32_t *sflags, const double *area,
              const int32_t *group, const double *delta, const int32_t *graph_of, const float *F, float flux_scale,
              const double *w, const double *z, double *theta, double *cell_sums) {
    if (group && !cell_sums) return -1;
    double4 *M = reinterpret_cast<double4 *>(cell_sums);
    hipLaunchKernelGGL(k_limit, dim3(blocks(N)), dim3(kB), 0, S(stream), N, slots, sflags, area, group, delta,
                       graph_of, F, flux_scale, w, z, theta, M);
    if (group)
        hipLaunchKernelGGL(k_limit_groups, dim3(blocks(N)), dim3(kB), 0, S(stream), N, group, delta, graph_of, M,
                           theta);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int gnn_protect(void *stream, int64_t N, const double *w, const double *z, double *uh, double *vh, double h_dry,
                double g, double froude_max) {
    if (!(g > 0.0) || froude_max < 0.0) return -1;
    hipLaunchKernelGGL(k_protect_dry, dim3(blocks(N)), dim3(kB), 0, S(stream), N, w, z, uh, vh, h_dry, g, froude_max);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int gnn_apply_bwd(void *stream, int64_t E, int64_t Epad, const int32_t *rec, const double *area,
                  const int32_t *group, const double *group_area, const double *delta, const int32_t *graph_of,
                  float flux_scale, const float *gq, void *dF) {
    if (group && !group_area) return -1;
    hipLaunchKernelGGL(k_apply_bwd, dim3(blocks(Epad)), dim3(kB), 0, S(stream), E, Epad, rec, area, group,
                       group_area, delta, graph_of, flux_scale, gq, static_cast<bf16 *>(dF));
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int gnn_adam(void *stream, int64_t n, float *param, const float *grad, float *m, float *v, float lr, float beta1,
             float beta2, float eps, int step) {
    const float c1 = 1.0f - powf(beta1, (float)step), c2 = 1.0f - powf(beta2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3(blocks(n)), dim3(kB), 0, S(stream), n, param, grad, m, v, lr, beta1, beta2, eps,
                       c1, c2);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int gnn_cast_weights(void *stream, int K, int Nout, const float *W, void *Wb, void *WTb) {
    hipLaunchKernelGGL(k_cast_weights, dim3(blocks((int64_t)K * Nout)), dim3(kB), 0, S(stream), K, Nout, W,
                       static_cast<bf16 *>(Wb), static_cast<bf16 *>(WTb));
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // extern "C"
