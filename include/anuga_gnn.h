/* anuga_gnn.h -- C ABI of the learned edge-flux surrogate's kernels (gfx950).
 *
 * A small edge graph network advances a shallow-water state by a time step
 * Delta spanning several CFL steps of the finite-volume solver: per edge it
 * predicts a mass flux, applied with opposite signs to the two triangles of
 * the edge (so the total mass changes only through open boundary edges, by
 * construction), and momentum contributions to each side.  Layers are dense
 * BF16 matrices on MFMA (v_mfma_f32_16x16x32_bf16) with FP32 accumulation;
 * the state update is accumulated in FP64.  These entry points have no
 * counterpart in the reference (its solver has no learned component); they
 * serve the learned rollout and its training (anuga_cuda_amd/learned.py).
 *
 * All pointers are device pointers; every launch goes to `stream` (a
 * hipStream_t, 0 for the default stream).  Matrices are row-major.  Row
 * counts M are multiples of 16 (callers pad), K is 32, 64 or 128, Nout is 32
 * or 64.  Returns 0, or -1 for an unsupported shape.  The edge outputs F are
 * FP32 rows of 8 (5 used); their gradient dF is BF16 rows of 32.
 */
#ifndef ANUGA_GNN_H
#define ANUGA_GNN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Edge features X0[E][32] (BF16) of the current state.  Edge e joins
 * triangles ea[e] and eb[e] (eb < 0: boundary edge, the outside state is the
 * inside one); (nx, ny) is its unit normal from ea to eb, len its length.
 * closed[e] != 0 marks a wall (reflected outside momentum).  The scales
 * normalise depths, momenta and lengths.  Rows E .. Epad-1 are zero. */
int gnn_features(void *stream, int64_t E, int64_t Epad, const int32_t *ea, const int32_t *eb, const float *nx,
                 const float *ny, const float *len, const uint8_t *closed, const double *w, const double *uh,
                 const double *vh, const double *z, float depth_scale, float momentum_scale, float length_scale,
                 void *X0);

/* Y[M][Nout] = act(X[M][K] W + b) (+ R).  WT = W transposed [Nout][K]
 * (BF16), b FP32.  act: 0 identity, 1 SiLU.  Zsave (may be NULL) receives the
 * pre-activation (BF16).  out_f32: Y is FP32 instead of BF16.  R (may be NULL)
 * is a BF16 residual added after the activation. */
int gnn_linear_fwd(void *stream, int64_t M, int K, int Nout, const void *X, const void *WT, const float *b,
                   const void *R, void *Zsave, void *Y, int act, int out_f32);

/* The layers below are persistent kernels: each wavefront keeps its
 * layer's weights in registers or LDS over many 16-row tiles. */

/* Features, physical part and encoder in one pass: H1[Epad][64] =
 * SiLU(X0 W + b) with X0 the edge features of gnn_features (WT [64][32];
 * bit for bit gnn_features, gnn_base_flux and gnn_linear_fwd),
 * and P[Epad][8] (may be NULL) the physical part of the edge outputs
 * (gnn_base_flux; rows E .. Epad-1 zero).  X0save and Zsave (may be NULL)
 * receive the features and the pre-activation. */
int gnn_feat_enc(void *stream, int64_t E, int64_t Epad, const int32_t *ea, const int32_t *eb, const float *nx,
                 const float *ny, const float *len, const uint8_t *closed, const double *w, const double *uh,
                 const double *vh, const double *z, float depth_scale, float momentum_scale, float length_scale,
                 const void *WT, const float *b, double g, double h_dry, float flux_scale, float *P, void *X0save,
                 void *Zsave, void *H1);

/* One message-passing layer with the node sums and the concatenation folded
 * into the matrix product: Hout = H + SiLU(H W1 + sum over the edges t of
 * both triangles of H[t] W2 + b) (W1, W2: the first and second 64 rows of W;
 * WT [64][128]), the edges listed per edge in nbr6[Epad][8] (int32: the
 * three edges of ea in tri_edges order, then those of eb, -1 for a boundary
 * edge's missing triangle and for padding rows, whose input row is zero).
 * This is [H, S[ea] + S[eb]] W + b with the node sums exact in FP32 (the
 * concatenation path rounds them to BF16 first).  Zsave (may be NULL)
 * receives the pre-activation, Hout (may be NULL if WTh is given) the rows.
 * With WTh (the head's weights, [32][64]) and bh, the head follows while the
 * rows are in registers: F as gnn_head computes it from them, bit for bit.
 * The persistent kernels need Epad a multiple of 16 (gnn_feat_enc: of 64). */
int gnn_mp_direct(void *stream, int64_t E, int64_t Epad, const int32_t *nbr6, const void *H, const void *WT,
                  const float *b, void *Zsave, void *Hout, const void *WTh, const float *bh, const float *P,
                  float *F);

/* Head: F[Epad][8] (FP32) = the first 8 columns of H4 W + b (WT [32][64]),
 * plus P[Epad][8] (may be NULL: the physical part).  Columns 0..4 are used
 * (gnn_apply). */
int gnn_head(void *stream, int64_t Epad, const void *H4, const void *WT, const float *b, const float *P, float *F);

/* dX[M][K] = dZ[M][Nout] W^T, with W [K][Nout] (BF16). */
int gnn_linear_bwd_x(void *stream, int64_t M, int K, int Nout, const void *dZ, const void *W, void *dX);

/* dW[K][Nout] (FP32) = X^T dZ and db[Nout] = column sums of dZ, reduced over
 * the M rows (partial is scratch of nparts*(K+1)*Nout floats, nparts = 256). */
int gnn_linear_bwd_w(void *stream, int64_t M, int K, int Nout, const void *X, const void *dZ, float *partial,
                     float *dW, float *db);

/* dZ = dY * SiLU'(Z), elementwise over n values (BF16). */
int gnn_silu_bwd(void *stream, int64_t n, const void *dY, const void *Z, void *dZ);

/* Node sums S[i][64] = sum over the (up to 3) edges of triangle i of H[e][64]
 * (tri_edges[i][3], -1 for none), then C[e][128] = [H[e], S[ea] + S[eb]]. */
int gnn_node_sum(void *stream, int64_t N, const int32_t *tri_edges, const void *H, void *S);
int gnn_concat(void *stream, int64_t E, int64_t Epad, const int32_t *ea, const int32_t *eb, const void *H,
               const void *S, void *C);

/* Backward of gnn_concat + gnn_node_sum for one layer:
 * dS[i] = sum over edges of i of dC[e][64:128];
 * dH[e] = dHnext[e] + dC[e][0:64] + dS[ea] + dS[eb]. */
int gnn_node_grad(void *stream, int64_t N, const int32_t *tri_edges, const void *dC, void *dS);
int gnn_edge_grad(void *stream, int64_t E, int64_t Epad, const int32_t *ea, const int32_t *eb, const void *dHnext,
                  const void *dC, const void *dS, void *dH);

/* The per-triangle kernels below read triangle i's edges per slot q (its
 * tri_edges order, s = 3i + q), streamed: slots[s][4] (int32: the edge, the
 * triangle across it, -1 on a boundary, the FP32 bits of its length and nx),
 * sny[s] (its ny, FP32; gnn_apply only), sflags[i] (bits 3q: an edge in the
 * slot, 3q + 1: closed, 3q + 2: i is its side a).
 *
 * Apply the head's outputs F[E][8] (FP32; columns 0: mass flux, 1-2: side-a
 * momentum (normal, tangential), 3-4: side-b momentum) over a step delta:
 * dq[i] = delta / area[i] * sum over edges of side-signed, rotated
 * contributions * len * flux_scale; a closed edge (a wall) carries no mass
 * flux.  group (may be NULL) merges small cells: group[i][4] lists the
 * members of triangle i's group (-1 padded; a lone triangle lists itself
 * only), group_area[i] the group's area; every member then takes
 * delta / group_area * (sum of the members' contributions, the mass flux of
 * edges inside the group left out).  theta (may be NULL) scales each edge's
 * mass flux by the factor of the cell it drains (gnn_limit).  Writes the new
 * state (w, uh, vh; may alias the inputs).  With targets (non-NULL dw_t, duh_t, dvh_t: the true
 * changes), also writes the area-weighted loss terms (area[i] / area_ref) *
 * sum over components of huber((dq - dq_true) / sigma) (quadratic up to 3,
 * linear beyond) and their gradient gq[i][3] with respect to dq, for the loss
 * sum(terms) / Ntri.  With groups, cell_sums (4 N doubles of scratch) holds
 * each grouped triangle's own sums between the two passes: first every
 * triangle sums its own edges, then a group's members add them up in member
 * order. */
int gnn_apply(void *stream, int64_t N, const int32_t *slots, const float *sny, const int32_t *sflags,
              const double *area, const int32_t *group, const double *group_area, const double *delta,
              const int32_t *graph_of, const float *F, float flux_scale, const double *theta, const double *w0, const double *uh0,
              const double *vh0, double *w1, double *uh1, double *vh1, const double *dw_t, const double *duh_t,
              const double *dvh_t, const float *sigma, double area_ref, float *gq, float *loss_terms,
              double *cell_sums);

/* Positivity limiter of the learned mass fluxes (draining time, Bollermann,
 * Chen, Kurganov and Noelle 2013): theta[i] = min(1, volume / outflow) of
 * triangle i's group (or of i), volume = sum of area * max(w - z, 0),
 * outflow = delta * the outgoing mass fluxes through the group's outer,
 * open edges.  Used at inference; training treats theta as 1.  With groups,
 * cell_sums (4 N doubles of scratch) holds each grouped triangle's volume and
 * outflow between the two passes. */
int gnn_limit(void *stream, int64_t N, const int32_t *slots, const int32_t *sflags, const double *area,
              const int32_t *group, const double *delta, const int32_t *graph_of, const float *F, float flux_scale,
              const double *w, const double *z, double *theta, double *cell_sums);

/* The physical part of the edge outputs into P[E][8] (columns 0..4, 5..7
 * zero; in the outputs' units: divided by flux_scale; gnn_head adds it to the
 * network's outputs): the first-order central-upwind
 * flux of the two centroid states after hydrostatic reconstruction, FP64
 * (mass flux from ea to eb; normal / tangential momentum change per unit
 * length of side ea, then of side eb, each side's normal momentum without its
 * own reconstructed pressure, which sums to zero around a cell).  Velocities
 * are zero where the depth is at most h_dry.  Exactly zero for a lake at
 * rest. */
int gnn_base_flux(void *stream, int64_t E, const int32_t *ea, const int32_t *eb, const float *nx, const float *ny,
                  const uint8_t *closed, const double *w, const double *uh, const double *vh, const double *z,
                  double g, double h_dry, float flux_scale, float *P);

/* uh = vh = 0 where w - z < h_dry (the solver's protect, without its mass
 * change); elsewhere, if froude_max > 0, the speed |(uh, vh)| / h capped at
 * froude_max * sqrt(g h) with its direction kept.  Returns -1 for g <= 0 or
 * froude_max < 0. */
int gnn_protect(void *stream, int64_t N, const double *w, const double *z, double *uh, double *vh, double h_dry,
                double g, double froude_max);

/* dF[Epad][32] (BF16) from gq: the transpose of gnn_apply's linear map (same
 * group arguments). */
int gnn_apply_bwd(void *stream, int64_t E, int64_t Epad, const int32_t *rec, const double *area,
                  const int32_t *group, const double *group_area, const double *delta, const int32_t *graph_of,
                  float flux_scale, const float *gq, void *dF);

/* Adam step on n FP32 parameters, then BF16 copies in both layouts:
 * W [K][Nout] -> Wb (same layout) and WTb ([Nout][K]); for biases pass K = 1
 * and WTb = NULL. */
int gnn_adam(void *stream, int64_t n, float *param, const float *grad, float *m, float *v, float lr, float beta1,
             float beta2, float eps, int step);
int gnn_cast_weights(void *stream, int K, int Nout, const float *W, void *Wb, void *WTb);

#ifdef __cplusplus
}
#endif
#endif
