// HIP kernels of the shallow-water step for gfx950.
//
// Two families share the numerics in sw_device.h:
//   * operator kernels, one per overridden ANUGA operator (the drop-in API:
//     each mirrors one reference kernel / C routine and its array contract);
//   * the edge-record fused step used by evolve() for mixed precision and for
//     meshes the vertex-record step of sw_step_vertex.h does not take (and on
//     request: it runs Euler, rk2, rk3 and every limiter too); per stage
//       flux       = compute_fluxes + boundary evaluation + gravity_wb; each
//                    wavefront stores the minimum of its CFL candidates
//                    (wavefront shuffles, a plain store);
//       timestep   = one workgroup per instance reduces those minima and does
//                    update_timestep and the time bookkeeping (k_timestep);
//       update     = update_conserved_quantities + protect + the next step's
//                    Manning friction;
//       distribute = extrapolate_second_order_sw + balance_deep_and_shallow +
//                    interpolate_from_vertices_to_edges + the next step's
//                    stage-gradient gravity term;
//     the explicit / semi-implicit updates are the only intermediates that go
//     through HBM and the host is not involved between steps.
//
// Data layout in HBM (one lane per triangle, 256-lane workgroups).  Per-edge
// and per-vertex data are slot-major: slot i (edge i faces vertex i) of
// triangle k lives at i*N + k, so one 64-lane load instruction reads whole
// 128-byte lines that no later instruction needs again:
//   * centroid state  cq[k]     = {w, uh, vh, z}  32-byte record
//   * edge state      eq[i*N+k] = {w, uh, vh, z}  32-byte record; a
//     neighbour's edge is a single 32-byte gather;
//   * edge[i*N+k] = {normal, length, m*N+n the slot of the neighbour's
//     matching edge or -(b+1) for boundary edge b}, a 32-byte record;
//   * flags[k]: bit i edge i is evaluated from the lower-numbered neighbour;
//     bit 3+i the evaluating triangle is full; bits 6-11 the order in which
//     the three edge fluxes are summed (2 bits per position); bits 12-13
//     number_of_boundaries;
//   * vertex coordinates (double2), vertex values, surrogate neighbours:
//     [3][N]; centroids (double2), {radius, area} (double2), updates: [N].
// Workgroup indices are remapped so each of the 8 XCDs sweeps one contiguous
// range of a locality-ordered mesh and neighbour gathers hit its own L2.
#pragma once

#include "sw_device.h"

namespace sw {

constexpr int kBlock = 256;

struct alignas(32) Q4 {
    double w, uh, vh, z;
};

struct DevParams {
    double g, epsilon, H0, h0, limiting_threshold, minimum_allowed_height, maximum_allowed_speed;
    double beta_w, beta_w_dry, beta_uh, beta_uh_dry, beta_vh, beta_vh_dry;
    double alpha_balance, CFL, evolve_max_timestep, evolve_min_timestep;
    int optimise_dry_cells, extrapolate_velocity_second_order, tight_slope_limiters;
    int use_centroid_velocities, use_sloped_mannings, max_smallsteps, default_order;
    int limiter;      // SW_LIMITER_*: distribute path
    double qbeta[3];  // Quantity.beta of stage, xmomentum, ymomentum (per-quantity limiters)
};

// Static geometry of edge i of triangle k, one 32-byte record at i*N + k.
struct alignas(32) EdgeRec {
    double nx, ny, len;  // outward normal and length
    int nb;              // neighbour's matching edge slot m*N+n, or -(b+1) for boundary edge b
    int tflags;          // the triangle's flags word (same in all three slots)
};

// FP32 edge geometry of the mixed-precision flux: normal, length, neighbour slot.
struct alignas(16) EdgeRecF {
    float nx, ny, len;
    int nb;
};

// Edge state of the mixed-precision loop, written by k_distribute and read by
// k_flux_mixed: FP32 depth (w - z formed in FP64) and momenta, and the depth
// class (depth_class) that carries the FP64 depth tests.
struct alignas(16) EdgeF {
    float h, uh, vh;
    int cls;
};

struct DevMesh {
    int N, Nb;
    int own_normals;  // shared-edge normals are exact negations, lengths equal
    const EdgeRec *edge;  // [3][N]
    const EdgeRecF *edgef;  // [3][N] the same in FP32 (mixed precision)
    const int *flags;     // [N]
    const int *sur;       // [3][N]
    const double2 *vc;    // [3][N] vertex coordinates
    const double2 *cc;    // [N]    centroid coordinates
    const double2 *ra;    // [N]    {radius, area}
    // mesh nodes (vertex-record step): node of vertex i of triangle k at
    // tri_node[i*N + k], node coordinates node_xy[n_nodes]
    const int *tri_node;
    const double2 *node_xy;
    const int *bcell, *bedge, *bkind;  // [Nb]
    const double *bdir;                // [Nb][3]
    // Ensembles: instance i owns device triangles [inst_off[i], inst_end[i]);
    // every inst_off is a multiple of kInstAlign, so no workgroup straddles
    // two instances.  wave_inst[w] is the instance of triangles 64w .. 64w+63.
    int n_inst;
    const int *inst_off, *inst_end;  // [n_inst]
    const int *wave_inst;            // [device triangles / 64, rounded up]
};

struct DevFields {
    Q4 *cq;                               // [N]     centroid w, uh, vh, z
    Q4 *cq0;                              // [N]     centroid values at the start of a Runge-Kutta step
    Q4 *eq;                               // [3][N]  edge w, uh, vh, z
    double *wv, *uhv, *vhv, *zv;          // [3][N]  vertex values
    double *ze;                           // [3][N]  bed edge values (static copy)
    double *wb, *uhb, *vhb, *zb;          // [Nb]    boundary values
    double *weu, *uheu, *vheu;            // [N]     explicit updates
    double *wsi, *uhsi, *vhsi;            // [N]     semi-implicit updates
    double *hc, *he, *hv, *hb;            // height    ([N], [3][N], [3][N], [Nb])
    double *uc, *ue, *uv, *ub;            // xvelocity
    double *vc, *ve, *vv, *vb;            // yvelocity
    double *friction, *zbar, *max_speed;  // [N]
    double *node_z;                       // [mesh nodes] bed at the node (vertex-record step)
    double2 *grav;                        // [N] -g*(wx, wy)*avg_h of the next step
    unsigned long long *wave_min;         // [kBlock/64 * blocks] CFL minimum of each wavefront (bits)
    EdgeF *eqf;                           // [3][N] mixed-precision edge state
};

// Scalar evolve state kept in device memory.  The first block of fields has
// the layout of sw_state in include/anuga_sw.h.
struct DevState {
    double time, starttime, finaltime, yieldtime, yieldstep, timestep, flux_timestep;
    double recorded_max_timestep, recorded_min_timestep;
    long long smallsteps, order, number_of_steps, number_of_first_order_steps;
    long long has_finaltime, event, error;
    // device-only
    long long halt, total_steps, bed_interp;
    // vertex-record step (sw_step_vertex.h): whether caller-set semi-implicit
    // terms are still to be applied; steps run since the step's records were
    // built; whether the bed edge values of every shared edge
    // are bitwise equal on both sides and every vertex slot of a mesh node
    // holds the same bed value (then the bed is read per node)
    long long csi_live, call_steps, bed_cont;
};

enum : long long { kErrNone = 0, kErrUpdate = 1, kErrSmallsteps = 2, kErrNeighbour = 4, kErrDiscontinuous = 8 };

SW_DEV int xcd_remap(int b, int nb) {
    int per = nb >> 3, rem = nb & 7, x = b & 7, idx = b >> 3;
    return x * per + (x < rem ? x : rem) + idx;
}

SW_DEV int tri_index() { return xcd_remap(blockIdx.x, gridDim.x) * kBlock + threadIdx.x; }

// Instance ranges start on multiples of this many triangles (the largest
// workgroup of any fused kernel).
constexpr int kInstAlign = 512;

// The triangle of this lane in the fused kernels, its instance and whether it
// is a real triangle of that instance (not padding).
struct Lane {
    int k, inst;
    bool valid;
};
template <int BLOCK = kBlock>
SW_DEV Lane lane_of(const DevMesh &M) {
    const int k0 = xcd_remap(blockIdx.x, gridDim.x) * BLOCK;
    const int inst = (M.n_inst == 1) ? 0 : M.wave_inst[k0 >> 6];
    const int end = (M.n_inst == 1) ? M.N : M.inst_end[inst];
    const int k = k0 + (int)threadIdx.x;
    return Lane{k, inst, k < end};
}

SW_DEV void set_error(DevState *st, long long code) {
    atomicOr((unsigned long long *)&st->error, (unsigned long long)code);
}

SW_DEV Q4 ldq(const Q4 *p) { return *p; }
SW_DEV EdgeRecF ld_edgef(const EdgeRecF *p) {
    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 a = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(p));
    return EdgeRecF{__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z), (int)a.w};
}
SW_DEV EdgeRec ld_edge(const EdgeRec *p) {
    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 *q = reinterpret_cast<const u32x4 *>(p);
    const u32x4 a = __builtin_nontemporal_load(q), b = __builtin_nontemporal_load(q + 1);
    EdgeRec r;
    r.nx = __hiloint2double((int)a.y, (int)a.x);
    r.ny = __hiloint2double((int)a.w, (int)a.z);
    r.len = __hiloint2double((int)b.y, (int)b.x);
    r.nb = (int)b.z;
    r.tflags = (int)b.w;
    return r;
}
SW_DEV int nob_of(int flags) { return (flags >> 12) & 3; }

// update_timestep (evolve.c:115-160 with the final-time clip as finaltime - time)
SW_DEV void timestep_logic(DevState *st, const DevParams &P) {
    double timestep = omin<double>(P.CFL * st->flux_timestep, P.evolve_max_timestep);
    st->recorded_max_timestep = omax<double>(timestep, st->recorded_max_timestep);
    st->recorded_min_timestep = omin<double>(timestep, st->recorded_min_timestep);
    if (timestep < P.evolve_min_timestep) {
        st->smallsteps += 1;
        if (st->smallsteps > P.max_smallsteps) {
            st->smallsteps = 0;
            if (st->order == 1)
                st->error |= kErrSmallsteps;
            else
                st->order = 1;
        }
    } else {
        st->smallsteps = 0;
        if (st->order == 1 && P.default_order == 2) st->order = 2;
    }
    if (st->has_finaltime && st->time + timestep > st->finaltime) timestep = st->finaltime - st->time;
    if (st->time + timestep > st->yieldtime) timestep = st->yieldtime - st->time;
    st->timestep = timestep;
}

// end-of-step bookkeeping (gpu_domain_advanced.py:2580, 2596-2618)
SW_DEV void finish_step(DevState *st) {
    st->number_of_steps += 1;
    st->total_steps += 1;
    if (st->order == 1) st->number_of_first_order_steps += 1;
    long long event = 0;
    if (st->has_finaltime && st->time >= st->finaltime - 1.0e-12) {
        st->time = st->finaltime;
        event = 2;
    } else if (st->time >= st->yieldtime) {
        event = 1;
    }
    st->event = event;
}

// Early exit of a fused kernel.  The first (or only) stage of a step halts an
// instance whose previous step ended at a yield, the final time or an error;
// a later Runge-Kutta stage runs only while its step runs (not halted, no
// error).
SW_DEV bool stage_skipped(DevState *st, int first) {
    if (first) {
        if (st->event || st->error) {
            if (threadIdx.x == 0) st->halt = 1;
            return true;
        }
        return false;
    }
    return st->halt || st->error;
}

// What k_timestep does in a fused step: the whole Euler step's bookkeeping;
// the first stage of a Runge-Kutta step (timestep and time advance: every
// stage uses the first stage's timestep, gpu_domain_advanced.py:2566-2573
// with ANUGA's evolve_one_rk2_step / evolve_one_rk3_step); the last stage of
// one (its flux_timestep and the end-of-step bookkeeping).
// kStageMid: a Runge-Kutta stage between the first and the last (rk3's
// second), which neither takes nor reports a timestep (vertex-record step).
enum : int { kStepEuler = 0, kStageFirst = 1, kStageLast = 2, kStageMid = 3 };

// ---------------------------------------------------------------------------
// compute_fluxes (shallow_water_ext.c:4519-4696).
//
// The C routine visits triangles in index order and computes every edge once,
// from the lower-numbered side, adding the flux to the neighbour with the
// opposite sign.  Here triangle k evaluates all three of its edges itself, but
// an edge shared with a lower-numbered neighbour n is evaluated exactly as the
// C evaluates it (left state from n, n's normal and length), and the sum is
// formed in the C's order: contributions of lower neighbours in ascending n,
// then k's own edges in local order (the order is precomputed per triangle),
// then the 1/area scaling.  The CFL candidates radii[k]/max_speed of all
// edges of k (gated by tri_full_flag of the evaluating triangle) give the same
// minimum as the C's edge-once loop.
//
// Each wavefront stores the minimum of its triangles' CFL candidates to
// wave_min (a plain store: no atomics, no fences); k_timestep reduces them.
//
// FUSED adds: boundary values evaluated on the fly (reflective / Dirichlet /
// transmissive) and gravity_wb.
//
// first: the first (or only) stage of a step.  A later Runge-Kutta stage runs
// only if its step is running: not halted and no error.
template <typename R, bool FUSED, int BLOCK = kBlock>
__global__ __launch_bounds__(BLOCK) void k_flux(DevMesh M, DevFields F, DevParams P, DevState *st_all, int first) {
    const Lane ln = lane_of<BLOCK>(M);
    DevState *st = st_all + ln.inst;
    if (FUSED && stage_skipped(st, first)) return;
    const int N = M.N;
    const int k = ln.k;
    double dtk = P.evolve_max_timestep;
    if (ln.valid) {
        // Load phase: every per-triangle and per-edge input is requested before
        // any of it is used, then the three neighbour edge records are gathered
        // together, so a wavefront waits for two memory round trips in all.
        EdgeRec er[3];
        Q4 own[3], nbq[3];
#pragma unroll
        for (int i = 0; i < 3; i++) er[i] = ld_edge(M.edge + i * N + k);
#pragma unroll
        for (int i = 0; i < 3; i++) own[i] = ldq(F.eq + i * N + k);
        const double2 ra = ld_s(M.ra + k);
        const double2 gr = FUSED ? ld_s(F.grav + k) : make_double2(0.0, 0.0);
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const int e = er[i].nb;
            nbq[i] = ldq(F.eq + (e >= 0 ? e : i * N + k));  // boundary edge: its own record
        }
        // (every loaded word is consumed, so no load result register is
        // recycled while the load is still in flight)
        const int flags = er[0].tflags | er[1].tflags | er[2].tflags;
        const double eps = P.epsilon;
        const double g = P.g;
        // The edge summed last in the C's order; the first two in that order
        // may be added as they come (0 + a + b is exact in either order).
        const int last_edge = (flags >> 10) & 3;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;   // flux sum of the first two edges
        double l0 = 0.0, l1 = 0.0, l2 = 0.0;   // signed flux of the last edge
        bool have_last = false;
        double sidex = 0.0, sidey = 0.0;        // gravity_wb side term, edge order
        double cfl_speed = 0.0;  // max over CFL-gated edges: min r/ms_i == r/max ms_i
        double max_speed_total = 0.0;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const int e = er[i].nb;
            const bool lower = (flags >> i) & 1;
            const double lk = er[i].len;
            if (FUSED) {
                const double hh = own[i].w - own[i].z;
                const double fact = -0.5 * g * hh * hh * lk;
                sidex = sidex + fact * er[i].nx;
                sidey = sidey + fact * er[i].ny;
            }
            double n1 = er[i].nx, n2 = er[i].ny, len = lk;
            double ql[3], qr[3], zl, zr;
            if (e >= 0) {
                if (lower) {
                    ql[0] = nbq[i].w; ql[1] = nbq[i].uh; ql[2] = nbq[i].vh; zl = nbq[i].z;
                    qr[0] = own[i].w; qr[1] = own[i].uh; qr[2] = own[i].vh; zr = own[i].z;
                    if (M.own_normals) {
                        n1 = -n1;
                        n2 = -n2;
                    } else {
                        const EdgeRec en = ld_edge(M.edge + e);
                        n1 = en.nx;
                        n2 = en.ny;
                        len = en.len;
                    }
                } else {
                    ql[0] = own[i].w; ql[1] = own[i].uh; ql[2] = own[i].vh; zl = own[i].z;
                    qr[0] = nbq[i].w; qr[1] = nbq[i].uh; qr[2] = nbq[i].vh; zr = nbq[i].z;
                }
            } else {
                const int m = -e - 1;
                ql[0] = nbq[i].w; ql[1] = nbq[i].uh; ql[2] = nbq[i].vh; zl = nbq[i].z;
                zr = zl;
                if (FUSED) {
                    const int kind = M.bkind[m];
                    double b0 = ql[0], b1 = ql[1], b2 = ql[2];  // transmissive: inside edge state
                    if (kind == 0) {  // reflective (evaluate_segment.cu:36-63)
                        reflect_momentum(ql[1], ql[2], n1, n2, b1, b2);
                    } else if (kind == 1) {  // Dirichlet constants
                        b0 = M.bdir[3 * m];
                        b1 = M.bdir[3 * m + 1];
                        b2 = M.bdir[3 * m + 2];
                    }
                    qr[0] = b0; qr[1] = b1; qr[2] = b2;
                } else {
                    qr[0] = F.wb[m]; qr[1] = F.uhb[m]; qr[2] = F.vhb[m];
                }
            }
        if (P.optimise_dry_cells && fabs(ql[0] - zl) < eps && fabs(qr[0] - zr) < eps) continue;
            if (fabs(zl - zr) > 1.0e-10) set_error(st, kErrDiscontinuous);
            double efd[3], ms;
            if constexpr (std::is_same<R, double>::value) {
                flux_central<double>(ql, qr, zl, zr, n1, n2, eps, P.h0, P.limiting_threshold, P.g, efd, ms);
            } else {
                flux_central_mixed(ql, qr, zl, zr, n1, n2, eps, P.h0, P.limiting_threshold, P.g, efd, ms);
            }
            // the lower-numbered triangle of the edge adds the flux, the other subtracts
            double f0 = efd[0] * len, f1 = efd[1] * len, f2 = efd[2] * len;
            if (!lower) {
                f0 = -f0;
                f1 = -f1;
                f2 = -f2;
            }
            if (i == last_edge) {
                l0 = f0;
                l1 = f1;
                l2 = f2;
                have_last = true;
            } else {
                s0 = s0 + f0;
                s1 = s1 + f1;
                s2 = s2 + f2;
            }
            if (((flags >> (3 + i)) & 1) && ms > eps) cfl_speed = omax<double>(cfl_speed, ms);
            if (!lower) max_speed_total = omax<double>(max_speed_total, ms);
        }
        if (have_last) {
            s0 = s0 + l0;
            s1 = s1 + l1;
            s2 = s2 + l2;
        }
        if (cfl_speed > 0.0) dtk = omin<double>(dtk, ra.x / cfl_speed);
        const double area = ra.y;
        const double inv_area = 1.0 / area;
        s0 *= inv_area;
        s1 *= inv_area;
        s2 *= inv_area;
        if (FUSED) {
            // gravity_wb (shallow_water_ext.c:1942-2037): the stage-gradient term
            // was formed by the previous distribute launch from the same values.
            s1 += gr.x;
            s2 += gr.y;
            s1 += -sidex / area;
            s2 += -sidey / area;
        }
        st_s(F.weu + k, s0);
        st_s(F.uheu + k, s1);
        st_s(F.vheu + k, s2);
        st_s(F.max_speed + k, max_speed_total);
    }
    const unsigned long long wm = wave_min_bits(dtk);
    if ((threadIdx.x & 63) == 0) F.wave_min[k >> 6] = wm;
}

// CFL minimum over the wavefront minima of k_flux: one workgroup of
// kReduceBlock threads per instance, over that instance's wavefronts.
// Unfused: sets flux_timestep only.  Fused: also update_timestep, the time
// advance and the end-of-step bookkeeping, unless the instance is halted.
constexpr int kReduceBlock = 1024;
template <bool FUSED>
__global__ __launch_bounds__(kReduceBlock) void k_timestep(DevMesh M, const unsigned long long *wave_min,
                                                           DevState *st_all, DevParams P, int mode) {
    const int inst = blockIdx.x;
    DevState *st = st_all + inst;
    // state and partial minima are requested together: one memory round trip
    DevState s = *st;
    const int w0 = M.inst_off[inst] >> 6, w1 = (M.inst_end[inst] + 63) >> 6;
    constexpr int U = 16;
    unsigned long long b = dbits(P.evolve_max_timestep);
    for (int base = w0; base < w1; base += U * kReduceBlock) {
        unsigned long long v[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int j = base + u * kReduceBlock + (int)threadIdx.x;
            v[u] = (j < w1) ? wave_min[j] : ~0ull;
        }
#pragma unroll
        for (int u = 0; u < U; u++) b = (v[u] < b) ? v[u] : b;
    }
    b = wave_min_bits(bitsd(b));
    __shared__ unsigned long long wmin[kReduceBlock / 64];
    if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = b;
    __syncthreads();
    if (threadIdx.x != 0) return;
    if (FUSED && (s.halt || (mode == kStageLast && s.error))) return;
#pragma unroll
    for (int w = 0; w < kReduceBlock / 64; w++) b = (wmin[w] < b) ? wmin[w] : b;
    s.flux_timestep = bitsd(b);
    if (FUSED) {
        if (mode != kStageLast) timestep_logic(&s, P);
        if (!(s.error & kErrSmallsteps)) {
            if (mode != kStageLast) s.time = s.time + s.timestep;
            if (mode != kStageFirst) finish_step(&s);
        }
    }
    *st = s;
}

// Fused flux of the mixed-precision loop: the k_flux<.., true> step with FP32
// edge state and geometry (16-byte records) and FP32 flux arithmetic; the
// depth tests, the explicit-update sums, gravity and the CFL candidates stay
// FP64.  The bed must be continuous across edges (k_check_bed), so a side's
// depth is its own edge depth.
template <int BLOCK = kBlock>
__global__ __launch_bounds__(BLOCK) void k_flux_mixed(DevMesh M, DevFields F, DevParams P, DevState *st_all,
                                                      int first) {
    const Lane ln = lane_of<BLOCK>(M);
    DevState *st = st_all + ln.inst;
    if (stage_skipped(st, first)) return;
    const int N = M.N;
    const int k = ln.k;
    double dtk = P.evolve_max_timestep;
    if (ln.valid) {
        EdgeRecF er[3];
        EdgeF own[3], nbe[3];
#pragma unroll
        for (int i = 0; i < 3; i++) er[i] = ld_edgef(M.edgef + i * N + k);
#pragma unroll
        for (int i = 0; i < 3; i++) own[i] = F.eqf[i * N + k];
        const int flags = ld_s(M.flags + k);
        const double2 ra = ld_s(M.ra + k);
        const double2 gr = ld_s(F.grav + k);
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const int e = er[i].nb;
            nbe[i] = F.eqf[e >= 0 ? e : i * N + k];  // boundary edge: its own record
        }
        const double eps = P.epsilon, g = P.g;
        const float gf = float(P.g), h0f = float(P.h0);
        const int last_edge = (flags >> 10) & 3;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, l0 = 0.0, l1 = 0.0, l2 = 0.0;
        bool have_last = false;
        double sidex = 0.0, sidey = 0.0, cfl_speed = 0.0, max_speed_total = 0.0;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const int e = er[i].nb;
            const bool lower = (flags >> i) & 1;
            {
                const double hh = double(own[i].h);
                const double fact = -0.5 * g * hh * hh * double(er[i].len);
                sidex = sidex + fact * double(er[i].nx);
                sidey = sidey + fact * double(er[i].ny);
            }
            float n1 = er[i].nx, n2 = er[i].ny, len = er[i].len;
            EdgeF L, R;
            if (e >= 0) {
                if (lower) {
                    L = nbe[i];
                    R = own[i];
                    if (M.own_normals) {
                        n1 = -n1;
                        n2 = -n2;
                    } else {
                        const EdgeRecF en = M.edgef[e];
                        n1 = en.nx;
                        n2 = en.ny;
                        len = en.len;
                    }
                } else {
                    L = own[i];
                    R = nbe[i];
                }
            } else {
                const int m = -e - 1;
                L = nbe[i];
                R = L;  // transmissive: inside edge state
                const int kind = M.bkind[m];
                if (kind == 0) {  // reflective (evaluate_segment.cu:36-63)
                    double b1, b2;
                    reflect_momentum(double(L.uh), double(L.vh), double(n1), double(n2), b1, b2);
                    R.uh = float(b1);
                    R.vh = float(b2);
                } else if (kind == 1) {  // Dirichlet constants
                    const double hb = M.bdir[3 * m] - F.ze[i * N + k];
                    R.h = float(hb);
                    R.cls = depth_class(hb, eps, P.limiting_threshold);
                    R.uh = float(M.bdir[3 * m + 1]);
                    R.vh = float(M.bdir[3 * m + 2]);
                }
            }
            if (P.optimise_dry_cells && L.cls == 0 && R.cls == 0) continue;
            double efd[3], ms;
            flux_mixed_core(L.h, R.h, L.cls, R.cls, R.h - L.h, L.uh, L.vh, R.uh, R.vh, n1, n2, gf, h0f, eps, efd,
                            ms);
            const double lend = double(len);
            double f0 = efd[0] * lend, f1 = efd[1] * lend, f2 = efd[2] * lend;
            if (!lower) {
                f0 = -f0;
                f1 = -f1;
                f2 = -f2;
            }
            if (i == last_edge) {
                l0 = f0;
                l1 = f1;
                l2 = f2;
                have_last = true;
            } else {
                s0 = s0 + f0;
                s1 = s1 + f1;
                s2 = s2 + f2;
            }
            if (((flags >> (3 + i)) & 1) && ms > eps) cfl_speed = omax<double>(cfl_speed, ms);
            if (!lower) max_speed_total = omax<double>(max_speed_total, ms);
        }
        if (have_last) {
            s0 = s0 + l0;
            s1 = s1 + l1;
            s2 = s2 + l2;
        }
        if (cfl_speed > 0.0) dtk = omin<double>(dtk, ra.x / cfl_speed);
        const double area = ra.y;
        const double inv_area = 1.0 / area;
        s0 *= inv_area;
        s1 *= inv_area;
        s2 *= inv_area;
        s1 += gr.x;
        s2 += gr.y;
        s1 += -sidex / area;
        s2 += -sidey / area;
        st_s(F.weu + k, s0);
        st_s(F.uheu + k, s1);
        st_s(F.vheu + k, s2);
        st_s(F.max_speed + k, max_speed_total);
    }
    const unsigned long long wm = wave_min_bits(dtk);
    if ((threadIdx.x & 63) == 0) F.wave_min[k >> 6] = wm;
}

// Mixed-precision edge state from the FP64 edge records (start of a mixed
// evolve call).
__global__ __launch_bounds__(kBlock) void k_pack_edges(DevMesh M, DevFields F, DevParams P) {
    const int N = M.N;
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= 3 * N) return;
    const Q4 e = F.eq[t];
    const double h = e.w - e.z;
    F.eqf[t] = EdgeF{float(h), float(e.uh), float(e.vh), depth_class(h, P.epsilon, P.limiting_threshold)};
}

// Bed continuity across shared edges (the mixed flux uses each side's own
// edge depth): flags the discontinuous-elevation error as k_flux does.
__global__ __launch_bounds__(kBlock) void k_check_bed(DevMesh M, DevFields F, DevState *st_all) {
    const int N = M.N;
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= 3 * N) return;
    const int e = M.edge[t].nb;
    if (e >= 0 && fabs(F.ze[t] - F.ze[e]) > 1.0e-10) {
        const int k = t % N;
        set_error(st_all + ((M.n_inst == 1) ? 0 : M.wave_inst[k >> 6]), kErrDiscontinuous);
    }
}

// gravity_wb as a separate operator (shallow_water_ext.c:1942-2037)
__global__ __launch_bounds__(kBlock) void k_gravity_wb(DevMesh M, DevFields F, DevParams P) {
    const int N = M.N;
    const int k = tri_index();
    if (k >= N) return;
    const double g = P.g;
    const Q4 c = ldq(F.cq + k);
    const double2 x0 = M.vc[k], x1 = M.vc[N + k], x2 = M.vc[2 * N + k];
    double wx, wy;
    gradient3(x0.x, x0.y, x1.x, x1.y, x2.x, x2.y, F.wv[k], F.wv[N + k], F.wv[2 * N + k], wx, wy);
    const double avg_h = c.w - c.z;
    double ex = F.uheu[k], ey = F.vheu[k];
    ex += -g * wx * avg_h;
    ey += -g * wy * avg_h;
    double sidex = 0.0, sidey = 0.0;
    for (int i = 0; i < 3; i++) {
        const Q4 eqi = ldq(F.eq + i * N + k);
        const double hh = eqi.w - eqi.z;
        const EdgeRec er = M.edge[i * N + k];
        const double fact = -0.5 * g * hh * hh * er.len;
        sidex = sidex + fact * er.nx;
        sidey = sidey + fact * er.ny;
    }
    const double area = M.ra[k].y;
    ex += -sidex / area;
    ey += -sidey / area;
    F.uheu[k] = ex;
    F.vheu[k] = ey;
}

// manning_friction_implicit (shallow_water_ext.c:1238-1317); accumulates into
// the semi-implicit updates like the C.
__global__ __launch_bounds__(kBlock) void k_manning(DevMesh M, DevFields F, DevParams P) {
    const int N = M.N;
    const int k = tri_index();
    if (k >= N) return;
    const double eta = F.friction[k], eps = P.minimum_allowed_height, g = P.g;
    if (!(eta > eps)) return;
    const double z0 = F.zv[k], z1 = F.zv[N + k], z2 = F.zv[2 * N + k];
    const Q4 c = ldq(F.cq + k);
    const double uh = c.uh, vh = c.vh;
    double zs = 1.0;
    if (P.use_sloped_mannings) {
        const double2 x0 = M.vc[k], x1 = M.vc[N + k], x2 = M.vc[2 * N + k];
        double zx, zy;
        gradient3(x0.x, x0.y, x1.x, x1.y, x2.x, x2.y, z0, z1, z2, zx, zy);
        zs = sqrt(1.0 + zx * zx + zy * zy);
    }
    const double z = (z0 + z1 + z2) / 3.0;
    const double h = c.w - z;
    if (h >= eps) {
        double S;
        if (P.use_sloped_mannings)
            S = -g * eta * eta * zs * sqrt((uh * uh + vh * vh));
        else
            S = -g * eta * eta * sqrt((uh * uh + vh * vh));
        S /= pow(h, 7.0 / 3);
        F.uhsi[k] += S * uh;
        F.vhsi[k] += S * vh;
    }
}


__global__ void k_update_timestep(DevState *st, DevParams P) { timestep_logic(st, P); }
__global__ void k_advance_time(DevState *st) { st->time = st->time + st->timestep; }
__global__ void k_finish_step(DevState *st) { finish_step(st); }

// Source terms of a step that depend only on the state the previous step left
// behind, formed as soon as that state is known (bit-identical to forming them
// at the start of the step: same inputs, same operations).
//
// Implicit Manning friction into the semi-implicit momentum updates
// (shallow_water_ext.c:1238-1317), from the centroid values.
SW_DEV void friction_si(const DevMesh &M, const DevFields &F, const DevParams &P, int k, const Q4 &c,
                        double &si_x, double &si_y) {
    si_x = 0.0;
    si_y = 0.0;
    const double g = P.g;
    const double eta = ld_s(F.friction + k);
    const double mah = P.minimum_allowed_height;
    if (!(eta > mah)) return;
    const double uh = c.uh, vh = c.vh;
    const double h = c.w - ld_s(F.zbar + k);
    if (h >= mah) {
        double S;
        if (P.use_sloped_mannings) {
            const int N = M.N;
            const double2 x0 = M.vc[k], x1 = M.vc[N + k], x2 = M.vc[2 * N + k];
            double zx, zy;
            gradient3(x0.x, x0.y, x1.x, x1.y, x2.x, x2.y, F.zv[k], F.zv[N + k], F.zv[2 * N + k], zx, zy);
            const double zs = sqrt(1.0 + zx * zx + zy * zy);
            S = -g * eta * eta * zs * sqrt((uh * uh + vh * vh));
        } else {
            S = -g * eta * eta * sqrt((uh * uh + vh * vh));
        }
        S /= pow(h, 7.0 / 3);
        si_x += S * uh;
        si_y += S * vh;
    }
}

// The stage-gradient part of gravity_wb (shallow_water_ext.c:1942-1990:
// -g * w_x * avg_h, w_x from the vertex stage), from the vertex stage values
// and the centroid depth avg_h = w_c - z_c.
SW_DEV double2 gravity_gradient(const double2 v[3], const double wv[3], double avg_h, double g) {
    double wx, wy;
    gradient3(v[0].x, v[0].y, v[1].x, v[1].y, v[2].x, v[2].y, wv[0], wv[1], wv[2], wx, wy);
    return make_double2(-g * wx * avg_h, -g * wy * avg_h);
}
SW_DEV double2 gravity_gradient_term(const double2 v[3], const double wv[3], const Q4 &c, double g) {
    return gravity_gradient(v, wv, c.w - c.z, g);
}

// saxpy_centroid_values (saxpy_centroid_values.cu:1-12) on the conserved
// quantities: q = a q + b q0
SW_DEV void saxpy_conserved(Q4 &c, const Q4 &c0, double a, double b) {
    c.w = a * c.w + b * c0.w;
    c.uh = a * c.uh + b * c0.uh;
    c.vh = a * c.vh + b * c0.vh;
}

// What k_update does besides the update in a fused step.
enum : int { kUpdateEuler = 0, kUpdateBackup = 1, kUpdateSaxpy = 2 };

// update_conserved_quantities (quantity.c:769-820), optionally followed by
// protect (the first stage of distribute_to_vertices_and_edges).  Fused
// Runge-Kutta stages: the first (kUpdateBackup) keeps the values it starts
// from in cq0 (backup_conserved_quantities); a later one (kUpdateSaxpy)
// combines its result with them, q = a q + b q0
// (saxpy_conserved_quantities), before protect.
template <bool FUSED, int BLOCK = kBlock>
__global__ __launch_bounds__(BLOCK) void k_update(DevMesh M, DevFields F, DevParams P, DevState *st_all, int mode,
                                                  double a, double b) {
    const Lane ln = lane_of<BLOCK>(M);
    DevState *st = st_all + ln.inst;
    if (FUSED && (st->halt || (st->error & (mode == kUpdateSaxpy ? ~0ll : kErrSmallsteps)))) return;
    const int k = ln.k;
    if (!ln.valid) return;
    const double dt = st->timestep;
    Q4 c = ldq(F.cq + k);
    if (FUSED && mode == kUpdateBackup) F.cq0[k] = c;
    double siw = ld_s(F.wsi + k), siu = ld_s(F.uhsi + k), siv = ld_s(F.vhsi + k);
    const bool stage_si = siw != 0.0;
    bool ok = update_one(c.w, ld_s(F.weu + k), siw, dt);
    ok = update_one(c.uh, ld_s(F.uheu + k), siu, dt) && ok;
    ok = update_one(c.vh, ld_s(F.vheu + k), siv, dt) && ok;
    if (!ok) set_error(st, kErrUpdate);
    if (FUSED && mode == kUpdateSaxpy) saxpy_conserved(c, ldq(F.cq0 + k), a, b);
    // a stage semi-implicit term is used once, as the update resets it
    // (quantity.c:769-820); nothing in the loop sets one, so this store only
    // happens after a caller uploaded one
    if (FUSED && stage_si) F.wsi[k] = 0.0;
    if (FUSED) {
        protect_one(c.w, c.z, c.uh, c.vh, P.minimum_allowed_height, P.maximum_allowed_speed, P.epsilon);
        // The centroid values are final for this step (distribute leaves them
        // alone), so the next step's Manning friction is formed here.
        double fx, fy;
        friction_si(M, F, P, k, c, fx, fy);
        st_s(F.uhsi + k, fx);
        st_s(F.vhsi + k, fy);
    } else {
        F.wsi[k] = 0.0;
        F.uhsi[k] = 0.0;
        F.vhsi[k] = 0.0;
    }
    F.cq[k] = c;
}

// backup_conserved_quantities (gpu_domain_advanced.py:1449-1475)
__global__ __launch_bounds__(kBlock) void k_backup(DevMesh M, DevFields F) {
    const int k = tri_index();
    if (k >= M.N) return;
    F.cq0[k] = ldq(F.cq + k);
}

// saxpy_conserved_quantities (gpu_domain_advanced.py:1481-1519)
__global__ __launch_bounds__(kBlock) void k_saxpy(DevMesh M, DevFields F, double a, double b) {
    const int k = tri_index();
    if (k >= M.N) return;
    Q4 c = ldq(F.cq + k);
    saxpy_conserved(c, ldq(F.cq0 + k), a, b);
    F.cq[k] = c;
}

__global__ __launch_bounds__(kBlock) void k_protect(DevMesh M, DevFields F, DevParams P) {
    const int k = tri_index();
    if (k >= M.N) return;
    Q4 c = ldq(F.cq + k);
    protect_one(c.w, c.z, c.uh, c.vh, P.minimum_allowed_height, P.maximum_allowed_speed, P.epsilon);
    F.cq[k] = c;
}

// _protect_swb2, the protect of the 'tsunami' flow algorithm
// (protect.cu:64-138): below a depth threshold that grows with the rise from
// the bed centroid to the highest bed edge midpoint (minimum_relative_height
// 0.1, at least minimum_allowed_height) the momenta become zero; a dry cell
// whose stage is below its lowest bed vertex less minimum_allowed_height has
// its stage and stage vertex values raised to that level.  max / min are the
// fmax / fmin the reference's kernel calls.  Each triangle reads and writes
// only its own values.
__global__ __launch_bounds__(kBlock) void k_protect_swb2(DevMesh M, DevFields F, DevParams P) {
    const int k = tri_index();
    const int N = M.N;
    if (k >= N) return;
    Q4 c = ldq(F.cq + k);
    double zv[3];
#pragma unroll
    for (int i = 0; i < 3; i++) zv[i] = ld_s(F.zv + i * N + k);
    const double mah = P.minimum_allowed_height;
    const double hc = c.w - c.z;
    const double bmax = 0.5 * fmax(zv[0] + zv[1], fmax(zv[1] + zv[2], zv[2] + zv[0]));
    if (!(hc < fmax(0.1 * (bmax - c.z), mah))) return;
    c.uh = 0.0;
    c.vh = 0.0;
    if (hc <= 0.0) {
        const double bmin = fmin(zv[0], fmin(zv[1], zv[2])) - mah;
        if (c.w < bmin) {
            c.w = fmax(c.w, bmin);
#pragma unroll
            for (int i = 0; i < 3; i++) F.wv[i * N + k] = fmax(F.wv[i * N + k], bmin);
        }
    }
    F.cq[k] = c;
}

// Centroid quantity as seen by the extrapolation: velocity when
// extrapolate_velocity_second_order (shallow_water_ext.c:2862-2874).
SW_DEV void centroid_wuv(const Q4 &c, const DevParams &P, double &w, double &u, double &v) {
    w = c.w;
    u = c.uh;
    v = c.vh;
    if (P.extrapolate_velocity_second_order == 1) {
        const double dk = omax<double>(c.w - c.z, P.minimum_allowed_height);
        u = u / dk;
        v = v / dk;
    }
}

// Inputs of the reconstruction of triangle k.  They are loaded in two batches
// (own data, then the surrogate neighbours' records) before any is used, so a
// wavefront waits for two memory round trips.
struct RecIn {
    Q4 c;           // own centroid record
    double2 xy;     // own centroid
    double2 v[3];   // vertex coordinates
    int sk[3];      // surrogate neighbours
    Q4 cs[3];       // their centroid records
    double2 xs[3];  // their centroids
    double zv[3];   // bed vertex values
    int flags;
};

SW_DEV void load_rec(const DevMesh &M, const DevFields &F, int k, bool second_order, RecIn &r) {
    const int N = M.N;
    r.c = ldq(F.cq + k);
#pragma unroll
    for (int i = 0; i < 3; i++) r.zv[i] = ld_s(F.zv + i * N + k);
    r.flags = ld_s(M.flags + k);
#pragma unroll
    for (int i = 0; i < 3; i++) r.v[i] = ld_s(M.vc + i * N + k);
    if (!second_order) return;
    r.xy = M.cc[k];
#pragma unroll
    for (int i = 0; i < 3; i++) r.sk[i] = ld_s(M.sur + i * N + k);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        r.cs[i] = ldq(F.cq + r.sk[i]);
        r.xs[i] = M.cc[r.sk[i]];
    }
}

// Previous vertex values of triangle k from the vertex arrays.
struct VertexArrays {
    const DevFields &F;
    int N, k;
    SW_DEV void operator()(int i, double &w, double &u, double &v) const {
        w = F.wv[i * N + k];
        u = F.uhv[i * N + k];
        v = F.vhv[i * N + k];
    }
};

// Second-order reconstruction of one conserved quantity of triangle k with the
// per-quantity limiters (extrapolate_second_order_and_limit_by_vertex /
// _by_edge, extrapolate_second_order_and_limit_by_vertex_or_edge.cu:694-800):
// the gradient from the surrogate neighbours' centroids (compute_gradients,
// :226-316; a boundary edge's surrogate is k itself; two boundary edges: the
// two-point gradient towards the one neighbour; three: none), the vertex
// values from the gradient (extrapolate_from_gradient, :172-224), then phi =
// min over the vertices (by_edge: the edge midpoints) of min(r beta, 1) against
// the range of qc and the neighbours' centroid values (:5-75, :77-170); by
// edge, the vertices are rebuilt as e_{i+1} + e_{i+2} - e_i.  qs are the
// surrogates' centroid values; the range over them equals the range over the
// true neighbours, a boundary surrogate being k (whose qc is in the range).
SW_DEV void reconstruct_limited(const RecIn &r, int k, int nob, double qc, const double qs[3], double beta,
                                bool by_edge, double qv[3]) {
    double a = 0.0, b = 0.0;
    if (nob < 2) {
        if (!(r.sk[0] == r.sk[1] || r.sk[1] == r.sk[2])) {
            const double x0 = r.xs[0].x, y0 = r.xs[0].y, x1 = r.xs[1].x, y1 = r.xs[1].y;
            const double x2 = r.xs[2].x, y2 = r.xs[2].y;
            const double q0 = qs[0], q1 = qs[1], q2 = qs[2];
            const double det = (y2 - y0) * (x1 - x0) - (y1 - y0) * (x2 - x0);
            a = ((y2 - y0) * (q1 - q0) - (y1 - y0) * (q2 - q0)) / det;
            b = ((x1 - x0) * (q2 - q0) - (x2 - x0) * (q1 - q0)) / det;
        }
    } else if (nob == 2) {
        // the one neighbour, chosen among values loaded beforehand: a
        // run-time index into the records (or a select of their addresses)
        // would put them in scratch memory
        const bool n0 = r.sk[0] != k, n1 = r.sk[1] != k, n2 = r.sk[2] != k;
        const double2 s0 = r.xs[0], s1 = r.xs[1], s2 = r.xs[2];
        const double qs0 = qs[0], qs1 = qs[1], qs2 = qs[2];
        if (n0 || n1 || n2) {
            const double x0 = n0 ? s0.x : n1 ? s1.x : s2.x;
            const double y0 = n0 ? s0.y : n1 ? s1.y : s2.y;
            const double x1 = r.xy.x, y1 = r.xy.y;
            const double q0 = n0 ? qs0 : n1 ? qs1 : qs2, q1 = qc;
            const double det = (x1 - x0) * (x1 - x0) + (y1 - y0) * (y1 - y0);
            a = (x1 - x0) * (q1 - q0) / det;
            b = (y1 - y0) * (q1 - q0) / det;
        }
    }
    const double x = r.xy.x, y = r.xy.y;
    double v[3], e[3];
#pragma unroll
    for (int i = 0; i < 3; i++) v[i] = qc + a * (r.v[i].x - x) + b * (r.v[i].y - y);
    e[0] = 0.5 * (v[1] + v[2]);
    e[1] = 0.5 * (v[2] + v[0]);
    e[2] = 0.5 * (v[0] + v[1]);
    double qmin = qc, qmax = qc;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        qmin = omin<double>(qmin, qs[i]);
        qmax = omax<double>(qmax, qs[i]);
    }
    double dqa[3], phi = 1.0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        double rr = 1.0;
        const double dq = (by_edge ? e[i] : v[i]) - qc;
        dqa[i] = dq;
        if (dq > 0.0) rr = (qmax - qc) / dq;
        if (dq < 0.0) rr = (qmin - qc) / dq;
        phi = omin<double>(omin<double>(rr * beta, 1.0), phi);
    }
    if (by_edge) {
#pragma unroll
        for (int i = 0; i < 3; i++) e[i] = qc + phi * dqa[i];
        qv[0] = e[1] + e[2] - e[0];
        qv[1] = e[2] + e[0] - e[1];
        qv[2] = e[0] + e[1] - e[2];
    } else {
#pragma unroll
        for (int i = 0; i < 3; i++) qv[i] = qc + phi * dqa[i];
    }
}

// The three conserved quantities with the per-quantity limiters.
SW_DEV void reconstruct_limited_all(const DevParams &P, int k, const RecIn &r, double wv[3], double uv[3],
                                    double vv[3]) {
    const int nob = nob_of(r.flags);
    const bool by_edge = P.limiter == 2;
    double qs[3];
#pragma unroll
    for (int i = 0; i < 3; i++) qs[i] = r.cs[i].w;
    reconstruct_limited(r, k, nob, r.c.w, qs, P.qbeta[0], by_edge, wv);
#pragma unroll
    for (int i = 0; i < 3; i++) qs[i] = r.cs[i].uh;
    reconstruct_limited(r, k, nob, r.c.uh, qs, P.qbeta[1], by_edge, uv);
#pragma unroll
    for (int i = 0; i < 3; i++) qs[i] = r.cs[i].vh;
    reconstruct_limited(r, k, nob, r.c.vh, qs, P.qbeta[2], by_edge, vv);
}

// extrapolate_second_order_sw for triangle k (shallow_water_ext.c:2717-3324),
// result in wv/uv/vv (momentum).  Returns false on the missing-neighbour error.
// stale(i, w, u, v) gives the previous vertex values of vertex i (a dry
// triangle keeps them).
template <typename R, typename Stale>
SW_DEV bool extrapolate_one(const DevParams &P, int k, const RecIn &r, double wv[3], double uv[3], double vv[3],
                            Stale stale) {
    const Q4 &ck = r.c;
    const int nob = nob_of(r.flags);
    double wc, uc, vc;
    centroid_wuv(ck, P, wc, uc, vc);
    bool ok = true;
    bool constant = false;
    if (nob == 3) {
        constant = true;
    } else {
        const double x = r.xy.x, y = r.xy.y;
        R dxv[3], dyv[3];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            dxv[i] = R(r.v[i].x - x);
            dyv[i] = R(r.v[i].y - y);
        }
        if (nob <= 1) {
            const double2 c0xy = r.xs[0], c1xy = r.xs[1], c2xy = r.xs[2];
            const double dx1 = c1xy.x - c0xy.x, dx2 = c2xy.x - c0xy.x;
            const double dy1 = c1xy.y - c0xy.y, dy2 = c2xy.y - c0xy.y;
            const double area2 = dy2 * dx1 - dy1 * dx2;
            if (area2 <= 0) {
                constant = true;
            } else {
                const Q4 &c0 = r.cs[0], &c1 = r.cs[1], &c2 = r.cs[2];
                double w0, u0, v0_, w1, u1, v1_, w2, u2, v2_;
                centroid_wuv(c0, P, w0, u0, v0_);
                centroid_wuv(c1, P, w1, u1, v1_);
                centroid_wuv(c2, P, w2, u2, v2_);
                const double hc = wc - ck.z;
                const double h0 = w0 - c0.z;
                const double h1 = w1 - c1.z;
                const double h2 = w2 - c2.z;
                const double hmin = omin<double>(omin<double>(h0, omin<double>(h1, h2)), hc);
                double hfactor = 0.0;
                if (hmin > 0.001) hfactor = (hmin - 0.001) / (hmin + 0.004);
                const double hmax = omax<double>(h0, omax<double>(h1, h2));
                if (P.optimise_dry_cells && hmax < P.epsilon) {
                    // dry: vertex values keep their previous contents
#pragma unroll
                    for (int i = 0; i < 3; i++) stale(i, wv[i], uv[i], vv[i]);
                } else {
                    const R inv_area2 = R(1.0) / R(area2);
                    R d[3];
                    R beta = R(P.beta_w_dry + (P.beta_w - P.beta_w_dry) * hfactor);
                    extrap_aux<R>(R(w0 - wc), R(w1 - w0), R(w2 - w0), R(dx1), R(dx2), R(dy1), R(dy2),
                                  inv_area2, dxv, dyv, beta, d);
                    wv[0] = wc + double(d[0]);
                    wv[1] = wc + double(d[1]);
                    wv[2] = wc + double(d[2]);
                    beta = R(P.beta_uh_dry + (P.beta_uh - P.beta_uh_dry) * hfactor);
                    extrap_aux<R>(R(u0 - uc), R(u1 - u0), R(u2 - u0), R(dx1), R(dx2), R(dy1), R(dy2),
                                  inv_area2, dxv, dyv, beta, d);
                    uv[0] = uc + double(d[0]);
                    uv[1] = uc + double(d[1]);
                    uv[2] = uc + double(d[2]);
                    beta = R(P.beta_vh_dry + (P.beta_vh - P.beta_vh_dry) * hfactor);
                    extrap_aux<R>(R(v0_ - vc), R(v1_ - v0_), R(v2_ - v0_), R(dx1), R(dx2), R(dy1), R(dy2),
                                  inv_area2, dxv, dyv, beta, d);
                    vv[0] = vc + double(d[0]);
                    vv[1] = vc + double(d[1]);
                    vv[2] = vc + double(d[2]);
                }
            }
        } else {
            // the first surrogate that is not k itself; the device copy of the
            // surrogate list holds it in slot 0
            if (r.sk[0] == k) {
                ok = false;
                constant = true;
            } else {
                const double2 c1xy = r.xs[0];
                const Q4 &c1 = r.cs[0];
                const double dx1 = c1xy.x - x, dy1 = c1xy.y - y;
                const double area2 = dx1 * dx1 + dy1 * dy1;
                double dx2 = 1.0 / area2;
                const double dy2 = dx2 * dy1;
                dx2 *= dx1;
                double w1, u1, v1_;
                centroid_wuv(c1, P, w1, u1, v1_);
                R d[3];
                const R bw = R(P.beta_w);
                extrap_one<R>(R(w1 - wc), R(dx2), R(dy2), dxv, dyv, bw, d);
                wv[0] = wc + double(d[0]);
                wv[1] = wc + double(d[1]);
                wv[2] = wc + double(d[2]);
                extrap_one<R>(R(u1 - uc), R(dx2), R(dy2), dxv, dyv, bw, d);
                uv[0] = uc + double(d[0]);
                uv[1] = uc + double(d[1]);
                uv[2] = uc + double(d[2]);
                extrap_one<R>(R(v1_ - vc), R(dx2), R(dy2), dxv, dyv, bw, d);
                vv[0] = vc + double(d[0]);
                vv[1] = vc + double(d[1]);
                vv[2] = vc + double(d[2]);
            }
        }
    }
    if (constant) {
        wv[0] = wv[1] = wv[2] = wc;
        uv[0] = uv[1] = uv[2] = uc;
        vv[0] = vv[1] = vv[2] = vc;
    }
    if (P.extrapolate_velocity_second_order == 1) {
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const double dv = omax<double>(wv[i] - r.zv[i], 0.);
            uv[i] = uv[i] * dv;
            vv[i] = vv[i] * dv;
        }
    }
    return ok;
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_extrapolate_second_order(DevMesh M, DevFields F, DevParams P,
                                                                     DevState *st) {
    const int N = M.N;
    const int k = tri_index();
    if (k >= N) return;
    RecIn r;
    load_rec(M, F, k, true, r);
    double wv[3], uv[3], vv[3];
    if (!extrapolate_one<R>(P, k, r, wv, uv, vv, VertexArrays{F, N, k})) set_error(st, kErrNeighbour);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        F.wv[i * N + k] = wv[i];
        F.uhv[i * N + k] = uv[i];
        F.vhv[i * N + k] = vv[i];
    }
}

// extrapolate_second_order_and_limit_by_vertex / _by_edge of the three
// conserved quantities (vertex and edge values; gpu_domain_advanced.py:1836-1859, 1941-1962)
__global__ __launch_bounds__(kBlock) void k_extrapolate_limited(DevMesh M, DevFields F, DevParams P) {
    const int N = M.N;
    const int k = tri_index();
    if (k >= N) return;
    RecIn r;
    load_rec(M, F, k, true, r);
    double wv[3], uv[3], vv[3];
    reconstruct_limited_all(P, k, r, wv, uv, vv);
    double ew[3], eu[3], ev[3];
    interpolate_one(wv, ew);
    interpolate_one(uv, eu);
    interpolate_one(vv, ev);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        F.wv[i * N + k] = wv[i];
        F.uhv[i * N + k] = uv[i];
        F.vhv[i * N + k] = vv[i];
        Q4 e = F.eq[i * N + k];
        e.w = ew[i];
        e.uh = eu[i];
        e.vh = ev[i];
        F.eq[i * N + k] = e;
    }
}

// extrapolate_first_order (extrapolate_first_order.cu:3-33) of the conserved quantities
__global__ __launch_bounds__(kBlock) void k_first_order_conserved(DevMesh M, DevFields F) {
    const int N = M.N;
    const int k = tri_index();
    if (k >= N) return;
    const Q4 c = ldq(F.cq + k);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        Q4 e = ldq(F.eq + i * N + k);
        e.w = c.w;
        e.uh = c.uh;
        e.vh = c.vh;
        F.eq[i * N + k] = e;
        F.wv[i * N + k] = c.w;
        F.uhv[i * N + k] = c.uh;
        F.vhv[i * N + k] = c.vh;
    }
}

__global__ __launch_bounds__(kBlock) void k_balance(DevMesh M, DevFields F, DevParams P) {
    const int N = M.N;
    const int k = tri_index();
    if (k >= N) return;
    double wv[3], zv[3], uv[3], vv[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        wv[i] = F.wv[i * N + k];
        zv[i] = F.zv[i * N + k];
        uv[i] = F.uhv[i * N + k];
        vv[i] = F.vhv[i * N + k];
    }
    const Q4 c = ldq(F.cq + k);
    BalanceParams bp{P.H0, P.alpha_balance, P.tight_slope_limiters, P.use_centroid_velocities};
    balance_one(c.w, c.z, wv, zv, c.uh, c.vh, uv, vv, bp);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        F.wv[i * N + k] = wv[i];
        F.uhv[i * N + k] = uv[i];
        F.vhv[i * N + k] = vv[i];
    }
}

// interpolate_from_vertices_to_edges for the three conserved quantities
__global__ __launch_bounds__(kBlock) void k_interpolate(DevMesh M, DevFields F) {
    const int N = M.N;
    const int k = tri_index();
    if (k >= N) return;
    double w[3], u[3], v[3], ew[3], eu[3], ev[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        w[i] = F.wv[i * N + k];
        u[i] = F.uhv[i * N + k];
        v[i] = F.vhv[i * N + k];
    }
    interpolate_one(w, ew);
    interpolate_one(u, eu);
    interpolate_one(v, ev);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        Q4 e = ldq(F.eq + i * N + k);
        e.w = ew[i];
        e.uh = eu[i];
        e.vh = ev[i];
        F.eq[i * N + k] = e;
    }
}

// Source terms of the first fused step of an evolve call.
__global__ __launch_bounds__(kBlock) void k_source_terms(DevMesh M, DevFields F, DevParams P) {
    const int N = M.N;
    const Lane ln = lane_of(M);
    const int k = ln.k;
    if (!ln.valid) return;
    const Q4 c = ldq(F.cq + k);
    const double2 v[3] = {M.vc[k], M.vc[N + k], M.vc[2 * N + k]};
    const double wv[3] = {F.wv[k], F.wv[N + k], F.wv[2 * N + k]};
    F.grav[k] = gravity_gradient_term(v, wv, c, P.g);
    double fx, fy;
    friction_si(M, F, P, k, c, fx, fy);
    // the forcing adds to what the caller may have set (zero otherwise)
    F.uhsi[k] += fx;
    F.vhsi[k] += fy;
}

// Fused distribute_to_vertices_and_edges after protect: reconstruction
// (second order, or first order when the evolve loop fell back to it),
// balance_deep_and_shallow and vertex->edge interpolation in registers.  The
// edge records are written whole; their bed value is recomputed from the bed
// vertex values when the bed edges are known to be that interpolation
// (st->bed_interp), else copied from the static bed edge array.  LIMITED
// selects the per-quantity limiters (P.limiter != SW_LIMITER_OPTIMISED) at
// compile time, so the default path carries none of their code.
template <typename R, bool FUSED, int BLOCK = kBlock, bool LIMITED = false>
__global__ __launch_bounds__(BLOCK) void k_distribute(DevMesh M, DevFields F, DevParams P, DevState *st_all) {
    const Lane ln = lane_of<BLOCK>(M);
    DevState *st = st_all + ln.inst;
    if (FUSED && (st->halt || (st->error & kErrSmallsteps))) return;
    const int N = M.N;
    const int k = ln.k;
    if (!ln.valid) return;
    const bool second_order = st->order != 1;
    RecIn r;
    load_rec(M, F, k, second_order, r);
    const Q4 &c = r.c;
    double wv[3], uv[3], vv[3];
    if (!second_order) {
        wv[0] = wv[1] = wv[2] = c.w;
        uv[0] = uv[1] = uv[2] = c.uh;
        vv[0] = vv[1] = vv[2] = c.vh;
    } else if (LIMITED) {
        reconstruct_limited_all(P, k, r, wv, uv, vv);
    } else {
        if (!extrapolate_one<R>(P, k, r, wv, uv, vv, VertexArrays{F, N, k})) set_error(st, kErrNeighbour);
    }
    BalanceParams bp{P.H0, P.alpha_balance, P.tight_slope_limiters, P.use_centroid_velocities};
    balance_one(c.w, c.z, wv, r.zv, c.uh, c.vh, uv, vv, bp);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        st_s(F.wv + i * N + k, wv[i]);
        st_s(F.uhv + i * N + k, uv[i]);
        st_s(F.vhv + i * N + k, vv[i]);
    }
    // next step's stage-gradient gravity term, from the final vertex stage
    if (FUSED) st_s(F.grav + k, gravity_gradient_term(r.v, wv, c, P.g));
    double ew[3], eu[3], ev[3], ez[3];
    interpolate_one(wv, ew);
    interpolate_one(uv, eu);
    interpolate_one(vv, ev);
    if (st->bed_interp) {
        interpolate_one(r.zv, ez);
    } else {
        ez[0] = F.ze[k];
        ez[1] = F.ze[N + k];
        ez[2] = F.ze[2 * N + k];
    }
    if (FUSED && !std::is_same<R, double>::value) {
        // mixed-precision loop: compact FP32 edge state for k_flux_mixed; the
        // FP64 edge records are rebuilt from the vertex values at the end of the call
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const double h = ew[i] - ez[i];
            F.eqf[i * N + k] = EdgeF{float(h), float(eu[i]), float(ev[i]),
                                     depth_class(h, P.epsilon, P.limiting_threshold)};
        }
        return;
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        Q4 e;
        e.w = ew[i];
        e.uh = eu[i];
        e.vh = ev[i];
        e.z = ez[i];
        F.eq[i * N + k] = e;
    }
}

// update_boundary (evaluate_segment.cu:3-119, set_boundary.cu:1-18).  Each
// branch only computes the boundary values; the stores are done once,
// unconditionally and in a fixed order (no divergent stores to different
// arrays for the compiler to merge across branches).
__global__ __launch_bounds__(kBlock) void k_update_boundary(DevMesh M, DevFields F) {
    const int m = blockIdx.x * kBlock + threadIdx.x;
    if (m >= M.Nb) return;
    const int vol = M.bcell[m], edge = M.bedge[m];
    const int e = edge * M.N + vol;
    const int kind = M.bkind[m];
    const Q4 q = ldq(F.eq + e);
    const double he = F.he[e], ue = F.ue[e], ve = F.ve[e];
    const EdgeRec er = M.edge[e];
    const double n1 = er.nx, n2 = er.ny;
    double wb = q.w, uhb = q.uh, vhb = q.vh, hb = he, ub = ue, vb = ve;  // transmissive
    if (kind == 0) {  // reflective
        reflect_momentum(q.uh, q.vh, n1, n2, uhb, vhb);
        const double r1 = ue * n1 + ve * n2;
        const double r2 = ue * n2 - ve * n1;
        ub = n1 * r1 - n2 * r2;
        vb = n2 * r1 + n1 * r2;
    } else if (kind == 1) {  // Dirichlet: conserved quantities fixed, others edge copies
        wb = M.bdir[3 * m];
        uhb = M.bdir[3 * m + 1];
        vhb = M.bdir[3 * m + 2];
    }
    F.zb[m] = q.z;
    F.wb[m] = wb;
    F.uhb[m] = uhb;
    F.vhb[m] = vhb;
    F.hb[m] = hb;
    F.ub[m] = ub;
    F.vb[m] = vb;
}

// update_other_quantities (update_centroids_of_velocities_and_height.cu:1-52
// + extrapolate_first_order of height / xvelocity / yvelocity)
__global__ __launch_bounds__(kBlock) void k_other_quantities(DevMesh M, DevFields F) {
    const int N = M.N;
    const int k = blockIdx.x * kBlock + threadIdx.x;
    const double H0 = 1.0e-8;
    if (k < N) {
        const Q4 c = ldq(F.cq + k);
        double h = c.w - c.z;
        if (h < 0) h = 0;
        const double factor = h / (h * h + H0);
        const double u = c.uh * factor, v = c.vh * factor;
        F.hc[k] = h;
        F.uc[k] = u;
        F.vc[k] = v;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            F.he[i * N + k] = h;
            F.hv[i * N + k] = h;
            F.ue[i * N + k] = u;
            F.uv[i * N + k] = u;
            F.ve[i * N + k] = v;
            F.vv[i * N + k] = v;
        }
    }
    if (k < M.Nb) {
        double h = F.wb[k] - F.zb[k];
        if (h < 0) h = 0;
        const double factor = h / (h * h + H0);
        F.hb[k] = h;
        F.ub[k] = F.uhb[k] * factor;
        F.vb[k] = F.vhb[k] * factor;
    }
}

__global__ __launch_bounds__(kBlock) void k_fill(double *p, long long n, double val) {
    const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = val;
}

// API transfers.  Host arrays are in the reference layout ([N] or [N][3]); the
// device keeps [N] records or slot-major [3][N] arrays.  `slots` is 1 for
// per-triangle arrays and 3 for per-edge / per-vertex arrays; `comp` selects a
// component of {w, uh, vh, z} records (or -1 for a plain double array).
// Host-order staging [n][slots] <-> device array of one instance.  A
// per-triangle array is slot-major with stride `stride` (the handle's device
// triangle count); the instance's triangles start at device index `off`.
// cal_of[d] is the caller's index of the instance's device triangle off + d,
// dev_of its inverse.  comp >= 0 addresses one component of the 32-byte
// {w, uh, vh, z} records.
__global__ __launch_bounds__(kBlock) void k_to_device(long long n, long long stride, long long off, int slots,
                                                      int comp, const int *cal_of, const double *src,
                                                      double *dst) {
    const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (t >= n * slots) return;
    const long long i = t / n, d = t - i * n;
    const long long pos = i * stride + off + d;
    const double v = src[(long long)cal_of[d] * slots + i];
    if (comp < 0)
        dst[pos] = v;
    else
        dst[4 * pos + comp] = v;
}
__global__ __launch_bounds__(kBlock) void k_to_host(long long n, long long stride, long long off, int slots,
                                                    int comp, const int *dev_of, const double *src, double *dst) {
    const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;  // host position
    if (t >= n * slots) return;
    const long long k = t / slots, i = t - k * slots;
    const long long pos = i * stride + off + dev_of[k];
    dst[t] = (comp < 0) ? src[pos] : src[4 * pos + comp];
}

// SWW output points.  Caller vertex 3k+j of an instance -> its device slot
// j*N + off + dev_of[k] in the [3][N] vertex arrays.
__global__ __launch_bounds__(kBlock) void k_sww_slots(int nv, long long N, long long off, const int *dev_of,
                                                      int *slot) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= nv) return;
    const int v = slot[t];
    slot[t] = (int)((v % 3) * N + off + dev_of[v / 3]);
}

// FP32 values of the SWW points out[4][P] (elevation, stage, xmomentum,
// ymomentum); lane t computes point order[t] (a permutation of 0 .. P-1).
// Point p is the mean of the vertex values in slot[start[p] ..
// start[p+1]), summed in that order and divided by the count
// (ANUGA quantity_ext.c _average_vertex_values, the smoothing of
// Quantity.get_vertex_values), then rounded to FP32; a unique-vertex point
// lists one slot.  Where stage - elevation < minimum_storable_height (an
// FP32 test, as the writer's) stage is stored as the elevation and the
// momenta as 0.  A point with no triangles gets zeros.
__global__ __launch_bounds__(kBlock) void k_sww_points(long long P, const int *order, const int *start,
                                                       const int *slot, const double *zv, const double *wv,
                                                       const double *uhv, const double *vhv, float min_storable,
                                                       float *out) {
    const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (t >= P) return;
    const long long p = order[t];
    const int a = start[p], b = start[p + 1];
    double sz = 0.0, sw = 0.0, su = 0.0, sv = 0.0;
    for (int i = a; i < b; i++) {
        const int s = slot[i];
        sz += zv[s];
        sw += wv[s];
        su += uhv[s];
        sv += vhv[s];
    }
    float z = 0.f, w = 0.f, u = 0.f, v = 0.f;
    if (b > a) {
        const double k = (double)(b - a);
        z = (float)(sz / k);
        w = (float)(sw / k);
        u = (float)(su / k);
        v = (float)(sv / k);
    }
    const bool storable = (w - z) >= min_storable;
    out[p] = z;
    out[P + p] = storable ? w : z;
    out[2 * P + p] = storable ? u : 0.f;
    out[3 * P + p] = storable ? v : 0.f;
}

// Mean bed, the check that the bed edge values are the interpolation of the
// bed vertex values and, given the neighbour words of the vertex-record step
// (nbw, else null), the check that both sides of every shared edge hold the
// same bed edge value, over device triangles [off, off + n) of one instance.
__global__ __launch_bounds__(kBlock) void k_bed_derived(int N, int off, int n, const double *zv, const double *ze,
                                                        double *zbar, const int *nbw, const int *tri_node,
                                                        const double *node_z, DevState *st) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    bool bad = false, disc = false;
    if (t < n) {
        const int k = off + t;
        const double v[3] = {zv[k], zv[N + k], zv[2 * N + k]};
        zbar[k] = (v[0] + v[1] + v[2]) / 3.0;
        double e[3];
        interpolate_one(v, e);
        bad = dbits(e[0]) != dbits(ze[k]) || dbits(e[1]) != dbits(ze[N + k]) || dbits(e[2]) != dbits(ze[2 * N + k]);
        if (nbw) {
            for (int i = 0; i < 3; i++) {
                const int w = nbw[i * N + k];  // (m << 29) | n, or negative on a boundary edge
                if (w >= 0) disc = disc || dbits(ze[i * N + k]) != dbits(ze[(w >> 29) * N + (w & 0x1FFFFFFF)]);
            }
            // every vertex slot of a node holds the node's value (k_node_bed
            // stored one of them): then the vertex-record step reads the bed at
            // the nodes
            for (int i = 0; i < 3; i++) disc = disc || dbits(v[i]) != dbits(node_z[tri_node[i * N + k]]);
        }
    }
    // one atomic per wavefront that found a mismatch
    if (__ballot(bad) != 0 && (threadIdx.x & 63) == 0) atomicAnd((unsigned long long *)&st->bed_interp, 0ull);
    if (__ballot(disc) != 0 && (threadIdx.x & 63) == 0) atomicAnd((unsigned long long *)&st->bed_cont, 0ull);
}

// The bed value of every node from one of its vertex slots (k_bed_derived
// then checks that all slots agree).
__global__ __launch_bounds__(kBlock) void k_node_bed(int N, int off, int n, const double *zv, const int *tri_node,
                                                     double *node_z) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n) return;
    const int k = off + t;
    for (int i = 0; i < 3; i++) node_z[tri_node[i * N + k]] = zv[i * N + k];
}

}  // namespace sw
