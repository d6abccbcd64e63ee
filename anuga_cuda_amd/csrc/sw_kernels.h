// HIP kernels of the shallow-water step for gfx950.
//
// Two families share the numerics in sw_device.h:
//   * operator kernels, one per overridden ANUGA operator (the drop-in API:
//     each mirrors one reference kernel / C routine and its array contract);
//   * fused step kernels used by evolve(): per Euler step
//       flux   = compute_fluxes + boundary evaluation + gravity_wb + Manning
//                friction + CFL minimum (wavefront/LDS reduction, one atomic
//                per workgroup),
//       tstep  = update_timestep and time bookkeeping (one lane),
//       update = update_conserved_quantities + protect,
//       distribute = extrapolate_second_order_sw + balance_deep_and_shallow +
//                interpolate_from_vertices_to_edges,
//     so the explicit / semi-implicit updates are the only intermediates that
//     round-trip through HBM and the host is not involved between steps.
//
// One lane per triangle; 256-lane workgroups (4 wavefronts of 64); the
// workgroup index is remapped so that each of the 8 XCDs works on one
// contiguous range of triangles and neighbour gathers of a locality-ordered
// mesh hit that XCD's own L2.
#pragma once

#include "sw_device.h"

namespace sw {

constexpr int kBlock = 256;

struct DevParams {
    double g, epsilon, H0, h0, limiting_threshold, minimum_allowed_height, maximum_allowed_speed;
    double beta_w, beta_w_dry, beta_uh, beta_uh_dry, beta_vh, beta_vh_dry;
    double alpha_balance, CFL, evolve_max_timestep, evolve_min_timestep;
    int optimise_dry_cells, extrapolate_velocity_second_order, tight_slope_limiters;
    int use_centroid_velocities, use_sloped_mannings, max_smallsteps, default_order;
};

struct DevMesh {
    int N, Nb;
    const int *nbr, *nbr_edge, *sur, *nob, *full;
    const double *normals, *len, *radii, *areas, *cc, *vc;
    const int *bcell, *bedge, *bkind;
    const double *bdir;
};

struct DevQ {
    double *c, *e, *v, *b, *eu, *si;
};

struct DevFields {
    DevQ w, uh, vh, z, h, u, v;
    double *friction;
    double *max_speed;
    double *zbar;  // (z0 + z1 + z2) / 3 of the bed vertex values, precomputed
};

// Scalar evolve state kept in device memory.  The first block of fields has
// the layout of sw_state in include/anuga_sw.h.
struct DevState {
    double time, starttime, finaltime, yieldtime, yieldstep, timestep, flux_timestep;
    double recorded_max_timestep, recorded_min_timestep;
    long long smallsteps, order, number_of_steps, number_of_first_order_steps;
    long long has_finaltime, event, error;
    // device-only
    long long halt, halt_next, total_steps;
    unsigned long long flux_min_bits;
};

enum : long long { kErrNone = 0, kErrUpdate = 1, kErrSmallsteps = 2, kErrNeighbour = 4, kErrDiscontinuous = 8 };

SW_DEV int xcd_remap(int b, int nb) {
    int per = nb >> 3, rem = nb & 7, x = b & 7, idx = b >> 3;
    return x * per + (x < rem ? x : rem) + idx;
}

SW_DEV int tri_index() { return xcd_remap(blockIdx.x, gridDim.x) * kBlock + threadIdx.x; }

SW_DEV void set_error(DevState *st, long long code) {
    atomicOr((unsigned long long *)&st->error, (unsigned long long)code);
}

// ---------------------------------------------------------------------------
// compute_fluxes (shallow_water_ext.c:4519-4696).
//
// The C routine visits triangles in index order and computes every edge once,
// from the lower-numbered side, adding the flux to the neighbour with the
// opposite sign.  Here triangle k evaluates all three of its edges itself, but
// an edge shared with a lower-numbered neighbour n is evaluated exactly as the
// C evaluates it (left state from n, n's normal and length), and the sum is
// formed in the C's order: contributions of lower neighbours in ascending n,
// then k's own edges in local order, then the 1/area scaling.  The CFL
// candidates radii[k]/max_speed of all edges of k (gated by tri_full_flag of
// the computing triangle) give the same minimum as the C's edge-once loop.
//
// FUSED adds: boundary values evaluated on the fly (reflective / Dirichlet /
// transmissive), gravity_wb, Manning friction into the semi-implicit update,
// and the halt guard of the device-side evolve loop.
template <typename R, bool FUSED>
__global__ __launch_bounds__(kBlock) void k_flux(DevMesh M, DevFields F, DevParams P, DevState *st) {
    if (FUSED && st->halt_next) return;
    const int k = tri_index();
    double dtk = P.evolve_max_timestep;
    if (k < M.N) {
        int nb[3], lowerflag[3];
        double ef[3][3];
        bool have[3];
        double max_speed_total = 0.0;
        const double eps = P.epsilon;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const int ki = 3 * k + i;
            const int n = M.nbr[ki];
            nb[i] = n;
            const bool lower = (n >= 0 && n < k);
            lowerflag[i] = lower;
            double ql[3], qr[3], zl, zr, n1, n2, len;
            int c;
            if (lower) {
                const int nm = 3 * n + M.nbr_edge[ki];
                ql[0] = F.w.e[nm];
                ql[1] = F.uh.e[nm];
                ql[2] = F.vh.e[nm];
                zl = F.z.e[nm];
                qr[0] = F.w.e[ki];
                qr[1] = F.uh.e[ki];
                qr[2] = F.vh.e[ki];
                zr = F.z.e[ki];
                n1 = M.normals[2 * nm];
                n2 = M.normals[2 * nm + 1];
                len = M.len[nm];
                c = n;
            } else {
                ql[0] = F.w.e[ki];
                ql[1] = F.uh.e[ki];
                ql[2] = F.vh.e[ki];
                zl = F.z.e[ki];
                n1 = M.normals[2 * ki];
                n2 = M.normals[2 * ki + 1];
                len = M.len[ki];
                c = k;
                if (n < 0) {
                    const int m = -n - 1;
                    if (FUSED) {
                        const int kind = M.bkind[m];
                        if (kind == 0) {  // reflective (evaluate_segment.cu:36-63)
                            qr[0] = ql[0];
                            reflect_momentum(ql[1], ql[2], n1, n2, qr[1], qr[2]);
                        } else if (kind == 1) {  // Dirichlet constants
                            qr[0] = M.bdir[3 * m];
                            qr[1] = M.bdir[3 * m + 1];
                            qr[2] = M.bdir[3 * m + 2];
                        } else {  // transmissive: inside edge state
                            qr[0] = ql[0];
                            qr[1] = ql[1];
                            qr[2] = ql[2];
                        }
                    } else {
                        qr[0] = F.w.b[m];
                        qr[1] = F.uh.b[m];
                        qr[2] = F.vh.b[m];
                    }
                    zr = zl;
                } else {
                    const int nm = 3 * n + M.nbr_edge[ki];
                    qr[0] = F.w.e[nm];
                    qr[1] = F.uh.e[nm];
                    qr[2] = F.vh.e[nm];
                    zr = F.z.e[nm];
                }
            }
            have[i] = false;
            ef[i][0] = ef[i][1] = ef[i][2] = 0.0;
            if (P.optimise_dry_cells && fabs(ql[0] - zl) < eps && fabs(qr[0] - zr) < eps) continue;
            if (fabs(zl - zr) > 1.0e-10) set_error(st, kErrDiscontinuous);
            R qlr[3] = {R(ql[0]), R(ql[1]), R(ql[2])};
            R qrr[3] = {R(qr[0]), R(qr[1]), R(qr[2])};
            R efr[3], msr;
            flux_central<R>(qlr, qrr, R(zl), R(zr), R(n1), R(n2), R(eps), R(P.h0),
                            R(P.limiting_threshold), R(P.g), efr, msr);
            const double ms = double(msr);
            ef[i][0] = double(efr[0]) * len;
            ef[i][1] = double(efr[1]) * len;
            ef[i][2] = double(efr[2]) * len;
            have[i] = true;
            if (M.full[c] == 1 && ms > eps) dtk = omin<double>(dtk, M.radii[k] / ms);
            if (!lower) max_speed_total = omax<double>(max_speed_total, ms);
        }
        // order: lower neighbours by ascending index, then own edges 0,1,2
        int o0 = 0, o1 = 1, o2 = 2;
        auto key = [&](int i) { return lowerflag[i] ? nb[i] : 0x7fffffff; };
        if (key(o1) < key(o0)) { int t = o0; o0 = o1; o1 = t; }
        if (key(o2) < key(o1)) { int t = o1; o1 = o2; o2 = t; }
        if (key(o1) < key(o0)) { int t = o0; o0 = o1; o1 = t; }
        const int ord[3] = {o0, o1, o2};
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int i = ord[j];
            if (lowerflag[i] && have[i]) {
                s0 += ef[i][0];
                s1 += ef[i][1];
                s2 += ef[i][2];
            }
        }
#pragma unroll
        for (int i = 0; i < 3; i++) {
            if (!lowerflag[i] && have[i]) {
                s0 -= ef[i][0];
                s1 -= ef[i][1];
                s2 -= ef[i][2];
            }
        }
        const double inv_area = 1.0 / M.areas[k];
        s0 *= inv_area;
        s1 *= inv_area;
        s2 *= inv_area;
        if (FUSED) {
            // gravity_wb (shallow_water_ext.c:1942-2037)
            const int k3 = 3 * k, k6 = 6 * k;
            const double g = P.g;
            double wx, wy;
            gradient3(M.vc[k6], M.vc[k6 + 1], M.vc[k6 + 2], M.vc[k6 + 3], M.vc[k6 + 4],
                      M.vc[k6 + 5], F.w.v[k3], F.w.v[k3 + 1], F.w.v[k3 + 2], wx, wy);
            const double wc = F.w.c[k];
            const double avg_h = wc - F.z.c[k];
            s1 += -g * wx * avg_h;
            s2 += -g * wy * avg_h;
            double sidex = 0.0, sidey = 0.0;
#pragma unroll
            for (int i = 0; i < 3; i++) {
                const double hh = F.w.e[k3 + i] - F.z.e[k3 + i];
                const double fact = -0.5 * g * hh * hh * M.len[k3 + i];
                sidex = sidex + fact * M.normals[k6 + 2 * i];
                sidey = sidey + fact * M.normals[k6 + 2 * i + 1];
            }
            const double area = M.areas[k];
            s1 += -sidex / area;
            s2 += -sidey / area;
            // manning_friction_implicit (shallow_water_ext.c:1238-1317)
            double si_x = 0.0, si_y = 0.0;
            const double eta = F.friction[k];
            const double mah = P.minimum_allowed_height;
            if (eta > mah) {
                const double uh = F.uh.c[k], vh = F.vh.c[k];
                const double h = wc - F.zbar[k];
                if (h >= mah) {
                    double S;
                    if (P.use_sloped_mannings) {
                        double zx, zy;
                        gradient3(M.vc[k6], M.vc[k6 + 1], M.vc[k6 + 2], M.vc[k6 + 3], M.vc[k6 + 4],
                                  M.vc[k6 + 5], F.z.v[k3], F.z.v[k3 + 1], F.z.v[k3 + 2], zx, zy);
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
            F.uh.si[k] = si_x;
            F.vh.si[k] = si_y;
        }
        F.w.eu[k] = s0;
        F.uh.eu[k] = s1;
        F.vh.eu[k] = s2;
        F.max_speed[k] = max_speed_total;
    }
    block_min_atomic<kBlock>(dtk, &st->flux_min_bits);
}

// gravity_wb as a separate operator (shallow_water_ext.c:1942-2037)
__global__ __launch_bounds__(kBlock) void k_gravity_wb(DevMesh M, DevFields F, DevParams P) {
    const int k = tri_index();
    if (k >= M.N) return;
    const int k3 = 3 * k, k6 = 6 * k;
    const double g = P.g;
    double wx, wy;
    gradient3(M.vc[k6], M.vc[k6 + 1], M.vc[k6 + 2], M.vc[k6 + 3], M.vc[k6 + 4], M.vc[k6 + 5],
              F.w.v[k3], F.w.v[k3 + 1], F.w.v[k3 + 2], wx, wy);
    const double avg_h = F.w.c[k] - F.z.c[k];
    double ex = F.uh.eu[k], ey = F.vh.eu[k];
    ex += -g * wx * avg_h;
    ey += -g * wy * avg_h;
    double sidex = 0.0, sidey = 0.0;
    for (int i = 0; i < 3; i++) {
        const double hh = F.w.e[k3 + i] - F.z.e[k3 + i];
        const double fact = -0.5 * g * hh * hh * M.len[k3 + i];
        sidex = sidex + fact * M.normals[k6 + 2 * i];
        sidey = sidey + fact * M.normals[k6 + 2 * i + 1];
    }
    const double area = M.areas[k];
    ex += -sidex / area;
    ey += -sidey / area;
    F.uh.eu[k] = ex;
    F.vh.eu[k] = ey;
}

// manning_friction_implicit (shallow_water_ext.c:1238-1317); accumulates into
// the semi-implicit updates like the C.
__global__ __launch_bounds__(kBlock) void k_manning(DevMesh M, DevFields F, DevParams P) {
    const int k = tri_index();
    if (k >= M.N) return;
    const double eta = F.friction[k], eps = P.minimum_allowed_height, g = P.g;
    if (!(eta > eps)) return;
    const int k3 = 3 * k, k6 = 6 * k;
    const double z0 = F.z.v[k3], z1 = F.z.v[k3 + 1], z2 = F.z.v[k3 + 2];
    const double uh = F.uh.c[k], vh = F.vh.c[k];
    double zs = 1.0;
    if (P.use_sloped_mannings) {
        double zx, zy;
        gradient3(M.vc[k6], M.vc[k6 + 1], M.vc[k6 + 2], M.vc[k6 + 3], M.vc[k6 + 4], M.vc[k6 + 5],
                  z0, z1, z2, zx, zy);
        zs = sqrt(1.0 + zx * zx + zy * zy);
    }
    const double z = (z0 + z1 + z2) / 3.0;
    const double h = F.w.c[k] - z;
    if (h >= eps) {
        double S;
        if (P.use_sloped_mannings)
            S = -g * eta * eta * zs * sqrt((uh * uh + vh * vh));
        else
            S = -g * eta * eta * sqrt((uh * uh + vh * vh));
        S /= pow(h, 7.0 / 3);
        F.uh.si[k] += S * uh;
        F.vh.si[k] += S * vh;
    }
}

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

__global__ void k_reset_flux_min(DevState *st, double evolve_max_timestep) {
    st->flux_min_bits = dbits(evolve_max_timestep);
}

__global__ void k_flux_finalize(DevState *st, double evolve_max_timestep) {
    st->flux_timestep = bitsd(st->flux_min_bits);
    st->flux_min_bits = dbits(evolve_max_timestep);
}

__global__ void k_update_timestep(DevState *st, DevParams P) { timestep_logic(st, P); }

// Unfused loop: time = initial_time + timestep, then counters and events.
__global__ void k_advance_time(DevState *st) { st->time = st->time + st->timestep; }
__global__ void k_finish_step(DevState *st) { finish_step(st); }

// Fused loop, one lane: flux minimum -> timestep -> time -> counters -> event.
// halt_next stops the next step's flux launch, halt the update/distribute
// launches of later steps; see sw_capi.hip for the launch sequence.
__global__ void k_step_timestep(DevState *st, DevParams P) {
    if (st->halt) return;
    if (st->halt_next || st->error) {
        st->halt = 1;
        st->halt_next = 1;
        return;
    }
    st->flux_timestep = bitsd(st->flux_min_bits);
    st->flux_min_bits = dbits(P.evolve_max_timestep);
    timestep_logic(st, P);
    if (st->error) {
        st->halt = 1;
        st->halt_next = 1;
        return;
    }
    const double initial_time = st->time;
    st->time = initial_time + st->timestep;
    finish_step(st);
    if (st->event) st->halt_next = 1;
}

// update_conserved_quantities (quantity.c:769-820), optionally followed by
// protect (the first stage of distribute_to_vertices_and_edges).
template <bool FUSED>
__global__ __launch_bounds__(kBlock) void k_update(DevMesh M, DevFields F, DevParams P, DevState *st) {
    if (FUSED && st->halt) return;
    const int k = tri_index();
    if (k >= M.N) return;
    const double dt = st->timestep;
    double w = F.w.c[k], uh = F.uh.c[k], vh = F.vh.c[k];
    double siw = F.w.si[k], siu = F.uh.si[k], siv = F.vh.si[k];
    bool ok = update_one(w, F.w.eu[k], siw, dt);
    ok = update_one(uh, F.uh.eu[k], siu, dt) && ok;
    ok = update_one(vh, F.vh.eu[k], siv, dt) && ok;
    if (!ok) set_error(st, kErrUpdate);
    if (FUSED) {
        protect_one(w, F.z.c[k], uh, vh, P.minimum_allowed_height, P.maximum_allowed_speed, P.epsilon);
    } else {
        F.w.si[k] = 0.0;
        F.uh.si[k] = 0.0;
        F.vh.si[k] = 0.0;
    }
    F.w.c[k] = w;
    F.uh.c[k] = uh;
    F.vh.c[k] = vh;
}

__global__ __launch_bounds__(kBlock) void k_protect(DevMesh M, DevFields F, DevParams P) {
    const int k = tri_index();
    if (k >= M.N) return;
    double w = F.w.c[k], uh = F.uh.c[k], vh = F.vh.c[k];
    protect_one(w, F.z.c[k], uh, vh, P.minimum_allowed_height, P.maximum_allowed_speed, P.epsilon);
    F.w.c[k] = w;
    F.uh.c[k] = uh;
    F.vh.c[k] = vh;
}

// Centroid quantity as seen by the extrapolation: velocity when
// extrapolate_velocity_second_order (shallow_water_ext.c:2862-2874).
SW_DEV void centroid_wuv(const DevFields &F, const DevParams &P, int j, double &w, double &u,
                         double &v) {
    w = F.w.c[j];
    u = F.uh.c[j];
    v = F.vh.c[j];
    if (P.extrapolate_velocity_second_order == 1) {
        const double dk = omax<double>(w - F.z.c[j], P.minimum_allowed_height);
        u = u / dk;
        v = v / dk;
    }
}

// extrapolate_second_order_sw for triangle k (shallow_water_ext.c:2717-3324),
// result in wv/uv/vv (momentum).  Returns false on the missing-neighbour error.
template <typename R>
SW_DEV bool extrapolate_one(const DevMesh &M, const DevFields &F, const DevParams &P, int k,
                            double wv[3], double uv[3], double vv[3]) {
    const int k3 = 3 * k, k6 = 6 * k;
    const int nob = M.nob[k];
    double wc, uc, vc;
    centroid_wuv(F, P, k, wc, uc, vc);
    bool ok = true;
    bool constant = false;
    if (nob == 3) {
        constant = true;
    } else {
        const double x = M.cc[2 * k], y = M.cc[2 * k + 1];
        R dxv[3], dyv[3];
        dxv[0] = R(M.vc[k6] - x);
        dxv[1] = R(M.vc[k6 + 2] - x);
        dxv[2] = R(M.vc[k6 + 4] - x);
        dyv[0] = R(M.vc[k6 + 1] - y);
        dyv[1] = R(M.vc[k6 + 3] - y);
        dyv[2] = R(M.vc[k6 + 5] - y);
        if (nob <= 1) {
            const int k0 = M.sur[k3], k1 = M.sur[k3 + 1], k2 = M.sur[k3 + 2];
            const double x0 = M.cc[2 * k0], y0 = M.cc[2 * k0 + 1];
            const double x1 = M.cc[2 * k1], y1 = M.cc[2 * k1 + 1];
            const double x2 = M.cc[2 * k2], y2 = M.cc[2 * k2 + 1];
            const double dx1 = x1 - x0, dx2 = x2 - x0, dy1 = y1 - y0, dy2 = y2 - y0;
            const double area2 = dy2 * dx1 - dy1 * dx2;
            if (area2 <= 0) {
                constant = true;
            } else {
                double w0, u0, v0, w1, u1, v1, w2, u2, v2;
                centroid_wuv(F, P, k0, w0, u0, v0);
                centroid_wuv(F, P, k1, w1, u1, v1);
                centroid_wuv(F, P, k2, w2, u2, v2);
                const double hc = wc - F.z.c[k];
                const double h0 = w0 - F.z.c[k0];
                const double h1 = w1 - F.z.c[k1];
                const double h2 = w2 - F.z.c[k2];
                const double hmin = omin<double>(omin<double>(h0, omin<double>(h1, h2)), hc);
                double hfactor = 0.0;
                if (hmin > 0.001) hfactor = (hmin - 0.001) / (hmin + 0.004);
                const double hmax = omax<double>(h0, omax<double>(h1, h2));
                if (P.optimise_dry_cells && hmax < P.epsilon) {
                    // dry: vertex values keep their previous contents
                    for (int i = 0; i < 3; i++) {
                        wv[i] = F.w.v[k3 + i];
                        uv[i] = F.uh.v[k3 + i];
                        vv[i] = F.vh.v[k3 + i];
                    }
                } else {
                    const R inv_area2 = R(1.0) / R(area2);
                    R d[3];
                    R beta = R(P.beta_w_dry + (P.beta_w - P.beta_w_dry) * hfactor);
                    extrap_aux<R>(R(wc), R(w0), R(w1), R(w2), R(dx1), R(dx2), R(dy1), R(dy2),
                                  inv_area2, dxv, dyv, beta, d);
                    for (int i = 0; i < 3; i++) wv[i] = wc + double(d[i]);
                    beta = R(P.beta_uh_dry + (P.beta_uh - P.beta_uh_dry) * hfactor);
                    extrap_aux<R>(R(uc), R(u0), R(u1), R(u2), R(dx1), R(dx2), R(dy1), R(dy2),
                                  inv_area2, dxv, dyv, beta, d);
                    for (int i = 0; i < 3; i++) uv[i] = uc + double(d[i]);
                    beta = R(P.beta_vh_dry + (P.beta_vh - P.beta_vh_dry) * hfactor);
                    extrap_aux<R>(R(vc), R(v0), R(v1), R(v2), R(dx1), R(dx2), R(dy1), R(dy2),
                                  inv_area2, dxv, dyv, beta, d);
                    for (int i = 0; i < 3; i++) vv[i] = vc + double(d[i]);
                }
            }
        } else {
            int e = -1;
            for (int i = 0; i < 3; i++)
                if (e < 0 && M.sur[k3 + i] != k) e = i;
            if (e < 0) {
                ok = false;
                constant = true;
            } else {
                const int k1 = M.sur[k3 + e];
                const double dx1 = M.cc[2 * k1] - x, dy1 = M.cc[2 * k1 + 1] - y;
                const double area2 = dx1 * dx1 + dy1 * dy1;
                double dx2 = 1.0 / area2;
                const double dy2 = dx2 * dy1;
                dx2 *= dx1;
                double w1, u1, v1;
                centroid_wuv(F, P, k1, w1, u1, v1);
                R d[3];
                const R bw = R(P.beta_w);
                extrap_one<R>(R(wc), R(w1), R(dx2), R(dy2), dxv, dyv, bw, d);
                for (int i = 0; i < 3; i++) wv[i] = wc + double(d[i]);
                extrap_one<R>(R(uc), R(u1), R(dx2), R(dy2), dxv, dyv, bw, d);
                for (int i = 0; i < 3; i++) uv[i] = uc + double(d[i]);
                extrap_one<R>(R(vc), R(v1), R(dx2), R(dy2), dxv, dyv, bw, d);
                for (int i = 0; i < 3; i++) vv[i] = vc + double(d[i]);
            }
        }
    }
    if (constant) {
        for (int i = 0; i < 3; i++) {
            wv[i] = wc;
            uv[i] = uc;
            vv[i] = vc;
        }
    }
    if (P.extrapolate_velocity_second_order == 1) {
        for (int i = 0; i < 3; i++) {
            const double dv = omax<double>(wv[i] - F.z.v[k3 + i], 0.);
            uv[i] = uv[i] * dv;
            vv[i] = vv[i] * dv;
        }
    }
    return ok;
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_extrapolate_second_order(DevMesh M, DevFields F, DevParams P,
                                                                     DevState *st) {
    const int k = tri_index();
    if (k >= M.N) return;
    double wv[3], uv[3], vv[3];
    if (!extrapolate_one<R>(M, F, P, k, wv, uv, vv)) set_error(st, kErrNeighbour);
    for (int i = 0; i < 3; i++) {
        F.w.v[3 * k + i] = wv[i];
        F.uh.v[3 * k + i] = uv[i];
        F.vh.v[3 * k + i] = vv[i];
    }
}

// extrapolate_first_order (extrapolate_first_order.cu:3-33) for one quantity
__global__ __launch_bounds__(kBlock) void k_first_order(int N, const double *c, double *e, double *v) {
    const int k = tri_index();
    if (k >= N) return;
    const double q = c[k];
    for (int i = 0; i < 3; i++) {
        e[3 * k + i] = q;
        v[3 * k + i] = q;
    }
}

__global__ __launch_bounds__(kBlock) void k_balance(DevMesh M, DevFields F, DevParams P) {
    const int k = tri_index();
    if (k >= M.N) return;
    const int k3 = 3 * k;
    double wv[3], zv[3], uv[3], vv[3];
    for (int i = 0; i < 3; i++) {
        wv[i] = F.w.v[k3 + i];
        zv[i] = F.z.v[k3 + i];
        uv[i] = F.uh.v[k3 + i];
        vv[i] = F.vh.v[k3 + i];
    }
    BalanceParams bp{P.H0, P.alpha_balance, P.tight_slope_limiters, P.use_centroid_velocities};
    balance_one(F.w.c[k], F.z.c[k], wv, zv, F.uh.c[k], F.vh.c[k], uv, vv, bp);
    for (int i = 0; i < 3; i++) {
        F.w.v[k3 + i] = wv[i];
        F.uh.v[k3 + i] = uv[i];
        F.vh.v[k3 + i] = vv[i];
    }
}

// interpolate_from_vertices_to_edges for the three conserved quantities
__global__ __launch_bounds__(kBlock) void k_interpolate(DevMesh M, DevFields F) {
    const int k = tri_index();
    if (k >= M.N) return;
    const int k3 = 3 * k;
    double v[3], e[3];
    DevQ *qs[3] = {&F.w, &F.uh, &F.vh};
    for (int q = 0; q < 3; q++) {
        for (int i = 0; i < 3; i++) v[i] = qs[q]->v[k3 + i];
        interpolate_one(v, e);
        for (int i = 0; i < 3; i++) qs[q]->e[k3 + i] = e[i];
    }
}

// Fused distribute_to_vertices_and_edges after protect: reconstruction
// (second order, or first order when the evolve loop fell back to it),
// balance_deep_and_shallow and vertex->edge interpolation in registers.
template <typename R, bool FUSED>
__global__ __launch_bounds__(kBlock) void k_distribute(DevMesh M, DevFields F, DevParams P, DevState *st) {
    if (FUSED && st->halt) return;
    const int k = tri_index();
    if (k >= M.N) return;
    const int k3 = 3 * k;
    double wv[3], uv[3], vv[3], zv[3];
    const double wc = F.w.c[k], uc = F.uh.c[k], vc = F.vh.c[k];
    if (st->order == 1) {
        for (int i = 0; i < 3; i++) {
            wv[i] = wc;
            uv[i] = uc;
            vv[i] = vc;
        }
    } else {
        if (!extrapolate_one<R>(M, F, P, k, wv, uv, vv)) set_error(st, kErrNeighbour);
    }
    for (int i = 0; i < 3; i++) zv[i] = F.z.v[k3 + i];
    BalanceParams bp{P.H0, P.alpha_balance, P.tight_slope_limiters, P.use_centroid_velocities};
    balance_one(wc, F.z.c[k], wv, zv, uc, vc, uv, vv, bp);
    double e[3];
    for (int i = 0; i < 3; i++) {
        F.w.v[k3 + i] = wv[i];
        F.uh.v[k3 + i] = uv[i];
        F.vh.v[k3 + i] = vv[i];
    }
    interpolate_one(wv, e);
    for (int i = 0; i < 3; i++) F.w.e[k3 + i] = e[i];
    interpolate_one(uv, e);
    for (int i = 0; i < 3; i++) F.uh.e[k3 + i] = e[i];
    interpolate_one(vv, e);
    for (int i = 0; i < 3; i++) F.vh.e[k3 + i] = e[i];
}

// update_boundary (evaluate_segment.cu:3-119, set_boundary.cu:1-18).  Each
// branch only computes the six boundary values; the stores are done once,
// unconditionally and in a fixed order (no divergent stores to different
// arrays for the compiler to merge across branches).
__global__ __launch_bounds__(kBlock) void k_update_boundary(DevMesh M, DevFields F) {
    const int m = blockIdx.x * kBlock + threadIdx.x;
    if (m >= M.Nb) return;
    const int vol = M.bcell[m], edge = M.bedge[m];
    const int e = 3 * vol + edge;
    const int kind = M.bkind[m];
    const double we = F.w.e[e], uhe = F.uh.e[e], vhe = F.vh.e[e];
    const double he = F.h.e[e], ue = F.u.e[e], ve = F.v.e[e];
    const double n1 = M.normals[6 * vol + 2 * edge], n2 = M.normals[6 * vol + 2 * edge + 1];
    double wb = we, uhb = uhe, vhb = vhe, hb = he, ub = ue, vb = ve;  // transmissive
    if (kind == 0) {  // reflective
        reflect_momentum(uhe, vhe, n1, n2, uhb, vhb);
        const double r1 = ue * n1 + ve * n2;
        const double r2 = ue * n2 - ve * n1;
        ub = n1 * r1 - n2 * r2;
        vb = n2 * r1 + n1 * r2;
    } else if (kind == 1) {  // Dirichlet: conserved quantities fixed, others edge copies
        wb = M.bdir[3 * m];
        uhb = M.bdir[3 * m + 1];
        vhb = M.bdir[3 * m + 2];
    }
    F.z.b[m] = F.z.e[e];
    F.w.b[m] = wb;
    F.uh.b[m] = uhb;
    F.vh.b[m] = vhb;
    F.h.b[m] = hb;
    F.u.b[m] = ub;
    F.v.b[m] = vb;
}

// update_other_quantities (update_centroids_of_velocities_and_height.cu:1-52
// + extrapolate_first_order of height / xvelocity / yvelocity)
__global__ __launch_bounds__(kBlock) void k_other_quantities(DevMesh M, DevFields F) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    const double H0 = 1.0e-8;
    if (k < M.N) {
        double h = F.w.c[k] - F.z.c[k];
        if (h < 0) h = 0;
        const double factor = h / (h * h + H0);
        const double u = F.uh.c[k] * factor, v = F.vh.c[k] * factor;
        F.h.c[k] = h;
        F.u.c[k] = u;
        F.v.c[k] = v;
        for (int i = 0; i < 3; i++) {
            F.h.e[3 * k + i] = h;
            F.h.v[3 * k + i] = h;
            F.u.e[3 * k + i] = u;
            F.u.v[3 * k + i] = u;
            F.v.e[3 * k + i] = v;
            F.v.v[3 * k + i] = v;
        }
    }
    if (k < M.Nb) {
        double h = F.w.b[k] - F.z.b[k];
        if (h < 0) h = 0;
        const double factor = h / (h * h + H0);
        F.h.b[k] = h;
        F.u.b[k] = F.uh.b[k] * factor;
        F.v.b[k] = F.vh.b[k] * factor;
    }
}

__global__ __launch_bounds__(kBlock) void k_fill(double *p, long long n, double val) {
    const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = val;
}

}  // namespace sw
