// Vertex-record Euler step of the FP64 loop (extrapolate_second_order_sw path).
//
// Per step, three kernels:
//   k_flux_v    compute_fluxes + boundary values on the fly + gravity_wb (its
//               stage-gradient term from the vertex records) + the CFL
//               minimum (one non-returning atomic per workgroup into one of
//               kSlots slots of the instance);
//   k_update_v  update_timestep and the time bookkeeping (every wavefront
//               reduces the slots itself: no separate launch), Manning
//               friction of the centroid values the step starts from, then
//               update_conserved_quantities + protect, in place;
//   k_extrap_v  extrapolate_second_order_sw + balance_deep_and_shallow of the
//               updated centroid values into the vertex records, and the
//               centroid depth of the next step's gravity term.
//
// What goes through HBM per triangle-step (FP64; node arrays gathered through
// the node indices, about half a node per triangle):
//   k_flux_v    reads node indices 12, node coordinates and bed ~12,
//               neighbour words 12, flags 4, {inscribed radius, area} 16, the
//               triangle's three vertex records 72, centroid depth 8; writes
//               the update record 24 and max_speed 8 (168 B);
//   k_update_v  reads centroid record 32, update record 24, Manning n 8,
//               mean bed 8; writes centroid record 32 (104 B);
//   k_extrap_v  reads centroid record 32, flags 4, centroid 16, node indices
//               12, surrogates 12, node coordinates and bed ~12; writes vertex
//               records 72, centroid depth 8 (168 B).
// Neighbour gathers (vertex records in k_flux_v; centroid records and
// centroids in k_extrap_v; node arrays in both) come from the L2 / last-level
// cache: the locality order keeps neighbours in nearby workgroups and the
// nodes are numbered in the order the triangles first reach them.
//
// Differences from the edge-record step of sw_kernels.h, none of which changes
// a result bit:
//   * no edge records: k_flux_v forms edge values from the vertex records
//     (e_i = (v_{i+1} + v_{i+2}) / 2, the interpolation of quantity.c:687-707),
//     so the edge state is never written; the edge bed is the interpolation of
//     the bed vertex values (required: bed_interp), and every vertex slot of a
//     mesh node holds the node's bed value (required: bed_cont), so both sides
//     of a shared edge see the same bed and the bed is stored per node;
//   * normals and edge lengths are recomputed from the vertex coordinates by
//     tri_geometry, the formulas the mesh is built with (inscribed radius and
//     area are read: they cost more to form than to load); sw_create checks
//     on the host that tri_geometry reproduces the caller's arrays bit for
//     bit and otherwise keeps the edge-record step;
//   * the stage-gradient gravity term is formed by the flux from the vertex
//     stage and the centroid depth the previous k_extrap_v stored (the same
//     operations on the same values as forming it there);
//   * the scalar state and the CFL slots are read at parity p and written at
//     p ^ 1, p the step parity (a kernel argument: step i of a captured graph
//     uses parity i % 2); a halted step stores nothing.  Errors found by lanes
//     go to a separate word per instance (errw) that is merged into the state
//     at the end of each batch of steps.
#pragma once

#include "sw_kernels.h"

namespace sw {

// Properties of an instance's static fields that let k_update_v skip loads
// (checked bitwise by k_step_begin, passed to the kernel as a launch
// argument): every triangle has the first triangle's Manning n; every
// triangle's centroid bed equals the mean of its bed vertex values.
constexpr unsigned kUniFriction = 1, kUniMeanBed = 2;

// CFL slots per instance and parity, one 128-byte line apart.
constexpr int kSlots = 64;
constexpr int kSlotStride = 16;  // unsigned long longs between slots
constexpr int kSlotBlock = kSlots * kSlotStride;

// Vertex record {w, uh, vh} of one vertex and update record {stage, xmom,
// ymom explicit update} of one triangle (24 bytes each).
struct V3 {
    double w, uh, vh;
};
struct E3 {
    double w, uh, vh;
};

struct StepFields {
    V3 *vr;                   // [3][N] vertex records
    E3 *eur;                  // [N] explicit updates of the step (max_speed goes to DevFields::max_speed)
    double *hc;               // [N] centroid depth w - z after the step (for the next step's gravity term)
    unsigned long long *slot; // [2][n_inst][kSlotBlock] CFL minima (bits of positive doubles)
    long long *errw;          // [n_inst] errors found by lanes
    unsigned long long *uni;  // [n_inst] kUniFriction | kUniMeanBed where they hold (k_step_begin)
    DevState *stv[2];         // [n_inst] scalar state; stv[0] is the handle's state
};

// Triangle geometry from its vertex coordinates, by the formulas of the mesh
// builder: centroid ((x0 + x1) + x2) / 3; area |(x1 y0 - x0 y1) + (x2 y1 - x1
// y2) + (x0 y2 - x2 y0)| / 2; edge i from vertex i+1 to vertex i+2 with length
// l = sqrt(ex^2 + ey^2) and outward normal (ey / l, -(ex / l)); radius the
// smallest centroid-to-edge-midpoint distance.  Host and device evaluate the
// same operations (no FMA contraction), so sw_create can check the caller's
// arrays against it.
struct TriGeom {
    double nx[3], ny[3], len[3];
    double radius, area, cx, cy;
};
__host__ __device__ inline double geo_min(double a, double b) { return (b < a) ? b : a; }
__host__ __device__ inline void tri_geometry(const double x[3], const double y[3], TriGeom &g) {
    g.cx = ((x[0] + x[1]) + x[2]) / 3.0;
    g.cy = ((y[0] + y[1]) + y[2]) / 3.0;
    g.area = fabs(((x[1] * y[0] - x[0] * y[1]) + (x[2] * y[1] - x[1] * y[2])) + (x[0] * y[2] - x[2] * y[0])) / 2.0;
    double d[3];
    for (int i = 0; i < 3; i++) {
        const int a = (i + 1) % 3, b = (i + 2) % 3;
        const double ex = x[b] - x[a], ey = y[b] - y[a];
        const double l = sqrt(ex * ex + ey * ey);
        g.len[i] = l;
        g.nx[i] = ey / l;
        g.ny[i] = -(ex / l);
        const double mx = (x[a] + x[b]) / 2.0, my = (y[a] + y[b]) / 2.0;
        const double dx = g.cx - mx, dy = g.cy - my;
        d[i] = sqrt(dx * dx + dy * dy);
    }
    g.radius = geo_min(geo_min(d[0], d[1]), d[2]);
}

SW_DEV unsigned long long wave_min_u64(unsigned long long b) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(b, off, 64);
        b = (o < b) ? o : b;
    }
    return b;
}

// Slot ks of instance inst at parity p.
SW_DEV unsigned long long *cfl_slot(const StepFields &S, int n_inst, int p, int inst, int ks) {
    return S.slot + ((size_t)(p * n_inst + inst) * kSlotBlock + (size_t)ks * kSlotStride);
}

// The instance of workgroup (logical block) b and the bounds of its triangles.
struct BlockInst {
    int inst, off, end;
};
template <int BLOCK>
SW_DEV BlockInst block_inst(const DevMesh &M, int b) {
    const int k0 = b * BLOCK;
    if (M.n_inst == 1) return BlockInst{0, 0, M.N};
    const int inst = M.wave_inst[k0 >> 6];
    return BlockInst{inst, M.inst_off[inst], M.inst_end[inst]};
}

// Edge state {w, uh, vh} of the edge between two vertex records, and the bed
// of that edge from the two bed vertex values.
SW_DEV Q4 edge_of(const V3 &a, const V3 &b, double za, double zb) {
    return Q4{0.5 * (a.w + b.w), 0.5 * (a.uh + b.uh), 0.5 * (a.vh + b.vh), 0.5 * (za + zb)};
}
SW_DEV int next3(int i) { return i == 2 ? 0 : i + 1; }
SW_DEV int prev3(int i) { return i == 0 ? 2 : i - 1; }

// compute_fluxes (shallow_water_ext.c:4519-4696) with boundary values and
// gravity_wb, as k_flux<double, true>, from vertex records.  nbw[i*N + k]
// encodes the neighbour across edge i: (m << 29) | n for edge m of triangle n
// (N < 2^29), or -(b+1) for boundary edge b.
#ifndef SW_FLUX_MINWAVES
#define SW_FLUX_MINWAVES 1
#endif
template <int BLOCK>
__global__ __launch_bounds__(BLOCK, SW_FLUX_MINWAVES) void k_flux_v(DevMesh M, DevFields F, StepFields S, const int *nbw,
                                                  DevParams P, int p) {
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const BlockInst bi = block_inst<BLOCK>(M, b);
    const DevState *st = S.stv[p] + bi.inst;
    // a halted step stores nothing
    const bool halted = st->event || st->error || st->halt || S.errw[bi.inst];
    const int N = M.N;
    const int k = b * BLOCK + (int)threadIdx.x;
    double dtk = P.evolve_max_timestep;
    if (k < bi.end) {
        double x[3], y[3], zv[3];
        V3 vq[3], na[3], nc[3];
        int nb[3], nd[3];
#pragma unroll
        for (int i = 0; i < 3; i++) nd[i] = ld_s(at32(M.tri_node, (uint32_t)(i * N + k)));
#pragma unroll
        for (int i = 0; i < 3; i++) vq[i] = *at32(S.vr, (uint32_t)(i * N + k));
#pragma unroll
        for (int i = 0; i < 3; i++) nb[i] = ld_s(at32(nbw, (uint32_t)(i * N + k)));
        const int flags = ld_s(at32(M.flags, (uint32_t)k));
        const double avg_h = ld_s(at32(S.hc, (uint32_t)k));
        // inscribed radius and area (the caller's arrays, which sw_create
        // checked tri_geometry reproduces: reading them is cheaper than the
        // three square roots and two divisions that form them)
        const double2 ra = ld_s(at32(M.ra, (uint32_t)k));
        // the coordinates and bed of the triangle's nodes (every vertex slot
        // of a node holds its value, checked when the bed is uploaded)
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const double2 v = *at32(M.node_xy, (uint32_t)nd[i]);
            x[i] = v.x;
            y[i] = v.y;
            zv[i] = *at32(F.node_z, (uint32_t)nd[i]);
        }
        // the two vertex records of the neighbour's matching edge (a boundary
        // edge reads its own, which are not used)
#pragma unroll
        for (int i = 0; i < 3; i++) {
            int n = k, m = i;
            if (nb[i] >= 0) {
                n = nb[i] & 0x1FFFFFFF;
                m = nb[i] >> 29;
            }
            na[i] = *at32(S.vr, (uint32_t)(next3(m) * N + n));
            nc[i] = *at32(S.vr, (uint32_t)(prev3(m) * N + n));
        }
        TriGeom gm;
        tri_geometry(x, y, gm);
        // gravity_wb's stage-gradient term from this step's vertex stage
        const double2 v3[3] = {make_double2(x[0], y[0]), make_double2(x[1], y[1]), make_double2(x[2], y[2])};
        const double wv3[3] = {vq[0].w, vq[1].w, vq[2].w};
        const double2 gr = gravity_gradient(v3, wv3, avg_h, P.g);
        const double eps = P.epsilon, g = P.g;
        const int last_edge = (flags >> 10) & 3;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, l0 = 0.0, l1 = 0.0, l2 = 0.0;
        bool have_last = false;
        double sidex = 0.0, sidey = 0.0, cfl_speed = 0.0, max_speed_total = 0.0;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const Q4 own = edge_of(vq[next3(i)], vq[prev3(i)], zv[next3(i)], zv[prev3(i)]);
            const bool lower = (flags >> i) & 1;
            const double lk = gm.len[i];
            {
                const double hh = own.w - own.z;
                const double fact = -0.5 * g * hh * hh * lk;
                sidex = sidex + fact * gm.nx[i];
                sidey = sidey + fact * gm.ny[i];
            }
            double n1 = gm.nx[i], n2 = gm.ny[i];
            double ql[3], qr[3], zl, zr;
            if (nb[i] >= 0) {
                // both sides hold the same bed value on a shared edge (bed_cont)
                const Q4 nbq = edge_of(na[i], nc[i], zv[next3(i)], zv[prev3(i)]);
                if (lower) {
                    // the neighbour evaluates this edge in the C: its state on
                    // the left, its normal (the negation of ours), its length
                    ql[0] = nbq.w; ql[1] = nbq.uh; ql[2] = nbq.vh; zl = nbq.z;
                    qr[0] = own.w; qr[1] = own.uh; qr[2] = own.vh; zr = own.z;
                    n1 = -n1;
                    n2 = -n2;
                } else {
                    ql[0] = own.w; ql[1] = own.uh; ql[2] = own.vh; zl = own.z;
                    qr[0] = nbq.w; qr[1] = nbq.uh; qr[2] = nbq.vh; zr = nbq.z;
                }
            } else {
                const int m = -nb[i] - 1;
                ql[0] = own.w; ql[1] = own.uh; ql[2] = own.vh; zl = own.z;
                zr = zl;
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
            }
            if (P.optimise_dry_cells && fabs(ql[0] - zl) < eps && fabs(qr[0] - zr) < eps) continue;
            if (fabs(zl - zr) > 1.0e-10 && !halted)
                atomicOr((unsigned long long *)(S.errw + bi.inst), (unsigned long long)kErrDiscontinuous);
            double efd[3], ms;
            flux_central<double>(ql, qr, zl, zr, n1, n2, eps, P.h0, P.limiting_threshold, g, efd, ms);
            double f0 = efd[0] * lk, f1 = efd[1] * lk, f2 = efd[2] * lk;
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
        if (!halted) {
            *at32(S.eur, (uint32_t)k) = E3{s0, s1, s2};
            st_s(at32(F.max_speed, (uint32_t)k), max_speed_total);
        }
    }
    // workgroup minimum of the CFL candidates, one atomic into a slot
    __shared__ unsigned long long red[BLOCK / 64];
    const unsigned long long wm = wave_min_bits(dtk);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = wm;
    __syncthreads();
    if (threadIdx.x == 0 && !halted) {
        unsigned long long m = red[0];
#pragma unroll
        for (int w = 1; w < BLOCK / 64; w++) m = (red[w] < m) ? red[w] : m;
        atomicMin(cfl_slot(S, M.n_inst, p, bi.inst, b & (kSlots - 1)), m);
    }
}


// update_one + protect_one of one triangle's centroid record (quantity.c:769-820,
// shallow_water_ext.c:1638-1705); the momenta's semi-implicit terms come as
// si_ratio.  Returns false on the semi-implicit error.
SW_DEV bool advance_centroid(Q4 &c, const E3 &eu, double siw, const double2 &sr, double dt, const DevParams &P) {
    double sw = siw;
    bool ok = update_one(c.w, eu.w, sw, dt);
    ok = update_ratio(c.uh, eu.uh, sr.x, dt) && ok;
    ok = update_ratio(c.vh, eu.vh, sr.y, dt) && ok;
    protect_one(c.w, c.z, c.uh, c.vh, P.minimum_allowed_height, P.maximum_allowed_speed, P.epsilon);
    return ok;
}

// A later Runge-Kutta stage: the update, then q = a q + b q0 with the values
// the step started from (saxpy_conserved_quantities), then protect -- the
// edge-record step's order (k_update, kUpdateSaxpy).
SW_DEV bool advance_centroid_saxpy(Q4 &c, const Q4 &c0, const E3 &eu, double siw, const double2 &sr, double dt,
                                   double a, double b, const DevParams &P) {
    double sw = siw;
    bool ok = update_one(c.w, eu.w, sw, dt);
    ok = update_ratio(c.uh, eu.uh, sr.x, dt) && ok;
    ok = update_ratio(c.vh, eu.vh, sr.y, dt) && ok;
    saxpy_conserved(c, c0, a, b);
    protect_one(c.w, c.z, c.uh, c.vh, P.minimum_allowed_height, P.maximum_allowed_speed, P.epsilon);
    return ok;
}

// Manning friction of the next step (shallow_water_ext.c:1238-1317) from the
// final centroid record, with the geometry in registers; the operations of
// friction_si.
SW_DEV double2 friction_next(const DevParams &P, double eta, const double x[3], const double y[3],
                             const double zv[3], const Q4 &c) {
    double si_x = 0.0, si_y = 0.0;
    const double mah = P.minimum_allowed_height, g = P.g;
    if (eta > mah) {
        const double uh = c.uh, vh = c.vh;
        const double zbar = (zv[0] + zv[1] + zv[2]) / 3.0;
        const double h = c.w - zbar;
        if (h >= mah) {
            double S;
            if (P.use_sloped_mannings) {
                double zx, zy;
                gradient3(x[0], y[0], x[1], y[1], x[2], y[2], zv[0], zv[1], zv[2], zx, zy);
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
    return make_double2(si_x, si_y);
}

// friction_next on a flat bed (use_sloped_mannings off) with the mean bed
// value zbar = (z0 + z1 + z2) / 3 precomputed: the same operations.
SW_DEV double2 friction_flat(const DevParams &P, double eta, double zbar, const Q4 &c) {
    double si_x = 0.0, si_y = 0.0;
    const double mah = P.minimum_allowed_height, g = P.g;
    if (eta > mah) {
        const double uh = c.uh, vh = c.vh;
        const double h = c.w - zbar;
        if (h >= mah) {
            double S = -g * eta * eta * sqrt((uh * uh + vh * vh));
            S /= pow(h, 7.0 / 3);
            si_x += S * uh;
            si_y += S * vh;
        }
    }
    return make_double2(si_x, si_y);
}

// update_timestep and the time advance (k_timestep<true>, Euler step) from
// the CFL slots of the step, evaluated identically by every wavefront; the
// state after the step goes to parity p ^ 1.  stop: nothing advances (a
// halted step, or the too-small-timestep error).
struct StepStart {
    DevState s2;
    bool stop;
};
SW_DEV StepStart step_start(const StepFields &S, const DevParams &P, const DevState &s, long long errw,
                            unsigned long long slotv, int mode) {
    // a step starts unless a yield, final time or error is pending; a later
    // stage runs unless the step halted or failed (stage_skipped)
    const bool halted = (mode == kStepEuler || mode == kStageFirst) ? (s.event || s.error || s.halt || errw)
                                                                     : (s.halt || s.error || errw);
    StepStart r{s, halted};
    if (!halted && mode != kStageMid) {
        // the first stage takes the timestep and advances the time; the last
        // reports its flux timestep and ends the step (k_timestep's modes)
        r.s2.flux_timestep = bitsd(wave_min_u64(slotv));
        if (mode != kStageLast) {
            timestep_logic(&r.s2, P);
            r.stop = (r.s2.error & kErrSmallsteps) != 0;
            if (!r.stop) {
                r.s2.time = r.s2.time + r.s2.timestep;
                r.s2.csi_live = 0;
                r.s2.call_steps += 1;
            }
        }
        if (!r.stop && mode != kStageFirst) finish_step(&r.s2);
    }
    // written on every path: a store on the halted path alone would be merged
    // with finish_step's store of the event into one store through a chosen
    // address, which keeps the state in scratch memory
    r.s2.halt = halted ? 1 : r.s2.halt;
    return r;
}

// The step of the centroid values, in place: update_timestep and the time
// bookkeeping (every wavefront reduces the CFL slots itself), Manning friction
// (shallow_water_ext.c:1238-1317) from the centroid values the step starts
// from (what the previous step's update left; nothing changes them in
// between), then update_conserved_quantities + protect (quantity.c:769-820,
// shallow_water_ext.c:1638-1705).  On the first step of a call that starts
// from caller-set semi-implicit terms (csi_live) those are added, as the
// forcing adds friction to them.  Runge-Kutta stages (mode): the first keeps
// the values it starts from in cq0; later ones combine their update with them
// (q = a q + b q0) before protect; only the first takes a timestep and only
// the last ends the step.
#ifndef SW_UPD_MINWAVES
#define SW_UPD_MINWAVES 1
#endif
template <int BLOCK>
__global__ __launch_bounds__(BLOCK, SW_UPD_MINWAVES) void k_update_v(DevMesh M, DevFields F, StepFields S,
                                                                     DevParams P, int p, unsigned uni, int mode,
                                                                     double sa, double sb) {
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const BlockInst bi = block_inst<BLOCK>(M, b);
    const int lane = (int)threadIdx.x & 63;
    const bool lead_wave = (b * BLOCK == bi.off) && threadIdx.x < 64;  // first wavefront of the instance
    const int k = b * BLOCK + (int)threadIdx.x;
    const bool valid = k < bi.end;
    const int kk = valid ? k : bi.off;  // padding lanes load a real triangle's data and store nothing

    const unsigned long long slotv = *cfl_slot(S, M.n_inst, p, bi.inst, lane);
    Q4 c = ldq(F.cq + kk);
    const E3 eu = S.eur[kk];
    const double eta = (uni & kUniFriction) ? F.friction[bi.off] : ld_s(F.friction + kk);
    const double zbar = (uni & kUniMeanBed) ? c.z : ld_s(F.zbar + kk);

    // the next step's slots start at +inf (k_flux_v of the next step takes
    // its minima there after this launch has ended)
    if (lead_wave) *cfl_slot(S, M.n_inst, p ^ 1, bi.inst, lane) = ~0ull;
    const DevState s = S.stv[p][bi.inst];
    const StepStart st = step_start(S, P, s, S.errw[bi.inst], slotv, mode);
    if (lead_wave && lane == 0) S.stv[p ^ 1][bi.inst] = st.s2;
    if (!valid || st.stop) return;
    const double dt = st.s2.timestep;
    double2 fr;
    if (P.use_sloped_mannings) {
        // the bed's slope from the node coordinates and bed values
        double x[3], y[3], zv[3];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const int nd = M.tri_node[i * M.N + k];
            const double2 v = M.node_xy[nd];
            x[i] = v.x;
            y[i] = v.y;
            zv[i] = F.node_z[nd];
        }
        fr = friction_next(P, eta, x, y, zv, c);
    } else {
        fr = friction_flat(P, eta, zbar, c);
    }
    double siw = 0.0;
    if (s.csi_live) {
        siw = F.wsi[k];
        fr.x = F.uhsi[k] + fr.x;
        fr.y = F.vhsi[k] + fr.y;
    }
    const double2 sr = make_double2(si_ratio(fr.x, c.uh), si_ratio(fr.y, c.vh));
    bool ok;
    if (mode == kStageMid || mode == kStageLast) {
        ok = advance_centroid_saxpy(c, ldq(F.cq0 + k), eu, siw, sr, dt, sa, sb, P);
    } else {
        if (mode == kStageFirst) F.cq0[k] = c;  // backup_conserved_quantities
        ok = advance_centroid(c, eu, siw, sr, dt, P);
    }
    if (!ok) atomicOr((unsigned long long *)(S.errw + bi.inst), (unsigned long long)kErrUpdate);
    F.cq[k] = c;
}

// extrapolate_second_order_sw + balance_deep_and_shallow of the updated
// centroid values (shallow_water_ext.c:2717-3324, :1462-1636): the vertex
// records of the next step's flux and the centroid depth of its gravity term.
// LIMITED: the per-quantity limiters instead of the optimised one
// (extrapolate_second_order_and_limit_by_vertex_or_edge.cu), chosen at
// compile time as in k_distribute.
#ifndef SW_EXT_MINWAVES
#define SW_EXT_MINWAVES 1
#endif
template <int BLOCK, bool LIMITED = false>
__global__ __launch_bounds__(BLOCK, SW_EXT_MINWAVES) void k_extrap_v(DevMesh M, DevFields F, StepFields S,
                                                                     DevParams P, int p) {
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const BlockInst bi = block_inst<BLOCK>(M, b);
    const int N = M.N;
    const int k = b * BLOCK + (int)threadIdx.x;
    const bool valid = k < bi.end;
    const int kk = valid ? k : bi.off;

    // own inputs, then the surrogate neighbours' updated centroid records and
    // centroids and the nodes' coordinates and bed, before anything is used
    RecIn r;
    r.c = ldq(at32(F.cq, (uint32_t)kk));
    r.flags = ld_s(at32(M.flags, (uint32_t)kk));
    r.xy = *at32(M.cc, (uint32_t)kk);
    int nd[3];
#pragma unroll
    for (int i = 0; i < 3; i++) nd[i] = ld_s(at32(M.tri_node, (uint32_t)(i * N + kk)));
#pragma unroll
    for (int i = 0; i < 3; i++) r.sk[i] = ld_s(at32(M.sur, (uint32_t)(i * N + kk)));
#pragma unroll
    for (int i = 0; i < 3; i++) {
        r.cs[i] = ldq(at32(F.cq, (uint32_t)r.sk[i]));
        r.xs[i] = *at32(M.cc, (uint32_t)r.sk[i]);
    }
    double x[3], y[3], zv[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const double2 v = *at32(M.node_xy, (uint32_t)nd[i]);
        x[i] = v.x;
        y[i] = v.y;
        zv[i] = *at32(F.node_z, (uint32_t)nd[i]);
    }
    const DevState &s2 = S.stv[p ^ 1][bi.inst];  // the state after this step's update
    if (!valid || s2.halt || s2.error) return;
    const Q4 c = r.c;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        r.v[i] = make_double2(x[i], y[i]);
        r.zv[i] = zv[i];
    }
    double wv[3], uv[3], vv[3];
    bool ok_nb = true;
    if (s2.order == 1) {
        wv[0] = wv[1] = wv[2] = c.w;
        uv[0] = uv[1] = uv[2] = c.uh;
        vv[0] = vv[1] = vv[2] = c.vh;
    } else if (LIMITED) {
        // the per-quantity limiters (limit by vertex / by edge), each
        // quantity with its own beta
        reconstruct_limited_all(P, k, r, wv, uv, vv);
    } else {
        // a dry triangle keeps its previous vertex values
        auto stale = [&](int i, double &w, double &u, double &v) {
            const V3 o = S.vr[i * N + k];
            w = o.w;
            u = o.uh;
            v = o.vh;
        };
        ok_nb = extrapolate_one<double>(P, k, r, wv, uv, vv, stale);
    }
    BalanceParams bp{P.H0, P.alpha_balance, P.tight_slope_limiters, P.use_centroid_velocities};
    balance_one(c.w, c.z, wv, zv, c.uh, c.vh, uv, vv, bp);
    if (!ok_nb) atomicOr((unsigned long long *)(S.errw + bi.inst), (unsigned long long)kErrNeighbour);
#pragma unroll
    for (int i = 0; i < 3; i++) S.vr[i * N + k] = V3{wv[i], uv[i], vv[i]};
    st_s(S.hc + k, c.w - c.z);
}

// Start of an evolve call on the vertex-record step: the vertex records from
// the vertex arrays, the centroid depths, and whether caller-set semi-implicit
// terms are live (the first step adds them).
__global__ __launch_bounds__(kBlock) void k_step_begin(DevMesh M, DevFields F, StepFields S) {
    const int N = M.N;
    const Lane ln = lane_of(M);
    const int k = ln.k;
    if (!ln.valid) return;
    const Q4 c = ldq(F.cq + k);
    for (int i = 0; i < 3; i++) S.vr[i * N + k] = V3{F.wv[i * N + k], F.uhv[i * N + k], F.vhv[i * N + k]};
    S.hc[k] = c.w - c.z;
    const bool live = F.wsi[k] != 0.0 || F.uhsi[k] != 0.0 || F.vhsi[k] != 0.0;
    if (__ballot(live) != 0 && (threadIdx.x & 63) == 0)
        atomicOr((unsigned long long *)&S.stv[0][ln.inst].csi_live, 1ull);
    const int off = (M.n_inst == 1) ? 0 : M.inst_off[ln.inst];
    const bool eta_differs = dbits(F.friction[k]) != dbits(F.friction[off]);
    const bool bed_differs = dbits(F.zbar[k]) != dbits(c.z);
    unsigned long long clear = 0;
    if (__ballot(eta_differs) != 0) clear |= kUniFriction;
    if (__ballot(bed_differs) != 0) clear |= kUniMeanBed;
    if (clear && (threadIdx.x & 63) == 0) atomicAnd(S.uni + ln.inst, ~clear);
}

// Start of an evolve call whose first step has parity p, per instance: both
// parities' slots at +inf.  A fresh start (k_step_begin follows, p = 0)
// clears the caller semi-implicit flag (k_step_begin sets it again) and the
// count of steps run since the records were built; a call that continues
// from the records of the previous one keeps both, and its first step reads
// the handle's state (stv[0], which the caller may have changed between the
// calls) at parity p.
__global__ void k_step_reset(StepFields S, int n_inst, int p, int fresh) {
    const int inst = blockIdx.x;
    for (int q = 0; q < 2; q++) *cfl_slot(S, n_inst, q, inst, threadIdx.x) = ~0ull;
    if (threadIdx.x == 0) {
        if (fresh) {
            S.stv[0][inst].csi_live = 0;
            S.stv[0][inst].call_steps = 0;
            S.uni[inst] = kUniFriction | kUniMeanBed;  // k_step_begin clears what does not hold
        }
        if (p) S.stv[1][inst] = S.stv[0][inst];
    }
}

// End of a batch of steps: the lanes' errors are merged into the state of
// parity p (the next step's), and the handle's state (stv[0]) is a copy of it.
__global__ void k_step_settle(StepFields S, int n_inst, int p) {
    const int inst = blockIdx.x * blockDim.x + threadIdx.x;
    if (inst >= n_inst) return;
    DevState s = S.stv[p][inst];
    s.error |= S.errw[inst];
    S.errw[inst] = 0;
    S.stv[p][inst] = s;
    S.stv[0][inst] = s;
}

// The operator API's arrays from the records of the vertex-record step, whose
// next step has parity p (run when something reads or writes those arrays
// after one or more evolve calls): the current centroid records into buffer
// 0, and the vertex records and update records into the arrays of the
// operator API (vertex values; edge values as their interpolation,
// quantity.c:687-707; explicit updates; max_speed).  The records stay valid,
// so a later call continues from them at parity p.
__global__ __launch_bounds__(kBlock) void k_step_end(DevMesh M, DevFields F, StepFields S, int p) {
    const int N = M.N;
    const Lane ln = lane_of(M);
    const int k = ln.k;
    if (!ln.valid) return;
    const DevState &st = S.stv[0][ln.inst];
    V3 v[3];
    for (int i = 0; i < 3; i++) {
        v[i] = S.vr[i * N + k];
        F.wv[i * N + k] = v[i].w;
        F.uhv[i * N + k] = v[i].uh;
        F.vhv[i * N + k] = v[i].vh;
    }
    for (int i = 0; i < 3; i++) {
        const Q4 e = edge_of(v[next3(i)], v[prev3(i)], 0.0, 0.0);  // the bed edge value is kept
        Q4 o = F.eq[i * N + k];
        o.w = e.w;
        o.uh = e.uh;
        o.vh = e.vh;
        F.eq[i * N + k] = o;
    }
    // the first step's flux always runs (a call starts with no pending event)
    const E3 u = S.eur[k];
    F.weu[k] = u.w;
    F.uheu[k] = u.uh;
    F.vheu[k] = u.vh;
    // a caller-set stage term is used by one step (quantity.c:769-820);
    // call_steps counts the steps run since k_step_begin built the records
    if (st.call_steps > 0) F.wsi[k] = 0.0;
}

}  // namespace sw
