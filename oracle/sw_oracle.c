/* CPU FP64 oracle for the shallow-water evolve() step -- TEST INFRASTRUCTURE.
 *
 * Not part of the product: only tests/, __graft_entry__.smoke() and bench.py's
 * CPU-baseline leg load this library, always as the checker.  Build:
 * oracle/Makefile (gcc -O2 -ffp-contract=off -fno-fast-math, serial).
 *
 * Every routine restates the reference's FP64 C with the same operation order
 * (expressions are parenthesised exactly as C evaluates the originals), so that
 * a HIP kernel compiled without FMA contraction can match it bit for bit.
 * Reference files are cited relative to the reference repository root.
 */
#include "sw_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* util_ext.h:42-56 -- plain comparisons, not fmax/fmin */
static inline double omax(double x, double y) { return (x > y) ? x : y; }
static inline double omin(double x, double y) { return (x < y) ? x : y; }

/* util_ext.h:66-109 */
static void gradient3(double x0, double y0, double x1, double y1, double x2, double y2,
                      double q0, double q1, double q2, double *a, double *b) {
    double det = (y2 - y0) * (x1 - x0) - (y1 - y0) * (x2 - x0);
    double aa = (y2 - y0) * (q1 - q0) - (y1 - y0) * (q2 - q0);
    aa /= det;
    double bb = (x1 - x0) * (q2 - q0) - (x2 - x0) * (q1 - q0);
    bb /= det;
    *a = aa;
    *b = bb;
}

/* shallow_water_ext.c:214-240 */
static double compute_speed(double *uh, double *h, double epsilon, double h0,
                            double limiting_threshold) {
    double u;
    if (*h < limiting_threshold) {
        if (*h < epsilon) {
            *h = 0.0;
            u = 0.0;
        } else {
            u = *uh / (*h + h0 / *h);
        }
        *uh = u * *h;
    } else {
        u = *uh / *h;
    }
    return u;
}

/* shallow_water_ext.c:302-442 (_flux_function_central), continuous-bed
 * version: the |z_l - z_r| > 1e-10 check is done by the caller. */
void or_flux_function_central(const double *q_left, const double *q_right,
                              double z_left, double z_right, double n1, double n2,
                              double epsilon, double h0, double limiting_threshold,
                              double g, double *edgeflux, double *max_speed) {
    double ql[3], qr[3], fl[3], fr[3];
    /* _rotate (shallow_water_ext.c:36-54): q1' = n1 q1 + n2 q2, q2' = -n2 q1 + n1 q2 */
    ql[0] = q_left[0];
    ql[1] = n1 * q_left[1] + n2 * q_left[2];
    ql[2] = -n2 * q_left[1] + n1 * q_left[2];
    qr[0] = q_right[0];
    qr[1] = n1 * q_right[1] + n2 * q_right[2];
    qr[2] = -n2 * q_right[1] + n1 * q_right[2];

    double z = 0.5 * (z_left + z_right);

    double w_left = ql[0];
    double h_left = w_left - z;
    double uh_left = ql[1];
    double u_left = compute_speed(&uh_left, &h_left, epsilon, h0, limiting_threshold);

    double w_right = qr[0];
    double h_right = w_right - z;
    double uh_right = qr[1];
    double u_right = compute_speed(&uh_right, &h_right, epsilon, h0, limiting_threshold);

    double vh_left = ql[2];
    double vh_right = qr[2];
    compute_speed(&vh_left, &h_left, epsilon, h0, limiting_threshold);
    compute_speed(&vh_right, &h_right, epsilon, h0, limiting_threshold);

    double soundspeed_left = sqrt(g * h_left);
    double soundspeed_right = sqrt(g * h_right);

    double s_max = omax(u_left + soundspeed_left, u_right + soundspeed_right);
    if (s_max < 0.0) s_max = 0.0;
    double s_min = omin(u_left - soundspeed_left, u_right - soundspeed_right);
    if (s_min > 0.0) s_min = 0.0;

    fl[0] = u_left * h_left;
    fl[1] = u_left * uh_left + 0.5 * g * h_left * h_left;
    fl[2] = u_left * vh_left;
    fr[0] = u_right * h_right;
    fr[1] = u_right * uh_right + 0.5 * g * h_right * h_right;
    fr[2] = u_right * vh_right;

    double denom = s_max - s_min;
    if (denom < epsilon) {
        edgeflux[0] = 0.0;
        edgeflux[1] = 0.0;
        edgeflux[2] = 0.0;
        *max_speed = 0.0;
    } else {
        double inverse_denominator = 1.0 / denom;
        for (int i = 0; i < 3; i++) {
            edgeflux[i] = s_max * fl[i] - s_min * fr[i];
            edgeflux[i] += s_max * s_min * (qr[i] - ql[i]);
            edgeflux[i] *= inverse_denominator;
        }
        *max_speed = omax(fabs(s_max), fabs(s_min));
        /* _rotate(edgeflux, n1, -n2) */
        double e1 = edgeflux[1], e2 = edgeflux[2];
        double m2 = -n2;
        edgeflux[1] = n1 * e1 + m2 * e2;
        edgeflux[2] = -m2 * e1 + n1 * e2;
    }
}

/* shallow_water_ext.c:4519-4696 (_compute_fluxes_central_structure): each
 * edge computed once, from the lower-numbered triangle, and added to the
 * neighbour with opposite sign.  Returns flux_timestep. */
double or_compute_fluxes(or_domain *D) {
    const int64_t N = D->number_of_elements;
    double timestep;
    double h0 = D->H0 * D->H0;
    double limiting_threshold = 10 * D->H0;
    double ql[3], qr[3], edgeflux[3], max_speed = 0.0, zl, zr;

    D->flux_call++;
    const int64_t call = D->flux_call;
    timestep = D->evolve_max_timestep;

    memset(D->stage_explicit_update, 0, N * sizeof(double));
    memset(D->xmom_explicit_update, 0, N * sizeof(double));
    memset(D->ymom_explicit_update, 0, N * sizeof(double));

    for (int64_t k = 0; k < N; k++) {
        double max_speed_total = 0.0;
        for (int i = 0; i < 3; i++) {
            int64_t ki = k * 3 + i, nm = 0;
            if (D->already_computed_flux[ki] == call) continue;
            ql[0] = D->stage_edge_values[ki];
            ql[1] = D->xmom_edge_values[ki];
            ql[2] = D->ymom_edge_values[ki];
            zl = D->bed_edge_values[ki];
            int64_t n = D->neighbours[ki];
            if (n < 0) {
                int64_t m = -n - 1;
                qr[0] = D->stage_boundary_values[m];
                qr[1] = D->xmom_boundary_values[m];
                qr[2] = D->ymom_boundary_values[m];
                zr = zl;
            } else {
                int64_t m = D->neighbour_edges[ki];
                nm = n * 3 + m;
                qr[0] = D->stage_edge_values[nm];
                qr[1] = D->xmom_edge_values[nm];
                qr[2] = D->ymom_edge_values[nm];
                zr = D->bed_edge_values[nm];
            }
            if (D->optimise_dry_cells) {
                if (fabs(ql[0] - zl) < D->epsilon && fabs(qr[0] - zr) < D->epsilon) {
                    D->already_computed_flux[ki] = call;
                    if (n >= 0) D->already_computed_flux[nm] = call;
                    max_speed = 0.0;
                    continue;
                }
            }
            int64_t ki2 = 2 * ki;
            or_flux_function_central(ql, qr, zl, zr, D->normals[ki2], D->normals[ki2 + 1],
                                     D->epsilon, h0, limiting_threshold, D->g, edgeflux,
                                     &max_speed);
            double length = D->edgelengths[ki];
            edgeflux[0] *= length;
            edgeflux[1] *= length;
            edgeflux[2] *= length;

            D->stage_explicit_update[k] -= edgeflux[0];
            D->xmom_explicit_update[k] -= edgeflux[1];
            D->ymom_explicit_update[k] -= edgeflux[2];
            D->already_computed_flux[ki] = call;
            if (n >= 0) {
                D->stage_explicit_update[n] += edgeflux[0];
                D->xmom_explicit_update[n] += edgeflux[1];
                D->ymom_explicit_update[n] += edgeflux[2];
                D->already_computed_flux[nm] = call;
            }
            if (D->tri_full_flag[k] == 1) {
                if (max_speed > D->epsilon) {
                    timestep = omin(timestep, D->radii[k] / max_speed);
                    if (n >= 0) timestep = omin(timestep, D->radii[n] / max_speed);
                }
            }
            max_speed_total = omax(max_speed_total, max_speed);
        }
        double inv_area = 1.0 / D->areas[k];
        D->stage_explicit_update[k] *= inv_area;
        D->xmom_explicit_update[k] *= inv_area;
        D->ymom_explicit_update[k] *= inv_area;
        D->max_speed[k] = max_speed_total;
    }
    D->flux_timestep = timestep;
    return timestep;
}

/* shallow_water_ext.c:1942-2037 (gravity_wb, the 'wb_2' gravity term) */
void or_gravity_wb(or_domain *D) {
    const int64_t N = D->number_of_elements;
    const double g = D->g;
    for (int64_t k = 0; k < N; k++) {
        int64_t k3 = 3 * k, k6 = 6 * k;
        double w0 = D->stage_vertex_values[k3 + 0];
        double w1 = D->stage_vertex_values[k3 + 1];
        double w2 = D->stage_vertex_values[k3 + 2];
        const double *X = D->vertex_coordinates + k6;
        double wx, wy;
        gradient3(X[0], X[1], X[2], X[3], X[4], X[5], w0, w1, w2, &wx, &wy);
        double avg_h = D->stage_centroid_values[k] - D->bed_centroid_values[k];
        D->xmom_explicit_update[k] += -g * wx * avg_h;
        D->ymom_explicit_update[k] += -g * wy * avg_h;
        double hh[3];
        hh[0] = D->stage_edge_values[k3 + 0] - D->bed_edge_values[k3 + 0];
        hh[1] = D->stage_edge_values[k3 + 1] - D->bed_edge_values[k3 + 1];
        hh[2] = D->stage_edge_values[k3 + 2] - D->bed_edge_values[k3 + 2];
        double sidex = 0.0, sidey = 0.0;
        for (int i = 0; i < 3; i++) {
            double n0 = D->normals[k6 + 2 * i];
            double n1 = D->normals[k6 + 2 * i + 1];
            double fact = -0.5 * g * hh[i] * hh[i] * D->edgelengths[k3 + i];
            sidex = sidex + fact * n0;
            sidey = sidey + fact * n1;
        }
        double area = D->areas[k];
        D->xmom_explicit_update[k] += -sidex / area;
        D->ymom_explicit_update[k] += -sidey / area;
    }
}

/* shallow_water_ext.c:1238-1270 (_manning_friction_flat), called by ANUGA's
 * manning_friction_implicit with eps = minimum_allowed_height and the
 * semi-implicit updates as output (gpu_domain_advanced.py:1296-1310). */
void or_manning_friction_flat(or_domain *D) {
    const int64_t N = D->number_of_elements;
    const double g = D->g, eps = D->minimum_allowed_height;
    const double *w = D->stage_centroid_values, *zv = D->bed_vertex_values;
    const double *uh = D->xmom_centroid_values, *vh = D->ymom_centroid_values;
    const double *eta = D->friction_centroid_values;
    double *xmom = D->xmom_semi_implicit_update, *ymom = D->ymom_semi_implicit_update;
    for (int64_t k = 0; k < N; k++) {
        if (eta[k] > eps) {
            int64_t k3 = 3 * k;
            double z = (zv[k3 + 0] + zv[k3 + 1] + zv[k3 + 2]) / 3.0;
            double h = w[k] - z;
            if (h >= eps) {
                double S = -g * eta[k] * eta[k] * sqrt((uh[k] * uh[k] + vh[k] * vh[k]));
                S /= pow(h, 7.0 / 3);
                xmom[k] += S * uh[k];
                ymom[k] += S * vh[k];
            }
        }
    }
}

/* shallow_water_ext.c:1272-1317 (_manning_friction_sloped) */
void or_manning_friction_sloped(or_domain *D) {
    const int64_t N = D->number_of_elements;
    const double g = D->g, eps = D->minimum_allowed_height;
    const double *w = D->stage_centroid_values, *zv = D->bed_vertex_values;
    const double *uh = D->xmom_centroid_values, *vh = D->ymom_centroid_values;
    const double *eta = D->friction_centroid_values;
    double *xmom = D->xmom_semi_implicit_update, *ymom = D->ymom_semi_implicit_update;
    for (int64_t k = 0; k < N; k++) {
        if (eta[k] > eps) {
            int64_t k3 = 3 * k, k6 = 6 * k;
            double z0 = zv[k3 + 0], z1 = zv[k3 + 1], z2 = zv[k3 + 2];
            const double *X = D->vertex_coordinates + k6;
            double zx, zy;
            gradient3(X[0], X[1], X[2], X[3], X[4], X[5], z0, z1, z2, &zx, &zy);
            double zs = sqrt(1.0 + zx * zx + zy * zy);
            double z = (z0 + z1 + z2) / 3.0;
            double h = w[k] - z;
            if (h >= eps) {
                double S = -g * eta[k] * eta[k] * zs * sqrt((uh[k] * uh[k] + vh[k] * vh[k]));
                S /= pow(h, 7.0 / 3);
                xmom[k] += S * uh[k];
                ymom[k] += S * vh[k];
            }
        }
    }
}

/* gpu_domain_advanced.py:1390-1405: forcing terms = [manning_friction_implicit] */
void or_compute_forcing_terms(or_domain *D) {
    if (D->use_sloped_mannings)
        or_manning_friction_sloped(D);
    else
        or_manning_friction_flat(D);
}

/* quantity.c:769-820 (_update) */
int or_update(int64_t N, double timestep, double *centroid_values, double *explicit_update,
              double *semi_implicit_update) {
    for (int64_t k = 0; k < N; k++) {
        double x = centroid_values[k];
        if (x == 0.0)
            semi_implicit_update[k] = 0.0;
        else
            semi_implicit_update[k] /= x;
    }
    for (int64_t k = 0; k < N; k++) centroid_values[k] += timestep * explicit_update[k];
    for (int64_t k = 0; k < N; k++) {
        double denominator = 1.0 - timestep * semi_implicit_update[k];
        if (denominator <= 0.0) return OR_ERR_UPDATE;
        centroid_values[k] /= denominator;
    }
    memset(semi_implicit_update, 0, N * sizeof(double));
    return OR_OK;
}

/* gpu_domain_advanced.py:1412-1442: update each conserved quantity */
int or_update_conserved_quantities(or_domain *D) {
    const int64_t N = D->number_of_elements;
    int rc;
    if ((rc = or_update(N, D->timestep, D->stage_centroid_values, D->stage_explicit_update,
                        D->stage_semi_implicit_update)))
        return rc;
    if ((rc = or_update(N, D->timestep, D->xmom_centroid_values, D->xmom_explicit_update,
                        D->xmom_semi_implicit_update)))
        return rc;
    return or_update(N, D->timestep, D->ymom_centroid_values, D->ymom_explicit_update,
                     D->ymom_semi_implicit_update);
}

/* shallow_water_ext.c:1638-1705 (_protect) */
void or_protect(or_domain *D) {
    const int64_t N = D->number_of_elements;
    const double mah = D->minimum_allowed_height, mas = D->maximum_allowed_speed;
    double *wc = D->stage_centroid_values, *zc = D->bed_centroid_values;
    double *xmomc = D->xmom_centroid_values, *ymomc = D->ymom_centroid_values;
    if (mas < D->epsilon) {
        for (int64_t k = 0; k < N; k++) {
            double hc = wc[k] - zc[k];
            if (hc < mah) {
                xmomc[k] = 0.0;
                ymomc[k] = 0.0;
                if (hc <= 0.0) wc[k] = zc[k];
            }
        }
    } else {
        for (int64_t k = 0; k < N; k++) {
            double hc = wc[k] - zc[k];
            if (hc < mah) {
                if (hc <= 0.0) {
                    wc[k] = zc[k];
                    xmomc[k] = 0.0;
                    ymomc[k] = 0.0;
                } else {
                    double u = xmomc[k] / hc;
                    if (fabs(u) > mas) {
                        double reduced_speed = mas * u / fabs(u);
                        xmomc[k] = reduced_speed * hc;
                    }
                    double v = ymomc[k] / hc;
                    if (fabs(v) > mas) {
                        double reduced_speed = mas * v / fabs(v);
                        ymomc[k] = reduced_speed * hc;
                    }
                }
            }
        }
    }
}

/* protect.cu:64-138 (_protect_swb2): the protect of the 'tsunami' flow
 * algorithm (gpu_domain_advanced.py:1014-1031).  The depth threshold scales
 * with the rise from the bed centroid to the highest bed edge midpoint
 * (minimum_relative_height 0.1); below it the momenta become zero, and a cell
 * that is dry with its stage below the lowest bed vertex less
 * minimum_allowed_height gets stage and stage vertex values raised to that
 * level.  The kernel's max / min of doubles are CUDA's fmax / fmin; the mass
 * it adds is summed per thread and discarded, so it is not formed here. */
void or_protect_swb2(or_domain *D) {
    const int64_t N = D->number_of_elements;
    const double mah = D->minimum_allowed_height;
    const double minimum_relative_height = 0.1;
    double *wc = D->stage_centroid_values, *wv = D->stage_vertex_values;
    const double *zc = D->bed_centroid_values, *zv = D->bed_vertex_values;
    double *xmomc = D->xmom_centroid_values, *ymomc = D->ymom_centroid_values;
    for (int64_t k = 0; k < N; k++) {
        const double *z = zv + 3 * k;
        double hc = wc[k] - zc[k];
        double bmax = 0.5 * fmax(z[0] + z[1], fmax(z[1] + z[2], z[2] + z[0]));
        if (hc < fmax(minimum_relative_height * (bmax - zc[k]), mah)) {
            xmomc[k] = 0.0;
            ymomc[k] = 0.0;
            if (hc <= 0.0) {
                double bmin = fmin(z[0], fmin(z[1], z[2])) - mah;
                if (wc[k] < bmin) {
                    wc[k] = fmax(wc[k], bmin);
                    for (int i = 0; i < 3; i++) wv[3 * k + i] = fmax(wv[3 * k + i], bmin);
                }
            }
        }
    }
}

/* shallow_water_ext.c:98-149 */
static void find_qmin_and_qmax(double dq0, double dq1, double dq2, double *qmin, double *qmax) {
    if (dq0 >= 0.0) {
        if (dq1 >= dq2) {
            *qmax = (dq1 >= 0.0) ? dq0 + dq1 : dq0;
            *qmin = dq0 + dq2;
            if (*qmin >= 0.0) *qmin = 0.0;
        } else {
            *qmax = (dq2 > 0) ? dq0 + dq2 : dq0;
            *qmin = dq0 + dq1;
            if (*qmin >= 0.0) *qmin = 0.0;
        }
    } else {
        if (dq1 <= dq2) {
            *qmin = (dq1 < 0.0) ? dq0 + dq1 : dq0;
            *qmax = dq0 + dq2;
            if (*qmax <= 0.0) *qmax = 0.0;
        } else {
            *qmin = (dq2 < 0.0) ? dq0 + dq2 : dq0;
            *qmax = dq0 + dq1;
            if (*qmax <= 0.0) *qmax = 0.0;
        }
    }
}

/* shallow_water_ext.c:151-183 -- note r0 carries over between iterations */
static void limit_gradient(double *dqv, double qmin, double qmax, double beta_w) {
    double r = 1000.0, r0 = 1.0, phi;
    const double TINY = 1.0e-100;
    for (int i = 0; i < 3; i++) {
        if (dqv[i] < -TINY) r0 = qmin / dqv[i];
        if (dqv[i] > TINY) r0 = qmax / dqv[i];
        r = omin(r0, r);
    }
    phi = omin(r * beta_w, 1.0);
    dqv[0] = dqv[0] * phi;
    dqv[1] = dqv[1] * phi;
    dqv[2] = dqv[2] * phi;
}

/* One quantity on the auxiliary triangle (shallow_water_ext.c:3023-3062). */
static void extrap_aux(const double *qc, int64_t k, int64_t k0, int64_t k1, int64_t k2,
                       double dx1, double dx2, double dy1, double dy2, double inv_area2,
                       double dxv0, double dxv1, double dxv2, double dyv0, double dyv1,
                       double dyv2, double beta, double *qv) {
    double dq0 = qc[k0] - qc[k];
    double dq1 = qc[k1] - qc[k0];
    double dq2 = qc[k2] - qc[k0];
    double a = dy2 * dq1 - dy1 * dq2;
    a *= inv_area2;
    double b = dx1 * dq2 - dx2 * dq1;
    b *= inv_area2;
    double dqv[3], qmin, qmax;
    dqv[0] = a * dxv0 + b * dyv0;
    dqv[1] = a * dxv1 + b * dyv1;
    dqv[2] = a * dxv2 + b * dyv2;
    find_qmin_and_qmax(dq0, dq1, dq2, &qmin, &qmax);
    limit_gradient(dqv, qmin, qmax, beta);
    qv[0] = qc[k] + dqv[0];
    qv[1] = qc[k] + dqv[1];
    qv[2] = qc[k] + dqv[2];
}

/* One quantity toward the single interior neighbour (shallow_water_ext.c:3174-3213). */
static void extrap_one(const double *qc, int64_t k, int64_t k1, double dx2, double dy2,
                       double dxv0, double dxv1, double dxv2, double dyv0, double dyv1,
                       double dyv2, double beta, double *qv) {
    double dq1 = qc[k1] - qc[k];
    double a = dq1 * dx2;
    double b = dq1 * dy2;
    double dqv[3], qmin, qmax;
    dqv[0] = a * dxv0 + b * dyv0;
    dqv[1] = a * dxv1 + b * dyv1;
    dqv[2] = a * dxv2 + b * dyv2;
    if (dq1 >= 0.0) {
        qmin = 0.0;
        qmax = dq1;
    } else {
        qmin = dq1;
        qmax = 0.0;
    }
    limit_gradient(dqv, qmin, qmax, beta);
    qv[0] = qc[k] + dqv[0];
    qv[1] = qc[k] + dqv[1];
    qv[2] = qc[k] + dqv[2];
}

/* shallow_water_ext.c:2717-3324 (_extrapolate_second_order_sw) */
int or_extrapolate_second_order_sw(or_domain *D) {
    const int64_t N = D->number_of_elements;
    const double eps = D->epsilon, mah = D->minimum_allowed_height;
    double *wc = D->stage_centroid_values, *zc = D->bed_centroid_values;
    double *uc = D->xmom_centroid_values, *vc = D->ymom_centroid_values;
    double *wv = D->stage_vertex_values, *uv = D->xmom_vertex_values, *vv = D->ymom_vertex_values;
    const double *C = D->centroid_coordinates, *X = D->vertex_coordinates;
    const int64_t *sn = D->surrogate_neighbours, *nob = D->number_of_boundaries;
    double *xstore = NULL, *ystore = NULL;
    int rc = OR_OK;

    if (D->extrapolate_velocity_second_order == 1) {
        xstore = (double *)malloc(N * sizeof(double));
        ystore = (double *)malloc(N * sizeof(double));
        for (int64_t k = 0; k < N; k++) {
            double dk = omax(wc[k] - zc[k], mah);
            xstore[k] = uc[k];
            uc[k] = uc[k] / dk;
            ystore[k] = vc[k];
            vc[k] = vc[k] / dk;
        }
    }

    for (int64_t k = 0; k < N; k++) {
        int64_t k3 = 3 * k, k6 = 6 * k;
        if (nob[k] == 3) {
            wv[k3] = wv[k3 + 1] = wv[k3 + 2] = wc[k];
            uv[k3] = uv[k3 + 1] = uv[k3 + 2] = uc[k];
            vv[k3] = vv[k3 + 1] = vv[k3 + 2] = vc[k];
            continue;
        }
        double xv0 = X[k6], yv0 = X[k6 + 1], xv1 = X[k6 + 2], yv1 = X[k6 + 3];
        double xv2 = X[k6 + 4], yv2 = X[k6 + 5];
        double x = C[2 * k], y = C[2 * k + 1];
        double dxv0 = xv0 - x, dxv1 = xv1 - x, dxv2 = xv2 - x;
        double dyv0 = yv0 - y, dyv1 = yv1 - y, dyv2 = yv2 - y;

        if (nob[k] <= 1) {
            int64_t k0 = sn[k3], k1 = sn[k3 + 1], k2 = sn[k3 + 2];
            double x0 = C[2 * k0], y0 = C[2 * k0 + 1];
            double x1 = C[2 * k1], y1 = C[2 * k1 + 1];
            double x2 = C[2 * k2], y2 = C[2 * k2 + 1];
            double dx1 = x1 - x0, dx2 = x2 - x0, dy1 = y1 - y0, dy2 = y2 - y0;
            double area2 = dy2 * dx1 - dy1 * dx2;
            if (area2 <= 0) {
                wv[k3] = wv[k3 + 1] = wv[k3 + 2] = wc[k];
                uv[k3] = uv[k3 + 1] = uv[k3 + 2] = uc[k];
                vv[k3] = vv[k3 + 1] = vv[k3 + 2] = vc[k];
                continue;
            }
            double hc = wc[k] - zc[k];
            double h0 = wc[k0] - zc[k0];
            double h1 = wc[k1] - zc[k1];
            double h2 = wc[k2] - zc[k2];
            double hmin = omin(omin(h0, omin(h1, h2)), hc);
            double hfactor = 0.0;
            if (hmin > 0.001) hfactor = (hmin - 0.001) / (hmin + 0.004);
            if (D->optimise_dry_cells) {
                double hmax = omax(h0, omax(h1, h2));
                if (hmax < eps) continue; /* vertex values stay as they were */
            }
            double inv_area2 = 1.0 / area2;
            double beta;
            beta = D->beta_w_dry + (D->beta_w - D->beta_w_dry) * hfactor;
            extrap_aux(wc, k, k0, k1, k2, dx1, dx2, dy1, dy2, inv_area2, dxv0, dxv1, dxv2, dyv0,
                       dyv1, dyv2, beta, wv + k3);
            beta = D->beta_uh_dry + (D->beta_uh - D->beta_uh_dry) * hfactor;
            extrap_aux(uc, k, k0, k1, k2, dx1, dx2, dy1, dy2, inv_area2, dxv0, dxv1, dxv2, dyv0,
                       dyv1, dyv2, beta, uv + k3);
            beta = D->beta_vh_dry + (D->beta_vh - D->beta_vh_dry) * hfactor;
            extrap_aux(vc, k, k0, k1, k2, dx1, dx2, dy1, dy2, inv_area2, dxv0, dxv1, dxv2, dyv0,
                       dyv1, dyv2, beta, vv + k3);
        } else {
            int64_t e;
            for (e = k3; e < k3 + 3; e++)
                if (sn[e] != k) break;
            if (e == k3 + 3) {
                rc = OR_ERR_NEIGHBOUR;
                break;
            }
            int64_t k1 = sn[e];
            double x1 = C[2 * k1], y1 = C[2 * k1 + 1];
            double dx1 = x1 - x, dy1 = y1 - y;
            double area2 = dx1 * dx1 + dy1 * dy1;
            double dx2 = 1.0 / area2;
            double dy2 = dx2 * dy1;
            dx2 *= dx1;
            extrap_one(wc, k, k1, dx2, dy2, dxv0, dxv1, dxv2, dyv0, dyv1, dyv2, D->beta_w, wv + k3);
            extrap_one(uc, k, k1, dx2, dy2, dxv0, dxv1, dxv2, dyv0, dyv1, dyv2, D->beta_w, uv + k3);
            extrap_one(vc, k, k1, dx2, dy2, dxv0, dxv1, dxv2, dyv0, dyv1, dyv2, D->beta_w, vv + k3);
        }
    }

    if (D->extrapolate_velocity_second_order == 1) {
        const double *zv = D->bed_vertex_values;
        for (int64_t k = 0; k < N; k++) {
            int64_t k3 = 3 * k;
            double dv0 = omax(wv[k3] - zv[k3], 0.);
            double dv1 = omax(wv[k3 + 1] - zv[k3 + 1], 0.);
            double dv2 = omax(wv[k3 + 2] - zv[k3 + 2], 0.);
            uc[k] = xstore[k];
            uv[k3] = uv[k3] * dv0;
            uv[k3 + 1] = uv[k3 + 1] * dv1;
            uv[k3 + 2] = uv[k3 + 2] * dv2;
            vc[k] = ystore[k];
            vv[k3] = vv[k3] * dv0;
            vv[k3 + 1] = vv[k3 + 1] * dv1;
            vv[k3 + 2] = vv[k3 + 2] * dv2;
        }
        free(xstore);
        free(ystore);
    }
    return rc;
}

/* extrapolate_first_order.c:8-40: edge = vertex = centroid */
/* extrapolate_second_order_and_limit_by_vertex / _by_edge for one quantity
 * (extrapolate_second_order_and_limit_by_vertex_or_edge.cu:694-800), per
 * triangle k:
 *  - gradient (a, b) from the centroid values (compute_gradients, :226-316):
 *    with fewer than two boundary edges the plane through the three surrogate
 *    neighbours' centroids; with two, the two-point gradient towards the one
 *    neighbour; with three, none (a = b = 0: the kernel leaves its gradient
 *    arrays untouched there, which for an isolated triangle means 0);
 *  - vertex values c + a (x_i - x) + b (y_i - y) and edge values as midpoint
 *    means (extrapolate_from_gradient, :172-224);
 *  - limiter against the range of the centroid value and the true
 *    neighbours' (limit_vertices_by_all_neighbours :5-75 on the vertex
 *    values, limit_edges_by_all_neighbours :77-170 on the edge values, the
 *    vertices then rebuilt as e_{i+1} + e_{i+2} - e_i): phi = min over the
 *    three of min(r beta, 1), r = (qmax - qc) / dq or (qmin - qc) / dq.
 * Edge values are left for the interpolation after balance. */
void or_extrapolate_second_order_and_limit(or_domain *D, const double *cv, double *vv, double *ev, double beta,
                                           int by_edge) {
    const int64_t N = D->number_of_elements;
    const double *cc = D->centroid_coordinates, *xv = D->vertex_coordinates;
    for (int64_t k = 0; k < N; k++) {
        const int64_t *sur = D->surrogate_neighbours + 3 * k;
        const int64_t nob = D->number_of_boundaries[k];
        double a = 0.0, b = 0.0;
        if (nob < 2) {
            const int64_t k0 = sur[0], k1 = sur[1], k2 = sur[2];
            if (!(k0 == k1 || k1 == k2)) {
                const double q0 = cv[k0], q1 = cv[k1], q2 = cv[k2];
                const double x0 = cc[2 * k0], y0 = cc[2 * k0 + 1];
                const double x1 = cc[2 * k1], y1 = cc[2 * k1 + 1];
                const double x2 = cc[2 * k2], y2 = cc[2 * k2 + 1];
                const double det = (y2 - y0) * (x1 - x0) - (y1 - y0) * (x2 - x0);
                a = (y2 - y0) * (q1 - q0) - (y1 - y0) * (q2 - q0);
                a /= det;
                b = (x1 - x0) * (q2 - q0) - (x2 - x0) * (q1 - q0);
                b /= det;
            }
        } else if (nob == 2) {
            int64_t k0 = k;
            for (int i = 0; i < 3 && k0 == k; i++) k0 = sur[i];
            if (k0 != k) {
                const double q0 = cv[k0], q1 = cv[k];
                const double x0 = cc[2 * k0], y0 = cc[2 * k0 + 1];
                const double x1 = cc[2 * k], y1 = cc[2 * k + 1];
                const double det = (x1 - x0) * (x1 - x0) + (y1 - y0) * (y1 - y0);
                a = (x1 - x0) * (q1 - q0) / det;
                b = (y1 - y0) * (q1 - q0) / det;
            }
        }
        const double x = cc[2 * k], y = cc[2 * k + 1], qc = cv[k];
        double v[3], e[3];
        for (int i = 0; i < 3; i++) v[i] = qc + a * (xv[6 * k + 2 * i] - x) + b * (xv[6 * k + 2 * i + 1] - y);
        e[0] = 0.5 * (v[1] + v[2]);
        e[1] = 0.5 * (v[2] + v[0]);
        e[2] = 0.5 * (v[0] + v[1]);
        double qmin = qc, qmax = qc;
        for (int i = 0; i < 3; i++) {
            const int64_t n = D->neighbours[3 * k + i];
            if (n >= 0) {
                qmin = omin(qmin, cv[n]);
                qmax = omax(qmax, cv[n]);
            }
        }
        const double *src = by_edge ? e : v;
        double dqa[3], phi = 1.0;
        for (int i = 0; i < 3; i++) {
            double r = 1.0;
            const double dq = src[i] - qc;
            dqa[i] = dq;
            if (dq > 0.0) r = (qmax - qc) / dq;
            if (dq < 0.0) r = (qmin - qc) / dq;
            phi = omin(omin(r * beta, 1.0), phi);
        }
        if (by_edge) {
            for (int i = 0; i < 3; i++) e[i] = qc + phi * dqa[i];
            v[0] = e[1] + e[2] - e[0];
            v[1] = e[2] + e[0] - e[1];
            v[2] = e[0] + e[1] - e[2];
        } else {
            for (int i = 0; i < 3; i++) v[i] = qc + phi * dqa[i];
        }
        for (int i = 0; i < 3; i++) {
            vv[3 * k + i] = v[i];
            ev[3 * k + i] = 0.5 * (v[(i + 1) % 3] + v[(i + 2) % 3]);
        }
    }
}

void or_extrapolate_first_order(int64_t N, const double *centroid, double *edge, double *vertex) {
    for (int64_t k = 0; k < N; k++) {
        int64_t k3 = 3 * k;
        edge[k3] = edge[k3 + 1] = edge[k3 + 2] = centroid[k];
        vertex[k3] = vertex[k3 + 1] = vertex[k3 + 2] = centroid[k];
    }
}

/* shallow_water_ext.c:1462-1636 (_balance_deep_and_shallow), H0 passed as
 * domain.H0 (gpu_domain_advanced.py:972-989). */
void or_balance_deep_and_shallow(or_domain *D) {
    const int64_t N = D->number_of_elements;
    const double H0 = D->H0, alpha_balance = D->alpha_balance;
    const double epsilon = 1.0e-6;
    double *wc = D->stage_centroid_values, *zc = D->bed_centroid_values;
    double *wv = D->stage_vertex_values, *zv = D->bed_vertex_values;
    double *xmomc = D->xmom_centroid_values, *ymomc = D->ymom_centroid_values;
    double *xmomv = D->xmom_vertex_values, *ymomv = D->ymom_vertex_values;
    for (int64_t k = 0; k < N; k++) {
        int64_t k3 = 3 * k;
        double hc_k = wc[k] - zc[k];
        double dz = 0.0, alpha, hv[3];
        if (D->tight_slope_limiters == 0) {
            for (int i = 0; i < 3; i++) dz = omax(dz, fabs(zv[k3 + i] - zc[k]));
        }
        hv[0] = wv[k3] - zv[k3];
        hv[1] = wv[k3 + 1] - zv[k3 + 1];
        hv[2] = wv[k3 + 2] - zv[k3 + 2];
        double hmin = omin(hv[0], omin(hv[1], hv[2]));
        if (D->tight_slope_limiters == 0) {
            if (dz > 0.0)
                alpha = omax(omin(alpha_balance * hmin / dz, 1.0), 0.0);
            else
                alpha = 1.0;
        } else {
            if (hmin < H0) {
                alpha = 1.0;
                for (int i = 0; i < 3; i++) {
                    double h_diff = hc_k - hv[i];
                    if (h_diff <= 0) {
                    } else {
                        alpha = omin(alpha, (hc_k - H0) / h_diff);
                    }
                }
                if (alpha > 1.0) alpha = 1.0;
                if (alpha < 0.0) alpha = 0.0;
            } else {
                alpha = 1.0;
            }
        }
        if (alpha < 1) {
            for (int i = 0; i < 3; i++) {
                wv[k3 + i] = zv[k3 + i] + (1 - alpha) * hc_k + alpha * hv[i];
                if (D->use_centroid_velocities == 1) {
                    double uc, vc;
                    if (hc_k > epsilon) {
                        uc = xmomc[k] / hc_k;
                        vc = ymomc[k] / hc_k;
                    } else {
                        uc = 0.0;
                        vc = 0.0;
                    }
                    hv[i] = wv[k3 + i] - zv[k3 + i];
                    xmomv[k3 + i] = uc * hv[i];
                    ymomv[k3 + i] = vc * hv[i];
                } else {
                    xmomv[k3 + i] = (1 - alpha) * xmomc[k] + alpha * xmomv[k3 + i];
                    ymomv[k3 + i] = (1 - alpha) * ymomc[k] + alpha * ymomv[k3 + i];
                }
            }
        }
    }
}

/* quantity.c:687-707 */
void or_interpolate_from_vertices_to_edges(int64_t N, const double *vertex, double *edge) {
    for (int64_t k = 0; k < N; k++) {
        int64_t k3 = 3 * k;
        double q0 = vertex[k3], q1 = vertex[k3 + 1], q2 = vertex[k3 + 2];
        edge[k3 + 0] = 0.5 * (q1 + q2);
        edge[k3 + 1] = 0.5 * (q0 + q2);
        edge[k3 + 2] = 0.5 * (q0 + q1);
    }
}

/* gpu_domain_advanced.py:1880-1992 (vertex-limiter branch of
 * distribute_to_vertices_and_edges, optimised_gradient_limiter=True) */
int or_distribute_to_vertices_and_edges(or_domain *D) {
    const int64_t N = D->number_of_elements;
    int rc = OR_OK;
    or_protect(D);
    if (D->_order_ == 1) {
        or_extrapolate_first_order(N, D->stage_centroid_values, D->stage_edge_values,
                                   D->stage_vertex_values);
        or_extrapolate_first_order(N, D->xmom_centroid_values, D->xmom_edge_values,
                                   D->xmom_vertex_values);
        or_extrapolate_first_order(N, D->ymom_centroid_values, D->ymom_edge_values,
                                   D->ymom_vertex_values);
    } else if (D->limiter == 1 || D->limiter == 2) {
        const int by_edge = D->limiter == 2;
        or_extrapolate_second_order_and_limit(D, D->stage_centroid_values, D->stage_vertex_values,
                                              D->stage_edge_values, D->beta_stage, by_edge);
        or_extrapolate_second_order_and_limit(D, D->xmom_centroid_values, D->xmom_vertex_values,
                                              D->xmom_edge_values, D->beta_xmom, by_edge);
        or_extrapolate_second_order_and_limit(D, D->ymom_centroid_values, D->ymom_vertex_values,
                                              D->ymom_edge_values, D->beta_ymom, by_edge);
    } else {
        rc = or_extrapolate_second_order_sw(D);
        if (rc) return rc;
    }
    or_balance_deep_and_shallow(D);
    or_interpolate_from_vertices_to_edges(N, D->stage_vertex_values, D->stage_edge_values);
    or_interpolate_from_vertices_to_edges(N, D->xmom_vertex_values, D->xmom_edge_values);
    or_interpolate_from_vertices_to_edges(N, D->ymom_vertex_values, D->ymom_edge_values);
    return rc;
}

/* evaluate_segment.cu:3-119 and gpu_domain_advanced.py:2006-2108, plus
 * set_boundary.cu:1-18 for the bed.  Transmissive copies the edge values
 * (ANUGA's Transmissive_boundary returns the inside edge state; the reference
 * driver has no kernel for it, gpu_domain_advanced.py:2099-2100). */
void or_update_boundary(or_domain *D) {
    const int64_t Nb = D->number_of_boundary_elements;
    for (int64_t m = 0; m < Nb; m++) {
        int64_t vol = D->boundary_cells[m], edge = D->boundary_edges[m];
        int64_t e = 3 * vol + edge;
        D->bed_boundary_values[m] = D->bed_edge_values[e];
        switch (D->boundary_type[m]) {
        case OR_BC_REFLECTIVE: {
            double n1 = D->normals[6 * vol + 2 * edge], n2 = D->normals[6 * vol + 2 * edge + 1];
            double q1 = D->xmom_edge_values[e], q2 = D->ymom_edge_values[e];
            double r1 = -q1 * n1 - q2 * n2;
            double r2 = -q1 * n2 + q2 * n1;
            D->stage_boundary_values[m] = D->stage_edge_values[e];
            D->height_boundary_values[m] = D->height_edge_values[e];
            D->xmom_boundary_values[m] = n1 * r1 - n2 * r2;
            D->ymom_boundary_values[m] = n2 * r1 + n1 * r2;
            q1 = D->xvelocity_edge_values[e];
            q2 = D->yvelocity_edge_values[e];
            r1 = q1 * n1 + q2 * n2;
            r2 = q1 * n2 - q2 * n1;
            D->xvelocity_boundary_values[m] = n1 * r1 - n2 * r2;
            D->yvelocity_boundary_values[m] = n2 * r1 + n1 * r2;
            break;
        }
        case OR_BC_DIRICHLET:
            D->height_boundary_values[m] = D->height_edge_values[e];
            D->xvelocity_boundary_values[m] = D->xvelocity_edge_values[e];
            D->yvelocity_boundary_values[m] = D->yvelocity_edge_values[e];
            D->stage_boundary_values[m] = D->boundary_dirichlet[3 * m + 0];
            D->xmom_boundary_values[m] = D->boundary_dirichlet[3 * m + 1];
            D->ymom_boundary_values[m] = D->boundary_dirichlet[3 * m + 2];
            break;
        default: /* transmissive */
            D->stage_boundary_values[m] = D->stage_edge_values[e];
            D->xmom_boundary_values[m] = D->xmom_edge_values[e];
            D->ymom_boundary_values[m] = D->ymom_edge_values[e];
            D->height_boundary_values[m] = D->height_edge_values[e];
            D->xvelocity_boundary_values[m] = D->xvelocity_edge_values[e];
            D->yvelocity_boundary_values[m] = D->yvelocity_edge_values[e];
            break;
        }
    }
}

/* update_centroids_of_velocities_and_height.cu:1-52 followed by first-order
 * extrapolation of height / xvelocity / yvelocity (gpu_domain_advanced.py:2114-2142) */
void or_update_other_quantities(or_domain *D) {
    const int64_t N = D->number_of_elements, Nb = D->number_of_boundary_elements;
    const double H0 = 1.0e-8;
    for (int64_t k = 0; k < N; k++) {
        double h = D->stage_centroid_values[k] - D->bed_centroid_values[k];
        if (h < 0) h = 0;
        D->height_centroid_values[k] = h;
        double factor = h / (h * h + H0);
        D->xvelocity_centroid_values[k] = D->xmom_centroid_values[k] * factor;
        D->yvelocity_centroid_values[k] = D->ymom_centroid_values[k] * factor;
    }
    for (int64_t k = 0; k < Nb; k++) {
        double h = D->stage_boundary_values[k] - D->bed_boundary_values[k];
        if (h < 0) h = 0;
        D->height_boundary_values[k] = h;
        double factor = h / (h * h + H0);
        D->xvelocity_boundary_values[k] = D->xmom_boundary_values[k] * factor;
        D->yvelocity_boundary_values[k] = D->ymom_boundary_values[k] * factor;
    }
    or_extrapolate_first_order(N, D->height_centroid_values, D->height_edge_values,
                               D->height_vertex_values);
    or_extrapolate_first_order(N, D->xvelocity_centroid_values, D->xvelocity_edge_values,
                               D->xvelocity_vertex_values);
    or_extrapolate_first_order(N, D->yvelocity_centroid_values, D->yvelocity_edge_values,
                               D->yvelocity_vertex_values);
}

/* ANUGA Generic_Domain.update_timestep as restated in evolve.c:115-160, with
 * the final-time clip written as finaltime - time (evolve.c:147 has a sign
 * slip).  Returns OR_ERR_SMALLSTEPS where ANUGA raises. */
int or_update_timestep(or_domain *D) {
    double timestep = omin(D->CFL * D->flux_timestep, D->evolve_max_timestep);
    D->recorded_max_timestep = omax(timestep, D->recorded_max_timestep);
    D->recorded_min_timestep = omin(timestep, D->recorded_min_timestep);
    if (timestep < D->evolve_min_timestep) {
        D->smallsteps += 1;
        if (D->smallsteps > D->max_smallsteps) {
            D->smallsteps = 0;
            if (D->_order_ == 1) return OR_ERR_SMALLSTEPS;
            D->_order_ = 1;
        }
    } else {
        D->smallsteps = 0;
        if (D->_order_ == 1 && D->default_order == 2) D->_order_ = 2;
    }
    if (D->has_finaltime && D->time + timestep > D->finaltime) timestep = D->finaltime - D->time;
    if (D->time + timestep > D->yieldtime) timestep = D->yieldtime - D->time;
    D->timestep = timestep;
    return OR_OK;
}

/* gpu_domain_advanced.py:2178-2191 (evolve_one_euler_step) */
int or_evolve_one_euler_step(or_domain *D) {
    or_compute_fluxes(D);
    or_gravity_wb(D);
    or_compute_forcing_terms(D);
    int rc = or_update_timestep(D);
    if (rc) return rc;
    return or_update_conserved_quantities(D);
}

/* backup_conserved_quantities (evolve.c:549-561; gpu_domain_advanced.py:1449-1475):
 * copies of the conserved centroid values (the reference's memcpy passes the
 * element count as the byte count; the whole arrays are meant) */
void or_backup_conserved_quantities(or_domain *D) {
    const size_t n = (size_t)D->number_of_elements * sizeof(double);
    memcpy(D->stage_centroid_backup, D->stage_centroid_values, n);
    memcpy(D->xmom_centroid_backup, D->xmom_centroid_values, n);
    memcpy(D->ymom_centroid_backup, D->ymom_centroid_values, n);
}

/* saxpy_conserved_quantities (evolve.c:566-603; saxpy_centroid_values.cu:1-12):
 * centroid = a * centroid + b * backup for each conserved quantity */
void or_saxpy_conserved_quantities(or_domain *D, double a, double b) {
    double *q[3] = {D->stage_centroid_values, D->xmom_centroid_values, D->ymom_centroid_values};
    const double *p[3] = {D->stage_centroid_backup, D->xmom_centroid_backup, D->ymom_centroid_backup};
    for (int c = 0; c < 3; c++)
        for (int64_t k = 0; k < D->number_of_elements; k++) q[c][k] = a * q[c][k] + b * p[c][k];
}

/* One Euler stage after the first: compute_fluxes, forcing terms and
 * update_conserved_quantities with the timestep of the first stage. */
static int or_rk_stage(or_domain *D) {
    or_compute_fluxes(D);
    or_gravity_wb(D);
    or_compute_forcing_terms(D);
    return or_update_conserved_quantities(D);
}

/* ANUGA's Generic_Domain.evolve_one_rk2_step / evolve_one_rk3_step, which
 * the reference drivers call for timestepping_method 'rk2' / 'rk3'
 * (gpu_domain_advanced.py:2198-2300, 2566-2573; the C stubs evolve.c:1093-1100
 * are empty).  Restated from ANUGA's published algorithm (the anuga package
 * is not in the reference): the first stage is the Euler step (with the
 * timestep and time advance) followed by distribute and update_boundary; rk2
 * adds one stage and Q = 0.5 Q + 0.5 Q0; rk3 adds Q = 0.25 Q + 0.75 Q0 (time
 * t0 + dt/2) and Q = 2/3 Q + 1/3 Q0 (time t0 + dt), each followed by
 * distribute and update_boundary.  Ends with time t0 + dt. */
int or_evolve_one_rk_step(or_domain *D, int stages) {
    const double t0 = D->time;
    or_backup_conserved_quantities(D);
    int rc = or_evolve_one_euler_step(D);
    if (rc) return rc;
    D->time = t0 + D->timestep;
    rc = or_distribute_to_vertices_and_edges(D);
    if (rc) return rc;
    or_update_boundary(D);
    rc = or_rk_stage(D);
    if (rc) return rc;
    if (stages == 2) {
        or_saxpy_conserved_quantities(D, 0.5, 0.5);
    } else {
        or_saxpy_conserved_quantities(D, 0.25, 0.75);
        D->time = t0 + D->timestep * 0.5;
        rc = or_distribute_to_vertices_and_edges(D);
        if (rc) return rc;
        or_update_boundary(D);
        rc = or_rk_stage(D);
        if (rc) return rc;
        or_saxpy_conserved_quantities(D, 2.0 / 3.0, 1.0 / 3.0);
        D->time = t0 + D->timestep;
    }
    rc = or_distribute_to_vertices_and_edges(D);
    if (rc) return rc;
    or_update_boundary(D);
    return OR_OK;
}

/* gpu_domain_advanced.py:2473-2545: evolve() preamble up to the first yield */
int or_evolve_prepare(or_domain *D, double yieldstep, double finaltime, int64_t has_finaltime) {
    int rc = or_distribute_to_vertices_and_edges(D);
    if (rc) return rc;
    D->time = D->starttime;
    D->yieldstep = yieldstep;
    D->_order_ = D->default_order;
    D->has_finaltime = has_finaltime;
    D->finaltime = finaltime;
    D->yieldtime = D->time + yieldstep;
    D->recorded_min_timestep = D->evolve_max_timestep;
    D->recorded_max_timestep = D->evolve_min_timestep;
    D->number_of_steps = 0;
    D->number_of_first_order_steps = 0;
    D->smallsteps = 0;
    rc = or_distribute_to_vertices_and_edges(D);
    if (rc) return rc;
    or_update_boundary(D);
    return OR_OK;
}

/* gpu_domain_advanced.py:2555-2663: one iteration of the evolve loop.
 * *event reports a yield (the caller then advances yieldtime) or the end. */
int or_evolve_step(or_domain *D, int64_t *event) {
    const double epsilon = 1.0e-12; /* anuga.config.epsilon */
    double initial_time = D->time;
    int rc;
    if (D->timestepping_method == 2 || D->timestepping_method == 3) {
        rc = or_evolve_one_rk_step(D, (int)D->timestepping_method);
        if (rc) return rc;
    } else {
        rc = or_evolve_one_euler_step(D);
        if (rc) return rc;
        D->time = initial_time + D->timestep;
        rc = or_distribute_to_vertices_and_edges(D);
        if (rc) return rc;
        or_update_boundary(D);
    }
    or_update_other_quantities(D);
    D->number_of_steps += 1;
    if (D->_order_ == 1) D->number_of_first_order_steps += 1;
    *event = OR_EVENT_NONE;
    if (D->has_finaltime && D->time >= D->finaltime - epsilon) {
        D->time = D->finaltime;
        *event = OR_EVENT_FINAL;
    } else if (D->time >= D->yieldtime) {
        *event = OR_EVENT_YIELD;
    }
    return OR_OK;
}

/* gpu_domain_advanced.py:2649-2656: bookkeeping when evolve() resumes after a yield */
void or_advance_yield(or_domain *D) {
    D->yieldtime += D->yieldstep;
    D->recorded_min_timestep = D->evolve_max_timestep;
    D->recorded_max_timestep = D->evolve_min_timestep;
    D->number_of_steps = 0;
    D->number_of_first_order_steps = 0;
}

int64_t or_sizeof_domain(void) { return (int64_t)sizeof(or_domain); }

/* Run steps until an event (yield / final time) or max_steps. */
int or_evolve_steps(or_domain *D, int64_t max_steps, int64_t *steps_done, int64_t *event) {
    *steps_done = 0;
    *event = OR_EVENT_NONE;
    while (*steps_done < max_steps) {
        int rc = or_evolve_step(D, event);
        if (rc) return rc;
        (*steps_done)++;
        if (*event != OR_EVENT_NONE) break;
    }
    return OR_OK;
}
