// Device-side numerics of the shallow-water step for gfx950 (CDNA4).
//
// Every helper evaluates the same IEEE operations, in the same order, as
// ANUGA's FP64 C kernels in the reference (src/anuga_HMPP/shallow_water_ext.c,
// quantity.c, util_ext.h).  The library is compiled with -ffp-contract=off so
// that no a*b+c is fused; FP64 division and square root lower to correctly
// rounded instruction sequences on gfx950, so a kernel built from these
// helpers reproduces the CPU results bit for bit (pow() in Manning friction is
// the one libm call whose last bit may differ).
//
// The helpers are templated on the arithmetic type R of the flux and the
// reconstruction: R = double is the FP64 path; R = float is the mixed path,
// where depths, wet/dry tests, accumulation and the timestep stay double.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#define SW_DEV __device__ __forceinline__

namespace sw {

// Loads / stores of data a kernel touches once (mesh geometry, per-triangle
// inputs and outputs): non-temporal, so they stream past the L2 and leave it to
// the edge and centroid records that neighbouring triangles gather.
template <typename T> SW_DEV T ld_s(const T *p) { return __builtin_nontemporal_load(p); }

// Element idx of an array as its base plus a 32-bit byte offset: with a
// uniform base the compiler can then use the global instructions' scalar-base
// form instead of forming a 64-bit address in every lane.  Every array here is
// far below 4 GiB.
template <typename T> SW_DEV T *at32(T *base, uint32_t idx) {
    using B = typename std::conditional<std::is_const<T>::value, const char, char>::type;
    return reinterpret_cast<T *>(reinterpret_cast<B *>(base) + (uint64_t)(idx * (uint32_t)sizeof(T)));
}
template <typename T> SW_DEV void st_s(T *p, T v) { __builtin_nontemporal_store(v, p); }
typedef double sw_f64x2 __attribute__((ext_vector_type(2)));
SW_DEV double2 ld_s(const double2 *p) {
    const sw_f64x2 v = __builtin_nontemporal_load(reinterpret_cast<const sw_f64x2 *>(p));
    return make_double2(v.x, v.y);
}
SW_DEV void st_s(double2 *p, double2 v) {
    sw_f64x2 w;
    w.x = v.x;
    w.y = v.y;
    __builtin_nontemporal_store(w, reinterpret_cast<sw_f64x2 *>(p));
}

// util_ext.h:42-56: comparisons, not fmax/fmin (NaN propagates the same way)
template <typename T> SW_DEV T omax(T x, T y) { return (x > y) ? x : y; }
template <typename T> SW_DEV T omin(T x, T y) { return (x < y) ? x : y; }

struct FluxParams {
    double epsilon, h0, limiting_threshold, g;
};

// _compute_speed (shallow_water_ext.c:214-240) for uh, then for vh on the
// depth the first call left (the C's two calls, shallow_water_ext.c:360-379):
// below limiting_threshold the speed is limited as the ANUGA manual has it,
// u = uh / (h + h0 / h), and the momentum made consistent, uh = u h; when
// h < epsilon the first call sets h = 0 and the second gives vh = 0 * 0;
// otherwise both divide by the same h + h0 / h, formed once here; above
// limiting_threshold u = uh / h, the second call's velocity is discarded and
// vh is unchanged.
template <typename R>
SW_DEV R compute_speed2(R &uh, R &vh, R &h, R epsilon, R h0, R limiting_threshold) {
    R u;
    if (h < limiting_threshold) {
        if (h < epsilon) {
            h = R(0);
            u = R(0);
            uh = u * h;
            vh = u * h;
        } else {
            const R d = h + h0 / h;
            u = uh / d;
            uh = u * h;
            vh = (vh / d) * h;
        }
    } else {
        u = uh / h;
    }
    return u;
}

// shallow_water_ext.c:302-442 (_flux_function_central, Kurganov-Noelle-Petrova
// central-upwind flux); ql/qr = (w, uh, vh) in x/y frame, result in x/y frame.
template <typename R>
SW_DEV void flux_central(const R ql_in[3], const R qr_in[3], R z_left, R z_right, R n1, R n2,
                         R epsilon, R h0, R limiting_threshold, R g, R ef[3], R &max_speed) {
    R ql0 = ql_in[0];
    R ql1 = n1 * ql_in[1] + n2 * ql_in[2];
    R ql2 = -n2 * ql_in[1] + n1 * ql_in[2];
    R qr0 = qr_in[0];
    R qr1 = n1 * qr_in[1] + n2 * qr_in[2];
    R qr2 = -n2 * qr_in[1] + n1 * qr_in[2];

    R z = R(0.5) * (z_left + z_right);

    R h_left = ql0 - z;
    R uh_left = ql1;
    R vh_left = ql2;
    R u_left = compute_speed2<R>(uh_left, vh_left, h_left, epsilon, h0, limiting_threshold);
    R h_right = qr0 - z;
    R uh_right = qr1;
    R vh_right = qr2;
    R u_right = compute_speed2<R>(uh_right, vh_right, h_right, epsilon, h0, limiting_threshold);

    R soundspeed_left = sqrt(g * h_left);
    R soundspeed_right = sqrt(g * h_right);

    R s_max = omax<R>(u_left + soundspeed_left, u_right + soundspeed_right);
    if (s_max < R(0)) s_max = R(0);
    R s_min = omin<R>(u_left - soundspeed_left, u_right - soundspeed_right);
    if (s_min > R(0)) s_min = R(0);

    R fl0 = u_left * h_left;
    R fl1 = u_left * uh_left + R(0.5) * g * h_left * h_left;
    R fl2 = u_left * vh_left;
    R fr0 = u_right * h_right;
    R fr1 = u_right * uh_right + R(0.5) * g * h_right * h_right;
    R fr2 = u_right * vh_right;

    R denom = s_max - s_min;
    if (denom < epsilon) {
        ef[0] = R(0);
        ef[1] = R(0);
        ef[2] = R(0);
        max_speed = R(0);
    } else {
        R inv = R(1) / denom;
        R smm = s_max * s_min;
        R e0 = s_max * fl0 - s_min * fr0;
        e0 += smm * (qr0 - ql0);
        e0 *= inv;
        R e1 = s_max * fl1 - s_min * fr1;
        e1 += smm * (qr1 - ql1);
        e1 *= inv;
        R e2 = s_max * fl2 - s_min * fr2;
        e2 += smm * (qr2 - ql2);
        e2 *= inv;
        max_speed = omax<R>(fabs(s_max), fabs(s_min));
        R m2 = -n2;
        ef[0] = e0;
        ef[1] = n1 * e1 + m2 * e2;
        ef[2] = -m2 * e1 + n1 * e2;
    }
}

// Mixed-precision flux.  The same KNP flux as flux_central with the velocity,
// wave-speed and flux arithmetic in FP32; the depth tests stay FP64: each side
// comes with its depth class, decided on the FP64 depth h = w - z,
//   0: |h| < eps (dry: the edge is skipped when both sides are dry, and
//      _compute_speed zeroes h and u), 1: h <= -eps (zeroed), 2: eps <= h <
//   limiting_threshold (damped velocity), 3: deeper.
SW_DEV int depth_class(double h, double eps, double limiting_threshold) {
    if (fabs(h) < eps) return 0;
    if (h < eps) return 1;
    if (h < limiting_threshold) return 2;
    return 3;
}

// _compute_speed (shallow_water_ext.c:214-240) with the branch given by the class.
SW_DEV float speed_f(float uh, float h, int cls, float h0, float &uh_out, float &h_out) {
    float u;
    if (cls <= 1) {
        h_out = 0.0f;
        u = 0.0f;
        uh_out = 0.0f;
    } else if (cls == 2) {
        h_out = h;
        u = uh / (h + h0 / h);
        uh_out = u * h;
    } else {
        h_out = h;
        u = uh / h;
        uh_out = uh;
    }
    return u;
}

// hl, hr: FP32 depths; dw: the stage jump w_r - w_l; (uhl, vhl), (uhr, vhr):
// momenta in the x/y frame.  Result in the x/y frame, FP64.
SW_DEV void flux_mixed_core(float hl, float hr, int cl, int cr, float dw, float uhl, float vhl, float uhr,
                            float vhr, float n1, float n2, float g, float h0, double epsilon, double ef[3],
                            double &max_speed) {
    const float ql1 = n1 * uhl + n2 * vhl, ql2 = -n2 * uhl + n1 * vhl;
    const float qr1 = n1 * uhr + n2 * vhr, qr2 = -n2 * uhr + n1 * vhr;
    float h_left, h_right, uh_left, uh_right, vh_left, vh_right, tmp;
    const float u_left = speed_f(ql1, hl, cl, h0, uh_left, h_left);
    const float u_right = speed_f(qr1, hr, cr, h0, uh_right, h_right);
    speed_f(ql2, hl, cl, h0, vh_left, tmp);
    speed_f(qr2, hr, cr, h0, vh_right, tmp);
    const float c_left = sqrtf(g * h_left), c_right = sqrtf(g * h_right);
    float s_max = omax<float>(u_left + c_left, u_right + c_right);
    if (s_max < 0.0f) s_max = 0.0f;
    float s_min = omin<float>(u_left - c_left, u_right - c_right);
    if (s_min > 0.0f) s_min = 0.0f;
    const float fl0 = u_left * h_left, fl1 = u_left * uh_left + 0.5f * g * h_left * h_left,
                fl2 = u_left * vh_left;
    const float fr0 = u_right * h_right, fr1 = u_right * uh_right + 0.5f * g * h_right * h_right,
                fr2 = u_right * vh_right;
    const float denom = s_max - s_min;
    if (double(denom) < epsilon) {
        ef[0] = ef[1] = ef[2] = 0.0;
        max_speed = 0.0;
        return;
    }
    const float inv = 1.0f / denom, smm = s_max * s_min;
    const float e0 = (s_max * fl0 - s_min * fr0 + smm * dw) * inv;
    const float e1 = (s_max * fl1 - s_min * fr1 + smm * (qr1 - ql1)) * inv;
    const float e2 = (s_max * fl2 - s_min * fr2 + smm * (qr2 - ql2)) * inv;
    max_speed = double(omax<float>(fabsf(s_max), fabsf(s_min)));
    ef[0] = double(e0);
    ef[1] = double(n1 * e1 - n2 * e2);
    ef[2] = double(n2 * e1 + n1 * e2);
}

// Mixed flux from FP64 edge states (w, uh, vh, z of each side).
SW_DEV void flux_central_mixed(const double ql[3], const double qr[3], double z_left, double z_right, double n1,
                               double n2, double epsilon, double h0, double limiting_threshold, double g,
                               double ef[3], double &max_speed) {
    const double z = 0.5 * (z_left + z_right);
    const double hl = ql[0] - z, hr = qr[0] - z;
    flux_mixed_core(float(hl), float(hr), depth_class(hl, epsilon, limiting_threshold),
                    depth_class(hr, epsilon, limiting_threshold), float(qr[0] - ql[0]), float(ql[1]), float(ql[2]),
                    float(qr[1]), float(qr[2]), float(n1), float(n2), float(g), float(h0), epsilon, ef, max_speed);
}

// util_ext.h:66-109 (_gradient)
SW_DEV void gradient3(double x0, double y0, double x1, double y1, double x2, double y2,
                      double q0, double q1, double q2, double &a, double &b) {
    double det = (y2 - y0) * (x1 - x0) - (y1 - y0) * (x2 - x0);
    a = (y2 - y0) * (q1 - q0) - (y1 - y0) * (q2 - q0);
    a /= det;
    b = (x1 - x0) * (q2 - q0) - (x2 - x0) * (q1 - q0);
    b /= det;
}

// shallow_water_ext.c:98-149 (find_qmin_and_qmax); results returned by value
// so they stay in registers.
template <typename R>
struct QRange {
    R qmin, qmax;
};

template <typename R>
SW_DEV QRange<R> find_qmin_and_qmax(R dq0, R dq1, R dq2) {
    R qmin, qmax;
    if (dq0 >= R(0)) {
        const bool a = dq1 >= dq2;
        const R hi = a ? dq1 : dq2;
        const R lo = a ? dq2 : dq1;
        // dq1 >= dq2: qmax = dq1 >= 0 ? dq0 + dq1 : dq0;  else qmax = dq2 > 0 ? dq0 + dq2 : dq0
        const bool take = a ? (hi >= R(0)) : (hi > R(0));
        qmax = take ? dq0 + hi : dq0;
        qmin = dq0 + lo;
        if (qmin >= R(0)) qmin = R(0);
    } else {
        const bool a = dq1 <= dq2;
        const R lo = a ? dq1 : dq2;
        const R hi = a ? dq2 : dq1;
        qmin = (lo < R(0)) ? dq0 + lo : dq0;
        qmax = dq0 + hi;
        if (qmax <= R(0)) qmax = R(0);
    }
    return {qmin, qmax};
}

// shallow_water_ext.c:151-183 (limit_gradient); r0 carries over between i.
template <typename R>
SW_DEV void limit_gradient(R dqv[3], R qmin, R qmax, R beta) {
    R r = R(1000.0), r0 = R(1.0);
    const R TINY = R(1.0e-100);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        if (dqv[i] < -TINY) r0 = qmin / dqv[i];
        if (dqv[i] > TINY) r0 = qmax / dqv[i];
        r = omin<R>(r0, r);
    }
    R phi = omin<R>(r * beta, R(1.0));
    dqv[0] = dqv[0] * phi;
    dqv[1] = dqv[1] * phi;
    dqv[2] = dqv[2] * phi;
}

// Auxiliary-triangle reconstruction of one quantity (shallow_water_ext.c:3023-3062).
// Returns the three limited jumps from the centroid.
// The differences dq0 = q0 - qk, dq1 = q1 - q0, dq2 = q2 - q0 are formed by the
// caller in FP64 (exactly the C's operations), then rounded to R: in the mixed
// path this keeps the cancellation of nearly equal stages out of FP32.
template <typename R>
SW_DEV void extrap_aux(R dq0, R dq1, R dq2, R dx1, R dx2, R dy1, R dy2, R inv_area2,
                       const R dxv[3], const R dyv[3], R beta, R dqv[3]) {
    R a = dy2 * dq1 - dy1 * dq2;
    a *= inv_area2;
    R b = dx1 * dq2 - dx2 * dq1;
    b *= inv_area2;
    dqv[0] = a * dxv[0] + b * dyv[0];
    dqv[1] = a * dxv[1] + b * dyv[1];
    dqv[2] = a * dxv[2] + b * dyv[2];
    const QRange<R> q = find_qmin_and_qmax<R>(dq0, dq1, dq2);
    limit_gradient<R>(dqv, q.qmin, q.qmax, beta);
}

// Reconstruction toward the single interior neighbour (shallow_water_ext.c:3174-3213).
// dq1 = q1 - qk formed by the caller in FP64.
template <typename R>
SW_DEV void extrap_one(R dq1, R dx2, R dy2, const R dxv[3], const R dyv[3], R beta, R dqv[3]) {
    R a = dq1 * dx2;
    R b = dq1 * dy2;
    dqv[0] = a * dxv[0] + b * dyv[0];
    dqv[1] = a * dxv[1] + b * dyv[1];
    dqv[2] = a * dxv[2] + b * dyv[2];
    const R qmin = (dq1 >= R(0)) ? R(0) : dq1;
    const R qmax = (dq1 >= R(0)) ? dq1 : R(0);
    limit_gradient<R>(dqv, qmin, qmax, beta);
}

// quantity.c:769-820 (_update) for one triangle.  Returns false when the
// semi-implicit denominator is <= 0 (the C returns -1 there).  With si = 0 the
// C's steps reduce to q + dt eu exactly (si / q = +-0, denominator 1 - dt * +-0
// = 1, q / 1 = q), so that case skips the two divisions; si is left for the
// caller to reset, as every caller does.
SW_DEV bool update_one(double &q, double eu, double &si, double dt) {
    if (si == 0.0) {
        q += dt * eu;
        return true;
    }
    double x = q;
    if (x == 0.0)
        si = 0.0;
    else
        si /= x;
    q += dt * eu;
    double denominator = 1.0 - dt * si;
    if (denominator <= 0.0) return false;
    q /= denominator;
    return true;
}

// The semi-implicit term divided by the centroid value, the first loop of
// quantity.c:769-820 (0 for a zero centroid value).
SW_DEV double si_ratio(double si, double x) { return (x == 0.0) ? 0.0 : si / x; }

// update_one with that division already done (sr = si_ratio(si, q)): the same
// operations, in the same order.
SW_DEV bool update_ratio(double &q, double eu, double sr, double dt) {
    q += dt * eu;
    if (sr == 0.0) return true;  // denominator 1 - dt * +-0 = 1
    const double denominator = 1.0 - dt * sr;
    if (denominator <= 0.0) return false;
    q /= denominator;
    return true;
}

// shallow_water_ext.c:1638-1705 (_protect) for one triangle.
SW_DEV void protect_one(double &wc, double zc, double &xmomc, double &ymomc, double mah,
                        double mas, double epsilon) {
    double hc = wc - zc;
    if (mas < epsilon) {
        if (hc < mah) {
            xmomc = 0.0;
            ymomc = 0.0;
            if (hc <= 0.0) wc = zc;
        }
    } else {
        if (hc < mah) {
            if (hc <= 0.0) {
                wc = zc;
                xmomc = 0.0;
                ymomc = 0.0;
            } else {
                double u = xmomc / hc;
                if (fabs(u) > mas) {
                    double reduced_speed = mas * u / fabs(u);
                    xmomc = reduced_speed * hc;
                }
                double v = ymomc / hc;
                if (fabs(v) > mas) {
                    double reduced_speed = mas * v / fabs(v);
                    ymomc = reduced_speed * hc;
                }
            }
        }
    }
}

struct BalanceParams {
    double H0, alpha_balance;
    int tight_slope_limiters, use_centroid_velocities;
};

// shallow_water_ext.c:1462-1636 (_balance_deep_and_shallow) for one triangle.
SW_DEV void balance_one(double wc, double zc, double wv[3], const double zv[3], double xmomc,
                        double ymomc, double xmomv[3], double ymomv[3], const BalanceParams &p) {
    const double epsilon = 1.0e-6;
    double hc_k = wc - zc;
    double dz = 0.0, alpha, hv[3];
    if (p.tight_slope_limiters == 0) {
        for (int i = 0; i < 3; i++) dz = omax<double>(dz, fabs(zv[i] - zc));
    }
    hv[0] = wv[0] - zv[0];
    hv[1] = wv[1] - zv[1];
    hv[2] = wv[2] - zv[2];
    double hmin = omin<double>(hv[0], omin<double>(hv[1], hv[2]));
    if (p.tight_slope_limiters == 0) {
        if (dz > 0.0)
            alpha = omax<double>(omin<double>(p.alpha_balance * hmin / dz, 1.0), 0.0);
        else
            alpha = 1.0;
    } else {
        if (hmin < p.H0) {
            alpha = 1.0;
            for (int i = 0; i < 3; i++) {
                double h_diff = hc_k - hv[i];
                if (!(h_diff <= 0)) alpha = omin<double>(alpha, (hc_k - p.H0) / h_diff);
            }
            if (alpha > 1.0) alpha = 1.0;
            if (alpha < 0.0) alpha = 0.0;
        } else {
            alpha = 1.0;
        }
    }
    if (alpha < 1) {
        for (int i = 0; i < 3; i++) {
            wv[i] = zv[i] + (1 - alpha) * hc_k + alpha * hv[i];
            if (p.use_centroid_velocities == 1) {
                double uc, vc;
                if (hc_k > epsilon) {
                    uc = xmomc / hc_k;
                    vc = ymomc / hc_k;
                } else {
                    uc = 0.0;
                    vc = 0.0;
                }
                hv[i] = wv[i] - zv[i];
                xmomv[i] = uc * hv[i];
                ymomv[i] = vc * hv[i];
            } else {
                xmomv[i] = (1 - alpha) * xmomc + alpha * xmomv[i];
                ymomv[i] = (1 - alpha) * ymomc + alpha * ymomv[i];
            }
        }
    }
}

// quantity.c:687-707 (_interpolate_from_vertices_to_edges) for one triangle.
SW_DEV void interpolate_one(const double v[3], double e[3]) {
    e[0] = 0.5 * (v[1] + v[2]);
    e[1] = 0.5 * (v[0] + v[2]);
    e[2] = 0.5 * (v[0] + v[1]);
}

// evaluate_segment.cu:36-63: reflective boundary momentum (normal component negated).
SW_DEV void reflect_momentum(double q1, double q2, double n1, double n2, double &b1, double &b2) {
    double r1 = -q1 * n1 - q2 * n2;
    double r2 = -q1 * n2 + q2 * n1;
    b1 = n1 * r1 - n2 * r2;
    b2 = n2 * r1 + n1 * r2;
}

// Positive doubles order like their bit patterns, so the CFL minimum is an
// unsigned 64-bit atomic min of the bits.
SW_DEV unsigned long long dbits(double x) { return (unsigned long long)__double_as_longlong(x); }
SW_DEV double bitsd(unsigned long long b) { return __longlong_as_double((long long)b); }

// Wavefront (64-lane) minimum of positive doubles as bit patterns, by
// cross-lane shuffles; the result is in every lane.
SW_DEV unsigned long long wave_min_bits(double v) {
    unsigned long long b = dbits(v);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(b, off, 64);
        b = (o < b) ? o : b;
    }
    return b;
}

}  // namespace sw
