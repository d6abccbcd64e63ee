// C ABI of the shallow-water library (declarations: include/anuga_sw.h).
//
// A handle owns every device buffer of one domain -- or of an ensemble of
// independent domains sharing one set of arrays -- and one HIP stream
// (default priority).  Operator entry points launch the operator kernels of
// sw_kernels.h; sw_evolve_steps runs the fused step (flux -> timestep ->
// update -> distribute) for every instance at once, with no host
// synchronisation between steps, replayed from a hipGraph of kGraphSteps
// steps, and stops at yields through device-side halt flags.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "../../include/anuga_sw.h"
#include "sw_kernels.h"
#include "sw_step_vertex.h"

using namespace sw;

namespace {

constexpr int kGraphSteps = 16;
constexpr int kBatchSteps = 8 * kGraphSteps;  // steps between host checks
constexpr int kNumTimed = 4;                  // flux, timestep, update, distribute

thread_local std::string g_create_error;

// One instance of a handle: its triangles are device triangles
// [off, off + n), its boundary edges [boff, boff + nb); its caller -> device
// permutation starts at dev_of[poff].
struct Inst {
    int64_t n = 0, off = 0, nb = 0, boff = 0, poff = 0;
};

// Output points of one instance (sw_set_sww_points): point p averages the
// vertex-value slots slot[start[p] .. start[p+1]); out holds the FP32 values.
// Lane t handles point order[t]: the points sorted by the device index of
// their first triangle, so neighbouring lanes gather neighbouring triangles.
struct SwwPoints {
    int64_t P = 0;
    int *start = nullptr, *slot = nullptr, *order = nullptr;
    float *out = nullptr;
};

}  // namespace

struct sw_handle {
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t N = 0, Nb = 0;  // device triangles (instances and their padding), boundary edges
    int n_inst = 1;
    int cur = -1;  // selected instance, -1 = all
    std::vector<Inst> inst;
    sw_params params{};
    DevParams dp{};
    DevMesh mesh{};
    DevFields f{};
    DevState *st = nullptr;  // [n_inst]
    std::vector<void *> allocations;
    std::string err;
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    bool timing = false;
    std::vector<hipEvent_t> ev;
    double times_ms[kNumTimed] = {};
    int64_t launches[kNumTimed] = {};
    double *staging = nullptr;  // [3 * largest instance] transfers
    int *cal_of = nullptr;      // [N] caller index (within its instance) of device triangle d
    int *dev_of = nullptr;      // [sum n] device index (within its instance) of caller triangle k
    int blk[3] = {kBlock, kBlock, kBlock};  // workgroup size of the fused flux, update, distribute kernels
    std::vector<SwwPoints> sww;             // [n_inst] output points
    // vertex-record Euler step (sw_step_vertex.h)
    StepFields sf{};
    int *nbw = nullptr;           // [3][N] neighbour words of k_flux_v
    bool vstep_geometry = false;  // caller's geometry equals tri_geometry and shared normals are negations
    int step_path = SW_STEP_AUTO; // caller's choice
    int vpar = 0;                 // parity of the next vertex-record step
    int graph_path = 0;           // step path the captured graph holds
    // After a vertex-record evolve call the step's records (vertex records,
    // centroid depths; the states at parity vpar) hold the state.  records_live: they are what k_step_begin would
    // build from the operator API's arrays, so the next call continues from
    // them.  api_pending: those arrays are behind the records (k_step_end and
    // the epilogue run when something reads or writes them: sync_api).
    bool records_live = false;
    bool api_pending = false;
    unsigned uni = 0;  // kUniFriction / kUniMeanBed bits that hold for every instance (k_update_v argument)
    // host copy of the instances' states (valid while st_host_ok) and the
    // pinned buffer they are read through
    std::vector<DevState> st_host;
    bool st_host_ok = false;
    DevState *st_pinned = nullptr;
    DevState *st_up = nullptr;  // [2][n_inst] pinned staging of state writes that are not waited for
};

namespace {

int fail(sw_handle *h, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (h)
        h->err = buf;
    else
        g_create_error = buf;
    return code;
}

#define HIPCHECK(h, call)                                                                    \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail((h), SW_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));     \
    } while (0)

template <typename T>
int dalloc(sw_handle *h, T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T));
    if (e != hipSuccess)
        return fail(h, SW_ERR_HIP, "hipMalloc(%zu bytes) failed: %s", count * sizeof(T),
                    hipGetErrorString(e));
    h->allocations.push_back(q);
    *p = static_cast<T *>(q);
    // zeroed on the handle's stream, so that its later work sees the zeros
    // (a null-stream memset is not ordered with a non-blocking stream)
    e = hipMemsetAsync(q, 0, count * sizeof(T), h->stream);
    if (e != hipSuccess) return fail(h, SW_ERR_HIP, "hipMemsetAsync failed: %s", hipGetErrorString(e));
    return SW_OK;
}

void dfree(sw_handle *h, void *p) {
    if (!p) return;
    for (auto &q : h->allocations)
        if (q == p) {
            (void)hipFree(p);
            q = h->allocations.back();
            h->allocations.pop_back();
            return;
        }
}

DevParams to_dev(const sw_params &p) {
    DevParams d{};
    d.g = p.g;
    d.epsilon = p.epsilon;
    d.H0 = p.H0;
    d.h0 = p.H0 * p.H0;
    d.limiting_threshold = 10 * p.H0;
    d.minimum_allowed_height = p.minimum_allowed_height;
    d.maximum_allowed_speed = p.maximum_allowed_speed;
    d.beta_w = p.beta_w;
    d.beta_w_dry = p.beta_w_dry;
    d.beta_uh = p.beta_uh;
    d.beta_uh_dry = p.beta_uh_dry;
    d.beta_vh = p.beta_vh;
    d.beta_vh_dry = p.beta_vh_dry;
    d.alpha_balance = p.alpha_balance;
    d.CFL = p.CFL;
    d.evolve_max_timestep = p.evolve_max_timestep;
    d.evolve_min_timestep = p.evolve_min_timestep;
    d.optimise_dry_cells = p.optimise_dry_cells;
    d.extrapolate_velocity_second_order = p.extrapolate_velocity_second_order;
    d.tight_slope_limiters = p.tight_slope_limiters;
    d.use_centroid_velocities = p.use_centroid_velocities;
    d.use_sloped_mannings = p.use_sloped_mannings;
    d.max_smallsteps = p.max_smallsteps;
    d.default_order = p.default_order;
    d.limiter = p.limiter;
    for (int i = 0; i < 3; i++) d.qbeta[i] = p.quantity_beta[i];
    return d;
}

int check_params(sw_handle *h, const sw_params *p) {
    if (!p) return fail(h, SW_ERR_ARG, "params is NULL");
    if (p->precision != SW_PRECISION_FP64 && p->precision != SW_PRECISION_MIXED)
        return fail(h, SW_ERR_ARG, "unknown precision %d", p->precision);
    if (p->locality_order != SW_ORDER_KEEP && p->locality_order != SW_ORDER_HILBERT)
        return fail(h, SW_ERR_ARG, "unknown locality order %d", p->locality_order);
    if (p->limiter < SW_LIMITER_OPTIMISED || p->limiter > SW_LIMITER_EDGE)
        return fail(h, SW_ERR_ARG, "unknown limiter %d", p->limiter);
    if (p->timestepping_method < 0 || p->timestepping_method > 3)
        return fail(h, SW_ERR_ARG, "unknown timestepping method %d (1 euler, 2 rk2, 3 rk3)", p->timestepping_method);
    if (p->default_order != 1 && p->default_order != 2)
        return fail(h, SW_ERR_ARG, "default_order must be 1 or 2");
    if (!(p->H0 > 0) || !(p->evolve_max_timestep > 0))
        return fail(h, SW_ERR_ARG, "H0 and evolve_max_timestep must be positive");
    return SW_OK;
}

unsigned grid_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

int launch_check(sw_handle *h, const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(h, SW_ERR_HIP, "launch of %s failed: %s", what, hipGetErrorString(e));
    return SW_OK;
}

// Turns the latched device errors of one instance's state into a return code.
int state_error(sw_handle *h, const DevState *out) {
    if (out->error) {
        if (out->error & kErrUpdate)
            return fail(h, SW_ERR_UPDATE, "update_conserved_quantities: semi-implicit denominator <= 0");
        if (out->error & kErrSmallsteps)
            return fail(h, SW_ERR_SMALLSTEPS,
                        "too small timestep reached even after %d steps of the first order scheme",
                        h->params.max_smallsteps);
        if (out->error & kErrNeighbour)
            return fail(h, SW_ERR_NEIGHBOUR, "extrapolate_second_order_sw: internal neighbour not found");
        if (out->error & kErrDiscontinuous)
            return fail(h, SW_ERR_ARG, "compute_fluxes: discontinuous elevation across an edge");
    }
    return SW_OK;
}

// Reads the device state of instance i and reports its latched errors.
int sync_state(sw_handle *h, int i, DevState *out) {
    HIPCHECK(h, hipMemcpyAsync(out, h->st + i, sizeof(DevState), hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    return state_error(h, out);
}

// Reports the first instance's latched error among the states `d`.
int states_error(sw_handle *h, std::vector<DevState> &d) {
    for (int i = 0; i < h->n_inst; i++)
        if (d[i].error) {
            int rc = state_error(h, &d[i]);
            if (h->n_inst > 1) h->err = "instance " + std::to_string(i) + ": " + h->err;
            return rc;
        }
    return SW_OK;
}

// Reads the states of all instances (through a pinned buffer); reports the
// first instance's error.
int sync_all(sw_handle *h, std::vector<DevState> &out) {
    out.resize(h->n_inst);
    HIPCHECK(h, hipMemcpyAsync(h->st_pinned, h->st, h->n_inst * sizeof(DevState), hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    std::memcpy(out.data(), h->st_pinned, h->n_inst * sizeof(DevState));
    return states_error(h, out);
}

// The states of all instances: the host's copy if it is current (kept by
// sw_evolve_steps and sw_advance_yield, which know what they leave on the
// device; dropped by every entry point that writes a state or runs device
// work on it), else read.
int current_states(sw_handle *h, std::vector<DevState> &out) {
    if (!h->st_host_ok) return sync_all(h, out);
    out = h->st_host;
    return states_error(h, out);
}

// A copy the host waits for, ordered on the handle's stream alone: a plain
// hipMemcpy runs on the null stream and so would also wait for whatever the
// caller has queued there (torch's default stream), e.g. learned steps that
// were meant to keep the GPU busy while the library's next call is set up.
hipError_t copy_wait(sw_handle *h, void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    return e;
}

int put_state(sw_handle *h, int i, const DevState &d) {
    h->st_host_ok = false;
    HIPCHECK(h, copy_wait(h, h->st + i, &d, sizeof d, hipMemcpyHostToDevice));
    return SW_OK;
}

// The instance that per-instance calls address: the selected one, or the only one.
int selected(sw_handle *h) {
    if (h->cur >= 0) return h->cur;
    if (h->n_inst == 1) return 0;
    return fail(h, SW_ERR_STATE, "ensemble handle: select an instance first (sw_select_instance)");
}

// Operator-by-operator entry points work on single-domain handles only.
int single(sw_handle *h, const char *what) {
    if (h->n_inst == 1) return SW_OK;
    return fail(h, SW_ERR_STATE, "%s: operator calls need a single-domain handle (this one holds an ensemble)", what);
}

// Where one (quantity, location) array of the API lives on the device: a
// per-triangle [N] array or a slot-major [3][N] array (slots = 1 or 3), either
// plain doubles or component `comp` of the {w, uh, vh, z} records.
struct Loc {
    double *ptr = nullptr;
    int slots = 1;
    int comp = -1;
    int64_t count = 0;  // values in the host array
    bool per_triangle = true;  // false: boundary array (caller's order)
};

// Array of instance `inst`: counts are the instance's; a boundary array's
// pointer is moved to the instance's first boundary edge.
bool locate(sw_handle *h, int inst, int q, int loc, Loc *out) {
    DevFields &F = h->f;
    const int64_t N = h->inst[inst].n, Nb = h->inst[inst].nb;
    Loc r;
    auto set = [&](void *p, int slots, int comp, int64_t count) {
        r.ptr = static_cast<double *>(p);
        r.slots = slots;
        r.comp = comp;
        r.count = count;
    };
    switch (q) {
        case SW_STAGE: case SW_XMOM: case SW_YMOM: case SW_ELEVATION: {
            const int comp = (q == SW_STAGE) ? 0 : (q == SW_XMOM) ? 1 : (q == SW_YMOM) ? 2 : 3;
            double *vert = (q == SW_STAGE) ? F.wv : (q == SW_XMOM) ? F.uhv : (q == SW_YMOM) ? F.vhv : F.zv;
            double *bnd = (q == SW_STAGE) ? F.wb : (q == SW_XMOM) ? F.uhb : (q == SW_YMOM) ? F.vhb : F.zb;
            double *eu = (q == SW_STAGE) ? F.weu : (q == SW_XMOM) ? F.uheu : (q == SW_YMOM) ? F.vheu : nullptr;
            double *si = (q == SW_STAGE) ? F.wsi : (q == SW_XMOM) ? F.uhsi : (q == SW_YMOM) ? F.vhsi : nullptr;
            switch (loc) {
                case SW_CENTROID: set(F.cq, 1, comp, N); break;
                case SW_EDGE: set(F.eq, 3, comp, 3 * N); break;
                case SW_VERTEX: set(vert, 3, -1, 3 * N); break;
                case SW_BOUNDARY: set(bnd, 1, -1, Nb); r.per_triangle = false; break;
                case SW_EXPLICIT_UPDATE: if (!eu) return false; set(eu, 1, -1, N); break;
                case SW_SEMI_IMPLICIT_UPDATE: if (!si) return false; set(si, 1, -1, N); break;
                default: return false;
            }
            break;
        }
        case SW_HEIGHT: case SW_XVELOCITY: case SW_YVELOCITY: {
            double *c = (q == SW_HEIGHT) ? F.hc : (q == SW_XVELOCITY) ? F.uc : F.vc;
            double *e = (q == SW_HEIGHT) ? F.he : (q == SW_XVELOCITY) ? F.ue : F.ve;
            double *v = (q == SW_HEIGHT) ? F.hv : (q == SW_XVELOCITY) ? F.uv : F.vv;
            double *b = (q == SW_HEIGHT) ? F.hb : (q == SW_XVELOCITY) ? F.ub : F.vb;
            switch (loc) {
                case SW_CENTROID: set(c, 1, -1, N); break;
                case SW_EDGE: set(e, 3, -1, 3 * N); break;
                case SW_VERTEX: set(v, 3, -1, 3 * N); break;
                case SW_BOUNDARY: set(b, 1, -1, Nb); r.per_triangle = false; break;
                default: return false;
            }
            break;
        }
        case SW_FRICTION:
            if (loc != SW_CENTROID) return false;
            set(F.friction, 1, -1, N);
            break;
        default: return false;
    }
    if (!r.per_triangle) r.ptr += h->inst[inst].boff;
    *out = r;
    return true;
}

__global__ void k_set_bed_interp(DevState *st, long long v) {
    st->bed_interp = v;
    st->bed_cont = v;
}

// Fused-kernel launches at one of the workgroup sizes the autotuner chooses
// from (128, 256, 512 lanes).
template <int B>
void launch_flux_b(sw_handle *h, bool mixed, int first) {
    const unsigned g = (unsigned)((h->N + B - 1) / B);
    if (mixed)
        hipLaunchKernelGGL((k_flux_mixed<B>), dim3(g), dim3(B), 0, h->stream, h->mesh, h->f, h->dp, h->st, first);
    else
        hipLaunchKernelGGL((k_flux<double, true, B>), dim3(g), dim3(B), 0, h->stream, h->mesh, h->f, h->dp, h->st,
                           first);
}
template <int B>
void launch_update_b(sw_handle *h, int mode, double a, double b) {
    const unsigned g = (unsigned)((h->N + B - 1) / B);
    hipLaunchKernelGGL((k_update<true, B>), dim3(g), dim3(B), 0, h->stream, h->mesh, h->f, h->dp, h->st, mode, a, b);
}
template <int B, bool LIMITED>
void launch_distribute_lb(sw_handle *h, bool mixed) {
    const unsigned g = (unsigned)((h->N + B - 1) / B);
    if (mixed)
        hipLaunchKernelGGL((k_distribute<float, true, B, LIMITED>), dim3(g), dim3(B), 0, h->stream, h->mesh, h->f,
                           h->dp, h->st);
    else
        hipLaunchKernelGGL((k_distribute<double, true, B, LIMITED>), dim3(g), dim3(B), 0, h->stream, h->mesh, h->f,
                           h->dp, h->st);
}
template <int B>
void launch_distribute_b(sw_handle *h, bool mixed) {
    if (h->params.limiter != SW_LIMITER_OPTIMISED)
        launch_distribute_lb<B, true>(h, mixed);
    else
        launch_distribute_lb<B, false>(h, mixed);
}

// Stages of a step: 1 (Euler), 2 (rk2) or 3 (rk3).
int rk_stages(const sw_params &p) {
    return p.timestepping_method == 2 ? 2 : p.timestepping_method == 3 ? 3 : 1;
}

// Weights (a, b) of q = a q + b q0 after stage `stage` (1-based, > 1) of an
// n-stage Runge-Kutta step (ANUGA's evolve_one_rk2_step / evolve_one_rk3_step).
void rk_weights(int stages, int stage, double *a, double *b) {
    if (stages == 2) {
        *a = 0.5, *b = 0.5;
    } else if (stage == 2) {
        *a = 0.25, *b = 0.75;
    } else {
        *a = 2.0 / 3.0, *b = 1.0 / 3.0;
    }
}

// Enqueues one fused step: flux, timestep, update (+ protect), distribute per
// stage; a Runge-Kutta step keeps its starting values in the first stage's
// update, combines them into the later stages' updates and finishes the step
// after the last stage's flux.  slot >= 0 records HIP events around the first
// stage's kernels into event group `slot` (timing mode).
int enqueue_fused_step(sw_handle *h, int slot) {
    const bool mixed = h->params.precision == SW_PRECISION_MIXED;
    const int stages = rk_stages(h->params);
    auto rec = [&](int idx) {
        if (slot >= 0) (void)hipEventRecord(h->ev[slot * (kNumTimed + 1) + idx], h->stream);
    };
    for (int stage = 1; stage <= stages; stage++) {
        const bool first = stage == 1;
        if (first) rec(0);
        switch (h->blk[0]) {
            case 128: launch_flux_b<128>(h, mixed, first); break;
            case 512: launch_flux_b<512>(h, mixed, first); break;
            default: launch_flux_b<256>(h, mixed, first); break;
        }
        if (first) rec(1);
        if (first || stage == stages) {
            const int mode = stages == 1 ? kStepEuler : first ? kStageFirst : kStageLast;
            hipLaunchKernelGGL((k_timestep<true>), dim3(h->n_inst), dim3(kReduceBlock), 0, h->stream, h->mesh,
                               h->f.wave_min, h->st, h->dp, mode);
        }
        if (first) rec(2);
        int mode = kUpdateEuler;
        double a = 0.0, b = 0.0;
        if (stages > 1 && first) mode = kUpdateBackup;
        if (!first) {
            mode = kUpdateSaxpy;
            rk_weights(stages, stage, &a, &b);
        }
        switch (h->blk[1]) {
            case 128: launch_update_b<128>(h, mode, a, b); break;
            case 512: launch_update_b<512>(h, mode, a, b); break;
            default: launch_update_b<256>(h, mode, a, b); break;
        }
        if (first) rec(3);
        switch (h->blk[2]) {
            case 128: launch_distribute_b<128>(h, mixed); break;
            case 512: launch_distribute_b<512>(h, mixed); break;
            default: launch_distribute_b<256>(h, mixed); break;
        }
        if (first) rec(4);
    }
    return launch_check(h, "fused step");
}

// Whether sw_evolve_steps runs the vertex-record step (sw_step_vertex.h) on the
// states `d` of the instances.
bool vstep_applies(const sw_handle *h, const std::vector<DevState> &d) {
    if (h->step_path == SW_STEP_EDGE || !h->vstep_geometry) return false;
    if (h->params.precision != SW_PRECISION_FP64) return false;
    for (const auto &s : d)
        if (!s.bed_interp || !s.bed_cont) return false;
    return true;
}

// Vertex-record step launches at one of the workgroup sizes the autotuner
// chooses from (the flux size for k_flux_v, the update size for k_update_v, the
// distribute size for k_extrap_v).
template <int B>
void launch_flux_v(sw_handle *h) {
    const unsigned g = (unsigned)((h->N + B - 1) / B);
    hipLaunchKernelGGL((k_flux_v<B>), dim3(g), dim3(B), 0, h->stream, h->mesh, h->f, h->sf, h->nbw, h->dp, h->vpar);
}
template <int B>
void launch_update_v(sw_handle *h, int mode, double a, double b) {
    const unsigned g = (unsigned)((h->N + B - 1) / B);
    hipLaunchKernelGGL((k_update_v<B>), dim3(g), dim3(B), 0, h->stream, h->mesh, h->f, h->sf, h->dp, h->vpar,
                       h->uni, mode, a, b);
}
template <int B>
void launch_extrap_v(sw_handle *h) {
    const unsigned g = (unsigned)((h->N + B - 1) / B);
    if (h->params.limiter != SW_LIMITER_OPTIMISED)
        hipLaunchKernelGGL((k_extrap_v<B, true>), dim3(g), dim3(B), 0, h->stream, h->mesh, h->f, h->sf, h->dp,
                           h->vpar);
    else
        hipLaunchKernelGGL((k_extrap_v<B>), dim3(g), dim3(B), 0, h->stream, h->mesh, h->f, h->sf, h->dp, h->vpar);
}

// Enqueues one vertex-record step (its Runge-Kutta stages, one parity each)
// starting at parity h->vpar; the timing slots are the first stage's flux
// (k_flux_v), update (k_update_v) and distribute (k_extrap_v).
int enqueue_vstep(sw_handle *h, int slot) {
    const int stages = rk_stages(h->params);
    auto rec = [&](int idx) {
        if (slot >= 0) (void)hipEventRecord(h->ev[slot * (kNumTimed + 1) + idx], h->stream);
    };
    for (int stage = 1; stage <= stages; stage++) {
        const bool first = stage == 1;
        int mode = kStepEuler;
        double a = 0.0, b = 0.0;
        if (stages > 1) {
            mode = first ? kStageFirst : stage == stages ? kStageLast : kStageMid;
            if (!first) rk_weights(stages, stage, &a, &b);
        }
        if (first) rec(0);
        switch (h->blk[0]) {
            case 128: launch_flux_v<128>(h); break;
            case 512: launch_flux_v<512>(h); break;
            default: launch_flux_v<256>(h); break;
        }
        if (first) rec(1);
        if (first) rec(2);
        switch (h->blk[1]) {
            case 128: launch_update_v<128>(h, mode, a, b); break;
            case 512: launch_update_v<512>(h, mode, a, b); break;
            default: launch_update_v<256>(h, mode, a, b); break;
        }
        if (first) rec(3);
        switch (h->blk[2]) {
            case 128: launch_extrap_v<128>(h); break;
            case 512: launch_extrap_v<512>(h); break;
            default: launch_extrap_v<256>(h); break;
        }
        if (first) rec(4);
        h->vpar ^= 1;
    }
    return launch_check(h, "vertex-record step");
}

int enqueue_step(sw_handle *h, bool vstep, int slot) { return vstep ? enqueue_vstep(h, slot) : enqueue_fused_step(h, slot); }

// Accumulates the kernel durations of event groups [0, nslots).  A kernel that
// returned at once because the loop was halted still counts as a launch.
int collect_timing(sw_handle *h, int nslots) {
    if (nslots <= 0) return SW_OK;
    HIPCHECK(h, hipEventSynchronize(h->ev[nslots * (kNumTimed + 1) - 1]));
    for (int s = 0; s < nslots; s++) {
        for (int i = 0; i < kNumTimed; i++) {
            float ms = 0;
            hipEvent_t *e = &h->ev[s * (kNumTimed + 1)];
            HIPCHECK(h, hipEventElapsedTime(&ms, e[i], e[i + 1]));
            h->times_ms[i] += ms;
            h->launches[i] += 1;
        }
    }
    return SW_OK;
}

void drop_graph(sw_handle *h);

// Captures kGraphSteps steps of the given path (an even number, so a
// vertex-record graph starts and ends at parity 0).
int build_graph(sw_handle *h, bool vstep) {
    const int path = vstep ? SW_STEP_VERTEX : SW_STEP_EDGE;
    if (h->graph_exec && h->graph_path == path) return SW_OK;
    drop_graph(h);
    h->graph_path = path;
    HIPCHECK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    int rc = SW_OK;
    for (int s = 0; s < kGraphSteps && rc == SW_OK; s++) rc = enqueue_step(h, vstep, -1);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(h->stream, &graph);
    if (rc) return rc;
    if (e != hipSuccess) return fail(h, SW_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    h->graph = graph;
    HIPCHECK(h, hipGraphInstantiate(&h->graph_exec, graph, nullptr, nullptr, 0));
    // the graph's device-side set-up happens now, not at its first launch
    HIPCHECK(h, hipGraphUpload(h->graph_exec, h->stream));
    return SW_OK;
}

void drop_graph(sw_handle *h) {
    if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
    if (h->graph) (void)hipGraphDestroy(h->graph);
    h->graph_exec = nullptr;
    h->graph = nullptr;
}

// Boundary arrays, other quantities and the semi-implicit reset that the
// fused loop defers to the end of a call (they do not feed back into the step).
int epilogue(sw_handle *h) {
    const unsigned g = grid_for(h->N);
    const unsigned gb = grid_for(h->Nb > h->N ? h->Nb : h->N);
    hipLaunchKernelGGL(k_update_boundary, dim3(grid_for(h->Nb) ? grid_for(h->Nb) : 1), dim3(kBlock), 0,
                       h->stream, h->mesh, h->f);
    hipLaunchKernelGGL(k_other_quantities, dim3(gb), dim3(kBlock), 0, h->stream, h->mesh, h->f);
    hipLaunchKernelGGL(k_fill, dim3(g), dim3(kBlock), 0, h->stream, h->f.uhsi, (long long)h->N, 0.0);
    hipLaunchKernelGGL(k_fill, dim3(g), dim3(kBlock), 0, h->stream, h->f.vhsi, (long long)h->N, 0.0);
    return launch_check(h, "evolve epilogue");
}

// Brings the operator API's arrays up to the vertex-record step's records
// (k_step_end, then the epilogue of the evolve call), if they are behind.
// Entry points that read those arrays call it first.
int sync_api(sw_handle *h) {
    if (!h->api_pending) return SW_OK;
    h->api_pending = false;
    hipLaunchKernelGGL(k_step_end, dim3(grid_for(h->N)), dim3(kBlock), 0, h->stream, h->mesh, h->f, h->sf, h->vpar);
    if (int rc = launch_check(h, "k_step_end")) return rc;
    return epilogue(h);
}

// Whether the operator API's array of (quantity, location) is current even
// while the others are behind the vertex-record step's records: the centroid
// values of the conserved quantities and of the bed live in the centroid
// records, which that step updates in place.
bool record_current(int quantity, int location) {
    return location == SW_CENTROID && (quantity == SW_STAGE || quantity == SW_XMOM || quantity == SW_YMOM ||
                                       quantity == SW_ELEVATION || quantity == SW_FRICTION);
}

// Entry points that may write the operator API's arrays or change what the
// step reads: the arrays are brought up to date first, and the next evolve
// call rebuilds the step's records from them.
int touch_api(sw_handle *h) {
    const int rc = sync_api(h);
    h->records_live = false;
    h->st_host_ok = false;
    return rc;
}

// Position of cell (x, y) along the Hilbert curve filling a 2^16 x 2^16 grid
// (the iterative xy -> d conversion with quadrant rotation).
uint64_t hilbert_index(uint32_t x, uint32_t y) {
    const uint32_t n = 1u << 16;
    uint64_t d = 0;
    for (uint32_t s = n / 2; s > 0; s /= 2) {
        const uint32_t rx = (x & s) ? 1u : 0u, ry = (y & s) ? 1u : 0u;
        d += (uint64_t)s * s * ((3u * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) {
                x = n - 1 - x;
                y = n - 1 - y;
            }
            std::swap(x, y);
        }
    }
    return d;
}

// cal_of = the triangles sorted by the Hilbert index of their centroids
// (ties by caller index), so neighbours are close in memory and in the
// workgroup / XCD sweep.
void hilbert_order(const double *centroids, int64_t N, std::vector<int> &cal_of) {
    double x0 = centroids[0], x1 = x0, y0 = centroids[1], y1 = y0;
    for (int64_t k = 1; k < N; k++) {
        x0 = std::min(x0, centroids[2 * k]);
        x1 = std::max(x1, centroids[2 * k]);
        y0 = std::min(y0, centroids[2 * k + 1]);
        y1 = std::max(y1, centroids[2 * k + 1]);
    }
    const double span = std::max(std::max(x1 - x0, y1 - y0), 1e-300);
    const double scale = 65535.0 / span;
    std::vector<std::pair<uint64_t, int>> key(N);
    for (int64_t k = 0; k < N; k++) {
        const double fx = (centroids[2 * k] - x0) * scale, fy = (centroids[2 * k + 1] - y0) * scale;
        const uint32_t ix = (uint32_t)std::min(std::max(fx, 0.0), 65535.0);
        const uint32_t iy = (uint32_t)std::min(std::max(fy, 0.0), 65535.0);
        key[k] = {hilbert_index(ix, iy), (int)k};
    }
    std::sort(key.begin(), key.end());
    for (int64_t d = 0; d < N; d++) cal_of[d] = key[d].second;
}

}  // namespace

extern "C" {

const char *sw_last_error(const sw_handle *h) { return h ? h->err.c_str() : g_create_error.c_str(); }

}  // extern "C"

namespace {

// Validates instance `m` and packs it into the device arrays at triangle offset
// I.off and boundary offset I.boff (device numbering of the whole handle).
bool pack_instance(const sw_mesh *m, const sw_params *p, const Inst &I, int64_t Ntot, int inst_index,
                   std::vector<EdgeRec> &edgev, std::vector<EdgeRecF> &edgefv, std::vector<int> &flagsv,
                   std::vector<int> &surv, std::vector<double> &vcv, std::vector<double> &ccv,
                   std::vector<double> &rav, std::vector<int> &cal_of, std::vector<int> &dev_of,
                   std::vector<int> &bcellv, std::vector<int> &bedgev, bool &own_normals,
                   std::vector<int> &nbwv, bool &geom_same) {
    const int64_t N = I.n, Nb = I.nb;
    auto bad = [&](const char *fmt, long long a, long long b = 0, long long c = 0) {
        char buf[256];
        snprintf(buf, sizeof buf, fmt, a, b, c);
        fail(nullptr, SW_ERR_ARG, "mesh %d: %s", inst_index, buf);
        return false;
    };
    const int64_t *arrays_i[] = {m->neighbours, m->neighbour_edges, m->surrogate_neighbours,
                                 m->number_of_boundaries, m->tri_full_flag};
    const double *arrays_f[] = {m->normals, m->edgelengths, m->radii, m->areas,
                                m->centroid_coordinates, m->vertex_coordinates};
    for (auto q : arrays_i)
        if (!q) return bad("a mesh index array is NULL", 0);
    for (auto q : arrays_f)
        if (!q) return bad("a mesh geometry array is NULL", 0);
    if (Nb > 0 && (!m->boundary_cells || !m->boundary_edges)) return bad("boundary arrays are NULL", 0);
    // device numbering within the instance: caller triangle k at I.off + loc[k]
    int *cal = cal_of.data() + I.off;
    int *loc = dev_of.data() + I.poff;
    if (p->locality_order == SW_ORDER_HILBERT) {
        std::vector<int> order(N);
        hilbert_order(m->centroid_coordinates, N, order);
        for (int64_t d = 0; d < N; d++) cal[d] = order[d];
    } else {
        for (int64_t d = 0; d < N; d++) cal[d] = (int)d;
    }
    for (int64_t d = 0; d < N; d++) loc[cal[d]] = (int)d;
    // Flags (which side evaluates an edge, summation order, full flag of the
    // evaluating triangle) come from the caller's numbering.
    for (int64_t k = 0; k < N; k++) {
        const int64_t d = I.off + loc[k];
        int flags = 0;
        int64_t key[3];
        int nbslot[3];
        for (int i = 0; i < 3; i++) {
            const int64_t ki = 3 * k + i;
            const int64_t n = m->neighbours[ki];
            int64_t computing = k;
            key[i] = (int64_t)1 << 62 | i;  // own edges after lower neighbours, in local order
            if (n < 0) {
                const int64_t b = -n - 1;
                if (b >= Nb || m->boundary_cells[b] != k || m->boundary_edges[b] != i)
                    return bad("neighbours[%lld] = %lld does not match boundary entry %lld", ki, n, b);
                nbslot[i] = (int)(-(I.boff + b) - 1);
                nbwv[i * Ntot + d] = nbslot[i];
            } else {
                const int64_t e = m->neighbour_edges[ki];
                if (n >= N || e < 0 || e > 2 || m->neighbours[3 * n + e] != k)
                    return bad("neighbour structure is not symmetric at edge %lld", ki);
                nbslot[i] = (int)(e * Ntot + I.off + loc[n]);
                nbwv[i * Ntot + d] = (int)((e << 29) | (I.off + loc[n]));
                if (n < k) {
                    flags |= 1 << i;
                    computing = n;
                    key[i] = n;
                }
                const int64_t ne = 3 * n + e;
                if (m->normals[2 * ne] != -m->normals[2 * ki] || m->normals[2 * ne + 1] != -m->normals[2 * ki + 1] ||
                    m->edgelengths[ne] != m->edgelengths[ki])
                    own_normals = false;
            }
            if (m->tri_full_flag[computing] == 1) flags |= 1 << (3 + i);
            const int64_t sn = m->surrogate_neighbours[ki];
            if (sn < 0 || sn >= N) return bad("surrogate_neighbours[%lld] out of range", ki);
            surv[i * Ntot + d] = (int)(I.off + loc[sn]);
            EdgeRec &er = edgev[i * Ntot + d];
            er.nx = m->normals[2 * ki];
            er.ny = m->normals[2 * ki + 1];
            er.len = m->edgelengths[ki];
            er.nb = nbslot[i];
            edgefv[i * Ntot + d] = EdgeRecF{(float)er.nx, (float)er.ny, (float)er.len, er.nb};
            vcv[2 * (i * Ntot + d)] = m->vertex_coordinates[6 * k + 2 * i];
            vcv[2 * (i * Ntot + d) + 1] = m->vertex_coordinates[6 * k + 2 * i + 1];
        }
        // summation order of the C edge-once loop: lower neighbours ascending, then own edges
        int ord[3] = {0, 1, 2};
        for (int a = 0; a < 3; a++)
            for (int b = a + 1; b < 3; b++)
                if (key[ord[b]] < key[ord[a]]) std::swap(ord[a], ord[b]);
        flags |= (ord[0] | ord[1] << 2 | ord[2] << 4) << 6;
        const int64_t nob = m->number_of_boundaries[k];
        if (nob < 0 || nob > 3) return bad("number_of_boundaries[%lld] = %lld", k, nob);
        flags |= (int)nob << 12;
        if (nob == 2) {
            // the reconstruction of a triangle with two boundary edges uses its
            // first surrogate neighbour that is not itself: keep it in slot 0
            for (int i = 0; i < 3; i++)
                if (surv[i * Ntot + d] != d) {
                    std::swap(surv[d], surv[i * Ntot + d]);
                    break;
                }
        }
        flagsv[d] = flags;
        for (int i = 0; i < 3; i++) edgev[i * Ntot + d].tflags = flags;
        if (geom_same) {
            // the vertex-record step recomputes the geometry from the vertex coordinates
            double x[3], y[3];
            for (int i = 0; i < 3; i++) {
                x[i] = m->vertex_coordinates[6 * k + 2 * i];
                y[i] = m->vertex_coordinates[6 * k + 2 * i + 1];
            }
            TriGeom g;
            tri_geometry(x, y, g);
            bool same = g.radius == m->radii[k] && g.area == m->areas[k] && g.cx == m->centroid_coordinates[2 * k] &&
                        g.cy == m->centroid_coordinates[2 * k + 1];
            for (int i = 0; i < 3; i++)
                same = same && g.nx[i] == m->normals[6 * k + 2 * i] && g.ny[i] == m->normals[6 * k + 2 * i + 1] &&
                       g.len[i] == m->edgelengths[3 * k + i];
            geom_same = same;
        }
        rav[2 * d] = m->radii[k];
        rav[2 * d + 1] = m->areas[k];
        ccv[2 * d] = m->centroid_coordinates[2 * k];
        ccv[2 * d + 1] = m->centroid_coordinates[2 * k + 1];
    }
    for (int64_t b = 0; b < Nb; b++) {
        const int64_t c = m->boundary_cells[b], e = m->boundary_edges[b];
        if (c < 0 || c >= N || e < 0 || e > 2) return bad("boundary entry %lld out of range", b);
        bcellv[I.boff + b] = (int)(I.off + loc[c]);
        bedgev[I.boff + b] = (int)e;
    }
    return true;
}

// Mesh nodes of the vertex-record step: the vertex slots of each instance whose
// coordinates are bitwise equal share a node.  Nodes are numbered in the
// order the device numbering first reaches them, so the node arrays follow
// the triangles' locality order.  tri_node[i * Ntot + d] is the node of
// vertex i of device triangle d (padding triangles: node 0); node_xy holds
// the coordinates.  Open addressing over the coordinate bits.
void build_nodes(const std::vector<double> &vcv, int64_t Ntot, const std::vector<Inst> &inst,
                 std::vector<int> &tri_node, std::vector<double> &node_xy) {
    tri_node.assign(3 * Ntot, 0);
    node_xy.clear();
    for (const Inst &I : inst) {
        size_t cap = 16;
        while (cap < (size_t)(3 * I.n) * 2) cap <<= 1;
        std::vector<int> table(cap, -1);  // node id, or -1
        for (int64_t d = I.off; d < I.off + I.n; d++) {
            for (int i = 0; i < 3; i++) {
                const double x = vcv[2 * (i * Ntot + d)], y = vcv[2 * (i * Ntot + d) + 1];
                uint64_t bx, by;
                std::memcpy(&bx, &x, 8);
                std::memcpy(&by, &y, 8);
                uint64_t hsh = (bx * 0x9E3779B97F4A7C15ull) ^ (by + 0x632BE59BD9B4E019ull + (bx << 6) + (bx >> 2));
                hsh ^= hsh >> 29;
                size_t s = (size_t)(hsh * 0xBF58476D1CE4E5B9ull) & (cap - 1);
                int node = -1;
                while (table[s] >= 0) {
                    const int c = table[s];
                    uint64_t cx, cy;
                    std::memcpy(&cx, &node_xy[2 * c], 8);
                    std::memcpy(&cy, &node_xy[2 * c + 1], 8);
                    if (cx == bx && cy == by) {
                        node = c;
                        break;
                    }
                    s = (s + 1) & (cap - 1);
                }
                if (node < 0) {
                    node = (int)(node_xy.size() / 2);
                    table[s] = node;
                    node_xy.push_back(x);
                    node_xy.push_back(y);
                }
                tri_node[i * Ntot + d] = node;
            }
        }
    }
    if (node_xy.empty()) node_xy.assign(2, 0.0);
}

sw_handle *create_impl(const sw_mesh *meshes, int n_inst, const sw_params *p, int device) {
    g_create_error.clear();
    if (!meshes || n_inst < 1) {
        fail(nullptr, SW_ERR_ARG, "no meshes");
        return nullptr;
    }
    if (check_params(nullptr, p)) return nullptr;
    // instance layout: each instance starts on a multiple of kInstAlign triangles
    std::vector<Inst> inst(n_inst);
    int64_t Ntot = 0, Nbtot = 0, Nsum = 0, Nmax = 0;
    for (int i = 0; i < n_inst; i++) {
        const int64_t n = meshes[i].number_of_elements, nb = meshes[i].number_of_boundary_elements;
        if (n <= 0 || nb < 0 || nb > 3 * n) {
            fail(nullptr, SW_ERR_ARG, "mesh %d: bad sizes N=%lld Nb=%lld", i, (long long)n, (long long)nb);
            return nullptr;
        }
        inst[i].off = (Ntot + kInstAlign - 1) / kInstAlign * kInstAlign;
        inst[i].n = n;
        inst[i].boff = Nbtot;
        inst[i].nb = nb;
        inst[i].poff = Nsum;
        Ntot = inst[i].off + n;
        Nbtot += nb;
        Nsum += n;
        Nmax = std::max(Nmax, n);
    }
    if (Ntot > (int64_t)0x2AAAA000) {
        fail(nullptr, SW_ERR_ARG, "too many triangles (%lld) for 32-bit edge slots", (long long)Ntot);
        return nullptr;
    }
    // padding triangles: no neighbours (nb = -1 marks them out of every edge test)
    std::vector<EdgeRec> edgev(3 * Ntot, EdgeRec{0.0, 0.0, 0.0, -1, 0});
    std::vector<EdgeRecF> edgefv(3 * Ntot, EdgeRecF{0.f, 0.f, 0.f, -1});
    std::vector<int> flagsv(Ntot, 0), surv(3 * Ntot, 0), cal_of(Ntot, 0), dev_of(Nsum, 0);
    std::vector<double> vcv(6 * Ntot, 0.0), ccv(2 * Ntot, 0.0), rav(2 * Ntot, 0.0);
    std::vector<int> bcellv(std::max<int64_t>(Nbtot, 1), 0), bedgev(std::max<int64_t>(Nbtot, 1), 0);
    bool own_normals = true, geom_same = true;
    std::vector<int> nbwv(3 * Ntot, -1);
    for (int i = 0; i < n_inst; i++)
        if (!pack_instance(&meshes[i], p, inst[i], Ntot, i, edgev, edgefv, flagsv, surv, vcv, ccv, rav, cal_of,
                           dev_of, bcellv, bedgev, own_normals, nbwv, geom_same))
            return nullptr;
    const int64_t nwaves = (Ntot + kInstAlign - 1) / kInstAlign * (kInstAlign / 64);
    std::vector<int> wave_inst(nwaves, n_inst - 1), inst_off(n_inst), inst_end(n_inst);
    for (int i = 0; i < n_inst; i++) {
        inst_off[i] = (int)inst[i].off;
        inst_end[i] = (int)(inst[i].off + inst[i].n);
        const int64_t w1 = (i + 1 < n_inst) ? inst[i + 1].off / 64 : nwaves;
        for (int64_t w = inst[i].off / 64; w < w1; w++) wave_inst[w] = i;
    }

    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) {
        fail(nullptr, SW_ERR_HIP, "hipSetDevice(%d): %s", device, hipGetErrorString(e));
        return nullptr;
    }
    sw_handle *h = new sw_handle();
    h->device = device;
    h->N = Ntot;
    h->Nb = Nbtot;
    h->n_inst = n_inst;
    h->inst = inst;
    h->sww.resize(n_inst);
    h->params = *p;
    h->dp = to_dev(*p);
    auto bail = [&](void) -> sw_handle * {
        g_create_error = h->err;
        sw_destroy(h);
        return nullptr;
    };
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        fail(h, SW_ERR_HIP, "hipStreamCreate failed");
        return bail();
    }
    auto put = [&](auto **dst, const auto &vec) -> int {
        if (dalloc(h, dst, vec.size())) return -1;
        if (copy_wait(h, *dst, vec.data(), vec.size() * sizeof(vec[0]), hipMemcpyHostToDevice) != hipSuccess)
            return fail(h, SW_ERR_HIP, "mesh upload failed");
        return 0;
    };
    DevMesh &M = h->mesh;
    M.N = (int)Ntot;
    M.Nb = (int)Nbtot;
    M.own_normals = own_normals ? 1 : 0;
    M.n_inst = n_inst;
    EdgeRec *dedge;
    EdgeRecF *dedgef;
    int *dflags, *dsur, *bcell, *bedge, *bkind, *dwave, *doff, *dend;
    double *dvc, *dcc, *dra, *bdir;
    if (put(&dedge, edgev) || put(&dedgef, edgefv) || put(&dflags, flagsv) || put(&dsur, surv) || put(&dvc, vcv) ||
        put(&dcc, ccv) || put(&dra, rav) || put(&h->cal_of, cal_of) || put(&h->dev_of, dev_of) ||
        put(&bcell, bcellv) || put(&bedge, bedgev) || put(&dwave, wave_inst) || put(&doff, inst_off) ||
        put(&dend, inst_end))
        return bail();
    if (dalloc(h, &bkind, Nbtot) || dalloc(h, &bdir, 3 * Nbtot)) return bail();
    M.edge = dedge;
    M.edgef = dedgef;
    M.flags = dflags;
    M.sur = dsur;
    M.vc = reinterpret_cast<const double2 *>(dvc);
    M.cc = reinterpret_cast<const double2 *>(dcc);
    M.ra = reinterpret_cast<const double2 *>(dra);
    M.bcell = bcell;
    M.bedge = bedge;
    M.bkind = bkind;
    M.bdir = bdir;
    M.inst_off = doff;
    M.inst_end = dend;
    M.wave_inst = dwave;

    DevFields &F = h->f;
    const int64_t N = Ntot, Nb = Nbtot;
    double **perN[] = {&F.weu, &F.uheu, &F.vheu, &F.wsi, &F.uhsi, &F.vhsi, &F.hc, &F.uc, &F.vc,
                       &F.friction, &F.zbar, &F.max_speed};
    double **per3N[] = {&F.wv, &F.uhv, &F.vhv, &F.zv, &F.ze, &F.he, &F.hv, &F.ue, &F.uv, &F.ve, &F.vv};
    double **perNb[] = {&F.wb, &F.uhb, &F.vhb, &F.zb, &F.hb, &F.ub, &F.vb};
    for (auto pp : perN)
        if (dalloc(h, pp, N)) return bail();
    for (auto pp : per3N)
        if (dalloc(h, pp, 3 * N)) return bail();
    for (auto pp : perNb)
        if (dalloc(h, pp, Nb)) return bail();
    if (dalloc(h, &F.cq, N) || dalloc(h, &F.cq0, N) || dalloc(h, &F.eq, 3 * N) || dalloc(h, &F.grav, N) ||
        dalloc(h, &F.wave_min, (size_t)nwaves) || dalloc(h, &F.eqf, 3 * N) || dalloc(h, &h->staging, 3 * Nmax))
        return bail();
    if (dalloc(h, &h->st, n_inst)) return bail();
    if (hipHostMalloc((void **)&h->st_pinned, n_inst * sizeof(DevState), hipHostMallocDefault) != hipSuccess) {
        h->st_pinned = nullptr;
        fail(h, SW_ERR_HIP, "hipHostMalloc of the state buffer failed");
        return bail();
    }
    if (hipHostMalloc((void **)&h->st_up, 2 * n_inst * sizeof(DevState), hipHostMallocDefault) != hipSuccess) {
        h->st_up = nullptr;
        fail(h, SW_ERR_HIP, "hipHostMalloc of the state staging buffer failed");
        return bail();
    }
    h->vstep_geometry = own_normals && geom_same && Ntot < ((int64_t)1 << 29);
    {
        std::vector<int> tri_node;
        std::vector<double> node_xy;
        build_nodes(vcv, Ntot, inst, tri_node, node_xy);
        int *dtn;
        double *dnxy;
        if (put(&dtn, tri_node) || put(&dnxy, node_xy) || dalloc(h, &F.node_z, node_xy.size() / 2)) return bail();
        M.tri_node = dtn;
        M.node_xy = reinterpret_cast<const double2 *>(dnxy);
    }
    {
        StepFields &S = h->sf;
        if (put(&h->nbw, nbwv) || dalloc(h, &S.vr, 3 * N) || dalloc(h, &S.eur, N) || dalloc(h, &S.hc, N) ||
            dalloc(h, &S.slot, (size_t)2 * n_inst * kSlotBlock) ||
            dalloc(h, &S.errw, n_inst) || dalloc(h, &S.stv[1], n_inst) || dalloc(h, &S.uni, n_inst))
            return bail();
        S.stv[0] = h->st;
    }
    std::vector<DevState> s0(n_inst);
    for (auto &d : s0) {
        d = DevState{};
        d.order = p->default_order;
        d.bed_interp = 1;
        d.bed_cont = 1;
    }
    if (copy_wait(h, h->st, s0.data(), n_inst * sizeof(DevState), hipMemcpyHostToDevice) != hipSuccess) {
        fail(h, SW_ERR_HIP, "state upload failed");
        return bail();
    }
    return h;
}

}  // namespace

extern "C" {

sw_handle *sw_create(const sw_mesh *m, const sw_params *p, int device) {
    if (!m) {
        g_create_error = "mesh is NULL";
        return nullptr;
    }
    return create_impl(m, 1, p, device);
}

sw_handle *sw_create_ensemble(const sw_mesh *meshes, int32_t n_instances, const sw_params *p, int device) {
    return create_impl(meshes, n_instances, p, device);
}

int sw_select_instance(sw_handle *h, int32_t instance) {
    if (!h) return SW_ERR_ARG;
    if (instance < -1 || instance >= h->n_inst)
        return fail(h, SW_ERR_ARG, "instance %d outside [-1, %d)", instance, h->n_inst);
    h->cur = instance;
    return SW_OK;
}

int sw_destroy(sw_handle *h) {
    if (!h) return SW_OK;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    drop_graph(h);
    for (auto e : h->ev)
        if (e) (void)hipEventDestroy(e);
    for (void *p : h->allocations) (void)hipFree(p);
    if (h->st_pinned) (void)hipHostFree(h->st_pinned);
    if (h->st_up) (void)hipHostFree(h->st_up);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return SW_OK;
}

int sw_set_params(sw_handle *h, const sw_params *p) {
    if (!h) return SW_ERR_ARG;
    if (int rc = touch_api(h)) return rc;
    if (int rc = check_params(h, p)) return rc;
    h->params = *p;
    h->dp = to_dev(*p);
    drop_graph(h);  // kernel arguments changed
    return SW_OK;
}

int sw_get_params(const sw_handle *h, sw_params *p) {
    if (!h || !p) return SW_ERR_ARG;
    *p = h->params;
    return SW_OK;
}

int sw_set_boundary(sw_handle *h, const int32_t *kind, const double *dirichlet) {
    if (!h) return SW_ERR_ARG;
    if (int rc = sync_api(h)) return rc;
    const int i = selected(h);
    if (i < 0) return i;
    const Inst &I = h->inst[i];
    if (I.nb == 0) return SW_OK;
    if (!kind) return fail(h, SW_ERR_ARG, "boundary kind array is NULL");
    std::vector<int> k(I.nb);
    for (int64_t b = 0; b < I.nb; b++) {
        if (kind[b] < 0 || kind[b] > 2) return fail(h, SW_ERR_ARG, "boundary kind[%lld]=%d", (long long)b, kind[b]);
        if (kind[b] == SW_BC_DIRICHLET && !dirichlet)
            return fail(h, SW_ERR_ARG, "Dirichlet boundary without values");
        k[b] = kind[b];
    }
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    HIPCHECK(h, copy_wait(h, (void *)(h->mesh.bkind + I.boff), k.data(), I.nb * sizeof(int), hipMemcpyHostToDevice));
    if (dirichlet)
        HIPCHECK(h, copy_wait(h, (void *)(h->mesh.bdir + 3 * I.boff), dirichlet, 3 * I.nb * sizeof(double),
                              hipMemcpyHostToDevice));
    return SW_OK;
}

int sw_upload(sw_handle *h, int quantity, int location, const double *host, int64_t count) {
    if (!h || !host) return SW_ERR_ARG;
    if (int rc = touch_api(h)) return rc;
    const int i = selected(h);
    if (i < 0) return i;
    const Inst &I = h->inst[i];
    Loc L;
    if (!locate(h, i, quantity, location, &L))
        return fail(h, SW_ERR_ARG, "no array for quantity %d location %d", quantity, location);
    if (count != L.count)
        return fail(h, SW_ERR_ARG, "upload of %lld values into an array of %lld", (long long)count, (long long)L.count);
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    if (L.count == 0) return SW_OK;
    if (!L.per_triangle) {
        HIPCHECK(h, copy_wait(h, L.ptr, host, L.count * sizeof(double), hipMemcpyHostToDevice));
    } else {
        HIPCHECK(h, copy_wait(h, h->staging, host, L.count * sizeof(double), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_to_device, dim3(grid_for(L.count)), dim3(kBlock), 0, h->stream, (long long)I.n,
                           (long long)h->N, (long long)I.off, L.slots, L.comp, h->cal_of + I.off, h->staging, L.ptr);
        if (int rc = launch_check(h, "k_to_device")) return rc;
        if (quantity == SW_ELEVATION && location == SW_EDGE) {
            hipLaunchKernelGGL(k_to_device, dim3(grid_for(L.count)), dim3(kBlock), 0, h->stream, (long long)I.n,
                               (long long)h->N, (long long)I.off, 3, -1, h->cal_of + I.off, h->staging, h->f.ze);
            if (int rc = launch_check(h, "k_to_device")) return rc;
        }
    }
    if (quantity == SW_ELEVATION && (location == SW_VERTEX || location == SW_EDGE)) {
        hipLaunchKernelGGL(k_set_bed_interp, dim3(1), dim3(1), 0, h->stream, h->st + i, 1ll);
        hipLaunchKernelGGL(k_node_bed, dim3(grid_for(I.n)), dim3(kBlock), 0, h->stream, (int)h->N, (int)I.off,
                           (int)I.n, (const double *)h->f.zv, h->mesh.tri_node, h->f.node_z);
        hipLaunchKernelGGL(k_bed_derived, dim3(grid_for(I.n)), dim3(kBlock), 0, h->stream, (int)h->N, (int)I.off,
                           (int)I.n, h->f.zv, h->f.ze, h->f.zbar, h->vstep_geometry ? h->nbw : nullptr,
                           h->mesh.tri_node, (const double *)h->f.node_z, h->st + i);
        if (int rc = launch_check(h, "k_bed_derived")) return rc;
    }
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    return SW_OK;
}

int sw_download(sw_handle *h, int quantity, int location, double *host, int64_t count) {
    if (!h || !host) return SW_ERR_ARG;
    if (!record_current(quantity, location))
        if (int rc = sync_api(h)) return rc;
    const int i = selected(h);
    if (i < 0) return i;
    const Inst &I = h->inst[i];
    Loc L;
    if (!locate(h, i, quantity, location, &L))
        return fail(h, SW_ERR_ARG, "no array for quantity %d location %d", quantity, location);
    if (count != L.count)
        return fail(h, SW_ERR_ARG, "download of %lld values from an array of %lld", (long long)count, (long long)L.count);
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    if (L.count == 0) return SW_OK;
    if (!L.per_triangle) {
        HIPCHECK(h, copy_wait(h, host, L.ptr, L.count * sizeof(double), hipMemcpyDeviceToHost));
    } else {
        hipLaunchKernelGGL(k_to_host, dim3(grid_for(L.count)), dim3(kBlock), 0, h->stream, (long long)I.n,
                           (long long)h->N, (long long)I.off, L.slots, L.comp, h->dev_of + I.poff, L.ptr, h->staging);
        if (int rc = launch_check(h, "k_to_host")) return rc;
        HIPCHECK(h, hipStreamSynchronize(h->stream));
        HIPCHECK(h, copy_wait(h, host, h->staging, L.count * sizeof(double), hipMemcpyDeviceToHost));
    }
    return SW_OK;
}

// Device-memory forms of sw_upload / sw_download: the caller's array (same
// layout and order as for the host forms) lives in device memory.
// Device-to-device transfers of one array between a caller's buffer (caller
// order) and the handle, enqueued on the handle's stream without waiting.
static int upload_async(sw_handle *h, int i, int quantity, int location, const double *dsrc, int64_t count) {
    const Inst &I = h->inst[i];
    Loc L;
    if (!locate(h, i, quantity, location, &L))
        return fail(h, SW_ERR_ARG, "no array for quantity %d location %d", quantity, location);
    if (count != L.count)
        return fail(h, SW_ERR_ARG, "upload of %lld values into an array of %lld", (long long)count, (long long)L.count);
    if (quantity == SW_ELEVATION)
        return fail(h, SW_ERR_ARG, "elevation is uploaded from host memory (sw_upload)");
    if (L.count == 0) return SW_OK;
    if (!L.per_triangle) {
        HIPCHECK(h, hipMemcpyAsync(L.ptr, dsrc, L.count * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
        return SW_OK;
    }
    hipLaunchKernelGGL(k_to_device, dim3(grid_for(L.count)), dim3(kBlock), 0, h->stream, (long long)I.n,
                       (long long)h->N, (long long)I.off, L.slots, L.comp, h->cal_of + I.off, dsrc, L.ptr);
    return launch_check(h, "k_to_device");
}

static int download_async(sw_handle *h, int i, int quantity, int location, double *ddst, int64_t count) {
    const Inst &I = h->inst[i];
    Loc L;
    if (!locate(h, i, quantity, location, &L))
        return fail(h, SW_ERR_ARG, "no array for quantity %d location %d", quantity, location);
    if (count != L.count)
        return fail(h, SW_ERR_ARG, "download of %lld values from an array of %lld", (long long)count, (long long)L.count);
    if (L.count == 0) return SW_OK;
    if (!L.per_triangle) {
        HIPCHECK(h, hipMemcpyAsync(ddst, L.ptr, L.count * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
        return SW_OK;
    }
    hipLaunchKernelGGL(k_to_host, dim3(grid_for(L.count)), dim3(kBlock), 0, h->stream, (long long)I.n,
                       (long long)h->N, (long long)I.off, L.slots, L.comp, h->dev_of + I.poff, L.ptr, ddst);
    return launch_check(h, "k_to_host");
}

int sw_upload_device(sw_handle *h, int quantity, int location, const double *dsrc, int64_t count) {
    if (!h || !dsrc) return SW_ERR_ARG;
    if (int rc = touch_api(h)) return rc;
    const int i = selected(h);
    if (i < 0) return i;
    if (int rc = upload_async(h, i, quantity, location, dsrc, count)) return rc;
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    return SW_OK;
}

int sw_download_device(sw_handle *h, int quantity, int location, double *ddst, int64_t count) {
    if (!h || !ddst) return SW_ERR_ARG;
    if (!record_current(quantity, location))
        if (int rc = sync_api(h)) return rc;
    const int i = selected(h);
    if (i < 0) return i;
    if (int rc = download_async(h, i, quantity, location, ddst, count)) return rc;
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    return SW_OK;
}

static int prepare_loop(sw_handle *h, double yieldstep, double finaltime, int32_t has_finaltime,
                        const double *restart_time);

int sw_restart_device(sw_handle *h, const double *stage, const double *xmom, const double *ymom, double time,
                      double yieldstep) {
    if (!h || !stage || !xmom || !ymom) return SW_ERR_ARG;
    if (int rc = touch_api(h)) return rc;
    if (h->n_inst != 1) return fail(h, SW_ERR_STATE, "sw_restart_device restarts a single domain, not an ensemble");
    const int64_t n = h->inst[0].n;
    if (int rc = upload_async(h, 0, SW_STAGE, SW_CENTROID, stage, n)) return rc;
    if (int rc = upload_async(h, 0, SW_XMOM, SW_CENTROID, xmom, n)) return rc;
    if (int rc = upload_async(h, 0, SW_YMOM, SW_CENTROID, ymom, n)) return rc;
    // the start time set and a latched error of an earlier run cleared with
    // the preamble's own state write
    return prepare_loop(h, yieldstep, 0.0, 0, &time);
}

int sw_download_conserved_device(sw_handle *h, double *stage, double *xmom, double *ymom) {
    if (!h || !stage || !xmom || !ymom) return SW_ERR_ARG;
    const int i = selected(h);
    if (i < 0) return i;
    const int64_t n = h->inst[i].n;
    if (int rc = download_async(h, i, SW_STAGE, SW_CENTROID, stage, n)) return rc;
    if (int rc = download_async(h, i, SW_XMOM, SW_CENTROID, xmom, n)) return rc;
    if (int rc = download_async(h, i, SW_YMOM, SW_CENTROID, ymom, n)) return rc;
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    return SW_OK;
}

int sw_download_max_speed(sw_handle *h, double *host, int64_t count) {
    if (!h || !host) return SW_ERR_ARG;
    if (int rc = sync_api(h)) return rc;
    const int i = selected(h);
    if (i < 0) return i;
    const Inst &I = h->inst[i];
    if (count != I.n) return fail(h, SW_ERR_ARG, "max_speed has %lld values", (long long)I.n);
    hipLaunchKernelGGL(k_to_host, dim3(grid_for(I.n)), dim3(kBlock), 0, h->stream, (long long)I.n, (long long)h->N,
                       (long long)I.off, 1, -1, h->dev_of + I.poff, h->f.max_speed, h->staging);
    if (int rc = launch_check(h, "k_to_host")) return rc;
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    HIPCHECK(h, copy_wait(h, host, h->staging, I.n * sizeof(double), hipMemcpyDeviceToHost));
    return SW_OK;
}

int sw_get_state(sw_handle *h, sw_state *s) {
    if (!h || !s) return SW_ERR_ARG;
    const int i = selected(h);
    if (i < 0) return i;
    DevState d;
    int rc = sync_state(h, i, &d);
    static_assert(sizeof(sw_state) <= sizeof(DevState), "layout");
    std::memcpy(s, &d, sizeof(sw_state));
    return rc;
}

int sw_set_state(sw_handle *h, const sw_state *s) {
    if (!h || !s) return SW_ERR_ARG;
    const int i = selected(h);
    if (i < 0) return i;
    DevState d;
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    HIPCHECK(h, copy_wait(h, &d, h->st + i, sizeof d, hipMemcpyDeviceToHost));
    std::memcpy(&d, s, sizeof(sw_state));
    d.halt = (s->event != SW_EVENT_NONE) ? 1 : 0;
    return put_state(h, i, d);
}

int sw_compute_fluxes(sw_handle *h, double *flux_timestep) {
    if (!h) return SW_ERR_ARG;
    if (int rc = touch_api(h)) return rc;
    if (int rc = single(h, "compute_fluxes")) return rc;
    const unsigned g = grid_for(h->N);
    if (h->params.precision == SW_PRECISION_MIXED)
        hipLaunchKernelGGL((k_flux<float, false>), dim3(g), dim3(kBlock), 0, h->stream, h->mesh, h->f, h->dp, h->st,
                           1);
    else
        hipLaunchKernelGGL((k_flux<double, false>), dim3(g), dim3(kBlock), 0, h->stream, h->mesh, h->f, h->dp,
                           h->st, 1);
    hipLaunchKernelGGL((k_timestep<false>), dim3(1), dim3(kReduceBlock), 0, h->stream, h->mesh, h->f.wave_min,
                       h->st, h->dp, kStepEuler);
    if (int rc = launch_check(h, "compute_fluxes")) return rc;
    DevState d;
    int rc = sync_state(h, 0, &d);
    if (flux_timestep) *flux_timestep = d.flux_timestep;
    return rc;
}

int sw_gravity_wb(sw_handle *h) {
    if (!h) return SW_ERR_ARG;
    if (int rc = touch_api(h)) return rc;
    if (int rc = single(h, "gravity_wb")) return rc;
    hipLaunchKernelGGL(k_gravity_wb, dim3(grid_for(h->N)), dim3(kBlock), 0, h->stream, h->mesh, h->f, h->dp);
    return launch_check(h, "gravity_wb");
}

int sw_manning_friction(sw_handle *h) {
    if (!h) return SW_ERR_ARG;
    if (int rc = touch_api(h)) return rc;
    if (int rc = single(h, "manning_friction")) return rc;
    hipLaunchKernelGGL(k_manning, dim3(grid_for(h->N)), dim3(kBlock), 0, h->stream, h->mesh, h->f, h->dp);
    return launch_check(h, "manning_friction");
}

int sw_update_timestep(sw_handle *h, double yieldstep, double finaltime, int32_t has_finaltime) {
    if (!h) return SW_ERR_ARG;
    if (int rc = touch_api(h)) return rc;
    if (int rc = single(h, "update_timestep")) return rc;
    (void)yieldstep;
    DevState d;
    if (int rc = sync_state(h, 0, &d)) return rc;
    d.finaltime = finaltime;
    d.has_finaltime = has_finaltime;
    HIPCHECK(h, copy_wait(h, h->st, &d, sizeof d, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_update_timestep, dim3(1), dim3(1), 0, h->stream, h->st, h->dp);
    if (int rc = launch_check(h, "update_timestep")) return rc;
    return sync_state(h, 0, &d);
}

int sw_update_conserved_quantities(sw_handle *h) {
    if (!h) return SW_ERR_ARG;
    if (int rc = touch_api(h)) return rc;
    if (int rc = single(h, "update_conserved_quantities")) return rc;
    hipLaunchKernelGGL((k_update<false>), dim3(grid_for(h->N)), dim3(kBlock), 0, h->stream, h->mesh, h->f, h->dp, h->st,
                       kUpdateEuler, 0.0, 0.0);
    if (int rc = launch_check(h, "update")) return rc;
    DevState d;
    return sync_state(h, 0, &d);
}

int sw_set_sww_points(sw_handle *h, int64_t number_of_points, const int64_t *number_of_triangles_per_node,
                      const int64_t *vertex_value_indices) {
    if (!h) return SW_ERR_ARG;
    if (int rc = sync_api(h)) return rc;
    const int i = selected(h);
    if (i < 0) return i;
    const Inst &I = h->inst[i];
    const int64_t nv = 3 * I.n;
    if (number_of_points < 0) return fail(h, SW_ERR_ARG, "number_of_points < 0");
    // unique vertices: point 3k+j is vertex j of triangle k
    const bool unique = number_of_points == 0;
    const int64_t P = unique ? nv : number_of_points;
    std::vector<int> start(P + 1), flat(nv);
    if (unique) {
        for (int64_t p = 0; p <= P; p++) start[p] = (int)p;
        for (int64_t v = 0; v < nv; v++) flat[v] = (int)v;
    } else {
        if (!number_of_triangles_per_node || !vertex_value_indices)
            return fail(h, SW_ERR_ARG, "node structure arrays are NULL");
        int64_t acc = 0;
        for (int64_t p = 0; p < P; p++) {
            if (number_of_triangles_per_node[p] < 0) return fail(h, SW_ERR_ARG, "negative triangle count at node %lld", (long long)p);
            start[p] = (int)acc;
            acc += number_of_triangles_per_node[p];
            if (acc > nv) break;
        }
        if (acc != nv)
            return fail(h, SW_ERR_ARG, "triangle counts per node sum to %lld, not 3 * %lld", (long long)acc, (long long)I.n);
        start[P] = (int)acc;
        std::vector<char> seen(nv, 0);
        for (int64_t v = 0; v < nv; v++) {
            const int64_t x = vertex_value_indices[v];
            if (x < 0 || x >= nv || seen[x]) return fail(h, SW_ERR_ARG, "vertex_value_indices is not a permutation of 0 .. 3N-1");
            seen[x] = 1;
            flat[v] = (int)x;
        }
    }
    // processing order: by the device index of each point's first triangle
    std::vector<int> dev_of(I.n);
    HIPCHECK(h, copy_wait(h, dev_of.data(), h->dev_of + I.poff, I.n * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<std::pair<int, int>> keyed(P);
    for (int64_t p = 0; p < P; p++)
        keyed[p] = {start[p] < start[p + 1] ? dev_of[flat[start[p]] / 3] : -1, (int)p};
    std::sort(keyed.begin(), keyed.end());
    std::vector<int> order(P);
    for (int64_t p = 0; p < P; p++) order[p] = keyed[p].second;
    SwwPoints &S = h->sww[i];
    dfree(h, S.start);
    dfree(h, S.slot);
    dfree(h, S.order);
    dfree(h, S.out);
    S = SwwPoints{};
    if (dalloc(h, &S.start, P + 1) || dalloc(h, &S.slot, nv) || dalloc(h, &S.order, P) || dalloc(h, &S.out, 4 * P))
        return SW_ERR_HIP;
    S.P = P;
    HIPCHECK(h, copy_wait(h, S.order, order.data(), P * sizeof(int), hipMemcpyHostToDevice));
    HIPCHECK(h, copy_wait(h, S.start, start.data(), (P + 1) * sizeof(int), hipMemcpyHostToDevice));
    HIPCHECK(h, copy_wait(h, S.slot, flat.data(), nv * sizeof(int), hipMemcpyHostToDevice));
    // caller vertex 3k+j -> device slot j*N + off + dev_of[k]
    hipLaunchKernelGGL(k_sww_slots, dim3(grid_for(nv)), dim3(kBlock), 0, h->stream, (int)nv, (long long)h->N,
                       (long long)I.off, h->dev_of + I.poff, S.slot);
    if (int rc = launch_check(h, "k_sww_slots")) return rc;
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    return SW_OK;
}

int sw_sww_values(sw_handle *h, double minimum_storable_height, float *out, int64_t count) {
    if (!h || !out) return SW_ERR_ARG;
    if (int rc = sync_api(h)) return rc;
    const int i = selected(h);
    if (i < 0) return i;
    const SwwPoints &S = h->sww[i];
    if (S.P == 0) return fail(h, SW_ERR_STATE, "no output points (call sw_set_sww_points)");
    if (count != 4 * S.P) return fail(h, SW_ERR_ARG, "sww values: %lld floats for 4 x %lld points", (long long)count, (long long)S.P);
    const DevFields &F = h->f;
    hipLaunchKernelGGL(k_sww_points, dim3(grid_for(S.P)), dim3(kBlock), 0, h->stream, (long long)S.P, S.order, S.start,
                       S.slot, F.zv, F.wv, F.uhv, F.vhv, (float)minimum_storable_height, S.out);
    if (int rc = launch_check(h, "k_sww_points")) return rc;
    HIPCHECK(h, hipMemcpyAsync(out, S.out, 4 * S.P * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    return SW_OK;
}

int sw_backup_conserved_quantities(sw_handle *h) {
    if (!h) return SW_ERR_ARG;
    if (int rc = touch_api(h)) return rc;
    if (int rc = single(h, "backup_conserved_quantities")) return rc;
    hipLaunchKernelGGL(k_backup, dim3(grid_for(h->N)), dim3(kBlock), 0, h->stream, h->mesh, h->f);
    return launch_check(h, "backup_conserved_quantities");
}

int sw_saxpy_conserved_quantities(sw_handle *h, double a, double b) {
    if (!h) return SW_ERR_ARG;
    if (int rc = touch_api(h)) return rc;
    if (int rc = single(h, "saxpy_conserved_quantities")) return rc;
    hipLaunchKernelGGL(k_saxpy, dim3(grid_for(h->N)), dim3(kBlock), 0, h->stream, h->mesh, h->f, a, b);
    return launch_check(h, "saxpy_conserved_quantities");
}

int sw_protect(sw_handle *h) {
    if (!h) return SW_ERR_ARG;
    if (int rc = touch_api(h)) return rc;
    if (int rc = single(h, "protect")) return rc;
    hipLaunchKernelGGL(k_protect, dim3(grid_for(h->N)), dim3(kBlock), 0, h->stream, h->mesh, h->f, h->dp);
    return launch_check(h, "protect");
}

int sw_protect_swb2(sw_handle *h) {
    if (!h) return SW_ERR_ARG;
    if (int rc = touch_api(h)) return rc;
    if (int rc = single(h, "protect_swb2")) return rc;
    hipLaunchKernelGGL(k_protect_swb2, dim3(grid_for(h->N)), dim3(kBlock), 0, h->stream, h->mesh, h->f, h->dp);
    return launch_check(h, "protect_swb2");
}

int sw_extrapolate_second_order_sw(sw_handle *h) {
    if (!h) return SW_ERR_ARG;
    if (int rc = touch_api(h)) return rc;
    if (int rc = single(h, "extrapolate_second_order_sw")) return rc;
    if (h->params.precision == SW_PRECISION_MIXED)
        hipLaunchKernelGGL((k_extrapolate_second_order<float>), dim3(grid_for(h->N)), dim3(kBlock), 0, h->stream, h->mesh, h->f, h->dp, h->st);
    else
        hipLaunchKernelGGL((k_extrapolate_second_order<double>), dim3(grid_for(h->N)), dim3(kBlock), 0, h->stream, h->mesh, h->f, h->dp, h->st);
    if (int rc = launch_check(h, "extrapolate_second_order_sw")) return rc;
    DevState d;
    return sync_state(h, 0, &d);
}

static int extrapolate_limited(sw_handle *h, int limiter, const char *what) {
    if (!h) return SW_ERR_ARG;
    if (int rc = touch_api(h)) return rc;
    if (int rc = single(h, what)) return rc;
    DevParams dp = h->dp;
    dp.limiter = limiter;
    hipLaunchKernelGGL(k_extrapolate_limited, dim3(grid_for(h->N)), dim3(kBlock), 0, h->stream, h->mesh, h->f, dp);
    return launch_check(h, what);
}

int sw_extrapolate_second_order_and_limit_by_vertex(sw_handle *h) {
    return extrapolate_limited(h, SW_LIMITER_VERTEX, "extrapolate_second_order_and_limit_by_vertex");
}

int sw_extrapolate_second_order_and_limit_by_edge(sw_handle *h) {
    return extrapolate_limited(h, SW_LIMITER_EDGE, "extrapolate_second_order_and_limit_by_edge");
}

int sw_extrapolate_first_order(sw_handle *h) {
    if (!h) return SW_ERR_ARG;
    if (int rc = touch_api(h)) return rc;
    if (int rc = single(h, "extrapolate_first_order")) return rc;
    hipLaunchKernelGGL(k_first_order_conserved, dim3(grid_for(h->N)), dim3(kBlock), 0, h->stream, h->mesh, h->f);
    return launch_check(h, "extrapolate_first_order");
}

int sw_balance_deep_and_shallow(sw_handle *h) {
    if (!h) return SW_ERR_ARG;
    if (int rc = touch_api(h)) return rc;
    if (int rc = single(h, "balance_deep_and_shallow")) return rc;
    hipLaunchKernelGGL(k_balance, dim3(grid_for(h->N)), dim3(kBlock), 0, h->stream, h->mesh, h->f, h->dp);
    return launch_check(h, "balance_deep_and_shallow");
}

int sw_interpolate_from_vertices_to_edges(sw_handle *h) {
    if (!h) return SW_ERR_ARG;
    if (int rc = touch_api(h)) return rc;
    if (int rc = single(h, "interpolate_from_vertices_to_edges")) return rc;
    hipLaunchKernelGGL(k_interpolate, dim3(grid_for(h->N)), dim3(kBlock), 0, h->stream, h->mesh, h->f);
    return launch_check(h, "interpolate_from_vertices_to_edges");
}

int sw_distribute_to_vertices_and_edges(sw_handle *h) {
    if (!h) return SW_ERR_ARG;
    if (int rc = touch_api(h)) return rc;
    if (int rc = single(h, "distribute_to_vertices_and_edges")) return rc;
    DevState d;
    if (int rc = sync_state(h, 0, &d)) return rc;
    if (int rc = sw_protect(h)) return rc;
    if (d.order == 1) {
        if (int rc = sw_extrapolate_first_order(h)) return rc;
    } else if (h->params.limiter == SW_LIMITER_VERTEX) {
        if (int rc = sw_extrapolate_second_order_and_limit_by_vertex(h)) return rc;
    } else if (h->params.limiter == SW_LIMITER_EDGE) {
        if (int rc = sw_extrapolate_second_order_and_limit_by_edge(h)) return rc;
    } else {
        if (int rc = sw_extrapolate_second_order_sw(h)) return rc;
    }
    if (int rc = sw_balance_deep_and_shallow(h)) return rc;
    return sw_interpolate_from_vertices_to_edges(h);
}

int sw_update_boundary(sw_handle *h) {
    if (!h) return SW_ERR_ARG;
    if (int rc = touch_api(h)) return rc;
    if (h->Nb == 0) return SW_OK;
    hipLaunchKernelGGL(k_update_boundary, dim3(grid_for(h->Nb)), dim3(kBlock), 0, h->stream, h->mesh, h->f);
    return launch_check(h, "update_boundary");
}

int sw_update_other_quantities(sw_handle *h) {
    if (!h) return SW_ERR_ARG;
    if (int rc = touch_api(h)) return rc;
    const int64_t n = h->Nb > h->N ? h->Nb : h->N;
    hipLaunchKernelGGL(k_other_quantities, dim3(grid_for(n)), dim3(kBlock), 0, h->stream, h->mesh, h->f);
    return launch_check(h, "update_other_quantities");
}

// distribute_to_vertices_and_edges of every instance (protect, then the
// reconstruction kernel with each instance's order), used by the evolve
// preamble.
// protect, then distribute_to_vertices_and_edges of the edge-record path on
// every instance (and the boundary values if `boundary`); d receives the
// states after it (one read).
int distribute_states(sw_handle *h, std::vector<DevState> &d, bool boundary, bool read_back = true) {
    const unsigned g = grid_for(h->N);
    hipLaunchKernelGGL(k_protect, dim3(g), dim3(kBlock), 0, h->stream, h->mesh, h->f, h->dp);
    const bool limited = h->params.limiter != SW_LIMITER_OPTIMISED;
    if (h->params.precision == SW_PRECISION_MIXED) {
        if (limited)
            hipLaunchKernelGGL((k_distribute<float, false, kBlock, true>), dim3(g), dim3(kBlock), 0, h->stream,
                               h->mesh, h->f, h->dp, h->st);
        else
            hipLaunchKernelGGL((k_distribute<float, false>), dim3(g), dim3(kBlock), 0, h->stream, h->mesh, h->f,
                               h->dp, h->st);
    } else {
        if (limited)
            hipLaunchKernelGGL((k_distribute<double, false, kBlock, true>), dim3(g), dim3(kBlock), 0, h->stream,
                               h->mesh, h->f, h->dp, h->st);
        else
            hipLaunchKernelGGL((k_distribute<double, false>), dim3(g), dim3(kBlock), 0, h->stream, h->mesh, h->f,
                               h->dp, h->st);
    }
    if (boundary && h->Nb > 0)
        hipLaunchKernelGGL(k_update_boundary, dim3(grid_for(h->Nb)), dim3(kBlock), 0, h->stream, h->mesh, h->f);
    if (int rc = launch_check(h, "distribute")) return rc;
    return read_back ? sync_all(h, d) : SW_OK;
}

// The host's states to the device through staging slot `slot` (0 or 1) of
// st_up, without waiting: the two slots let one call make two such writes
// before its next wait, and every call that writes ends with a wait on the
// stream, so a slot is free again when the next call fills it.
static int put_states_async(sw_handle *h, const std::vector<DevState> &d, int slot) {
    DevState *stage = h->st_up + slot * h->n_inst;
    std::memcpy(stage, d.data(), h->n_inst * sizeof(DevState));
    HIPCHECK(h, hipMemcpyAsync(h->st, stage, h->n_inst * sizeof(DevState), hipMemcpyHostToDevice, h->stream));
    h->st_host_ok = false;
    return SW_OK;
}

// evolve()'s preamble (gpu_domain_advanced.py:2473-2545): distribute with the
// default order, reset the time bookkeeping, distribute again and evaluate
// the boundary.  With restart_time the instances also start at that time
// with no event, error or halt latched (sw_restart_device).  One state read
// at the start and one at the end: the first distribute can only latch the
// neighbour error, which the reset clears and the second distribute, on the
// same values, latches again.
static int prepare_loop(sw_handle *h, double yieldstep, double finaltime, int32_t has_finaltime,
                        const double *restart_time) {
    std::vector<DevState> d;
    if (restart_time) {
        // as they stand: a latched error of an earlier run is what a restart
        // clears, so it is not reported
        if (h->st_host_ok) {
            d = h->st_host;
        } else {
            d.resize(h->n_inst);
            HIPCHECK(h, hipMemcpyAsync(h->st_pinned, h->st, h->n_inst * sizeof(DevState), hipMemcpyDeviceToHost,
                                       h->stream));
            HIPCHECK(h, hipStreamSynchronize(h->stream));
            std::memcpy(d.data(), h->st_pinned, h->n_inst * sizeof(DevState));
        }
    } else if (int rc = sync_all(h, d)) {
        return rc;
    }
    for (auto &s : d) {
        s.order = h->params.default_order;
        if (restart_time) {
            s.time = *restart_time;
            s.starttime = *restart_time;
            s.event = SW_EVENT_NONE;
            s.error = 0;
            s.halt = 0;
        }
    }
    if (int rc = put_states_async(h, d, 0)) return rc;
    if (int rc = distribute_states(h, d, false, false)) return rc;
    for (auto &s : d) {
        s.time = s.starttime;
        s.yieldstep = yieldstep;
        s.order = h->params.default_order;
        s.has_finaltime = has_finaltime;
        s.finaltime = finaltime;
        s.yieldtime = s.time + yieldstep;
        s.recorded_min_timestep = h->params.evolve_max_timestep;
        s.recorded_max_timestep = h->params.evolve_min_timestep;
        s.number_of_steps = 0;
        s.number_of_first_order_steps = 0;
        s.smallsteps = 0;
        s.event = 0;
        s.error = 0;
        s.halt = 0;
    }
    if (int rc = put_states_async(h, d, 1)) return rc;
    return distribute_states(h, d, true);
}

int distribute_all(sw_handle *h) {
    std::vector<DevState> d;
    return distribute_states(h, d, false);
}

int sw_evolve_prepare(sw_handle *h, double yieldstep, double finaltime, int32_t has_finaltime) {
    if (!h) return SW_ERR_ARG;
    if (int rc = touch_api(h)) return rc;
    if (h->n_inst > 1 && h->cur >= 0)
        return fail(h, SW_ERR_STATE, "an ensemble is prepared as a whole (select instance -1)");
    return prepare_loop(h, yieldstep, finaltime, has_finaltime, nullptr);
}

// Steps to launch before the next host check: the estimated number of steps
// to the nearest yield or final time of the running instances (at the CFL
// timestep of the last step, +5% and two steps of margin), at most
// kBatchSteps.  Steps queued past a yield only run the halted early exits, so
// this bounds that waste; a low estimate costs another check.
static int64_t batch_estimate(const sw_handle *h, const std::vector<DevState> &d) {
    int64_t est = 0;
    for (int i = 0; i < h->n_inst && i < (int)d.size(); i++) {
        const DevState &s = d[i];
        if (s.event != SW_EVENT_NONE || s.halt || s.error) continue;
        const double dt = std::min(h->params.CFL * s.flux_timestep, h->params.evolve_max_timestep);
        if (!(dt > 0.0)) return kBatchSteps;
        double target = s.yieldtime;
        if (s.has_finaltime && s.finaltime < target) target = s.finaltime;
        const double n = (target - s.time) / dt;
        if (!(n < (double)kBatchSteps)) return kBatchSteps;
        est = std::max<int64_t>(est, (int64_t)(1.05 * n) + 2);
    }
    // whole graphs for the bulk and single launches for the rest (the caller
    // does both), rather than whole graphs whose last steps would only find
    // the yield and halt; a short estimate costs one more batch
    if (est == 0) est = kGraphSteps;
    return std::min<int64_t>(est, kBatchSteps);
}

int sw_evolve_steps(sw_handle *h, int64_t max_steps, int64_t *steps_done, int32_t *event) {
    if (!h || max_steps < 0) return SW_ERR_ARG;
    std::vector<DevState> d0, d;
    if (int rc = current_states(h, d0)) return rc;
    h->st_host_ok = false;
    for (const auto &s : d0)
        if (s.event != SW_EVENT_NONE || s.halt)
            return fail(h, SW_ERR_STATE, "evolve_steps called while a yield is pending (call sw_advance_yield)");
    int64_t launched = 0;
    int rc = SW_OK;
    const bool mixed = h->params.precision == SW_PRECISION_MIXED;
    const bool vstep = vstep_applies(h, d0) && max_steps > 0;
    // the other paths step the operator API's arrays themselves
    if (!vstep)
        if (int rc0 = touch_api(h)) return rc0;
    if (vstep) {
        // continue from the records of the previous call, or build them
        const bool fresh = !h->records_live;
        if (fresh) h->vpar = 0;
        hipLaunchKernelGGL(k_step_reset, dim3(h->n_inst), dim3(kSlots), 0, h->stream, h->sf, h->n_inst, h->vpar,
                           fresh ? 1 : 0);
        if (fresh) {
            hipLaunchKernelGGL(k_step_begin, dim3(grid_for(h->N)), dim3(kBlock), 0, h->stream, h->mesh, h->f, h->sf);
            rc = launch_check(h, "k_step_begin");
            // which loads k_update_v can skip: a launch argument, so a change
            // means a new step graph
            std::vector<unsigned long long> u(h->n_inst);
            if (rc == SW_OK) {
                HIPCHECK(h, hipMemcpyAsync(u.data(), h->sf.uni, h->n_inst * sizeof(u[0]), hipMemcpyDeviceToHost,
                                           h->stream));
                HIPCHECK(h, hipStreamSynchronize(h->stream));
                unsigned all = kUniFriction | kUniMeanBed;
                for (auto x : u) all &= (unsigned)x;
                if (all != h->uni) {
                    h->uni = all;
                    drop_graph(h);
                }
            }
        }
        if (rc == SW_OK) rc = launch_check(h, "k_step_reset");
        // from here on the records hold the state; the operator API's arrays
        // follow when something reads them (sync_api)
        h->records_live = true;
        h->api_pending = true;
    } else if (max_steps > 0) {
        if (mixed) {
            const unsigned g3 = grid_for(3 * h->N);
            hipLaunchKernelGGL(k_check_bed, dim3(g3), dim3(kBlock), 0, h->stream, h->mesh, h->f, h->st);
            hipLaunchKernelGGL(k_pack_edges, dim3(g3), dim3(kBlock), 0, h->stream, h->mesh, h->f, h->dp);
        }
        hipLaunchKernelGGL(k_source_terms, dim3(grid_for(h->N)), dim3(kBlock), 0, h->stream, h->mesh, h->f, h->dp);
        rc = launch_check(h, "k_source_terms");
    }
    // Launch in batches between host checks so a yield early in a long request
    // does not leave many idle steps queued.
    d = d0;
    while (launched < max_steps && rc == SW_OK) {
        const int64_t batch = batch_estimate(h, d);
        int64_t this_batch = max_steps - launched < batch ? max_steps - launched : batch;
        int64_t done_here = 0;
        if (h->timing) {
            for (; done_here < this_batch && rc == SW_OK; done_here++) rc = enqueue_step(h, vstep, (int)done_here);
            if (rc == SW_OK) rc = collect_timing(h, (int)done_here);
        } else {
            // a vertex-record graph is captured and replayed at parity 0
            if (vstep && h->vpar && this_batch > kGraphSteps) rc = enqueue_step(h, vstep, -1), done_here++;
            const bool graphs = !vstep || h->vpar == 0;
            if (rc == SW_OK && graphs) rc = build_graph(h, vstep);
            for (; rc == SW_OK && graphs && done_here + kGraphSteps <= this_batch; done_here += kGraphSteps) {
                hipError_t e = hipGraphLaunch(h->graph_exec, h->stream);
                if (e != hipSuccess) rc = fail(h, SW_ERR_HIP, "hipGraphLaunch: %s", hipGetErrorString(e));
            }
            for (; rc == SW_OK && done_here < this_batch; done_here++) rc = enqueue_step(h, vstep, -1);
        }
        launched += done_here;
        if (vstep) {
            // the next step's state, errors merged, is also the handle's state
            hipLaunchKernelGGL(k_step_settle, dim3((h->n_inst + 63) / 64), dim3(64), 0, h->stream, h->sf, h->n_inst,
                               h->vpar);
            if (rc == SW_OK) rc = launch_check(h, "k_step_settle");
        }
        if (rc) break;
        rc = sync_all(h, d);
        if (rc) break;
        // stop once every instance has reached a yield, its final time or an error
        bool all_stopped = true;
        for (const auto &s : d) all_stopped = all_stopped && (s.event != SW_EVENT_NONE || s.halt);
        if (all_stopped) break;
    }
    if (mixed && launched > 0) {
        // the mixed loop keeps FP32 edge state: rebuild the FP64 edge records
        hipLaunchKernelGGL(k_interpolate, dim3(grid_for(h->N)), dim3(kBlock), 0, h->stream, h->mesh, h->f);
    }
    // The vertex-record step's records stay authoritative (k_step_end and the
    // epilogue are deferred to sync_api); the states read after the last
    // batch are current unless the loop stopped on a launch error.
    int rc2 = SW_OK, rc3 = SW_OK;
    if (!vstep) rc2 = epilogue(h);
    if (!vstep || rc != SW_OK || launched == 0) rc3 = sync_all(h, d);
    // nothing runs on the states after this read: the host's copy is current
    if (rc == SW_OK && rc2 == SW_OK && rc3 == SW_OK) {
        h->st_host = d;
        h->st_host_ok = true;
    }
    int64_t steps = 0;
    int32_t ev = SW_EVENT_NONE;
    for (int i = 0; i < h->n_inst && i < (int)d.size(); i++) {
        steps = std::max<int64_t>(steps, d[i].total_steps - d0[i].total_steps);
        ev = std::max<int32_t>(ev, (int32_t)d[i].event);
    }
    if (steps_done) *steps_done = steps;
    if (event) *event = ev;
    if (rc) return rc;
    if (rc2) return rc2;
    return rc3;
}

int sw_advance_yield(sw_handle *h) {
    if (!h) return SW_ERR_ARG;
    std::vector<DevState> d;
    if (int rc = current_states(h, d)) return rc;
    for (int i = 0; i < h->n_inst; i++) {
        if (h->cur >= 0 && i != h->cur) continue;
        DevState &s = d[i];
        s.yieldtime += s.yieldstep;
        s.recorded_min_timestep = h->params.evolve_max_timestep;
        s.recorded_max_timestep = h->params.evolve_min_timestep;
        s.number_of_steps = 0;
        s.number_of_first_order_steps = 0;
        s.event = 0;
        s.halt = 0;
    }
    h->st_host_ok = false;
    HIPCHECK(h, copy_wait(h, h->st, d.data(), h->n_inst * sizeof(DevState), hipMemcpyHostToDevice));
    h->st_host = d;
    h->st_host_ok = true;
    return SW_OK;
}

int sw_evolve_one_step_unfused(sw_handle *h, int32_t *event) {
    if (!h) return SW_ERR_ARG;
    if (int rc = touch_api(h)) return rc;
    if (int rc = single(h, "evolve_one_step_unfused")) return rc;
    DevState d;
    if (int rc = sync_state(h, 0, &d)) return rc;
    if (d.event != SW_EVENT_NONE) return fail(h, SW_ERR_STATE, "a yield is pending (call sw_advance_yield)");
    // Runge-Kutta stages after the first reuse the first stage's timestep
    // (ANUGA's evolve_one_rk2_step / evolve_one_rk3_step).  rk3's intermediate
    // time t0 + dt/2 is read by no operator here (the boundaries and forcing
    // terms do not depend on time), so the time goes from t0 to t0 + dt.
    const int stages = rk_stages(h->params);
    if (stages > 1)
        if (int rc = sw_backup_conserved_quantities(h)) return rc;
    for (int stage = 1; stage <= stages; stage++) {
        double ft;
        if (int rc = sw_compute_fluxes(h, &ft)) return rc;
        if (int rc = sw_gravity_wb(h)) return rc;
        if (int rc = sw_manning_friction(h)) return rc;
        if (stage == 1) hipLaunchKernelGGL(k_update_timestep, dim3(1), dim3(1), 0, h->stream, h->st, h->dp);
        if (int rc = sw_update_conserved_quantities(h)) return rc;
        if (stage == 1) {
            hipLaunchKernelGGL(k_advance_time, dim3(1), dim3(1), 0, h->stream, h->st);
        } else {
            double a, b;
            rk_weights(stages, stage, &a, &b);
            if (int rc = sw_saxpy_conserved_quantities(h, a, b)) return rc;
        }
        if (int rc = sw_distribute_to_vertices_and_edges(h)) return rc;
        if (int rc = sw_update_boundary(h)) return rc;
    }
    if (int rc = sw_update_other_quantities(h)) return rc;
    hipLaunchKernelGGL(k_finish_step, dim3(1), dim3(1), 0, h->stream, h->st);
    if (int rc = launch_check(h, "step bookkeeping")) return rc;
    int rc = sync_state(h, 0, &d);
    if (d.event != SW_EVENT_NONE) {
        d.halt = 1;
        if (int rc2 = put_state(h, 0, d)) return rc2;
    }
    if (event) *event = (int32_t)d.event;
    return rc;
}

int sw_autotune(sw_handle *h, int32_t steps, int32_t *chosen) {
    if (!h || steps < 1) return SW_ERR_ARG;
    static const int cand[3] = {128, 256, 512};
    const int timed_slot[3] = {0, 2, 3};  // flux, update, distribute in the timing table
    double best_t[3] = {1e300, 1e300, 1e300};
    int best_b[3] = {h->blk[0], h->blk[1], h->blk[2]};
    const bool was_timing = h->timing;
    if (int rc = sw_enable_timing(h, 1)) return rc;
    int rc = SW_OK;
    for (int c = 0; c < 3 && rc == SW_OK; c++) {
        for (int k = 0; k < 3; k++) h->blk[k] = cand[c];
        for (int i = 0; i < kNumTimed; i++) h->times_ms[i] = 0, h->launches[i] = 0;
        int64_t done = 0;
        int32_t ev = SW_EVENT_NONE;
        rc = sw_evolve_steps(h, steps, &done, &ev);
        if (rc) break;
        for (int k = 0; k < 3; k++) {
            const int i = timed_slot[k];
            if (h->launches[i] == 0) continue;
            const double t = h->times_ms[i] / h->launches[i];
            if (t < best_t[k]) best_t[k] = t, best_b[k] = cand[c];
        }
        if (ev != SW_EVENT_NONE || done < steps) break;  // the run reached a yield: keep what was measured
    }
    for (int k = 0; k < 3; k++) h->blk[k] = best_b[k];
    for (int i = 0; i < kNumTimed; i++) h->times_ms[i] = 0, h->launches[i] = 0;
    drop_graph(h);  // launch shapes changed
    h->timing = was_timing;
    if (chosen)
        for (int k = 0; k < 3; k++) chosen[k] = h->blk[k];
    return rc;
}

int sw_set_step_path(sw_handle *h, int32_t path) {
    if (!h) return SW_ERR_ARG;
    if (int rc = touch_api(h)) return rc;
    if (path != SW_STEP_AUTO && path != SW_STEP_EDGE)
        return fail(h, SW_ERR_ARG, "step path %d (SW_STEP_AUTO or SW_STEP_EDGE)", path);
    h->step_path = path;
    drop_graph(h);
    return SW_OK;
}

int sw_get_step_path(sw_handle *h, int32_t *path) {
    if (!h || !path) return SW_ERR_ARG;
    std::vector<DevState> d;
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    d.resize(h->n_inst);
    HIPCHECK(h, copy_wait(h, d.data(), h->st, h->n_inst * sizeof(DevState), hipMemcpyDeviceToHost));
    *path = vstep_applies(h, d) ? SW_STEP_VERTEX : SW_STEP_EDGE;
    return SW_OK;
}

int sw_synchronize(sw_handle *h) {
    if (!h) return SW_ERR_ARG;
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    return SW_OK;
}

int sw_enable_timing(sw_handle *h, int32_t enable) {
    if (!h) return SW_ERR_ARG;
    if (enable && h->ev.empty()) {
        h->ev.resize((size_t)kBatchSteps * (kNumTimed + 1), nullptr);
        for (auto &e : h->ev) HIPCHECK(h, hipEventCreate(&e));
    }
    h->timing = enable != 0;
    return SW_OK;
}

int sw_kernel_times(sw_handle *h, double *ms, int64_t *launches, int32_t nkernels) {
    if (!h || nkernels < 0) return SW_ERR_ARG;
    for (int i = 0; i < nkernels; i++) {
        if (i < kNumTimed) {
            if (ms) ms[i] = h->times_ms[i];
            if (launches) launches[i] = h->launches[i];
            h->times_ms[i] = 0;
            h->launches[i] = 0;
        } else {
            if (ms) ms[i] = 0;
            if (launches) launches[i] = 0;
        }
    }
    return SW_OK;
}

}  // extern "C"
