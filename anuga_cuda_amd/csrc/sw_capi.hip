// C ABI of the shallow-water library (declarations: include/anuga_sw.h).
//
// A handle owns every device buffer of one domain and one HIP stream (default
// priority).  Operator entry points launch the operator kernels of
// sw_kernels.h; sw_evolve_steps runs the fused step (flux -> timestep ->
// update -> distribute) with no host synchronisation between steps, replayed
// from a hipGraph of kGraphSteps steps, and stops at yields through device-side
// halt flags.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/anuga_sw.h"
#include "sw_kernels.h"

using namespace sw;

namespace {

constexpr int kGraphSteps = 16;
constexpr int kBatchSteps = 8 * kGraphSteps;  // steps between host checks
constexpr int kNumTimed = 4;                  // flux, timestep, update, distribute

thread_local std::string g_create_error;

}  // namespace

struct sw_handle {
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t N = 0, Nb = 0;
    sw_params params{};
    DevParams dp{};
    DevMesh mesh{};
    DevFields f{};
    DevState *st = nullptr;
    std::vector<void *> allocations;
    std::string err;
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    bool timing = false;
    std::vector<hipEvent_t> ev;
    double times_ms[kNumTimed] = {0, 0, 0, 0};
    int64_t launches[kNumTimed] = {0, 0, 0, 0};
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
    e = hipMemset(q, 0, count * sizeof(T));
    if (e != hipSuccess) return fail(h, SW_ERR_HIP, "hipMemset failed: %s", hipGetErrorString(e));
    return SW_OK;
}

int upload_i32(sw_handle *h, int **dst, const int64_t *src, size_t count, int64_t lo, int64_t hi,
               const char *name) {
    if (!src) return fail(h, SW_ERR_ARG, "mesh array %s is NULL", name);
    std::vector<int> tmp(count ? count : 1);
    for (size_t i = 0; i < count; i++) {
        if (src[i] < lo || src[i] > hi)
            return fail(h, SW_ERR_ARG, "%s[%zu]=%lld outside [%lld, %lld]", name, i,
                        (long long)src[i], (long long)lo, (long long)hi);
        tmp[i] = (int)src[i];
    }
    int rc = dalloc(h, dst, count);
    if (rc) return rc;
    HIPCHECK(h, hipMemcpy(*dst, tmp.data(), count * sizeof(int), hipMemcpyHostToDevice));
    return SW_OK;
}

int upload_f64(sw_handle *h, double **dst, const double *src, size_t count, const char *name) {
    if (!src) return fail(h, SW_ERR_ARG, "mesh array %s is NULL", name);
    int rc = dalloc(h, dst, count);
    if (rc) return rc;
    HIPCHECK(h, hipMemcpy(*dst, src, count * sizeof(double), hipMemcpyHostToDevice));
    return SW_OK;
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
    return d;
}

int check_params(sw_handle *h, const sw_params *p) {
    if (!p) return fail(h, SW_ERR_ARG, "params is NULL");
    if (p->precision != SW_PRECISION_FP64 && p->precision != SW_PRECISION_MIXED)
        return fail(h, SW_ERR_ARG, "unknown precision %d", p->precision);
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

// Reads the device state and turns latched device errors into a return code.
int sync_state(sw_handle *h, DevState *out) {
    HIPCHECK(h, hipMemcpyAsync(out, h->st, sizeof(DevState), hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
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

double *field_ptr(sw_handle *h, int q, int loc, int64_t *count) {
    DevQ *Q = nullptr;
    switch (q) {
        case SW_STAGE: Q = &h->f.w; break;
        case SW_XMOM: Q = &h->f.uh; break;
        case SW_YMOM: Q = &h->f.vh; break;
        case SW_ELEVATION: Q = &h->f.z; break;
        case SW_HEIGHT: Q = &h->f.h; break;
        case SW_XVELOCITY: Q = &h->f.u; break;
        case SW_YVELOCITY: Q = &h->f.v; break;
        case SW_FRICTION:
            if (loc != SW_CENTROID) return nullptr;
            *count = h->N;
            return h->f.friction;
        default: return nullptr;
    }
    switch (loc) {
        case SW_CENTROID: *count = h->N; return Q->c;
        case SW_EDGE: *count = 3 * h->N; return Q->e;
        case SW_VERTEX: *count = 3 * h->N; return Q->v;
        case SW_BOUNDARY: *count = h->Nb; return Q->b;
        case SW_EXPLICIT_UPDATE: *count = h->N; return Q->eu;
        case SW_SEMI_IMPLICIT_UPDATE: *count = h->N; return Q->si;
        default: return nullptr;
    }
}

__global__ void k_zbar(int N, const double *zv, double *zbar) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k < N) zbar[k] = (zv[3 * k] + zv[3 * k + 1] + zv[3 * k + 2]) / 3.0;
}

// Enqueues one fused step.  slot >= 0 records HIP events around every kernel
// into event group `slot` (timing mode).
int enqueue_fused_step(sw_handle *h, int slot) {
    const unsigned g = grid_for(h->N);
    const bool mixed = h->params.precision == SW_PRECISION_MIXED;
    auto rec = [&](int idx) {
        if (slot >= 0) (void)hipEventRecord(h->ev[slot * (kNumTimed + 1) + idx], h->stream);
    };
    rec(0);
    if (mixed)
        hipLaunchKernelGGL((k_flux<float, true>), dim3(g), dim3(kBlock), 0, h->stream, h->mesh, h->f, h->dp, h->st);
    else
        hipLaunchKernelGGL((k_flux<double, true>), dim3(g), dim3(kBlock), 0, h->stream, h->mesh, h->f, h->dp, h->st);
    rec(1);
    hipLaunchKernelGGL(k_step_timestep, dim3(1), dim3(1), 0, h->stream, h->st, h->dp);
    rec(2);
    hipLaunchKernelGGL((k_update<true>), dim3(g), dim3(kBlock), 0, h->stream, h->mesh, h->f, h->dp, h->st);
    rec(3);
    if (mixed)
        hipLaunchKernelGGL((k_distribute<float, true>), dim3(g), dim3(kBlock), 0, h->stream, h->mesh, h->f, h->dp, h->st);
    else
        hipLaunchKernelGGL((k_distribute<double, true>), dim3(g), dim3(kBlock), 0, h->stream, h->mesh, h->f, h->dp, h->st);
    rec(4);
    return launch_check(h, "fused step");
}

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

int build_graph(sw_handle *h) {
    if (h->graph_exec) return SW_OK;
    HIPCHECK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    int rc = SW_OK;
    for (int s = 0; s < kGraphSteps && rc == SW_OK; s++) rc = enqueue_fused_step(h, -1);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(h->stream, &graph);
    if (rc) return rc;
    if (e != hipSuccess) return fail(h, SW_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    h->graph = graph;
    HIPCHECK(h, hipGraphInstantiate(&h->graph_exec, graph, nullptr, nullptr, 0));
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
    hipLaunchKernelGGL(k_fill, dim3(g), dim3(kBlock), 0, h->stream, h->f.uh.si, (long long)h->N, 0.0);
    hipLaunchKernelGGL(k_fill, dim3(g), dim3(kBlock), 0, h->stream, h->f.vh.si, (long long)h->N, 0.0);
    return launch_check(h, "evolve epilogue");
}

}  // namespace

extern "C" {

const char *sw_last_error(const sw_handle *h) { return h ? h->err.c_str() : g_create_error.c_str(); }

sw_handle *sw_create(const sw_mesh *m, const sw_params *p, int device) {
    g_create_error.clear();
    if (!m) {
        fail(nullptr, SW_ERR_ARG, "mesh is NULL");
        return nullptr;
    }
    if (check_params(nullptr, p)) return nullptr;
    const int64_t N = m->number_of_elements, Nb = m->number_of_boundary_elements;
    if (N <= 0 || N > (int64_t)0x2AAAAAAA || Nb < 0 || Nb > 3 * N) {
        fail(nullptr, SW_ERR_ARG, "bad sizes N=%lld Nb=%lld", (long long)N, (long long)Nb);
        return nullptr;
    }
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) {
        fail(nullptr, SW_ERR_HIP, "hipSetDevice(%d): %s", device, hipGetErrorString(e));
        return nullptr;
    }
    sw_handle *h = new sw_handle();
    h->device = device;
    h->N = N;
    h->Nb = Nb;
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
    DevMesh &M = h->mesh;
    M.N = (int)N;
    M.Nb = (int)Nb;
    int *nbr, *nbr_edge, *sur, *nob, *full, *bcell, *bedge, *bkind;
    double *normals, *len, *radii, *areas, *cc, *vc, *bdir;
    if (upload_i32(h, &nbr, m->neighbours, 3 * N, -Nb, N - 1, "neighbours") ||
        upload_i32(h, &nbr_edge, m->neighbour_edges, 3 * N, -1, 2, "neighbour_edges") ||
        upload_i32(h, &sur, m->surrogate_neighbours, 3 * N, 0, N - 1, "surrogate_neighbours") ||
        upload_i32(h, &nob, m->number_of_boundaries, N, 0, 3, "number_of_boundaries") ||
        upload_i32(h, &full, m->tri_full_flag, N, -1, 1, "tri_full_flag") ||
        upload_f64(h, &normals, m->normals, 6 * N, "normals") ||
        upload_f64(h, &len, m->edgelengths, 3 * N, "edgelengths") ||
        upload_f64(h, &radii, m->radii, N, "radii") || upload_f64(h, &areas, m->areas, N, "areas") ||
        upload_f64(h, &cc, m->centroid_coordinates, 2 * N, "centroid_coordinates") ||
        upload_f64(h, &vc, m->vertex_coordinates, 6 * N, "vertex_coordinates"))
        return bail();
    if (Nb > 0) {
        if (upload_i32(h, &bcell, m->boundary_cells, Nb, 0, N - 1, "boundary_cells") ||
            upload_i32(h, &bedge, m->boundary_edges, Nb, 0, 2, "boundary_edges"))
            return bail();
    } else {
        if (dalloc(h, &bcell, 1) || dalloc(h, &bedge, 1)) return bail();
    }
    // every boundary neighbour index must refer to a listed boundary edge
    for (int64_t i = 0; i < 3 * N; i++) {
        const int64_t n = m->neighbours[i];
        if (n < 0) {
            const int64_t b = -n - 1;
            if (m->boundary_cells[b] != i / 3 || m->boundary_edges[b] != i % 3) {
                fail(h, SW_ERR_ARG, "neighbours[%lld] = %lld does not match boundary entry %lld",
                     (long long)i, (long long)n, (long long)b);
                return bail();
            }
        } else if (m->neighbour_edges[i] < 0) {
            fail(h, SW_ERR_ARG, "neighbour_edges[%lld] is negative for an interior edge", (long long)i);
            return bail();
        }
    }
    if (dalloc(h, &bkind, Nb) || dalloc(h, &bdir, 3 * Nb)) return bail();
    M.nbr = nbr;
    M.nbr_edge = nbr_edge;
    M.sur = sur;
    M.nob = nob;
    M.full = full;
    M.normals = normals;
    M.len = len;
    M.radii = radii;
    M.areas = areas;
    M.cc = cc;
    M.vc = vc;
    M.bcell = bcell;
    M.bedge = bedge;
    M.bkind = bkind;
    M.bdir = bdir;

    DevFields &F = h->f;
    DevQ *conserved[3] = {&F.w, &F.uh, &F.vh};
    for (DevQ *Q : conserved) {
        if (dalloc(h, &Q->c, N) || dalloc(h, &Q->e, 3 * N) || dalloc(h, &Q->v, 3 * N) ||
            dalloc(h, &Q->b, Nb) || dalloc(h, &Q->eu, N) || dalloc(h, &Q->si, N))
            return bail();
    }
    DevQ *others[4] = {&F.z, &F.h, &F.u, &F.v};
    for (DevQ *Q : others) {
        if (dalloc(h, &Q->c, N) || dalloc(h, &Q->e, 3 * N) || dalloc(h, &Q->v, 3 * N) ||
            dalloc(h, &Q->b, Nb))
            return bail();
        Q->eu = Q->si = nullptr;
    }
    if (dalloc(h, &F.friction, N) || dalloc(h, &F.max_speed, N) || dalloc(h, &F.zbar, N)) return bail();
    if (dalloc(h, &h->st, 1)) return bail();
    DevState s0{};
    s0.order = p->default_order;
    s0.flux_min_bits = 0;
    std::memcpy(&s0.flux_min_bits, &p->evolve_max_timestep, sizeof(double));
    if (hipMemcpy(h->st, &s0, sizeof s0, hipMemcpyHostToDevice) != hipSuccess) {
        fail(h, SW_ERR_HIP, "state upload failed");
        return bail();
    }
    return h;
}

int sw_destroy(sw_handle *h) {
    if (!h) return SW_OK;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    drop_graph(h);
    for (auto e : h->ev)
        if (e) (void)hipEventDestroy(e);
    for (void *p : h->allocations) (void)hipFree(p);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return SW_OK;
}

int sw_set_params(sw_handle *h, const sw_params *p) {
    if (!h) return SW_ERR_ARG;
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
    if (h->Nb == 0) return SW_OK;
    if (!kind) return fail(h, SW_ERR_ARG, "boundary kind array is NULL");
    std::vector<int> k(h->Nb);
    for (int64_t i = 0; i < h->Nb; i++) {
        if (kind[i] < 0 || kind[i] > 2) return fail(h, SW_ERR_ARG, "boundary kind[%lld]=%d", (long long)i, kind[i]);
        if (kind[i] == SW_BC_DIRICHLET && !dirichlet)
            return fail(h, SW_ERR_ARG, "Dirichlet boundary without values");
        k[i] = kind[i];
    }
    HIPCHECK(h, hipMemcpy((void *)h->mesh.bkind, k.data(), h->Nb * sizeof(int), hipMemcpyHostToDevice));
    if (dirichlet)
        HIPCHECK(h, hipMemcpy((void *)h->mesh.bdir, dirichlet, 3 * h->Nb * sizeof(double), hipMemcpyHostToDevice));
    return SW_OK;
}

int sw_upload(sw_handle *h, int quantity, int location, const double *host, int64_t count) {
    if (!h || !host) return SW_ERR_ARG;
    int64_t n = 0;
    double *d = field_ptr(h, quantity, location, &n);
    if (!d) return fail(h, SW_ERR_ARG, "no array for quantity %d location %d", quantity, location);
    if (count != n) return fail(h, SW_ERR_ARG, "upload of %lld values into an array of %lld", (long long)count, (long long)n);
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    if (n) HIPCHECK(h, hipMemcpy(d, host, n * sizeof(double), hipMemcpyHostToDevice));
    if (quantity == SW_ELEVATION && location == SW_VERTEX) {
        hipLaunchKernelGGL(k_zbar, dim3(grid_for(h->N)), dim3(kBlock), 0, h->stream, (int)h->N, h->f.z.v, h->f.zbar);
        if (int rc = launch_check(h, "k_zbar")) return rc;
        HIPCHECK(h, hipStreamSynchronize(h->stream));
    }
    return SW_OK;
}

int sw_download(sw_handle *h, int quantity, int location, double *host, int64_t count) {
    if (!h || !host) return SW_ERR_ARG;
    int64_t n = 0;
    double *d = field_ptr(h, quantity, location, &n);
    if (!d) return fail(h, SW_ERR_ARG, "no array for quantity %d location %d", quantity, location);
    if (count != n) return fail(h, SW_ERR_ARG, "download of %lld values from an array of %lld", (long long)count, (long long)n);
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    if (n) HIPCHECK(h, hipMemcpy(host, d, n * sizeof(double), hipMemcpyDeviceToHost));
    return SW_OK;
}

int sw_download_max_speed(sw_handle *h, double *host, int64_t count) {
    if (!h || !host || count != h->N) return SW_ERR_ARG;
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    HIPCHECK(h, hipMemcpy(host, h->f.max_speed, h->N * sizeof(double), hipMemcpyDeviceToHost));
    return SW_OK;
}

int sw_get_state(sw_handle *h, sw_state *s) {
    if (!h || !s) return SW_ERR_ARG;
    DevState d;
    int rc = sync_state(h, &d);
    static_assert(sizeof(sw_state) <= sizeof(DevState), "layout");
    std::memcpy(s, &d, sizeof(sw_state));
    return rc;
}

int sw_set_state(sw_handle *h, const sw_state *s) {
    if (!h || !s) return SW_ERR_ARG;
    DevState d;
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    HIPCHECK(h, hipMemcpy(&d, h->st, sizeof d, hipMemcpyDeviceToHost));
    std::memcpy(&d, s, sizeof(sw_state));
    d.halt = d.halt_next = 0;
    if (s->event != SW_EVENT_NONE) d.halt = d.halt_next = 1;
    HIPCHECK(h, hipMemcpy(h->st, &d, sizeof d, hipMemcpyHostToDevice));
    return SW_OK;
}

int sw_compute_fluxes(sw_handle *h, double *flux_timestep) {
    if (!h) return SW_ERR_ARG;
    const unsigned g = grid_for(h->N);
    hipLaunchKernelGGL(k_reset_flux_min, dim3(1), dim3(1), 0, h->stream, h->st, h->dp.evolve_max_timestep);
    if (h->params.precision == SW_PRECISION_MIXED)
        hipLaunchKernelGGL((k_flux<float, false>), dim3(g), dim3(kBlock), 0, h->stream, h->mesh, h->f, h->dp, h->st);
    else
        hipLaunchKernelGGL((k_flux<double, false>), dim3(g), dim3(kBlock), 0, h->stream, h->mesh, h->f, h->dp, h->st);
    hipLaunchKernelGGL(k_flux_finalize, dim3(1), dim3(1), 0, h->stream, h->st, h->dp.evolve_max_timestep);
    if (int rc = launch_check(h, "compute_fluxes")) return rc;
    DevState d;
    int rc = sync_state(h, &d);
    if (flux_timestep) *flux_timestep = d.flux_timestep;
    return rc;
}

int sw_gravity_wb(sw_handle *h) {
    if (!h) return SW_ERR_ARG;
    hipLaunchKernelGGL(k_gravity_wb, dim3(grid_for(h->N)), dim3(kBlock), 0, h->stream, h->mesh, h->f, h->dp);
    return launch_check(h, "gravity_wb");
}

int sw_manning_friction(sw_handle *h) {
    if (!h) return SW_ERR_ARG;
    hipLaunchKernelGGL(k_manning, dim3(grid_for(h->N)), dim3(kBlock), 0, h->stream, h->mesh, h->f, h->dp);
    return launch_check(h, "manning_friction");
}

int sw_update_timestep(sw_handle *h, double yieldstep, double finaltime, int32_t has_finaltime) {
    if (!h) return SW_ERR_ARG;
    (void)yieldstep;
    DevState d;
    if (int rc = sync_state(h, &d)) return rc;
    d.finaltime = finaltime;
    d.has_finaltime = has_finaltime;
    HIPCHECK(h, hipMemcpy(h->st, &d, sizeof d, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_update_timestep, dim3(1), dim3(1), 0, h->stream, h->st, h->dp);
    if (int rc = launch_check(h, "update_timestep")) return rc;
    return sync_state(h, &d);
}

int sw_update_conserved_quantities(sw_handle *h) {
    if (!h) return SW_ERR_ARG;
    hipLaunchKernelGGL((k_update<false>), dim3(grid_for(h->N)), dim3(kBlock), 0, h->stream, h->mesh, h->f, h->dp, h->st);
    if (int rc = launch_check(h, "update")) return rc;
    DevState d;
    return sync_state(h, &d);
}

int sw_protect(sw_handle *h) {
    if (!h) return SW_ERR_ARG;
    hipLaunchKernelGGL(k_protect, dim3(grid_for(h->N)), dim3(kBlock), 0, h->stream, h->mesh, h->f, h->dp);
    return launch_check(h, "protect");
}

int sw_extrapolate_second_order_sw(sw_handle *h) {
    if (!h) return SW_ERR_ARG;
    if (h->params.precision == SW_PRECISION_MIXED)
        hipLaunchKernelGGL((k_extrapolate_second_order<float>), dim3(grid_for(h->N)), dim3(kBlock), 0, h->stream, h->mesh, h->f, h->dp, h->st);
    else
        hipLaunchKernelGGL((k_extrapolate_second_order<double>), dim3(grid_for(h->N)), dim3(kBlock), 0, h->stream, h->mesh, h->f, h->dp, h->st);
    if (int rc = launch_check(h, "extrapolate_second_order_sw")) return rc;
    DevState d;
    return sync_state(h, &d);
}

int sw_extrapolate_first_order(sw_handle *h) {
    if (!h) return SW_ERR_ARG;
    const unsigned g = grid_for(h->N);
    DevQ *qs[3] = {&h->f.w, &h->f.uh, &h->f.vh};
    for (DevQ *Q : qs)
        hipLaunchKernelGGL(k_first_order, dim3(g), dim3(kBlock), 0, h->stream, (int)h->N, Q->c, Q->e, Q->v);
    return launch_check(h, "extrapolate_first_order");
}

int sw_balance_deep_and_shallow(sw_handle *h) {
    if (!h) return SW_ERR_ARG;
    hipLaunchKernelGGL(k_balance, dim3(grid_for(h->N)), dim3(kBlock), 0, h->stream, h->mesh, h->f, h->dp);
    return launch_check(h, "balance_deep_and_shallow");
}

int sw_interpolate_from_vertices_to_edges(sw_handle *h) {
    if (!h) return SW_ERR_ARG;
    hipLaunchKernelGGL(k_interpolate, dim3(grid_for(h->N)), dim3(kBlock), 0, h->stream, h->mesh, h->f);
    return launch_check(h, "interpolate_from_vertices_to_edges");
}

int sw_distribute_to_vertices_and_edges(sw_handle *h) {
    if (!h) return SW_ERR_ARG;
    DevState d;
    if (int rc = sync_state(h, &d)) return rc;
    if (int rc = sw_protect(h)) return rc;
    if (d.order == 1) {
        if (int rc = sw_extrapolate_first_order(h)) return rc;
    } else {
        if (int rc = sw_extrapolate_second_order_sw(h)) return rc;
    }
    if (int rc = sw_balance_deep_and_shallow(h)) return rc;
    return sw_interpolate_from_vertices_to_edges(h);
}

int sw_update_boundary(sw_handle *h) {
    if (!h) return SW_ERR_ARG;
    if (h->Nb == 0) return SW_OK;
    hipLaunchKernelGGL(k_update_boundary, dim3(grid_for(h->Nb)), dim3(kBlock), 0, h->stream, h->mesh, h->f);
    return launch_check(h, "update_boundary");
}

int sw_update_other_quantities(sw_handle *h) {
    if (!h) return SW_ERR_ARG;
    const int64_t n = h->Nb > h->N ? h->Nb : h->N;
    hipLaunchKernelGGL(k_other_quantities, dim3(grid_for(n)), dim3(kBlock), 0, h->stream, h->mesh, h->f);
    return launch_check(h, "update_other_quantities");
}

int sw_evolve_prepare(sw_handle *h, double yieldstep, double finaltime, int32_t has_finaltime) {
    if (!h) return SW_ERR_ARG;
    DevState d;
    if (int rc = sync_state(h, &d)) return rc;
    d.order = h->params.default_order;
    HIPCHECK(h, hipMemcpy(h->st, &d, sizeof d, hipMemcpyHostToDevice));
    if (int rc = sw_distribute_to_vertices_and_edges(h)) return rc;
    d.time = d.starttime;
    d.yieldstep = yieldstep;
    d.order = h->params.default_order;
    d.has_finaltime = has_finaltime;
    d.finaltime = finaltime;
    d.yieldtime = d.time + yieldstep;
    d.recorded_min_timestep = h->params.evolve_max_timestep;
    d.recorded_max_timestep = h->params.evolve_min_timestep;
    d.number_of_steps = 0;
    d.number_of_first_order_steps = 0;
    d.smallsteps = 0;
    d.event = 0;
    d.error = 0;
    d.halt = d.halt_next = 0;
    std::memcpy(&d.flux_min_bits, &h->params.evolve_max_timestep, sizeof(double));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    HIPCHECK(h, hipMemcpy(h->st, &d, sizeof d, hipMemcpyHostToDevice));
    if (int rc = sw_distribute_to_vertices_and_edges(h)) return rc;
    if (int rc = sw_update_boundary(h)) return rc;
    return sync_state(h, &d);
}

int sw_evolve_steps(sw_handle *h, int64_t max_steps, int64_t *steps_done, int32_t *event) {
    if (!h || max_steps < 0) return SW_ERR_ARG;
    DevState d0, d;
    if (int rc = sync_state(h, &d0)) return rc;
    if (d0.event != SW_EVENT_NONE || d0.halt || d0.halt_next)
        return fail(h, SW_ERR_STATE, "evolve_steps called while a yield is pending (call sw_advance_yield)");
    int64_t launched = 0;
    int rc = SW_OK;
    // Launch in batches between host checks so a yield early in a long request
    // does not leave many idle steps queued.
    const int64_t batch = kBatchSteps;
    while (launched < max_steps && rc == SW_OK) {
        int64_t this_batch = max_steps - launched < batch ? max_steps - launched : batch;
        int64_t done_here = 0;
        if (h->timing) {
            for (; done_here < this_batch && rc == SW_OK; done_here++)
                rc = enqueue_fused_step(h, (int)done_here);
            if (rc == SW_OK) rc = collect_timing(h, (int)done_here);
        } else {
            rc = build_graph(h);
            for (; rc == SW_OK && done_here + kGraphSteps <= this_batch; done_here += kGraphSteps) {
                hipError_t e = hipGraphLaunch(h->graph_exec, h->stream);
                if (e != hipSuccess) rc = fail(h, SW_ERR_HIP, "hipGraphLaunch: %s", hipGetErrorString(e));
            }
            for (; rc == SW_OK && done_here < this_batch; done_here++) rc = enqueue_fused_step(h, -1);
        }
        launched += done_here;
        if (rc) break;
        rc = sync_state(h, &d);
        if (rc || d.event != SW_EVENT_NONE || d.halt_next) break;
    }
    int rc2 = epilogue(h);
    int rc3 = sync_state(h, &d);
    if (steps_done) *steps_done = d.total_steps - d0.total_steps;
    if (event) *event = (int32_t)d.event;
    if (rc) return rc;
    if (rc2) return rc2;
    return rc3;
}

int sw_advance_yield(sw_handle *h) {
    if (!h) return SW_ERR_ARG;
    DevState d;
    if (int rc = sync_state(h, &d)) return rc;
    d.yieldtime += d.yieldstep;
    d.recorded_min_timestep = h->params.evolve_max_timestep;
    d.recorded_max_timestep = h->params.evolve_min_timestep;
    d.number_of_steps = 0;
    d.number_of_first_order_steps = 0;
    d.event = 0;
    d.halt = d.halt_next = 0;
    HIPCHECK(h, hipMemcpy(h->st, &d, sizeof d, hipMemcpyHostToDevice));
    return SW_OK;
}

int sw_evolve_one_step_unfused(sw_handle *h, int32_t *event) {
    if (!h) return SW_ERR_ARG;
    DevState d;
    if (int rc = sync_state(h, &d)) return rc;
    if (d.event != SW_EVENT_NONE) return fail(h, SW_ERR_STATE, "a yield is pending (call sw_advance_yield)");
    double ft;
    if (int rc = sw_compute_fluxes(h, &ft)) return rc;
    if (int rc = sw_gravity_wb(h)) return rc;
    if (int rc = sw_manning_friction(h)) return rc;
    hipLaunchKernelGGL(k_update_timestep, dim3(1), dim3(1), 0, h->stream, h->st, h->dp);
    if (int rc = sw_update_conserved_quantities(h)) return rc;
    hipLaunchKernelGGL(k_advance_time, dim3(1), dim3(1), 0, h->stream, h->st);
    if (int rc = sw_distribute_to_vertices_and_edges(h)) return rc;
    if (int rc = sw_update_boundary(h)) return rc;
    if (int rc = sw_update_other_quantities(h)) return rc;
    hipLaunchKernelGGL(k_finish_step, dim3(1), dim3(1), 0, h->stream, h->st);
    if (int rc = launch_check(h, "step bookkeeping")) return rc;
    int rc = sync_state(h, &d);
    if (d.event != SW_EVENT_NONE) {
        d.halt = d.halt_next = 1;
        HIPCHECK(h, hipMemcpy(h->st, &d, sizeof d, hipMemcpyHostToDevice));
    }
    if (event) *event = (int32_t)d.event;
    return rc;
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
    for (int i = 0; i < nkernels && i < kNumTimed; i++) {
        if (ms) ms[i] = h->times_ms[i];
        if (launches) launches[i] = h->launches[i];
        h->times_ms[i] = 0;
        h->launches[i] = 0;
    }
    return SW_OK;
}

}  // extern "C"
