// mesh_build.hip -- mesh ingestion on the device: the derived geometry and
// the neighbour structure of a triangle mesh (gfx950).
//
// The reference builds these arrays on the host with ANUGA's General_mesh /
// Mesh (the anuga package, not part of the reference) and uploads them
// (gpu_domain_advanced.py:582-733; layout sw_domain.h:60-180).  Here one
// launch per triangle computes the oriented triangle and all its geometry,
// a radix sort of the edge keys pairs every interior edge with its partner,
// and one launch per sorted edge writes the neighbour arrays.  Every FP64
// value is formed with the same operations, in the same order, as the host
// build (anuga_cuda_amd/mesh.py), so both give identical bits (the library
// is compiled without FMA contraction; +, -, *, / and sqrt are correctly
// rounded).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/anuga_sw.h"

namespace {

constexpr int kMeshBlock = 256;

enum : int { kBadIndex = 1, kDegenerate = 2, kZeroEdge = 4, kNonManifold = 8 };

// Per triangle: orientation (v1 <-> v2 where the input is clockwise), vertex
// coordinates, centroid, area, outward normals, edge lengths and midpoints,
// radius (smallest centroid-to-edge-midpoint distance), and the three edge
// keys lo * P + hi with their flat edge index 3k + i (edge i opposite vertex i).
__global__ __launch_bounds__(kMeshBlock) void k_mesh_triangles(long long N, long long P, const double *pts,
                                                               const long long *tri_in, long long *tri,
                                                               double *vc, double *cc, double *area,
                                                               double *normals, double *len, double *mid,
                                                               double *radii, unsigned long long *key,
                                                               int *flat, int *err) {
    const long long k = (long long)blockIdx.x * kMeshBlock + threadIdx.x;
    if (k >= N) return;
    long long v0 = tri_in[3 * k], v1 = tri_in[3 * k + 1], v2 = tri_in[3 * k + 2];
    if (v0 < 0 || v1 < 0 || v2 < 0 || v0 >= P || v1 >= P || v2 >= P) {
        // valid, unpaired edge records (the host stops before pairing)
#pragma unroll
        for (int i = 0; i < 3; i++) {
            key[3 * k + i] = ~0ull;
            flat[3 * k + i] = (int)(3 * k + i);
        }
        atomicOr(err, kBadIndex);
        return;
    }
    double x0 = pts[2 * v0], y0 = pts[2 * v0 + 1];
    double x1 = pts[2 * v1], y1 = pts[2 * v1 + 1];
    double x2 = pts[2 * v2], y2 = pts[2 * v2 + 1];
    const double cross = (x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0);
    if (cross == 0.0) atomicOr(err, kDegenerate);
    if (cross < 0.0) {
        long long tv = v1;
        v1 = v2;
        v2 = tv;
        double t = x1;
        x1 = x2;
        x2 = t;
        t = y1;
        y1 = y2;
        y2 = t;
    }
    tri[3 * k] = v0;
    tri[3 * k + 1] = v1;
    tri[3 * k + 2] = v2;
    vc[6 * k] = x0;
    vc[6 * k + 1] = y0;
    vc[6 * k + 2] = x1;
    vc[6 * k + 3] = y1;
    vc[6 * k + 4] = x2;
    vc[6 * k + 5] = y2;
    const double cx = ((x0 + x1) + x2) / 3.0;
    const double cy = ((y0 + y1) + y2) / 3.0;
    cc[2 * k] = cx;
    cc[2 * k + 1] = cy;
    area[k] = fabs(((x1 * y0 - x0 * y1) + (x2 * y1 - x1 * y2)) + (x0 * y2 - x2 * y0)) / 2.0;
    const double ex[3] = {x2 - x1, x0 - x2, x1 - x0};
    const double ey[3] = {y2 - y1, y0 - y2, y1 - y0};
    const double mx[3] = {(x1 + x2) / 2, (x0 + x2) / 2, (x1 + x0) / 2};
    const double my[3] = {(y1 + y2) / 2, (y0 + y2) / 2, (y1 + y0) / 2};
    const long long va[3] = {v1, v2, v0}, vb[3] = {v2, v0, v1};
    double d[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const double l = sqrt(ex[i] * ex[i] + ey[i] * ey[i]);
        if (l == 0.0) atomicOr(err, kZeroEdge);
        len[3 * k + i] = l;
        normals[6 * k + 2 * i] = ey[i] / l;
        normals[6 * k + 2 * i + 1] = -(ex[i] / l);
        mid[6 * k + 2 * i] = mx[i];
        mid[6 * k + 2 * i + 1] = my[i];
        const double dx = cx - mx[i], dy = cy - my[i];
        d[i] = sqrt(dx * dx + dy * dy);
        const long long lo = va[i] < vb[i] ? va[i] : vb[i], hi = va[i] < vb[i] ? vb[i] : va[i];
        key[3 * k + i] = (unsigned long long)(lo * P + hi);
        flat[3 * k + i] = (int)(3 * k + i);
    }
    radii[k] = fmin(fmin(d[0], d[1]), d[2]);
}

__global__ __launch_bounds__(kMeshBlock) void k_mesh_fill(long long n, long long *a, long long v) {
    const long long i = (long long)blockIdx.x * kMeshBlock + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void k_mesh_clear_flag(int *err) { *err = 0; }

// Sorted edge keys: the first edge of each run of equal keys pairs itself
// with the second; a run longer than two is an edge shared by more than two
// triangles.
__global__ __launch_bounds__(kMeshBlock) void k_mesh_pairs(long long E, const unsigned long long *ks,
                                                           const int *fs, long long *nb, long long *ne, int *err) {
    const long long i = (long long)blockIdx.x * kMeshBlock + threadIdx.x;
    if (i + 1 >= E) return;
    const unsigned long long k = ks[i];
    if (ks[i + 1] != k || (i > 0 && ks[i - 1] == k)) return;
    if (i + 2 < E && ks[i + 2] == k) atomicOr(err, kNonManifold);
    const int a = fs[i], b = fs[i + 1];
    nb[a] = b / 3;
    ne[a] = b % 3;
    nb[b] = a / 3;
    ne[b] = a % 3;
}

unsigned grid_of(long long n) { return (unsigned)((n + kMeshBlock - 1) / kMeshBlock); }

struct DevBufs {
    std::vector<void *> p;
    ~DevBufs() {
        for (void *q : p) (void)hipFree(q);
    }
    template <typename T>
    bool alloc(T **out, long long count) {
        void *q = nullptr;
        if (hipMalloc(&q, (size_t)(count > 0 ? count : 1) * sizeof(T)) != hipSuccess) return false;
        p.push_back(q);
        *out = static_cast<T *>(q);
        return true;
    }
};

thread_local std::string g_mesh_error;

int mesh_fail(const char *msg) {
    g_mesh_error = msg;
    return SW_ERR_ARG;
}

// Scratch of sw_build_mesh_noalloc, one copy per GPU for the life of the
// process: the edge keys and flat edge indices before and after the radix
// sort, the error flags and the sort's temporaries, for meshes of up to
// SW_NOALLOC_MAX_TRIANGLES triangles.  Host code reaches it only through the
// address hipGetSymbolAddress returns on each device (the array's name is a
// host-side stand-in).
constexpr long long kNaEdges = 3ll * SW_NOALLOC_MAX_TRIANGLES;
constexpr size_t kNaKeyBytes = (size_t)kNaEdges * sizeof(unsigned long long);
constexpr size_t kNaFlatBytes = (size_t)kNaEdges * sizeof(int);
constexpr size_t kNaErrBytes = 256;
constexpr size_t kNaSortBytes = (size_t)8 << 20;
constexpr size_t kNaBytes = 2 * kNaKeyBytes + 2 * kNaFlatBytes + kNaErrBytes + kNaSortBytes;
static_assert(kNaKeyBytes % 256 == 0 && kNaFlatBytes % 256 == 0, "scratch sections stay 256-byte aligned");
__device__ __align__(256) unsigned char scratch[kNaBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
unsigned char *g_scratch_ptr[kMaxDevices] = {};

// The device address of this GPU's scratch (the device is current), looked up
// once per device.
int scratch_address(int device, unsigned char **out) {
    if (device < 0 || device >= kMaxDevices) return mesh_fail("sw_build_mesh_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            char buf[256];
            snprintf(buf, sizeof buf, "sw_build_mesh_noalloc: hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
            g_mesh_error = buf;
            return SW_ERR_HIP;
        }
        g_scratch_ptr[device] = static_cast<unsigned char *>(ptr);
    }
    *out = g_scratch_ptr[device];
    return SW_OK;
}

}  // namespace

extern "C" {

const char *sw_mesh_last_error(void) { return g_mesh_error.c_str(); }

int sw_build_mesh(int device, int64_t number_of_points, const double *points, int64_t number_of_elements,
                  const int64_t *triangles, sw_mesh_geometry *out) {
    const long long P = number_of_points, N = number_of_elements, E = 3 * N;
    if (!points || !triangles || !out || P < 3 || N < 1) return mesh_fail("sw_build_mesh: empty mesh or NULL array");
    if (!out->triangles || !out->vertex_coordinates || !out->centroid_coordinates || !out->areas || !out->normals ||
        !out->edgelengths || !out->edge_midpoint_coordinates || !out->radii || !out->neighbours ||
        !out->neighbour_edges)
        return mesh_fail("sw_build_mesh: NULL output array");
    if (E >= (1ll << 31) || P >= (1ll << 31)) return mesh_fail("sw_build_mesh: mesh too large for 32-bit edge indices");
    char buf[256];
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) {
        snprintf(buf, sizeof buf, "hipSetDevice(%d): %s", device, hipGetErrorString(e));
        g_mesh_error = buf;
        return SW_ERR_HIP;
    }
    DevBufs B;
    double *dp, *vc, *cc, *area, *nrm, *len, *mid, *rad;
    long long *tin, *tri, *nb, *ne;
    unsigned long long *key, *key_s;
    int *flat, *flat_s, *err;
    if (!B.alloc(&dp, 2 * P) || !B.alloc(&tin, 3 * N) || !B.alloc(&tri, 3 * N) || !B.alloc(&vc, 6 * N) ||
        !B.alloc(&cc, 2 * N) || !B.alloc(&area, N) || !B.alloc(&nrm, 6 * N) || !B.alloc(&len, 3 * N) ||
        !B.alloc(&mid, 6 * N) || !B.alloc(&rad, N) || !B.alloc(&nb, E) || !B.alloc(&ne, E) || !B.alloc(&key, E) ||
        !B.alloc(&key_s, E) || !B.alloc(&flat, E) || !B.alloc(&flat_s, E) || !B.alloc(&err, 1)) {
        g_mesh_error = "sw_build_mesh: device allocation failed";
        return SW_ERR_HIP;
    }
    hipStream_t s = nullptr;
#define MCHECK(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) {                                                               \
            snprintf(buf, sizeof buf, "sw_build_mesh: %s: %s", #call, hipGetErrorString(e_)); \
            g_mesh_error = buf;                                                               \
            return SW_ERR_HIP;                                                                \
        }                                                                                     \
    } while (0)
    MCHECK(hipMemcpy(dp, points, 2 * P * sizeof(double), hipMemcpyHostToDevice));
    MCHECK(hipMemcpy(tin, triangles, 3 * N * sizeof(long long), hipMemcpyHostToDevice));
    MCHECK(hipMemset(err, 0, sizeof(int)));
    hipLaunchKernelGGL(k_mesh_triangles, dim3(grid_of(N)), dim3(kMeshBlock), 0, s, N, P, dp, tin, tri, vc, cc, area,
                       nrm, len, mid, rad, key, flat, err);
    hipLaunchKernelGGL(k_mesh_fill, dim3(grid_of(E)), dim3(kMeshBlock), 0, s, E, nb, -1ll);
    hipLaunchKernelGGL(k_mesh_fill, dim3(grid_of(E)), dim3(kMeshBlock), 0, s, E, ne, -1ll);
    MCHECK(hipGetLastError());
    int herr = 0;
    MCHECK(hipMemcpy(&herr, err, sizeof(int), hipMemcpyDeviceToHost));
    if (herr & kBadIndex) return mesh_fail("triangle vertex index out of range");
    // keys are below P^2: sort only the bits that can be set
    const unsigned long long kmax = (unsigned long long)P * (unsigned long long)P - 1;
    const int end_bit = 64 - __builtin_clzll(kmax);
    size_t temp = 0;
    MCHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, temp, key, key_s, flat, flat_s, (int)E, 0, end_bit, s));
    void *tmp = nullptr;
    if (!B.alloc((char **)&tmp, (long long)temp)) {
        g_mesh_error = "sw_build_mesh: sort scratch allocation failed";
        return SW_ERR_HIP;
    }
    MCHECK(hipcub::DeviceRadixSort::SortPairs(tmp, temp, key, key_s, flat, flat_s, (int)E, 0, end_bit, s));
    hipLaunchKernelGGL(k_mesh_pairs, dim3(grid_of(E)), dim3(kMeshBlock), 0, s, E, key_s, flat_s, nb, ne, err);
    MCHECK(hipGetLastError());
    MCHECK(hipMemcpy(&herr, err, sizeof(int), hipMemcpyDeviceToHost));
    if (herr & kDegenerate) return mesh_fail("degenerate triangle (zero area)");
    if (herr & kZeroEdge) return mesh_fail("zero-length edge");
    if (herr & kNonManifold) return mesh_fail("non-manifold mesh: an edge is shared by more than two triangles");
    MCHECK(hipMemcpy(out->triangles, tri, 3 * N * sizeof(long long), hipMemcpyDeviceToHost));
    MCHECK(hipMemcpy(out->vertex_coordinates, vc, 6 * N * sizeof(double), hipMemcpyDeviceToHost));
    MCHECK(hipMemcpy(out->centroid_coordinates, cc, 2 * N * sizeof(double), hipMemcpyDeviceToHost));
    MCHECK(hipMemcpy(out->areas, area, N * sizeof(double), hipMemcpyDeviceToHost));
    MCHECK(hipMemcpy(out->normals, nrm, 6 * N * sizeof(double), hipMemcpyDeviceToHost));
    MCHECK(hipMemcpy(out->edgelengths, len, 3 * N * sizeof(double), hipMemcpyDeviceToHost));
    MCHECK(hipMemcpy(out->edge_midpoint_coordinates, mid, 6 * N * sizeof(double), hipMemcpyDeviceToHost));
    MCHECK(hipMemcpy(out->radii, rad, N * sizeof(double), hipMemcpyDeviceToHost));
    MCHECK(hipMemcpy(out->neighbours, nb, E * sizeof(long long), hipMemcpyDeviceToHost));
    MCHECK(hipMemcpy(out->neighbour_edges, ne, E * sizeof(long long), hipMemcpyDeviceToHost));
#undef MCHECK
    return SW_OK;
}

int sw_build_mesh_noalloc(int device, int64_t number_of_points, const double *points, int64_t number_of_elements,
                          const int64_t *triangles, const sw_mesh_geometry *out) {
    const long long P = number_of_points, N = number_of_elements, E = 3 * N;
    if (!points || !triangles || !out || P < 3 || N < 1)
        return mesh_fail("sw_build_mesh_noalloc: empty mesh or NULL array");
    if (!out->triangles || !out->vertex_coordinates || !out->centroid_coordinates || !out->areas || !out->normals ||
        !out->edgelengths || !out->edge_midpoint_coordinates || !out->radii || !out->neighbours ||
        !out->neighbour_edges)
        return mesh_fail("sw_build_mesh_noalloc: NULL output array");
    if (N > SW_NOALLOC_MAX_TRIANGLES) {
        char buf[160];
        snprintf(buf, sizeof buf, "sw_build_mesh_noalloc: %lld triangles exceed the scratch limit of %d", N,
                 SW_NOALLOC_MAX_TRIANGLES);
        g_mesh_error = buf;
        return SW_ERR_ARG;
    }
    if (P >= (1ll << 31)) return mesh_fail("sw_build_mesh_noalloc: too many points for 32-bit edge keys");
    char buf[256];
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) {
        snprintf(buf, sizeof buf, "hipSetDevice(%d): %s", device, hipGetErrorString(e));
        g_mesh_error = buf;
        return SW_ERR_HIP;
    }
    unsigned char *base = nullptr;
    if (int rc = scratch_address(device, &base)) return rc;
    unsigned long long *key = reinterpret_cast<unsigned long long *>(base);
    unsigned long long *key_s = reinterpret_cast<unsigned long long *>(base + kNaKeyBytes);
    int *flat = reinterpret_cast<int *>(base + 2 * kNaKeyBytes);
    int *flat_s = reinterpret_cast<int *>(base + 2 * kNaKeyBytes + kNaFlatBytes);
    int *err = reinterpret_cast<int *>(base + 2 * kNaKeyBytes + 2 * kNaFlatBytes);
    void *sort_tmp = base + 2 * kNaKeyBytes + 2 * kNaFlatBytes + kNaErrBytes;
    hipStream_t s = nullptr;
#define MCHECK(call)                                                                                  \
    do {                                                                                              \
        hipError_t e_ = (call);                                                                       \
        if (e_ != hipSuccess) {                                                                       \
            snprintf(buf, sizeof buf, "sw_build_mesh_noalloc: %s: %s", #call, hipGetErrorString(e_)); \
            g_mesh_error = buf;                                                                       \
            return SW_ERR_HIP;                                                                        \
        }                                                                                             \
    } while (0)
    // keys are below P^2: sort only the bits that can be set
    const unsigned long long kmax = (unsigned long long)P * (unsigned long long)P - 1;
    const int end_bit = 64 - __builtin_clzll(kmax);
    // the sort passes alternate between the two key / index arrays of the
    // scratch, so its temporaries are only its histograms
    hipcub::DoubleBuffer<unsigned long long> keys(key, key_s);
    hipcub::DoubleBuffer<int> flats(flat, flat_s);
    size_t temp = 0;
    MCHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, temp, keys, flats, (int)E, 0, end_bit, s));
    if (temp > kNaSortBytes) return mesh_fail("sw_build_mesh_noalloc: the radix sort needs more scratch than reserved");
    // the flags start at zero on every call (the array keeps the last call's contents)
    hipLaunchKernelGGL(k_mesh_clear_flag, dim3(1), dim3(1), 0, s, err);
    hipLaunchKernelGGL(k_mesh_triangles, dim3(grid_of(N)), dim3(kMeshBlock), 0, s, N, P, points,
                       reinterpret_cast<const long long *>(triangles), reinterpret_cast<long long *>(out->triangles),
                       out->vertex_coordinates, out->centroid_coordinates, out->areas, out->normals, out->edgelengths,
                       out->edge_midpoint_coordinates, out->radii, key, flat, err);
    hipLaunchKernelGGL(k_mesh_fill, dim3(grid_of(E)), dim3(kMeshBlock), 0, s, E,
                       reinterpret_cast<long long *>(out->neighbours), -1ll);
    hipLaunchKernelGGL(k_mesh_fill, dim3(grid_of(E)), dim3(kMeshBlock), 0, s, E,
                       reinterpret_cast<long long *>(out->neighbour_edges), -1ll);
    MCHECK(hipGetLastError());
    int herr = 0;
    MCHECK(hipMemcpy(&herr, err, sizeof(int), hipMemcpyDeviceToHost));
    if (herr & kBadIndex) return mesh_fail("triangle vertex index out of range");
    MCHECK(hipcub::DeviceRadixSort::SortPairs(sort_tmp, temp, keys, flats, (int)E, 0, end_bit, s));
    hipLaunchKernelGGL(k_mesh_pairs, dim3(grid_of(E)), dim3(kMeshBlock), 0, s, E, keys.Current(), flats.Current(),
                       reinterpret_cast<long long *>(out->neighbours),
                       reinterpret_cast<long long *>(out->neighbour_edges), err);
    MCHECK(hipGetLastError());
    MCHECK(hipMemcpy(&herr, err, sizeof(int), hipMemcpyDeviceToHost));
    if (herr & kDegenerate) return mesh_fail("degenerate triangle (zero area)");
    if (herr & kZeroEdge) return mesh_fail("zero-length edge");
    if (herr & kNonManifold) return mesh_fail("non-manifold mesh: an edge is shared by more than two triangles");
#undef MCHECK
    return SW_OK;
}

}  // extern "C"
