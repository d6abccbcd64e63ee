/* anuga_sw.h -- C ABI of the MI355X (gfx950) shallow-water evolve() library.
 *
 * Drop-in boundary for ANUGA's shallow-water step operators.  Every entry point
 * below replaces one operator of the reference GPU domain driver
 * (src/anuga_cuda/gpu_domain_advanced.py) or the reference's C evolve loop
 * (src/anuga_HMPP/evolve.c); the citation on each declaration names the
 * interface it replaces.  Plain pointers and sizes only: the library owns its
 * device buffers; host pointers are borrowed for the duration of a call.
 *
 * Conventions
 *   - All arithmetic is IEEE FP64 unless the handle was created with
 *     SW_PRECISION_MIXED (FP32 flux and reconstruction, FP64 accumulation,
 *     timestep and wet/dry tests).
 *   - Arrays use the reference layout (src/anuga_HMPP/sw_domain.h:60-180):
 *     per-triangle [N], per-edge / per-vertex [N][3], normals and vertex
 *     coordinates [N][6], centroids [N][2], boundary [Nb].
 *   - Index arrays are int64 on the host (as the reference's `long*`); they
 *     are stored as int32 on the device.
 *   - Every function returns 0 on success or a negative SW_ERR_* code;
 *     sw_last_error() gives the message.  A handle is not re-entrant.
 */
#ifndef ANUGA_SW_H
#define ANUGA_SW_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sw_handle sw_handle;

/* Error codes. */
#define SW_OK 0
#define SW_ERR_UPDATE (-1)       /* semi-implicit denominator <= 0 (quantity.c:803-805) */
#define SW_ERR_SMALLSTEPS (-2)   /* timestep below evolve_min_timestep at first order */
#define SW_ERR_NEIGHBOUR (-3)    /* two-boundary triangle without interior neighbour */
#define SW_ERR_ARG (-10)         /* invalid argument */
#define SW_ERR_HIP (-11)         /* HIP runtime error */
#define SW_ERR_STATE (-12)       /* call out of sequence */

/* Quantity ids, locations and boundary kinds. */
enum sw_quantity {
    SW_STAGE = 0, SW_XMOM = 1, SW_YMOM = 2, SW_ELEVATION = 3, SW_FRICTION = 4,
    SW_HEIGHT = 5, SW_XVELOCITY = 6, SW_YVELOCITY = 7
};
enum sw_location {
    SW_CENTROID = 0, SW_EDGE = 1, SW_VERTEX = 2, SW_BOUNDARY = 3,
    SW_EXPLICIT_UPDATE = 4, SW_SEMI_IMPLICIT_UPDATE = 5
};
enum sw_boundary_kind { SW_BC_REFLECTIVE = 0, SW_BC_DIRICHLET = 1, SW_BC_TRANSMISSIVE = 2 };
enum sw_precision { SW_PRECISION_FP64 = 0, SW_PRECISION_MIXED = 1 };
/* Device numbering of the triangles (read by sw_create only).  The arithmetic
 * never depends on it: which side evaluates a shared edge and the order in
 * which edge fluxes are summed follow the caller's numbering, as in the C. */
enum sw_locality_order { SW_ORDER_KEEP = 0, SW_ORDER_HILBERT = 1 };
enum sw_limiter { SW_LIMITER_OPTIMISED = 0, SW_LIMITER_VERTEX = 1, SW_LIMITER_EDGE = 2 };
enum sw_event { SW_EVENT_NONE = 0, SW_EVENT_YIELD = 1, SW_EVENT_FINAL = 2 };
/* Kernels of the fused Euler step (sw_set_step_path): SW_STEP_AUTO picks the
 * vertex-record step wherever it applies, SW_STEP_EDGE keeps the edge-record one;
 * sw_get_step_path reports SW_STEP_VERTEX or SW_STEP_EDGE. */
enum sw_step_path { SW_STEP_AUTO = 0, SW_STEP_EDGE = 1, SW_STEP_VERTEX = 2 };

/* Domain parameters; names follow struct domain (src/anuga_HMPP/sw_domain.h:16-60)
 * and the attributes the reference driver passes to its kernels. */
typedef struct sw_params {
    double g, epsilon, H0, minimum_allowed_height, maximum_allowed_speed;
    double beta_w, beta_w_dry, beta_uh, beta_uh_dry, beta_vh, beta_vh_dry;
    double alpha_balance, CFL, evolve_max_timestep, evolve_min_timestep;
    int32_t optimise_dry_cells, extrapolate_velocity_second_order;
    int32_t tight_slope_limiters, use_centroid_velocities, use_sloped_mannings;
    int32_t max_smallsteps, default_order, precision;
    int32_t locality_order;  /* SW_ORDER_*; arrays at the ABI stay in the caller's order */
    /* timestepping_method (sw_domain.h:49; gpu_domain_advanced.py:2566-2573):
     * 1 (or 0) Euler, 2 rk2, 3 rk3 */
    int32_t timestepping_method;
    /* distribute path (gpu_domain_advanced.py:1814-1990): SW_LIMITER_OPTIMISED
     * (default: extrapolate_second_order_sw, optimised_gradient_limiter),
     * SW_LIMITER_VERTEX (per-quantity extrapolate_second_order_and_limit_by_vertex),
     * SW_LIMITER_EDGE (use_edge_limiter: ..._limit_by_edge) */
    int32_t limiter;
    /* Quantity.beta of stage, xmomentum, ymomentum for the per-quantity
     * limiters (ANUGA's default 1.0; 0 limits to first order); unused by
     * SW_LIMITER_OPTIMISED */
    double quantity_beta[3];
} sw_params;

/* Mesh arrays (host, borrowed during sw_create). */
typedef struct sw_mesh {
    int64_t number_of_elements;          /* N */
    int64_t number_of_boundary_elements; /* Nb */
    const int64_t *neighbours;           /* [N][3], boundary edges hold -(m+1) */
    const int64_t *neighbour_edges;      /* [N][3] */
    const int64_t *surrogate_neighbours; /* [N][3] */
    const int64_t *number_of_boundaries; /* [N] */
    const int64_t *tri_full_flag;        /* [N] */
    const double *normals;               /* [N][6] */
    const double *edgelengths;           /* [N][3] */
    const double *radii;                 /* [N] */
    const double *areas;                 /* [N] */
    const double *centroid_coordinates;  /* [N][2] */
    const double *vertex_coordinates;    /* [N][6] */
    const int64_t *boundary_cells;       /* [Nb] vol_id of boundary edge m */
    const int64_t *boundary_edges;       /* [Nb] edge_id of boundary edge m */
} sw_mesh;

/* Scalar evolve state (time bookkeeping of ANUGA's Generic_Domain, restated
 * in src/anuga_HMPP/sw_domain.h:29-58). */
typedef struct sw_state {
    double time, starttime, finaltime, yieldtime, yieldstep, timestep, flux_timestep;
    double recorded_max_timestep, recorded_min_timestep;
    int64_t smallsteps, order, number_of_steps, number_of_first_order_steps;
    int64_t has_finaltime, event, error;
} sw_state;

/* Handle lifetime.  Replaces allocate_device_array + asynchronous_transfer
 * (gpu_domain_advanced.py:582-906) and the HMPP get_python_domain marshalling
 * (src/anuga_HMPP/sw_domain_fun.h:57-329).  Returns NULL on failure
 * (message via sw_last_error(NULL)). */
sw_handle *sw_create(const sw_mesh *mesh, const sw_params *params, int device);
/* Ensemble of independent domains in one handle: every fused kernel launch
 * covers all instances (each with its own timestep, time, yields and order)
 * and one hipGraph replays the step for the whole ensemble.  The reference
 * runs one domain per driver object (gpu_domain_advanced.py:47-180); this has
 * no counterpart there.  Per-instance calls (sw_upload, sw_download,
 * sw_get_state, sw_set_state, sw_set_boundary, sw_download_max_speed) address
 * the instance chosen with sw_select_instance; sw_evolve_prepare,
 * sw_evolve_steps and sw_advance_yield act on every instance (or, for
 * sw_advance_yield, on the selected one).  The operator-by-operator entry
 * points need a single-domain handle (SW_ERR_STATE otherwise). */
sw_handle *sw_create_ensemble(const sw_mesh *meshes, int32_t n_instances, const sw_params *params, int device);
/* instance in [0, n_instances), or -1 for all */
int sw_select_instance(sw_handle *h, int32_t instance);
int sw_destroy(sw_handle *h);
const char *sw_last_error(const sw_handle *h);

/* Mesh ingestion on the device.  The reference receives these arrays from
 * ANUGA's host-side General_mesh / Mesh and uploads them
 * (gpu_domain_advanced.py:582-733; layout sw_domain.h:60-180); here one
 * call builds them on GPU `device` from the points and triangles and returns
 * them in host arrays the caller allocates (row-major):
 *   triangles[N][3]     the input, v1 and v2 swapped where it is clockwise;
 *   vertex_coordinates[N][6], centroid_coordinates[N][2], areas[N];
 *   normals[N][6]       outward unit normal of edge i (opposite vertex i);
 *   edgelengths[N][3], edge_midpoint_coordinates[N][6];
 *   radii[N]            smallest centroid-to-edge-midpoint distance;
 *   neighbours[N][3], neighbour_edges[N][3]: the triangle across edge i and
 *                       that triangle's edge index, -1 on boundary edges.
 * Returns SW_ERR_ARG (message via sw_mesh_last_error) for a vertex index out
 * of range, a degenerate triangle, a zero-length edge or an edge shared by
 * more than two triangles. */
typedef struct {
    int64_t *triangles;
    double *vertex_coordinates;
    double *centroid_coordinates;
    double *areas;
    double *normals;
    double *edgelengths;
    double *edge_midpoint_coordinates;
    double *radii;
    int64_t *neighbours;
    int64_t *neighbour_edges;
} sw_mesh_geometry;
int sw_build_mesh(int device, int64_t number_of_points, const double *points, int64_t number_of_elements,
                  const int64_t *triangles, sw_mesh_geometry *out);

/* sw_build_mesh on device memory, allocating nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * points, triangles and every array of *out are device pointers on GPU
 * `device` (out itself is a host struct); the layouts, results and errors are
 * those of sw_build_mesh.  The edge keys, their sort and the error flags use
 * a scratch array the library holds on each GPU for the life of the process
 * (about 84 MB), so meshes are limited to SW_NOALLOC_MAX_TRIANGLES triangles
 * (larger ones return SW_ERR_ARG).  Calls that overlap on one GPU race on
 * that scratch: serialise them.  The call returns when the results are in
 * *out (it synchronises the null stream). */
#define SW_NOALLOC_MAX_TRIANGLES 1048576
int sw_build_mesh_noalloc(int device, int64_t number_of_points, const double *points, int64_t number_of_elements,
                          const int64_t *triangles, const sw_mesh_geometry *out);
const char *sw_mesh_last_error(void);
int sw_set_params(sw_handle *h, const sw_params *params);
int sw_get_params(const sw_handle *h, sw_params *params);

/* Boundary objects: per boundary edge kind and Dirichlet values [Nb][3]
 * (stage, xmom, ymom).  Replaces set_boundary + boundary_map dispatch
 * (gpu_domain_advanced.py:2006-2108).  dirichlet may be NULL if unused. */
int sw_set_boundary(sw_handle *h, const int32_t *kind, const double *dirichlet);

/* Host <-> device transfer of one quantity array (length N, 3N or Nb).
 * Replaces the per-array cpy_back / memcpy_htod of the reference driver
 * (src/utilities/utility.py:134-243). */
int sw_upload(sw_handle *h, int quantity, int location, const double *host, int64_t count);
int sw_download(sw_handle *h, int quantity, int location, double *host, int64_t count);
/* The same transfers with the caller's array in device memory (same layout
 * and order; elevation only through sw_upload).  For callers that keep their
 * own device copies of the state, such as the learned surrogate's rollout;
 * no counterpart in the reference. */
int sw_upload_device(sw_handle *h, int quantity, int location, const double *device_src, int64_t count);
int sw_download_device(sw_handle *h, int quantity, int location, double *device_dst, int64_t count);
/* Restart of a single domain from device buffers: uploads the centroid
 * values of stage, xmomentum and ymomentum (caller order), sets time and
 * starttime to `time`, clears a latched event or error and prepares the
 * evolve loop for yields every `yieldstep` (sw_evolve_prepare, no final
 * time); vertex values of cells in dry neighbourhoods keep their contents.
 * The matching download of the three conserved centroid arrays waits once.
 * For drivers that restart the solver from states held on the device, such
 * as the learned surrogate's judged rollout; no counterpart in the
 * reference. */
int sw_restart_device(sw_handle *h, const double *stage, const double *xmomentum, const double *ymomentum,
                      double time, double yieldstep);
int sw_download_conserved_device(sw_handle *h, double *stage, double *xmomentum, double *ymomentum);
int sw_get_state(sw_handle *h, sw_state *state);
int sw_set_state(sw_handle *h, const sw_state *state);
int sw_download_max_speed(sw_handle *h, double *host, int64_t count);

/* SWW output at yields.  The reference's store_timestep
 * (gpu_domain_advanced.py:2167-2172) hands over to ANUGA's SWW writer after
 * copy_back_necessary_data, which copies nothing (:734-737); here the device
 * reduces the state to the FP32 point values the writer stores, and only
 * those cross PCIe.
 *
 * Output points of the selected instance: with number_of_points > 0, mesh
 * nodes (ANUGA's smoothed output): node p averages the vertex values whose
 * flat caller indices 3k+j are listed, in this order, in
 * vertex_value_indices[start_p .. start_p + number_of_triangles_per_node[p])
 * (start_p the running sum of the counts; ANUGA's
 * build_inverted_triangle_structure arrays, the indices a permutation of
 * 0 .. 3N-1).  With number_of_points = 0: unique vertices, 3N points, point
 * 3k+j = vertex j of triangle k (both arrays may be NULL). */
int sw_set_sww_points(sw_handle *h, int64_t number_of_points, const int64_t *number_of_triangles_per_node,
                      const int64_t *vertex_value_indices);
/* FP32 point values out[4][P] (count = 4 P) of elevation, stage, xmomentum,
 * ymomentum: per point the mean of its vertex values (summed in the listed
 * order, divided by the count, rounded to FP32; ANUGA's
 * Quantity.get_vertex_values(smooth=True)); where stage - elevation <
 * minimum_storable_height (FP32 test) stage is stored as the elevation and
 * the momenta as 0 (ANUGA's sww writer). */
int sw_sww_values(sw_handle *h, double minimum_storable_height, float *out, int64_t count);

/* Step operators, one per overridden ANUGA method. */
/* compute_fluxes_central_structure (gpu_domain_advanced.py:1612-1646;
 * src/anuga_HMPP/shallow_water_ext.c:4519-4696).  Writes explicit updates and
 * max_speed; *flux_timestep receives the CFL minimum (host sync). */
int sw_compute_fluxes(sw_handle *h, double *flux_timestep);
/* gravity_wb (gpu_domain_advanced.py:1656-1673; shallow_water_ext.c:1942-2037) */
int sw_gravity_wb(sw_handle *h);
/* manning_friction_implicit (gpu_domain_advanced.py:1255-1324;
 * shallow_water_ext.c:1238-1317), flat or sloped per params */
int sw_manning_friction(sw_handle *h);
/* update_timestep (ANUGA Generic_Domain; src/anuga_HMPP/evolve.c:115-160) */
int sw_update_timestep(sw_handle *h, double yieldstep, double finaltime, int32_t has_finaltime);
/* update_conserved_quantities (gpu_domain_advanced.py:1412-1442; quantity.c:769-820) */
int sw_update_conserved_quantities(sw_handle *h);
/* backup_conserved_quantities (gpu_domain_advanced.py:1449-1475; evolve.c:549-561):
 * keeps a copy of the stage / xmomentum / ymomentum centroid values */
int sw_backup_conserved_quantities(sw_handle *h);
/* saxpy_conserved_quantities (gpu_domain_advanced.py:1481-1519;
 * saxpy_centroid_values.cu:1-12): centroid = a * centroid + b * backup for the
 * conserved quantities */
int sw_saxpy_conserved_quantities(sw_handle *h, double a, double b);
/* protect_against_infinitesimal_and_negative_heights
 * (gpu_domain_advanced.py:1003-1055; shallow_water_ext.c:1638-1705) */
int sw_protect(sw_handle *h);
/* _protect_swb2 (protect.cu:64-138), the protect that
 * protect_against_infinitesimal_and_negative_heights runs for the 'tsunami'
 * flow algorithm (gpu_domain_advanced.py:1014-1031): momenta zero below
 * max(0.1 * (highest bed edge midpoint - bed centroid), minimum_allowed_height)
 * of depth; a dry cell below its lowest bed vertex less minimum_allowed_height
 * gets stage and stage vertex values raised to that level.  The mass this adds
 * is not reported (the reference's kernel discards it). */
int sw_protect_swb2(sw_handle *h);
/* extrapolate_second_order_sw (gpu_domain_advanced.py:1061-1161;
 * shallow_water_ext.c:2717-3324) */
int sw_extrapolate_second_order_sw(sw_handle *h);
/* extrapolate_second_order_and_limit_by_vertex / _by_edge of stage,
 * xmomentum and ymomentum with their quantity_beta
 * (gpu_domain_advanced.py:1836-1859, 1941-1962;
 * extrapolate_second_order_and_limit_by_vertex_or_edge.cu:694-800): vertex
 * and edge values */
int sw_extrapolate_second_order_and_limit_by_vertex(sw_handle *h);
int sw_extrapolate_second_order_and_limit_by_edge(sw_handle *h);
/* extrapolate_first_order for the conserved quantities
 * (gpu_domain_advanced.py:1888-1907; extrapolate_first_order.cu:3-33) */
int sw_extrapolate_first_order(sw_handle *h);
/* balance_deep_and_shallow (gpu_domain_advanced.py:957-993;
 * shallow_water_ext.c:1462-1636) */
int sw_balance_deep_and_shallow(sw_handle *h);
/* interpolate_from_vertices_to_edges for the conserved quantities
 * (gpu_domain_advanced.py:1975-1986; quantity.c:687-707) */
int sw_interpolate_from_vertices_to_edges(sw_handle *h);
/* distribute_to_vertices_and_edges (gpu_domain_advanced.py:1744-1992) */
int sw_distribute_to_vertices_and_edges(sw_handle *h);
/* update_boundary (gpu_domain_advanced.py:2006-2108; evaluate_segment.cu:3-119;
 * set_boundary.cu:1-18) */
int sw_update_boundary(sw_handle *h);
/* update_other_quantities (gpu_domain_advanced.py:2114-2142;
 * update_centroids_of_velocities_and_height.cu:1-52) */
int sw_update_other_quantities(sw_handle *h);

/* evolve(): preamble up to the first yield (gpu_domain_advanced.py:2473-2545). */
int sw_evolve_prepare(sw_handle *h, double yieldstep, double finaltime, int32_t has_finaltime);
/* evolve() loop body (gpu_domain_advanced.py:2555-2663 / evolve.c:1466-1937):
 * runs up to max_steps steps (Euler, rk2 or rk3 per
 * params.timestepping_method, gpu_domain_advanced.py:2566-2573) on the device
 * with fused kernels replayed from a hipGraph, stopping at the first yield or
 * the final time.  Returns the number of steps taken and the event
 * (SW_EVENT_*). */
int sw_evolve_steps(sw_handle *h, int64_t max_steps, int64_t *steps_done, int32_t *event);
/* Bookkeeping when evolve() resumes after a yield (gpu_domain_advanced.py:2649-2656). */
int sw_advance_yield(sw_handle *h);
/* Unfused reference-order step (one launch per operator, as the reference
 * driver's evolve_one_euler_step + distribute + update_boundary; for rk2 /
 * rk3 with backup_conserved_quantities and saxpy_conserved_quantities between
 * the stages). */
int sw_evolve_one_step_unfused(sw_handle *h, int32_t *event);

/* Autotuner of the fused step's launch shapes (no counterpart in the
 * reference): runs `steps` real Euler steps with each candidate workgroup
 * size (128, 256, 512 lanes) for the flux, update and distribute kernels,
 * timing every launch with HIP events, and keeps the fastest size per
 * kernel.  The tuning steps advance the simulation like any other steps (the
 * results do not depend on the workgroup size).  chosen[3] (may be NULL)
 * receives the sizes kept for flux, update, distribute. */
int sw_autotune(sw_handle *h, int32_t steps, int32_t *chosen);

/* Synchronise the handle's stream. */
int sw_synchronize(sw_handle *h);
/* Kernel timing of the fused step: when enabled, sw_evolve_steps records HIP
 * events around every launch on the handle's stream; sw_kernel_times returns
 * accumulated milliseconds and launch counts per fused kernel (0 flux,
 * 1 timestep reduction, 2 update, 3 distribute) and resets them. */
int sw_enable_timing(sw_handle *h, int32_t enable);
int sw_kernel_times(sw_handle *h, double *ms, int64_t *launches, int32_t nkernels);

/* Kernels of the fused step: no counterpart in the reference (its drivers
 * launch one kernel per operator).  The vertex-record step (flux from the
 * vertex records; timestep, update, protect and the next step's friction;
 * extrapolation and balance into the vertex records) applies to FP64 Euler,
 * rk2 and rk3 steps with the optimised limiter on meshes whose geometry arrays equal those
 * recomputed from the vertex coordinates, whose bed edge values are the
 * interpolation of the bed vertex values and whose bed is single-valued at
 * every mesh node; it gives the same bits as the edge-record step (flux,
 * timestep, update, distribute).  In timing mode the vertex-record step's
 * kernels are reported as flux, update and distribute; the timestep entry
 * counts an empty interval (the update kernel reduces the timestep). */
int sw_set_step_path(sw_handle *h, int32_t path);
/* The step sw_evolve_steps would run now (SW_STEP_VERTEX or SW_STEP_EDGE). */
int sw_get_step_path(sw_handle *h, int32_t *path);

#ifdef __cplusplus
}
#endif
#endif
