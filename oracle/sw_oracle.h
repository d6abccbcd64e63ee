/* CPU FP64 oracle for the shallow-water evolve() step -- TEST INFRASTRUCTURE.
 *
 * This library is the correctness checker for the HIP path.  Only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg load it; the product
 * (anuga_cuda_amd) never links, loads or calls it.
 *
 * It restates, operation for operation, ANUGA's FP64 C kernels as shipped in
 * the reference (src/anuga_HMPP/shallow_water_ext.c, quantity.c, util_ext.h)
 * plus the per-step orchestration of the reference GPU driver
 * (src/anuga_cuda/gpu_domain_advanced.py:2296-2698) and ANUGA's
 * update_timestep (restated by the reference in src/anuga_HMPP/evolve.c:115-160).
 * Each function cites the lines it follows in sw_oracle.c.
 *
 * Parity pinning (how this restatement is checked, since the reference C does
 * not build in this image: it needs numpy/noprefix.h, removed in numpy 2, and
 * the Python-2 C API): tests/test_oracle_*.py check closed forms (lake at
 * rest, constant-state fluxes, the KNP flux on Riemann data, the semi-implicit
 * update formula, Manning's S), invariants (mass conservation, flux
 * antisymmetry, limiter bounds), brute-force recomputation on tiny meshes and
 * the reference's own published step counts on the Merimbula mesh
 * (docs/profilling/CUDA/profile_on_GTX480_prepared_{50,300fs,500fs}_*).
 */
#ifndef SW_ORACLE_H
#define SW_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_BC_REFLECTIVE = 0, OR_BC_DIRICHLET = 1, OR_BC_TRANSMISSIVE = 2 };

/* Field names follow struct domain in src/anuga_HMPP/sw_domain.h:16-180. */
typedef struct or_domain {
    int64_t number_of_elements;
    int64_t number_of_boundary_elements;
    /* parameters */
    double epsilon, H0, g, minimum_allowed_height, maximum_allowed_speed;
    double beta_w, beta_w_dry, beta_uh, beta_uh_dry, beta_vh, beta_vh_dry;
    double alpha_balance, CFL, evolve_max_timestep, evolve_min_timestep;
    int64_t optimise_dry_cells, extrapolate_velocity_second_order;
    int64_t tight_slope_limiters, use_centroid_velocities, use_sloped_mannings;
    int64_t max_smallsteps, default_order;
    /* evolve state */
    double time, starttime, finaltime, yieldtime, yieldstep, timestep, flux_timestep;
    double recorded_max_timestep, recorded_min_timestep;
    int64_t smallsteps, _order_, number_of_steps, number_of_first_order_steps;
    int64_t has_finaltime;
    /* mesh (int64 indices as in the reference) */
    const int64_t *neighbours, *neighbour_edges, *surrogate_neighbours;
    const int64_t *number_of_boundaries, *tri_full_flag;
    const double *normals, *edgelengths, *radii, *areas;
    const double *centroid_coordinates, *vertex_coordinates;
    const int64_t *boundary_cells, *boundary_edges, *boundary_type;
    const double *boundary_dirichlet; /* [Nb][3] stage, xmom, ymom */
    /* conserved quantities */
    double *stage_centroid_values, *stage_edge_values, *stage_vertex_values, *stage_boundary_values;
    double *stage_explicit_update, *stage_semi_implicit_update;
    double *xmom_centroid_values, *xmom_edge_values, *xmom_vertex_values, *xmom_boundary_values;
    double *xmom_explicit_update, *xmom_semi_implicit_update;
    double *ymom_centroid_values, *ymom_edge_values, *ymom_vertex_values, *ymom_boundary_values;
    double *ymom_explicit_update, *ymom_semi_implicit_update;
    /* static quantities */
    double *bed_centroid_values, *bed_edge_values, *bed_vertex_values, *bed_boundary_values;
    double *friction_centroid_values;
    /* other quantities */
    double *height_centroid_values, *height_edge_values, *height_vertex_values, *height_boundary_values;
    double *xvelocity_centroid_values, *xvelocity_edge_values, *xvelocity_vertex_values, *xvelocity_boundary_values;
    double *yvelocity_centroid_values, *yvelocity_edge_values, *yvelocity_vertex_values, *yvelocity_boundary_values;
    double *max_speed;
    int64_t *already_computed_flux; /* [N][3] work array, zero-initialised */
    int64_t flux_call;              /* counter behind already_computed_flux */
    /* timestepping_method (sw_domain.h:49): 1 euler, 2 rk2, 3 rk3; backups
     * of the conserved centroid values for the multi-stage steps */
    int64_t timestepping_method;
    double *stage_centroid_backup, *xmom_centroid_backup, *ymom_centroid_backup;
    /* distribute path (gpu_domain_advanced.py:1814-1990): 0 the optimised
     * vertex limiter (extrapolate_second_order_sw), 1 per-quantity limit by
     * vertex (optimised_gradient_limiter False), 2 per-quantity limit by edge
     * (use_edge_limiter); Quantity.beta of stage, xmomentum, ymomentum */
    int64_t limiter;
    double beta_stage, beta_xmom, beta_ymom;
} or_domain;

/* status codes */
#define OR_OK 0
#define OR_ERR_UPDATE (-1)        /* semi-implicit denominator <= 0 */
#define OR_ERR_SMALLSTEPS (-2)    /* too small timestep even at first order */
#define OR_ERR_NEIGHBOUR (-3)     /* number_of_boundaries==2 without interior neighbour */

/* evolve events */
#define OR_EVENT_NONE 0
#define OR_EVENT_YIELD 1
#define OR_EVENT_FINAL 2

double or_compute_fluxes(or_domain *D);
void or_gravity_wb(or_domain *D);
void or_manning_friction_flat(or_domain *D);
void or_manning_friction_sloped(or_domain *D);
void or_compute_forcing_terms(or_domain *D);
int or_update(int64_t N, double timestep, double *centroid_values,
              double *explicit_update, double *semi_implicit_update);
int or_update_conserved_quantities(or_domain *D);
void or_protect(or_domain *D);
void or_protect_swb2(or_domain *D);
int or_extrapolate_second_order_sw(or_domain *D);
void or_extrapolate_second_order_and_limit(or_domain *D, const double *centroid, double *vertex, double *edge,
                                           double beta, int by_edge);
void or_extrapolate_first_order(int64_t N, const double *centroid, double *edge, double *vertex);
void or_balance_deep_and_shallow(or_domain *D);
void or_interpolate_from_vertices_to_edges(int64_t N, const double *vertex, double *edge);
int or_distribute_to_vertices_and_edges(or_domain *D);
void or_update_boundary(or_domain *D);
void or_update_other_quantities(or_domain *D);
int or_update_timestep(or_domain *D);
int or_evolve_one_euler_step(or_domain *D);
void or_backup_conserved_quantities(or_domain *D);
void or_saxpy_conserved_quantities(or_domain *D, double a, double b);
int or_evolve_one_rk_step(or_domain *D, int stages);
int or_evolve_prepare(or_domain *D, double yieldstep, double finaltime, int64_t has_finaltime);
int or_evolve_step(or_domain *D, int64_t *event);
int or_evolve_steps(or_domain *D, int64_t max_steps, int64_t *steps_done, int64_t *event);
void or_advance_yield(or_domain *D);
int64_t or_sizeof_domain(void);
void or_flux_function_central(const double *q_left, const double *q_right,
                              double z_left, double z_right, double n1, double n2,
                              double epsilon, double h0, double limiting_threshold,
                              double g, double *edgeflux, double *max_speed);

#ifdef __cplusplus
}
#endif
#endif
