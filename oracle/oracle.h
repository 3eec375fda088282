/* CPU oracle for the SPH particle-interaction hot path of mfree_iwf.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only the test
 * suite, the smoke check and the CPU-baseline leg of the benchmark may load
 * it.  The product library (libmfree_iwf_amd.so) never links, loads or
 * falls back to it.
 *
 * It restates, in plain FP64 C++, the reference's algorithm in the
 * reference's own operation order (every function cites the reference
 * file:line it follows).  Particle arrays are SoA in idx order; the oracle
 * sorts internally exactly as the reference does each stage
 * (src/body.h:134-147) and un-sorts before returning.
 *
 * Parity pinning, per function (see tests/test_oracle_*.py):
 *   orc_cubic_spline        bit-exact vs the compiled reference kernel.cpp
 *                           (oracle/_ref) and committed golden vectors.
 *   orc_grid_geometry,
 *   orc_verlet_pairs        bit-exact pair lists (and emission order, in
 *                           ORC_ORDER_REFERENCE) vs the compiled reference
 *                           grid.h + std::sort, and golden vectors.
 *   orc_find_neighbors      bit-exact vs compiled neighbor_search.h.
 *   orc_ul_derivatives,
 *   orc_tl_*                pinned by closed forms and invariants only
 *                           (the reference's physics headers include glm,
 *                           which is absent, so they cannot be compiled).
 *   orc_ul_step/orc_tl_step RK4 pinned by 4th-order convergence and by
 *                           exact rigid translation.
 */
#ifndef MFREE_ORACLE_H_
#define MFREE_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_OK 0
#define ORC_ERR_ARG (-1)
#define ORC_ERR_NBH_OVERFLOW (-3)
#define ORC_ERR_NO_NEIGHBOR (-4)
#define ORC_ERR_NONFINITE (-5)
#define ORC_ERR_CELLS (-6)
#define ORC_ERR_CAPACITY (-9)

/* Within-cell order after the per-stage sort.
 * REFERENCE: std::sort (unstable) of the idx-ordered pointer array by hash,
 *            exactly as src/body.h:137 -> same permutation as the reference;
 *            the pair loops evaluate every stored pair, as the reference
 *            does (the CPU-baseline mode).
 * STABLE:    a stable sort (counting sort) -> within a cell particles appear
 *            in idx order (the order the GPU library produces); the pair
 *            loops skip stored pairs with q >= 2, whose terms are exact +-0
 *            (the validation mode). */
#define ORC_ORDER_REFERENCE 0
#define ORC_ORDER_STABLE 1

typedef struct orc_consts {
  double E, nu, rho0;                 /* physical_constants, src/simulation_data.h:25-40 */
  double wdeltap, corr_exp, mghn_eps; /* constants_monaghan, src/simulation_data.h:42-54 */
  double alpha, beta, eta;            /* constants_artificial_viscosity :56-68 */
  double xsph_eps;                    /* correction_constants :70-87 */
} orc_consts;

/* SoA particle state in idx order (caller-owned arrays of length n).
 * Fields follow src/particle.h:42-60 (+ particle_monaghan R, particle_tl P). */
typedef struct orc_soa {
  double *x, *y, *X, *Y, *m, *rho, *h, *vx, *vy, *Sxx, *Sxy, *Syy, *fcx, *fcy;
} orc_soa;

typedef struct orc_geom {
  double bbmin_x, bbmin_y, bbmax_x, bbmax_y, cell;
  uint32_t nx, ny, num_cell;
} orc_geom;

/* src/kernel.cpp:20-55.  out = {w, w_x, w_y}. */
void orc_cubic_spline(double xi, double yi, double xj, double yj, double h, double out[3]);

/* src/simulation_data.cpp:32-47 */
double orc_G(double E, double nu);
double orc_K(double E, double nu);

/* src/grid.h:83-115 (update_geometry).  Returns ORC_ERR_CELLS if
 * num_cell >= n (the reference asserts num_cell < n, src/grid.h:114). */
int orc_grid_geometry(uint32_t n, const double* x, const double* y, const double* h,
                      double kernel_width, orc_geom* g);

/* src/grid.h:35-41,48-81,117-177 + src/body.h:134-137.
 * Emits each particle's Verlet list in idx space: row_ptr[n+1], col[] holds
 * the idx of neighbour j, rows indexed by idx i, each row in the reference's
 * emission order (3x3 block scanned ni-major, nj-minor, j in sorted order).
 * max_nbh: a row with >= max_nbh entries is ORC_ERR_NBH_OVERFLOW (the
 * reference asserts num_nbh < MAX_NBH=200, src/grid.h:167); a row with no
 * entry is ORC_ERR_NO_NEIGHBOR (src/grid.h:168).  cap = length of col. */
int orc_verlet_pairs(uint32_t n, const double* x, const double* y, const double* h,
                     int order, uint32_t max_nbh, uint32_t* row_ptr, uint32_t* col,
                     uint64_t cap, orc_geom* g);

/* src/benchmarks/neighbor_search.h:116-147 with distance_r2 (:32-41):
 * O(N^2), constant h, j scanned in idx order. */
int orc_find_neighbors(uint32_t n, const double* x, const double* y, double h,
                       uint32_t max_nbh, uint32_t* row_ptr, uint32_t* col, uint64_t cap);

/* One Gray-Monaghan UL derivative evaluation: src/body.h:125-148 with the
 * action list of src/benchmarks/rings.cpp:86-96.
 * d8 (n*8, idx order): x_t, y_t, vx_t, vy_t, rho_t, Sxx_t, Sxy_t, Syy_t.
 * aux (n*16 or NULL): quad_weight, p, Rxx, Rxy, Ryy, Sxx_x, Sxy_y, Sxy_x,
 *                     Syy_y, vx_x, vx_y, vy_x, vy_y, num_nbh, 0, 0. */
int orc_ul_derivatives(const orc_consts* c, uint32_t n, const orc_soa* s, int order,
                       uint32_t max_nbh, double* d8, double* aux);

/* RK4 (src/rk4.h:52-108) over nsteps steps of the UL body; state updated in
 * place.  threads: OpenMP threads for the per-particle loops (results are
 * independent of the thread count). */
int orc_ul_step(const orc_consts* c, uint32_t n, orc_soa* s, double dt, uint32_t nsteps,
                int order, uint32_t max_nbh, int threads);

/* The same for a chosen UL action list:
 * ORC_METHOD_MONAGHAN  rings_monaghan (src/benchmarks/rings.cpp:86-96);
 * ORC_METHOD_GODUNOV   rings_godunov (src/benchmarks/rings.cpp:151-158):
 *   quadrature weight, EOS, derive_velocity, Jaumann rate, godunov_continuity,
 *   godunov_momentum (src/godunov.h:41-189), advection; particle_ul, so no
 *   artificial stress, viscosity or XSPH.  aux[2..8] are 0 for it.
 * fcx, fcy (contact forces, may be NULL) enter both momentum equations as
 * f/m (src/cont_mech.h:44-45, src/godunov.h:148-152). */
#define ORC_METHOD_MONAGHAN 0
#define ORC_METHOD_GODUNOV 1
int orc_ul_derivatives_m(const orc_consts* c, int method, uint32_t n, const orc_soa* s, int order,
                         uint32_t max_nbh, double* d8, double* aux);
int orc_ul_step_m(const orc_consts* c, int method, uint32_t n, orc_soa* s, double dt, uint32_t nsteps,
                  int order, uint32_t max_nbh, int threads);

/* ---------------- Total-Lagrangian (Randles-Libersky) ------------------ */

/* Frozen neighbour lists + corrected kernel values of a TL body.
 * Created from the reference configuration (x = X at t = 0). */
typedef struct orc_tl orc_tl;

/* Builds the TL neighbour lists and the Randles-Libersky corrected kernel
 * (src/precomp_shape_functions.h:231-290).
 * search: 0 = O(N^2) find_neighbors with h_search (neighbor_search.h:116-147),
 *         1 = same predicate evaluated through a cell list (identical set
 *             and idx order; usable at sizes where O(N^2) is too slow).
 * contrib_only: if 1, pairs with q >= 2 (exact zero kernel) are dropped. */
int orc_tl_create(uint32_t n, const double* x, const double* y, const double* m,
                  const double* rho, const double* h, double h_search, int search,
                  int contrib_only, uint32_t max_nbh, orc_tl** out);
/* The same for a chosen TL action list:
 * ORC_TL_RL            tensile_tl (src/benchmarks/tensile.cpp:213-289);
 * ORC_TL_COROTATIONAL  tensile_corot (src/benchmarks/tensile.cpp:137-211): the plain
 *                      kernel (precomp_sph), co-rotated displacement gradient,
 *                      Cauchy-Green stress, nominal stress, co-rotated stress
 *                      divergence, viscosity, advection;
 * ORC_TL_POTENTIAL     tensile_potential (src/benchmarks/tensile.cpp:291-361): the
 *                      Randles-Libersky kernel and the reverse pairs' kernels
 *                      (precomp_randles_libersky_symmetric), Green-St Venant stress,
 *                      elastic-potential forces, viscosity, advection.
 * Derivatives of the variants (orc_tl_derivatives) carry rho_t = S_t = 0 and
 * the aux layouts:
 *   co-rotated: Hxx, Hxy, Hyx, Hyy, Rxx, Rxy, Ryx, Ryy, Pxx, Pxy, Pyx, Pyy, Sxx, Sxy, Syy, 0;
 *   potential:  Hxx, Hxy, Hyx, Hyy, Fe00, Fe10, Fe01, Fe11, 0, 0, 0, 0, Sxx, Sxy, Syy, 0
 *               (Fe = -2 V J S). */
#define ORC_TL_RL 0
#define ORC_TL_COROTATIONAL 1
#define ORC_TL_POTENTIAL 2
int orc_tl_create_m(int variant, uint32_t n, const double* x, const double* y, const double* m,
                    const double* rho, const double* h, double h_search, int search,
                    int contrib_only, uint32_t max_nbh, orc_tl** out);
void orc_tl_destroy(orc_tl* t);
uint64_t orc_tl_num_pairs(const orc_tl* t);
/* row_ptr[n+1], col[npairs], w[3*npairs] (w, w_x, w_y corrected). */
void orc_tl_get(const orc_tl* t, uint32_t* row_ptr, uint32_t* col, double* w);
/* Randles-Libersky diagnostics per particle: C (Shepard sum) and B (2x2,
 * column-major [B00,B01,B10,B11] in glm indexing) before inversion. */
void orc_tl_get_moments(const orc_tl* t, double* C, double* B4);

/* Dirichlet velocity strips (src/boundary_conditions.h:27-51 and
 * src/benchmarks/tensile.cpp:20-42): particles in idx_left get
 * x = X - v_bc*t, vx = -v_bc; idx_right x = X + v_bc*t, vx = +v_bc;
 * y = Y, vy = 0, vx_t = vy_t = 0. */
int orc_tl_set_bc(orc_tl* t, const uint32_t* idx_left, uint32_t nl,
                  const uint32_t* idx_right, uint32_t nr, double v_bc);

/* One TL derivative evaluation (src/body.h:110-123, action list of
 * src/benchmarks/tensile.cpp:276-286) at simulation time `time`.
 * The state s is modified by the boundary conditions exactly as the
 * reference does (they act on the stage copy).
 * d8 as in orc_ul_derivatives; aux (n*16 or NULL): Hxx, Hxy, Hyx, Hyy,
 * Fdotxx, Fdotxy, Fdotyx, Fdotyy, Pxx, Pxy, Pyx, Pyy, Pxx_x, Pxy_x, Pyx_y, Pyy_y. */
int orc_tl_derivatives(const orc_consts* c, orc_tl* t, uint32_t n, orc_soa* s,
                       double time, double* d8, double* aux);

/* RK4 over nsteps (src/rk4.h:71-108), BCs applied per stage and after the
 * combine; *time advanced by dt per step (src/rk4.h:107). */
int orc_tl_step(const orc_consts* c, orc_tl* t, uint32_t n, orc_soa* s, double dt,
                uint32_t nsteps, double* time, int threads);

/* ---------------- Penalty contact between bodies ------------------------ */

/* convex_hull (src/convex_hull.h:62-95): Andrew's monotone chain over the
 * points sorted by (x, y), turning test ccw, popping clockwise turns only
 * (collinear points stay); writes the hull's point indexes (counter-clockwise
 * from the lowest-leftmost point, the first not repeated) and returns their
 * count.  idx_out holds up to 2n entries. */
uint32_t orc_convex_hull(const double* x, const double* y, uint32_t n, uint32_t* idx_out);

/* body::update_boundary + boundary::update_normals (src/body.h:152-165,
 * src/boundary.h:47-73): boundary point k takes the position of particle
 * idx[k]; its normal is the mean of the outward normals of the edges to
 * its neighbours k-1 and k+1 (cyclic), each normalised as glm::normalize.
 * pts: 4 per point (x, y, nx, ny). */
void orc_boundary_update(uint32_t nb, const uint32_t* idx, const double* x, const double* y, double* pts);

/* contact_apply_body_to_boundary (src/contact.h:36-80) of one slave body
 * against one master outline: every slave particle's contact force is
 * zeroed, then set to -m (n (p - c).n) alpha / dt^2 (alpha = 0.1) where the
 * nearest master boundary point c with normal n has (p - c).n <= 0.  The
 * nearest point is found by exhaustive search; of points at exactly the
 * same squared distance the first in boundary order is taken. */
void orc_contact(uint32_t nb, const double* pts, uint32_t ns, const double* x, const double* y, const double* m,
                 double dt, double* fcx, double* fcy);

#ifdef __cplusplus
}
#endif
#endif
