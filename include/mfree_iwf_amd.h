/* mfree_iwf_amd -- MI355X (gfx950) implementation of mfree_iwf's SPH
 * particle-interaction hot path, behind a C ABI.
 *
 * The reference exposes this path only as C++ templates (no FFI of its own):
 *   body<T>(T** particles, n, simulation_data, dt, bc)   src/body.h:216-233
 *   body<T>::add_action(void(*)(body<T>&))               src/body.h:219-221
 *   body<T>::step() -> rk4<T>::step -> step_strong       src/body.h:223-225,
 *                                                        src/rk4.h:71-138
 *   compute_time_derivatives (UL / TL overloads)         src/body.h:110-148
 *   grid<T> cell list -> Verlet lists                    src/grid.h:35-177
 *   find_neighbors (TL init)                             src/benchmarks/neighbor_search.h:116-147
 *   precomp_randles_libersky (TL init)                   src/precomp_shape_functions.h:231-290
 *   boundary_conditions<T>                               src/boundary_conditions.h:27-51
 *   simulation_time singleton (time, dt)                 src/simulation_time.cpp:20-38
 * Each entry point below names the reference interface it replaces.  The
 * C++ surface with the reference's own names (body<T>, rk4<T>, particle
 * classes, actions) is the header-only adapter in include/mfree/, which is
 * written on top of these functions.
 *
 * Conventions
 *  - A context owns all device memory and one HIP stream (created on the
 *    device given in the options).  Host buffers are caller-owned and used
 *    synchronously.  A context is not thread-safe; distinct contexts may be
 *    used from different threads.
 *  - Particle arrays are SoA, in each body's idx order: element i of a
 *    body's arrays is the particle with idx == i (the reference's drivers
 *    create particle i with idx i, src/benchmarks/rings.cpp:44-52).
 *  - Errors are returned as negative mfree_status codes; they replace the
 *    reference's asserts (src/grid.h:114,167-168,
 *    src/benchmarks/neighbor_search.h:142-143) and its FE_INVALID /
 *    FE_OVERFLOW traps (src/mfree_iwf.cpp:40).
 *  - There is no CPU fallback: without a usable gfx950 device every compute
 *    entry point returns MFREE_ERR_HIP.
 */
#ifndef MFREE_IWF_AMD_H_
#define MFREE_IWF_AMD_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MFREE_ABI_VERSION 1

typedef int32_t mfree_status;
#define MFREE_OK 0
#define MFREE_ERR_ARG (-1)          /* bad argument / call order */
#define MFREE_ERR_HIP (-2)          /* HIP runtime error or no device */
#define MFREE_ERR_NBH_OVERFLOW (-3) /* >= max_nbh neighbours (src/grid.h:167) */
#define MFREE_ERR_NO_NEIGHBOR (-4)  /* particle without neighbour (src/grid.h:168) */
#define MFREE_ERR_NONFINITE (-5)    /* NaN/Inf in a derivative (src/mfree_iwf.cpp:40) */
#define MFREE_ERR_CELLS (-6)        /* num_cell >= n (src/grid.h:114) */
#define MFREE_ERR_PIPELINE (-7)     /* action list not recognised */
#define MFREE_ERR_CAPACITY (-9)     /* caller buffer too small */
#define MFREE_ERR_UNSUPPORTED (-10) /* e.g. non-zero contact forces */
#define MFREE_ERR_SCRATCH (-11)     /* input larger than an allocation-free entry point's fixed scratch */
#define MFREE_ERR_SKIN (-12)        /* a particle moved more than half the neighbour-list skin in one step,
                                       or the fastest particle moves more than h/2 per step */

/* Methods (selected in the reference by the particle type, src/body.h:60-148). */
#define MFREE_METHOD_UL_MONAGHAN 0 /* particle_monaghan, rings_monaghan action list */
#define MFREE_METHOD_TL_RL 1       /* particle_tl, tensile_tl action list */
#define MFREE_METHOD_UL_GODUNOV 2  /* particle_ul, rings_godunov action list (src/benchmarks/rings.cpp:151-158):
                                      Parshikov's Godunov-type continuity and momentum (src/godunov.h:41-189);
                                      FP64 fused pipeline only */
#define MFREE_METHOD_TL_COROTATIONAL 3 /* particle_corotational, tensile_corot action list
                                          (src/benchmarks/tensile.cpp:202-208): co-rotated displacement
                                          gradient, Cauchy-Green stress, co-rotated stress divergence
                                          (src/corotational.h:32-85, src/derivatives.h:202-258,521-588);
                                          plain kernel frozen at set-up (precomp_sph); FP64 only */
#define MFREE_METHOD_TL_POTENTIAL 4    /* particle_potential, tensile_potential action list
                                          (src/benchmarks/tensile.cpp:354-358): Green-St Venant stress,
                                          elastic-potential forces (src/potential.h:33-85); Randles-Libersky
                                          kernel of both pair directions frozen at set-up
                                          (precomp_randles_libersky_symmetric); FP64 only */

/* Pipelines of a UL FP64 context. */
#define MFREE_PIPELINE_FUSED 0     /* fused per-stage kernels (default) */
#define MFREE_PIPELINE_REFERENCE 1 /* the reference's structure on the GPU: materialised
                                      MAX_NBH neighbour and kernel lists per stage, one
                                      kernel per action (src/body.h:125-148); the
                                      "plain FP64 GPU" baseline */

/* Arithmetic of the pair sums. */
#define MFREE_PRECISION_FP64 0  /* everything FP64 (reference arithmetic) */
#define MFREE_PRECISION_MIXED 1 /* FP32 pair sums with compensated accumulation,
                                   FP64 per-particle algebra and RK4 state */

/* Per-body constants: simulation_data (src/simulation_data.h:25-100) + dt.
 * Mirrors physical_constants(nu, E, rho0), constants_monaghan(wdeltap,
 * corr_exp, eps), constants_artificial_viscosity(alpha, beta, eta),
 * correction_constants(..., xsph_eps) and rk4(n, dt). */
typedef struct mfree_params {
  double E, nu, rho0;
  double wdeltap, corr_exp, mghn_eps;
  double alpha, beta, eta;
  double xsph_eps;
  double dt;   /* time step of this body (rk4 ctor, src/rk4.h:242-243) */
  double time; /* initial simulation time of this body */
} mfree_params;

typedef struct mfree_options {
  int32_t method;           /* MFREE_METHOD_* (all bodies of a context) */
  int32_t precision;        /* MFREE_PRECISION_* */
  int32_t device;           /* HIP device ordinal */
  uint32_t max_nbh;         /* reference MAX_NBH check (src/particle.h:31); 0 = off */
  int32_t use_graph;        /* capture RK4 steps in a HIP graph */
  int32_t deterministic;    /* 1: within-cell order by idx (bit-reproducible) */
  double tl_h_search;       /* TL: h passed to find_neighbors; <= 0: per-particle h */
  int32_t tile_cells;       /* mixed UL: cells per LDS pair tile; 0 = autotune, < 0 = no tiles */
  int32_t pipeline;         /* MFREE_PIPELINE_* (REFERENCE: UL + FP64 only) */
  uint32_t xcd_chunk;       /* UL pair passes: particle tiles per XCD chunk (consecutive tiles that
                               one XCD takes in turn); 0 = library default (64); a value of at
                               least tiles/8 gives each XCD one contiguous range */
  double cell_scale;        /* mixed UL: cell side in units of 2 h_max (>= 1); 0 = autotuned with
                               tile_cells (tile_cells = 0) or 1; FP64 contexts always use 1 */
  int32_t neighbor_list;    /* mixed UL RK4 steps: 0 = default (on), 1 = on, -1 = off.  On: one
                               binning and one neighbour list per step (cells 2 h_max + skin, skin
                               >= 4 dt max|v| of the body), every stage evaluates the pairs
                               r < 2 h_i of its own positions from the list; off: the reference's
                               per-stage binning (as FP64 contexts always do) */
  int32_t list_lanes;       /* neighbour-list stages: lanes per particle (1, 2 or 4; 0 = default) */
  int32_t isolate_bodies;   /* UL, fused pipeline: 1 = a body that fails (>= max_nbh neighbours, a
                               non-finite derivative, no neighbour, its cell grid or skin) is frozen
                               at the state it started the failing step with, its time stops, and
                               the other bodies keep stepping: mfree_step returns MFREE_OK and
                               mfree_body_status reports the body's error; mfree_set_state on the
                               body thaws it.  0 (default): the first failure fails the step, as
                               the reference's assertions stop its run */
  int32_t list_staging;     /* neighbour-list stages with one lane per particle: 0 = default (on),
                               1 = on: each workgroup of 256 sorted particles copies the three
                               contiguous storage runs that hold its candidates into LDS and
                               evaluates its lists from there (workgroups whose runs exceed 1024
                               particles or that span two bodies gather from global memory);
                               -1 = off: every neighbour gathered from global memory.  Both give
                               the same sums in the same order.  Applies with list_lanes 1 (the
                               default); with 2 or 4 lanes every stage gathers from global memory */
} mfree_options;

/* SoA host view of particles (src/particle.h:42-60), idx order.  X, Y
 * (reference configuration) are read for TL bodies; fcx, fcy (contact
 * forces) may be NULL (zero); mfree_get_state fills them when given. */
typedef struct mfree_particles {
  double *x, *y, *X, *Y;
  double *m, *rho, *h;
  double *vx, *vy;
  double *Sxx, *Sxy, *Syy;
  double *fcx, *fcy;
} mfree_particles;

typedef struct mfree_ctx mfree_ctx;

/* Library identity. */
int32_t mfree_abi_version(void);
const char* mfree_status_string(mfree_status s);
/* The reference's kernel function cubic_spline (src/kernel.h:29,
 * src/kernel.cpp:20-55), host side: out = {w, w_x, w_y}.  Drivers use it for
 * the artificial-stress normalisation wdeltap = W(dx) (src/benchmarks/rings.cpp:74-75). */
void mfree_cubic_spline(double xi, double yi, double xj, double yj, double h, double out[3]);

/* Number of usable HIP devices (0 on a machine without a GPU). */
int32_t mfree_device_count(void);
void mfree_default_options(mfree_options* o);

/* Context = the set of bodies a driver steps (src/mfree_iwf.cpp:45-89 keeps
 * them in a std::vector<body<T>>).  Allocates nothing on the device yet. */
mfree_status mfree_create(const mfree_options* opt, mfree_ctx** out);
void mfree_destroy(mfree_ctx* ctx);
const char* mfree_last_error(const mfree_ctx* ctx);

/* Replaces body<T>::body(particles, n, simulation_data, dt, bc)
 * (src/body.h:216-233): registers a body and copies its particles in.
 * Returns the body id (0, 1, ...) in *body_id.  Bodies can be added until
 * the first compute call. */
mfree_status mfree_add_body(mfree_ctx* ctx, const mfree_params* p, const mfree_particles* host,
                            uint32_t n, uint32_t* body_id);

/* Replaces boundary_conditions<T>::add_boundary_condition (src/boundary_conditions.h:32-34)
 * with the TL strip functions of src/benchmarks/tensile.cpp:20-42:
 * idx_left: x = X - v_bc*t, vx = -v_bc; idx_right: x = X + v_bc*t, vx = +v_bc;
 * y = Y, vy = 0, vx_t = vy_t = 0.  TL bodies only. */
mfree_status mfree_set_strip_bc(mfree_ctx* ctx, uint32_t body, const uint32_t* idx_left, uint32_t nl,
                                const uint32_t* idx_right, uint32_t nr, double v_bc);

/* Replaces find_neighbors + precomp_randles_libersky (src/benchmarks/tensile.cpp:271-272):
 * builds the frozen TL neighbour lists and corrected kernel gradients on the
 * device from the current positions.  Called implicitly by the first
 * step of a TL context. */
mfree_status mfree_tl_precompute(mfree_ctx* ctx);

/* Overwrites / reads back a body's state (idx order).  Replaces direct
 * access to *particles[i] (src/body.h:196-198, get_particles). */
mfree_status mfree_set_state(mfree_ctx* ctx, uint32_t body, const mfree_particles* host);
mfree_status mfree_get_state(mfree_ctx* ctx, uint32_t body, mfree_particles* host);

/* Energy of one body in FP64 (no counterpart in the reference: the path's
 * energy-conservation check).  energy[0] = kinetic sum m|v|^2/2,
 * energy[1] = volumetric sum m c0^2 (ln(rho/rho0) + rho0/rho - 1) (the work
 * of the EOS p = c0^2 (rho - rho0) of src/material.h:37-64), energy[2] =
 * deviatoric sum (m/rho) S:S/(4G).  Artificial viscosity dissipates, so the
 * total decays; the check compares its trajectory across precision modes. */
mfree_status mfree_energy(mfree_ctx* ctx, uint32_t body, double* energy /* 3 */);

/* Allocation-free variant of mfree_energy.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Inputs are DEVICE arrays of one body on `device` (m, rho, vx, vy, Sxx, Sxy,
 * Syy of `dev`; the other fields are ignored), n particles, constants from
 * p (E, nu, rho0); the three energies (as mfree_energy) are written to the
 * DEVICE array energy_dev[3].  Work is enqueued on `stream` (a hipStream_t;
 * NULL = the null stream) and not waited for.  No memory is allocated or
 * freed: the per-chunk partial sums live in a fixed file-scope device array
 * of 96 KiB per GPU, so n is limited to 16,777,216 (larger n returns
 * MFREE_ERR_SCRATCH before any device work).  Calls that overlap on one GPU
 * share that array and race.  For the same arrays in the same order the
 * result is bit-identical to mfree_energy's. */
mfree_status mfree_energy_noalloc(int32_t device, void* stream, const mfree_params* p, const mfree_particles* dev,
                                  uint32_t n, double* energy_dev);
mfree_status mfree_get_time(mfree_ctx* ctx, uint32_t body, double* time);
/* The status of a body of a context with isolate_bodies: MFREE_OK, or the
 * error (MFREE_ERR_NBH_OVERFLOW, ...) that froze it.  No counterpart in the
 * reference, which stops the run at the first failure; for batches of
 * independent bodies.  Always MFREE_OK without isolate_bodies. */
mfree_status mfree_body_status(mfree_ctx* ctx, uint32_t body, mfree_status* status);

/* Results files.  Replaces vtk_writer_write (src/vtk_writer.h:27-127), which
 * src/mfree_iwf.cpp:99-103 calls after every 10th step with
 * results/out_%06d.vtk: a legacy ASCII VTK unstructured grid of every body
 * (bodies in order, particles in idx order), POINTS (x, y, 0), one vertex
 * cell per particle, POINT_DATA h, density, Sxx, Sxy, Syy (scalars) and
 * velocity (vector), every value printed with %f, the same text as the
 * reference's writer for the same state.
 *
 * mfree_vtk_write_async snapshots the current state on the context's stream
 * (a gather into idx order, into a staging buffer of its own), copies it to
 * pinned host memory on a second stream and formats the file on a host
 * thread; it returns once the snapshot is enqueued, so steps may follow at
 * once.  One file is in flight per context: the next call, mfree_vtk_wait
 * and mfree_destroy wait for it; mfree_vtk_wait returns its status.
 * mfree_vtk_write writes synchronously. */
mfree_status mfree_vtk_write(mfree_ctx* ctx, const char* path);
mfree_status mfree_vtk_write_async(mfree_ctx* ctx, const char* path);
mfree_status mfree_vtk_wait(mfree_ctx* ctx);
/* The same file from host arrays (x, y, h, rho, Sxx, Sxy, Syy, vx, vy of
 * each body are read): nbodies bodies, body k with counts[k] particles.
 * Needs no device. */
mfree_status mfree_vtk_write_host(const char* path, uint32_t nbodies, const mfree_particles* bodies,
                                  const uint32_t* counts);

/* ---------------------------------------------------------------------
 * Penalty contact between bodies (src/contact.h:36-80), run by the main
 * loop after every step (src/mfree_iwf.cpp:84-97): each body steps, its
 * outline follows its particles, then every ordered pair of bodies
 * (master i, slave j != i) sets the slave's contact forces, which its
 * next step reads as f_c / m in the momentum equations.
 */

/* Replaces boundary<T>(indexes, particles) + body<T>::set_boundary
 * (src/boundary.h:31-45, src/body.h:200-202): the outline of a body, its
 * particle idx in order around the body (the drivers take the convex hull,
 * mfree_convex_hull).  n = 0 removes it.  Forces become part of every step
 * of the context from then on. */
mfree_status mfree_set_boundary(mfree_ctx* ctx, uint32_t body, const uint32_t* idx, uint32_t n);
/* Replaces body<T>::update_boundary of every body with an outline
 * (src/body.h:152-165, src/boundary.h:47-73): outline points take their
 * particles' current positions; each normal is the mean of the outward
 * normals of the edges to its two neighbours (glm::normalize each). */
mfree_status mfree_update_boundaries(mfree_ctx* ctx);
/* The body's outline points as (x, y, nx, ny) (cap points); *n = their count
 * (pts = NULL: only the count). */
mfree_status mfree_get_boundary(mfree_ctx* ctx, uint32_t body, double* pts, uint32_t cap, uint32_t* n);
/* Replaces contact_apply_body_to_boundary(master, slave) (src/contact.h:36-80):
 * unless one of the bodies has no outline, every particle of the slave body
 * gets f_c = -m n ((p - c).n) alpha / dt^2 (alpha = 0.1, dt the slave's)
 * where c is the master outline point closest to it and (p - c).n <= 0,
 * and f_c = 0 otherwise.  The contexts may differ (on one device); the
 * master's outline must be current (mfree_update_boundaries).  The closest
 * point is exact; of points at exactly the same squared distance the first
 * in outline order is taken (the reference's k-d tree takes the first it
 * visits). */
mfree_status mfree_contact(mfree_ctx* master, uint32_t master_body, mfree_ctx* slave, uint32_t slave_body);
/* The main loop's contact pass over the bodies of one context
 * (src/mfree_iwf.cpp:91-97): mfree_contact(i, j) for i = 0.., j = 0.., i != j. */
mfree_status mfree_contact_all(mfree_ctx* ctx);
/* Replaces convex_hull (src/convex_hull.h:62-95), host only: the idx of the
 * points' convex hull, counter-clockwise from the lowest-leftmost point,
 * collinear points on the hull kept; returns their count (idx_out holds up
 * to 2n entries). */
uint32_t mfree_convex_hull(const double* x, const double* y, uint32_t n, uint32_t* idx_out);

/* Replaces body<T>::step() (src/body.h:223-225 -> src/rk4.h:71-108) for
 * every body: nsteps RK4 steps.  Each body advances its own time by its dt
 * per step. */
mfree_status mfree_step(mfree_ctx* ctx, uint32_t nsteps);

/* One derivative evaluation of every body's current state without
 * stepping: compute_time_derivatives (src/body.h:110-148).  Parity helper.
 * d8 (n_total*8, bodies concatenated, idx order): x_t, y_t, vx_t, vy_t,
 * rho_t, Sxx_t, Sxy_t, Syy_t.  aux may be NULL; otherwise n_total*16:
 * UL: quad_weight, p, Rxx, Rxy, Ryy, Sxx_x, Sxy_y, Sxy_x, Syy_y, vx_x, vx_y,
 *     vy_x, vy_y, num_nbh, 0, 0
 * TL: Hxx, Hxy, Hyx, Hyy, Fdotxx, Fdotxy, Fdotyx, Fdotyy, Pxx, Pxy, Pyx, Pyy,
 *     Pxx_x, Pxy_x, Pyx_y, Pyy_y
 * TL co-rotated: Hxx, Hxy, Hyx, Hyy, Rxx, Rxy, Ryx, Ryy, Pxx, Pxy, Pyx, Pyy, Sxx, Sxy, Syy, 0
 * TL potential:  Hxx, Hxy, Hyx, Hyy, Fe00, Fe10, Fe01, Fe11 (Fe = -2 V F S, glm [column][row]),
 *                0, 0, 0, 0, Sxx, Sxy, Syy, 0
 * (the stage's stresses of the co-rotated and potential lists, recomputed from H; those lists do
 * not integrate S or rho, so the state's S and rho keep their initial values) */
mfree_status mfree_eval_derivatives(mfree_ctx* ctx, double* d8, double* aux);

/* The Verlet lists the reference would build for a body's current state
 * (src/grid.h:117-177, predicate r^2 <= 4h_i^2 + 1e-6), in idx space,
 * rows by idx, each row in the reference's emission order.  row_ptr has
 * n+1 entries; col has cap entries; *npairs receives the total (also when
 * cap is too small, then MFREE_ERR_CAPACITY).  For TL bodies: the frozen
 * lists (find_neighbors order). */
mfree_status mfree_get_pairs(mfree_ctx* ctx, uint32_t body, uint32_t* row_ptr, uint32_t* col, uint64_t cap,
                             uint64_t* npairs);

/* Timing of the dominant (pair) kernel, measured with HIP events on the
 * context's stream while enabled (HIP graphs are not used while enabled):
 * the per-stage pair pass, or on the neighbour-list path the list-stage
 * kernel of stages 2-4. */
mfree_status mfree_profile_enable(mfree_ctx* ctx, int32_t on);
mfree_status mfree_profile_read(mfree_ctx* ctx, double* pair_ms_total, uint64_t* pair_launches,
                                double* step_ms_total, uint64_t* steps);

/* HBM calibration (no counterpart in the reference): streams `bytes` of
 * doubles through a plain grid-stride device copy kernel (k_calib_copy)
 * `reps` times on `device`; *gbps = 2*bytes*reps / elapsed.  Used to scale
 * the FETCH_SIZE / WRITE_SIZE counters of the pair kernels. */
mfree_status mfree_calibrate_copy(int32_t device, uint64_t bytes, uint32_t reps, double* gbps);

/* Tile size (cells per LDS tile) the mixed UL pair pass uses, after the
 * autotuner ran (first compute call); 0 when tiles are not used. */
int32_t mfree_tile_cells(const mfree_ctx* ctx);

/* Cell side of the mixed UL pair pass in units of the reference's 2 h_max
 * (1 unless the autotuner chose larger cells); no counterpart in the
 * reference, whose grid is fixed at 2 h_max (src/grid.h:83-115). */
double mfree_cell_scale(const mfree_ctx* ctx);

/* Total particle count and per-body counts. */
uint64_t mfree_num_particles(const mfree_ctx* ctx);
uint32_t mfree_num_bodies(const mfree_ctx* ctx);

/* ---------------------------------------------------------------------
 * Scene generators: parametrised re-creations of the reference drivers.
 */

/* rings_monaghan (src/benchmarks/rings.cpp:20-99): square lattice of
 * nbox x nbox points at -ro + i*dx, dx = 2*ro/(nbox-1), kept if
 * ri <= |p| < ro, each emitted twice, interleaved (left ring at x - centre:
 * even idx; right ring at x + centre: odd idx).  rho = rho0, h = hdx*dx,
 * m = dx*dx*rho0, vx = +v0 if x < 0 else -v0.  Optionally shuffled
 * (SplitMix64(seed) + Fisher-Yates; idx = new position) and with
 * N(0, jitter)-distributed velocity jitter (Box-Muller, same stream). */
typedef struct mfree_rings_spec {
  double ri, ro, centre;
  uint32_t nbox;
  double hdx, rho0, v0;
  int32_t shuffle;
  uint64_t seed;
  double jitter;
} mfree_rings_spec;

/* Returns the particle count (0 on bad spec). */
uint32_t mfree_scene_rings_count(const mfree_rings_spec* spec);
/* Fills caller arrays (capacity cap); X, Y get the initial positions.
 * Returns MFREE_ERR_CAPACITY if cap is too small. */
mfree_status mfree_scene_rings(const mfree_rings_spec* spec, mfree_particles* out, uint32_t cap,
                               uint32_t* n_out);

/* Plate with a central hole for the TL method (geometry of the tension
 * benchmark; lattice like src/benchmarks/tensile.cpp:232-257): points
 * (i*dx, j*dx), i < nx = round(L/dx)+1, j < ny = round(W/dx)+1, dropped if
 * closer than r_hole to the plate centre.  rho = rho0, h = hdx*dx,
 * m = dx*dx*rho0, X = x, Y = y.  Strips: i < strip (left), i >= nx-strip
 * (right) -> idx lists. */
typedef struct mfree_plate_spec {
  double L, W, r_hole, dx, hdx, rho0;
  uint32_t strip;
} mfree_plate_spec;

uint32_t mfree_scene_plate_count(const mfree_plate_spec* spec, uint32_t* n_left, uint32_t* n_right);
mfree_status mfree_scene_plate(const mfree_plate_spec* spec, mfree_particles* out, uint32_t cap,
                               uint32_t* idx_left, uint32_t* idx_right, uint32_t* n_out);

/* ---------------------------------------------------------------------
 * Learned stepper.  No counterpart in the reference: its only integrator is
 * the RK4 of src/rk4.h:71-108, which these calls stand in for, 4 steps at a
 * time, on UL (Gray-Monaghan) contexts.
 *
 * One learned step (tau = 4 dt per body):
 *   D1 = derivatives at the state (the stage-0 pass of the UL pipeline);
 *   D2 = derivatives after rho and S were advanced over tau with D1;
 *   base update: S += tau*dS1, rho *= exp(tau*drho1/rho), v += tau*a2,
 *                x += tau*xdot1 + tau^2/2*a2        (FP64)
 *   plus the network's correction of all 8 fields.  The network is a
 *   3-layer message-passing net (hidden 128, SiLU) on the radius-2h
 *   neighbour graph (edge features: kernel weight and gradient, i.e.
 *   functions of the relative position and distance), with 16
 *   dimensionless node features (current and previous velocity as their
 *   difference, stress, density, D1, D2, kernel sum, Courant number).
 *   Forward and backward run as BF16 MFMA kernels with FP32 accumulation.
 */
typedef struct mfree_surrogate mfree_surrogate;

typedef struct mfree_surrogate_options {
  int32_t device;
  uint64_t seed;             /* weight initialisation and minibatch order */
  uint32_t max_batch_nodes;  /* training minibatch cap (whole bodies of one snapshot) */
  float max_compression;     /* snapshots of bodies with max |rho/rho0 - 1| above this are not trained on */
  int32_t base_order;        /* base step of a learned step: 4 (RK4 over 4 dt, default) or 3 (Kutta) */
} mfree_surrogate_options;

void mfree_surrogate_default_options(mfree_surrogate_options* o);
mfree_status mfree_surrogate_create(const mfree_surrogate_options* o, mfree_surrogate** out);
void mfree_surrogate_destroy(mfree_surrogate* s);
const char* mfree_surrogate_last_error(const mfree_surrogate* s);

/* Training data from rollouts of ctx (any number of bodies): the
 * velocities are recorded, 4 plain RK4 steps, then per capture the
 * learned-step inputs at the current state and 4 plain steps whose result is
 * the target (the next capture follows directly, or after `gap` more steps
 * and a fresh 4-step velocity record).  One snapshot = one body at one capture; snapshots whose
 * body exceeds max_compression are kept out of training (counted by
 * mfree_surrogate_snapshots are the usable ones).  Recomputes the feature /
 * target normalisation over the usable snapshots. */
mfree_status mfree_surrogate_collect(mfree_surrogate* s, mfree_ctx* ctx, uint32_t ncapture, uint32_t gap);
uint64_t mfree_surrogate_snapshots(const mfree_surrogate* s);
uint64_t mfree_surrogate_dataset_nodes(const mfree_surrogate* s);

/* `iters` Adam steps (cosine-decayed from lr), each on the whole bodies of
 * one capture up to max_batch_nodes nodes.  loss: iters floats (normalised
 * MSE per minibatch) or NULL. */
mfree_status mfree_surrogate_train(mfree_surrogate* s, uint32_t iters, float lr, float* loss);

/* Records the current velocities as "one learned step ago" (the learned
 * step's previous-velocity feature). */
mfree_status mfree_surrogate_record_velocity(mfree_surrogate* s, mfree_ctx* ctx);

/* One learned step of every body of ctx: the base step (one RK step over
 * 4 dt, options.base_order) plus the network's correction, projected per body
 * onto zero mass-weighted mean of the x, y, vx, vy increments.  use_net = 0:
 * the base step only (the method without the network).  Like a plain step it
 * moves ctx's storage to its other buffer. */
mfree_status mfree_surrogate_step(mfree_surrogate* s, mfree_ctx* ctx, int32_t use_net);

/* Rollout accounting: learned_steps counts the learned steps whose result
 * was kept (4 dt each), plain_steps the plain RK4 steps, so
 * 4 learned_steps + plain_steps = nsteps.  max_deviation / last_deviation:
 * the judge's position deviation (below), largest over bodies;
 * max_ratio_*: the largest network-to-base error ratios of velocities,
 * density and stresses; worst_body: the body that decided the last
 * rejection; numeric_fallbacks: groups rolled back because their learned
 * steps failed numerically (non-finite values, neighbour overflow). */
typedef struct mfree_judge_stats {
  uint32_t groups, fallbacks, learned_steps, plain_steps;
  double max_deviation, last_deviation;
  double ms_learned, ms_judge, ms_fallback;
  double max_ratio_vel, max_ratio_rho, max_ratio_stress;
  uint32_t worst_body, numeric_fallbacks;
} mfree_judge_stats;

/* nsteps time steps (dt units) of ctx with the learned stepper: 4 plain
 * warm-up steps, then groups of `group` learned steps; the last one of each
 * group is judged against 4 plain RK4 steps from the same state (the plain
 * result is kept).  The judge works per body on the increments over those
 * 4 dt from the judged state: the position deviation is the relative L2
 * error of the learned displacement against the plain one; for velocities,
 * density and stresses it is the ratio of the learned step's error against
 * the plain steps to that of the same step's base step without the network
 * (the latter floored at 1e-5 c0, rho0 and K per particle, where the base
 * step is nearly exact).
 * The group is rolled back and redone with plain steps if, for any body, the
 * position deviation exceeds `threshold` or a ratio exceeds 2 (the network
 * made the step twice as wrong as no network), or if a learned step fails
 * numerically; threshold = +inf accepts every group whose learned steps
 * complete.  The last group may be shorter (at least 2 learned steps);
 * fewer than 8 remaining steps run plain. */
mfree_status mfree_surrogate_rollout(mfree_surrogate* s, mfree_ctx* ctx, uint32_t nsteps, uint32_t group,
                                     double threshold, mfree_judge_stats* stats);

/* Network parameters: flat FP32 (mfree_mpnn_param_count floats, tensors in
 * mfree_mpnn_layout order, each row-major out x in) and normalisation
 * (2*16 + 2*8 floats: feature mean, feature std, target mean, target std). */
uint64_t mfree_mpnn_param_count(void);
uint32_t mfree_mpnn_layout(int32_t* shapes /* 2 per tensor, or NULL */);
mfree_status mfree_surrogate_get_params(mfree_surrogate* s, float* w, float* norm);
mfree_status mfree_surrogate_set_params(mfree_surrogate* s, const float* w, const float* norm);

/* The learned step's graph at the current state of ctx, rows and columns by
 * global particle index (CSR), edge3 = (w, gx, gy) per edge. */
mfree_status mfree_surrogate_graph(mfree_surrogate* s, mfree_ctx* ctx, uint32_t* row_ptr, uint32_t* col,
                                   float* edge3, uint64_t cap, uint64_t* nedges);

/* Gradient check: one forward + backward of the network with weights w on a
 * host graph (CSR, at most 64 edges per node) over nodes at pos3 = (x, y, h) per
 * node (the edge features are evaluated from these as in a learned step),
 * normalised node features x (n x 16) and targets t (n x 8).  The graph
 * must be symmetric.  y: n x 8 outputs; loss: MSE; grad: dloss/dw.  With
 * grad == NULL only the inference forward runs (the fused layer kernels a
 * learned step uses): t may be NULL and loss is left untouched. */
mfree_status mfree_mpnn_forward_backward(int32_t device, uint32_t n, const uint32_t* row_ptr, const uint32_t* col,
                                         const float* pos3, const float* x, const float* t, const float* w,
                                         float* y, float* loss, float* grad);

#ifdef __cplusplus
}
#endif
#endif /* MFREE_IWF_AMD_H_ */
