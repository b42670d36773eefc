/*
 * vt_api.h -- C ABI of the MI355X (gfx950) implementation of Vortex-Transport's
 * DG right-hand side and RK4 loop.
 *
 * Plain C types only.  Host arrays use the reference's element-major order:
 *   vertices   [n][3][2]  (Element.H:27, Meshgeneration.H:9)
 *   neighbors  [n][3]     right, left, vertical  (Meshgeneration.H:10, .cpp:246-252)
 *   type       [n]        0 right angle down, 1 up (Meshgeneration.H:11)
 *   U          [n][Np][4] rho, rho u, rho v, rho E (Element.H:31)
 *   F          [n][Np][2][4]                       (Element.H:32)
 * Np = (p+1)(p+2)/2.  Every function returns a vt_status (0 = VT_OK) unless
 * noted.  A context is single-threaded and stream-ordered; separate contexts
 * are independent.  Device memory is owned by the context; host arrays are
 * owned by the caller and are only read or written during the call.
 */
#ifndef VT_API_H
#define VT_API_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum vt_status {
    VT_OK = 0,
    VT_ERR_ARG = 1,            /* bad size, pointer or connectivity */
    VT_ERR_UNSUPPORTED_P = 2,  /* p outside 0..3 (Lagrangebasis.cpp:66-69) */
    VT_ERR_QUAD = 3,           /* integration order outside 0..19 (Quadraturerule.cpp:164-167) */
    VT_ERR_NONPHYSICAL = 4,    /* rho <= 0, p <= 0 or c2 <= 0 in the Roe flux (Numericalflux.cpp:51-124) */
    VT_ERR_HIP = 5,            /* HIP runtime error, see vt_last_hip_error() */
    VT_ERR_INIT_TYPE = 6,      /* initialisation type not 0 or 1 (Element.cpp:327-330) */
    VT_ERR_METHOD = 7,         /* stepping method not 0 or 1 (main.cpp:166-169) */
    VT_ERR_MESH = 8            /* mesh sizes not supported (ny < 2) */
} vt_status;

typedef struct vt_ctx vt_ctx;

/* Text for a status code (static storage). */
const char *vt_status_string(int status);
/* Last HIP error string recorded by the library (static storage). */
const char *vt_last_hip_error(void);

/* ---- set-up (host) ------------------------------------------------------ */

/* Structured periodic mesh; replaces generate_mesh (Meshgeneration.H:15,
 * Meshgeneration.cpp:16-182) without the grid/ dump.  Cell size and origin are
 * computed in float as the reference does; perturb_grid enters through the
 * reference's sin/cos(pow()) term (NaN coordinates result for half-widths
 * above 6, as in the reference).  Arrays hold 2*nx*ny elements. */
int vt_mesh_generate(int nx, int ny, float domain_x, float domain_y, float perturb_grid,
                     double *vertices, int *neighbors, int *type);
/* Same connectivity and grid with every interior vertex moved by
 * amplitude * h * U(-1,1) per component (splitmix64 stream from seed).  No
 * reference counterpart: stands in for generate_mesh where its perturbation
 * cannot be used (domains wider than 12). */
int vt_mesh_generate_seeded(int nx, int ny, float domain_x, float domain_y, double amplitude,
                            uint64_t seed, double *vertices, int *neighbors, int *type);
/* Reference nodes [Np][2]; replaces generate_nodes_reference_space (Meshgeneration.H:17). */
int vt_nodes_reference(int p, double *nodes);
/* Reference operators [Np][Np], [Np][Np], [2][Np][Np]; replaces
 * inverse_mass_matrix_reference_space and sitffness_matrix_reference_space
 * (Preevolve.H:10,19).  Any pointer may be NULL. */
int vt_reference_operators(int p, int integration_order, double *M, double *Minv, double *S);
/* Quadrature sizes and points; replaces gauss_line_integral / gauss_area_integral
 * (Quadraturerule.H:7,10).  line_xw [nq][2], area_xyw [na][3]; NULL to query sizes. */
int vt_quadrature(int integration_order, int *nq_line, double *line_xw, int *n_area, double *area_xyw);

/* ---- solver context ----------------------------------------------------- */

/* Creates a context on the current HIP device and uploads geometry and
 * connectivity.  Replaces the element set-up loop of main.cpp:115-147: the
 * Element constructor, build_jacobians, build_mass_matrix_inverse,
 * build_stiffness_matrix (Element.cpp:21-116) and the Evolve_element
 * constructor with evaluate_basis_in_quadrature_poits (Evolve.cpp:24-132).
 * Any element numbering is accepted; with the reference's own (the two
 * triangles of a cell numbered 2c, 2c + 1, Meshgeneration.cpp:63-117) each
 * cell diagonal's flux is evaluated once for both triangles. */
int vt_create(int n_elem, const double *vertices, const int *neighbors, const int *type, int p,
              int integration_order, vt_ctx **out);
/* Same, with the operators the reference passes to each element instead of an
 * integration order: the line rule line_xw [nq][2] ({sigma, weight}, the
 * Evolve_element constructor argument, Evolve.H:15), the reference inverse mass
 * matrix Minv_ref [Np][Np] and stiffness matrices S_ref [2][Np][Np] (the
 * arguments of Element::build_mass_matrix_inverse / build_stiffness_matrix,
 * Element.H:49-50).  Used by the C++ drop-in classes. */
int vt_create_with_operators(int n_elem, const double *vertices, const int *neighbors, const int *type, int p,
                             int nq, const double *line_xw, const double *Minv_ref, const double *S_ref,
                             vt_ctx **out);
void vt_destroy(vt_ctx *ctx);
int vt_num_elements(const vt_ctx *ctx);
int vt_num_nodes(const vt_ctx *ctx);

/* Initial state; replaces Element::initialize_hydrodinamics (Element.cpp:119-333).
 * init_type 0 = nodal interpolation, 1 = least-squares projection.  The
 * reference hard-codes epsilon = 0.3, x0 = y0 = 0 (Element.cpp:136,142-143). */
int vt_init_vortex(vt_ctx *ctx, int init_type, double epsilon, double x0, double y0);
/* Exact vortex at the nodes at time t (the init_type 0 formula with t > 0). */
int vt_exact_vortex(const vt_ctx *ctx, double t, double epsilon, double x0, double y0, double *U);
/* vt_init_vortex computed on the device (Element.cpp:119-326 as the set-up of
 * large meshes and ensembles; the reference computes it on the host): no host
 * state array and no transfer.  The formula's operations are those of
 * vt_init_vortex in the same order and rounding except exp and pow, which are
 * the device library's, so values can differ from vt_init_vortex's (which
 * equal the reference's bits) in the last ulps. */
int vt_init_vortex_device(vt_ctx *ctx, int init_type, double epsilon, double x0, double y0);
/* Sets the objective's target to the nodal density of the type 0 vortex
 * (epsilon, x0, y0) at time t, computed on the device when the objective is
 * evaluated: vt_mass_l2_objective and vt_adjoint then accept rho_target = NULL
 * (no reference counterpart). */
int vt_set_target_vortex(vt_ctx *ctx, double t, double epsilon, double x0, double y0);

int vt_set_state(vt_ctx *ctx, const double *U);
int vt_get_state(vt_ctx *ctx, double *U);
/* Nodal flux of the current state, computed as Timestepping.cpp:59-76 refreshes
 * Element::hidrodynamics_vector_F after every state update. */
int vt_get_flux(vt_ctx *ctx, double *F);
/* Simulation time (Element::time, advanced by dt each step as Timestepping.cpp:91,315). */
double vt_get_time(const vt_ctx *ctx);
int vt_set_time(vt_ctx *ctx, double t);

/* dU/dt of the current state [n][Np][4]; replaces evolve_elem
 * (Timestepping.cpp:18-36) and Evolve_element's six RHS methods
 * (Evolve.cpp:135-307). */
int vt_rhs(vt_ctx *ctx, double *dUdt);
/* nsteps steps of forward Euler (method 0, Timestepping.cpp:39-82) or RK4
 * (method 1, Timestepping.cpp:85-305); replaces the time loop body of
 * main.cpp:154-170.  Steps are replayed from captured HIP graphs.  Returns
 * VT_ERR_NONPHYSICAL at the first failing flux evaluation, which
 * vt_last_error() locates; the state is then undefined (the reference exits). */
int vt_step(vt_ctx *ctx, int method, double dt, int nsteps);
/* Captures and instantiates the HIP graphs that vt_step(ctx, method, dt,
 * nsteps) replays, without running them, so that a later vt_step of that
 * length (a timed one, say) does no graph construction.  No reference
 * counterpart: the reference has no graphs. */
int vt_prepare_steps(vt_ctx *ctx, int method, double dt, int nsteps);
/* Location of the last non-physical state: element, quadrature point, side
 * (0..2), code (1 left, 2 right, 3 Roe c2, as Numericalflux.cpp:52,85,121),
 * step (1-based within the vt_step call) and stage (1..4). */
int vt_last_error(const vt_ctx *ctx, int *info /* [6] */);

/* ---- measurement and checks ---------------------------------------------- */

/* Same as vt_step, timed with HIP events recorded on the context's stream
 * around the graph launches: *ms = elapsed milliseconds. */
int vt_step_timed(vt_ctx *ctx, int method, double dt, int nsteps, double *ms);
/* Launches the four RK4 stage kernels of nsteps steps one by one, each
 * between its own pair of HIP events on the context's stream;
 * ms_per_stage[s] = average duration of stage s+1. */
int vt_stage_kernel_times(vt_ctx *ctx, double dt, int nsteps, double *ms_per_stage /* [4] */);
/* Device Roe flux on n independent inputs (records of UL[4] UR[4] n[2]);
 * replaces numerical_flux (Numericalflux.H:7) for checking.  codes[i] is 0 or
 * the reference's error number. */
int vt_numerical_flux(const double *records, double *flux /* [n][4] */, int *codes, int n);

/* Most records one vt_numerical_flux_noalloc call takes. */
#define VT_NOALLOC_MAX_RECORDS 32768

/* vt_numerical_flux without any allocation: no reference counterpart.  For
 * callers that cannot allocate device memory (tight serving loops, harnesses
 * that forbid hipMalloc).  Inputs, outputs and codes are staged through one
 * static device array of 3.8 MB (32768 records; each GPU the library runs on
 * holds it until the process ends) instead of per-call buffers; calls no
 * hipMalloc or hipFree.  n above VT_NOALLOC_MAX_RECORDS (or negative) returns
 * VT_ERR_ARG before any device call.  Calls that overlap on one GPU share the
 * array and race; calls from one thread, or serialised, do not.  Same results
 * as vt_numerical_flux. */
int vt_numerical_flux_noalloc(const double *records, double *flux /* [n][4] */, int *codes, int n);

/* L2 state-error metric of plottools/l2stateerror.py:31-99 between the current
 * state and U_ref [n][Np][4] (the script's first dump): err[k] =
 * sqrt(I_k / 25), I_k = sum_e detJ_e sum_i w_i sum_j phi_j(x_i) (U - U_ref)_jk^2
 * over the 25 points of the order-10 area rule.  The script hard-codes the
 * p = 3 basis; here phi is the context's basis (the same at p = 3).  Computed
 * on the device from full-precision values (the script reads 6-decimal text). */
int vt_l2_state_error(vt_ctx *ctx, const double *U_ref, double err[4]);

/* Launch shape of the stage kernels: block = elements per workgroup (64, the
 * default, 128 or 256); persistent != 0 runs 8 * slots_per_xcd resident workgroups that loop
 * over the element tiles instead of one workgroup per tile.  The order in which
 * the workgroups take the tiles (one contiguous band per XCD, or tile b for
 * workgroup b where the mesh's rows hold a multiple of 8 tiles) is chosen by
 * vt_create from the connectivity.  Results do not depend on either.  No
 * reference counterpart (the reference's OpenMP loop, Evolve.cpp:7-9, has no
 * launch shape). */
int vt_set_launch(vt_ctx *ctx, int block, int persistent, int slots_per_xcd);
int vt_get_launch(const vt_ctx *ctx, int *cfg /* [3] block, persistent, slots_per_xcd */);
/* Small-mesh path of vt_step (the reference's own 8 x 8-cell case): a mesh
 * of at most 128 elements whose sides each face back across the same side
 * (the reference's meshes do) is advanced by up to 64 steps per launch of one
 * workgroup with its stage state in LDS, each shared side's Roe flux evaluated
 * once per Gauss point, and the RK update in the reference's operation order
 * (Timestepping.cpp:85-305).  On by default where the mesh qualifies;
 * enable = 0 returns vt_step to the per-stage kernels, enable = 1 on a mesh
 * that does not qualify returns VT_ERR_MESH.  No reference counterpart. */
int vt_set_small_path(vt_ctx *ctx, int enable);
int vt_get_small_path(const vt_ctx *ctx);

/* ---- discrete adjoint ---------------------------------------------------- */
/* Derivatives of the discrete map of vt_step (method 1): no reference
 * counterpart (the reference computes none).  f(U) = dU/dt is the right-hand
 * side of vt_rhs (Evolve.cpp:135-307 assembled by Timestepping.cpp:18-36).  */

/* g = f'(U)^T w at the current state (w, g: [n][Np][4]). */
int vt_rhs_vjp(vt_ctx *ctx, const double *w, double *g);
/* Recorded forward run: nsteps RK4 steps of dt from the current state, as
 * vt_step, keeping the state at the start of every `every` steps on the
 * device (ceil(nsteps/every) states).  The state ends at U_T. */
int vt_adjoint_forward(vt_ctx *ctx, double dt, int nsteps, int every);
/* Identifies the recording vt_adjoint_backward would sweep: *id = the
 * number of vt_adjoint_forward calls (vt_adjoint makes one) made on the
 * context so far when the last of them completed, 0 when there is no
 * complete recording.  A caller that records several runs on one context
 * keeps the id of each and re-records a run whose id is no longer current
 * before sweeping it. */
int vt_adjoint_recording(const vt_ctx *ctx, unsigned long long *id);
/* Reverse sweep over the last recorded forward run: lam0 = (dU_T/dU_0)^T lamT.
 * Stage states are recomputed one segment (`every` steps) at a time from the
 * checkpoints; 4*every + 1 states of device memory.  The current state is
 * not changed. */
int vt_adjoint_backward(vt_ctx *ctx, const double *lamT, double *lam0);
/* Frees the context's adjoint memory (checkpoints, segment stage states,
 * adjoint planes, objective buffers), which it otherwise keeps for re-use
 * until vt_destroy; the recording is gone (vt_adjoint_recording gives 0 and
 * vt_adjoint_backward returns VT_ERR_ARG until the next vt_adjoint_forward).
 * No reference counterpart. */
int vt_adjoint_release(vt_ctx *ctx);
/* on != 0 (the default): the reverse sweep recomputes the next segment's
 * stage states on a second stream while the adjoint kernels sweep the
 * current one, in a second set of 4*every states allocated when it fits
 * beside everything else with 2 GiB to spare (else the sweep runs on one
 * stream).  Results are the same either way (no reference counterpart). */
int vt_set_adjoint_overlap(vt_ctx *ctx, int on);
/* dU0/d(epsilon, x0, y0) of vt_init_vortex type 0 at time t: dU [3][n][Np][4]. */
int vt_vortex_tangent(const vt_ctx *ctx, double t, double epsilon, double x0, double y0, double *dU);
/* J = sum_e detJ_e (rho_e - rho_target_e)^T M_ref (rho_e - rho_target_e) of the
 * current state, M_ref the reference mass matrix of Preevolve.cpp:43-54;
 * rho_target [n][Np], or NULL for the target of vt_set_target_vortex (else
 * VT_ERR_ARG); dJdU [n][Np][4] may be NULL. */
int vt_mass_l2_objective(vt_ctx *ctx, const double *rho_target, double *J, double *dJdU);
/* vt_adjoint_forward, then J and dJ/dU_T of vt_mass_l2_objective at U_T, then
 * vt_adjoint_backward: dJdU0 = dJ/dU_0. */
int vt_adjoint(vt_ctx *ctx, double dt, int nsteps, int every, const double *rho_target, double *J, double *dJdU0);

/* ---- ensemble ------------------------------------------------------------ */
/* Many independent problems advanced together: no reference counterpart (the
 * reference runs one mesh per process, main.cpp:38-176).  Each instance is
 * what vt_create would build from its own arrays; every RK stage of all
 * instances is one launch per quadrature size.  Instances share dt, step count
 * and time. */
typedef struct vt_batch vt_batch;

/* n_elem[i], p[i], order[i] per instance; vertices [sum n][3][2], neighbors
 * [sum n][3] (indices local to the instance) and type [sum n] concatenated in
 * instance order. */
int vt_batch_create(int n_inst, const int *n_elem, const int *p, const int *order, const double *vertices,
                    const int *neighbors, const int *type, vt_batch **out);
void vt_batch_destroy(vt_batch *b);
int vt_batch_num_instances(const vt_batch *b);
/* Kernel launches per RK stage: distinct quadrature sizes (at most 4 orders p
 * each), or (p, order) groups in the concurrent launch mode. */
int vt_batch_num_launches(const vt_batch *b);
/* Vortex initial state per instance (vt_init_vortex with epsilon[i], x0[i], y0[i]). */
int vt_batch_init_vortex(vt_batch *b, int init_type, const double *epsilon, const double *x0, const double *y0);
/* The same computed on the device, one launch per instance
 * (vt_init_vortex_device). */
int vt_batch_init_vortex_device(vt_batch *b, int init_type, const double *epsilon, const double *x0,
                                const double *y0);
/* Target of every instance's objective: the type 0 vortex (epsilon[i], x0[i],
 * y0[i]) at time t, computed on the device when the objectives are evaluated;
 * the ensemble objective and adjoint calls then accept rho_target = NULL. */
int vt_batch_set_target_vortex(vt_batch *b, double t, const double *epsilon, const double *x0, const double *y0);
int vt_batch_set_state(vt_batch *b, int inst, const double *U /* [n_elem[inst]][Np][4] */);
int vt_batch_get_state(vt_batch *b, int inst, double *U);
double vt_batch_get_time(const vt_batch *b);
int vt_batch_set_time(vt_batch *b, double t);
/* vt_step over all instances; VT_ERR_NONPHYSICAL at the first failing flux
 * evaluation of any instance, located by vt_batch_last_error. */
int vt_batch_step(vt_batch *b, int method, double dt, int nsteps);
/* vt_prepare_steps for an ensemble (no reference counterpart). */
int vt_batch_prepare_steps(vt_batch *b, int method, double dt, int nsteps);
int vt_batch_step_timed(vt_batch *b, int method, double dt, int nsteps, double *ms);
/* instance, element (local), quadrature point, side, code, step, stage. */
int vt_batch_last_error(const vt_batch *b, int *info /* [7] */);
/* Launch shape of the ensemble's stage launches: block 64, 128 or 256;
 * persistent 0: one launch per quadrature size, one workgroup per tile;
 * 1: the same launch with 8 * slots_per_xcd resident workgroups;
 * 2: one launch per (p, order) group, the groups run concurrently on streams
 * of their own (slots_per_xcd ignored). */
int vt_batch_set_launch(vt_batch *b, int block, int persistent, int slots_per_xcd);
int vt_batch_get_launch(const vt_batch *b, int *cfg /* [3] */);
/* Times trial_steps (1..64) RK4 steps of dt for each candidate launch shape,
 * from the current state each time, keeps the fastest and restores state
 * and time.  best_cfg [3] and best_ms (per step) may be NULL. */
int vt_batch_autotune(vt_batch *b, double dt, int trial_steps, int *best_cfg, double *best_ms);
/* Type 0 vortex state of instance inst at time t (vt_exact_vortex). */
int vt_batch_exact_vortex(const vt_batch *b, int inst, double t, double epsilon, double x0, double y0, double *U);

/* ---- Ensemble adjoint (no reference counterpart) ----------------------------
 * Every instance i has its own objective J_i of the vt_mass_l2_objective form
 * and gradient dJ_i/dU0_i.  Instances with the same (p, integration order)
 * share every launch of the forward, objective and reverse sweep, as they
 * share the launches of vt_batch_step.  Concatenated arrays list the instances
 * in order: rho_target [sum n_i][Np_i] (or NULL for the targets of
 * vt_batch_set_target_vortex), dJdU [sum n_i][Np_i][4]. */
/* J [n_inst] and, when dJdU is not NULL, dJ_i/dU_i at the current state. */
int vt_batch_mass_l2_objective(vt_batch *b, const double *rho_target, double *J, double *dJdU);
/* nsteps RK4 steps of dt recorded with a checkpoint every `every` steps
 * (the state advances to U_T as with vt_batch_step), the objectives J [n_inst]
 * at U_T, and the reverse sweep: dJdU0 = dJ_i/dU0_i.  Device memory per
 * (p, order) group of E elements: (ceil(nsteps/every) + 4 every + 8) states
 * of E * Np * 32 bytes (4 every more when the overlapped sweep fits), one
 * group's at a time: a group's is released before another group records, and
 * the last group's is kept for re-use by the next call until
 * vt_batch_adjoint_release or vt_batch_destroy.
 * VT_ERR_NONPHYSICAL as vt_batch_step. */
int vt_batch_adjoint(vt_batch *b, double dt, int nsteps, int every, const double *rho_target, double *J,
                     double *dJdU0);
/* Frees the adjoint memory vt_batch_adjoint keeps for re-use. */
int vt_batch_adjoint_release(vt_batch *b);
/* grad [n_inst][3] = dJ_i/d(epsilon_i, x0_i, y0_i) of type 0 vortex initial
 * states set at time t, from dJdU0 of vt_batch_adjoint (the tangent of
 * vt_vortex_tangent contracted with dJ_i/dU0_i). */
int vt_batch_vortex_gradient(const vt_batch *b, double t, const double *epsilon, const double *x0, const double *y0,
                             const double *dJdU0, double *grad);
/* vt_batch_adjoint followed by vt_batch_vortex_gradient without dJdU0 leaving
 * the device: the objectives J [n_inst] and grad [n_inst][3] = dJ_i/d(epsilon_i,
 * x0_i, y0_i) of type 0 vortex initial states set at the batch's time at the
 * call, the tangent contracted with dJ_i/dU0_i on the device (block sums added
 * in block order).  Same memory and errors as vt_batch_adjoint. */
int vt_batch_adjoint_vortex(vt_batch *b, double dt, int nsteps, int every, const double *rho_target,
                            const double *epsilon, const double *x0, const double *y0, double *J, double *grad);

#ifdef __cplusplus
}
#endif
#endif
