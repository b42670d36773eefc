/*
 * vt_oracle.h -- CPU restatement of Vortex-Transport's DG right-hand side and
 * RK4 loop, used as the parity checker.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or as the timed CPU baseline).  The product path never calls it.
 *
 * Each function follows the reference expression order so that, built with
 * gcc -O2 -ffp-contract=off on x86-64, it reproduces the reference bit for bit
 * (checked against the reference's own translation units in oracle/_ref).
 *
 * Layouts (all row-major, reference order):
 *   vertices  [ne][3][2]   vertex 0 = right-angle vertex, then counter-clockwise
 *   bnd       [ne][3]      right, left, vertical neighbour (Meshgeneration.cpp:200-252)
 *   U         [ne][Np][4]  rho, rho u, rho v, rho E            (Element.H:31)
 *   F         [ne][Np][2][4] nodal flux, x then y               (Element.H:32)
 */
#ifndef VT_ORACLE_H
#define VT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes; 1..3 are the three non-physical-state exits of
 * Numericalflux.cpp:51-63 (left), :84-88 (right), :120-124 (Roe c2). */
enum { ORO_OK = 0, ORO_NONPHYS_LEFT = 1, ORO_NONPHYS_RIGHT = 2, ORO_NONPHYS_C2 = 3,
       ORO_ERR_ARG = 10, ORO_ERR_P = 11, ORO_ERR_QUAD = 12, ORO_ERR_INIT = 13 };

int oro_np(int p);
/* Gauss-Legendre rule on [0,1] for an integration order (Quadraturerule.cpp:13-181).
 * xw: [nq][2] = {sigma, weight}; returns nq or -1 (unsupported order). */
int oro_line_rule(int order, double *xw);
/* Dunavant rule on the reference triangle (Quadraturerule.cpp:183-661).
 * xyw: [n][3] = {xi, eta, weight}; returns n or -1. */
int oro_area_rule(int order, double *xyw);
/* Equispaced nodes, eta outer, xi inner (Meshgeneration.cpp:188-219). */
int oro_nodes_reference(int p, double *nodes);
/* Lagrange basis and gradient, p = 0..3 (Lagrangebasis.cpp:12-147). Return 0 or ORO_ERR_P. */
int oro_basis(int p, double xi, double eta, double *phi);
int oro_basis_grad(int p, double xi, double eta, double *grad /*[Np][2]*/);
/* Affine P1 map (Lagrangebasis.cpp:150-165). */
void oro_ref_to_phys(double xi, double eta, const double *v /*[3][2]*/, double *r);

/* Reference mass matrix by area quadrature (Preevolve.cpp:43-54). */
int oro_mass_matrix_ref(int p, int order, double *M);
/* Gauss-Jordan inverse with partial pivoting (stands where Preevolve.cpp:57-65 calls
 * Eigen::MatrixXd::inverse()).  Returns 0, or -1 if singular. */
int oro_invert(int n, const double *A, double *Ainv);
int oro_inverse_mass_matrix_ref(int p, int order, double *Minv);
/* S[0] = d/dxi, S[1] = d/deta: S_ij = sum_q grad phi_i phi_j w_q (Preevolve.cpp:86-144). */
int oro_stiffness_matrix_ref(int p, int order, double *S /*[2][Np][Np]*/);

/* Structured periodic mesh (Meshgeneration.cpp:16-182), float cell size and
 * float origin as the reference computes them. */
int oro_generate_mesh(int nx, int ny, float domain_x, float domain_y, float perturb_grid,
                      double *vertices, int *bnd, int *type);

typedef struct oro_solver oro_solver;

/* Builds per-element geometry and operators (Element.cpp:21-116) and the edge
 * basis tables (Evolve.cpp:76-132).  Minv_ref may be NULL (computed here). */
oro_solver *oro_create(int ne, const double *vertices, const int *bnd, const int *type,
                       int p, int integration_order, const double *Minv_ref);
void oro_destroy(oro_solver *s);
int oro_num_elements(const oro_solver *s);
int oro_num_nodes(const oro_solver *s);

/* Isentropic vortex initial condition, type 0 nodal interpolation and
 * type 1 least-squares projection (Element.cpp:119-333).  The reference
 * hard-codes epsilon = 0.3, x0 = y0 = 0 (Element.cpp:134-143). */
int oro_init_vortex(oro_solver *s, int init_type, double epsilon, double x0, double y0);
/* Exact vortex at time t evaluated at the nodes (the init_type 0 formula with t > 0). */
void oro_exact_vortex(const oro_solver *s, double t, double epsilon, double x0, double y0, double *U);

void oro_set_state(oro_solver *s, const double *U);   /* U then F refreshed from U */
void oro_set_state_flux(oro_solver *s, const double *U, const double *F);
void oro_get_state(const oro_solver *s, double *U);
void oro_get_flux(const oro_solver *s, double *F);
double oro_get_time(const oro_solver *s, int elem);
/* Per element: jac[4], detJ, invJ[4], detInvJ, normals[3][2], lengths[3], node coords [Np][2]. */
void oro_get_geometry(const oro_solver *s, int elem, double *out);

/* One evaluation of evolve_elem (Timestepping.cpp:18-36); dUdt may be NULL. */
int oro_evolve_elem(oro_solver *s, double *dUdt);
/* forward_euler (Timestepping.cpp:39-82) and rk4 (Timestepping.cpp:85-305). */
int oro_forward_euler(oro_solver *s, double dt);
int oro_rk4(oro_solver *s, double dt);
/* method 0 forward Euler, 1 RK4 (main.cpp:154-169).  Returns the first error
 * code and stops after that step. */
int oro_run(oro_solver *s, int method, double dt, int nsteps);
/* First error location: elem, quadrature point, side (0..2), code, step
 * (1-based within the last oro_run), stage (1..4). */
void oro_last_error(const oro_solver *s, int *info /*[6]*/);

/* Roe flux (Numericalflux.cpp:13-156). Returns 0 or the error code. */
int oro_numerical_flux(const double *UL, const double *UR, const double *n, double *flux);

/* L2 state error of plottools/l2stateerror.py:31-99: p = 3 basis, Dunavant
 * order-10 rule, 1/25 normalisation; inputs as the post-processor reads them. */
void oro_l2_state_error(int ne, const double *vertices, const double *U_first,
                        const double *U_last, double *err /*[4]*/);

/* Number of OpenMP threads the solver uses (<=0: runtime default). */
void oro_set_threads(int n);

/* ---- derivatives (vt_oracle_adjoint.c) ----------------------------------
 * Tangents are forward mode, adjoints reverse mode, written separately.  The
 * right-hand side f(U) = dU/dt is taken at the solver's current state. */
void oro_roe_flux_jvp(const double *UL, const double *UR, const double *n, const double *dUL,
                      const double *dUR, double *dF);
void oro_roe_flux_vjp(const double *UL, const double *UR, const double *n, const double *bF,
                      double *bUL, double *bUR);
void oro_rhs_jvp(const oro_solver *s, const double *dU, double *df);  /* df = f'(U) dU */
void oro_rhs_vjp(const oro_solver *s, const double *w, double *g);    /* g = f'(U)^T w */
/* nsteps RK4 steps of dt from the current state (which ends at U_T):
 * dUT = (dU_T/dU_0) dU0, lam0 = (dU_T/dU_0)^T lamT. */
int oro_rk4_tangent(oro_solver *s, double dt, int nsteps, const double *dU0, double *dUT);
int oro_rk4_adjoint(oro_solver *s, double dt, int nsteps, const double *lamT, double *lam0);
/* d U0 / d(epsilon, x0, y0) of the type 0 vortex state at time t: [3][ne][Np][4]. */
void oro_vortex_tangent(const oro_solver *s, double t, double epsilon, double x0, double y0, double *dU);
/* J = sum_e detJ_e drho_e^T M_ref drho_e with drho = rho(current) - rho_target
 * ([ne][Np]); dJdU ([ne][Np][4], may be NULL) receives its gradient. */
double oro_mass_l2_objective(const oro_solver *s, const double *rho_target, double *dJdU);

#ifdef __cplusplus
}
#endif
#endif
