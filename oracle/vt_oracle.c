/*
 * vt_oracle.c -- CPU restatement of Vortex-Transport's DG RHS + RK4 path.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker and the CPU baseline); the
 * product path never links or calls this file.
 *
 * Every arithmetic expression keeps the operand order of the reference file
 * cited beside it, so that gcc -O2 -ffp-contract=off on x86-64 gives the same
 * bits as the reference translation units built the same way.
 *
 * Pinning: the reference's own translation units (oracle/Makefile target ref)
 * on fixtures for every function of the forward path, plus closed forms and
 * invariants (tests/test_oracle_pins.py).  One part is parity unpinned at the
 * last ulp: oro_invert stands where the reference calls Eigen's .inverse()
 * (Preevolve.cpp:57-65), which is absent here; it is checked against the exact
 * rational inverse of M_ref (~1e-14), and the fixtures feed both sides the
 * same inverse.
 */
#define _DEFAULT_SOURCE
#include "vt_oracle.h"
#include "oro_internal.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oro_tables.h"


static int g_threads = 0;
void oro_set_threads(int n) { g_threads = n; }

int oro_np(int p) { return (p + 1) * (p + 2) / 2; }

int oro_line_rule(int order, double *xw) {
    const double (*t)[2] = 0;
    int n = ORO_line_rule(order, &t);
    if (n < 0) return -1;
    if (xw)
        for (int i = 0; i < n; ++i) { xw[2 * i] = t[i][0]; xw[2 * i + 1] = t[i][1]; }
    return n;
}

int oro_area_rule(int order, double *xyw) {
    const double (*t)[3] = 0;
    int n = ORO_area_rule(order, &t);
    if (n < 0) return -1;
    if (xyw)
        for (int i = 0; i < n; ++i)
            for (int c = 0; c < 3; ++c) xyw[3 * i + c] = t[i][c];
    return n;
}

/* Meshgeneration.cpp:188-219 */
int oro_nodes_reference(int p, double *nodes) {
    if (p < 0 || p > 3) return ORO_ERR_P;
    if (p == 0) { nodes[0] = 0; nodes[1] = 0; return 0; }
    double h = 1.0 / p;
    int c = 0;
    for (int j = 0; j < p + 1; ++j)
        for (int i = 0; i < p + 1; ++i)
            if (i + j < p + 1) { nodes[2 * c] = h * i; nodes[2 * c + 1] = h * j; c++; }
    return 0;
}

/* Lagrangebasis.cpp:24-63 (same polynomial expressions and operand order). */
int oro_basis(int p, double xi, double eta, double *phi) {
    switch (p) {
    case 0: phi[0] = 1.0; break;
    case 1:
        phi[0] = 1.0 - xi - eta; phi[1] = xi; phi[2] = eta; break;
    case 2:
        phi[0] = 1.0 - 3.0*xi - 3.0*eta + 2.0*xi*xi + 4.0*xi*eta + 2.0*eta*eta;
        phi[1] = 4.0*xi - 4.0*xi*xi - 4.0*xi*eta;
        phi[2] = -xi + 2.0*xi*xi;
        phi[3] = 4.0*eta - 4.0*xi*eta - 4.0*eta*eta;
        phi[4] = 4.0*xi*eta;
        phi[5] = -eta + 2.0*eta*eta;
        break;
    case 3:
        phi[0] = 1.0 - 11.0/2.0*xi - 11.0/2.0*eta + 9.0*xi*xi + 18.0*xi*eta + 9.0*eta*eta
                 - 9.0/2.0*xi*xi*xi - 27.0/2.0*xi*xi*eta - 27.0/2.0*xi*eta*eta - 9.0/2.0*eta*eta*eta;
        phi[1] = 9.0*xi - 45.0/2.0*xi*xi - 45.0/2.0*xi*eta + 27.0/2.0*xi*xi*xi + 27.0*xi*xi*eta
                 + 27.0/2.0*xi*eta*eta;
        phi[2] = -9.0/2.0*xi + 18.0*xi*xi + 9.0/2.0*xi*eta - 27.0/2.0*xi*xi*xi - 27.0/2.0*xi*xi*eta;
        phi[3] = xi - 9.0/2.0*xi*xi + 9.0/2.0*xi*xi*xi;
        phi[4] = 9.0*eta - 45.0/2.0*xi*eta - 45.0/2.0*eta*eta + 27.0/2.0*xi*xi*eta + 27.0*xi*eta*eta
                 + 27.0/2.0*eta*eta*eta;
        phi[5] = 27.0*xi*eta - 27.0*xi*xi*eta - 27.0*xi*eta*eta;
        phi[6] = -9.0/2.0*xi*eta + 27.0/2.0*xi*xi*eta;
        phi[7] = -9.0/2.0*eta + 9.0/2.0*xi*eta + 18.0*eta*eta - 27.0/2.0*xi*eta*eta - 27.0/2.0*eta*eta*eta;
        phi[8] = -9.0/2.0*xi*eta + 27.0/2.0*xi*eta*eta;
        phi[9] = eta - 9.0/2.0*eta*eta + 9.0/2.0*eta*eta*eta;
        break;
    default: return ORO_ERR_P;
    }
    return 0;
}

/* Lagrangebasis.cpp:66-147 */
int oro_basis_grad(int p, double xi, double eta, double *g) {
#define GX(i) g[2 * (i)]
#define GY(i) g[2 * (i) + 1]
    switch (p) {
    case 0: GX(0) = 0.0; GY(0) = 0.0; break;
    case 1:
        GX(0) = -1.0; GX(1) = 1.0; GX(2) = 0.0;
        GY(0) = -1.0; GY(1) = 0.0; GY(2) = 1.0;
        break;
    case 2:
        GX(0) = -3.0 + 4.0*xi + 4.0*eta;
        GX(1) = 4.0 - 8.0*xi - 4.0*eta;
        GX(2) = -1.0 + 4.0*xi;
        GX(3) = -4.0 * eta;
        GX(4) = 4.0 * eta;
        GX(5) = 0.0;
        GY(0) = -3.0 + 4.0*xi + 4.0*eta;
        GY(1) = -4.0*xi;
        GY(2) = 0.0;
        GY(3) = 4.0 - 4.0*xi - 8.0*eta;
        GY(4) = 4.0*xi;
        GY(5) = -1.0 + 4.0*eta;
        break;
    case 3:
        GX(0) = -11.0/2.0 + 18.0*xi + 18.0*eta - 27.0/2.0*xi*xi - 27.0*xi*eta - 27.0/2.0*eta*eta;
        GX(1) = 9.0 - 45.0*xi - 45.0/2.0*eta + 81.0/2.0*xi*xi + 54.0*xi*eta + 27.0/2.0*eta*eta;
        GX(2) = -9.0/2.0 + 36.0*xi + 9.0/2.0*eta - 81.0/2.0*xi*xi - 27.0*xi*eta;
        GX(3) = 1.0 - 9.0*xi + 27.0/2.0*xi*xi;
        GX(4) = -45.0/2.0*eta + 27.0*xi*eta + 27.0*eta*eta;
        GX(5) = 27.0*eta - 54.0*xi*eta - 27.0*eta*eta;
        GX(6) = -9.0/2.0*eta + 27.0*xi*eta;
        GX(7) = 9.0/2.0*eta - 27.0/2.0*eta*eta;
        GX(8) = -9.0/2.0*eta + 27.0/2.0*eta*eta;
        GX(9) = 0.0;
        GY(0) = -11.0/2.0 + 18.0*xi + 18.0*eta - 27.0/2.0*xi*xi - 27.0*xi*eta - 27.0/2.0*eta*eta;
        GY(1) = -45.0/2.0*xi + 27.0*xi*xi + 27.0*xi*eta;
        GY(2) = 9.0/2.0*xi - 27.0/2.0*xi*xi;
        GY(3) = 0.0;
        GY(4) = 9.0 - 45.0/2.0*xi - 45.0*eta + 27.0/2.0*xi*xi + 54.0*xi*eta + 81.0/2.0*eta*eta;
        GY(5) = 27.0*xi - 27.0*xi*xi - 54.0*xi*eta;
        GY(6) = -9.0/2.0*xi + 27.0/2.0*xi*xi;
        GY(7) = -9.0/2.0 + 9.0/2.0*xi + 36.0*eta - 27.0*xi*eta - 81.0/2.0*eta*eta;
        GY(8) = -9.0/2.0*xi + 27.0*xi*eta;
        GY(9) = 1.0 - 9.0*eta + 27.0/2.0*eta*eta;
        break;
    default: return ORO_ERR_P;
    }
#undef GX
#undef GY
    return 0;
}

/* Lagrangebasis.cpp:150-165 */
void oro_ref_to_phys(double xi, double eta, const double *v, double *r) {
    double phi[3];
    oro_basis(1, xi, eta, phi);
    r[0] = v[0] * phi[0] + v[2] * phi[1] + v[4] * phi[2];
    r[1] = v[1] * phi[0] + v[3] * phi[1] + v[5] * phi[2];
}

/* Preevolve.cpp:15-54 */
int oro_mass_matrix_ref(int p, int order, double *M) {
    if (p < 0 || p > 3) return ORO_ERR_P;
    double q[73 * 3];
    int n = oro_area_rule(order, q);
    if (n < 0) return ORO_ERR_QUAD;
    int np = oro_np(p);
    double *phi = malloc(sizeof(double) * np * n);
    for (int i = 0; i < n; ++i) {
        double b[MAXNP];
        oro_basis(p, q[3 * i], q[3 * i + 1], b);
        for (int j = 0; j < np; ++j) phi[j * n + i] = b[j];
    }
    for (int i = 0; i < np; ++i)
        for (int j = 0; j < np; ++j) {
            double m = 0;
            for (int k = 0; k < n; ++k) m += phi[i * n + k] * phi[j * n + k] * q[3 * k + 2];
            M[i * np + j] = m;
        }
    free(phi);
    return 0;
}

/* Gauss-Jordan elimination with partial pivoting on [A | I]. */
int oro_invert(int n, const double *A, double *Ainv) {
    double *a = malloc(sizeof(double) * n * 2 * n);
    int w = 2 * n;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < w; ++j) a[i * w + j] = j < n ? A[i * n + j] : (j - n == i ? 1.0 : 0.0);
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (fabs(a[r * w + c]) > fabs(a[piv * w + c])) piv = r;
        if (a[piv * w + c] == 0.0) { free(a); return -1; }
        if (piv != c)
            for (int j = 0; j < w; ++j) { double t = a[c * w + j]; a[c * w + j] = a[piv * w + j]; a[piv * w + j] = t; }
        double d = a[c * w + c];
        for (int j = 0; j < w; ++j) a[c * w + j] /= d;
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            double f = a[r * w + c];
            if (f == 0.0) continue;
            for (int j = 0; j < w; ++j) a[r * w + j] -= f * a[c * w + j];
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) Ainv[i * n + j] = a[i * w + n + j];
    free(a);
    return 0;
}

int oro_inverse_mass_matrix_ref(int p, int order, double *Minv) {
    double M[MAXNP * MAXNP];
    int rc = oro_mass_matrix_ref(p, order, M);
    if (rc) return rc;
    return oro_invert(oro_np(p), M, Minv) ? ORO_ERR_ARG : 0;
}

/* Preevolve.cpp:86-144 */
int oro_stiffness_matrix_ref(int p, int order, double *S) {
    if (p < 0 || p > 3) return ORO_ERR_P;
    double q[73 * 3];
    int n = oro_area_rule(order, q);
    if (n < 0) return ORO_ERR_QUAD;
    int np = oro_np(p);
    double *phi = malloc(sizeof(double) * np * n);
    double *gr = malloc(sizeof(double) * np * n * 2);
    for (int i = 0; i < n; ++i) {
        double b[MAXNP], g[2 * MAXNP];
        oro_basis(p, q[3 * i], q[3 * i + 1], b);
        oro_basis_grad(p, q[3 * i], q[3 * i + 1], g);
        for (int j = 0; j < np; ++j) {
            phi[j * n + i] = b[j];
            gr[(j * n + i) * 2] = g[2 * j];
            gr[(j * n + i) * 2 + 1] = g[2 * j + 1];
        }
    }
    for (int i = 0; i < np; ++i)
        for (int j = 0; j < np; ++j) {
            double s0 = 0, s1 = 0;
            for (int k = 0; k < n; ++k) {
                s0 += gr[(i * n + k) * 2] * phi[j * n + k] * q[3 * k + 2];
                s1 += gr[(i * n + k) * 2 + 1] * phi[j * n + k] * q[3 * k + 2];
            }
            S[i * np + j] = s0;
            S[np * np + i * np + j] = s1;
        }
    free(phi);
    free(gr);
    return 0;
}

/* Meshgeneration.cpp:16-182 (the grid/ text dump is not reproduced). */
int oro_generate_mesh(int nx, int ny, float domain_x, float domain_y, float perturb_grid,
                      double *V, int *bnd, int *type) {
    if (nx < 1 || ny < 1) return ORO_ERR_ARG;
    int ne = 2 * nx * ny;
    double *gx = malloc(sizeof(double) * (nx + 1)), *gy = malloc(sizeof(double) * (ny + 1));
    double element_size_x = domain_x / nx; /* float division, then widened */
    double element_size_y = domain_y / ny;
    for (int i = 0; i < nx + 1; ++i) gx[i] = -1 * domain_x / 2 + element_size_x * i;
    for (int i = 0; i < ny + 1; ++i) gy[i] = -1 * domain_y / 2 + element_size_y * i;
    int npts = (nx + 1) * (ny + 1);
    double *px = malloc(sizeof(double) * npts), *py = malloc(sizeof(double) * npts);
    int c = 0;
    for (int i = 0; i < ny + 1; ++i)
        for (int j = 0; j < nx + 1; ++j) {
            px[c] = gx[j] + (i == 0 ? 0 : 1) * (i == ny ? 0 : 1) * (j == 0 ? 0 : 1) * (j == nx ? 0 : 1)
                    * perturb_grid * 0.2 * element_size_x * sin(pow(gx[j] + 6, gy[i] + 6));
            py[c] = gy[i] + (i == 0 ? 0 : 1) * (i == ny ? 0 : 1) * (j == 0 ? 0 : 1) * (j == nx ? 0 : 1)
                    * perturb_grid * 0.2 * element_size_y * cos(pow(gy[i] + 6, gx[j] + 6));
            c++;
        }
    int *n1 = malloc(sizeof(int) * ne), *n2 = malloc(sizeof(int) * ne), *n3 = malloc(sizeof(int) * ne);
    c = 0;
    for (int i = 1; i < ny + 1; ++i)
        for (int j = 0; j < nx; ++j) {
            n1[c++] = j + (nx + 1) * (i - 1);
            n1[c++] = (j + (nx + 1) * i) + 1;
        }
    for (int i = 0; i < ne; ++i) n2[i] = n1[i] + (i % 2 == 0 ? 1 : -1);
    for (int i = 0; i < ne; ++i) n3[i] = n2[(i % 2 == 0 ? i + 1 : i - 1)];
    int rc = 0;
    for (int i = 0; i < ne; ++i) {
        type[i] = (i % 2 == 0 ? 0 : 1);
        int left = (i % (2 * nx) == 0 ? i - 1 + 2 * nx : i - 1);
        int right = ((i + 1) % (2 * nx) == 0 ? i + 1 - 2 * nx : i + 1);
        int vert;
        if (i >= 2 * nx * (ny - 1))
            vert = (i % 2 == 0 ? i + 1 - 2 * nx : i - 1 - 2 * nx * (ny - 1));
        else if (i >= 2 * nx)
            vert = (i % 2 == 0 ? i + 1 - 2 * nx : i - 1 + 2 * nx);
        else
            vert = (i % 2 == 0 ? i + 1 + 2 * nx * (ny - 1) : i - 1 + 2 * nx);
        bnd[3 * i] = right; bnd[3 * i + 1] = left; bnd[3 * i + 2] = vert;
        for (int s = 0; s < 3; ++s)
            if (bnd[3 * i + s] < 0 || bnd[3 * i + s] >= ne) rc = ORO_ERR_ARG;
        V[6 * i + 0] = px[n1[i]]; V[6 * i + 1] = py[n1[i]];
        V[6 * i + 2] = px[n2[i]]; V[6 * i + 3] = py[n2[i]];
        V[6 * i + 4] = px[n3[i]]; V[6 * i + 5] = py[n3[i]];
    }
    free(gx); free(gy); free(px); free(py); free(n1); free(n2); free(n3);
    return rc;
}

int oro_num_elements(const oro_solver *s) { return s->ne; }
int oro_num_nodes(const oro_solver *s) { return s->np; }

/* Element.cpp:54-93 */
static void build_jacobians(const double *v, double *g, double *len) {
    double *jac = g + GEO_JAC, *inv = g + GEO_INVJ, *nrm = g + GEO_N;
    jac[0] = v[2] - v[0];
    jac[1] = v[4] - v[0];
    jac[2] = v[3] - v[1];
    jac[3] = v[5] - v[1];
    g[GEO_DET] = jac[0] * jac[3] - jac[2] * jac[1];
    double d = g[GEO_DET];
    inv[0] = (1 / d) * (v[5] - v[1]);
    inv[1] = (1 / d) * (v[0] - v[4]);
    inv[2] = (1 / d) * (v[1] - v[3]);
    inv[3] = (1 / d) * (v[2] - v[0]);
    g[GEO_DETINV] = inv[0] * inv[3] - inv[2] * inv[1];
    nrm[0] = (v[3] - v[1]) / pow(pow(v[2] - v[0], 2) + pow(v[3] - v[1], 2), 0.5);
    nrm[2] = (v[5] - v[3]) / pow(pow(v[4] - v[2], 2) + pow(v[5] - v[3], 2), 0.5);
    nrm[4] = (v[1] - v[5]) / pow(pow(v[0] - v[4], 2) + pow(v[1] - v[5], 2), 0.5);
    nrm[1] = -1 * (v[2] - v[0]) / pow(pow(v[2] - v[0], 2) + pow(v[3] - v[1], 2), 0.5);
    nrm[3] = -1 * (v[4] - v[2]) / pow(pow(v[4] - v[2], 2) + pow(v[5] - v[3], 2), 0.5);
    nrm[5] = -1 * (v[0] - v[4]) / pow(pow(v[0] - v[4], 2) + pow(v[1] - v[5], 2), 0.5);
    len[0] = pow(pow(v[2] - v[0], 2) + pow(v[3] - v[1], 2), 0.5);
    len[1] = pow(pow(v[4] - v[2], 2) + pow(v[5] - v[3], 2), 0.5);
    len[2] = pow(pow(v[0] - v[4], 2) + pow(v[1] - v[5], 2), 0.5);
}

oro_solver *oro_create(int ne, const double *vertices, const int *bnd, const int *type, int p,
                       int integration_order, const double *Minv_ref) {
    if (ne < 1 || p < 0 || p > 3) return 0;
    int nq = oro_line_rule(integration_order, 0);
    if (nq < 0 || oro_area_rule(integration_order, 0) < 0) return 0;
    /* An exactly singular M_ref (a one-point area rule at p >= 1) has no
     * inverse; the reference's Eigen call does not check and steps NaN states.
     * Such configurations are refused, as the library refuses them. */
    double minv_probe[MAXNP * MAXNP];
    if (!Minv_ref && oro_inverse_mass_matrix_ref(p, integration_order, minv_probe)) return 0;
    for (int e = 0; e < ne; ++e)
        for (int k = 0; k < 3; ++k)
            if (bnd[3 * e + k] < 0 || bnd[3 * e + k] >= ne) return 0;
    oro_solver *s = calloc(1, sizeof(*s));
    s->ne = ne; s->p = p; s->np = oro_np(p); s->nq = nq; s->order = integration_order;
    oro_line_rule(integration_order, &s->xw[0][0]);
    int np = s->np;
    s->vert = malloc(sizeof(double) * 6 * ne);
    memcpy(s->vert, vertices, sizeof(double) * 6 * ne);
    s->bnd = malloc(sizeof(int) * 3 * ne);
    memcpy(s->bnd, bnd, sizeof(int) * 3 * ne);
    s->type = malloc(sizeof(int) * ne);
    memcpy(s->type, type, sizeof(int) * ne);
    s->nodes_phys = malloc(sizeof(double) * ne * np * 2);
    s->geo = malloc(sizeof(double) * ne * GEO_STRIDE);
    s->len = malloc(sizeof(double) * ne * 3);
    s->Minv_phys = malloc(sizeof(double) * ne * np * np);
    s->S_phys = malloc(sizeof(double) * ne * 2 * np * np);
    s->U = calloc((size_t)ne * np * 4, sizeof(double));
    s->F = calloc((size_t)ne * np * 8, sizeof(double));
    s->dU = calloc((size_t)ne * np * 4, sizeof(double));
    s->time = calloc(ne, sizeof(double));
    if (Minv_ref) memcpy(s->Minv_ref, Minv_ref, sizeof(double) * np * np);
    else oro_inverse_mass_matrix_ref(p, integration_order, s->Minv_ref);
    oro_stiffness_matrix_ref(p, integration_order, s->S_ref);
    double nodes[2 * MAXNP];
    oro_nodes_reference(p, nodes);
    for (int e = 0; e < ne; ++e) {
        const double *v = s->vert + 6 * e;
        for (int i = 0; i < np; ++i) oro_ref_to_phys(nodes[2 * i], nodes[2 * i + 1], v, s->nodes_phys + (e * np + i) * 2);
        double *g = s->geo + e * GEO_STRIDE;
        build_jacobians(v, g, s->len + 3 * e);
        /* Element.cpp:96-116 */
        double d = g[GEO_DET];
        double *inv = g + GEO_INVJ;
        for (int i = 0; i < np; ++i)
            for (int j = 0; j < np; ++j) {
                s->Minv_phys[(size_t)e * np * np + i * np + j] = s->Minv_ref[i * np + j] / d;
                double a = s->S_ref[i * np + j], b = s->S_ref[np * np + i * np + j];
                s->S_phys[(size_t)e * 2 * np * np + i * np + j] = d * (a * inv[0] + b * inv[2]);
                s->S_phys[(size_t)e * 2 * np * np + np * np + i * np + j] = d * (a * inv[1] + b * inv[3]);
            }
    }
    /* Evolve.cpp:76-132 */
    for (int q = 0; q < nq; ++q) {
        double sg = s->xw[q][0], b[6][MAXNP];
        oro_basis(p, sg, 0.0, b[0]);
        oro_basis(p, 1.0 - sg, 0.0, b[1]);
        oro_basis(p, 1.0 - sg, sg, b[2]);
        oro_basis(p, sg, 1.0 - sg, b[3]);
        oro_basis(p, 0.0, 1.0 - sg, b[4]);
        oro_basis(p, 0.0, sg, b[5]);
        for (int j = 0; j < np; ++j)
            for (int side = 0; side < 3; ++side) {
                s->plus[side][j][q] = b[2 * side][j];
                s->minus[side][j][q] = b[2 * side + 1][j];
            }
    }
    return s;
}

void oro_destroy(oro_solver *s) {
    if (!s) return;
    free(s->vert); free(s->bnd); free(s->type); free(s->nodes_phys); free(s->geo); free(s->len);
    free(s->Minv_phys); free(s->S_phys); free(s->U); free(s->F); free(s->dU); free(s->time);
    free(s);
}

/* Nodal flux refresh, repeated at Timestepping.cpp:59-76, 141-158, 183-200,
 * 225-242, 283-300 and in Element.cpp:306-323. */
static inline void refresh_flux(const double *U, double *F) {
    double gamma = 1.4;
    double rho = U[0];
    double u = U[1] / rho;
    double v = U[2] / rho;
    double E = U[3] / rho;
    double p = rho * (gamma - 1) * (E - (pow(u, 2) + pow(v, 2)) / 2);
    double H = E + p / rho;
    F[0] = rho * u;
    F[1] = rho * pow(u, 2) + p;
    F[2] = rho * u * v;
    F[3] = rho * u * H;
    F[4] = rho * v;
    F[5] = rho * u * v;
    F[6] = rho * pow(v, 2) + p;
    F[7] = rho * v * H;
}

/* Element.cpp:134-160 */
static void vortex_point(double x, double y, double t, double epsilon, double x0, double y0,
                         double *out /* rho u v p E H */) {
    double rho_infty = 1.0, rc = 1.0, gamma = 1.4, M_infty = 0.5, p_infty = 20.0 / 7.0;
    double U_infty = 1.0 / pow(2.0, 0.5), V_infty = 1.0 / pow(2.0, 0.5);
    double f0 = 1.0 - (pow(x - x0 - U_infty * t, 2.0) + pow(y - y0 - V_infty * t, 2.0)) / pow(rc, 2.0);
    double f1 = 1.0 - pow(epsilon, 2.0) * (gamma - 1) * pow(M_infty, 2.0) * exp(f0) / (8.0 * pow(M_PI, 2.0));
    double f2 = epsilon * (pow(U_infty, 2.0) + pow(V_infty, 2.0)) * exp(f0 / 2.0) / (2.0 * M_PI * rc);
    double rho = rho_infty * pow(f1, 1.0 / (gamma - 1.0));
    double u = U_infty - f2 * (y - y0 - V_infty * t);
    double v = V_infty + f2 * (x - x0 - U_infty * t);
    double p = p_infty * pow(f1, gamma / (gamma - 1.0));
    double E = p / (rho * (gamma - 1.0)) + (pow(u, 2.0) + pow(v, 2.0)) / 2.0;
    double H = E + p / rho;
    out[0] = rho; out[1] = u; out[2] = v; out[3] = p; out[4] = E; out[5] = H;
}

int oro_init_vortex(oro_solver *s, int init_type, double epsilon, double x0, double y0) {
    int np = s->np;
    if (init_type == 0) {
        for (int e = 0; e < s->ne; ++e)
            for (int i = 0; i < np; ++i) {
                double w[6];
                const double *xy = s->nodes_phys + (e * np + i) * 2;
                vortex_point(xy[0], xy[1], s->time[e], epsilon, x0, y0, w);
                double rho = w[0], u = w[1], v = w[2], p = w[3], E = w[4], H = w[5];
                double *U = s->U + ((size_t)e * np + i) * 4, *F = s->F + ((size_t)e * np + i) * 8;
                U[0] = rho; U[1] = rho * u; U[2] = rho * v; U[3] = rho * E;
                F[0] = rho * u; F[1] = rho * pow(u, 2) + p; F[2] = rho * u * v; F[3] = rho * u * H;
                F[4] = rho * v; F[5] = rho * u * v; F[6] = rho * pow(v, 2) + p; F[7] = rho * v * H;
            }
        return 0;
    }
    if (init_type == 1) { /* Element.cpp:180-326 */
        double q[73 * 3];
        int nqa = oro_area_rule(s->order, q);
        double *phi = malloc(sizeof(double) * np * nqa);
        double *ue = malloc(sizeof(double) * nqa * 4);
        for (int e = 0; e < s->ne; ++e) {
            const double *v = s->vert + 6 * e;
            double detJ = s->geo[e * GEO_STRIDE + GEO_DET];
            for (int i = 0; i < nqa; ++i) {
                double b[MAXNP], r[2], w[6];
                oro_basis(s->p, q[3 * i], q[3 * i + 1], b);
                for (int j = 0; j < np; ++j) phi[j * nqa + i] = b[j];
                oro_ref_to_phys(q[3 * i], q[3 * i + 1], v, r);
                vortex_point(r[0], r[1], s->time[e], epsilon, x0, y0, w);
                ue[4 * i] = w[0]; ue[4 * i + 1] = w[0] * w[1]; ue[4 * i + 2] = w[0] * w[2]; ue[4 * i + 3] = w[0] * w[4];
            }
            double bvec[MAXNP][4];
            for (int i = 0; i < np; ++i)
                for (int k = 0; k < 4; ++k) {
                    bvec[i][k] = 0.0;
                    for (int j = 0; j < nqa; ++j) bvec[i][k] += detJ * phi[i * nqa + j] * ue[4 * j + k] * q[3 * j + 2];
                }
            for (int i = 0; i < np; ++i) {
                double *U = s->U + ((size_t)e * np + i) * 4;
                for (int k = 0; k < 4; ++k) {
                    U[k] = 0.0;
                    for (int j = 0; j < np; ++j) U[k] += s->Minv_phys[(size_t)e * np * np + i * np + j] * bvec[j][k];
                }
                refresh_flux(U, s->F + ((size_t)e * np + i) * 8);
            }
        }
        free(phi);
        free(ue);
        return 0;
    }
    return ORO_ERR_INIT;
}

void oro_exact_vortex(const oro_solver *s, double t, double epsilon, double x0, double y0, double *Uout) {
    for (int e = 0; e < s->ne; ++e)
        for (int i = 0; i < s->np; ++i) {
            double w[6];
            const double *xy = s->nodes_phys + (e * s->np + i) * 2;
            vortex_point(xy[0], xy[1], t, epsilon, x0, y0, w);
            double *U = Uout + ((size_t)e * s->np + i) * 4;
            U[0] = w[0]; U[1] = w[0] * w[1]; U[2] = w[0] * w[2]; U[3] = w[0] * w[4];
        }
}

void oro_set_state(oro_solver *s, const double *U) {
    size_t n = (size_t)s->ne * s->np;
    memcpy(s->U, U, sizeof(double) * n * 4);
    for (size_t i = 0; i < n; ++i) refresh_flux(s->U + 4 * i, s->F + 8 * i);
}
void oro_set_state_flux(oro_solver *s, const double *U, const double *F) {
    size_t n = (size_t)s->ne * s->np;
    memcpy(s->U, U, sizeof(double) * n * 4);
    memcpy(s->F, F, sizeof(double) * n * 8);
}
void oro_get_state(const oro_solver *s, double *U) { memcpy(U, s->U, sizeof(double) * s->ne * s->np * 4); }
void oro_get_flux(const oro_solver *s, double *F) { memcpy(F, s->F, sizeof(double) * s->ne * s->np * 8); }
double oro_get_time(const oro_solver *s, int e) { return s->time[e]; }
void oro_get_geometry(const oro_solver *s, int e, double *out) {
    memcpy(out, s->geo + e * GEO_STRIDE, sizeof(double) * 16);
    memcpy(out + 16, s->len + 3 * e, sizeof(double) * 3);
    memcpy(out + 19, s->nodes_phys + (size_t)e * s->np * 2, sizeof(double) * s->np * 2);
}
void oro_last_error(const oro_solver *s, int *info) { memcpy(info, s->err, sizeof(s->err)); }

/* Numericalflux.cpp:22-166 */
int oro_numerical_flux(const double *u_left, const double *u_right, const double *nv, double *flux) {
    int err = 0;
    double gamma = 1.4;
    double gmi = gamma - 1.0;
    double rL = u_left[0];
    double uL = u_left[1] / rL;
    double vL = u_left[2] / rL;
    double unL = uL * nv[0] + vL * nv[1];
    double qL = pow(pow(u_left[1], 2) + pow(u_left[2], 2), 0.5) / rL;
    double pL = (gamma - 1) * (u_left[3] - 0.5 * rL * pow(qL, 2));
    if ((pL <= 0) || (rL <= 0)) err = ORO_NONPHYS_LEFT;
    double rHL = u_left[3] + pL;
    double HL = rHL / rL;
    double FL[4];
    FL[0] = rL * unL;
    FL[1] = u_left[1] * unL + pL * nv[0];
    FL[2] = u_left[2] * unL + pL * nv[1];
    FL[3] = rHL * unL;
    double rR = u_right[0];
    double uR = u_right[1] / rR;
    double vR = u_right[2] / rR;
    double unR = uR * nv[0] + vR * nv[1];
    double qR = pow(pow(u_right[1], 2) + pow(u_right[2], 2), 0.5) / rR;
    double pR = (gamma - 1) * (u_right[3] - 0.5 * rR * pow(qR, 2));
    if (!err && ((pR <= 0) || (rR <= 0))) err = ORO_NONPHYS_RIGHT;
    double rHR = u_right[3] + pR;
    double HR = rHR / rR;
    double FR[4];
    FR[0] = rR * unR;
    FR[1] = u_right[1] * unR + pR * nv[0];
    FR[2] = u_right[2] * unR + pR * nv[1];
    FR[3] = rHR * unR;
    double du[4];
    du[0] = u_right[0] - u_left[0];
    du[1] = u_right[1] - u_left[1];
    du[2] = u_right[2] - u_left[2];
    du[3] = u_right[3] - u_left[3];
    double di = pow(rR / rL, 0.5);
    double d1 = 1.0 / (1.0 + di);
    double ui = (di * uR + uL) * d1;
    double vi = (di * vR + vL) * d1;
    double Hi = (di * HR + HL) * d1;
    double af = 0.5 * (ui * ui + vi * vi);
    double ucp = ui * nv[0] + vi * nv[1];
    double c2 = gmi * (Hi - af);
    if (!err && c2 <= 0) err = ORO_NONPHYS_C2;
    double ci = pow(c2, 2);
    double ci1 = 1.0 / ci;
    double l[3];
    l[0] = ucp + ci;
    l[1] = ucp - ci;
    l[2] = ucp;
    double epsilon = ci * 0.1;
    for (int i = 0; i < 3; ++i)
        if ((l[i] < epsilon) && (l[i] > -1 * epsilon)) l[i] = 0.5 * (epsilon + l[i] * l[i] / epsilon);
    for (int i = 0; i < 3; ++i) l[i] = fabs(l[i]);
    double l3 = l[2];
    double s1 = 0.5 * (l[0] + l[1]);
    double s2 = 0.5 * (l[0] - l[1]);
    double G1 = gmi * (af * du[0] - ui * du[1] - vi * du[2] + du[3]);
    double G2 = -1 * ucp * du[0] + du[1] * nv[0] + du[2] * nv[1];
    double C1 = G1 * (s1 - l3) * ci1 * ci1 + G2 * s2 * ci1;
    double C2 = G1 * s2 * ci1 + G2 * (s1 - l3);
    flux[0] = 0.5 * (FL[0] + FR[0]) - 0.5 * (l3 * du[0] + C1);
    flux[1] = 0.5 * (FL[1] + FR[1]) - 0.5 * (l3 * du[1] + C1 * ui + C2 * nv[0]);
    flux[2] = 0.5 * (FL[2] + FR[2]) - 0.5 * (l3 * du[2] + C1 * vi + C2 * nv[1]);
    flux[3] = 0.5 * (FL[3] + FR[3]) - 0.5 * (l3 * du[3] + C1 * Hi + C2 * ucp);
    return err;
}

/* Evolve.cpp:135-307 for one element; returns an error key or UINT64_MAX. */
static uint64_t rhs_element(const oro_solver *s, int e, double *dU) {
    const int np = s->np, nq = s->nq;
    const double *U = s->U;
    double Up[3][MAXNQ][4], Um[3][MAXNQ][4], fl[3][MAXNQ][4];
    int nb[3];
    nb[0] = s->bnd[3 * e + 2]; /* side 1: vertical */
    if (s->type[e] == 0) { nb[1] = s->bnd[3 * e]; nb[2] = s->bnd[3 * e + 1]; }
    else { nb[1] = s->bnd[3 * e + 1]; nb[2] = s->bnd[3 * e]; }
    const double *Ue = U + (size_t)e * np * 4;
    /* compute_U_plus_minus (Evolve.cpp:135-204) */
    for (int i = 0; i < nq; ++i) {
        for (int side = 0; side < 3; ++side)
            for (int k = 0; k < 4; ++k) { Up[side][i][k] = 0.0; Um[side][i][k] = 0.0; }
        for (int k = 0; k < 4; ++k)
            for (int j = 0; j < np; ++j)
                for (int side = 0; side < 3; ++side) Up[side][i][k] += s->plus[side][j][i] * Ue[j * 4 + k];
        if (s->type[e] == 0 || s->type[e] == 1)
            for (int k = 0; k < 4; ++k)
                for (int j = 0; j < np; ++j)
                    for (int side = 0; side < 3; ++side)
                        Um[side][i][k] += s->minus[side][j][i] * U[((size_t)nb[side] * np + j) * 4 + k];
    }
    /* compute_numerical_flux (Evolve.cpp:207-222) */
    uint64_t key = UINT64_MAX;
    const double *nrm = s->geo + e * GEO_STRIDE + GEO_N;
    for (int i = 0; i < nq; ++i)
        for (int side = 0; side < 3; ++side) {
            int err = oro_numerical_flux(Up[side][i], Um[side][i], nrm + 2 * side, fl[side][i]);
            if (err && key == UINT64_MAX) key = (((uint64_t)e * 16 + i) * 4 + side) * 4 + err;
        }
    /* integrate_numerical_flux (Evolve.cpp:225-251) */
    const double *len = s->len + 3 * e;
    double L[MAXNP][4], V[MAXNP][4], R[MAXNP][4];
    for (int i = 0; i < np; ++i)
        for (int k = 0; k < 4; ++k) {
            L[i][k] = 0.0;
            for (int j = 0; j < nq; ++j) {
                L[i][k] += len[0] * s->plus[0][i][j] * fl[0][j][k] * s->xw[j][1];
                L[i][k] += len[1] * s->plus[1][i][j] * fl[1][j][k] * s->xw[j][1];
                L[i][k] += len[2] * s->plus[2][i][j] * fl[2][j][k] * s->xw[j][1];
            }
        }
    /* compute_stiffness_vector (Evolve.cpp:254-272) */
    const double *Sx = s->S_phys + (size_t)e * 2 * np * np, *Sy = Sx + np * np;
    const double *Fe = s->F + (size_t)e * np * 8;
    for (int i = 0; i < np; ++i)
        for (int k = 0; k < 4; ++k) {
            V[i][k] = 0.0;
            for (int j = 0; j < np; ++j) V[i][k] += Sx[i * np + j] * Fe[j * 8 + k] + Sy[i * np + j] * Fe[j * 8 + 4 + k];
        }
    /* compute_residual_vector (Evolve.cpp:275-287) */
    for (int i = 0; i < np; ++i)
        for (int k = 0; k < 4; ++k) R[i][k] = V[i][k] - L[i][k];
    /* compute_time_derivative_U (Evolve.cpp:290-307) */
    const double *Mi = s->Minv_phys + (size_t)e * np * np;
    for (int i = 0; i < np; ++i)
        for (int k = 0; k < 4; ++k) {
            double a = 0.0;
            for (int j = 0; j < np; ++j) a += Mi[i * np + j] * R[j][k];
            dU[i * 4 + k] = a;
        }
    return key;
}

static void set_err(oro_solver *s, uint64_t key) {
    if (key == UINT64_MAX || s->err[3] != 0) return; /* keep the first failure */
    s->err[4] = s->cur_step;
    s->err[5] = s->cur_stage;
    uint64_t t = key;
    s->err[3] = (int)(t % 4); t /= 4;
    s->err[2] = (int)(t % 4); t /= 4;
    s->err[1] = (int)(t % 16); t /= 16;
    s->err[0] = (int)t;
}

/* Timestepping.cpp:18-36 */
static int evolve(oro_solver *s) {
    uint64_t best = UINT64_MAX;
    s->cur_stage++;
    const int ne = s->ne;
#ifdef _OPENMP
    int nt = g_threads > 0 ? g_threads : omp_get_max_threads();
#pragma omp parallel for schedule(static) reduction(min : best) num_threads(nt)
#endif
    for (int e = 0; e < ne; ++e) {
        uint64_t k = rhs_element(s, e, s->dU + (size_t)e * s->np * 4);
        if (k < best) best = k;
    }
    set_err(s, best);
    return best == UINT64_MAX ? 0 : (int)(best % 4);
}

int oro_evolve_elem(oro_solver *s, double *dUdt) {
    s->cur_stage = 0;
    int rc = evolve(s);
    if (dUdt) memcpy(dUdt, s->dU, sizeof(double) * s->ne * s->np * 4);
    return rc;
}

/* Timestepping.cpp:39-82 */
int oro_forward_euler(oro_solver *s, double t_step) {
    s->cur_stage = 0;
    int rc = evolve(s);
    size_t n = (size_t)s->ne * s->np;
    for (size_t i = 0; i < n; ++i) {
        for (int k = 0; k < 4; ++k) s->U[4 * i + k] += t_step * s->dU[4 * i + k];
        refresh_flux(s->U + 4 * i, s->F + 8 * i);
    }
    for (int e = 0; e < s->ne; ++e) s->time[e] += t_step;
    return rc;
}

/* Timestepping.cpp:85-305 */
int oro_rk4(oro_solver *s, double t_step) {
    size_t n = (size_t)s->ne * s->np * 4;
    double *k1 = malloc(sizeof(double) * n), *k2 = malloc(sizeof(double) * n);
    double *k3 = malloc(sizeof(double) * n), *k4 = malloc(sizeof(double) * n);
    double *U = malloc(sizeof(double) * n);
    memcpy(U, s->U, sizeof(double) * n);
    int rc = 0, r;
    size_t nn = n / 4;
    s->cur_stage = 0;
    r = evolve(s); if (!rc) rc = r;
    for (size_t i = 0; i < nn; ++i) {
        for (int k = 0; k < 4; ++k) {
            k1[4 * i + k] = t_step * s->dU[4 * i + k];
            s->U[4 * i + k] = U[4 * i + k] + k1[4 * i + k] / 2.0;
        }
        refresh_flux(s->U + 4 * i, s->F + 8 * i);
    }
    r = evolve(s); if (!rc) rc = r;
    for (size_t i = 0; i < nn; ++i) {
        for (int k = 0; k < 4; ++k) {
            k2[4 * i + k] = t_step * s->dU[4 * i + k];
            s->U[4 * i + k] = U[4 * i + k] + k2[4 * i + k] / 2.0;
        }
        refresh_flux(s->U + 4 * i, s->F + 8 * i);
    }
    r = evolve(s); if (!rc) rc = r;
    for (size_t i = 0; i < nn; ++i) {
        for (int k = 0; k < 4; ++k) {
            k3[4 * i + k] = t_step * s->dU[4 * i + k];
            s->U[4 * i + k] = U[4 * i + k] + k3[4 * i + k];
        }
        refresh_flux(s->U + 4 * i, s->F + 8 * i);
    }
    r = evolve(s); if (!rc) rc = r;
    for (size_t i = 0; i < n; ++i) k4[i] = t_step * s->dU[i];
    for (size_t i = 0; i < nn; ++i) {
        for (int k = 0; k < 4; ++k) {
            size_t m = 4 * i + k;
            s->U[m] = U[m] + (1.0 / 6.0) * (k1[m] + 2 * k2[m] + 2 * k3[m] + k4[m]);
        }
        refresh_flux(s->U + 4 * i, s->F + 8 * i);
    }
    for (int e = 0; e < s->ne; ++e) s->time[e] += t_step;
    free(k1); free(k2); free(k3); free(k4); free(U);
    return rc;
}

/* Stops after the step in which the first non-physical state occurs, where
 * the reference exits (Numericalflux.cpp:62,87,123). */
int oro_run(oro_solver *s, int method, double dt, int nsteps) {
    memset(s->err, 0, sizeof(s->err));
    for (int a = 0; a < nsteps; ++a) {
        int r;
        s->cur_step = a + 1;
        if (method == 0) r = oro_forward_euler(s, dt);
        else if (method == 1) r = oro_rk4(s, dt);
        else return ORO_ERR_ARG;
        if (r) return r;
    }
    return 0;
}

/* plottools/l2stateerror.py:31-99 with the Dunavant order-10 table of
 * plotutilities.py:51-80 (the same values as the order-10 area rule). */
void oro_l2_state_error(int ne, const double *V, const double *Ui, const double *Uf, double *err) {
    double q[25 * 3];
    oro_area_rule(10, q);
    double bq[10][25];
    for (int i = 0; i < 25; ++i) {
        double b[10];
        oro_basis(3, q[3 * i], q[3 * i + 1], b);
        for (int j = 0; j < 10; ++j) bq[j][i] = b[j];
    }
    double integral[4] = {0, 0, 0, 0};
    for (int e = 0; e < ne; ++e) {
        const double *v = V + 6 * e;
        double j00 = v[2] - v[0], j01 = v[4] - v[0], j10 = v[3] - v[1], j11 = v[5] - v[1];
        double det = j00 * j11 - j10 * j01;
        double du[10][4];
        for (int j = 0; j < 10; ++j)
            for (int k = 0; k < 4; ++k) {
                double d = Uf[(e * 10 + j) * 4 + k] - Ui[(e * 10 + j) * 4 + k];
                du[j][k] = d * d;
            }
        for (int k = 0; k < 4; ++k) {
            double sum = 0.0;
            for (int i = 0; i < 25; ++i) {
                double dq = 0.0;
                for (int j = 0; j < 10; ++j) dq += du[j][k] * bq[j][i];
                sum += q[3 * i + 2] * dq;
            }
            integral[k] += det * sum;
        }
    }
    for (int k = 0; k < 4; ++k) err[k] = sqrt((1.0 / 25.0) * integral[k]);
}
