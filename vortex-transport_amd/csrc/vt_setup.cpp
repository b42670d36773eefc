// vt_setup.cpp -- host set-up (see vt_setup.hpp).  Built with
// g++ -O2 -ffp-contract=off: each expression keeps the reference's operand
// order so the set-up data (mesh, operators, initial state) has its bits.
#include "vt_setup.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "vt_device.hpp"
#include "vt_tables.h"

namespace vt {

int line_rule(int order, std::vector<double>* xw) {
    const double(*t)[2] = nullptr;
    int n = VT_line_rule(order, &t);
    if (n < 0) return -1;
    if (xw) {
        xw->resize(2 * n);
        for (int i = 0; i < n; ++i) { (*xw)[2 * i] = t[i][0]; (*xw)[2 * i + 1] = t[i][1]; }
    }
    return n;
}

int area_rule(int order, std::vector<double>* xyw) {
    const double(*t)[3] = nullptr;
    int n = VT_area_rule(order, &t);
    if (n < 0) return -1;
    if (xyw) {
        xyw->resize(3 * n);
        for (int i = 0; i < n; ++i)
            for (int c = 0; c < 3; ++c) (*xyw)[3 * i + c] = t[i][c];
    }
    return n;
}

// Meshgeneration.cpp:188-219: spacing 1.0/p, eta outer, xi inner, i + j <= p.
bool reference_nodes(int p, double* nodes) {
    if (p < 0 || p > kMaxP) return false;
    if (p == 0) { nodes[0] = 0.0; nodes[1] = 0.0; return true; }
    const double h = 1.0 / p;
    int c = 0;
    for (int j = 0; j <= p; ++j)
        for (int i = 0; i + j <= p; ++i) { nodes[2 * c] = h * i; nodes[2 * c + 1] = h * j; ++c; }
    return true;
}

// Lagrangebasis.cpp:37-63.  The monomial expressions (and their evaluation
// order) are those of the reference so that edge tables and operators match
// its values to the last bit.
bool basis(int p, double x, double y, double* f) {
    switch (p) {
    case 0: f[0] = 1.0; return true;
    case 1: f[0] = 1.0 - x - y; f[1] = x; f[2] = y; return true;
    case 2:
        f[0] = 1.0 - 3.0*x - 3.0*y + 2.0*x*x + 4.0*x*y + 2.0*y*y;
        f[1] = 4.0*x - 4.0*x*x - 4.0*x*y;
        f[2] = -x + 2.0*x*x;
        f[3] = 4.0*y - 4.0*x*y - 4.0*y*y;
        f[4] = 4.0*x*y;
        f[5] = -y + 2.0*y*y;
        return true;
    case 3:
        f[0] = 1.0 - 11.0/2.0*x - 11.0/2.0*y + 9.0*x*x + 18.0*x*y + 9.0*y*y - 9.0/2.0*x*x*x
               - 27.0/2.0*x*x*y - 27.0/2.0*x*y*y - 9.0/2.0*y*y*y;
        f[1] = 9.0*x - 45.0/2.0*x*x - 45.0/2.0*x*y + 27.0/2.0*x*x*x + 27.0*x*x*y + 27.0/2.0*x*y*y;
        f[2] = -9.0/2.0*x + 18.0*x*x + 9.0/2.0*x*y - 27.0/2.0*x*x*x - 27.0/2.0*x*x*y;
        f[3] = x - 9.0/2.0*x*x + 9.0/2.0*x*x*x;
        f[4] = 9.0*y - 45.0/2.0*x*y - 45.0/2.0*y*y + 27.0/2.0*x*x*y + 27.0*x*y*y + 27.0/2.0*y*y*y;
        f[5] = 27.0*x*y - 27.0*x*x*y - 27.0*x*y*y;
        f[6] = -9.0/2.0*x*y + 27.0/2.0*x*x*y;
        f[7] = -9.0/2.0*y + 9.0/2.0*x*y + 18.0*y*y - 27.0/2.0*x*y*y - 27.0/2.0*y*y*y;
        f[8] = -9.0/2.0*x*y + 27.0/2.0*x*y*y;
        f[9] = y - 9.0/2.0*y*y + 9.0/2.0*y*y*y;
        return true;
    default: return false;
    }
}

// Lagrangebasis.cpp:93-149 ({d/dxi, d/deta} per node).
bool basis_gradient(int p, double x, double y, double* g) {
    double* gx = g;
    auto set = [&](int i, double a, double b) { gx[2 * i] = a; gx[2 * i + 1] = b; };
    switch (p) {
    case 0: set(0, 0.0, 0.0); return true;
    case 1: set(0, -1.0, -1.0); set(1, 1.0, 0.0); set(2, 0.0, 1.0); return true;
    case 2:
        set(0, -3.0 + 4.0*x + 4.0*y, -3.0 + 4.0*x + 4.0*y);
        set(1, 4.0 - 8.0*x - 4.0*y, -4.0*x);
        set(2, -1.0 + 4.0*x, 0.0);
        set(3, -4.0 * y, 4.0 - 4.0*x - 8.0*y);
        set(4, 4.0 * y, 4.0*x);
        set(5, 0.0, -1.0 + 4.0*y);
        return true;
    case 3:
        set(0, -11.0/2.0 + 18.0*x + 18.0*y - 27.0/2.0*x*x - 27.0*x*y - 27.0/2.0*y*y,
               -11.0/2.0 + 18.0*x + 18.0*y - 27.0/2.0*x*x - 27.0*x*y - 27.0/2.0*y*y);
        set(1, 9.0 - 45.0*x - 45.0/2.0*y + 81.0/2.0*x*x + 54.0*x*y + 27.0/2.0*y*y,
               -45.0/2.0*x + 27.0*x*x + 27.0*x*y);
        set(2, -9.0/2.0 + 36.0*x + 9.0/2.0*y - 81.0/2.0*x*x - 27.0*x*y, 9.0/2.0*x - 27.0/2.0*x*x);
        set(3, 1.0 - 9.0*x + 27.0/2.0*x*x, 0.0);
        set(4, -45.0/2.0*y + 27.0*x*y + 27.0*y*y,
               9.0 - 45.0/2.0*x - 45.0*y + 27.0/2.0*x*x + 54.0*x*y + 81.0/2.0*y*y);
        set(5, 27.0*y - 54.0*x*y - 27.0*y*y, 27.0*x - 27.0*x*x - 54.0*x*y);
        set(6, -9.0/2.0*y + 27.0*x*y, -9.0/2.0*x + 27.0/2.0*x*x);
        set(7, 9.0/2.0*y - 27.0/2.0*y*y, -9.0/2.0 + 9.0/2.0*x + 36.0*y - 27.0*x*y - 81.0/2.0*y*y);
        set(8, -9.0/2.0*y + 27.0/2.0*y*y, -9.0/2.0*x + 27.0*x*y);
        set(9, 0.0, 1.0 - 9.0*y + 27.0/2.0*y*y);
        return true;
    default: return false;
    }
}

bool invert(int n, const double* A, double* Ainv) {
    std::vector<double> a(static_cast<size_t>(n) * 2 * n);
    const int w = 2 * n;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < w; ++j) a[i * w + j] = j < n ? A[i * n + j] : (j - n == i ? 1.0 : 0.0);
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(a[r * w + c]) > std::fabs(a[piv * w + c])) piv = r;
        if (a[piv * w + c] == 0.0) return false;
        if (piv != c)
            for (int j = 0; j < w; ++j) std::swap(a[c * w + j], a[piv * w + j]);
        const double d = a[c * w + c];
        for (int j = 0; j < w; ++j) a[c * w + j] /= d;
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            const double f = a[r * w + c];
            if (f == 0.0) continue;
            for (int j = 0; j < w; ++j) a[r * w + j] -= f * a[c * w + j];
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) Ainv[i * n + j] = a[i * w + n + j];
    return true;
}

// Preevolve.cpp:43-54 (mass) and :128-137 (stiffness): products taken as
// (a * b) * w and summed from zero in quadrature order.
bool reference_operators(int p, int order, double* M, double* Minv, double* S) {
    if (p < 0 || p > kMaxP) return false;
    std::vector<double> q;
    const int n = area_rule(order, &q);
    if (n < 0) return false;
    const int np = num_nodes(p);
    std::vector<double> phi(np * n), gr(2 * np * n);
    for (int k = 0; k < n; ++k) {
        double b[kMaxNp], g[2 * kMaxNp];
        basis(p, q[3 * k], q[3 * k + 1], b);
        basis_gradient(p, q[3 * k], q[3 * k + 1], g);
        for (int j = 0; j < np; ++j) {
            phi[j * n + k] = b[j];
            gr[(j * n + k) * 2] = g[2 * j];
            gr[(j * n + k) * 2 + 1] = g[2 * j + 1];
        }
    }
    std::vector<double> m(np * np);
    for (int i = 0; i < np; ++i)
        for (int j = 0; j < np; ++j) {
            double a = 0, s0 = 0, s1 = 0;
            for (int k = 0; k < n; ++k) {
                a += phi[i * n + k] * phi[j * n + k] * q[3 * k + 2];
                s0 += gr[(i * n + k) * 2] * phi[j * n + k] * q[3 * k + 2];
                s1 += gr[(i * n + k) * 2 + 1] * phi[j * n + k] * q[3 * k + 2];
            }
            m[i * np + j] = a;
            if (S) { S[i * np + j] = s0; S[np * np + i * np + j] = s1; }
        }
    if (M) std::memcpy(M, m.data(), sizeof(double) * np * np);
    if (Minv && !invert(np, m.data(), Minv)) return false;
    return true;
}

// Meshgeneration.cpp:16-182.  Cell size is float/int in float, the origin
// (-1 * L) / 2 in float, both widened to double as the reference does.
static int mesh_common(int nx, int ny, const double* px, const double* py, double* V, int* bnd, int* type) {
    const int ne = 2 * nx * ny;
    for (int i = 0; i < ne; ++i) {
        // Triangle 2c (type 0): LL, LR, UL; triangle 2c+1 (type 1): UR, UL, LR.
        const int cell = i / 2, r = cell / nx + 1, j = cell % nx;
        int n1 = (i % 2 == 0) ? j + (nx + 1) * (r - 1) : (j + (nx + 1) * r) + 1;
        int n2 = n1 + (i % 2 == 0 ? 1 : -1);
        // node 3 is node 2 of the partner triangle in the same cell
        int pn1 = (i % 2 == 0) ? (j + (nx + 1) * r) + 1 : j + (nx + 1) * (r - 1);
        int n3 = pn1 + (i % 2 == 0 ? -1 : 1);
        type[i] = i % 2;
        const int left = (i % (2 * nx) == 0 ? i - 1 + 2 * nx : i - 1);
        const int right = ((i + 1) % (2 * nx) == 0 ? i + 1 - 2 * nx : i + 1);
        int vert;
        if (i >= 2 * nx * (ny - 1)) vert = (i % 2 == 0 ? i + 1 - 2 * nx : i - 1 - 2 * nx * (ny - 1));
        else if (i >= 2 * nx) vert = (i % 2 == 0 ? i + 1 - 2 * nx : i - 1 + 2 * nx);
        else vert = (i % 2 == 0 ? i + 1 + 2 * nx * (ny - 1) : i - 1 + 2 * nx);
        bnd[3 * i] = right; bnd[3 * i + 1] = left; bnd[3 * i + 2] = vert;
        V[6 * i + 0] = px[n1]; V[6 * i + 1] = py[n1];
        V[6 * i + 2] = px[n2]; V[6 * i + 3] = py[n2];
        V[6 * i + 4] = px[n3]; V[6 * i + 5] = py[n3];
    }
    return 0;
}

int generate_mesh(int nx, int ny, float domain_x, float domain_y, float perturb_grid, double* V, int* bnd, int* type) {
    if (nx < 1 || ny < 2) return 1;
    std::vector<double> gx(nx + 1), gy(ny + 1);
    const double hx = domain_x / nx, hy = domain_y / ny;
    for (int i = 0; i < nx + 1; ++i) gx[i] = -1 * domain_x / 2 + hx * i;
    for (int i = 0; i < ny + 1; ++i) gy[i] = -1 * domain_y / 2 + hy * i;
    std::vector<double> px((nx + 1) * (ny + 1)), py((nx + 1) * (ny + 1));
    int c = 0;
    for (int i = 0; i < ny + 1; ++i)
        for (int j = 0; j < nx + 1; ++j, ++c) {
            const int interior = (i == 0 ? 0 : 1) * (i == ny ? 0 : 1) * (j == 0 ? 0 : 1) * (j == nx ? 0 : 1);
            // Meshgeneration.cpp:47-48: the sin/cos(pow()) factor is always
            // evaluated, and is NaN once a half-width exceeds 6.
            px[c] = gx[j] + interior * perturb_grid * 0.2 * hx * std::sin(std::pow(gx[j] + 6, gy[i] + 6));
            py[c] = gy[i] + interior * perturb_grid * 0.2 * hy * std::cos(std::pow(gy[i] + 6, gx[j] + 6));
        }
    return mesh_common(nx, ny, px.data(), py.data(), V, bnd, type);
}

static inline uint64_t splitmix64(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

int generate_mesh_seeded(int nx, int ny, float domain_x, float domain_y, double amplitude, uint64_t seed,
                         double* V, int* bnd, int* type) {
    if (nx < 1 || ny < 2) return 1;
    std::vector<double> gx(nx + 1), gy(ny + 1);
    const double hx = domain_x / nx, hy = domain_y / ny;
    for (int i = 0; i < nx + 1; ++i) gx[i] = -1 * domain_x / 2 + hx * i;
    for (int i = 0; i < ny + 1; ++i) gy[i] = -1 * domain_y / 2 + hy * i;
    std::vector<double> px((nx + 1) * (ny + 1)), py((nx + 1) * (ny + 1));
    uint64_t st = seed;
    int c = 0;
    for (int i = 0; i < ny + 1; ++i)
        for (int j = 0; j < nx + 1; ++j, ++c) {
            const double ux = 2.0 * (double)(splitmix64(&st) >> 11) * 0x1.0p-53 - 1.0;
            const double uy = 2.0 * (double)(splitmix64(&st) >> 11) * 0x1.0p-53 - 1.0;
            const bool interior = i > 0 && i < ny && j > 0 && j < nx;
            px[c] = gx[j] + (interior ? amplitude * hx * ux : 0.0);
            py[c] = gy[i] + (interior ? amplitude * hy * uy : 0.0);
        }
    return mesh_common(nx, ny, px.data(), py.data(), V, bnd, type);
}

void side_neighbours(int ne, const int* bnd, const int* type, int* ns) {
    for (int e = 0; e < ne; ++e) {
        ns[3 * e] = bnd[3 * e + 2];
        ns[3 * e + 1] = type[e] == 0 ? bnd[3 * e] : bnd[3 * e + 1];
        ns[3 * e + 2] = type[e] == 0 ? bnd[3 * e + 1] : bnd[3 * e];
    }
}

// Element.cpp:134-160, constants as the reference sets them.
void vortex_point(double x, double y, double t, double epsilon, double x0, double y0, double* out) {
    const double rho_infty = 1.0, rc = 1.0, gamma = 1.4, M_infty = 0.5, p_infty = 20.0 / 7.0;
    const double U_infty = 1.0 / std::pow(2.0, 0.5), V_infty = 1.0 / std::pow(2.0, 0.5);
    const double f0 = 1.0 - (std::pow(x - x0 - U_infty * t, 2.0) + std::pow(y - y0 - V_infty * t, 2.0)) / std::pow(rc, 2.0);
    const double f1 = 1.0 - std::pow(epsilon, 2.0) * (gamma - 1) * std::pow(M_infty, 2.0) * std::exp(f0) / (8.0 * std::pow(M_PI, 2.0));
    const double f2 = epsilon * (std::pow(U_infty, 2.0) + std::pow(V_infty, 2.0)) * std::exp(f0 / 2.0) / (2.0 * M_PI * rc);
    const double rho = rho_infty * std::pow(f1, 1.0 / (gamma - 1.0));
    const double u = U_infty - f2 * (y - y0 - V_infty * t);
    const double v = V_infty + f2 * (x - x0 - U_infty * t);
    const double p = p_infty * std::pow(f1, gamma / (gamma - 1.0));
    const double E = p / (rho * (gamma - 1.0)) + (std::pow(u, 2.0) + std::pow(v, 2.0)) / 2.0;
    out[0] = rho; out[1] = u; out[2] = v; out[3] = p; out[4] = E; out[5] = E + p / rho;
}

void vortex_constants(double epsilon, double* k) {
    const double rc = 1.0, gamma = 1.4, M_infty = 0.5;
    const double U_infty = 1.0 / std::pow(2.0, 0.5), V_infty = 1.0 / std::pow(2.0, 0.5);
    k[0] = U_infty;
    k[1] = V_infty;
    k[2] = std::pow(epsilon, 2.0) * (gamma - 1) * std::pow(M_infty, 2.0);
    k[3] = 8.0 * std::pow(M_PI, 2.0);
    k[4] = epsilon * (std::pow(U_infty, 2.0) + std::pow(V_infty, 2.0));
    k[5] = 2.0 * M_PI * rc;
    k[6] = 1.0 / (gamma - 1.0);
    k[7] = gamma / (gamma - 1.0);
}

// Lagrangebasis.cpp:150-165 applied to each reference node (Element.cpp:47-50).
void element_nodes(int p, const double* v, double* xy) {
    double nodes[2 * kMaxNp];
    reference_nodes(p, nodes);
    for (int i = 0; i < num_nodes(p); ++i) {
        const double xi = nodes[2 * i], eta = nodes[2 * i + 1];
        const double f0 = 1.0 - xi - eta, f1 = xi, f2 = eta;
        xy[2 * i] = v[0] * f0 + v[2] * f1 + v[4] * f2;
        xy[2 * i + 1] = v[1] * f0 + v[3] * f1 + v[5] * f2;
    }
}

void nodal_flux(const double* U, double* F) {
    const double gamma = 1.4;
    const double rho = U[0], u = U[1] / rho, v = U[2] / rho, E = U[3] / rho;
    const double p = rho * (gamma - 1) * (E - (std::pow(u, 2) + std::pow(v, 2)) / 2);
    const double H = E + p / rho;
    F[0] = rho * u; F[1] = rho * std::pow(u, 2) + p; F[2] = rho * u * v; F[3] = rho * u * H;
    F[4] = rho * v; F[5] = rho * u * v; F[6] = rho * std::pow(v, 2) + p; F[7] = rho * v * H;
}

void nodal_flux_n(size_t n, const double* U, double* F) {
    const long long nn = static_cast<long long>(n);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < nn; ++i) nodal_flux(U + 4 * i, F + 8 * i);
}

bool initial_state(int ne, const double* V, int p, int order, int init_type, double epsilon, double x0, double y0,
                   double t, double* U) {
    const int np = num_nodes(p);
    if (init_type == 0) {
#pragma omp parallel for schedule(static)
        for (int e = 0; e < ne; ++e) {
            double xy[2 * kMaxNp];
            element_nodes(p, V + 6 * e, xy);
            for (int i = 0; i < np; ++i) {
                double w[6];
                vortex_point(xy[2 * i], xy[2 * i + 1], t, epsilon, x0, y0, w);
                double* u = U + (static_cast<size_t>(e) * np + i) * 4;
                u[0] = w[0]; u[1] = w[0] * w[1]; u[2] = w[0] * w[2]; u[3] = w[0] * w[4];
            }
        }
        return true;
    }
    if (init_type != 1) return false;
    // Least-squares projection (Element.cpp:180-304): U = M_phys^-1 b with
    // b_i = sum_q detJ phi_i(q) U_exact(q) w_q and M_phys^-1 = M_ref^-1 / detJ.
    std::vector<double> q, Minv(np * np);
    const int nqa = area_rule(order, &q);
    if (nqa < 0 || !reference_operators(p, order, nullptr, Minv.data(), nullptr)) return false;
    std::vector<double> phi(np * nqa);
    for (int k = 0; k < nqa; ++k) {
        double b[kMaxNp];
        basis(p, q[3 * k], q[3 * k + 1], b);
        for (int j = 0; j < np; ++j) phi[j * nqa + k] = b[j];
    }
#pragma omp parallel for schedule(static)
    for (int e = 0; e < ne; ++e) {
        const double* v = V + 6 * e;
        const double detJ = (v[2] - v[0]) * (v[5] - v[1]) - (v[3] - v[1]) * (v[4] - v[0]);
        std::vector<double> ue(4 * nqa);
        for (int k = 0; k < nqa; ++k) {
            const double xi = q[3 * k], eta = q[3 * k + 1];
            const double f0 = 1.0 - xi - eta;
            const double x = v[0] * f0 + v[2] * xi + v[4] * eta, y = v[1] * f0 + v[3] * xi + v[5] * eta;
            double w[6];
            vortex_point(x, y, t, epsilon, x0, y0, w);
            ue[4 * k] = w[0]; ue[4 * k + 1] = w[0] * w[1]; ue[4 * k + 2] = w[0] * w[2]; ue[4 * k + 3] = w[0] * w[4];
        }
        double bv[kMaxNp][4];
        for (int i = 0; i < np; ++i)
            for (int c = 0; c < 4; ++c) {
                double s = 0.0;
                for (int k = 0; k < nqa; ++k) s += detJ * phi[i * nqa + k] * ue[4 * k + c] * q[3 * k + 2];
                bv[i][c] = s;
            }
        for (int i = 0; i < np; ++i)
            for (int c = 0; c < 4; ++c) {
                double s = 0.0;
                for (int j = 0; j < np; ++j) s += Minv[i * np + j] / detJ * bv[j][c];
                U[(static_cast<size_t>(e) * np + i) * 4 + c] = s;
            }
    }
    return true;
}

const int* edge_nodes(int p, int side) {
    switch (p) {
    case 0: return kEdge0[side];
    case 1: return kEdge1[side];
    case 2: return kEdge2[side];
    case 3: return kEdge3[side];
    default: return nullptr;
    }
}

// Edge tables (Evolve.cpp:93-131): plus points (sigma,0), (1-sigma,sigma),
// (0,1-sigma); minus points (1-sigma,0), (sigma,1-sigma), (0,sigma).
bool build_operator_table(int p, int nq, const double* xw, const double* Minv, const double* S, OperatorTable* out) {
    if (p < 0 || p > kMaxP || nq < 1 || nq > kMaxNq) return false;
    const int np = num_nodes(p), ne1 = p + 1;
    OperatorTable& t = *out;
    t.p = p; t.np = np; t.nq = nq;
    t.data.assign(tab_size(p, nq), 0.0);
    std::memcpy(t.data.data() + tab_minv(p), Minv, sizeof(double) * np * np);
    for (int i = 0; i < np; ++i)
        for (int j = 0; j < np; ++j) {
            t.data[tab_minvt(p) + j * np + i] = Minv[i * np + j];
            t.data[tab_s(p) + (j * np + i) * 2] = S[i * np + j];
            t.data[tab_s(p) + (j * np + i) * 2 + 1] = S[np * np + i * np + j];
        }
    t.max_offedge = 0;
    for (int q = 0; q < nq; ++q) {
        const double sg = xw[2 * q], w = xw[2 * q + 1];
        const double pts[3][2][2] = {{{sg, 0.0}, {1.0 - sg, 0.0}},
                                     {{1.0 - sg, sg}, {sg, 1.0 - sg}},
                                     {{0.0, 1.0 - sg}, {0.0, sg}}};
        for (int s = 0; s < 3; ++s) {
            double bp[kMaxNp], bm[kMaxNp];
            basis(p, pts[s][0][0], pts[s][0][1], bp);
            basis(p, pts[s][1][0], pts[s][1][1], bm);
            const int* en = edge_nodes(p, s);
            for (int m = 0; m < ne1; ++m) {
                t.data[tab_tp(p) + (s * nq + q) * ne1 + m] = bp[en[m]];
                t.data[tab_tm(p, nq) + (s * nq + q) * ne1 + m] = bm[en[m]];
                t.data[tab_lift(p, nq) + (s * nq + q) * ne1 + m] = bp[en[m]] * w;
            }
            for (int j = 0; j < np; ++j) {
                bool on = false;
                for (int m = 0; m < ne1; ++m) on = on || en[m] == j;
                if (!on) t.max_offedge = std::fmax(t.max_offedge, std::fmax(std::fabs(bp[j]), std::fabs(bm[j])));
            }
        }
    }
    return true;
}

bool build_operator_table(int p, int order, OperatorTable* out) {
    std::vector<double> xw;
    const int nq = line_rule(order, &xw);
    if (p < 0 || p > kMaxP || nq < 0 || area_rule(order, nullptr) < 0) return false;
    const int np = num_nodes(p);
    std::vector<double> Minv(np * np), S(2 * np * np);
    if (!reference_operators(p, order, nullptr, Minv.data(), S.data())) return false;
    if (!build_operator_table(p, nq, xw.data(), Minv.data(), S.data(), out)) return false;
    out->order = order;
    return true;
}


void vortex_tangent(int ne, const double* V, int p, double t, double epsilon, double x0, double y0, double* dU) {
    const int np = num_nodes(p);
    const double gamma = 1.4, M_infty = 0.5, p_infty = 20.0 / 7.0, rc = 1.0;
    const double U_infty = 1.0 / std::sqrt(2.0), V_infty = 1.0 / std::sqrt(2.0);
    const double A = (gamma - 1) * M_infty * M_infty / (8.0 * M_PI * M_PI);
    const double B = (U_infty * U_infty + V_infty * V_infty) / (2.0 * M_PI * rc);
    const size_t plane = static_cast<size_t>(ne) * np * 4;
#pragma omp parallel for schedule(static)
    for (int e = 0; e < ne; ++e) {
        double xy[2 * kMaxNp];
        element_nodes(p, V + 6 * e, xy);
        for (int i = 0; i < np; ++i) {
            const double dx = xy[2 * i] - x0 - U_infty * t, dy = xy[2 * i + 1] - y0 - V_infty * t;
            const double f0 = 1.0 - (dx * dx + dy * dy) / (rc * rc);
            const double ex = std::exp(f0), eh = std::exp(0.5 * f0);
            const double f1 = 1.0 - epsilon * epsilon * A * ex;
            const double f2 = epsilon * B * eh;
            const double rho = std::pow(f1, 1.0 / (gamma - 1.0)), pr = p_infty * std::pow(f1, gamma / (gamma - 1.0));
            const double u = U_infty - f2 * dy, v = V_infty + f2 * dx;
            // d f0 / d(eps, x0, y0) = (0, 2 dx, 2 dy); d dx/dx0 = -1, d dy/dy0 = -1
            const double df0[3] = {0.0, 2.0 * dx / (rc * rc), 2.0 * dy / (rc * rc)};
            for (int a = 0; a < 3; ++a) {
                const double df1 = -(a == 0 ? 2.0 * epsilon : 0.0) * A * ex - epsilon * epsilon * A * ex * df0[a];
                const double df2 = (a == 0 ? B * eh : 0.0) + f2 * 0.5 * df0[a];
                const double drho = rho / ((gamma - 1.0) * f1) * df1;
                const double dp = pr * gamma / ((gamma - 1.0) * f1) * df1;
                const double ddx = a == 1 ? -1.0 : 0.0, ddy = a == 2 ? -1.0 : 0.0;
                const double du = -df2 * dy - f2 * ddy, dv = df2 * dx + f2 * ddx;
                double* o = dU + a * plane + (static_cast<size_t>(e) * np + i) * 4;
                o[0] = drho;
                o[1] = drho * u + rho * du;
                o[2] = drho * v + rho * dv;
                o[3] = dp / (gamma - 1.0) + 0.5 * drho * (u * u + v * v) + rho * (u * du + v * dv);
            }
        }
    }
}

void vortex_gradient(int ne, const double* V, int p, double t, double epsilon, double x0, double y0,
                     const double* lam, double* grad) {
    // Chunks of elements in parallel, each summed in element order; the chunk
    // sums are then added in chunk order, so the result does not depend on
    // the thread count.
    constexpr int kChunk = 1024;
    const int np = num_nodes(p);
    const int nchunk = (ne + kChunk - 1) / kChunk;
    std::vector<double> part(3 * static_cast<size_t>(nchunk), 0.0);
#pragma omp parallel
    {
        std::vector<double> dU(static_cast<size_t>(3) * kChunk * np * 4);
#pragma omp for schedule(static)
        for (int c = 0; c < nchunk; ++c) {
            const int e0 = c * kChunk, n = std::min(kChunk, ne - e0);
            const size_t plane = static_cast<size_t>(n) * np * 4;
            vortex_tangent(n, V + 6 * static_cast<size_t>(e0), p, t, epsilon, x0, y0, dU.data());
            const double* l = lam + static_cast<size_t>(e0) * np * 4;
            for (int a = 0; a < 3; ++a) {
                double g = 0.0;
                for (size_t i = 0; i < plane; ++i) g += dU[a * plane + i] * l[i];
                part[3 * static_cast<size_t>(c) + a] = g;
            }
        }
    }
    double g[3] = {0.0, 0.0, 0.0};
    for (int c = 0; c < nchunk; ++c)
        for (int a = 0; a < 3; ++a) g[a] += part[3 * static_cast<size_t>(c) + a];
    for (int a = 0; a < 3; ++a) grad[a] = g[a];
}

// Blocks of 512 elements: one block's element-major slice (160 KB at p = 3)
// stays in cache while its 4 Np plane segments are written or read.
constexpr int kConvBlock = 512;

void aos_to_planes(int n, int np, int ld, int e0, const double* U, double* planes) {
    const int nb = (n + kConvBlock - 1) / kConvBlock;
#pragma omp parallel for schedule(static)
    for (int b = 0; b < nb; ++b) {
        const int lo = b * kConvBlock, hi = std::min(n, lo + kConvBlock);
        for (int k = 0; k < 4; ++k)
            for (int j = 0; j < np; ++j) {
                double* dst = planes + static_cast<size_t>(k * np + j) * ld + e0;
                for (int e = lo; e < hi; ++e) dst[e] = U[(static_cast<size_t>(e) * np + j) * 4 + k];
            }
    }
}

void planes_to_aos(int n, int np, int ld, int e0, const double* planes, double* U) {
    const int nb = (n + kConvBlock - 1) / kConvBlock;
#pragma omp parallel for schedule(static)
    for (int b = 0; b < nb; ++b) {
        const int lo = b * kConvBlock, hi = std::min(n, lo + kConvBlock);
        for (int k = 0; k < 4; ++k)
            for (int j = 0; j < np; ++j) {
                const double* src = planes + static_cast<size_t>(k * np + j) * ld + e0;
                for (int e = lo; e < hi; ++e) U[(static_cast<size_t>(e) * np + j) * 4 + k] = src[e];
            }
    }
}

void density_to_planes(int n, int np, int ld, const double* rho, double* planes) {
    const int nb = (n + kConvBlock - 1) / kConvBlock;
#pragma omp parallel for schedule(static)
    for (int b = 0; b < nb; ++b) {
        const int lo = b * kConvBlock, hi = std::min(n, lo + kConvBlock);
        for (int j = 0; j < np; ++j)
            for (int e = lo; e < hi; ++e) planes[static_cast<size_t>(j) * ld + e] = rho[static_cast<size_t>(e) * np + j];
    }
}

}  // namespace vt
