// vt_setup.hpp -- host-side set-up for the MI355X DG solver: quadrature,
// Lagrange basis, reference operators, structured periodic mesh, isentropic
// vortex initial state and the edge tables the kernels consume.
//
// Compiled with g++ -O2 -ffp-contract=off so that every set-up quantity has
// the same bits as the reference computes on the host (expression order kept
// where the reference fixes it; files cited per function).
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

namespace vt {

constexpr int kMaxP = 3;
constexpr int kMaxNp = 10;
constexpr int kMaxNq = 10;

inline int num_nodes(int p) { return (p + 1) * (p + 2) / 2; }

// Gauss-Legendre rule on [0,1] for integration order 0..19 (Quadraturerule.cpp:13-181):
// xw = {sigma_0, w_0, sigma_1, w_1, ...}; returns nq or -1.
int line_rule(int order, std::vector<double>* xw);
// Dunavant triangle rule (Quadraturerule.cpp:183-661): {xi, eta, w} triples; returns n or -1.
int area_rule(int order, std::vector<double>* xyw);

// Equispaced reference nodes, eta outer and xi inner (Meshgeneration.cpp:188-219).
bool reference_nodes(int p, double* nodes /*[Np][2]*/);
// P0..P3 Lagrange basis and gradient (Lagrangebasis.cpp:12-147).
bool basis(int p, double xi, double eta, double* phi);
bool basis_gradient(int p, double xi, double eta, double* grad /*[Np][2]*/);

// Reference-triangle operators by area quadrature (Preevolve.cpp:15-144):
// M, its inverse (Gauss-Jordan, partial pivoting) and S[2][Np][Np] (d/dxi, d/deta).
bool reference_operators(int p, int order, double* M, double* Minv, double* S);
bool invert(int n, const double* A, double* Ainv);

// Structured periodic mesh with the reference's connectivity and float cell
// size (Meshgeneration.cpp:16-182).  bnd[e] = {right, left, vertical}.
// Returns 0, or 1 for invalid sizes (ny < 2 gives out-of-range neighbours in
// the reference's formulas and is rejected).
int generate_mesh(int nx, int ny, float domain_x, float domain_y, float perturb_grid,
                  double* vertices, int* bnd, int* type);
// Same connectivity and grid, interior vertices displaced by
// amplitude * h * U(-1,1) per component from a splitmix64 stream (seed).
int generate_mesh_seeded(int nx, int ny, float domain_x, float domain_y, double amplitude,
                         uint64_t seed, double* vertices, int* bnd, int* type);

// Neighbour across each side: side 1 -> vertical, side 2 -> right (type 0) or
// left (type 1), side 3 -> the other (Evolve.cpp:167-202).
void side_neighbours(int ne, const int* bnd, const int* type, int* nbr_side /*[ne][3]*/);

// Isentropic vortex (Element.cpp:134-160) at one point: rho, u, v, p, E, H.
void vortex_point(double x, double y, double t, double epsilon, double x0, double y0, double* out);
// The point-independent factors of vortex_point, computed by the same host
// expressions so that a device evaluation starts from the same bits:
// k = {U_infty, V_infty, pow(eps,2)(gamma-1)pow(M,2), 8 pow(pi,2),
//      eps (pow(U,2) + pow(V,2)), 2 pi rc, 1/(gamma-1), gamma/(gamma-1)}.
void vortex_constants(double epsilon, double* k /*[8]*/);
// Nodal physical coordinates of an element (Element.cpp:47-50).
void element_nodes(int p, const double* v /*[3][2]*/, double* xy /*[Np][2]*/);
// Initial state, type 0 (nodal) or 1 (least-squares) (Element.cpp:119-333),
// U in [ne][Np][4] order.  Returns false for an unknown init type.
bool initial_state(int ne, const double* vertices, int p, int order, int init_type, double epsilon,
                   double x0, double y0, double t, double* U);
// Derivative of the type 0 nodal vortex state at time t with respect to
// (epsilon, x0, y0): dU [3][ne][Np][4] (closed form of Element.cpp:134-166).
void vortex_tangent(int ne, const double* vertices, int p, double t, double epsilon, double x0, double y0,
                    double* dU);
// grad [3] = sum over elements, nodes and components of dU/d(epsilon, x0, y0)
// (vortex_tangent at time t) times lam [ne][Np][4], element by element in
// order, without forming the tangent.
void vortex_gradient(int ne, const double* vertices, int p, double t, double epsilon, double x0, double y0,
                     const double* lam, double* grad);
// Element-major [n][Np][4] state <-> structure-of-arrays planes: value k of
// node j of element e lives at planes[(k * Np + j) * ld + e0 + e].  A density
// field [n][Np] goes to Np planes (j * ld + e).  Threaded over element blocks.
void aos_to_planes(int n, int np, int ld, int e0, const double* U, double* planes);
void planes_to_aos(int n, int np, int ld, int e0, const double* planes, double* U);
void density_to_planes(int n, int np, int ld, const double* rho, double* planes);
// Nodal flux from a state, the refresh of Timestepping.cpp:59-76.
void nodal_flux(const double* U /*[4]*/, double* F /*[2][4]*/);
// nodal_flux of n nodes, U [n][4] -> F [n][2][4], on all host threads.
void nodal_flux_n(std::size_t n, const double* U, double* F);

// Edge node lists: the P+1 nodes of the reference triangle lying on each side.
// Basis functions of all other nodes vanish on that side (their restriction is
// a degree-P polynomial with P+1 zeros), so traces and lifts need only these.
const int* edge_nodes(int p, int side);

// Device operator table for one (p, integration order), packed as the kernels read it.
struct OperatorTable {
    int p = 0, np = 0, nq = 0, order = 0;
    std::vector<double> data;  // layout: see kTab* offsets in vt_device.hpp
    double max_offedge = 0;    // largest |phi| dropped by the edge-node restriction
};
bool build_operator_table(int p, int order, OperatorTable* out);
// Same from explicit operators: line rule xw [nq][2], Minv [Np][Np], S [2][Np][Np].
bool build_operator_table(int p, int nq, const double* xw, const double* Minv, const double* S, OperatorTable* out);

}  // namespace vt
