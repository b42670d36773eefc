// vt_dropin.cpp -- the reference's Parameters / mesh / Element /
// Evolve_element / Timestepping call surface (same names, argument meaning
// and exit behaviour) implemented over the C ABI of libvt_hip.so.
//
// Set-up (parameters, mesh, quadrature, basis, reference operators, element
// geometry, initial state, edge tables) runs on the host with the reference's
// arithmetic; rk4 / forward_euler / evolve_elem / numerical_flux run on the
// GPU.  There is no CPU evaluation of the right-hand side in this file.
#include <sys/stat.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "Element.H"
#include "Evolve.H"
#include "Lagrangebasis.H"
#include "Meshgeneration.H"
#include "Numericalflux.H"
#include "Parameters.H"
#include "Preevolve.H"
#include "Quadraturerule.H"
#include "Timestepping.H"
#include "Utilities.H"
#include "vt_api.h"
#include "vt_setup.hpp"

namespace {

int num_nodes(int p) { return (p + 1) * (p + 2) / 2; }

[[noreturn]] void die_hip(const char* where, int rc) {
    std::cerr << "GPU error in " << where << ": " << vt_status_string(rc) << " " << vt_last_hip_error() << std::endl;
    std::exit(EXIT_FAILURE);
}

// Reproduces the reference's exit on a non-physical state (Numericalflux.cpp:51-124).
[[noreturn]] void die_nonphysical(const int* info) {
    std::cerr << info[3] << ". Error: Non-physical state! " << std::endl;
    std::cout << " element " << info[0] << ", side " << info[2] + 1 << ", quadrature point " << info[1]
              << ", step " << info[4] << ", stage " << info[5] << std::endl;
    std::exit(EXIT_FAILURE);
}

// The reference operators handed to Element::build_mass_matrix_inverse /
// build_stiffness_matrix, kept per element array (the array starts `number`
// elements before an element; main.cpp:115 allocates it with new[]), so the
// classes keep the reference's object layout.
struct ArrayOperators {
    std::vector<double> minv, stiff;  // [Np][Np], [2][Np][Np]
};
std::mutex g_ops_mu;
std::map<const Element*, ArrayOperators> g_ops;

void record_operators(const Element* e, const std::vector<double>* minv, const std::vector<double>* stiff) {
    std::lock_guard<std::mutex> lock(g_ops_mu);
    ArrayOperators& o = g_ops[e - e->number];
    if (minv && o.minv.empty()) o.minv = *minv;
    if (stiff && o.stiff.empty()) o.stiff = *stiff;
}

// One GPU context per element array, created on first use from what the
// elements hold (vertices, neighbours, type, the reference operators handed
// to build_*), and the line rule of the Evolve_element array.
struct Binding {
    vt_ctx* ctx = nullptr;
    int n = 0, np = 0;
};
std::unordered_map<const Element*, Binding> g_bind;

Binding& bind(Element* el, Evolve_element* ev, int n) {
    auto it = g_bind.find(el);
    if (it != g_bind.end() && it->second.n == n) return it->second;
    if (it != g_bind.end()) { vt_destroy(it->second.ctx); g_bind.erase(it); }
    const int p = el[0].p, np = num_nodes(p);
    std::vector<double> v(6 * static_cast<size_t>(n));
    std::vector<int> nb(3 * static_cast<size_t>(n)), ty(n);
    for (int i = 0; i < n; ++i) {
        for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 2; ++c) v[6 * static_cast<size_t>(i) + 2 * a + c] = el[i].vertices_coords_phys_space[a][c];
        nb[3 * static_cast<size_t>(i)] = el[i].right_element;
        nb[3 * static_cast<size_t>(i) + 1] = el[i].left_element;
        nb[3 * static_cast<size_t>(i) + 2] = el[i].vertical_element;
        ty[i] = el[i].type;
    }
    ArrayOperators ops;
    {
        std::lock_guard<std::mutex> lock(g_ops_mu);
        auto o = g_ops.find(el);
        if (o != g_ops.end()) ops = o->second;
    }
    if (ops.minv.size() != static_cast<size_t>(np * np) || ops.stiff.size() != static_cast<size_t>(2 * np * np)) {
        std::cerr << "Element::build_mass_matrix_inverse / build_stiffness_matrix must run before time stepping" << std::endl;
        std::exit(EXIT_FAILURE);
    }
    const int nq = static_cast<int>(ev[0].gau_integ_line.size());
    std::vector<double> xw(2 * static_cast<size_t>(nq));
    for (int q = 0; q < nq; ++q) { xw[2 * q] = ev[0].gau_integ_line[q][0]; xw[2 * q + 1] = ev[0].gau_integ_line[q][1]; }
    Binding b;
    int rc = vt_create_with_operators(n, v.data(), nb.data(), ty.data(), p, nq, xw.data(), ops.minv.data(), ops.stiff.data(),
                                      &b.ctx);
    if (rc) die_hip("vt_create_with_operators", rc);
    b.n = n;
    b.np = np;
    return g_bind[el] = b;
}

void upload(Binding& b, const Element* el) {
    std::vector<double> U(static_cast<size_t>(b.n) * b.np * 4);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < b.n; ++i)
        for (int j = 0; j < b.np; ++j)
            for (int k = 0; k < 4; ++k) U[(static_cast<size_t>(i) * b.np + j) * 4 + k] = el[i].hidrodynamics_vector_U[j][k];
    int rc = vt_set_state(b.ctx, U.data());
    if (rc) die_hip("vt_set_state", rc);
}

// Copies the device state into the elements and refreshes the nodal flux the
// way the reference does after every update (Timestepping.cpp:59-76).
void download(Binding& b, Element* el) {
    std::vector<double> U(static_cast<size_t>(b.n) * b.np * 4);
    int rc = vt_get_state(b.ctx, U.data());
    if (rc) die_hip("vt_get_state", rc);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < b.n; ++i)
        for (int j = 0; j < b.np; ++j) {
            const double* u = &U[(static_cast<size_t>(i) * b.np + j) * 4];
            double F[8];
            vt::nodal_flux(u, F);
            for (int k = 0; k < 4; ++k) {
                el[i].hidrodynamics_vector_U[j][k] = u[k];
                el[i].hidrodynamics_vector_F[j][0][k] = F[k];
                el[i].hidrodynamics_vector_F[j][1][k] = F[4 + k];
            }
        }
}

void step_on_gpu(Element* el, Evolve_element* ev, int method, double dt, int nsteps, int n) {
    Binding& b = bind(el, ev, n);
    upload(b, el);
    int rc = vt_step(b.ctx, method, dt, nsteps);
    if (rc == VT_ERR_NONPHYSICAL) {
        int info[6];
        vt_last_error(b.ctx, info);
        die_nonphysical(info);
    }
    if (rc) die_hip("vt_step", rc);
    download(b, el);
    for (int i = 0; i < n; ++i)
        for (int s = 0; s < nsteps; ++s) el[i].time += dt;
}

}  // namespace

// ---------------------------------------------------------------------------
// Parameters.cpp:12-64
// ---------------------------------------------------------------------------
parameters read_input_files(int argc, char* argv[]) {
    if (argc != 2) {
        std::cerr << "Usage: " << argv[0] << " <input_file>" << std::endl;
        exit(EXIT_FAILURE);
    }
    std::ifstream in(argv[1]);
    if (!in.is_open()) {
        std::cerr << "Error: Unable to open input file '" << argv[1] << "'." << std::endl;
        exit(EXIT_FAILURE);
    }
    std::unordered_map<std::string, std::string> kv;
    std::string line;
    while (std::getline(in, line)) {
        const size_t pos = line.find('=');
        if (pos != std::string::npos) kv[line.substr(0, pos)] = line.substr(pos + 1);
    }
    parameters prm;
    prm.p = std::stoi(kv["p"]);
    prm.domain_x = std::stof(kv["domain_x"]);
    prm.domain_y = std::stof(kv["domain_y"]);
    prm.num_element_in_x = std::stoi(kv["num_element_in_x"]);
    prm.num_element_in_y = std::stoi(kv["num_element_in_y"]);
    prm.perturb_grid = std::stoi(kv["perturb_grid"]);
    prm.integration_order = std::stoi(kv["integration_order"]);
    prm.simulation_time = std::stof(kv["simulation_time"]);
    prm.number_time_steps = std::stoi(kv["number_time_steps"]);
    prm.time_step = std::stod(kv["simulation_time"]) / std::stod(kv["number_time_steps"]);
    prm.write_every_steps = std::stoi(kv["write_every_steps"]);
    prm.U_initialization_type = std::stoi(kv["U_initialization_type"]);
    prm.stepping_method = std::stoi(kv["stepping_method"]);
    return prm;
}

// ---------------------------------------------------------------------------
// Utilities.cpp:15-76
// ---------------------------------------------------------------------------
void writeToFile(const std::string& filename, const std::vector<std::string>& lines) {
    std::ofstream out(filename);
    if (!out.is_open()) {
        std::cerr << "Error opening file!" << std::endl;
        return;
    }
    for (const std::string& l : lines) out << l << std::endl;
}

void clean_create_directory(const std::string& dirname) {
    if (std::system(("rm -rf " + dirname).c_str()) != 0) {
        std::cerr << "Failed to clean directory: " << dirname << std::endl;
        exit(EXIT_FAILURE);
    }
    if (std::system(("mkdir -p " + dirname).c_str()) != 0) {
        std::cerr << "Failed to create directory: " << dirname << std::endl;
        exit(EXIT_FAILURE);
    }
}

void write_output(Element* elements, const int& n, const int& step_num) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) elements[i].write_data(step_num);
}

// ---------------------------------------------------------------------------
// Meshgeneration.cpp:16-219
// ---------------------------------------------------------------------------
mesh generate_mesh(const parameters& prm) {
    const int nx = prm.num_element_in_x, ny = prm.num_element_in_y, n = 2 * nx * ny;
    std::vector<double> v(6 * static_cast<size_t>(n));
    std::vector<int> nb(3 * static_cast<size_t>(n)), ty(n);
    int rc = vt_mesh_generate(nx, ny, prm.domain_x, prm.domain_y, prm.perturb_grid, v.data(), nb.data(), ty.data());
    if (rc) {
        std::cerr << "generate_mesh: " << vt_status_string(rc) << std::endl;
        exit(EXIT_FAILURE);
    }
    mesh m;
    m.element_number.resize(n);
    m.element_coordinates.assign(n, std::vector<std::vector<double>>(3, std::vector<double>(2)));
    m.elements_at_boundary.assign(n, std::vector<int>(3));
    m.element_type = ty;
    for (int i = 0; i < n; ++i) {
        m.element_number[i] = i;
        for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 2; ++c) m.element_coordinates[i][a][c] = v[6 * static_cast<size_t>(i) + 2 * a + c];
        for (int s = 0; s < 3; ++s) m.elements_at_boundary[i][s] = nb[3 * static_cast<size_t>(i) + s];
    }
    clean_create_directory("grid");
    for (int i = 0; i < n; ++i) {
        std::vector<std::string> lines(11);
        lines[0] = "element_number=" + std::to_string(i);
        lines[1] = "coordinate_1_x=" + std::to_string(m.element_coordinates[i][0][0]);
        lines[2] = "coordinate_1_y=" + std::to_string(m.element_coordinates[i][0][1]);
        lines[3] = "coordinate_2_x=" + std::to_string(m.element_coordinates[i][1][0]);
        lines[4] = "coordinate_2_y=" + std::to_string(m.element_coordinates[i][1][1]);
        lines[5] = "coordinate_3_x=" + std::to_string(m.element_coordinates[i][2][0]);
        lines[6] = "coordinate_3_y=" + std::to_string(m.element_coordinates[i][2][1]);
        lines[7] = "type=" + std::to_string(m.element_type[i]);
        lines[8] = "right_element=" + std::to_string(m.elements_at_boundary[i][0]);
        lines[9] = "left_element=" + std::to_string(m.elements_at_boundary[i][1]);
        lines[10] = "vertical_element=" + std::to_string(m.elements_at_boundary[i][2]);
        writeToFile("grid/element_" + std::to_string(i) + ".txt", lines);
    }
    return m;
}

std::vector<std::vector<double>> generate_nodes_reference_space(const parameters& prm) {
    const int np = num_nodes(prm.p);
    std::vector<double> x(2 * static_cast<size_t>(np));
    if (vt_nodes_reference(prm.p, x.data())) {
        std::cerr << "ERROR: Unsupported p order" << std::endl;
        exit(EXIT_FAILURE);
    }
    std::vector<std::vector<double>> out(np, std::vector<double>(2));
    for (int i = 0; i < np; ++i) { out[i][0] = x[2 * i]; out[i][1] = x[2 * i + 1]; }
    return out;
}

// ---------------------------------------------------------------------------
// Quadraturerule.cpp:13-661
// ---------------------------------------------------------------------------
std::vector<std::vector<double>> gauss_line_integral(const int& order) {
    std::vector<double> xw;
    const int n = vt::line_rule(order, &xw);
    if (n < 0) {
        printf("ERROR: Unsupported quadrature order \n");
        exit(EXIT_FAILURE);
    }
    std::vector<std::vector<double>> out(n, std::vector<double>(2));
    for (int i = 0; i < n; ++i) { out[i][0] = xw[2 * i]; out[i][1] = xw[2 * i + 1]; }
    return out;
}

std::vector<std::vector<double>> gauss_area_integral(const int& order) {
    std::vector<double> q;
    const int n = vt::area_rule(order, &q);
    if (n < 0) {
        printf("ERROR: Unsupported quadrature order \n");
        exit(EXIT_FAILURE);
    }
    std::vector<std::vector<double>> out(n, std::vector<double>(3));
    for (int i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) out[i][c] = q[3 * i + c];
    return out;
}

// ---------------------------------------------------------------------------
// Lagrangebasis.cpp:12-165
// ---------------------------------------------------------------------------
std::vector<double> lagrange_basis_reference_space(const int& p, const std::vector<double>& c) {
    std::vector<double> phi(num_nodes(p));
    if (!vt::basis(p, c[0], c[1], phi.data())) {
        printf("ERROR: Unsupported p order for lagrange_basis_reference_space");
        exit(EXIT_FAILURE);
    }
    return phi;
}

std::vector<std::vector<double>> lagrange_basis_gradient_reference_space(const int& p, const std::vector<double>& c) {
    std::vector<double> g(2 * static_cast<size_t>(num_nodes(p)));
    if (!vt::basis_gradient(p, c[0], c[1], g.data())) {
        printf("ERROR: Unsupported p order for lagrange_basis_gradient_reference_space");
        exit(EXIT_FAILURE);
    }
    std::vector<std::vector<double>> out(num_nodes(p), std::vector<double>(2));
    for (int i = 0; i < num_nodes(p); ++i) { out[i][0] = g[2 * i]; out[i][1] = g[2 * i + 1]; }
    return out;
}

std::vector<double> reference_to_physical_space(const std::vector<double>& c, const std::vector<std::vector<double>>& v) {
    std::vector<double> phi = lagrange_basis_reference_space(1, c);
    std::vector<double> r(2);
    r[0] = v[0][0] * phi[0] + v[1][0] * phi[1] + v[2][0] * phi[2];
    r[1] = v[0][1] * phi[0] + v[1][1] * phi[1] + v[2][1] * phi[2];
    return r;
}

// ---------------------------------------------------------------------------
// Preevolve.cpp:15-144 (mass matrix by the area rule, its inverse by
// Gauss-Jordan with partial pivoting, stiffness matrices)
// ---------------------------------------------------------------------------
static void operators_from_rule(int p, const std::vector<std::vector<double>>& q, std::vector<double>* M,
                                std::vector<double>* S) {
    const int np = num_nodes(p), n = static_cast<int>(q.size());
    M->assign(np * np, 0.0);
    S->assign(2 * np * np, 0.0);
    std::vector<double> phi(np * n), gr(2 * np * n);
    for (int k = 0; k < n; ++k) {
        double b[10], g[20];
        if (!vt::basis(p, q[k][0], q[k][1], b) || !vt::basis_gradient(p, q[k][0], q[k][1], g)) {
            printf("ERROR: Unsupported p order for lagrange_basis_reference_space");
            exit(EXIT_FAILURE);
        }
        for (int j = 0; j < np; ++j) { phi[j * n + k] = b[j]; gr[(j * n + k) * 2] = g[2 * j]; gr[(j * n + k) * 2 + 1] = g[2 * j + 1]; }
    }
    for (int i = 0; i < np; ++i)
        for (int j = 0; j < np; ++j) {
            double a = 0, s0 = 0, s1 = 0;
            for (int k = 0; k < n; ++k) {
                a += phi[i * n + k] * phi[j * n + k] * q[k][2];
                s0 += gr[(i * n + k) * 2] * phi[j * n + k] * q[k][2];
                s1 += gr[(i * n + k) * 2 + 1] * phi[j * n + k] * q[k][2];
            }
            (*M)[i * np + j] = a;
            (*S)[i * np + j] = s0;
            (*S)[np * np + i * np + j] = s1;
        }
}

std::vector<std::vector<double>> inverse_mass_matrix_reference_space(const int& p, const std::vector<std::vector<double>>& q) {
    std::vector<double> M, S, Mi(num_nodes(p) * num_nodes(p));
    operators_from_rule(p, q, &M, &S);
    vt::invert(num_nodes(p), M.data(), Mi.data());
    std::vector<std::vector<double>> out(num_nodes(p), std::vector<double>(num_nodes(p)));
    for (int i = 0; i < num_nodes(p); ++i)
        for (int j = 0; j < num_nodes(p); ++j) out[i][j] = Mi[i * num_nodes(p) + j];
    return out;
}

std::vector<std::vector<std::vector<double>>> sitffness_matrix_reference_space(const int& p, const std::vector<std::vector<double>>& q) {
    std::vector<double> M, S;
    operators_from_rule(p, q, &M, &S);
    const int np = num_nodes(p);
    std::vector<std::vector<std::vector<double>>> out(2, std::vector<std::vector<double>>(np, std::vector<double>(np)));
    for (int d = 0; d < 2; ++d)
        for (int i = 0; i < np; ++i)
            for (int j = 0; j < np; ++j) out[d][i][j] = S[d * np * np + i * np + j];
    return out;
}

// ---------------------------------------------------------------------------
// Element (Element.cpp:19-425)
// ---------------------------------------------------------------------------
Element::Element() {}

Element::Element(const int& ele_num, const mesh& m, const std::vector<std::vector<double>>& nods, const int& p_lagrange)
    : time(0.0), p(p_lagrange), number(ele_num), type(m.element_type[ele_num]),
      right_element(m.elements_at_boundary[ele_num][0]), left_element(m.elements_at_boundary[ele_num][1]),
      vertical_element(m.elements_at_boundary[ele_num][2]), vertices_coords_phys_space(m.element_coordinates[ele_num]),
      nods_coords_refe_space(nods), nods_coords_phys_space(num_nodes(p_lagrange), std::vector<double>(2)),
      hidrodynamics_vector_U(num_nodes(p_lagrange), std::vector<double>(4)),
      hidrodynamics_vector_F(num_nodes(p_lagrange), std::vector<std::vector<double>>(2, std::vector<double>(4))),
      jacobian(2, std::vector<double>(2)), determinant_jacobian(0), inverse_jacobian(2, std::vector<double>(2)),
      determinant_inverse_jacobian(0),
      inverse_mass_matrix_physical_space(num_nodes(p_lagrange), std::vector<double>(num_nodes(p_lagrange))),
      stiffness_matrix_physical_space(2, std::vector<std::vector<double>>(num_nodes(p_lagrange), std::vector<double>(num_nodes(p_lagrange)))),
      units_vectors_perpendicular_to_element_boundary(3, std::vector<double>(2)), sides_lenght(3) {
    for (int i = 0; i < num_nodes(p); ++i) nods_coords_phys_space[i] = reference_to_physical_space(nods_coords_refe_space[i], vertices_coords_phys_space);
}

void Element::build_jacobians() {
    const auto& v = vertices_coords_phys_space;
    jacobian[0][0] = v[1][0] - v[0][0];
    jacobian[0][1] = v[2][0] - v[0][0];
    jacobian[1][0] = v[1][1] - v[0][1];
    jacobian[1][1] = v[2][1] - v[0][1];
    determinant_jacobian = jacobian[0][0] * jacobian[1][1] - jacobian[1][0] * jacobian[0][1];
    inverse_jacobian[0][0] = (1 / determinant_jacobian) * (v[2][1] - v[0][1]);
    inverse_jacobian[0][1] = (1 / determinant_jacobian) * (v[0][0] - v[2][0]);
    inverse_jacobian[1][0] = (1 / determinant_jacobian) * (v[0][1] - v[1][1]);
    inverse_jacobian[1][1] = (1 / determinant_jacobian) * (v[1][0] - v[0][0]);
    determinant_inverse_jacobian = inverse_jacobian[0][0] * inverse_jacobian[1][1] - inverse_jacobian[1][0] * inverse_jacobian[0][1];
    for (int s = 0; s < 3; ++s) {
        const int a = s, b = (s + 1) % 3;
        const double dx = v[b][0] - v[a][0], dy = v[b][1] - v[a][1];
        const double len = std::pow(std::pow(dx, 2) + std::pow(dy, 2), 0.5);
        units_vectors_perpendicular_to_element_boundary[s][0] = dy / len;
        units_vectors_perpendicular_to_element_boundary[s][1] = -1 * dx / len;
        sides_lenght[s] = len;
    }
}

void Element::build_mass_matrix_inverse(const std::vector<std::vector<double>>& inv) {
    const int np = num_nodes(p);
    std::vector<double> ref(np * np);
    for (int i = 0; i < np; ++i)
        for (int j = 0; j < np; ++j) {
            inverse_mass_matrix_physical_space[i][j] = inv[i][j] / determinant_jacobian;
            ref[i * np + j] = inv[i][j];
        }
    record_operators(this, &ref, nullptr);
}

void Element::build_stiffness_matrix(const std::vector<std::vector<std::vector<double>>>& S) {
    const int np = num_nodes(p);
    std::vector<double> ref(2 * np * np);
    for (int i = 0; i < np; ++i)
        for (int j = 0; j < np; ++j) {
            stiffness_matrix_physical_space[0][i][j] = determinant_jacobian * (S[0][i][j] * inverse_jacobian[0][0] + S[1][i][j] * inverse_jacobian[1][0]);
            stiffness_matrix_physical_space[1][i][j] = determinant_jacobian * (S[0][i][j] * inverse_jacobian[0][1] + S[1][i][j] * inverse_jacobian[1][1]);
            ref[i * np + j] = S[0][i][j];
            ref[np * np + i * np + j] = S[1][i][j];
        }
    record_operators(this, nullptr, &ref);
}

void Element::initialize_hydrodinamics(const int& ini_type, const std::vector<std::vector<double>>& gau_area_int) {
    const int np = num_nodes(p);
    if (ini_type == 0) {
        for (int i = 0; i < np; ++i) {
            double w[6];
            vt::vortex_point(nods_coords_phys_space[i][0], nods_coords_phys_space[i][1], time, 0.3, 0.0, 0.0, w);
            const double rho = w[0], u = w[1], v = w[2], pr = w[3], E = w[4], H = w[5];
            hidrodynamics_vector_U[i] = {rho, rho * u, rho * v, rho * E};
            hidrodynamics_vector_F[i][0] = {rho * u, rho * std::pow(u, 2) + pr, rho * u * v, rho * u * H};
            hidrodynamics_vector_F[i][1] = {rho * v, rho * u * v, rho * std::pow(v, 2) + pr, rho * v * H};
        }
        return;
    }
    if (ini_type == 1) {  // least-squares projection, Element.cpp:180-326
        const int nqa = static_cast<int>(gau_area_int.size());
        std::vector<double> phi(np * nqa), ue(4 * nqa);
        for (int k = 0; k < nqa; ++k) {
            std::vector<double> xi = {gau_area_int[k][0], gau_area_int[k][1]};
            std::vector<double> b = lagrange_basis_reference_space(p, xi);
            for (int j = 0; j < np; ++j) phi[j * nqa + k] = b[j];
            std::vector<double> r = reference_to_physical_space(xi, vertices_coords_phys_space);
            double w[6];
            vt::vortex_point(r[0], r[1], time, 0.3, 0.0, 0.0, w);
            ue[4 * k] = w[0]; ue[4 * k + 1] = w[0] * w[1]; ue[4 * k + 2] = w[0] * w[2]; ue[4 * k + 3] = w[0] * w[4];
        }
        std::vector<std::vector<double>> bv(np, std::vector<double>(4));
        for (int i = 0; i < np; ++i)
            for (int c = 0; c < 4; ++c) {
                bv[i][c] = 0.0;
                for (int k = 0; k < nqa; ++k) bv[i][c] += determinant_jacobian * phi[i * nqa + k] * ue[4 * k + c] * gau_area_int[k][2];
            }
        for (int i = 0; i < np; ++i) {
            hidrodynamics_vector_U[i] = {0.0, 0.0, 0.0, 0.0};
            for (int c = 0; c < 4; ++c)
                for (int j = 0; j < np; ++j) hidrodynamics_vector_U[i][c] += inverse_mass_matrix_physical_space[i][j] * bv[j][c];
            double F[8];
            vt::nodal_flux(hidrodynamics_vector_U[i].data(), F);
            hidrodynamics_vector_F[i][0] = {F[0], F[1], F[2], F[3]};
            hidrodynamics_vector_F[i][1] = {F[4], F[5], F[6], F[7]};
        }
        return;
    }
    printf("ERROR: Unsupported initialization type\n0 : direct interpolation\n1 : least squere projection\n");
    exit(EXIT_FAILURE);
}

// Element.cpp:336-425: the text format plottools/ reads.
void Element::write_data(const int& step_num) {
    const int np = num_nodes(p);
    std::vector<std::string> lines(1 + np);
    lines[0] = "node_number time x y u0 u1 u2 u3 fx0 fx1 fx2 fx3 fy0 fy1 fy2 fy3";
    for (int i = 0; i < np; ++i) {
        std::string& l = lines[i + 1];
        l = std::to_string(i) + " " + std::to_string(time) + " " + std::to_string(nods_coords_phys_space[i][0]) + " " +
            std::to_string(nods_coords_phys_space[i][1]) + " ";
        for (int k = 0; k < 4; ++k) l += std::to_string(hidrodynamics_vector_U[i][k]) + " ";
        for (int d = 0; d < 2; ++d)
            for (int k = 0; k < 4; ++k) l += std::to_string(hidrodynamics_vector_F[i][d][k]) + " ";
    }
    writeToFile("output/step_" + std::to_string(step_num) + "/element_" + std::to_string(number) + ".txt", lines);
    if (step_num != 0) return;
    auto row = [&](const std::vector<double>& r) {
        std::string s = "[ " + std::to_string(r[0]);
        for (size_t j = 1; j < r.size(); ++j) s += ", " + std::to_string(r[j]);
        return s + " ] ,";
    };
    std::vector<std::string> L;
    L.push_back("jacobian");
    L.push_back("[ [ " + std::to_string(jacobian[0][0]) + " , " + std::to_string(jacobian[0][1]) + " ] , ");
    L.push_back("  [ " + std::to_string(jacobian[1][0]) + " , " + std::to_string(jacobian[1][1]) + " ] ]");
    L.push_back("determinant of jacobian");
    L.push_back(std::to_string(determinant_jacobian));
    L.push_back("inverse jacobian");
    L.push_back("[ [ " + std::to_string(inverse_jacobian[0][0]) + " , " + std::to_string(inverse_jacobian[0][1]) + " ] , ");
    L.push_back("  [ " + std::to_string(inverse_jacobian[1][0]) + " , " + std::to_string(inverse_jacobian[1][1]) + " ] ]");
    L.push_back("determinant of inverse jacobian");
    L.push_back(std::to_string(determinant_inverse_jacobian));
    auto block = [&](const char* title, const std::vector<std::vector<double>>& m) {
        L.push_back(title);
        const size_t first = L.size();
        for (int i = 0; i < np; ++i) L.push_back(row(m[i]));
        L[first] = "[ " + L[first];
        L.back().erase(L.back().size() - 1);
        L.back() += "]";
    };
    block("inverse mass matrix", inverse_mass_matrix_physical_space);
    block("stiffness matrix : x component", stiffness_matrix_physical_space[0]);
    block("stiffness matrix : y component", stiffness_matrix_physical_space[1]);
    for (int s = 0; s < 3; ++s)
        L.push_back("Unit vector perpendicular to edge " + std::to_string(s + 1) + " : [ " +
                    std::to_string(units_vectors_perpendicular_to_element_boundary[s][0]) + " , " +
                    std::to_string(units_vectors_perpendicular_to_element_boundary[s][1]) + " ]");
    writeToFile("output/step_" + std::to_string(step_num) + "/JMS_element_" + std::to_string(number) + ".txt", L);
}

// ---------------------------------------------------------------------------
// Evolve_element (Evolve.cpp:22-132); RHS stages run fused on the GPU
// ---------------------------------------------------------------------------
Evolve_element::Evolve_element() {}

Evolve_element::Evolve_element(Element* t, Element* r, Element* l, Element* v, const std::vector<std::vector<double>>& g, const int& pp)
    : p(pp), element_this(t), element_right(r), element_left(l), element_vertical(v), gau_integ_line(g) {
    const int np = num_nodes(p), nq = static_cast<int>(g.size());
    for (auto* m : {&plus_phi_in_quadrature_points_side_1, &plus_phi_in_quadrature_points_side_2, &plus_phi_in_quadrature_points_side_3,
                    &minus_phi_in_quadrature_points_side_1, &minus_phi_in_quadrature_points_side_2, &minus_phi_in_quadrature_points_side_3})
        m->assign(np, std::vector<double>(nq));
    for (auto* m : {&U_plus_side_1, &U_plus_side_2, &U_plus_side_3, &U_minus_side_1, &U_minus_side_2, &U_minus_side_3,
                    &numerical_flux_side_1, &numerical_flux_side_2, &numerical_flux_side_3})
        m->assign(nq, std::vector<double>(4));
    for (auto* m : {&DG_numerical_flux_integration, &DG_stiffness_vector, &DG_residual_vector, &DG_time_derivative_U})
        m->assign(np, std::vector<double>(4));
}

void Evolve_element::evaluate_basis_in_quadrature_poits() {
    const int np = num_nodes(p), nq = static_cast<int>(gau_integ_line.size());
    for (int q = 0; q < nq; ++q) {
        const double sg = gau_integ_line[q][0];
        const double pts[6][2] = {{sg, 0.0}, {1.0 - sg, 0.0}, {1.0 - sg, sg}, {sg, 1.0 - sg}, {0.0, 1.0 - sg}, {0.0, sg}};
        std::vector<std::vector<double>>* tabs[6] = {&plus_phi_in_quadrature_points_side_1, &minus_phi_in_quadrature_points_side_1,
                                                     &plus_phi_in_quadrature_points_side_2, &minus_phi_in_quadrature_points_side_2,
                                                     &plus_phi_in_quadrature_points_side_3, &minus_phi_in_quadrature_points_side_3};
        for (int t = 0; t < 6; ++t) {
            std::vector<double> b = lagrange_basis_reference_space(p, {pts[t][0], pts[t][1]});
            for (int j = 0; j < np; ++j) (*tabs[t])[j][q] = b[j];
        }
    }
}

void Evolve_element::compute_U_plus_minus() {}
void Evolve_element::compute_numerical_flux() {}
void Evolve_element::integrate_numerical_flux() {}
void Evolve_element::compute_stiffness_vector() {}
void Evolve_element::compute_residual_vector() {}

void Evolve_element::compute_time_derivative_U() {
    // The element array starts number elements before this one; its
    // Evolve_element array likewise (main.cpp:115,137 allocate both with new[]).
    Element* base = element_this - element_this->number;
    Evolve_element* ebase = this - element_this->number;
    auto it = g_bind.find(base);
    const int n = it != g_bind.end() ? it->second.n : -1;
    if (n < 0) {
        std::cerr << "compute_time_derivative_U: call evolve_elem (or rk4) on the element array first" << std::endl;
        exit(EXIT_FAILURE);
    }
    evolve_elem(ebase, n);
}

// ---------------------------------------------------------------------------
// Timestepping.cpp:18-305, on the GPU
// ---------------------------------------------------------------------------
void evolve_elem(Evolve_element* ev, int n) {
    Element* el = ev[0].element_this;
    Binding& b = bind(el, ev, n);
    upload(b, el);
    std::vector<double> d(static_cast<size_t>(n) * b.np * 4);
    int rc = vt_rhs(b.ctx, d.data());
    if (rc == VT_ERR_NONPHYSICAL) {
        int info[6];
        vt_last_error(b.ctx, info);
        die_nonphysical(info);
    }
    if (rc) die_hip("vt_rhs", rc);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < b.np; ++j)
            for (int k = 0; k < 4; ++k) ev[i].DG_time_derivative_U[j][k] = d[(static_cast<size_t>(i) * b.np + j) * 4 + k];
}

void forward_euler(Element* el, Evolve_element* ev, double& dt, const int& n) { step_on_gpu(el, ev, 0, dt, 1, n); }

void rk4(Element* el, Evolve_element* ev, double& dt, const int& n) { step_on_gpu(el, ev, 1, dt, 1, n); }

void gpu_steps(Element* el, Evolve_element* ev, int method, double dt, int nsteps, const int& n) {
    if (method != 0 && method != 1) {
        printf("ERROR: Unsupported time stepping method\n0 : forward euler\n1 : rk4\n");
        exit(EXIT_FAILURE);
    }
    step_on_gpu(el, ev, method, dt, nsteps, n);
}

// ---------------------------------------------------------------------------
// Numericalflux.cpp:13-166 on the GPU
// ---------------------------------------------------------------------------
void numerical_flux(const std::vector<double>& ul, const std::vector<double>& ur, const std::vector<double>& nv, std::vector<double>& flux) {
    double rec[10] = {ul[0], ul[1], ul[2], ul[3], ur[0], ur[1], ur[2], ur[3], nv[0], nv[1]};
    double f[4];
    int code = 0;
    int rc = vt_numerical_flux(rec, f, &code, 1);
    if (rc) die_hip("vt_numerical_flux", rc);
    if (code) {
        std::cerr << code << ". Error: Non-physical state! " << std::endl;
        exit(EXIT_FAILURE);
    }
    for (int k = 0; k < 4; ++k) flux[k] = f[k];
}
