// ref_harness.cpp -- drives the reference's own translation units (built
// unmodified from /root/reference by oracle/Makefile) to produce parity data.
//
// TEST INFRASTRUCTURE ONLY.  This follows the initialisation order of the
// reference driver (main.cpp:96-147) and its time loop (main.cpp:154-179),
// with two differences, both because Preevolve.cpp cannot be compiled here
// (it needs Eigen 3, which is not installed):
//   * the reference inverse mass matrix and stiffness matrices are read from
//     files written by the oracle (they are the arguments of
//     Element::build_mass_matrix_inverse / build_stiffness_matrix,
//     Element.H:49-50), and
//   * the mesh may be read from a file instead of generate_mesh (for meshes
//     the reference mesher cannot produce, e.g. domains wider than 12).
// No reference source is modified or copied; output is raw little-endian
// binary, and nothing is written under grid/ or output/ unless generate_mesh
// is used (it writes grid/ into the current directory, Meshgeneration.cpp:160-179).
//
// Usage:
//   vt_ref_run run  <input> <out.bin> --minv F --stiff F [--mesh F] [--steps N]
//                   [--rhs-at-end] [--no-state] [--write-dumps]
// --write-dumps writes the reference's text output into the current directory
// as its driver does (main.cpp:119-121,136,171-177): output/step_0 after the
// set-up and output/step_N every write_every_steps steps (write_output,
// Utilities.cpp:65-76, Element::write_data, Element.cpp:336-425).
//   vt_ref_run flux <in.bin> <out.bin>      records of UL[4] UR[4] n[2] -> flux[4]
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "Element.H"
#include "Evolve.H"
#include "Meshgeneration.H"
#include "Numericalflux.H"
#include "Parameters.H"
#include "Quadraturerule.H"
#include "Timestepping.H"
#include "Utilities.H"

// Defined in Timestepping.cpp:18 with external linkage but declared in no header.
void evolve_elem(Evolve_element* evo_elemts, int n_elements);

static std::vector<double> read_doubles(const char* path) {
    FILE* f = std::fopen(path, "rb");
    if (!f) { std::fprintf(stderr, "cannot open %s\n", path); std::exit(2); }
    std::vector<double> v;
    double x;
    while (std::fread(&x, sizeof(double), 1, f) == 1) v.push_back(x);
    std::fclose(f);
    return v;
}

static void put(FILE* f, const std::vector<double>& v) { std::fwrite(v.data(), sizeof(double), v.size(), f); }

static int run_flux(const char* in, const char* out) {
    std::vector<double> d = read_doubles(in);
    size_t n = d.size() / 10;
    FILE* f = std::fopen(out, "wb");
    std::vector<double> ul(4), ur(4), nv(2), fl(4);
    for (size_t r = 0; r < n; ++r) {
        for (int k = 0; k < 4; ++k) { ul[k] = d[10 * r + k]; ur[k] = d[10 * r + 4 + k]; }
        nv[0] = d[10 * r + 8]; nv[1] = d[10 * r + 9];
        numerical_flux(ul, ur, nv, fl);
        put(f, fl);
    }
    std::fclose(f);
    return 0;
}

int main(int argc, char* argv[]) {
    if (argc >= 4 && std::strcmp(argv[1], "flux") == 0) return run_flux(argv[2], argv[3]);
    if (argc < 4 || std::strcmp(argv[1], "run") != 0) {
        std::fprintf(stderr, "usage: %s run <input> <out.bin> --minv F --stiff F [--mesh F] [--steps N] "
                             "[--rhs-at-end] [--no-state]\n       %s flux <in.bin> <out.bin>\n", argv[0], argv[0]);
        return 2;
    }
    const char *minv_path = 0, *stiff_path = 0, *mesh_path = 0;
    int steps = -1;
    bool rhs_at_end = false, no_state = false, write_dumps = false;
    for (int i = 4; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "--minv" && i + 1 < argc) minv_path = argv[++i];
        else if (a == "--stiff" && i + 1 < argc) stiff_path = argv[++i];
        else if (a == "--mesh" && i + 1 < argc) mesh_path = argv[++i];
        else if (a == "--steps" && i + 1 < argc) steps = std::atoi(argv[++i]);
        else if (a == "--rhs-at-end") rhs_at_end = true;
        else if (a == "--no-state") no_state = true;
        else if (a == "--write-dumps") write_dumps = true;
        else { std::fprintf(stderr, "unknown option %s\n", argv[i]); return 2; }
    }
    if (!minv_path || !stiff_path) { std::fprintf(stderr, "--minv and --stiff are required\n"); return 2; }

    char* pargv[2] = {argv[0], argv[2]};
    parameters parms = read_input_files(2, pargv);
    if (steps < 0) steps = parms.number_time_steps;
    const int np = (parms.p + 1) * (parms.p + 2) / 2;

    mesh simulation_mesh;
    if (mesh_path) {
        std::vector<double> m = read_doubles(mesh_path);  // ne, then per element 6 coords, 3 nbrs, type
        int ne = (int)m[0];
        simulation_mesh.element_number.resize(ne);
        simulation_mesh.element_coordinates.assign(ne, std::vector<std::vector<double>>(3, std::vector<double>(2)));
        simulation_mesh.elements_at_boundary.assign(ne, std::vector<int>(3));
        simulation_mesh.element_type.resize(ne);
        for (int e = 0; e < ne; ++e) {
            const double* r = &m[1 + 10 * e];
            simulation_mesh.element_number[e] = e;
            for (int v = 0; v < 3; ++v)
                for (int c = 0; c < 2; ++c) simulation_mesh.element_coordinates[e][v][c] = r[2 * v + c];
            for (int s = 0; s < 3; ++s) simulation_mesh.elements_at_boundary[e][s] = (int)r[6 + s];
            simulation_mesh.element_type[e] = (int)r[9];
        }
    } else {
        simulation_mesh = generate_mesh(parms);
    }
    const int ne = (int)simulation_mesh.element_number.size();

    std::vector<std::vector<double>> gauss_integral_line = gauss_line_integral(parms.integration_order);
    std::vector<std::vector<double>> gauss_integral_area = gauss_area_integral(parms.integration_order);
    std::vector<std::vector<double>> nodes_reference_space = generate_nodes_reference_space(parms);

    std::vector<double> mi = read_doubles(minv_path), st = read_doubles(stiff_path);
    if ((int)mi.size() != np * np || (int)st.size() != 2 * np * np) { std::fprintf(stderr, "operator size mismatch\n"); return 2; }
    std::vector<std::vector<double>> inverse_mass_matrix(np, std::vector<double>(np));
    std::vector<std::vector<std::vector<double>>> stiffness_matrix(2, std::vector<std::vector<double>>(np, std::vector<double>(np)));
    for (int i = 0; i < np; ++i)
        for (int j = 0; j < np; ++j) {
            inverse_mass_matrix[i][j] = mi[i * np + j];
            stiffness_matrix[0][i][j] = st[i * np + j];
            stiffness_matrix[1][i][j] = st[np * np + i * np + j];
        }

    Element* elements = new Element[ne];
#pragma omp parallel for schedule(static)
    for (int i = 0; i < ne; ++i) {
        elements[i] = Element(i, simulation_mesh, nodes_reference_space, parms.p);
        elements[i].build_jacobians();
        elements[i].build_mass_matrix_inverse(inverse_mass_matrix);
        elements[i].build_stiffness_matrix(stiffness_matrix);
        elements[i].initialize_hydrodinamics(parms.U_initialization_type, gauss_integral_area);
    }
    if (write_dumps) {
        clean_create_directory("output");
        clean_create_directory("output/step_0");
        write_output(elements, ne, 0);
    }
    Evolve_element* evolve_elements = new Evolve_element[ne];
#pragma omp parallel for schedule(static)
    for (int i = 0; i < ne; ++i) {
        evolve_elements[i] = Evolve_element(&elements[i], &elements[elements[i].right_element], &elements[elements[i].left_element],
                                            &elements[elements[i].vertical_element], gauss_integral_line, parms.p);
        evolve_elements[i].evaluate_basis_in_quadrature_poits();
    }

    FILE* out = std::fopen(argv[3], "wb");
    auto dump_state = [&]() {
        std::vector<double> u(4 * np), f(8 * np);
        for (int e = 0; e < ne; ++e) {
            for (int i = 0; i < np; ++i)
                for (int k = 0; k < 4; ++k) u[4 * i + k] = elements[e].hidrodynamics_vector_U[i][k];
            put(out, u);
        }
        for (int e = 0; e < ne; ++e) {
            for (int i = 0; i < np; ++i)
                for (int d = 0; d < 2; ++d)
                    for (int k = 0; k < 4; ++k) f[8 * i + 4 * d + k] = elements[e].hidrodynamics_vector_F[i][d][k];
            put(out, f);
        }
    };
    // Setup record: per element type, R, L, V, 6 vertices, jac4, detJ, invJ4, detInvJ, n6, len3, nodes.
    if (!no_state) {
        std::vector<double> rec;
        for (int e = 0; e < ne; ++e) {
            const Element& el = elements[e];
            rec.assign({(double)el.type, (double)el.right_element, (double)el.left_element, (double)el.vertical_element});
            for (int v = 0; v < 3; ++v)
                for (int c = 0; c < 2; ++c) rec.push_back(el.vertices_coords_phys_space[v][c]);
            rec.insert(rec.end(), {el.jacobian[0][0], el.jacobian[0][1], el.jacobian[1][0], el.jacobian[1][1], el.determinant_jacobian,
                                   el.inverse_jacobian[0][0], el.inverse_jacobian[0][1], el.inverse_jacobian[1][0],
                                   el.inverse_jacobian[1][1], el.determinant_inverse_jacobian});
            for (int s = 0; s < 3; ++s)
                for (int c = 0; c < 2; ++c) rec.push_back(el.units_vectors_perpendicular_to_element_boundary[s][c]);
            for (int s = 0; s < 3; ++s) rec.push_back(el.sides_lenght[s]);
            for (int i = 0; i < np; ++i)
                for (int c = 0; c < 2; ++c) rec.push_back(el.nods_coords_phys_space[i][c]);
            put(out, rec);
        }
        dump_state();
    }

    auto t0 = std::chrono::steady_clock::now();
    for (int a = 1; a < steps + 1; ++a) {
        if (parms.stepping_method == 0) forward_euler(elements, evolve_elements, parms.time_step, ne);
        else if (parms.stepping_method == 1) rk4(elements, evolve_elements, parms.time_step, ne);
        else { std::printf("ERROR: Unsupported time stepping method\n"); return 1; }
        if (write_dumps && a % parms.write_every_steps == 0) {
            clean_create_directory("output/step_" + std::to_string(a));
            write_output(elements, ne, a);
        }
    }
    auto t1 = std::chrono::steady_clock::now();
    double sec = std::chrono::duration<double>(t1 - t0).count();

    if (!no_state) {
        dump_state();
        std::vector<double> tm(1, elements[0].time);
        put(out, tm);
        if (rhs_at_end) {
            evolve_elem(evolve_elements, ne);
            std::vector<double> d(4 * np);
            for (int e = 0; e < ne; ++e) {
                for (int i = 0; i < np; ++i)
                    for (int k = 0; k < 4; ++k) d[4 * i + k] = evolve_elements[e].DG_time_derivative_U[i][k];
                put(out, d);
            }
        }
    }
    std::fclose(out);
    std::printf("[TIMING] elements=%d p=%d steps=%d stepping_sec=%.6f\n", ne, parms.p, steps, sec);
    return 0;
}
