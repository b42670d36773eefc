// Host-side check of the C++ drop-in: builds the element array the way the
// reference driver does (main.cpp:96-131) through the drop-in API and dumps
// per element: type, right, left, vertical, vertices[6], jacobian[4], detJ,
// invJ[4], detInvJ, normals[6], lengths[3], nodes[2 Np], then U0 and F0, as
// raw doubles.  With --write it also writes the step-0 text output into the
// current directory as the reference driver does (main.cpp:119-121,136).
// No GPU call is made.
//   dropin_check <input_file> <out.bin> [--write]
#include <cstdio>
#include <cstring>
#include <vector>

#include "Element.H"
#include "Meshgeneration.H"
#include "Parameters.H"
#include "Preevolve.H"
#include "Quadraturerule.H"
#include "Utilities.H"

int main(int argc, char* argv[]) {
    char* pargv[2] = {argv[0], argv[1]};
    parameters prm = read_input_files(2, pargv);
    mesh m = generate_mesh(prm);
    auto area = gauss_area_integral(prm.integration_order);
    auto nodes = generate_nodes_reference_space(prm);
    auto minv = inverse_mass_matrix_reference_space(prm.p, area);
    auto stiff = sitffness_matrix_reference_space(prm.p, area);
    const int n = 2 * prm.num_element_in_x * prm.num_element_in_y;
    std::vector<Element> el(n);
    for (int i = 0; i < n; ++i) {
        el[i] = Element(i, m, nodes, prm.p);
        el[i].build_jacobians();
        el[i].build_mass_matrix_inverse(minv);
        el[i].build_stiffness_matrix(stiff);
        el[i].initialize_hydrodinamics(prm.U_initialization_type, area);
    }
    if (argc > 3 && std::strcmp(argv[3], "--write") == 0) {
        clean_create_directory("output");
        clean_create_directory("output/step_0");
        write_output(el.data(), n, 0);
    }
    FILE* f = std::fopen(argv[2], "wb");
    auto put = [&](double x) { std::fwrite(&x, sizeof(double), 1, f); };
    for (int i = 0; i < n; ++i) {
        const Element& e = el[i];
        put(e.type); put(e.right_element); put(e.left_element); put(e.vertical_element);
        for (auto& v : e.vertices_coords_phys_space) { put(v[0]); put(v[1]); }
        put(e.jacobian[0][0]); put(e.jacobian[0][1]); put(e.jacobian[1][0]); put(e.jacobian[1][1]);
        put(e.determinant_jacobian);
        put(e.inverse_jacobian[0][0]); put(e.inverse_jacobian[0][1]); put(e.inverse_jacobian[1][0]); put(e.inverse_jacobian[1][1]);
        put(e.determinant_inverse_jacobian);
        for (auto& nv : e.units_vectors_perpendicular_to_element_boundary) { put(nv[0]); put(nv[1]); }
        for (double l : e.sides_lenght) put(l);
        for (auto& x : e.nods_coords_phys_space) { put(x[0]); put(x[1]); }
    }
    for (int i = 0; i < n; ++i)
        for (auto& u : el[i].hidrodynamics_vector_U)
            for (double x : u) put(x);
    for (int i = 0; i < n; ++i)
        for (auto& fj : el[i].hidrodynamics_vector_F)
            for (auto& fd : fj)
                for (double x : fd) put(x);
    std::fclose(f);
    return 0;
}
