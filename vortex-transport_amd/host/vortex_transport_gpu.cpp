// vortex_transport_gpu -- driver with the flow of the reference's main.cpp
// (main.cpp:81-197: [PHASE] markers, set-up, time loop with text output every
// write_every_steps, [TIMING] lines), using the drop-in classes.  Between two
// output steps the state stays in HBM (gpu_steps), so only output steps copy
// it back to the elements.
//
//   vortex_transport_gpu <input_file>
#include <chrono>
#include <cstdio>
#include <iostream>
#include <string>
#include <vector>

#include "Element.H"
#include "Evolve.H"
#include "Meshgeneration.H"
#include "Parameters.H"
#include "Preevolve.H"
#include "Quadraturerule.H"
#include "Timestepping.H"
#include "Utilities.H"

int main(int argc, char* argv[]) {
    const auto t_start = std::chrono::steady_clock::now();
    std::cout << "[PHASE] initialization begin" << std::endl;
    parameters parms = read_input_files(argc, argv);
    mesh simulation_mesh = generate_mesh(parms);
    std::vector<std::vector<double>> line = gauss_line_integral(parms.integration_order);
    std::vector<std::vector<double>> area = gauss_area_integral(parms.integration_order);
    std::vector<std::vector<double>> nodes = generate_nodes_reference_space(parms);
    std::vector<std::vector<double>> minv = inverse_mass_matrix_reference_space(parms.p, area);
    std::vector<std::vector<std::vector<double>>> stiff = sitffness_matrix_reference_space(parms.p, area);
    const int n = 2 * parms.num_element_in_x * parms.num_element_in_y;
    Element* elements = new Element[n];
    clean_create_directory("output");
    clean_create_directory("output/step_0");
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        elements[i] = Element(i, simulation_mesh, nodes, parms.p);
        elements[i].build_jacobians();
        elements[i].build_mass_matrix_inverse(minv);
        elements[i].build_stiffness_matrix(stiff);
        elements[i].initialize_hydrodinamics(parms.U_initialization_type, area);
    }
    write_output(elements, n, 0);
    Evolve_element* evolve = new Evolve_element[n];
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        evolve[i] = Evolve_element(&elements[i], &elements[elements[i].right_element], &elements[elements[i].left_element],
                                   &elements[elements[i].vertical_element], line, parms.p);
        evolve[i].evaluate_basis_in_quadrature_poits();
    }
    const auto t_init = std::chrono::steady_clock::now();
    std::cout << "[PHASE] initialization end" << std::endl;
    std::cout << "[PHASE] timestepping begin" << std::endl;
    if (parms.stepping_method != 0 && parms.stepping_method != 1) {
        printf("ERROR: Unsupported time stepping method\n0 : forward euler\n1 : rk4\n");
        return 1;
    }
    int a = 0;
    while (a < parms.number_time_steps) {
        int chunk = parms.write_every_steps - a % parms.write_every_steps;
        if (a + chunk > parms.number_time_steps) chunk = parms.number_time_steps - a;
        gpu_steps(elements, evolve, parms.stepping_method, parms.time_step, chunk, n);
        a += chunk;
        std::cout << "Step : " << a << std::endl;
        if (a % parms.write_every_steps == 0) {
            clean_create_directory("output/step_" + std::to_string(a));
            std::cout << "Writing Step : " << a << std::endl;
            write_output(elements, n, a);
        }
    }
    const auto t_step = std::chrono::steady_clock::now();
    std::cout << "[PHASE] timestepping end" << std::endl;
    const auto t_end = std::chrono::steady_clock::now();
    std::printf("[TIMING] total_time_sec = %.6f\n", std::chrono::duration<double>(t_end - t_start).count());
    std::printf("[TIMING] init_time_sec = %.6f\n", std::chrono::duration<double>(t_init - t_start).count());
    std::printf("[TIMING] timestepping_time_sec = %.6f\n", std::chrono::duration<double>(t_step - t_init).count());
    std::printf("[TIMING] teardown_time_sec = %.6f\n", std::chrono::duration<double>(t_end - t_step).count());
    return 0;
}
