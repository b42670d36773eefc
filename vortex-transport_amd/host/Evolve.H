// Drop-in declaration of the per-element RHS class (Evolve.H:8-83): the same
// public members and methods.  On the GPU path the six RHS stages are fused
// into one device kernel run by evolve_elem(); see Evolve_element below.
#ifndef EVOLVE_H
#define EVOLVE_H

#include <vector>

#include "Element.H"

class Evolve_element {
public:
    Evolve_element();
    Evolve_element(Element* this_elem, Element* right_elem, Element* left_elem, Element* vertical_elem,
                   const std::vector<std::vector<double>>& gau_int_l, const int& p_parms);

    int p;
    Element* element_this;
    Element* element_right;
    Element* element_left;
    Element* element_vertical;
    std::vector<std::vector<double>> gau_integ_line;

    std::vector<std::vector<double>> plus_phi_in_quadrature_points_side_1;
    std::vector<std::vector<double>> plus_phi_in_quadrature_points_side_2;
    std::vector<std::vector<double>> plus_phi_in_quadrature_points_side_3;
    std::vector<std::vector<double>> minus_phi_in_quadrature_points_side_1;
    std::vector<std::vector<double>> minus_phi_in_quadrature_points_side_2;
    std::vector<std::vector<double>> minus_phi_in_quadrature_points_side_3;

    std::vector<std::vector<double>> U_plus_side_1;
    std::vector<std::vector<double>> U_plus_side_2;
    std::vector<std::vector<double>> U_plus_side_3;
    std::vector<std::vector<double>> U_minus_side_1;
    std::vector<std::vector<double>> U_minus_side_2;
    std::vector<std::vector<double>> U_minus_side_3;

    std::vector<std::vector<double>> numerical_flux_side_1;
    std::vector<std::vector<double>> numerical_flux_side_2;
    std::vector<std::vector<double>> numerical_flux_side_3;

    std::vector<std::vector<double>> DG_numerical_flux_integration;
    std::vector<std::vector<double>> DG_stiffness_vector;
    std::vector<std::vector<double>> DG_residual_vector;
    std::vector<std::vector<double>> DG_time_derivative_U;

    // Host set-up of the edge basis tables, as the reference (Evolve.cpp:76-132).
    void evaluate_basis_in_quadrature_poits();
    // The reference's RHS stages.  On the GPU they run fused, for all elements
    // at once, inside evolve_elem(); the first five are kept for source
    // compatibility and leave their members untouched, and
    // compute_time_derivative_U() fills DG_time_derivative_U from one device
    // evaluation of the whole mesh this element belongs to.
    void compute_U_plus_minus();
    void compute_numerical_flux();
    void integrate_numerical_flux();
    void compute_stiffness_vector();
    void compute_residual_vector();
    void compute_time_derivative_U();
};

#endif  // EVOLVE_H
