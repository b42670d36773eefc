// Drop-in declaration of the element class (Element.H:10-53): the same public
// members, in the same order and types (so code built against either header
// sees one object layout), and the same methods.
#ifndef ELEMENT_H
#define ELEMENT_H

#include <string>
#include <vector>

#include "Meshgeneration.H"
#include "Parameters.H"

class Element {
public:
    Element();
    Element(const int& ele_num, const mesh& mesh_info, const std::vector<std::vector<double>>& nods_ref_spa, const int& p_lagrange);

    double time;
    int p;
    int number;
    int type;
    int right_element;
    int left_element;
    int vertical_element;
    std::vector<std::vector<double>> vertices_coords_phys_space;
    std::vector<std::vector<double>> nods_coords_refe_space;
    std::vector<std::vector<double>> nods_coords_phys_space;

    std::vector<std::vector<double>> hidrodynamics_vector_U;                // [Np][4]
    std::vector<std::vector<std::vector<double>>> hidrodynamics_vector_F;   // [Np][2][4]

    std::vector<std::vector<double>> jacobian;
    double determinant_jacobian;
    std::vector<std::vector<double>> inverse_jacobian;
    double determinant_inverse_jacobian;

    std::vector<std::vector<double>> inverse_mass_matrix_physical_space;
    std::vector<std::vector<std::vector<double>>> stiffness_matrix_physical_space;
    std::vector<std::vector<double>> units_vectors_perpendicular_to_element_boundary;
    std::vector<double> sides_lenght;

    void initialize_hydrodinamics(const int& ini_type, const std::vector<std::vector<double>>& gau_area_int);
    void write_data(const int& step_num);
    void build_mass_matrix_inverse(const std::vector<std::vector<double>>& inv_mass_matrix);
    void build_stiffness_matrix(const std::vector<std::vector<std::vector<double>>>& stiff_matrix);
    void build_jacobians();
};

#endif  // ELEMENT_H
