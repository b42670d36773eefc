// Drop-in declaration of the reference-operator interface (Preevolve.H:10-19).
#ifndef PREEVOLVE_H
#define PREEVOLVE_H
#include <vector>

std::vector<std::vector<double>> inverse_mass_matrix_reference_space(const int& p, const std::vector<std::vector<double>>& gauss_area_int);
std::vector<std::vector<std::vector<double>>> sitffness_matrix_reference_space(const int& p, const std::vector<std::vector<double>>& gauss_area_int);

#endif  // PREEVOLVE_H
