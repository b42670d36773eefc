// Drop-in declaration of the basis interface (Lagrangebasis.H:5-11).
#ifndef LAGRANGEBASIS_H
#define LAGRANGEBASIS_H
#include <vector>

std::vector<double> lagrange_basis_reference_space(const int& p, const std::vector<double>& coords_ref_spa);
std::vector<std::vector<double>> lagrange_basis_gradient_reference_space(const int& p, const std::vector<double>& coords_ref_spa);
std::vector<double> reference_to_physical_space(const std::vector<double>& coord_ref_spa,
                                                const std::vector<std::vector<double>>& vertex_phys_spa);

#endif  // LAGRANGEBASIS_H
