// Drop-in declaration of the quadrature interface (Quadraturerule.H:7-10).
#ifndef QUADRATURERULE_H
#define QUADRATURERULE_H
#include <vector>

// [nq][2] = {sigma, weight} on [0,1]
std::vector<std::vector<double>> gauss_line_integral(const int& IntOrder);
// [n][3] = {xi, eta, weight} on the reference triangle
std::vector<std::vector<double>> gauss_area_integral(const int& IntOrder);

#endif  // QUADRATURERULE_H
