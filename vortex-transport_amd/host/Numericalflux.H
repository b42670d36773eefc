// Drop-in declaration of the Roe flux (Numericalflux.H:7).  Evaluated on the
// GPU; exits on a non-physical state with the reference's message.
#ifndef NUMERICALFLUX_H
#define NUMERICALFLUX_H
#include <vector>

void numerical_flux(const std::vector<double>& u_left, const std::vector<double>& u_right,
                    const std::vector<double>& normal_vector, std::vector<double>& flux);

#endif  // NUMERICALFLUX_H
