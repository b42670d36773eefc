// Drop-in declaration of the mesh interface (Meshgeneration.H:7-17).
#ifndef MESHGENERATION_H
#define MESHGENERATION_H
#include <vector>

#include "Parameters.H"

struct mesh {
    std::vector<int> element_number;
    std::vector<std::vector<std::vector<double>>> element_coordinates;  // [n][3][2]
    std::vector<std::vector<int>> elements_at_boundary;                 // [n][3]: right, left, vertical
    std::vector<int> element_type;                                      // [n]: 0 or 1
};

// Structured periodic mesh with the reference connectivity; also writes the
// grid/ text files the reference writes (Meshgeneration.cpp:160-179).
mesh generate_mesh(const parameters& parms);

std::vector<std::vector<double>> generate_nodes_reference_space(const parameters& parms);

#endif  // MESHGENERATION_H
