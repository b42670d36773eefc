// Drop-in declaration of the time integration interface (Timestepping.H:9-10
// and evolve_elem of Timestepping.cpp:18).  All three run on the GPU.
#ifndef TIMESTEPPING_H
#define TIMESTEPPING_H

#include "Element.H"
#include "Evolve.H"

void forward_euler(Element* elemts, Evolve_element* evo_elemts, double& t_step, const int& n_elements);
void rk4(Element* elemts, Evolve_element* evo_elemts, double& t_step, const int& n_elements);
void evolve_elem(Evolve_element* evo_elemts, int n_elements);

// Resident stepping (no counterpart in the reference): nsteps steps of
// `method` (0 Euler, 1 RK4) with the state kept in HBM and copied back to the
// elements once at the end; used by the driver between output steps.
void gpu_steps(Element* elemts, Evolve_element* evo_elemts, int method, double t_step, int nsteps, const int& n_elements);

#endif  // TIMESTEPPING_H
