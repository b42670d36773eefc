// Drop-in declaration of the reference's input parameters (Parameters.H:5-22);
// same field names, types and meaning.
#ifndef PARAMETERS_H_
#define PARAMETERS_H_

struct parameters {
    int p;
    float domain_x;
    float domain_y;
    int num_element_in_x;
    int num_element_in_y;
    float perturb_grid;
    int integration_order;
    float simulation_time;
    int number_time_steps;
    double time_step;
    int write_every_steps;
    int U_initialization_type;
    int stepping_method;
};

// key=value input file -> parameters (Parameters.cpp:12-64); exits with the
// reference's messages on a wrong argument count or an unreadable file.
parameters read_input_files(int argc, char* argv[]);

#endif  // PARAMETERS_H_
