// Drop-in declaration of the output helpers (Utilities.H:8-16).
#ifndef UTILITIES_H
#define UTILITIES_H
#include <string>
#include <vector>

void writeToFile(const std::string& filename, const std::vector<std::string>& lines);
void clean_create_directory(const std::string& dirname);

class Element;
void write_output(Element* elements, const int& total_num_elements, const int& step_num);

#endif  // UTILITIES_H
