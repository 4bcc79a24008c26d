// The luxum executable: the reference's driver (src/solvers/main.cc) on the
// C++ facade.  Reads parameters.yaml from the working directory, runs the
// solver the YAML selects and writes the reference's HDF5 output there.
//   luxum [parameters.yaml] [--print-every N]
#include <cstdlib>
#include <cstring>
#include <iostream>

#include "luxum.hpp"

int main(int argc, char **argv) {
    const char *path = "parameters.yaml";
    int print_every = 1;
    for (int i = 1; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--print-every") && i + 1 < argc) print_every = std::atoi(argv[++i]);
        else path = argv[i];
    }
    try {
        luxum::reader inputData(path);
        luxum::parallel mpi(inputData);
        luxum::grid gridData(inputData, mpi);
        luxum::maxwell solver(inputData, mpi, gridData);
        solver.print_every = print_every;
        solver.solve();
    } catch (const luxum::error &e) {
        std::cerr << "luxum: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
