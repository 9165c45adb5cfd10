// panSim command-line driver (src/main.cpp:18-48): parse the reference's
// flags, build the simulation, print the statistics header and one line per
// simulated day, run to the end.
#include <cstdio>
#include <cstring>
#include <iostream>

#include "../../../include/pansim_sim.h"
#include "scenario.h"

int main(int argc, char** argv) {
    for (int i = 1; i < argc; ++i) {
        if (!std::strcmp(argv[i], "-h") || !std::strcmp(argv[i], "--help")) {
            std::cout << ps::options_help();
            return 0;
        }
        if (!std::strcmp(argv[i], "--version")) {
            std::cout << "panSim (MI355X) 0.1\n";
            return 0;
        }
    }
    pansim_sim* sim = nullptr;
    int rc = pansim_sim_create(&sim, argc - 1, (const char* const*)(argv + 1), 1);
    if (rc != 0) {
        std::cerr << pansim_sim_create_error() << '\n';
        return 1;
    }
    rc = pansim_sim_run(sim);
    if (rc != 0) std::cerr << pansim_sim_error(sim) << '\n';
    pansim_sim_destroy(sim);
    return rc == 0 ? 0 : 1;
}
