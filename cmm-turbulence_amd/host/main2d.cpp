// cmm2d: command-line entry of the MI355X CMM engine.
//   cmm2d [param_file=<path>] [COMMAND=VALUE ...] [save_params_only=<path>]
// The parameter file is read first, then the command-line values override it
// (Cuda-CMM/src/main2d.cpp:66-105).  save_params_only=<path> writes the resulting parameter file
// and exits without touching the GPU.
#include <iostream>
#include <string>

#include "driver.hpp"
#include "settings.hpp"

int main(int argc, char **argv) {
    cmmhost::Settings s;
    for (int i = 0; i < argc; ++i) {
        std::string a = argv[i];
        size_t eq = a.find('=');
        if (eq != std::string::npos && a.substr(0, eq) == "param_file") cmmhost::load_param_file(s, a.substr(eq + 1));
    }
    s.apply_commands(argc, argv);
    for (int i = 0; i < argc; ++i) {
        std::string a = argv[i];
        if (a.rfind("save_params_only=", 0) == 0) {
            cmmhost::save_param_file(s, a.substr(17));
            return 0;
        }
    }
    // both loops share the driver (Cuda-CMM/src/main2d.cpp:92-95 dispatches to cuda_euler_2d or cuda_vlasov_1d)
    if (s.simulation_type == "cmm_euler_2d" || s.simulation_type == "cmm_vlasov_poisson_1d") return cmmhost::run_euler(s);
    std::cout << "Simulation type not recognized, currently supported:" << std::endl;
    std::cout << "\tcmm_euler_2d" << std::endl;
    std::cout << "\tcmm_vlasov_poisson_1d" << std::endl;
    return 0;
}
