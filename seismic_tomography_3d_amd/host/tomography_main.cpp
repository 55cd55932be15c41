// tomography_main.cpp -- inversion driver (src/tomography/tomography_main.cpp:7-43)
#include <chrono>
#include <iostream>

#include "tomography.hpp"

int main(int argc, char **argv) {
    if (argc < 2) {
        std::cerr << "usage: " << argv[0] << " parameters.txt\n";
        return 1;
    }
    try {
        const auto ti = std::chrono::steady_clock::now();
        st_host::Least_squares t;
        t.parameters = argv[1];
        t.set_parameters();
        t.import_obs_data();
        while (true) {
            t.forward_modeling();
            t.import_cal_data();
            if (t.converged()) break;
            t.optimization();
            t.model_update();
        }
        t.export_convergency();
        const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - ti).count();
        std::cout << "\nTomography run time: " << s << " s." << std::endl;
    } catch (const std::exception &ex) {
        std::cerr << ex.what() << "\n";
        return 2;
    }
    return 0;
}
