// eikonal_main.cpp -- eikonal driver (src/eikonal/eikonal_main.cpp:10-50)
// on the device FIM.  Usage: st_eikonal parameters.txt [--batched]
#include <chrono>
#include <cstring>
#include <iostream>
#include <vector>

#include "block_fim.hpp"
#include "file_manager.hpp"

int main(int argc, char **argv) {
    if (argc < 2) {
        std::cerr << "usage: " << argv[0] << " parameters.txt [--batched]\n";
        return 1;
    }
    const bool batched = argc > 2 && std::strcmp(argv[2], "--batched") == 0;
    try {
        st_host::Block_FIM e;
        e.parameters = argv[1];
        e.set_parameters();
        e.prepare_volumes();
        const auto ti = std::chrono::steady_clock::now();
        if (batched) {
            std::vector<float> picks((size_t)e.total_shots * e.total_nodes);
            st_fim_stats st;
            e.solve_all_shots(picks.data(), nullptr, &st);
            for (int i = 0; i < e.total_shots; i++) {
                e.shot_id = i;
                std::copy(picks.begin() + (size_t)i * e.total_nodes,
                          picks.begin() + (size_t)(i + 1) * e.total_nodes, e.first_arrival);
                e.write_first_arrival();
            }
            std::cout << "rounds " << st.rounds << ", block visits " << st.visits << "\n";
        } else {
            for (int i = 0; i < e.total_shots; i++) {
                e.shot_id = i;
                e.info_message();
                e.eikonal_equation();
                e.write_time_volume();
                e.write_first_arrival();
            }
        }
        const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - ti).count();
        std::cout << "Run time = " << s << "s.\n";
        e.destroy_volumes();
    } catch (const std::exception &ex) {
        std::cerr << ex.what() << "\n";
        return 2;
    }
    return 0;
}
