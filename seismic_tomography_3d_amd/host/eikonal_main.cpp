// eikonal_main.cpp -- eikonal driver (src/eikonal/eikonal_main.cpp:10-50):
// the parameter file's eikonal_type picks the solver (0 Podvin & Lecomte,
// 1 block FIM, 2 Accurate_FSM; all on the device).  Usage: st_eikonal parameters.txt
// [--batched] -- with --batched all shots are solved in device batches.
#include <chrono>
#include <cstring>
#include <iostream>
#include <memory>
#include <vector>

#include "accurate_fsm.hpp"
#include "block_fim.hpp"
#include "classic.hpp"
#include "file_manager.hpp"

int main(int argc, char **argv) {
    if (argc < 2) {
        std::cerr << "usage: " << argv[0] << " parameters.txt [--batched]\n";
        return 1;
    }
    const bool batched = argc > 2 && std::strcmp(argv[2], "--batched") == 0;
    try {
        const int type = std::stoi(st_host::catch_parameter("eikonal_type", argv[1]));
        std::unique_ptr<st_host::DeviceEikonal> e;
        if (type == 0)
            e.reset(new st_host::Classic());
        else if (type == 1)
            e.reset(new st_host::Block_FIM());
        else if (type == 2)
            e.reset(new st_host::Accurate_FSM());
        else
            throw std::invalid_argument("eikonal_type " + std::to_string(type) +
                                        ": 0 (Podvin), 1 (block FIM) or 2 (Accurate_FSM)");
        e->parameters = argv[1];
        e->set_parameters();
        e->prepare_volumes();
        const auto ti = std::chrono::steady_clock::now();
        if (batched) {
            std::vector<float> picks((size_t)e->total_shots * e->total_nodes);
            if (type == 1) {
                st_fim_stats st;
                static_cast<st_host::Block_FIM *>(e.get())->solve_all_shots(picks.data(), nullptr,
                                                                            &st);
                std::cout << "rounds " << st.rounds << ", block visits " << st.visits << "\n";
            } else {
                const double ms =
                    type == 2
                        ? static_cast<st_host::Accurate_FSM *>(e.get())->solve_all_shots(picks.data())
                        : static_cast<st_host::Classic *>(e.get())->solve_all_shots(picks.data());
                std::cout << "device time " << ms << " ms\n";
            }
            for (int i = 0; i < e->total_shots; i++) {
                e->shot_id = i;
                std::copy(picks.begin() + (size_t)i * e->total_nodes,
                          picks.begin() + (size_t)(i + 1) * e->total_nodes, e->first_arrival);
                e->write_first_arrival();
            }
        } else {
            for (int i = 0; i < e->total_shots; i++) {
                e->shot_id = i;
                e->info_message();
                e->eikonal_equation();
                e->write_time_volume();
                e->write_first_arrival();
            }
        }
        const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - ti).count();
        std::cout << "Run time = " << s << "s.\n";
        e->destroy_volumes();
    } catch (const std::exception &ex) {
        std::cerr << ex.what() << "\n";
        return 2;
    }
    return 0;
}
