// block_fim.cpp -- see block_fim.hpp.
#include "block_fim.hpp"

namespace st_host {

Block_FIM::Block_FIM() {
    name = "fim";
    st_fim_default_opts(&opts);
}

void Block_FIM::prepare_volumes() {
    const int pad = 4;
    nxx = nx + (pad - nx % pad) % pad;
    nyy = ny + (pad - ny % pad) % pad;
    nzz = nz + (pad - nz % pad) % pad;
    nPointsB = nxx * nyy * nzz;
    upload_model();
}

void Block_FIM::eikonal_equation() {
    const std::vector<float> s = source_xyz(shot_id, 1);
    const std::vector<float> r = node_xyz();
    st_fim_opts o = opts;
    o.batch = 1;
    check(st_fim_solve(ctx_, s.data(), 1, r.data(), total_nodes, first_arrival, travel_time, &o,
                       nullptr),
          ctx_);
}

void Block_FIM::solve_all_shots(float *picks, const st_fim_opts *o, st_fim_stats *stats) {
    const std::vector<float> s = source_xyz(0, total_shots);
    const std::vector<float> r = node_xyz();
    check(st_fim_solve(ctx_, s.data(), total_shots, r.data(), total_nodes, picks, nullptr,
                       o ? o : &opts, stats),
          ctx_);
}

}  // namespace st_host
