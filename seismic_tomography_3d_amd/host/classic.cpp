// classic.cpp -- see classic.hpp.
#include "classic.hpp"

namespace st_host {

Classic::Classic() { name = "pod"; }

void Classic::prepare_volumes() {
    nxx = nx + 2;
    nyy = ny + 2;
    nzz = nz + 2;
    nPointsB = nxx * nyy * nzz;
    upload_model();
}

void Classic::eikonal_equation() {
    const std::vector<float> s = source_xyz(shot_id, 1);
    const std::vector<float> r = node_xyz();
    check(st_podvin_solve(ctx_, s.data(), 1, r.data(), total_nodes, first_arrival, travel_time, 1,
                          nullptr),
          ctx_);
}

double Classic::solve_all_shots(float *picks, int batch) {
    const std::vector<float> s = source_xyz(0, total_shots);
    const std::vector<float> r = node_xyz();
    double ms = 0.0;
    check(st_podvin_solve(ctx_, s.data(), total_shots, r.data(), total_nodes, picks, nullptr,
                          batch, &ms),
          ctx_);
    return ms;
}

}  // namespace st_host
