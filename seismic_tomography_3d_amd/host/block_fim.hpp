// block_fim.hpp -- Block_FIM (src/eikonal/block_FIM/block_FIM.hpp:11-39) as
// a drop-in Eikonal over the C ABI of include/st_api.h.
#pragma once

#include <vector>

#include "eikonal.hpp"
#include "st_api.h"

namespace st_host {

class Block_FIM : public Eikonal {
  public:
    Block_FIM();
    ~Block_FIM() override;

    // block_FIM.cpp:71-140: model, geometry, reciprocity, output flags
    void set_parameters() override;
    // block_FIM.cpp:142-188: the padded block layout lives on the device;
    // this uploads the slowness (again after a model update)
    void prepare_volumes() override;
    void info_message() override;
    // block_FIM.cpp:212-220 for shot_id: travel_time (z fastest) and the
    // receiver picks in first_arrival
    void eikonal_equation() override;
    void write_time_volume() override;    // :222-226
    void write_first_arrival() override;  // :228-262 (file only; picks are computed on the device)
    void destroy_volumes() override;

    // all shots in batched persistent solves: picks[total_shots*total_nodes]
    void solve_all_shots(float *picks, const st_fim_opts *opts, st_fim_stats *stats);

    st_ctx *context() { return ctx_; }
    std::vector<float> source_xyz(int first, int count) const;
    std::vector<float> node_xyz() const;
    st_fim_opts opts;

  private:
    st_ctx *ctx_ = nullptr;
    std::vector<float> slow_, times_, picks_;
};

void check(int status, st_ctx *ctx);

}  // namespace st_host
