// block_fim.hpp -- Block_FIM (src/eikonal/block_FIM/block_FIM.hpp:11-39) as
// a drop-in Eikonal over the C ABI of include/st_api.h.
#pragma once

#include "device_eikonal.hpp"

namespace st_host {

class Block_FIM : public DeviceEikonal {
  public:
    Block_FIM();

    // block_FIM.cpp:142-188: the padded block layout lives on the device
    void prepare_volumes() override;
    // block_FIM.cpp:212-220 for shot_id: travel_time (z fastest) and the
    // receiver picks in first_arrival
    void eikonal_equation() override;

    // all shots in batched persistent solves: picks[total_shots*total_nodes]
    void solve_all_shots(float *picks, const st_fim_opts *opts, st_fim_stats *stats);

    st_fim_opts opts;

  protected:
    const char *method() const override { return "block FIM"; }
};

}  // namespace st_host
