// accurate_fsm.hpp -- Accurate_FSM (src/eikonal/accurate_FSM/accurate_FSM.hpp)
// as a drop-in Eikonal: the reference's default CPU eikonal, solved on the
// device bit-identically (st_fsm_solve).
#pragma once

#include "device_eikonal.hpp"

namespace st_host {

class Accurate_FSM : public DeviceEikonal {
  public:
    Accurate_FSM();

    // accurate_FSM.cpp:145-156: one-cell padding (held on the device)
    void prepare_volumes() override;
    // accurate_FSM.cpp:577-650 + write_first_arrival's picks for shot_id
    void eikonal_equation() override;

    // all shots, batched on the device: picks[total_shots*total_nodes];
    // returns the device time in ms
    double solve_all_shots(float *picks, int batch = 0);

  protected:
    const char *method() const override { return "Noble, Gesret and Belayouni (2014) FSM"; }
};

}  // namespace st_host
