// classic.hpp -- Classic (src/eikonal/classic/classic.hpp), the Podvin &
// Lecomte (1991) eikonal, as a drop-in Eikonal solved on the device
// bit-identically (st_podvin_solve).
#pragma once

#include "device_eikonal.hpp"

namespace st_host {

class Classic : public DeviceEikonal {
  public:
    Classic();

    // classic.cpp:145-160: one-cell padding (held on the device)
    void prepare_volumes() override;
    // classic.cpp:184-1842 + write_first_arrival's picks for shot_id
    void eikonal_equation() override;

    // all shots, batched on the device: picks[total_shots*total_nodes];
    // returns the device time in ms
    double solve_all_shots(float *picks, int batch = 0);

  protected:
    const char *method() const override { return "Podvin & Lecomte (1991)"; }
};

}  // namespace st_host
