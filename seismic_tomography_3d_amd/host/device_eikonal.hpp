// device_eikonal.hpp -- what the reference's device-backed solvers share:
// the parameter-file set-up (block_FIM.cpp:71-140 and accurate_FSM.cpp:
// 74-143 are the same code), the geometry helpers and the output files,
// over one C-ABI context (include/st_api.h).
#pragma once

#include <vector>

#include "eikonal.hpp"
#include "st_api.h"

namespace st_host {

void check(int status, st_ctx *ctx);

class DeviceEikonal : public Eikonal {
  public:
    ~DeviceEikonal() override;

    // model, geometry, reciprocity, output flags
    void set_parameters() override;
    void info_message() override;
    void write_time_volume() override;    // block_FIM.cpp:222-226
    void write_first_arrival() override;  // :228-262 (file only; picks come from the device)
    void destroy_volumes() override;

    st_ctx *context() { return ctx_; }
    std::vector<float> source_xyz(int first, int count) const;
    std::vector<float> node_xyz() const;

  protected:
    // uploads the slowness (again after a model update)
    void upload_model();
    virtual const char *method() const = 0;
    st_ctx *ctx_ = nullptr;
    std::vector<float> slow_, times_, picks_;
};

}  // namespace st_host
