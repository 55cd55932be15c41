// tomography.hpp -- inversion interface of the reference
// (src/tomography/tomography.hpp:11-56) and its least-squares implementation
// (src/tomography/least_squares/least_squares.hpp) over the C ABI: the
// forward model, ray tracing, CGLS and model update run on the device.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "accurate_fsm.hpp"
#include "block_fim.hpp"
#include "classic.hpp"

namespace st_host {

class Tomography {
  protected:
    int max_iteration = 0;
    int window = 5;
    float stdv = 2.0f;
    bool smooth = false;
    std::vector<float> dm, dobs, dcal;
    int n_data = 0, n_model = 0;
    std::string obs_data_folder, cal_data_folder, convergency_folder, estimated_model_folder;
    std::vector<float> residuo;
    std::unique_ptr<DeviceEikonal> eikonal;  // chosen by eikonal_type
    virtual void info_message() = 0;

  public:
    int iteration = 0;
    std::string parameters;
    virtual ~Tomography() = default;
    virtual void set_parameters() = 0;
    virtual void import_obs_data() = 0;
    virtual void forward_modeling() = 0;
    virtual void import_cal_data() = 0;
    virtual bool converged() = 0;
    virtual void optimization() = 0;
    virtual void model_update() = 0;
    virtual void export_convergency() = 0;
};

class Least_squares : public Tomography {
  public:
    void set_parameters() override;     // least_squares.cpp:83-142 (solver: :85-104)
    void import_obs_data() override;    // :144-161
    void forward_modeling() override;   // :163-190 (+ ray tracing :192-280)
    void import_cal_data() override;    // :282-299 (picks stay in memory)
    bool converged() override;          // :301-319 (the reference's FP32 sum)
    void optimization() override;       // :321-356 (+ tk_reg_matrix, sparse_cgls)
    void model_update() override;       // :498-570
    void export_convergency() override; // :572-579
    const std::vector<float> &residuals() const { return residuo; }
    const std::vector<float> &calculated() const { return dcal; }

  private:
    void info_message() override;
    int tk_order = 2;
    float lambda = 0.f;
    int cg_iterations = 10;  // least_squares.cpp:427
    float dx_tomo = 0, dy_tomo = 0, dz_tomo = 0;
    bool write_files = true;
};

}  // namespace st_host
