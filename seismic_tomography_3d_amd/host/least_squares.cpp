// least_squares.cpp -- see tomography.hpp.
#include <cmath>
#include <fstream>
#include <iostream>

#include "file_manager.hpp"
#include "tomography.hpp"

namespace st_host {

void Least_squares::info_message() {
    if (iteration == max_iteration)
        std::cout << "------- Checking final residuo ------------\n";
    else
        std::cout << "------- Computing iteration " << iteration + 1 << " of " << max_iteration
                  << " ------------\n";
    if (iteration > 0) std::cout << "Previous iteration residuo: " << residuo[iteration - 1] << "\n";
}

void Least_squares::set_parameters() {
    // least_squares.cpp:85-104: 0 Classic, 1 Block_FIM, 2 Accurate_FSM,
    // anything else Accurate_FSM
    const int type = std::stoi(catch_parameter("eikonal_type", parameters));
    if (type == 0)
        eikonal.reset(new Classic());
    else if (type == 1)
        eikonal.reset(new Block_FIM());
    else
        eikonal.reset(new Accurate_FSM());
    eikonal->parameters = parameters;
    eikonal->set_parameters();
    n_data = eikonal->total_shots * eikonal->total_nodes;
    max_iteration = std::stoi(catch_parameter("max_iteration", parameters));
    lambda = std::stof(catch_parameter("tk_param", parameters));
    tk_order = (int)std::stof(catch_parameter("tk_order", parameters));
    smooth = str2bool(catch_parameter("smooth", parameters));
    window = std::stoi(catch_parameter("filter_samples", parameters));
    stdv = std::stof(catch_parameter("standard_deviation", parameters));
    dx_tomo = std::stof(catch_parameter("dx_tomo", parameters));
    dy_tomo = std::stof(catch_parameter("dy_tomo", parameters));
    dz_tomo = std::stof(catch_parameter("dz_tomo", parameters));
    const std::string it = catch_parameter("cgls_iterations", parameters);  // optional key
    if (!it.empty()) cg_iterations = std::stoi(it);
    obs_data_folder = catch_parameter("dobs_folder", parameters);
    cal_data_folder = catch_parameter("first_arrival_folder", parameters);
    convergency_folder = catch_parameter("convergency_folder", parameters);
    estimated_model_folder = catch_parameter("estimated_models_folder", parameters);
    iteration = 0;
    dobs.assign(n_data, 0.f);
    dcal.assign(n_data, 0.f);
    dm.assign(eikonal->nPoints, 0.f);
    eikonal->prepare_volumes();
    int nt[3];
    check(st_tomo_set_eikonal(eikonal->context(), type), eikonal->context());
    check(st_tomo_set_grid(eikonal->context(), dx_tomo, dy_tomo, dz_tomo, 0), eikonal->context());
    check(st_tomo_dims(eikonal->context(), &nt[0], &nt[1], &nt[2]), eikonal->context());
    n_model = nt[0] * nt[1] * nt[2];
}

void Least_squares::import_obs_data() {
    const int nr = eikonal->total_nodes;
    for (int s = 0; s < eikonal->total_shots; s++)
        read_binary_float(obs_data_folder + "obsData_" + std::to_string(nr) + "_samples_shot_" +
                              std::to_string(s + 1) + ".bin",
                          dobs.data() + (long)s * nr, nr);
}

void Least_squares::forward_modeling() {
    info_message();
    st_ctx *ctx = eikonal->context();
    const std::vector<float> s = eikonal->source_xyz(0, eikonal->total_shots);
    const std::vector<float> r = eikonal->node_xyz();
    std::vector<float> illum(eikonal->nPoints);
    st_fim_stats st;
    const Block_FIM *fim = dynamic_cast<const Block_FIM *>(eikonal.get());
    check(st_forward(ctx, s.data(), eikonal->total_shots, r.data(), eikonal->total_nodes,
                     dcal.data(), illum.data(), fim ? &fim->opts : nullptr, &st),
          ctx);
    if (write_files) {
        const int nr = eikonal->total_nodes;
        for (int k = 0; k < eikonal->total_shots; k++)
            write_binary_float(cal_data_folder + eikonal->name + "_times_nr" + std::to_string(nr) +
                                   "_shot_" + std::to_string(k + 1) + ".bin",
                               dcal.data() + (long)k * nr, nr);
        write_binary_float("../outputs/illumination/" + eikonal->name + "_illumination_iteration_" +
                               std::to_string(iteration) + ".bin",
                           illum.data(), eikonal->nPoints);
    }
}

void Least_squares::import_cal_data() {
    // the picks of forward_modeling are already in dcal: no disk round trip
}

bool Least_squares::converged() {
    // the reference's sequential FP32 sum of powf(d, 2.0f) (which its
    // optimised build evaluates as d * d); st_residual_norm gives the FP64
    // value of the same norm
    float r = 0.0f;
    for (int i = 0; i < n_data; i++) {
        const float d = dobs[i] - dcal[i];
        r += d * d;
    }
    residuo.push_back(std::sqrt(r));
    if (iteration >= max_iteration) {
        std::cout << "\nFinal residuo: " << std::sqrt(r) << std::endl;
        return true;
    }
    iteration += 1;
    return false;
}

void Least_squares::optimization() {
    std::cout << "Solving linear system using Tikhonov regularization with order " << tk_order
              << "\n";
    check(st_optimize(eikonal->context(), dobs.data(), dcal.data(), n_data, tk_order, lambda,
                      cg_iterations, nullptr),
          eikonal->context());
}

void Least_squares::model_update() {
    std::vector<float> velocity(eikonal->nPoints);
    check(st_model_update(eikonal->context(), nullptr, smooth ? 1 : 0, window, stdv, dm.data(),
                          velocity.data()),
          eikonal->context());
    check(st_get_model(eikonal->context(), eikonal->slowness), eikonal->context());
    if (write_files)
        write_binary_float(estimated_model_folder + eikonal->name + "_model_iteration_" +
                               std::to_string(iteration) + ".bin",
                           velocity.data(), eikonal->nPoints);
}

void Least_squares::export_convergency() {
    std::ofstream f(convergency_folder + eikonal->name + "_convergency.txt");
    for (float r : residuo) f << r << "\n";
}

}  // namespace st_host
