// device_eikonal.cpp -- see device_eikonal.hpp.
#include "device_eikonal.hpp"

#include <iostream>
#include <stdexcept>
#include <utility>

#include "file_manager.hpp"

namespace st_host {

void check(int status, st_ctx *ctx) {
    if (status != ST_OK)
        throw std::runtime_error(std::string("st status ") + std::to_string(status) + ": " +
                                 st_last_error(ctx));
}

DeviceEikonal::~DeviceEikonal() { destroy_volumes(); }

void DeviceEikonal::set_parameters() {
    nx = std::stoi(catch_parameter("x_samples", parameters));
    ny = std::stoi(catch_parameter("y_samples", parameters));
    nz = std::stoi(catch_parameter("z_samples", parameters));
    nPoints = nx * ny * nz;
    dx = std::stof(catch_parameter("x_spacing", parameters));
    dy = std::stof(catch_parameter("y_spacing", parameters));
    dz = std::stof(catch_parameter("z_spacing", parameters));
    slow_.assign(nPoints, 0.f);
    slowness = slow_.data();
    read_binary_float(catch_parameter("vp_file", parameters), slowness, nPoints);
    for (int i = 0; i < nPoints; i++) slowness[i] = 1.0f / slowness[i];

    const int st = std::stoi(catch_parameter("shots_geometry_type", parameters));
    const int nt = std::stoi(catch_parameter("nodes_geometry_type", parameters));
    if (st != 0 || nt != 0)
        throw std::invalid_argument("only regular geometry (type 0) is supported");
    shots = new Regular();
    nodes = new Regular();
    shots->set_geometry(parameters, "shots");
    nodes->set_geometry(parameters, "nodes");
    reciprocity = str2bool(catch_parameter("reciprocity", parameters));
    if (reciprocity) {  // block_FIM.cpp:118-127
        std::swap(shots->x, nodes->x);
        std::swap(shots->y, nodes->y);
        std::swap(shots->z, nodes->z);
        std::swap(shots->all, nodes->all);
    }
    total_shots = shots->all;
    total_nodes = nodes->all;
    export_time_volume = str2bool(catch_parameter("export_time_volume", parameters));
    export_first_arrival = str2bool(catch_parameter("export_first_arrival", parameters));
    time_volume_folder = catch_parameter("time_volume_folder", parameters);
    first_arrival_folder = catch_parameter("first_arrival_folder", parameters);
    times_.assign(nPoints, 0.f);
    travel_time = times_.data();
    picks_.assign(total_nodes, 0.f);
    first_arrival = picks_.data();
}

void DeviceEikonal::upload_model() {
    if (!ctx_) check(st_ctx_create(0, &ctx_), nullptr);
    check(st_set_model(ctx_, slowness, nx, ny, nz, dx, dy, dz), ctx_);
}

void DeviceEikonal::info_message() {
    std::cout << "3D eikonal equation solver (" << method() << ", on the GPU)\n"
              << "Shot " << shot_id + 1 << " of " << total_shots << " at ("
              << shots->x[shot_id] << ", " << shots->y[shot_id] << ", " << shots->z[shot_id]
              << ") m\n";
}

std::vector<float> DeviceEikonal::source_xyz(int first, int count) const {
    std::vector<float> s(3 * count);
    for (int i = 0; i < count; i++) {
        s[3 * i] = shots->x[first + i];
        s[3 * i + 1] = shots->y[first + i];
        s[3 * i + 2] = shots->z[first + i];
    }
    return s;
}

std::vector<float> DeviceEikonal::node_xyz() const {
    std::vector<float> r(3 * total_nodes);
    for (int i = 0; i < total_nodes; i++) {
        r[3 * i] = nodes->x[i];
        r[3 * i + 1] = nodes->y[i];
        r[3 * i + 2] = nodes->z[i];
    }
    return r;
}

void DeviceEikonal::write_time_volume() {
    if (export_time_volume)
        write_binary_float(time_volume_folder + name + "_eikonal_nz" + std::to_string(nz) + "_nx" +
                               std::to_string(nx) + "_ny" + std::to_string(ny) + "_shot_" +
                               std::to_string(shot_id + 1) + ".bin",
                           travel_time, nPoints);
}

void DeviceEikonal::write_first_arrival() {
    if (export_first_arrival)
        write_binary_float(first_arrival_folder + name + "_times_nr" + std::to_string(total_nodes) +
                               "_shot_" + std::to_string(shot_id + 1) + ".bin",
                           first_arrival, total_nodes);
}

void DeviceEikonal::destroy_volumes() {
    if (ctx_) st_ctx_destroy(ctx_);
    ctx_ = nullptr;
    delete shots;
    delete nodes;
    shots = nodes = nullptr;
}

}  // namespace st_host
