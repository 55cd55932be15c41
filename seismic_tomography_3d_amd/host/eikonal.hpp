// eikonal.hpp -- the solver plugin interface of the reference
// (src/eikonal/eikonal.hpp:15-58): same public fields and virtual methods,
// so a caller of the reference's Eikonal* drives this build unchanged.
#pragma once

#include <memory>
#include <string>

#include "geometry.hpp"

namespace st_host {

class Eikonal {
  protected:
    bool reciprocity = false;
    bool export_time_volume = false;
    bool export_first_arrival = false;
    std::string time_volume_folder;
    std::string first_arrival_folder;

  public:
    float dx = 0, dy = 0, dz = 0;
    int nx = 0, ny = 0, nz = 0, nPoints = 0;
    int nxx = 0, nyy = 0, nzz = 0, nPointsB = 0;

    float *slowness = nullptr;
    float *travel_time = nullptr;
    float *first_arrival = nullptr;

    Geometry *shots = nullptr;
    Geometry *nodes = nullptr;

    std::string name;
    std::string parameters;

    int shot_id = 0;
    int total_shots = 0;
    int total_nodes = 0;

    virtual ~Eikonal() = default;
    virtual void set_parameters() = 0;
    virtual void prepare_volumes() = 0;
    virtual void info_message() = 0;
    virtual void eikonal_equation() = 0;
    virtual void write_time_volume() = 0;
    virtual void write_first_arrival() = 0;
    virtual void destroy_volumes() = 0;
};

}  // namespace st_host
