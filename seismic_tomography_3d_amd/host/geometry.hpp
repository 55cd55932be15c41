// geometry.hpp -- acquisition geometry (src/geometry/geometry.hpp:8-30 and
// src/geometry/regular/regular.hpp).  Regular grids only: every workload of
// this path uses them (circular rings are outside its scope).
#pragma once

#include <string>
#include <vector>

namespace st_host {

class Geometry {
  public:
    int all = 0;
    std::vector<float> x, y, z;
    virtual ~Geometry() = default;
    virtual void set_geometry(const std::string &parameters, const std::string &name) = 0;
};

class Regular : public Geometry {
  public:
    // regular.cpp:29-56: corners {name}_SW/NW/SE, {name}_n_xline/_n_yline,
    // flat elevation or topography file, geometry written to
    // {name}_output_file as "x, y, z" lines.
    void set_geometry(const std::string &parameters, const std::string &name) override;
    // regular.cpp:6-27
    static std::vector<float> linspace(float xi, float xf, int n);
    // regular.cpp:93-100: x[k + j*n_yline] = x_line[j], y[...] = y_line[k]
    void build(float swx, float swy, float nwy, float sex, int n_xline, int n_yline);
};

}  // namespace st_host
