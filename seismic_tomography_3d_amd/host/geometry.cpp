// geometry.cpp -- see geometry.hpp.
#include "geometry.hpp"

#include <fstream>
#include <iostream>
#include <stdexcept>

#include "file_manager.hpp"

namespace st_host {

std::vector<float> Regular::linspace(float xi, float xf, int n) {
    std::vector<float> v;
    if (n <= 0) return v;
    if (n == 1) {
        v.push_back(xi);
        return v;
    }
    v.reserve(n);
    const float delta = (xf - xi) / (n - 1);
    for (int i = 0; i < n; i++) v.push_back(xi + (float)(delta * i));
    return v;
}

void Regular::build(float swx, float swy, float nwy, float sex, int n_xline, int n_yline) {
    all = n_xline * n_yline;
    x.assign(all, 0.f);
    y.assign(all, 0.f);
    z.assign(all, 0.f);
    const std::vector<float> xt = linspace(swx, sex, n_xline);
    const std::vector<float> yt = linspace(swy, nwy, n_yline);
    for (int k = 0; k < (int)yt.size(); k++)
        for (int j = 0; j < (int)xt.size(); j++) {
            x[k + j * n_yline] = xt[j];
            y[k + j * n_yline] = yt[k];
        }
}

void Regular::set_geometry(const std::string &p, const std::string &name) {
    const float elevation = std::stof(catch_parameter(name + "_elevation", p));
    const bool topography = str2bool(catch_parameter(name + "_topography", p));
    const std::string topo_file = catch_parameter(name + "_topography_file", p);
    const std::string output_file = catch_parameter(name + "_output_file", p);
    const int n_xline = std::stoi(catch_parameter(name + "_n_xline", p));
    const int n_yline = std::stoi(catch_parameter(name + "_n_yline", p));
    auto pair = [&](const std::string &key) {
        const std::vector<std::string> s = split(catch_parameter(name + key, p), ',');
        if (s.size() < 2) throw std::invalid_argument("bad coordinate pair " + name + key);
        return std::make_pair(std::stof(s[0]), std::stof(s[1]));
    };
    const auto SW = pair("_SW"), NW = pair("_NW"), SE = pair("_SE");
    build(SW.first, SW.second, NW.second, SE.first, n_xline, n_yline);
    if (topography) {
        std::vector<std::string> t;
        read_text_file(topo_file, t);
        if ((int)t.size() < all) throw std::invalid_argument("topography file too short");
        for (int i = 0; i < all; i++) z[i] = std::stof(t[i]);
    } else {
        for (int i = 0; i < all; i++) z[i] = elevation;
    }
    if (!output_file.empty()) {
        std::ofstream f(output_file);
        if (!f.is_open()) throw std::invalid_argument("Error: file could not be opened: " + output_file);
        for (int i = 0; i < all; i++) f << x[i] << ", " << y[i] << ", " << z[i] << "\n";
        std::cout << "Geometry file " << output_file << " was succesfully written." << std::endl;
    }
}

}  // namespace st_host

// C entry for tests: Regular::set_geometry for prefix name, returns `all`
extern "C" int sth_regular(const char *parameters, const char *name, float *x, float *y, float *z,
                           int cap) {
    st_host::Regular g;
    g.set_geometry(parameters, name);
    for (int i = 0; i < g.all && i < cap; i++) {
        x[i] = g.x[i];
        y[i] = g.y[i];
        z[i] = g.z[i];
    }
    return g.all;
}
