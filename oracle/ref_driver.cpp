// ref_driver.cpp -- C entry points over the reference's OWN CPU code
// (accurate_FSM.cpp, classic.cpp, trilinear.cpp, gaussian.cpp, regular.cpp,
// file_manager.cpp), compiled where it lies under the reference checkout by
// oracle/build_ref.sh into oracle/_ref/.  Test infrastructure only: it pins
// the C restatement in oracle.c and can serve as the CPU baseline.
//
// It replaces nothing of the reference: it includes the reference headers,
// instantiates the reference classes and calls their public methods in the
// order eikonal_main.cpp:24-41 does (minus Block_FIM, whose CUDA cannot be
// built here).

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <string>
#include <thread>
#include <vector>

#include "eikonal/accurate_FSM/accurate_FSM.hpp"
#include "eikonal/classic/classic.hpp"
#include "geometry/regular/regular.hpp"
#include "utils/file_manager/file_manager.hpp"
#include "utils/interpolation/trilinear.hpp"
#include "utils/smoothing/gaussian.hpp"
#include "utils/wave/acoustic.hpp"

namespace {

// Derived classes reach the protected output flags that set_parameters()
// would otherwise read from a parameter file.
template <class Base> struct Driven : public Base {
    void outputs(const std::string &folder) {
        this->export_time_volume = true;
        this->export_first_arrival = true;
        this->time_volume_folder = folder;
        this->first_arrival_folder = folder;
        this->reciprocity = false;
    }
};

Geometry *make_points(const float *x, const float *y, const float *z, int n) {
    Regular *g = new Regular();
    g->all = n;
    g->x = new float[n];
    g->y = new float[n];
    g->z = new float[n];
    for (int i = 0; i < n; i++) { g->x[i] = x[i]; g->y[i] = y[i]; g->z[i] = z[i]; }
    return g;
}

void free_points(Geometry *g) {
    delete[] g->x; delete[] g->y; delete[] g->z;
    delete g;
}

template <class Solver>
int solve_one(const float *slowness, int nx, int ny, int nz, float dx, float dy,
              float dz, const float *sxyz, int nshot, const float *rx,
              const float *ry, const float *rz, int nr, float *T_out,
              float *picks_out, const char *folder) {
    Driven<Solver> e;
    e.outputs(folder ? std::string(folder) : std::string("./"));
    e.nx = nx; e.ny = ny; e.nz = nz; e.nPoints = nx * ny * nz;
    e.dx = dx; e.dy = dy; e.dz = dz;
    e.slowness = new float[e.nPoints];
    std::memcpy(e.slowness, slowness, sizeof(float) * e.nPoints);
    std::vector<float> sx(nshot), sy(nshot), sz(nshot);
    for (int s = 0; s < nshot; s++) { sx[s] = sxyz[3 * s]; sy[s] = sxyz[3 * s + 1]; sz[s] = sxyz[3 * s + 2]; }
    e.shots = make_points(sx.data(), sy.data(), sz.data(), nshot);
    e.nodes = make_points(rx, ry, rz, nr);
    e.total_shots = nshot;
    e.total_nodes = nr;
    e.travel_time = new float[e.nPoints]();
    e.first_arrival = new float[nr]();
    e.prepare_volumes();
    for (int s = 0; s < nshot; s++) {
        e.shot_id = s;
        e.eikonal_equation();
        if (T_out) std::memcpy(T_out + (long)s * e.nPoints, e.travel_time, sizeof(float) * e.nPoints);
        if (picks_out) {
            if (folder) {
                e.write_first_arrival();  // the reference's own pick + file
                read_binary_float(std::string(folder) + e.name + "_times_nr" + std::to_string(nr) +
                                      "_shot_" + std::to_string(s + 1) + ".bin",
                                  picks_out + (long)s * nr, nr);
            } else {
                for (int r = 0; r < nr; r++) {
                    float x = rx[r], y = ry[r], z = rz[r];
                    float x0 = floorf(x / dx) * dx, y0 = floorf(y / dy) * dy, z0 = floorf(z / dz) * dz;
                    float x1 = x0 + dx, y1 = y0 + dy, z1 = z0 + dz;
                    int id = ((int)(z / dz)) + ((int)(x / dx)) * nz + ((int)(y / dy)) * nx * nz;
                    const float *t = e.travel_time;
                    picks_out[(long)s * nr + r] = trilinear(t[id], t[id + 1], t[id + nz], t[id + 1 + nz],
                                                            t[id + nx * nz], t[id + 1 + nx * nz],
                                                            t[id + nz + nx * nz], t[id + 1 + nz + nx * nz],
                                                            x0, x1, y0, y1, z0, z1, x, y, z);
                }
            }
        }
    }
    e.destroy_volumes();
    delete[] e.slowness;
    delete[] e.travel_time;
    delete[] e.first_arrival;
    free_points(e.shots);
    free_points(e.nodes);
    return 0;
}

}  // namespace

extern "C" {

// type 0 = Classic (Podvin & Lecomte), 2 = Accurate_FSM (eikonal_main.cpp:15-22).
// folder != NULL routes picks through the reference's write_first_arrival().
int ref_eikonal(int type, const float *slowness, int nx, int ny, int nz, float dx,
                float dy, float dz, const float *sxyz, int nshot, const float *rx,
                const float *ry, const float *rz, int nr, float *T_out,
                float *picks_out, const char *folder) {
    if (type == 2)
        return solve_one<Accurate_FSM>(slowness, nx, ny, nz, dx, dy, dz, sxyz, nshot, rx, ry, rz,
                                       nr, T_out, picks_out, folder);
    if (type == 0)
        return solve_one<Classic>(slowness, nx, ny, nz, dx, dy, dz, sxyz, nshot, rx, ry, rz, nr,
                                  T_out, picks_out, folder);
    return -1;
}

// CPU-baseline timing: shots of the reference Accurate_FSM spread over
// nthreads host threads, one solver instance per thread.
int ref_fsm_threads(const float *slowness, int nx, int ny, int nz, float dx, float dy,
                    float dz, const float *sxyz, int nshot, const float *rx, const float *ry,
                    const float *rz, int nr, float *picks_out, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; t++) {
        pool.emplace_back([=]() {
            for (int s = t; s < nshot; s += nthreads)
                solve_one<Accurate_FSM>(slowness, nx, ny, nz, dx, dy, dz, sxyz + 3 * s, 1, rx, ry,
                                        rz, nr, nullptr, picks_out + (long)s * nr, nullptr);
        });
    }
    for (auto &th : pool) th.join();
    return 0;
}

// Full reference parameter-file path for one solver type: set_parameters
// (catch_parameter, Regular geometry, reciprocity, model read), then the
// shot loop with the reference's own file outputs.
int ref_run_params(const char *param_file, int type, int *nshot, int *nnode) {
    Eikonal *e = nullptr;
    if (type == 2) e = new Accurate_FSM();
    else if (type == 0) e = new Classic();
    else return -1;
    e->parameters = std::string(param_file);
    e->set_parameters();
    e->prepare_volumes();
    for (int i = 0; i < e->total_shots; i++) {
        e->shot_id = i;
        e->eikonal_equation();
        e->write_time_volume();
        e->write_first_arrival();
    }
    *nshot = e->total_shots;
    *nnode = e->total_nodes;
    e->destroy_volumes();
    return 0;
}

float ref_trilinear(float c000, float c001, float c100, float c101, float c010, float c011,
                    float c110, float c111, float x0, float x1, float y0, float y1, float z0,
                    float z1, float x, float y, float z) {
    return trilinear(c000, c001, c100, c101, c010, c011, c110, c111, x0, x1, y0, y1, z0, z1, x,
                     y, z);
}

void ref_gaussian(float *in, float *out, int nx, int ny, int nz, int samples, float stdv) {
    gaussian(in, out, nx, ny, nz, samples, stdv);
}

// catch_parameter (file_manager.cpp:74-120); returns the value length.
int ref_catch_parameter(const char *target, const char *file, char *out, int cap) {
    std::string v = catch_parameter(std::string(target), std::string(file));
    int n = (int)v.size();
    if (n >= cap) n = cap - 1;
    std::memcpy(out, v.data(), n);
    out[n] = 0;
    return (int)v.size();
}

// Regular::set_geometry (regular.cpp:29-56) for prefix "shots"/"nodes".
int ref_regular(const char *param_file, const char *name, float *x, float *y, float *z, int cap) {
    Regular g;
    g.set_geometry(std::string(param_file), std::string(name));
    int n = g.all < cap ? g.all : cap;
    for (int i = 0; i < n; i++) { x[i] = g.x[i]; y[i] = g.y[i]; z[i] = g.z[i]; }
    int all = g.all;
    delete[] g.x; delete[] g.y; delete[] g.z;
    return all;
}

// One shot of the acoustic modelling loop of wave_main.cpp:129-166 on caller
// arrays, through the reference's own acoustic.cpp functions: expand (nb
// replicated cells), dampers (Cerjan factor), then per time step
// applyWavelet, wavePropagation, wavefieldUpdate, buildSeismogram.  Receiver
// and source grid indices as wave_main.cpp:105-110,135-139.  seismogram is
// [nrec][nt] (time fastest).
int ref_acoustic_shot(const float *vel, int nx, int ny, int nz, float dx, float dy, float dz,
                      int nb, float factor, const float *wavelet, int nt, float dt, float sx,
                      float sy, float sz, const float *rx, const float *ry, const float *rz,
                      int nrec, float *seismogram, float *ufut) {
    const int nxx = nx + 2 * nb, nyy = ny + 2 * nb, nzz = nz + 2 * nb;
    const int nPoints = nxx * nyy * nzz;
    float *V = expand(const_cast<float *>(vel), nx, ny, nz, nb);
    std::vector<float> U_pas(nPoints), U_pre(nPoints), U_fut(nPoints);
    std::vector<float> d1(nb, 0.f), d2(nb * nb, 0.f), d3(nb * nb * nb, 0.f);
    dampers(d1.data(), d2.data(), d3.data(), factor, nb);
    std::vector<int> rIdx(nrec), rIdy(nrec), rIdz(nrec);
    for (int r = 0; r < nrec; r++) {
        rIdx[r] = (int)(rx[r] / dx) + nb;
        rIdy[r] = (int)(ry[r] / dy) + nb;
        rIdz[r] = (int)(rz[r] / dz) + nb;
    }
    const int sId = ((int)(sz / dz) + nb) + ((int)(sx / dx) + nb) * nzz + ((int)(sy / dy) + nb) * nzz * nxx;
    std::vector<float> w(wavelet, wavelet + nt);
    setWaveField(U_pas.data(), U_pre.data(), U_fut.data(), nPoints);
    for (int t = 0; t < nt; t++) {
        applyWavelet(U_pre.data(), w.data(), t, sId);
        wavePropagation(V, U_pas.data(), U_pre.data(), U_fut.data(), d1.data(), d2.data(), d3.data(),
                        nb, nxx, nyy, nzz, dx, dy, dz, dt);
        wavefieldUpdate(U_pas.data(), U_pre.data(), U_fut.data(), nPoints);
        buildSeismogram(U_fut.data(), nxx, nyy, nzz, seismogram, t, nt, rIdx.data(), rIdy.data(),
                        rIdz.data(), nrec);
    }
    if (ufut) std::copy(U_fut.begin(), U_fut.end(), ufut);
    delete[] V;
    return 0;
}

// The reference's Ricker wavelet (acoustic.cpp:81-98).
void ref_ricker(float delay, float fmax, float dt, int nt, float *out) {
    float *r = rickerGeneration(delay, fmax, dt, nt);
    std::copy(r, r + nt, out);
    delete[] r;
}

// The reference's damping profiles (acoustic.cpp:100-137).
void ref_dampers(float *d1, float *d2, float *d3, float factor, int nb) {
    std::fill(d1, d1 + nb, 0.f);
    std::fill(d2, d2 + nb * nb, 0.f);
    std::fill(d3, d3 + nb * nb * nb, 0.f);
    dampers(d1, d2, d3, factor, nb);
}

}  // extern "C"
