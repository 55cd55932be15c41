// wave.hip -- data synthesis on the device: the acoustic finite-difference
// modelling of src/utils/wave (acoustic.cpp, wave_main.cpp) and the
// first-break auto-picker of src/utils/auto_picking/auto_picking.cpp.
//
// Results are bit-identical to the reference's serial CPU build (the file is
// compiled without FMA contraction, with IEEE division and denormals):
//
// * The reference's propagation loop damps U_pre and U_pas in place while it
//   sweeps the cells in index order (z fastest, then x, then y), so a cell
//   reads its neighbours at a lower index (i-1..i-4, j-1..j-4, k-1..k-4)
//   already multiplied by their damper and the others not yet.  The device
//   step reads the undamped field and applies the damper to exactly those
//   twelve neighbours; U_pas is d * (U_pre of the previous step), formed
//   when read.  Cells outside the stencil interior (the 4-cell shell) are
//   never written and hold +0, so multiplying them by a damper changes
//   nothing.
// * The wavelet is added when a value is loaded (U_pre at step t carries
//   wavelet[t] at the source cell, U_pas the damped wavelet[t-1]); the
//   fields in memory never hold it, so the recorded sample of a step is the
//   value the reference's buildSeismogram reads (its U_fut, before the next
//   injection).
// * powf(v, 2.0f) is v * v in the reference's optimised builds (gcc and
//   clang fold it at -O1 and above; the libm call of an -O0 build rounds
//   some exact ties the other way), and is squared here.  expf is a true
//   libm call: the dampers are made on the host.
//
// Layout in HBM: padded volumes nzz x nxx x nyy (z fastest, the reference's
// index i + j nzz + k nxx nzz); V^2 dt^2 and the damper as two such volumes
// shared by all shots; per shot three rotating wavefields.  A step reads
// U_pre, U_pas, V^2 dt^2 and the damper once per cell and writes U_fut
// once: 20 bytes per cell and step for one shot; a launch that advances G
// shots reads the two model volumes once, 8 / G + 12 bytes per shot-cell.
#include <cmath>
#include <cstring>

#include "st_common.hpp"

namespace st {

namespace {

constexpr int kWz = 64;          // tile extent along z (one wave)
constexpr int kWaves = 4;        // waves per workgroup, along x

// -------------------------------------------------------------------------
// model set-up kernels

// expand (acoustic.cpp:17-79): the padded volume replicates the edge cells
// of the model outwards, i.e. reads the model at the clamped index; what is
// stored is the coefficient V^2 dt^2 of the update (acoustic.cpp:306)
__global__ void wave_expand(const float *__restrict__ m, float *__restrict__ out, int nx, int ny,
                            int nz, int nb, int nxx, int nyy, int nzz, float dt2) {
    const long long N = (long long)nxx * nyy * nzz;
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < N;
         q += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(q % nzz);
        const long long r = q / nzz;
        const int j = (int)(r % nxx), k = (int)(r / nxx);
        const int zi = min(max(i - nb, 0), nz - 1), xi = min(max(j - nb, 0), nx - 1),
                  yi = min(max(k - nb, 0), ny - 1);
        const float v = m[zi + (long long)xi * nz + (long long)yi * nx * nz];
        out[q] = v * v * dt2;
    }
}

// depth into the absorbing layer along one axis (0 at the outer face,
// nb - 1 next to the model), -1 outside the layer
__device__ __forceinline__ int layer_index(int i, int n, int nb) {
    if (i < nb) return i;
    if (i >= n - nb) return nb - (i - (n - nb)) - 1;
    return -1;
}

// damper volume: 1 in the model, the 1D profile where one axis lies in the
// layer, the 2D one where two do, the 3D one in the corners -- the case
// list of acoustic.cpp:192-302, one branch per number of layered axes
__global__ void wave_damper_volume(const float *__restrict__ d1, const float *__restrict__ d2,
                                   const float *__restrict__ d3, float *__restrict__ out, int nb,
                                   int nxx, int nyy, int nzz) {
    const long long N = (long long)nxx * nyy * nzz;
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < N;
         q += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(q % nzz);
        const long long r = q / nzz;
        const int j = (int)(r % nxx), k = (int)(r / nxx);
        const int lz = layer_index(i, nzz, nb), lx = layer_index(j, nxx, nb),
                  ly = layer_index(k, nyy, nb);
        const int n = (lz >= 0) + (lx >= 0) + (ly >= 0);
        float d = 1.0f;
        if (n == 1)
            d = d1[lz >= 0 ? lz : (lx >= 0 ? lx : ly)];
        else if (n == 2)
            d = lz < 0 ? d2[lx + ly * nb] : (lx < 0 ? d2[lz + ly * nb] : d2[lz + lx * nb]);
        else if (n == 3)
            d = d3[lz + lx * nb + (long long)ly * nb * nb];
        out[q] = d;
    }
}

// -------------------------------------------------------------------------
// the time step

struct StepArgs {
    const float *cv;       // V^2 dt^2, padded volume
    const float *damp;     // damper, padded volume
    const float *prev;     // undamped U_pre of step t-1, shot stride N
    const float *cur;      // U_pre of step t (without the wavelet)
    float *out;            // U_fut of step t
    const long long *sid;  // source cell per shot (-1: none)
    long long N;           // cells per padded volume
    int nzz, nxx, nyy;
    int ntz, ntx;          // tiles along z and x
    int ychunk;            // planes per workgroup
    float wcur, wprev;     // wavelet[t], wavelet[t-1] (0 at t = 0)
    float den_x, den_y, den_z;   // 5040 d^2
};

// One workgroup advances a 64 (z) x 4 RX (x) tile of G shots through
// `ychunk` planes of y, marching a register queue of the y-neighbours
// (four damped behind, four ahead) and staging the current plane with its
// z/x halo in LDS twice: undamped (read by the neighbours at higher index)
// and damped (read by those at lower index).
template <int RX, int G>
__global__ __launch_bounds__(256) void wave_step(StepArgs a) {
    constexpr int TX = kWaves * RX, HZ = kWz + 8, HX = TX + 8;
    __shared__ float tb[G][HX][HZ];
    __shared__ float td[G][HX][HZ];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int bz = blockIdx.x % a.ntz, bx = blockIdx.x / a.ntz;
    const int z0 = 4 + bz * kWz, x0 = 4 + bx * TX;
    const int kb = 4 + blockIdx.y * a.ychunk;
    const int ke = min(kb + a.ychunk, a.nyy - 4);
    const long long NXZ = (long long)a.nzz * a.nxx;
    const long long gofs = (long long)blockIdx.z * G * a.N;
    const float *cur = a.cur + gofs;
    const float *prev = a.prev + gofs;
    float *out = a.out + gofs;
    long long sid[G];
#pragma unroll
    for (int g = 0; g < G; g++) sid[g] = a.sid[blockIdx.z * G + g];

    const int z = z0 + lane;
    const bool zin = z < a.nzz, zc = z < a.nzz - 4;
    long long col[RX];
    bool in[RX], comp[RX];
#pragma unroll
    for (int r = 0; r < RX; r++) {
        const int x = x0 + w * RX + r;
        in[r] = zin && x < a.nxx;
        comp[r] = zc && x < a.nxx - 4;
        col[r] = z + (long long)x * a.nzz;
    }
    auto ld = [&](int g, long long q) {   // U_pre with the wavelet at the source
        const float v = cur[g * a.N + q];
        return q == sid[g] ? v + a.wcur : v;
    };

    float qd[G][RX][4], qc[G][RX], qf[G][RX][4];
#pragma unroll
    for (int r = 0; r < RX; r++) {
#pragma unroll
        for (int m = 0; m < 4; m++) {
            const long long qb = (long long)(kb - 4 + m) * NXZ + col[r];
            const long long qa = (long long)(kb + 1 + m) * NXZ + col[r];
            const float d = in[r] ? a.damp[qb] : 0.0f;
#pragma unroll
            for (int g = 0; g < G; g++) {
                qd[g][r][m] = in[r] ? ld(g, qb) * d : 0.0f;
                qf[g][r][m] = in[r] ? ld(g, qa) : 0.0f;
            }
        }
#pragma unroll
        for (int g = 0; g < G; g++) qc[g][r] = in[r] ? ld(g, (long long)kb * NXZ + col[r]) : 0.0f;
    }

    for (int k = kb; k < ke; k++) {
        const long long pk = (long long)k * NXZ;
        float dk[RX];
#pragma unroll
        for (int r = 0; r < RX; r++) dk[r] = in[r] ? a.damp[pk + col[r]] : 0.0f;
        __syncthreads();   // the previous plane's stencil reads are done
#pragma unroll
        for (int g = 0; g < G; g++)
#pragma unroll
            for (int r = 0; r < RX; r++) {
                tb[g][w * RX + r + 4][lane + 4] = qc[g][r];
                td[g][w * RX + r + 4][lane + 4] = qc[g][r] * dk[r];
            }
        // x halo: the 4 rows before the tile (lower index: damped) and the 4
        // after it (undamped), 64 lanes each
        {
            const int xl = x0 - 4 + w, xr = x0 + TX + w;
            if (zin) {
                const long long ql = pk + z + (long long)xl * a.nzz;
                const float d = a.damp[ql];
#pragma unroll
                for (int g = 0; g < G; g++) td[g][w][lane + 4] = ld(g, ql) * d;
                if (xr < a.nxx) {
                    const long long qr = pk + z + (long long)xr * a.nzz;
#pragma unroll
                    for (int g = 0; g < G; g++) tb[g][TX + 4 + w][lane + 4] = ld(g, qr);
                }
            }
        }
        // z halo: 4 cells below (damped) and 4 above (undamped) per tile row
        if (threadIdx.x < 8 * TX) {
            const int row = threadIdx.x >> 3, m = threadIdx.x & 7, x = x0 + row;
            if (x < a.nxx) {
                if (m < 4) {
                    const long long q = pk + (z0 - 4 + m) + (long long)x * a.nzz;
                    const float d = a.damp[q];
#pragma unroll
                    for (int g = 0; g < G; g++) td[g][row + 4][m] = ld(g, q) * d;
                } else if (z0 + kWz + m - 4 < a.nzz) {
                    const long long q = pk + (z0 + kWz + m - 4) + (long long)x * a.nzz;
#pragma unroll
                    for (int g = 0; g < G; g++) tb[g][row + 4][kWz + m] = ld(g, q);
                }
            }
        }
        __syncthreads();

#pragma unroll
        for (int r = 0; r < RX; r++) {
            if (!comp[r]) continue;
            const long long c = pk + col[r];
            const float cv = a.cv[c];
            const int ix = w * RX + r + 4, iz = lane + 4;
#pragma unroll
            for (int g = 0; g < G; g++) {
                const float u = qc[g][r];
                float sx = -9.0f * (td[g][ix - 4][iz] + tb[g][ix + 4][iz]);
                sx = sx + 128.0f * (td[g][ix - 3][iz] + tb[g][ix + 3][iz]);
                sx = sx - 1008.0f * (td[g][ix - 2][iz] + tb[g][ix + 2][iz]);
                sx = sx + 8064.0f * (td[g][ix - 1][iz] + tb[g][ix + 1][iz]);
                sx = sx - 14350.0f * u;
                float sy = -9.0f * (qd[g][r][0] + qf[g][r][3]);
                sy = sy + 128.0f * (qd[g][r][1] + qf[g][r][2]);
                sy = sy - 1008.0f * (qd[g][r][2] + qf[g][r][1]);
                sy = sy + 8064.0f * (qd[g][r][3] + qf[g][r][0]);
                sy = sy - 14350.0f * u;
                float sz = -9.0f * (td[g][ix][iz - 4] + tb[g][ix][iz + 4]);
                sz = sz + 128.0f * (td[g][ix][iz - 3] + tb[g][ix][iz + 3]);
                sz = sz - 1008.0f * (td[g][ix][iz - 2] + tb[g][ix][iz + 2]);
                sz = sz + 8064.0f * (td[g][ix][iz - 1] + tb[g][ix][iz + 1]);
                sz = sz - 14350.0f * u;
                const float lap = (sx / a.den_x + sy / a.den_y) + sz / a.den_z;
                float p = prev[g * a.N + c];
                if (c == sid[g]) p = p + a.wprev;
                const float pas = p * dk[r];
                float f = cv * lap + 2.0f * u;
                f = f - pas;
                out[g * a.N + c] = f * dk[r];
            }
        }
        // advance the y queue: the centre (damped) moves behind
        const bool more = k + 1 < ke;
#pragma unroll
        for (int r = 0; r < RX; r++) {
            const long long qn = pk + 5 * NXZ + col[r];
#pragma unroll
            for (int g = 0; g < G; g++) {
                qd[g][r][0] = qd[g][r][1];
                qd[g][r][1] = qd[g][r][2];
                qd[g][r][2] = qd[g][r][3];
                qd[g][r][3] = qc[g][r] * dk[r];
                qc[g][r] = qf[g][r][0];
                qf[g][r][0] = qf[g][r][1];
                qf[g][r][1] = qf[g][r][2];
                qf[g][r][2] = qf[g][r][3];
                qf[g][r][3] = (more && in[r]) ? ld(g, qn) : 0.0f;
            }
        }
    }
}

// buildSeismogram (acoustic.cpp:325-334): sample t of every receiver of G
// shots, stored time-major rec[t][g][r]
__global__ void wave_record(const float *__restrict__ out, long long N,
                            const long long *__restrict__ rcell, int nrec, int G, float *rec,
                            int t) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x, g = blockIdx.y;
    if (r >= nrec) return;
    rec[((long long)t * G + g) * nrec + r] = out[g * N + rcell[r]];
}

// rec[t][g][r] -> seis[g][r][t] (the reference's seismogram[t + r nt])
__global__ void wave_transpose(const float *__restrict__ rec, float *__restrict__ seis, int nt,
                               int G, int nrec) {
    __shared__ float tile[32][33];
    const int g = blockIdx.z;
    const int t0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 256 threads: 32 x 8
    for (int j = ty; j < 32; j += 8) {
        const int t = t0 + j, r = r0 + tx;
        if (t < nt && r < nrec) tile[j][tx] = rec[((long long)t * G + g) * nrec + r];
    }
    __syncthreads();
    for (int j = ty; j < 32; j += 8) {
        const int r = r0 + j, t = t0 + tx;
        if (t < nt && r < nrec) seis[((long long)g * nrec + r) * nt + t] = tile[tx][j];
    }
}

// -------------------------------------------------------------------------
// auto-picker

// the sample std::max_element returns: the largest under operator<, the
// first of equals; NaN never replaces a number (nor, on element 0, is it
// replaced -- handled by the final step)
struct Arg {
    float v;
    long long i;
};
__device__ __forceinline__ Arg better(Arg a, Arg b) {
    if (isnan(b.v)) return a;
    if (isnan(a.v)) return b;
    if (a.v > b.v || (a.v == b.v && a.i < b.i)) return a;
    return b;
}
__device__ Arg block_arg_max(Arg a) {
    __shared__ float sv[8];
    __shared__ long long si[8];
    for (int o = 32; o > 0; o >>= 1) {
        Arg b{__shfl_down(a.v, o), __shfl_down(a.i, o)};
        a = better(a, b);
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) { sv[wv] = a.v; si[wv] = a.i; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < (int)(blockDim.x >> 6); q++) a = better(a, Arg{sv[q], si[q]});
        sv[0] = a.v;
        si[0] = a.i;
    }
    __syncthreads();
    return Arg{sv[0], si[0]};
}

// max over a file of n samples per file, first pass: one (value, index)
// per workgroup
__global__ __launch_bounds__(256) void pick_max_partial(const float *__restrict__ x, long long n,
                                                        float *pv, long long *pi) {
    const long long f0 = (long long)blockIdx.y * n;
    Arg a{NAN, 0x7fffffffffffffffLL};
    for (long long q = blockIdx.x * 256LL + threadIdx.x; q < n; q += (long long)gridDim.x * 256)
        a = better(a, Arg{x[f0 + q], q});
    a = block_arg_max(a);
    if (threadIdx.x == 0) {
        pv[blockIdx.y * gridDim.x + blockIdx.x] = a.v;
        pi[blockIdx.y * gridDim.x + blockIdx.x] = a.i;
    }
}

// second pass: scale = 1 / max per file (auto_picking.cpp:35-38)
__global__ __launch_bounds__(256) void pick_max_final(const float *__restrict__ x, long long n,
                                                      const float *pv, const long long *pi,
                                                      int parts, float *scale) {
    Arg a{NAN, 0x7fffffffffffffffLL};
    for (int q = threadIdx.x; q < parts; q += 256)
        a = better(a, Arg{pv[blockIdx.x * parts + q], pi[blockIdx.x * parts + q]});
    a = block_arg_max(a);
    if (threadIdx.x == 0) {
        const float x0 = x[(long long)blockIdx.x * n];
        const float mx = isnan(x0) ? x0 : a.v;
        scale[blockIdx.x] = 1.0f / mx;
    }
}

// one workgroup per trace (auto_picking.cpp:42-76): the energy ratio of
// the iw samples after and before each time, its maximum, and the first
// time its normalised value crosses amp_cut
__global__ __launch_bounds__(256) void pick_trace(const float *__restrict__ seis, int ntr_file,
                                                  int nt, const float *__restrict__ scale, int iw,
                                                  float dt, float amp_cut, float shift,
                                                  float *__restrict__ picks) {
    extern __shared__ float sm[];
    float *x = sm, *S = sm + nt;
    const long long tr = blockIdx.x;
    const float inv = scale[tr / ntr_file];
    const float *s = seis + tr * nt;
    for (int q = threadIdx.x; q < nt; q += 256) {
        x[q] = s[q] * inv;
        S[q] = 0.0f;
    }
    __syncthreads();
    for (int t = iw + threadIdx.x; t < nt - iw; t += 256) {
        float A = 0.0f, B = 0.0f;
        for (int k = 0; k < iw; k++) {
            A += x[t - k] + 1e-15f;
            B += x[t + k] + 1e-15f;
        }
        float v = fabsf(B / A * (B - A));
        const float wt = (float)(iw + t);
        v *= (B + A) * (B - A) / (wt * wt);
        S[t] = v;
    }
    __syncthreads();
    Arg a{NAN, 0x7fffffffffffffffLL};
    for (int q = threadIdx.x; q < nt; q += 256) a = better(a, Arg{S[q], q});
    a = block_arg_max(a);
    const float ia = 1.0f / a.v;
    int first = 0x7fffffff;
    for (int q = threadIdx.x; q < nt; q += 256)
        if (S[q] * ia > amp_cut) { first = q; break; }
    for (int o = 32; o > 0; o >>= 1) first = min(first, __shfl_down(first, o));
    __shared__ int fw[4];
    if ((threadIdx.x & 63) == 0) fw[threadIdx.x >> 6] = first;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < 4; q++) first = min(first, fw[q]);
        picks[tr] = first == 0x7fffffff ? 0.0f : (float)first * dt + shift;
    }
}

}  // namespace

// -------------------------------------------------------------------------
// host side

// dampers (acoustic.cpp:100-137), on the host libm as the reference
void acoustic_dampers(float *d1, float *d2, float *d3, float factor, int nb) {
    std::memset(d1, 0, sizeof(float) * nb);
    std::memset(d2, 0, sizeof(float) * nb * nb);
    std::memset(d3, 0, sizeof(float) * (size_t)nb * nb * nb);
    for (int i = 0; i < nb; i++) {
        const float a = factor * (float)(nb - i);
        d1[i] = expf(-(a * a));
    }
    for (int i = 0; i < nb; i++)
        for (int j = 0; j < nb; j++) {
            d2[j + i * nb] += d1[i];
            d2[i + j * nb] += d1[i];
        }
    for (int i = 0; i < nb; i++)
        for (int j = 0; j < nb; j++)
            for (int k = 0; k < nb; k++) {
                float &c = d3[i + j * nb + (size_t)k * nb * nb];
                c += d2[i + j * nb];
                c += d2[j + k * nb];
                c += d2[i + k * nb];
            }
    for (int q = 0; q < nb * nb; q++) d2[q] -= 1.0f;
    for (size_t q = 0; q < (size_t)nb * nb * nb; q++) d3[q] -= 5.0f;
}

namespace {

struct WaveState {
    bool has_model = false;
    int nx = 0, ny = 0, nz = 0, nb = 0, nxx = 0, nyy = 0, nzz = 0;
    float dx = 0, dy = 0, dz = 0, dt = 0, den[3] = {0, 0, 0};
    long long N = 0;
    DevBuf cv, damp, fields, sid, rcell, rec, seis, pv, pi, scale, picks;
    hipEvent_t ev[2] = {nullptr, nullptr};
};

WaveState &wave_state(st_ctx *ctx) {
    if (!ctx->wave) {
        auto *w = new WaveState();
        ctx->wave = w;
        ST_HIP(hipEventCreate(&w->ev[0]));
        ST_HIP(hipEventCreate(&w->ev[1]));
    }
    return *static_cast<WaveState *>(ctx->wave);
}

template <class T> void upload(st_ctx *ctx, DevBuf &b, const T *h, size_t n) {
    b.reserve(n * sizeof(T));
    ST_HIP(hipMemcpyAsync(b.ptr, h, n * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
}

int grid_for(long long n) { return (int)std::min<long long>((n + 255) / 256, 65536); }

// cells of points (wave_main.cpp:100-105, 124-128): (int)(p / d) + nb per
// axis; points must lie in the model
std::vector<long long> cells(const WaveState &w, const float *p, int n, const char *what) {
    std::vector<long long> c(n);
    const float xmax = (w.nx - 1) * w.dx, ymax = (w.ny - 1) * w.dy, zmax = (w.nz - 1) * w.dz;
    for (int q = 0; q < n; q++) {
        const float x = p[3 * q], y = p[3 * q + 1], z = p[3 * q + 2];
        const bool ok = std::isfinite(x) && std::isfinite(y) && std::isfinite(z) && x >= 0 &&
                        y >= 0 && z >= 0 && x <= xmax && y <= ymax && z <= zmax;
        ST_REQUIRE(ok, ST_ERR_ARG, std::string(what) + " " + std::to_string(q) +
                                       " lies outside the model");
        const int i = (int)(z / w.dz) + w.nb, j = (int)(x / w.dx) + w.nb,
                  k = (int)(y / w.dy) + w.nb;
        c[q] = i + (long long)j * w.nzz + (long long)k * w.nxx * w.nzz;
    }
    return c;
}

template <int RX, int G> void launch_step(hipStream_t st, const StepArgs &a) {
    dim3 grid(a.ntz * a.ntx, (a.nyy - 8 + a.ychunk - 1) / a.ychunk, 1);
    hipLaunchKernelGGL((wave_step<RX, G>), grid, dim3(256), 0, st, a);
}

}  // namespace

void wave_release(st_ctx *ctx) {
    if (!ctx->wave) return;
    auto *w = static_cast<WaveState *>(ctx->wave);
    for (auto *b : {&w->cv, &w->damp, &w->fields, &w->sid, &w->rcell, &w->rec, &w->seis,
                    &w->pv, &w->pi, &w->scale, &w->picks})
        b->release();
    for (auto &e : w->ev)
        if (e) (void)hipEventDestroy(e);
    delete w;
    ctx->wave = nullptr;
}

void wave_set_model(st_ctx *ctx, const float *vel, int nx, int ny, int nz, float dx, float dy,
                    float dz, int nb, float factor, float dt) {
    ST_REQUIRE(vel && nx > 0 && ny > 0 && nz > 0, ST_ERR_ARG, "bad acoustic model");
    ST_REQUIRE(dx > 0 && dy > 0 && dz > 0 && dt > 0, ST_ERR_ARG, "bad spacing or time step");
    ST_REQUIRE(nb >= 4, ST_ERR_ARG, "the absorbing layer needs nb >= 4 cells");
    WaveState &w = wave_state(ctx);
    w.has_model = false;
    w.nx = nx; w.ny = ny; w.nz = nz; w.nb = nb;
    w.nxx = nx + 2 * nb; w.nyy = ny + 2 * nb; w.nzz = nz + 2 * nb;
    w.dx = dx; w.dy = dy; w.dz = dz; w.dt = dt;
    w.N = (long long)w.nxx * w.nyy * w.nzz;
    w.den[0] = 5040.0f * (dx * dx);
    w.den[1] = 5040.0f * (dy * dy);
    w.den[2] = 5040.0f * (dz * dz);
    const long long M = (long long)nx * ny * nz;
    std::vector<float> d1(nb), d2((size_t)nb * nb), d3((size_t)nb * nb * nb);
    acoustic_dampers(d1.data(), d2.data(), d3.data(), factor, nb);
    DevBuf tmp, t1, t2, t3;
    upload(ctx, tmp, vel, M);
    upload(ctx, t1, d1.data(), d1.size());
    upload(ctx, t2, d2.data(), d2.size());
    upload(ctx, t3, d3.data(), d3.size());
    w.cv.reserve(sizeof(float) * w.N);
    w.damp.reserve(sizeof(float) * w.N);
    hipLaunchKernelGGL(wave_expand, dim3(grid_for(w.N)), dim3(256), 0, ctx->stream,
                       tmp.as<float>(), w.cv.as<float>(), nx, ny, nz, nb, w.nxx, w.nyy, w.nzz,
                       dt * dt);
    hipLaunchKernelGGL(wave_damper_volume, dim3(grid_for(w.N)), dim3(256), 0, ctx->stream,
                       t1.as<float>(), t2.as<float>(), t3.as<float>(), w.damp.as<float>(), nb,
                       w.nxx, w.nyy, w.nzz);
    ST_HIP(hipGetLastError());
    ST_HIP(hipStreamSynchronize(ctx->stream));
    tmp.release(); t1.release(); t2.release(); t3.release();
    w.has_model = true;
}

namespace {

// picker over `files` files of ntr traces x nt samples resident at seis
void pick_device(st_ctx *ctx, WaveState &w, const float *seis, int files, int ntr, int nt,
                 float dt, float window, float amp_cut, float shift, float *dpicks) {
    const int iw = (int)(window / dt) + 1;
    ST_REQUIRE(std::isfinite(window) && window >= 0.0f && iw >= 1, ST_ERR_ARG,
               "picker: bad window");
    const long long n = (long long)ntr * nt;
    const int parts = (int)std::min<long long>((n + 255) / 256, 1024);
    w.pv.reserve(sizeof(float) * parts * files);
    w.pi.reserve(sizeof(long long) * parts * files);
    w.scale.reserve(sizeof(float) * files);
    hipLaunchKernelGGL(pick_max_partial, dim3(parts, files), dim3(256), 0, ctx->stream, seis, n,
                       w.pv.as<float>(), w.pi.as<long long>());
    hipLaunchKernelGGL(pick_max_final, dim3(files), dim3(256), 0, ctx->stream, seis, n,
                       w.pv.as<float>(), w.pi.as<long long>(), parts, w.scale.as<float>());
    hipLaunchKernelGGL(pick_trace, dim3((unsigned)files * ntr), dim3(256),
                       sizeof(float) * 2 * nt, ctx->stream, seis, ntr, nt, w.scale.as<float>(), iw,
                       dt, amp_cut, shift, dpicks);
    ST_HIP(hipGetLastError());
}

constexpr int kMaxPickSamples = 8192;   // 2 nt floats of LDS per trace (64 KiB)

}  // namespace

void wave_shots(st_ctx *ctx, const float *wavelet, int nt, const float *src_xyz, int nshot,
                const float *rcv_xyz, int nrec, int group, float *seis_out, float *ufut_out,
                const float *pick_params, float *picks_out, double *device_ms) {
    WaveState &w = wave_state(ctx);
    ST_REQUIRE(w.has_model, ST_ERR_NO_MODEL, "st_acoustic_set_model first");
    ST_REQUIRE(wavelet && nt > 0 && src_xyz && nshot > 0 && rcv_xyz && nrec > 0, ST_ERR_ARG,
               "bad acoustic shot arguments");
    if (picks_out) ST_REQUIRE(nt <= kMaxPickSamples, ST_ERR_ARG, "picker: nt above 8192");
    const std::vector<long long> sc = cells(w, src_xyz, nshot, "source");
    const std::vector<long long> rc = cells(w, rcv_xyz, nrec, "receiver");
    int G = group > 0 ? group : (nshot >= 2 ? 2 : 1);
    ST_REQUIRE(G == 1 || G == 2 || G == 4, ST_ERR_ARG, "shots per launch must be 1, 2 or 4");
    G = std::min(G, nshot == 1 ? 1 : G);
    const int ngroups = (nshot + G - 1) / G;
    const long long N = w.N;
    w.fields.reserve(sizeof(float) * 3 * G * N);
    w.rcell.reserve(sizeof(long long) * nrec);
    w.sid.reserve(sizeof(long long) * G);
    w.rec.reserve(sizeof(float) * (size_t)nt * G * nrec);
    w.seis.reserve(sizeof(float) * (size_t)nt * G * nrec);
    if (picks_out) w.picks.reserve(sizeof(float) * (size_t)G * nrec);
    ST_HIP(hipMemcpyAsync(w.rcell.ptr, rc.data(), sizeof(long long) * nrec,
                          hipMemcpyHostToDevice, ctx->stream));
    float *F[3] = {w.fields.as<float>(), w.fields.as<float>() + G * N,
                   w.fields.as<float>() + 2 * G * N};
    StepArgs a{};
    a.cv = w.cv.as<float>();
    a.damp = w.damp.as<float>();
    a.sid = w.sid.as<long long>();
    a.N = N;
    a.nzz = w.nzz; a.nxx = w.nxx; a.nyy = w.nyy;
    const int RX = G == 4 ? 2 : 4;
    a.ntz = (w.nzz - 8 + kWz - 1) / kWz;
    a.ntx = (w.nxx - 8 + kWaves * RX - 1) / (kWaves * RX);
    // about 2048 workgroups (8 per CU): split y into chunks when the tiles of
    // a plane are fewer; a chunk re-reads 8 planes of the queue
    const long long tiles = (long long)a.ntz * a.ntx;
    const long long chunks = std::max<long long>(1, (2048 + tiles - 1) / tiles);
    a.ychunk = (int)std::max<long long>(16, (w.nyy - 8 + chunks - 1) / chunks);
    a.ychunk = std::min(a.ychunk, w.nyy - 8);
    a.den_x = w.den[0]; a.den_y = w.den[1]; a.den_z = w.den[2];
    ST_HIP(hipEventRecord(w.ev[0], ctx->stream));
    for (int gi = 0; gi < ngroups; gi++) {
        std::vector<long long> sid(G, -1);
        for (int g = 0; g < G && gi * G + g < nshot; g++) sid[g] = sc[gi * G + g];
        ST_HIP(hipMemcpyAsync(w.sid.ptr, sid.data(), sizeof(long long) * G,
                              hipMemcpyHostToDevice, ctx->stream));
        ST_HIP(hipMemsetAsync(w.fields.ptr, 0, sizeof(float) * 3 * G * N, ctx->stream));
        int ip = 0, ic = 1, io = 2;   // U_pas source, U_pre, U_fut
        for (int t = 0; t < nt; t++) {
            a.prev = F[ip];
            a.cur = F[ic];
            a.out = F[io];
            a.wcur = wavelet[t];
            a.wprev = t > 0 ? wavelet[t - 1] : 0.0f;
            if (G == 1) launch_step<4, 1>(ctx->stream, a);
            else if (G == 2) launch_step<4, 2>(ctx->stream, a);
            else launch_step<2, 4>(ctx->stream, a);
            hipLaunchKernelGGL(wave_record, dim3((nrec + 255) / 256, G), dim3(256), 0, ctx->stream,
                               F[io], N, w.rcell.as<long long>(), nrec, G, w.rec.as<float>(), t);
            const int tmp = ip;
            ip = ic; ic = io; io = tmp;
        }
        ST_HIP(hipGetLastError());
        hipLaunchKernelGGL(wave_transpose, dim3((nt + 31) / 32, (nrec + 31) / 32, G), dim3(256), 0,
                           ctx->stream, w.rec.as<float>(), w.seis.as<float>(), nt, G, nrec);
        const int gs = std::min(G, nshot - gi * G);
        if (picks_out)
            pick_device(ctx, w, w.seis.as<float>(), gs, nrec, nt, w.dt, pick_params[0],
                        pick_params[1], pick_params[2], w.picks.as<float>());
        if (seis_out)
            ST_HIP(hipMemcpyAsync(seis_out + (size_t)gi * G * nrec * nt, w.seis.ptr,
                                  sizeof(float) * (size_t)gs * nrec * nt, hipMemcpyDeviceToHost,
                                  ctx->stream));
        if (picks_out)
            ST_HIP(hipMemcpyAsync(picks_out + (size_t)gi * G * nrec, w.picks.ptr,
                                  sizeof(float) * (size_t)gs * nrec, hipMemcpyDeviceToHost,
                                  ctx->stream));
        if (ufut_out && gi == ngroups - 1)   // U_fut of the last shot (wave_main.cpp:160)
            ST_HIP(hipMemcpyAsync(ufut_out, F[ic] + (size_t)(gs - 1) * N, sizeof(float) * N,
                                  hipMemcpyDeviceToHost, ctx->stream));
    }
    ST_HIP(hipEventRecord(w.ev[1], ctx->stream));
    ST_HIP(hipStreamSynchronize(ctx->stream));
    if (device_ms) {
        float ms = 0;
        ST_HIP(hipEventElapsedTime(&ms, w.ev[0], w.ev[1]));
        *device_ms = ms;
    }
}

void wave_pick(st_ctx *ctx, const float *seis, int ntraces, int nt, float dt, float window,
               float amp_cut, float shift, float *picks) {
    ST_REQUIRE(seis && picks && ntraces > 0 && nt > 0 && dt > 0, ST_ERR_ARG,
               "bad picker arguments");
    ST_REQUIRE(nt <= kMaxPickSamples, ST_ERR_ARG, "picker: nt above 8192");
    WaveState &w = wave_state(ctx);
    const size_t n = (size_t)ntraces * nt;
    upload(ctx, w.seis, seis, n);
    w.picks.reserve(sizeof(float) * ntraces);
    pick_device(ctx, w, w.seis.as<float>(), 1, ntraces, nt, dt, window, amp_cut, shift,
                w.picks.as<float>());
    ST_HIP(hipMemcpyAsync(picks, w.picks.ptr, sizeof(float) * ntraces, hipMemcpyDeviceToHost,
                          ctx->stream));
    ST_HIP(hipStreamSynchronize(ctx->stream));
}

}  // namespace st
