// podvin.hip -- Podvin & Lecomte (1991) first arrivals on the device
// (src/eikonal/classic/classic.cpp, the reference's eikonal_type 0),
// bit-identical to the reference's CPU code.
//
// The reference iterates a Jacobi step nItEikonal times (classic.cpp:284-
// 1833): every active cell takes the lowest of its old time and the arrivals
// of the 1D head-wave, edge and corner diffraction, 2D and 3D Huygens
// operators computed from the old times; a cell that did not improve turns
// inactive, and the 26 neighbours of every cell that did become the next
// active set.  Each step reads only old times, so the order of the cells is
// free: here the active (shot, cell) pairs of a batch of shots live in a
// device list, one thread per pair.  A step is three device passes:
// evaluate every listed pair (new value into a side array); commit the
// changed values and stamp their interior neighbours with the next step's
// number; select, in index order, the pairs stamped for the next step (a
// rocPRIM select over the stamps), so that the lanes of a wave hold
// consecutive cells and their neighbour loads coalesce.  Nothing is read
// back to the host between steps.  The same FP32 operations as the CPU code
// (this file is built with -ffp-contract=off and IEEE division and square
// root), and the minimum of the same candidate set, give the same times bit
// for bit.
//
// Layout: the padded slowness (one cell of replication per side, the
// reference's expand_model) shared by all shots; per shot a padded time
// volume and a stamp volume; one list of shot * N + cell indices and one
// array of new values.
#include <cmath>

#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

#include "st_common.hpp"

namespace st {

void launch_picks_padvol(hipStream_t st, const float *T, long long vol, int nshot,
                         const float *rcv, int nr, float *out, int nx, int ny, int nz, float dx,
                         float dy, float dz);

namespace {

// Operator offsets (z, x, y) of classic.cpp:301-1773 in the reference's
// order; S cells are the voxels whose lower corner is the given node.
// (T offset, four S cells): 1D head waves, classic.cpp:301-341
const int kHead[6][5][3] = {
    {{0, -1, 0}, {0, -1, 0}, {-1, -1, 0}, {0, -1, -1}, {-1, -1, -1}},
    {{0, 1, 0}, {0, 0, 0}, {-1, 0, 0}, {0, 0, -1}, {-1, 0, -1}},
    {{0, 0, -1}, {0, 0, -1}, {0, -1, -1}, {-1, 0, -1}, {-1, -1, -1}},
    {{0, 0, 1}, {0, 0, 0}, {-1, 0, 0}, {0, -1, 0}, {-1, -1, 0}},
    {{-1, 0, 0}, {-1, 0, 0}, {-1, -1, 0}, {-1, 0, -1}, {-1, -1, -1}},
    {{1, 0, 0}, {0, 0, 0}, {0, -1, 0}, {0, 0, -1}, {0, -1, -1}},
};
// (T offset, two S cells): edge diffractions, classic.cpp:343-395
const int kDiff[12][3][3] = {
    {{-1, -1, 0}, {-1, -1, 0}, {-1, -1, -1}}, {{-1, 1, 0}, {-1, 0, 0}, {-1, 0, -1}},
    {{1, -1, 0}, {0, -1, 0}, {0, -1, -1}},    {{1, 1, 0}, {0, 0, 0}, {0, 0, -1}},
    {{-1, 0, -1}, {-1, 0, -1}, {-1, -1, -1}}, {{-1, 0, 1}, {-1, 0, 0}, {-1, -1, 0}},
    {{1, 0, -1}, {0, 0, -1}, {0, -1, -1}},    {{1, 0, 1}, {0, 0, 0}, {0, -1, 0}},
    {{0, -1, -1}, {0, -1, -1}, {-1, -1, -1}}, {{0, -1, 1}, {0, -1, 0}, {-1, -1, 0}},
    {{0, 1, -1}, {0, 0, -1}, {-1, 0, -1}},    {{0, 1, 1}, {0, 0, 0}, {-1, 0, 0}},
};
// (T offset, S cell): corner diffractions, classic.cpp:397-429
const int kCorner[8][2][3] = {
    {{-1, -1, -1}, {-1, -1, -1}}, {{-1, -1, 1}, {-1, -1, 0}}, {{1, -1, -1}, {0, -1, -1}},
    {{1, -1, 1}, {0, -1, 0}},     {{-1, 1, -1}, {-1, 0, -1}}, {{-1, 1, 1}, {-1, 0, 0}},
    {{1, 1, -1}, {0, 0, -1}},     {{1, 1, 1}, {0, 0, 0}},
};
// (two S cells of Sref, T1, T2, T1, T2): 2D operators, classic.cpp:431-765
const int kQuad[12][6][3] = {
    {{-1, -1, 0}, {-1, -1, -1}, {0, -1, 0}, {-1, -1, 0}, {-1, 0, 0}, {-1, -1, 0}},
    {{0, -1, 0}, {0, -1, -1}, {0, -1, 0}, {1, -1, 0}, {1, 0, 0}, {1, -1, 0}},
    {{0, 0, 0}, {0, 0, -1}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {1, 1, 0}},
    {{-1, 0, 0}, {-1, 0, -1}, {0, 1, 0}, {-1, 1, 0}, {-1, 0, 0}, {-1, 1, 0}},
    {{-1, 0, -1}, {-1, -1, -1}, {0, 0, -1}, {-1, 0, -1}, {-1, 0, 0}, {-1, 0, -1}},
    {{0, 0, -1}, {0, -1, -1}, {0, 0, -1}, {1, 0, -1}, {1, 0, 0}, {1, 0, -1}},
    {{0, 0, 0}, {0, -1, 0}, {1, 0, 0}, {1, 0, 1}, {0, 0, 1}, {1, 0, 1}},
    {{-1, 0, 0}, {-1, -1, 0}, {0, 0, 1}, {-1, 0, 1}, {-1, 0, 0}, {-1, 0, 1}},
    {{0, -1, -1}, {-1, -1, -1}, {0, -1, 0}, {0, -1, -1}, {0, 0, -1}, {0, -1, -1}},
    {{0, -1, 0}, {-1, -1, 0}, {0, -1, 0}, {0, -1, 1}, {0, 0, 1}, {0, -1, 1}},
    {{0, 0, 0}, {-1, 0, 0}, {0, 0, 1}, {0, 1, 1}, {0, 1, 0}, {0, 1, 1}},
    {{0, 0, -1}, {-1, 0, -1}, {0, 1, 0}, {0, 1, -1}, {0, 0, -1}, {0, 1, -1}},
};
// (S cell, then M, N, P, Q of three faces): 3D operators, classic.cpp:767-1773
const int kOct[8][13][3] = {
    {{-1, -1, -1}, {-1, -1, -1}, {-1, -1, 0}, {-1, 0, -1}, {-1, 0, 0}, {-1, -1, -1}, {-1, -1, 0},
     {0, -1, -1}, {0, -1, 0}, {-1, -1, -1}, {0, -1, -1}, {-1, 0, -1}, {0, 0, -1}},
    {{-1, 0, -1}, {-1, 1, -1}, {-1, 0, -1}, {-1, 1, 0}, {-1, 0, 0}, {-1, 1, -1}, {-1, 1, 0},
     {0, 1, -1}, {0, 1, 0}, {-1, 1, -1}, {-1, 0, -1}, {0, 1, -1}, {0, 0, -1}},
    {{-1, 0, 0}, {-1, 1, 1}, {-1, 1, 0}, {-1, 0, 1}, {-1, 0, 0}, {-1, 1, 1}, {0, 1, 1},
     {-1, 1, 0}, {0, 1, 0}, {-1, 1, 1}, {-1, 0, 1}, {0, 1, 1}, {0, 0, 1}},
    {{-1, -1, 0}, {-1, -1, 1}, {-1, 0, 1}, {-1, -1, 0}, {-1, 0, 0}, {-1, -1, 1}, {-1, -1, 0},
     {0, -1, 1}, {0, -1, 0}, {-1, -1, 1}, {0, -1, 1}, {-1, 0, 1}, {0, 0, 1}},
    {{0, -1, -1}, {1, -1, -1}, {1, 0, -1}, {1, -1, 0}, {1, 0, 0}, {1, -1, -1}, {1, -1, 0},
     {0, -1, -1}, {0, -1, 0}, {1, -1, -1}, {0, -1, -1}, {1, 0, -1}, {0, 0, -1}},
    {{0, 0, -1}, {1, 1, -1}, {1, 1, 0}, {1, 0, -1}, {1, 0, 0}, {1, 1, -1}, {0, 1, -1},
     {1, 1, 0}, {0, 1, 0}, {1, 1, -1}, {1, 0, -1}, {0, 1, -1}, {0, 0, -1}},
    {{0, -1, 0}, {1, -1, 1}, {1, -1, 0}, {1, 0, 1}, {1, 0, 0}, {1, -1, 1}, {0, -1, 1},
     {1, -1, 0}, {0, -1, 0}, {1, -1, 1}, {1, 0, 1}, {0, -1, 1}, {0, 0, 1}},
    {{0, 0, 0}, {1, 1, 1}, {1, 0, 1}, {1, 1, 0}, {1, 0, 0}, {1, 1, 1}, {1, 1, 0},
     {0, 1, 1}, {0, 1, 0}, {1, 1, 1}, {0, 1, 1}, {1, 0, 1}, {0, 0, 1}},
};

// the offsets as cell-index differences for one padded geometry (kernel
// argument: wave-uniform, read through the scalar unit)
struct PodTab {
    int head[6][5];
    int diff[12][3];
    int corner[8][2];
    int quad[12][6];
    int oct[8][13];
    int nbr[26];              // the 26 neighbours (activation): index difference
    int nbz[26], nbx[26], nby[26];   // and the same as (z, x, y) steps
};

PodTab make_tab(int NZ, int NXZ) {
    auto lin = [&](const int *o) { return o[0] + o[1] * NZ + o[2] * NXZ; };
    PodTab t{};
    for (int q = 0; q < 6; q++)
        for (int e = 0; e < 5; e++) t.head[q][e] = lin(kHead[q][e]);
    for (int q = 0; q < 12; q++)
        for (int e = 0; e < 3; e++) t.diff[q][e] = lin(kDiff[q][e]);
    for (int q = 0; q < 8; q++)
        for (int e = 0; e < 2; e++) t.corner[q][e] = lin(kCorner[q][e]);
    for (int q = 0; q < 12; q++)
        for (int e = 0; e < 6; e++) t.quad[q][e] = lin(kQuad[q][e]);
    for (int q = 0; q < 8; q++)
        for (int e = 0; e < 13; e++) t.oct[q][e] = lin(kOct[q][e]);
    int n = 0;
    for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++)
            for (int dz = -1; dz <= 1; dz++)
                if (dz || dx || dy) {
                    t.nbz[n] = dz;
                    t.nbx[n] = dx;
                    t.nby[n] = dy;
                    t.nbr[n++] = dz + dx * NZ + dy * NXZ;
                }
    return t;
}

__device__ __forceinline__ float smin(float a, float b) { return b < a ? b : a; }   // std::min

// the lowest of T[c] and every operator's arrival (classic.cpp:297-1773)
__device__ float pod_lowest(const float *__restrict__ T, const float *__restrict__ S, long long c,
                            const PodTab &t, float h, float sqrt2, float sqrt3) {
    float lowest = T[c], Tijk;
#pragma unroll
    for (int q = 0; q < 6; q++) {
        const float m = smin(S[c + t.head[q][1]],
                             smin(S[c + t.head[q][2]], smin(S[c + t.head[q][3]], S[c + t.head[q][4]])));
        Tijk = T[c + t.head[q][0]] + h * m;
        if (Tijk < lowest) lowest = Tijk;
    }
#pragma unroll
    for (int q = 0; q < 12; q++) {
        Tijk = T[c + t.diff[q][0]] + h * sqrt2 * smin(S[c + t.diff[q][1]], S[c + t.diff[q][2]]);
        if (Tijk < lowest) lowest = Tijk;
    }
#pragma unroll
    for (int q = 0; q < 8; q++) {
        Tijk = T[c + t.corner[q][0]] + h * sqrt3 * S[c + t.corner[q][1]];
        if (Tijk < lowest) lowest = Tijk;
    }
#pragma unroll
    for (int q = 0; q < 12; q++) {
        const float Sref = smin(S[c + t.quad[q][0]], S[c + t.quad[q][1]]);
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const float T1 = T[c + t.quad[q][2 + 2 * e]], T2 = T[c + t.quad[q][3 + 2 * e]];
            if ((T1 - T2) > 0.0f && (T1 - T2) < h * Sref / sqrt2) {
                Tijk = T1 + sqrtf(h * h * Sref * Sref - (T1 - T2) * (T1 - T2));
                if (Tijk < lowest) lowest = Tijk;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const float Sref = S[c + t.oct[q][0]];
        const float hs2 = h * h * Sref * Sref;
#pragma unroll
        for (int f = 0; f < 3; f++) {
            const float M = T[c + t.oct[q][1 + 4 * f]], N = T[c + t.oct[q][2 + 4 * f]],
                        P = T[c + t.oct[q][3 + 4 * f]], Q = T[c + t.oct[q][4 + 4 * f]];
            if ((M <= N) && (M <= P) && ((2.0f * (P - M) * (P - M) + (N - M) * (N - M)) <= hs2) &&
                ((2.0f * (N - M) * (N - M) + (P - M) * (P - M)) <= hs2) &&
                ((N - M) * (N - M) + (P - M) * (P - M) + (N - M) * (P - M) >= 0.5f * hs2)) {
                Tijk = N + P - M + sqrtf(hs2 - (N - M) * (N - M) - (P - M) * (P - M));
                if (Tijk < lowest) lowest = Tijk;
            }
            if ((N <= Q) && (P <= Q) &&
                ((Q - N) * (Q - N) + (Q - P) * (Q - P) + (Q - N) * (Q - P) <= 0.5f * hs2)) {
                Tijk = Q + sqrtf(hs2 - (Q - N) * (Q - N) - (Q - P) * (Q - P));
                if (Tijk < lowest) lowest = Tijk;
            }
            if ((N - M >= 0) && (N - M <= Q - N) && (2 * (Q - N) * (Q - N) + (N - M) * (N - M) <= hs2)) {
                Tijk = Q + sqrtf(hs2 - (Q - N) * (Q - N) - (N - M) * (N - M));
                if (Tijk < lowest) lowest = Tijk;
            }
            if ((P - M >= 0) && (P - M <= Q - P) && (2 * (Q - P) * (Q - P) + (P - M) * (P - M) <= hs2)) {
                Tijk = Q + sqrtf(hs2 - (Q - P) * (Q - P) - (P - M) * (P - M));
                if (Tijk < lowest) lowest = Tijk;
            }
        }
    }
    return lowest;
}

struct PodGeo {
    int nzz, nxx, nyy;
    long long N;   // padded cells per shot
};

// T = 1e6 everywhere, stamps = -1 (classic.cpp:198-219)
__global__ void pod_fill(float *__restrict__ T, int *__restrict__ stamp, long long n) {
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (long long)gridDim.x * blockDim.x) {
        T[q] = 1e6f;
        stamp[q] = -1;
    }
}

// the source node, and its 26 neighbours stamped for step 0
// (classic.cpp:186-246); one thread per shot
__global__ void pod_seed(float *__restrict__ T, int *__restrict__ stamp, const float *__restrict__ S,
                         const float *__restrict__ src, int nshot, float dx, float dy, float dz,
                         PodGeo g, const int *__restrict__ nit, PodTab t) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nshot) return;
    const float sx = src[3 * s], sy = src[3 * s + 1], sz = src[3 * s + 2];
    const int sidx = (int)(sx / dx) + 1, sidy = (int)(sy / dy) + 1, sidz = (int)(sz / dz) + 1;
    const long long sId = sidz + (long long)sidx * g.nzz + (long long)sidy * g.nxx * g.nzz;
    const float gx = floorf(sx / dx) * dx, gy = floorf(sy / dy) * dy, gz = floorf(sz / dz) * dz;
    const float ex = sx - gx, ey = sy - gy, ez = sz - gz;
    const float dist = sqrtf(ex * ex + ey * ey + ez * ez);
    float *Ts = T + (long long)s * g.N;
    Ts[sId] = dist * S[sId];
    if (nit[s] <= 0) return;
    for (int e = 0; e < 26; e++) {
        const int i = sidz + t.nbz[e], j = sidx + t.nbx[e], k = sidy + t.nby[e];
        if (i < 1 || i > g.nzz - 2 || j < 1 || j > g.nxx - 2 || k < 1 || k > g.nyy - 2) continue;
        stamp[(long long)s * g.N + sId + t.nbr[e]] = 0;
    }
}

// one Jacobi step, part 1: the lowest arrival of every listed (shot, cell)
__global__ __launch_bounds__(256) void pod_eval(const int *__restrict__ list,
                                                const int *__restrict__ count,
                                                const float *__restrict__ T,
                                                const float *__restrict__ S,
                                                float *__restrict__ newv, PodGeo g, PodTab t,
                                                float h, float sqrt2, float sqrt3) {
    const int n = *count;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int q = list[i];
        const int s = q / (int)g.N, c = q - s * (int)g.N;
        newv[i] = pod_lowest(T + (long long)s * g.N, S, c, t, h, sqrt2, sqrt3);
    }
}

// part 2: commit the changed cells (the others turn inactive) and stamp the
// interior neighbours of the changed ones for the next step
// (classic.cpp:1775-1832)
__global__ __launch_bounds__(256) void pod_commit(const int *__restrict__ list,
                                                  const int *__restrict__ count,
                                                  const float *__restrict__ newv,
                                                  float *__restrict__ T, int *__restrict__ stamp,
                                                  const int *__restrict__ nit, int step, PodGeo g,
                                                  PodTab t) {
    const int n = *count;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int q = list[i];
        const int s = q / (int)g.N, c = q - s * (int)g.N;
        const float v = newv[i];
        float *Ts = T + (long long)s * g.N;
        if (v == Ts[c]) continue;
        Ts[c] = v;
        if (step + 1 >= nit[s]) continue;
        const int ck = c / (g.nxx * g.nzz), r = c - ck * g.nxx * g.nzz;
        const int cj = r / g.nzz, ci = r - cj * g.nzz;
        int *st = stamp + (long long)s * g.N;
#pragma unroll
        for (int e = 0; e < 26; e++) {
            const int i2 = ci + t.nbz[e], j2 = cj + t.nbx[e], k2 = ck + t.nby[e];
            if (i2 >= 1 && i2 <= g.nzz - 2 && j2 >= 1 && j2 <= g.nxx - 2 && k2 >= 1 &&
                k2 <= g.nyy - 2)
                st[c + t.nbr[e]] = step + 1;
        }
    }
}

// flag of the select: stamped for the step
struct StampIs {
    int step;
    __host__ __device__ bool operator()(int v) const { return v == step; }
};

// padded volumes of a batch -> cropped (nz, nx, ny) volumes (reduce_model,
// classic.cpp:62-72)
__global__ void pod_crop(const float *__restrict__ T, float *__restrict__ out, int nx, int ny,
                         int nz, long long N, int nshot) {
    const long long per = (long long)nx * ny * nz, n = per * nshot;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (long long)gridDim.x * blockDim.x) {
        const long long s = q / per, r = q - s * per;
        const int z = (int)(r % nz);
        const long long r2 = r / nz;
        const int x = (int)(r2 % nx), y = (int)(r2 / nx);
        out[q] = T[s * N + (z + 1) + (long long)(x + 1) * (nz + 2) +
                   (long long)(y + 1) * (nx + 2) * (nz + 2)];
    }
}

int grid_for(long long n) { return (int)std::min<long long>((n + 255) / 256, 65536); }

}  // namespace

// nItEikonal of classic.cpp:248-273: the largest source-to-corner distance in
// padded nodes (powf(x, 2.0f) is the square the reference's optimised build
// computes)
int podvin_iterations(int sidx, int sidy, int sidz, int nxx, int nyy, int nzz) {
    const float fx = (float)sidx, fy = (float)sidy, fz = (float)sidz;
    const float gx = (float)(nxx - sidx), gy = (float)(nyy - sidy), gz = (float)(nzz - sidz);
    const float c[8][3] = {{fx, fy, fz}, {gx, fy, fz}, {fx, gy, fz}, {fx, fy, gz},
                           {fx, gy, gz}, {gx, fy, gz}, {gx, gy, fz}, {gx, gy, gz}};
    int n = 0;
    for (int q = 0; q < 8; q++) {
        const int aux = (int)sqrtf(c[q][0] * c[q][0] + c[q][1] * c[q][1] + c[q][2] * c[q][2]);
        if (aux > n) n = aux;
    }
    return n;
}

// bytes of device memory one shot of a batch needs
size_t podvin_bytes_per_shot(const st_ctx *ctx) {
    const long long N = (long long)(ctx->nx + 2) * (ctx->ny + 2) * (ctx->nz + 2);
    // times, stamps, the list and the new values (the list can hold every
    // cell of the shot)
    return (size_t)N * (4 + 4 + 4 + 4);
}

// Solves nshot shots (device sources dsrc, host copies hsrc) with the
// padded slowness S into the buffers given; returns the device time in ms.
double podvin_batch(st_ctx *ctx, const float *dsrc, const float *hsrc, int nshot, const float *S,
                    float *T, int *stamp, int *list, float *newv, int *count, int *nit_dev,
                    DevBuf &temp) {
    PodGeo g;
    g.nzz = ctx->nz + 2;
    g.nxx = ctx->nx + 2;
    g.nyy = ctx->ny + 2;
    g.N = (long long)g.nzz * g.nxx * g.nyy;
    const long long total = g.N * nshot;
    ST_REQUIRE(total < (1LL << 31), ST_ERR_ARG, "Podvin batch too large");
    const PodTab t = make_tab(g.nzz, g.nxx * g.nzz);
    std::vector<int> nit(nshot);
    int maxit = 0;
    for (int s = 0; s < nshot; s++) {
        const int sidx = (int)(hsrc[3 * s] / ctx->dx) + 1, sidy = (int)(hsrc[3 * s + 1] / ctx->dy) + 1,
                  sidz = (int)(hsrc[3 * s + 2] / ctx->dz) + 1;
        nit[s] = podvin_iterations(sidx, sidy, sidz, g.nxx, g.nyy, g.nzz);
        maxit = std::max(maxit, nit[s]);
    }
    ST_HIP(hipMemcpyAsync(nit_dev, nit.data(), sizeof(int) * nshot, hipMemcpyHostToDevice,
                          ctx->stream));
    const auto ids = rocprim::counting_iterator<int>(0);
    auto flags = [&](int step) { return rocprim::make_transform_iterator(stamp, StampIs{step}); };
    size_t tbytes = 0;
    ST_HIP(rocprim::select(nullptr, tbytes, ids, flags(0), list, count, (size_t)total, ctx->stream));
    temp.reserve(std::max<size_t>(tbytes, 256));
    ST_HIP(hipEventRecord(ctx->ev[2], ctx->stream));
    hipLaunchKernelGGL(pod_fill, dim3(grid_for(total)), dim3(256), 0, ctx->stream, T, stamp, total);
    hipLaunchKernelGGL(pod_seed, dim3((nshot + 63) / 64), dim3(64), 0, ctx->stream, T, stamp, S,
                       dsrc, nshot, ctx->dx, ctx->dy, ctx->dz, g, nit_dev, t);
    ST_HIP(hipGetLastError());
    const float h = ctx->dx, sqrt2 = sqrtf(2.0f), sqrt3 = sqrtf(3.0f);
    int dev = 0, ncu = 0;
    ST_HIP(hipGetDevice(&dev));
    ST_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int grid = std::max(1, ncu) * 16;
    for (int step = 0; step < maxit; step++) {
        size_t tb = temp.bytes;
        ST_HIP(rocprim::select(temp.ptr, tb, ids, flags(step), list, count, (size_t)total,
                               ctx->stream));
        hipLaunchKernelGGL(pod_eval, dim3(grid), dim3(256), 0, ctx->stream, list, count, T, S, newv,
                           g, t, h, sqrt2, sqrt3);
        hipLaunchKernelGGL(pod_commit, dim3(grid), dim3(256), 0, ctx->stream, list, count, newv, T,
                           stamp, nit_dev, step, g, t);
    }
    ST_HIP(hipGetLastError());
    ST_HIP(hipEventRecord(ctx->ev[3], ctx->stream));
    ST_HIP(hipEventSynchronize(ctx->ev[3]));
    float ms = 0.f;
    ST_HIP(hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]));
    return ms;
}

void podvin_picks(st_ctx *ctx, const float *T, int nshot, const float *drcv, int nrcv,
                  float *dpicks) {
    const long long N = (long long)(ctx->nx + 2) * (ctx->ny + 2) * (ctx->nz + 2);
    launch_picks_padvol(ctx->stream, T, N, nshot, drcv, nrcv, dpicks, ctx->nx, ctx->ny, ctx->nz,
                        ctx->dx, ctx->dy, ctx->dz);
}

void podvin_crop_volumes(st_ctx *ctx, const float *T, float *dout, int nshot) {
    const long long N = (long long)(ctx->nx + 2) * (ctx->ny + 2) * (ctx->nz + 2);
    const long long n = (long long)ctx->nx * ctx->ny * ctx->nz * nshot;
    hipLaunchKernelGGL(pod_crop, dim3(grid_for(n)), dim3(256), 0, ctx->stream, T, dout, ctx->nx,
                       ctx->ny, ctx->nz, N, nshot);
    ST_HIP(hipGetLastError());
}

}  // namespace st
