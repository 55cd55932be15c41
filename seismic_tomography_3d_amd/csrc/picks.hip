// picks.hip -- receiver first arrivals by trilinear interpolation.
//
// write_first_arrival (src/eikonal/block_FIM/block_FIM.cpp:228-262) with
// trilinear (src/utils/interpolation/trilinear.cpp:3-20), one thread per
// (shot, receiver).  This file is compiled with -ffp-contract=off and
// correctly rounded division so a pick is bit-identical to the CPU
// formula on the same volume.  Where the reference reads past the volume
// (a receiver on the last node, weight exactly 0) the +1 corner is clamped.
#include "st_common.hpp"

namespace st {

struct ZxyLayout {  // caller volume, z fastest
    int nx, nz;
    __device__ long long operator()(int x, int y, int z) const {
        return z + (long long)x * nz + (long long)y * nx * nz;
    }
};

struct PadLayout {  // sweeping solver: one-cell padded volume, z fastest, NS shots per cell
    int nxx, nzz, NS;
    __device__ long long operator()(int x, int y, int z) const {
        return ((z + 1) + (long long)(x + 1) * nzz + (long long)(y + 1) * nxx * nzz) * NS;
    }
};

struct BlkLayout {  // solver block layout
    int B, nbx, nby;
    __device__ long long operator()(int x, int y, int z) const {
        return blk_addr(x, y, z, B, nbx, nby);
    }
};

__device__ __forceinline__ float trilinear_ref(float c000, float c001, float c100, float c101,
                                               float c010, float c011, float c110, float c111,
                                               float x0, float x1, float y0, float y1, float z0,
                                               float z1, float x, float y, float z) {
    const float xd = (x - x0) / (x1 - x0);
    const float yd = (y - y0) / (y1 - y0);
    const float zd = (z - z0) / (z1 - z0);
    const float c00 = c000 * (1 - xd) + c100 * xd;
    const float c01 = c001 * (1 - xd) + c101 * xd;
    const float c10 = c010 * (1 - xd) + c110 * xd;
    const float c11 = c011 * (1 - xd) + c111 * xd;
    const float c0 = c00 * (1 - yd) + c10 * yd;
    const float c1 = c01 * (1 - yd) + c11 * yd;
    return c0 * (1 - zd) + c1 * zd;
}

template <class L>
__global__ void picks_kernel(const float *__restrict__ T, long long vol, int nshot,
                             const float *__restrict__ rcv, int nr, float *__restrict__ out,
                             int nx, int ny, int nz, float dx, float dy, float dz, L lay) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)nshot * nr) return;
    const long long s = i / nr;
    const int r = (int)(i - s * nr);
    const float *Ts = T + s * vol;
    const float x = rcv[3 * r], y = rcv[3 * r + 1], z = rcv[3 * r + 2];
    const float x0 = floorf(x / dx) * dx;
    const float y0 = floorf(y / dy) * dy;
    const float z0 = floorf(z / dz) * dz;
    const float x1 = floorf(x / dx) * dx + dx;
    const float y1 = floorf(y / dy) * dy + dy;
    const float z1 = floorf(z / dz) * dz + dz;
    const int iz = (int)(z / dz), ix = (int)(x / dx), iy = (int)(y / dy);
    const int jz = iz + 1 < nz ? iz + 1 : iz;
    const int jx = ix + 1 < nx ? ix + 1 : ix;
    const int jy = iy + 1 < ny ? iy + 1 : iy;
    const float c000 = Ts[lay(ix, iy, iz)];
    const float c001 = Ts[lay(ix, iy, jz)];
    const float c100 = Ts[lay(jx, iy, iz)];
    const float c101 = Ts[lay(jx, iy, jz)];
    const float c010 = Ts[lay(ix, jy, iz)];
    const float c011 = Ts[lay(ix, jy, jz)];
    const float c110 = Ts[lay(jx, jy, iz)];
    const float c111 = Ts[lay(jx, jy, jz)];
    out[i] = trilinear_ref(c000, c001, c100, c101, c010, c011, c110, c111, x0, x1, y0, y1, z0, z1,
                           x, y, z);
}

void launch_picks_blk(hipStream_t st, const float *T, long long vol, int nshot, const float *rcv,
                      int nr, float *out, int nx, int ny, int nz, float dx, float dy, float dz,
                      int B, int nbx, int nby) {
    const long long n = (long long)nshot * nr;
    if (n == 0) return;
    hipLaunchKernelGGL(picks_kernel<BlkLayout>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                       T, vol, nshot, rcv, nr, out, nx, ny, nz, dx, dy, dz,
                       BlkLayout{B, nbx, nby});
    ST_HIP(hipGetLastError());
}

void launch_picks_zxy(hipStream_t st, const float *T, long long vol, int nshot, const float *rcv,
                      int nr, float *out, int nx, int ny, int nz, float dx, float dy, float dz) {
    const long long n = (long long)nshot * nr;
    if (n == 0) return;
    hipLaunchKernelGGL(picks_kernel<ZxyLayout>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                       T, vol, nshot, rcv, nr, out, nx, ny, nz, dx, dy, dz, ZxyLayout{nx, nz});
    ST_HIP(hipGetLastError());
}

void launch_picks_pad(hipStream_t st, const float *T, int NS, int nshot, const float *rcv, int nr,
                      float *out, int nx, int ny, int nz, float dx, float dy, float dz) {
    const long long n = (long long)nshot * nr;
    if (n == 0) return;
    // shot s of cell c is T[c * NS + s]: a shot's stride is 1
    hipLaunchKernelGGL(picks_kernel<PadLayout>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                       T, 1LL, nshot, rcv, nr, out, nx, ny, nz, dx, dy, dz,
                       PadLayout{nx + 2, nz + 2, NS});
    ST_HIP(hipGetLastError());
}

void launch_picks_padvol(hipStream_t st, const float *T, long long vol, int nshot,
                         const float *rcv, int nr, float *out, int nx, int ny, int nz, float dx,
                         float dy, float dz) {
    const long long n = (long long)nshot * nr;
    if (n == 0) return;
    // one padded volume per shot, shot stride vol
    hipLaunchKernelGGL(picks_kernel<PadLayout>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                       T, vol, nshot, rcv, nr, out, nx, ny, nz, dx, dy, dz,
                       PadLayout{nx + 2, nz + 2, 1});
    ST_HIP(hipGetLastError());
}

}  // namespace st
