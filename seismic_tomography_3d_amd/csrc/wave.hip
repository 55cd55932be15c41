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
// Layout in HBM: padded volumes nzz x nxx x nyy in the reference's order
// (z fastest, then x, then y; acoustic.cpp:166-168) with the z stride
// rounded up to nzp, a multiple of 32, so that rows start on 128-byte lines
// and every tile row segment is whole lines; the cells z >= nzz of a row are
// padding (zero, never read).
// V^2 dt^2 and the damper are two such volumes shared by all shots; per
// shot three rotating wavefields.  A cell's damper is formed from the 1D
// profile (nb floats) in the reference's accumulation order, not stored.  A
// step reads U_pre, U_pas and V^2 dt^2 once per cell and writes U_fut once:
// 16 bytes per cell and step.
#include <cmath>
#include <cstring>
#include <type_traits>
#include <utility>

#include "st_common.hpp"

namespace st {

namespace {

constexpr int kWaves = 4;        // waves per workgroup, along x

// -------------------------------------------------------------------------
// model set-up kernels

// expand (acoustic.cpp:17-79): the padded volume replicates the edge cells
// of the model outwards, i.e. reads the model at the clamped index; what is
// stored is the coefficient V^2 dt^2 of the update (acoustic.cpp:306)
__global__ void wave_expand(const float *__restrict__ m, float *__restrict__ out, int nx, int ny,
                            int nz, int nb, int nxx, int nyy, int nzz, int nzp, float dt2) {
    const long long N = (long long)nxx * nyy * nzp;
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < N;
         q += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(q % nzp);
        const long long r = q / nzp;
        if (i >= nzz) {
            out[q] = 0.0f;
            continue;
        }
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

// A cell's damper from the 1D profile: 1 in the model, d1[l] where one axis
// lies in the layer, (d1[a] + d1[b]) - 1 where two do, and in the corners
// ((d1z + d1x) + (d1x + d1y)) + (d1z + d1y) - 5 -- the 2D and 3D tables of
// acoustic.cpp:108-136 entry for entry (each 2D entry is one sum of two
// profile values, each 3D entry three 2D sums added in that order).  v* is
// the profile value of an axis, i* whether it lies in the layer.
__device__ __forceinline__ float damper_of(float vz, bool iz, float vx, bool ix, float vy,
                                           bool iy) {
    const float szx = vz + vx, sxy = vx + vy, szy = vz + vy;
    const float d3 = ((szx + sxy) + szy) - 5.0f;
    const float d2 = (iz ? (ix ? szx : szy) : sxy) - 1.0f;
    const float d1v = iz ? vz : (ix ? vx : vy);
    const int n = (int)iz + (int)ix + (int)iy;
    return n == 0 ? 1.0f : (n == 1 ? d1v : (n == 2 ? d2 : d3));
}

// -------------------------------------------------------------------------
// the time step

struct StepArgs {
    const float *cv;       // V^2 dt^2, padded volume
    const float *d1;       // 1D damping profile (nb)
    const float *prev;     // undamped U_pre of step t-1, shot stride N
    const float *cur;      // U_pre of step t (without the wavelet)
    float *out;            // U_fut of step t
    const long long *sid;  // source cell per shot
    long long N;           // cells per padded volume
    int nzz, nzp, nzr, nxx, nyy;   // nzr: nzz rounded up to 4 (the last quad)
    int nb;
    int ntz, ntx;          // tiles along z and x
    int nchunks, ychunk;   // chunks of y planes and their length
    float wcur, wprev;     // wavelet[t], wavelet[t-1] (0 at t = 0)
    float den_x, den_y, den_z;   // 5040 d^2
};

typedef float f4 __attribute__((ext_vector_type(4)));

// Buffer resource over the cells [first, N) of one volume (raw buffer,
// 32-bit byte offsets; a load past the end returns 0 instead of faulting)
__device__ __forceinline__ __amdgpu_buffer_rsrc_t vol_rsrc(const float *base, long long first,
                                                           long long N) {
    const long long bytes = (N - first) * 4;
    const unsigned num = bytes > 0xffffffffLL ? 0xffffffffu : (unsigned)bytes;
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(base + first), (short)0, (int)num,
                                             0x00020000);
}
__device__ __forceinline__ f4 bld4(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(f4, __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0));
}
__device__ __forceinline__ void bst1(__amdgpu_buffer_rsrc_t r, unsigned off, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)off, 0, 0);
}
__device__ __forceinline__ void bst4(__amdgpu_buffer_rsrc_t r, unsigned off, f4 v) {
    __builtin_amdgcn_raw_buffer_store_b128(
        __builtin_bit_cast(__attribute__((ext_vector_type(4))) unsigned, v), r, (int)off, 0, 0);
}

template <class F, int... J>
__device__ __forceinline__ void unrolled(F &&f, std::integer_sequence<int, J...>) {
    (f(std::integral_constant<int, J>{}), ...);
}

// One workgroup advances a tile of 4 LZ (z) x TX (x) cells of one shot
// through a chunk of y planes.  A thread holds four consecutive z cells of
// one x row (one 16-byte load per field); a wave covers 64 / LZ rows.  The
// y-neighbours live in a ring of ten register planes (four damped behind,
// the centre, four ahead and the one being loaded), indexed at compile time
// by unrolling the plane loop ten times.  The current plane is staged with
// its z/x halo in LDS twice -- undamped (read by the neighbours at higher
// index) and damped (read by those at lower index) -- double buffered, one
// barrier per plane; the loads of plane k + 1 are issued before the stencil
// of plane k.  A thread's own cells come from registers: its z-neighbours
// take two 16-byte LDS reads, its x-neighbours eight.  Workgroups are
// numbered so that each XCD (workgroup id mod 8) receives a contiguous run
// of tiles: x-neighbours, which share halo rows, run on the same L2.
template <int LZ>
__global__ __launch_bounds__(256) void wave_step(StepArgs a) {
    constexpr int RW = 64 / LZ;            // x rows per wave
    constexpr int TX = kWaves * RW;        // x rows per tile
    constexpr int TZ = 4 * LZ;             // z cells per tile
    constexpr int HX = TX + 8, HZ4 = LZ + 2;
    constexpr int NXH = 8 * LZ;            // x-halo quads (128 or 64)
    __shared__ f4 tb[2][HX][HZ4];
    __shared__ f4 td[2][HX][HZ4];

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int total = gridDim.x, pblk = blockIdx.x, full = total & ~7;
    const int logical = pblk < full ? (pblk & 7) * (full >> 3) + (pblk >> 3) : pblk;
    const int ntiles = a.ntz * a.ntx, per_shot = ntiles * a.nchunks;
    const int shot = logical / per_shot, rem = logical - shot * per_shot;
    const int chunk = rem / ntiles, tile = rem - chunk * ntiles;
    const int bz = tile % a.ntz, bx = tile / a.ntz;
    const int z0 = bz * TZ, x0 = 4 + bx * TX;
    const int kb = 4 + chunk * a.ychunk, ke = min(kb + a.ychunk, a.nyy - 4);
    const int nzp = a.nzp;
    const long long NXZ = (long long)nzp * a.nxx, N = a.N;
    const long long so = (long long)shot * N;
    const float *cur = a.cur + so, *prev = a.prev + so;
    float *out = a.out + so;
    // the source: plane ks, byte offset cs in the plane, cell sj of its quad
    const long long sq = a.sid[shot];
    const int ks = (int)(sq / NXZ);
    const unsigned sc = (unsigned)(sq - (long long)ks * NXZ);
    const unsigned cs = (sc & ~3u) * 4u;
    const int sj = (int)(sc & 3u);

    // Every load reads a 16-byte quad inside the array (coordinates
    // clamped); a value outside it lands in LDS or the ring but is never
    // read by a computed cell, whose neighbours all lie inside.
    const int zq = lane % LZ, row = w * RW + lane / LZ;
    const int z = z0 + 4 * zq, x = x0 + row;
    const unsigned col = (unsigned)(min(z, a.nzr - 4) + min(x, a.nxx - 1) * nzp) * 4u;
    // quads are aligned and the interior starts at z = 4: a quad is either
    // below it or starts inside it (and may end past it)
    bool comp[4];
#pragma unroll
    for (int j = 0; j < 4; j++) comp[j] = z >= 4 && z + j < a.nzz - 4 && x < a.nxx - 4;
    const bool any = comp[0];
    const bool all4 = comp[3];
    // halo quads: x halo (4 rows before the tile, damped; 4 after) by waves
    // 0 and 1 (LZ = 16) or wave 0 (LZ = 8); z halo (one quad below each row,
    // damped; one above) by wave 2
    const bool xh_on = tid < NXH;
    const int xhr = tid / LZ, xhz = tid % LZ;          // halo row 0..7
    const bool xleft = xhr < 4;
    const int xhx = xleft ? x0 - 4 + xhr : min(x0 + TX + xhr - 4, a.nxx - 1);
    const unsigned oxh = (unsigned)(min(z0 + 4 * xhz, a.nzr - 4) + xhx * nzp) * 4u;
    const int xlrow = xleft ? xhr : TX + xhr;
    const bool zh_on = w == 2 && lane < 2 * TX;
    const bool zbelow = lane < TX;
    const int zhrow = zbelow ? lane : lane - TX;
    const int zhz = zbelow ? max(z0 - 4, 0) : min(z0 + TZ, a.nzr - 4);
    const unsigned ozh = (unsigned)(zhz + min(x0 + min(zhrow, TX - 1), a.nxx - 1) * nzp) * 4u;
    const int zhcol = zbelow ? 0 : LZ + 1;

    // profile values of the fixed axes: own row and quad, the x-halo row
    // and quad, the z-halo quad below
    const int nb = a.nb;
    auto lidx = [&](int i, int n) { return i < nb ? i : (i >= n - nb ? n - 1 - i : -1); };
    auto prof = [&](int l) { return l >= 0 ? a.d1[l] : 0.0f; };
    const int lxo = lidx(min(x, a.nxx - 1), a.nxx);
    const float vxo = prof(lxo);
    float vzo[4], vzh[4], vzb[4];
    int lzo[4], lzh[4], lzb[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        lzo[j] = lidx(min(z, a.nzr - 4) + j, a.nzz);
        vzo[j] = prof(lzo[j]);
        lzh[j] = lidx(min(z0 + 4 * xhz, a.nzr - 4) + j, a.nzz);
        vzh[j] = prof(lzh[j]);
        lzb[j] = lidx(max(z0 - 4, 0) + j, a.nzz);
        vzb[j] = prof(lzb[j]);
    }
    const int lxh = lidx(xhx, a.nxx), lxb = lidx(min(x0 + min(zhrow, TX - 1), a.nxx - 1), a.nxx);
    const float vxh = prof(lxh), vxb = prof(lxb);
    // dampers of a quad in plane k
    auto quad_damper = [&](const float *vz, const int *lz, float vx, int lx, int k) {
        const int ly = lidx(k, a.nyy);
        const float vy = prof(ly);
        f4 d;
#pragma unroll
        for (int j = 0; j < 4; j++) d[j] = damper_of(vz[j], lz[j] >= 0, vx, lx >= 0, vy, ly >= 0);
        return d;
    };

    struct Pf {
        f4 c, p, xh, zh;
    };
    Pf pf[2];
    auto fetch = [&](Pf &f, int k) {
        const long long first = (long long)k * NXZ;
        const auto Rb = vol_rsrc(cur, first, N);
        f.c = bld4(vol_rsrc(a.cv, first, N), col);
        f.p = bld4(vol_rsrc(prev, first, N), col);
        if (xh_on) f.xh = bld4(Rb, oxh);
        if (zh_on) f.zh = bld4(Rb, ozh);
    };
    // add the wavelet to the source cell of a quad of U_pre of plane k
    auto inject = [&](f4 v, int k, unsigned off, float wv) {
        if (k == ks && off == cs) v[sj] = v[sj] + wv;
        return v;
    };

    f4 q[10];   // ring of y planes
#pragma unroll
    for (int m = 0; m < 9; m++) {
        const int k = kb - 4 + m;
        const long long first = (long long)k * NXZ;
        f4 v = inject(bld4(vol_rsrc(cur, first, N), col), k, col, a.wcur);
        q[m] = m < 4 ? v * quad_damper(vzo, lzo, vxo, lxo, k) : v;
    }
    fetch(pf[0], kb);

    for (int kbase = kb; kbase < ke; kbase += 10) {
        unrolled([&](auto Jc) {
            constexpr int J = decltype(Jc)::value;
            constexpr int b = J & 1, C = (J + 4) % 10, NEW = (J + 9) % 10;
            const int k = kbase + J;
            if (k >= ke) return;
            Pf &cu = pf[b];
            if (k == ks) {   // uniform: the source plane
                if (xh_on) cu.xh = inject(cu.xh, k, oxh, a.wcur);
                if (zh_on) cu.zh = inject(cu.zh, k, ozh, a.wcur);
                cu.p = inject(cu.p, k, col, a.wprev);
            }
            const f4 u = q[C];
            const f4 dk = quad_damper(vzo, lzo, vxo, lxo, k);
            const f4 ud = u * dk;
            tb[b][row + 4][zq + 1] = u;
            td[b][row + 4][zq + 1] = ud;
            if (xh_on) {
                if (xleft) td[b][xlrow][xhz + 1] = cu.xh * quad_damper(vzh, lzh, vxh, lxh, k);
                else tb[b][xlrow][xhz + 1] = cu.xh;
            }
            if (zh_on) {
                if (zbelow) td[b][zhrow + 4][zhcol] = cu.zh * quad_damper(vzb, lzb, vxb, lxb, k);
                else tb[b][zhrow + 4][zhcol] = cu.zh;
            }
            __syncthreads();
            const bool more = k + 1 < ke;
            if (more) {
                fetch(pf[b ^ 1], k + 1);
                q[NEW] = bld4(vol_rsrc(cur, (long long)(k + 5) * NXZ, N), col);
            }
            if (any) {
                const f4 zlo = td[b][row + 4][zq], zhi = tb[b][row + 4][zq + 2];
                const f4 xm1 = td[b][row + 3][zq + 1], xp1 = tb[b][row + 5][zq + 1];
                const f4 xm2 = td[b][row + 2][zq + 1], xp2 = tb[b][row + 6][zq + 1];
                const f4 xm3 = td[b][row + 1][zq + 1], xp3 = tb[b][row + 7][zq + 1];
                const f4 xm4 = td[b][row][zq + 1], xp4 = tb[b][row + 8][zq + 1];
                // z-neighbours of cell j: lower index damped (own quad or the
                // quad below), higher index undamped (own quad or above)
                float zn[12], zp[12];
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    zn[e] = zlo[e];
                    zn[4 + e] = ud[e];
                    zp[4 + e] = u[e];
                    zp[8 + e] = zhi[e];
                }
                f4 res;
#pragma unroll
                for (int jj = 0; jj < 4; jj++) {
                    const float uc = u[jj];
                    float sx = -9.0f * (xm4[jj] + xp4[jj]);
                    sx = sx + 128.0f * (xm3[jj] + xp3[jj]);
                    sx = sx - 1008.0f * (xm2[jj] + xp2[jj]);
                    sx = sx + 8064.0f * (xm1[jj] + xp1[jj]);
                    sx = sx - 14350.0f * uc;
                    float sy = -9.0f * (q[J % 10][jj] + q[(J + 8) % 10][jj]);
                    sy = sy + 128.0f * (q[(J + 1) % 10][jj] + q[(J + 7) % 10][jj]);
                    sy = sy - 1008.0f * (q[(J + 2) % 10][jj] + q[(J + 6) % 10][jj]);
                    sy = sy + 8064.0f * (q[(J + 3) % 10][jj] + q[(J + 5) % 10][jj]);
                    sy = sy - 14350.0f * uc;
                    float sz = -9.0f * (zn[4 + jj - 4] + zp[4 + jj + 4]);
                    sz = sz + 128.0f * (zn[4 + jj - 3] + zp[4 + jj + 3]);
                    sz = sz - 1008.0f * (zn[4 + jj - 2] + zp[4 + jj + 2]);
                    sz = sz + 8064.0f * (zn[4 + jj - 1] + zp[4 + jj + 1]);
                    sz = sz - 14350.0f * uc;
                    const float lap = (sx / a.den_x + sy / a.den_y) + sz / a.den_z;
                    const float pas = cu.p[jj] * dk[jj];
                    float f = cu.c[jj] * lap + 2.0f * uc;
                    f = f - pas;
                    res[jj] = f * dk[jj];
                }
                const auto Ro = vol_rsrc(out, (long long)k * NXZ, N);
                if (all4) {
                    bst4(Ro, col, res);
                } else {
#pragma unroll
                    for (int jj = 0; jj < 4; jj++)
                        if (comp[jj]) bst1(Ro, col + 4u * jj, res[jj]);
                }
            }
            // the centre moves behind (damped); the loaded plane gets its wavelet
            q[C] = ud;
            if (more && k + 5 == ks) q[NEW] = inject(q[NEW], k + 5, col, a.wcur);
        }, std::make_integer_sequence<int, 10>{});
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

// rickerGeneration (acoustic.cpp:81-98), on the host, in the same mixed
// precision: the last factor of the exponent is a double (pow of pi)
void acoustic_ricker(float delay, float fmax, float dt, int nt, float *out) {
    const float pi = 4.0f * atanf(1.0f);
    const float fc = fmax / (3.0f * sqrtf(pi));
    for (int t = 0; t < nt; t++) {
        const float u = (float)t * dt - delay;
        const float aux1 = 1.0f - 2.0f * pi * (u * u) * (fc * fc) * (pi * pi);
        const double e = (double)(-pi * (u * u) * (fc * fc)) * ((double)pi * (double)pi);
        const float aux2 = expf((float)e);
        out[t] = aux1 * aux2;
    }
}

namespace {

struct WaveState {
    bool has_model = false;
    int nx = 0, ny = 0, nz = 0, nb = 0, nxx = 0, nyy = 0, nzz = 0, nzp = 0;
    float dx = 0, dy = 0, dz = 0, dt = 0, den[3] = {0, 0, 0};
    long long N = 0;
    DevBuf cv, d1, fields, sid, rcell, rec, seis, pv, pi, scale, picks;
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
        c[q] = i + (long long)j * w.nzp + (long long)k * w.nxx * w.nzp;
    }
    return c;
}

template <int LZ> void launch_step(hipStream_t st, const StepArgs &a, int shots) {
    const long long n = (long long)shots * a.ntz * a.ntx * a.nchunks;
    ST_REQUIRE(n < (1LL << 31), ST_ERR_ARG, "acoustic model too large for one launch");
    hipLaunchKernelGGL((wave_step<LZ>), dim3((unsigned)n), dim3(256), 0, st, a);
}

}  // namespace

void wave_release(st_ctx *ctx) {
    if (!ctx->wave) return;
    auto *w = static_cast<WaveState *>(ctx->wave);
    for (auto *b : {&w->cv, &w->d1, &w->fields, &w->sid, &w->rcell, &w->rec, &w->seis,
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
    w.nzp = (w.nzz + 31) & ~31;
    w.N = (long long)w.nxx * w.nyy * w.nzp;
    w.den[0] = 5040.0f * (dx * dx);
    w.den[1] = 5040.0f * (dy * dy);
    w.den[2] = 5040.0f * (dz * dz);
    const long long M = (long long)nx * ny * nz;
    // the 1D profile of dampers (acoustic.cpp:102-106), on the host libm;
    // the kernels form the 2D and 3D entries from it
    std::vector<float> d1(nb), d2((size_t)nb * nb), d3((size_t)nb * nb * nb);
    acoustic_dampers(d1.data(), d2.data(), d3.data(), factor, nb);
    DevBuf tmp;
    upload(ctx, tmp, vel, M);
    upload(ctx, w.d1, d1.data(), d1.size());
    w.cv.reserve(sizeof(float) * w.N);
    hipLaunchKernelGGL(wave_expand, dim3(grid_for(w.N)), dim3(256), 0, ctx->stream,
                       tmp.as<float>(), w.cv.as<float>(), nx, ny, nz, nb, w.nxx, w.nyy, w.nzz,
                       w.nzp, dt * dt);
    ST_HIP(hipGetLastError());
    ST_HIP(hipStreamSynchronize(ctx->stream));
    tmp.release();
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
    ST_REQUIRE(group >= 0, ST_ERR_ARG, "shots per launch must be >= 0");
    const long long N = w.N;
    StepArgs a{};
    a.cv = w.cv.as<float>();
    a.d1 = w.d1.as<float>();
    a.nb = w.nb;
    a.N = N;
    a.nzz = w.nzz; a.nzp = w.nzp; a.nzr = (w.nzz + 3) & ~3; a.nxx = w.nxx; a.nyy = w.nyy;
    // z tiles of 64 cells (16 lanes of four cells per x row), or of 32 when
    // that leaves markedly fewer idle lanes; tiles start at z = 0 (line
    // aligned) and cover the interior 4 <= z < nzz - 4
    const int nzi = w.nzz - 8, nxi = w.nxx - 8, nyi = w.nyy - 8;
    auto eff = [&](int tz) { return (double)nzi / (double)(((w.nzz - 4 + tz - 1) / tz) * tz); };
    const int LZ = eff(32) > eff(64) + 0.05 ? 8 : 16;
    const int TZ = 4 * LZ, TXT = kWaves * (64 / LZ);
    a.ntz = (w.nzz - 4 + TZ - 1) / TZ;
    a.ntx = (nxi + TXT - 1) / TXT;
    const long long tiles = (long long)a.ntz * a.ntx;
    // shots per launch: enough workgroups (about 2048, 8 per CU) with chunks
    // of at least 32 planes; three wavefields per shot, at most 24 GiB
    int G = group;
    if (G == 0) {
        const long long wg1 = tiles * std::max(1, nyi / 32);
        G = (int)std::max<long long>(1, (2048 + wg1 - 1) / wg1);
    }
    G = std::min(G, nshot);
    G = (int)std::max<long long>(1, std::min<long long>(G, (24LL << 30) / (12 * N)));
    const long long chunks = std::max<long long>(1, (2048 + G * tiles - 1) / (G * tiles));
    a.ychunk = (int)std::min<long long>(nyi, std::max<long long>(32, (nyi + chunks - 1) / chunks));
    a.nchunks = (nyi + a.ychunk - 1) / a.ychunk;
    a.den_x = w.den[0]; a.den_y = w.den[1]; a.den_z = w.den[2];
    const int ngroups = (nshot + G - 1) / G;
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
    a.sid = w.sid.as<long long>();
    ST_HIP(hipEventRecord(w.ev[0], ctx->stream));
    for (int gi = 0; gi < ngroups; gi++) {
        const int gs = std::min(G, nshot - gi * G);
        ST_HIP(hipMemcpyAsync(w.sid.ptr, sc.data() + (size_t)gi * G, sizeof(long long) * gs,
                              hipMemcpyHostToDevice, ctx->stream));
        ST_HIP(hipMemsetAsync(w.fields.ptr, 0, sizeof(float) * 3 * G * N, ctx->stream));
        int ip = 0, ic = 1, io = 2;   // U_pas source, U_pre, U_fut
        for (int t = 0; t < nt; t++) {
            a.prev = F[ip];
            a.cur = F[ic];
            a.out = F[io];
            a.wcur = wavelet[t];
            a.wprev = t > 0 ? wavelet[t - 1] : 0.0f;
            if (LZ == 16) launch_step<16>(ctx->stream, a, gs);
            else launch_step<8>(ctx->stream, a, gs);
            hipLaunchKernelGGL(wave_record, dim3((nrec + 255) / 256, gs), dim3(256), 0, ctx->stream,
                               F[io], N, w.rcell.as<long long>(), nrec, G, w.rec.as<float>(), t);
            const int tmp = ip;
            ip = ic; ic = io; io = tmp;
        }
        ST_HIP(hipGetLastError());
        hipLaunchKernelGGL(wave_transpose, dim3((nt + 31) / 32, (nrec + 31) / 32, gs), dim3(256), 0,
                           ctx->stream, w.rec.as<float>(), w.seis.as<float>(), nt, G, nrec);
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
            ST_HIP(hipMemcpy2DAsync(ufut_out, sizeof(float) * w.nzz, F[ic] + (size_t)(gs - 1) * N,
                                    sizeof(float) * w.nzp, sizeof(float) * w.nzz,
                                    (size_t)w.nxx * w.nyy, hipMemcpyDeviceToHost, ctx->stream));
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
