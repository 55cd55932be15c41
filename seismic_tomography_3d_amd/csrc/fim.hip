// fim.hip -- block Fast Iterative Method (Jeong & Whitaker 2008) for gfx950.
//
// Replaces the reference's host-driven CUDA loop (block_FIM_solver,
// src/eikonal/block_FIM/cuda_kernel.cu:17-215) and its kernels run_solver
// (:250-417), run_reduction (:419-457) and run_check_neighbor (:459-619).
//
// Design (MI355X-first, not a translation):
//  * one wave64 per block visit, the block staged with its 6 face halos in
//    LDS; 4^3 blocks put one cell on each lane, 8^3 blocks put a z column of
//    8 cells on each lane (Gauss-Seidel along z inside the lane);
//  * a visit iterates the Godunov update (get_time_eikonal, :217-248) until
//    no cell of the block changes by more than eps (wave ballot), instead of
//    100 fixed sweeps;
//  * the active list never leaves the device: a visit sets the schedule bit
//    of each neighbour whose shared face changed (and of itself if it hit
//    the iteration cap) with atomicOr, and the wave appends the newly
//    scheduled blocks to the next round's list with one atomicAdd and a
//    ballot/popcount prefix;
//  * items are (shot, block) pairs, so all shots of a batch share every
//    round launch; rounds are kernel launches separated only by stream
//    order (no host sync), the host polls the list length every few rounds.
//
// Fixed point: the update is T <- min(T, G(T)) with G monotone, so every
// order of visits converges to the same discrete solution as the reference
// (whose missing barrier after the write, :408, already mixes Jacobi and
// Gauss-Seidel).  Out-of-domain neighbours are kInf (the reference's
// "halo = own value", :294-297, gives the same result under min).

#include "st_common.hpp"

namespace st {

struct FimArgs {
    const float *S;            // slowness, block layout (shared by shots)
    float *T;                  // travel times, shot-major block layout
    unsigned *flags;           // per (shot, block): bit p = scheduled for parity p
    unsigned *list0, *list1;   // active lists by round parity
    unsigned *cnt;             // [0..2] list lengths (rotating), [3] rounds run
    unsigned long long *stats; // [0] visits, [1] inner iterations
    int nbx, nby, nbz;
    int n4x, n4y, n4z;
    long long vol, nblk;
    float h, eps;
    int max_inner;
    int round;                 // round index modulo 6
};

constexpr int kShotShift = 22;
constexpr unsigned kBlkMask = (1u << kShotShift) - 1u;

__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// First-order Godunov upwind update from the smallest neighbour of each axis
// (a, b, c), hs = h*s.  Same branch structure as get_time_eikonal
// (cuda_kernel.cu:217-248); the 3D discriminant is written as
// 3(hs)^2 - [(a-b)^2 + (b-c)^2 + (c-a)^2], which equals the reference's
// 2(a(b-a)+b(c-b)+c(a-c)) + 3(hs)^2 without its cancellation at large T.
__device__ __forceinline__ float godunov(float a, float b, float c, float hs, float hs2) {
    const float hi = fmaxf(fmaxf(a, b), c);
    const float lo = fminf(fminf(a, b), c);
    const float mid = __builtin_amdgcn_fmed3f(a, b, c);
    float ret = kInf;
    if (lo < kInf) {
        ret = lo + hs;
        if (ret > mid) {
            const float d = mid - lo;
            const float t2 = 0.5f * ((mid + lo) + fsqrt(2.0f * hs2 - d * d));
            if (t2 > mid) ret = t2;
            if (ret > hi) {
                const float e1 = hi - mid, e2 = mid - lo, e3 = hi - lo;
                const float q = 3.0f * hs2 - (e1 * e1 + e2 * e2 + e3 * e3);
                const float t3 = ((hi + mid + lo) + fsqrt(q)) * (1.0f / 3.0f);
                if (t3 > hi) ret = t3;
            }
        }
    }
    return ret;
}

// Schedules the candidates held by lanes 0..6 for the next round and
// appends the newly scheduled ones to the next list.
__device__ __forceinline__ void schedule(const FimArgs &A, unsigned shot, unsigned *fl,
                                         unsigned target, unsigned bit_nxt, unsigned *nxt,
                                         unsigned *nxt_cnt, int lane) {
    bool fresh = false;
    if (target != 0xFFFFFFFFu) {
        const unsigned old = atomicOr(&fl[target], bit_nxt);
        fresh = !(old & bit_nxt);
    }
    const unsigned long long fm = __ballot(fresh);
    if (fm) {
        unsigned base = 0;
        if (lane == 0) base = atomicAdd(nxt_cnt, (unsigned)__popcll(fm));
        base = __shfl(base, 0);
        if (fresh) {
            const unsigned rank = (unsigned)__popcll(fm & ((1ull << lane) - 1ull));
            nxt[base + rank] = (shot << kShotShift) | target;
        }
    }
}

// ---------------------------------------------------------------------------
// 4^3 blocks: lane = tx + 4 ty + 16 tz, one cell per lane, Jacobi sweeps.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void fim_round4(FimArgs A) {
    __shared__ float tile[216];  // (4+2)^3, index x + 6 y + 36 z
    const int lane = threadIdx.x;
    const int p = A.round & 1;
    const unsigned count = A.cnt[A.round % 3];
    if (blockIdx.x == 0 && lane == 0) {
        A.cnt[(A.round + 2) % 3] = 0;
        if (count) atomicAdd(&A.cnt[3], 1u);
    }
    if (count == 0) return;
    const unsigned *cur = p ? A.list1 : A.list0;
    unsigned *nxt = p ? A.list0 : A.list1;
    unsigned *nxt_cnt = &A.cnt[(A.round + 1) % 3];
    const unsigned bit_cur = 1u << p, bit_nxt = 1u << (p ^ 1);

    const int tx = lane & 3, ty = (lane >> 2) & 3, tz = lane >> 4;
    const int tc = (tx + 1) + 6 * (ty + 1) + 36 * (tz + 1);
    constexpr unsigned long long kXm = 0x1111111111111111ull, kXp = 0x8888888888888888ull;
    constexpr unsigned long long kYm = 0x000F000F000F000Full, kYp = 0xF000F000F000F000ull;
    constexpr unsigned long long kZm = 0x000000000000FFFFull, kZp = 0xFFFF000000000000ull;
    const int nbxy = A.nbx * A.nby;
    unsigned long long my_visits = 0, my_iters = 0;

    for (unsigned i = blockIdx.x; i < count; i += gridDim.x) {
        const unsigned item = cur[i];
        const unsigned shot = item >> kShotShift, b = item & kBlkMask;
        float *Tsh = A.T + (long long)shot * A.vol;
        unsigned *fl = A.flags + (long long)shot * A.nblk;
        if (lane == 0) atomicAnd(&fl[b], ~bit_cur);
        const int bx = b % A.nbx, rest = b / A.nbx, by = rest % A.nby, bz = rest / A.nby;
        const long long base = (long long)b * 64;

        float v = Tsh[base + lane];
        const float s = A.S[base + lane];
        const bool valid = (bx * 4 + tx < A.n4x) && (by * 4 + ty < A.n4y) && (bz * 4 + tz < A.n4z);
        tile[tc] = v;
        for (int hid = lane; hid < 96; hid += 64) {
            const int f = hid >> 4, k = hid & 15, u = k & 3, w = k >> 2;
            int ti;
            float hv = kInf;
            switch (f) {
            case 0: ti = 6 * (u + 1) + 36 * (w + 1);
                    if (bx > 0) hv = Tsh[base - 64 + 3 + 4 * u + 16 * w]; break;
            case 1: ti = 5 + 6 * (u + 1) + 36 * (w + 1);
                    if (bx < A.nbx - 1) hv = Tsh[base + 64 + 4 * u + 16 * w]; break;
            case 2: ti = (u + 1) + 36 * (w + 1);
                    if (by > 0) hv = Tsh[base - 64LL * A.nbx + u + 12 + 16 * w]; break;
            case 3: ti = (u + 1) + 30 + 36 * (w + 1);
                    if (by < A.nby - 1) hv = Tsh[base + 64LL * A.nbx + u + 16 * w]; break;
            case 4: ti = (u + 1) + 6 * (w + 1);
                    if (bz > 0) hv = Tsh[base - 64LL * nbxy + u + 4 * w + 48]; break;
            default: ti = (u + 1) + 6 * (w + 1) + 180;
                    if (bz < A.nbz - 1) hv = Tsh[base + 64LL * nbxy + u + 4 * w]; break;
            }
            tile[ti] = hv;
        }
        __syncthreads();

        const float hs = A.h * s, hs2 = hs * hs;
        const float v0 = v;
        int it = 0;
        bool more = true;
        while (it < A.max_inner) {
            const float a = fminf(tile[tc - 1], tile[tc + 1]);
            const float bb = fminf(tile[tc - 6], tile[tc + 6]);
            const float c = fminf(tile[tc - 36], tile[tc + 36]);
            const float t = godunov(a, bb, c, hs, hs2);
            const float nv = valid ? fminf(v, t) : v;
            const bool ch = (v - nv) > A.eps;
            __syncthreads();
            tile[tc] = nv;
            v = nv;
            ++it;
            more = __ballot(ch) != 0ull;
            __syncthreads();
            if (!more) break;
        }

        const float dl = v0 - v;
        if (__ballot(dl > 0.0f)) Tsh[base + lane] = v;
        const unsigned long long m = __ballot(dl > A.eps);
        unsigned target = 0xFFFFFFFFu;
        switch (lane) {
        case 0: if (more) target = b; break;
        case 1: if ((m & kXm) && bx > 0) target = b - 1; break;
        case 2: if ((m & kXp) && bx < A.nbx - 1) target = b + 1; break;
        case 3: if ((m & kYm) && by > 0) target = b - A.nbx; break;
        case 4: if ((m & kYp) && by < A.nby - 1) target = b + A.nbx; break;
        case 5: if ((m & kZm) && bz > 0) target = b - nbxy; break;
        case 6: if ((m & kZp) && bz < A.nbz - 1) target = b + nbxy; break;
        default: break;
        }
        schedule(A, shot, fl, target, bit_nxt, nxt, nxt_cnt, lane);
        my_visits += 1;
        my_iters += it;
        __syncthreads();
    }
    if (lane == 0 && A.stats) {
        atomicAdd(&A.stats[0], my_visits);
        atomicAdd(&A.stats[1], my_iters);
    }
}

// ---------------------------------------------------------------------------
// 8^3 blocks: lane = tx + 8 ty owns the z column tz = 0..7 in registers.
// x/y neighbours come through the LDS tile, z neighbours from registers
// (Gauss-Seidel along z, alternating direction each sweep).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void fim_round8(FimArgs A) {
    __shared__ float tile[1000];  // (8+2)^3, index x + 10 y + 100 z
    const int lane = threadIdx.x;
    const int p = A.round & 1;
    const unsigned count = A.cnt[A.round % 3];
    if (blockIdx.x == 0 && lane == 0) {
        A.cnt[(A.round + 2) % 3] = 0;
        if (count) atomicAdd(&A.cnt[3], 1u);
    }
    if (count == 0) return;
    const unsigned *cur = p ? A.list1 : A.list0;
    unsigned *nxt = p ? A.list0 : A.list1;
    unsigned *nxt_cnt = &A.cnt[(A.round + 1) % 3];
    const unsigned bit_cur = 1u << p, bit_nxt = 1u << (p ^ 1);

    const int tx = lane & 7, ty = lane >> 3;
    const int tcol = (tx + 1) + 10 * (ty + 1);  // tile index of (tx, ty, z=-1)+100
    const int nbxy = A.nbx * A.nby;
    unsigned long long my_visits = 0, my_iters = 0;

    for (unsigned i = blockIdx.x; i < count; i += gridDim.x) {
        const unsigned item = cur[i];
        const unsigned shot = item >> kShotShift, b = item & kBlkMask;
        float *Tsh = A.T + (long long)shot * A.vol;
        unsigned *fl = A.flags + (long long)shot * A.nblk;
        if (lane == 0) atomicAnd(&fl[b], ~bit_cur);
        const int bx = b % A.nbx, rest = b / A.nbx, by = rest % A.nby, bz = rest / A.nby;
        const long long base = (long long)b * 512;
        const bool vxy = (bx * 8 + tx < A.n4x) && (by * 8 + ty < A.n4y);

        // vz[0] / vz[9]: z halos below / above the column; vz[1..8]: the cells
        float vz[10], v0[8], hs[8];
        #pragma unroll
        for (int k = 0; k < 8; k++) {
            vz[k + 1] = Tsh[base + lane + 64 * k];
            hs[k] = A.h * A.S[base + lane + 64 * k];
            v0[k] = vz[k + 1];
            tile[tcol + 100 * (k + 1)] = vz[k + 1];
        }
        vz[0] = (bz > 0) ? Tsh[base - 512LL * nbxy + lane + 64 * 7] : kInf;
        vz[9] = (bz < A.nbz - 1) ? Tsh[base + 512LL * nbxy + lane] : kInf;
        // x / y face halos: 4 faces x 64 cells (u = in-face coordinate, w = z)
        for (int hid = lane; hid < 256; hid += 64) {
            const int f = hid >> 6, k = hid & 63, u = k & 7, w = k >> 3;
            int ti;
            float hv = kInf;
            switch (f) {
            case 0: ti = 10 * (u + 1) + 100 * (w + 1);
                    if (bx > 0) hv = Tsh[base - 512 + 7 + 8 * u + 64 * w]; break;
            case 1: ti = 9 + 10 * (u + 1) + 100 * (w + 1);
                    if (bx < A.nbx - 1) hv = Tsh[base + 512 + 8 * u + 64 * w]; break;
            case 2: ti = (u + 1) + 100 * (w + 1);
                    if (by > 0) hv = Tsh[base - 512LL * A.nbx + u + 56 + 64 * w]; break;
            default: ti = (u + 1) + 90 + 100 * (w + 1);
                    if (by < A.nby - 1) hv = Tsh[base + 512LL * A.nbx + u + 64 * w]; break;
            }
            tile[ti] = hv;
        }
        __syncthreads();

        int it = 0;
        bool more = true;
        while (it < A.max_inner) {
            bool ch = false;
            // one plane of the column: x/y neighbours from the tile (Jacobi
            // across lanes), z neighbours from registers (Gauss-Seidel)
            auto plane = [&](const int k) {
                const int ti = tcol + 100 * (k + 1);
                const float a = fminf(tile[ti - 1], tile[ti + 1]);
                const float bb = fminf(tile[ti - 10], tile[ti + 10]);
                const float c = fminf(vz[k], vz[k + 2]);
                const float t = godunov(a, bb, c, hs[k], hs[k] * hs[k]);
                const bool valid = vxy && (bz * 8 + k < A.n4z);
                const float nv = valid ? fminf(vz[k + 1], t) : vz[k + 1];
                ch |= (vz[k + 1] - nv) > A.eps;
                vz[k + 1] = nv;
                __syncthreads();
                tile[ti] = nv;
                __syncthreads();
            };
            if ((it & 1) == 0) {
                #pragma unroll
                for (int k = 0; k < 8; k++) plane(k);
            } else {
                #pragma unroll
                for (int k = 7; k >= 0; k--) plane(k);
            }
            ++it;
            more = __ballot(ch) != 0ull;
            if (!more) break;
        }
        float v[8];
        #pragma unroll
        for (int k = 0; k < 8; k++) v[k] = vz[k + 1];

        bool anydec = false, fz0 = false, fz7 = false, fall = false;
        #pragma unroll
        for (int k = 0; k < 8; k++) {
            const float dl = v0[k] - v[k];
            anydec |= dl > 0.0f;
            const bool c = dl > A.eps;
            fall |= c;
            if (k == 0) fz0 = c;
            if (k == 7) fz7 = c;
        }
        if (__ballot(anydec)) {
            #pragma unroll
            for (int k = 0; k < 8; k++) Tsh[base + lane + 64 * k] = v[k];
        }
        const unsigned long long m = __ballot(fall);
        const bool zm_ch = __ballot(fz0) != 0ull, zp_ch = __ballot(fz7) != 0ull;
        constexpr unsigned long long kXm = 0x0101010101010101ull, kXp = 0x8080808080808080ull;
        constexpr unsigned long long kYm = 0x00000000000000FFull, kYp = 0xFF00000000000000ull;
        unsigned target = 0xFFFFFFFFu;
        switch (lane) {
        case 0: if (more) target = b; break;
        case 1: if ((m & kXm) && bx > 0) target = b - 1; break;
        case 2: if ((m & kXp) && bx < A.nbx - 1) target = b + 1; break;
        case 3: if ((m & kYm) && by > 0) target = b - A.nbx; break;
        case 4: if ((m & kYp) && by < A.nby - 1) target = b + A.nbx; break;
        case 5: if (zm_ch && bz > 0) target = b - nbxy; break;
        case 6: if (zp_ch && bz < A.nbz - 1) target = b + nbxy; break;
        default: break;
        }
        schedule(A, shot, fl, target, bit_nxt, nxt, nxt_cnt, lane);
        my_visits += 1;
        my_iters += it;
        __syncthreads();
    }
    if (lane == 0 && A.stats) {
        atomicAdd(&A.stats[0], my_visits);
        atomicAdd(&A.stats[1], my_iters);
    }
}

// ---------------------------------------------------------------------------
// set-up kernels
// ---------------------------------------------------------------------------

// Block-layout slowness for the pad-to-B extent; cells past the model are
// the clamp-to-edge replica of expand_model (block_FIM.cpp:8-57).
__global__ void pack_model(const float *__restrict__ s, float *__restrict__ S, int nx, int ny,
                           int nz, int B, int nbx, int nby, long long total) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int B3 = B * B * B;
    const long long b = i / B3;
    const int t = (int)(i - b * B3);
    const int tx = t % B, ty = (t / B) % B, tz = t / (B * B);
    const int bx = (int)(b % nbx);
    const long long rest = b / nbx;
    const int by = (int)(rest % nby), bz = (int)(rest / nby);
    const int x = min(bx * B + tx, nx - 1), y = min(by * B + ty, ny - 1), z = min(bz * B + tz, nz - 1);
    S[i] = s[z + (long long)x * nz + (long long)y * nx * nz];
}

__global__ void fill_inf(float4 *__restrict__ T, long long n4) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n4) T[i] = make_float4(kInf, kInf, kInf, kInf);
}

// apply_source_time (block_FIM.cpp:295-369): the source's grid node gets
// S*distance, its block is the only active one.
__global__ void seed_sources(const float *__restrict__ src, int nshot, const float *__restrict__ S,
                             float *T, unsigned *flags, unsigned *list0, unsigned *cnt,
                             float dx, float dy, float dz, int B, int nbx, int nby,
                             long long vol, long long nblk) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j == 0) { cnt[0] = nshot; cnt[1] = 0; cnt[2] = 0; cnt[3] = 0; }
    if (j >= nshot) return;
    const float sx = src[3 * j], sy = src[3 * j + 1], sz = src[3 * j + 2];
    const int ix = (int)(sx / dx), iy = (int)(sy / dy), iz = (int)(sz / dz);
    const long long a = blk_addr(ix, iy, iz, B, nbx, nby);
    const float ex = (float)(ix * dx) - sx, ey = (float)(iy * dy) - sy, ez = (float)(iz * dz) - sz;
    T[(long long)j * vol + a] = S[a] * sqrtf(ex * ex + ey * ey + ez * ez);
    const unsigned b = (unsigned)(a / (B * B * B));
    flags[(long long)j * nblk + b] = 1u;
    list0[j] = ((unsigned)j << kShotShift) | b;
}

// block layout -> z-fastest model extent (extract_solution + reduce_model,
// block_FIM.cpp:371-401 and :59-69)
__global__ void extract_volume(const float *__restrict__ T, float *__restrict__ out, int nx, int ny,
                               int nz, int B, int nbx, int nby, long long vol, int nshot) {
    const long long npts = (long long)nx * ny * nz;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts * nshot) return;
    const long long sh = i / npts, c = i - sh * npts;
    const int y = (int)(c / ((long long)nx * nz));
    const long long r = c - (long long)y * nx * nz;
    const int x = (int)(r / nz), z = (int)(r - (long long)x * nz);
    out[i] = T[sh * vol + blk_addr(x, y, z, B, nbx, nby)];
}

}  // namespace st

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace st {

static int blk_index(int B) { return B == 8 ? 1 : 0; }

void ensure_block_model(st_ctx *ctx, int B) {
    const int bi = blk_index(B);
    if (ctx->blk_valid[bi]) return;
    const BlockGeo g = make_geo(ctx->nx, ctx->ny, ctx->nz, B);
    ctx->slow_blk[bi].reserve(sizeof(float) * g.vol);
    const int tpb = 256;
    hipLaunchKernelGGL(pack_model, dim3((unsigned)((g.vol + tpb - 1) / tpb)), dim3(tpb), 0,
                       ctx->stream, ctx->slow_zxy.as<float>(), ctx->slow_blk[bi].as<float>(),
                       ctx->nx, ctx->ny, ctx->nz, B, g.nbx, g.nby, g.vol);
    ST_HIP(hipGetLastError());
    ctx->blk_valid[bi] = true;
}

const float *block_model(st_ctx *ctx, int B) { return ctx->slow_blk[blk_index(B)].as<float>(); }

// Device-resident FIM over one batch of shots already uploaded to dsrc.
// Leaves the travel times in ctx->T (shot-major block layout).
void fim_batch(st_ctx *ctx, const float *dsrc, int nshot, const st_fim_opts &o,
               st_fim_stats *stats, bool warm_loaded) {
    const int B = o.block;
    const BlockGeo g = make_geo(ctx->nx, ctx->ny, ctx->nz, B);
    unsigned *cnt = ctx->counters.as<unsigned>();
    unsigned long long *dstats = reinterpret_cast<unsigned long long *>(cnt + 8);
    const int tpb = 256;

    if (!warm_loaded) {
        const long long n4 = (long long)nshot * g.vol / 4;
        hipLaunchKernelGGL(fill_inf, dim3((unsigned)((n4 + tpb - 1) / tpb)), dim3(tpb), 0,
                           ctx->stream, ctx->T.as<float4>(), n4);
    }
    ST_HIP(hipMemsetAsync(ctx->flags.ptr, 0, sizeof(unsigned) * g.nblk * nshot, ctx->stream));
    hipLaunchKernelGGL(seed_sources, dim3((nshot + 63) / 64), dim3(64), 0, ctx->stream, dsrc,
                       nshot, block_model(ctx, B), ctx->T.as<float>(), ctx->flags.as<unsigned>(),
                       ctx->list[0].as<unsigned>(), cnt, ctx->dx, ctx->dy, ctx->dz, B, g.nbx,
                       g.nby, g.vol, g.nblk);
    ST_HIP(hipGetLastError());

    FimArgs A;
    A.S = block_model(ctx, B);
    A.T = ctx->T.as<float>();
    A.flags = ctx->flags.as<unsigned>();
    A.list0 = ctx->list[0].as<unsigned>();
    A.list1 = ctx->list[1].as<unsigned>();
    A.cnt = cnt;
    A.stats = dstats;
    A.nbx = g.nbx; A.nby = g.nby; A.nbz = g.nbz;
    A.n4x = g.n4x; A.n4y = g.n4y; A.n4z = g.n4z;
    A.vol = g.vol; A.nblk = g.nblk;
    A.h = ctx->dx;
    A.eps = o.eps;
    A.max_inner = o.max_inner;

    int grid = o.grid;
    if (grid <= 0) grid = 256 * 16;
    const int K = o.rounds_per_sync > 0 ? o.rounds_per_sync : 6;
    unsigned *hcnt = static_cast<unsigned *>(ctx->host_pinned);
    const long long max_rounds = 4LL * (g.nbx + g.nby + g.nbz) * 64 + 4096;

    ST_HIP(hipEventRecord(ctx->ev[2], ctx->stream));
    long long r = 0;
    int slot = 0, pending = -1;
    for (;;) {
        for (int k = 0; k < K; k++, r++) {
            A.round = (int)(r % 6);
            if (B == 8)
                hipLaunchKernelGGL(fim_round8, dim3(grid), dim3(64), 0, ctx->stream, A);
            else
                hipLaunchKernelGGL(fim_round4, dim3(grid), dim3(64), 0, ctx->stream, A);
        }
        ST_HIP(hipGetLastError());
        ST_HIP(hipMemcpyAsync(&hcnt[slot], &cnt[r % 3], sizeof(unsigned), hipMemcpyDeviceToHost,
                              ctx->stream));
        ST_HIP(hipEventRecord(ctx->ev[slot], ctx->stream));
        if (pending >= 0) {
            ST_HIP(hipEventSynchronize(ctx->ev[pending]));
            if (hcnt[pending] == 0) break;
        }
        pending = slot;
        slot ^= 1;
        ST_REQUIRE(r < max_rounds, ST_ERR_HIP, "FIM did not terminate within the round cap");
    }
    ST_HIP(hipEventRecord(ctx->ev[3], ctx->stream));
    if (stats) {
        ST_HIP(hipMemcpyAsync(hcnt + 4, cnt + 3, sizeof(unsigned), hipMemcpyDeviceToHost, ctx->stream));
        ST_HIP(hipMemcpyAsync(hcnt + 8, dstats, 2 * sizeof(unsigned long long),
                              hipMemcpyDeviceToHost, ctx->stream));
        ST_HIP(hipStreamSynchronize(ctx->stream));
        const unsigned long long *hs = reinterpret_cast<const unsigned long long *>(hcnt + 8);
        stats->rounds += hcnt[4];
        stats->visits += (long long)hs[0];
        stats->inner_iters += (long long)hs[1];
        stats->batches += 1;
        float ms = 0.f;
        ST_HIP(hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]));
        stats->solve_ms += ms;
        ST_HIP(hipMemsetAsync(dstats, 0, 2 * sizeof(unsigned long long), ctx->stream));
    }
}

void launch_extract(st_ctx *ctx, float *dout, int nshot, int B) {
    const BlockGeo g = make_geo(ctx->nx, ctx->ny, ctx->nz, B);
    const long long n = (long long)ctx->nx * ctx->ny * ctx->nz * nshot;
    hipLaunchKernelGGL(extract_volume, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream,
                       ctx->T.as<float>(), dout, ctx->nx, ctx->ny, ctx->nz, B, g.nbx, g.nby, g.vol,
                       nshot);
    ST_HIP(hipGetLastError());
}

}  // namespace st
