// fsm.hip -- the reference's CPU eikonal, Accurate_FSM (Noble, Gesret and
// Belayouni 2014 operators; src/eikonal/accurate_FSM/accurate_FSM.cpp), on
// the GPU for a batch of shots, bit-identical to the CPU code.
//
// Why a parallel schedule can reproduce the sequential sweeps exactly: the
// update of cell (i, j, k) in a sweep with directions (sz, sx, sy) reads T
// only at offsets (-sz or 0, -sx or 0, -sy or 0) -- its own value and seven
// upwind neighbours (accurate_FSM.cpp:168-175).  Number the cells of a
// sweep's box by p = a + b + c, their distances from the box's start corner
// along z, x, y.  Every upwind neighbour has a smaller p, and the loop nest
// (k outer, j, i inner) also visits every cell after all cells of smaller p
// that it reads.  So updating the planes p = 0, 1, 2, ... one after another,
// all cells of a plane at once (they read nothing from each other), gives
// each update the same inputs, and with the same arithmetic the same
// result.  One kernel launch per plane (kernel boundaries order the planes);
// a launch covers that plane for every shot of the batch.
//
// Layout: travel times are stored shot-fastest, T[cell * NS + shot] (NS =
// the batch rounded up to 64), so a wave updates one cell for 64 shots: its
// T loads are contiguous 256-byte rows and every slowness load and index
// is wave-uniform.  That needs one cell per plane for all shots, while the
// eight source-bounded sweeps of init_sweep start at each shot's own source:
// planes are therefore numbered from the corner of the union of the shots'
// boxes, and a lane skips cells outside its shot's box.  Within one shot's
// box the absolute plane number is its own p plus a constant, so the order
// argument above holds per shot.
//
// Compiled with -ffp-contract=off and HIP's correctly rounded division and
// square root; min/max follow std::min/std::max operand order.  The seeds'
// powf(x, 2.0f) are computed as x*x, the form the reference's compiler
// emits for them (exact squares; the oracle's CPU build agrees bit for bit).
#include "st_common.hpp"

namespace st {

void launch_picks_pad(hipStream_t st, const float *T, int NS, int nshot, const float *rcv, int nr,
                      float *out, int nx, int ny, int nz, float dx, float dy, float dz);

__device__ __forceinline__ float smin(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float smax(float a, float b) { return (a < b) ? b : a; }

// Geometry and the constants eikonal_equation derives from the spacing
// (accurate_FSM.cpp:632-644), all computed on the host in FP32.
struct FsmGeo {
    int nxx, nyy, nzz;      // padded extent (model + 2)
    long long nB;           // nxx * nyy * nzz
    float dx, dy, dz;
    float dz2i, dx2i, dy2i, dz2dx2, dz2dy2, dx2dy2, dsum;
};

// One sweep: directions (1 forward, -1 backward) for z (i), x (j), y (k),
// and whether it is one of the eight source-bounded sweeps of init_sweep.
struct FsmSweep {
    int tz, tx, ty;
    int init;
};

// Start and extent of a sweep along one axis (n = padded length, sid = the
// source's padded index): init_sweep (accurate_FSM.cpp:267-388) starts the
// forward loops at max(1, sid) and the backward loops at sid + 1;
// full_sweep (:390-511) runs 1 .. n-1 forward and n-2 .. 0 backward.  For a
// source on the last node the reference's backward start sid + 1 = n - 1
// reads one cell past the array; the start is clamped to n - 2 there (the
// CPU oracle does the same).
__host__ __device__ inline void sweep_axis(int t, int init, int n, int sid, int &start, int &len) {
    if (t > 0) {
        start = init ? (sid > 1 ? sid : 1) : 1;
        len = n - start;
    } else {
        const int s = init ? (sid < n - 3 ? sid : n - 3) + 1 : n - 2;
        start = s;
        len = s + 1;
    }
}

// The update of one cell, accurate_FSM.cpp:159-265 (inner_sweep).
__device__ __forceinline__ void fsm_update(float *__restrict__ T, int NS,
                                           const float *__restrict__ S, const FsmGeo &g, int i,
                                           int j, int k, int sgntz, int sgntx, int sgnty) {
    const int sgnvz = (sgntz + 1) >> 1, sgnvx = (sgntx + 1) >> 1, sgnvy = (sgnty + 1) >> 1;
    const int nzz = g.nzz, nxx = g.nxx, nyy = g.nyy;
    const long long NZ = nzz, NXZ = (long long)nxx * nzz;
    const float dx = g.dx, dy = g.dy, dz = g.dz;
    const int i1 = i - sgnvz, j1 = j - sgnvx, k1 = k - sgnvy;
    const int jm = max(j - 1, 1), jc = min(j, nxx - 1);
    const int km = max(k - 1, 1), kc = min(k, nyy - 1);
    const int im = max(i - 1, 1), ic = min(i, nzz - 1);

    // T points at this lane's shot; cells are NS floats apart
    auto tat = [&](long long c) { return T[c * NS]; };
    const float tv = tat((i - sgntz) + j * NZ + k * NXZ);
    const float te = tat(i + (j - sgntx) * NZ + k * NXZ);
    const float tn = tat(i + j * NZ + (k - sgnty) * NXZ);
    const float tev = tat((i - sgntz) + (j - sgntx) * NZ + k * NXZ);
    const float ten = tat(i + (j - sgntx) * NZ + (k - sgnty) * NXZ);
    const float tnv = tat((i - sgntz) + j * NZ + (k - sgnty) * NXZ);
    const float tnve = tat((i - sgntz) + (j - sgntx) * NZ + (k - sgnty) * NXZ);
    const long long ijk = (i + j * NZ + k * NXZ) * NS;

    // 1D operators: upwind time plus spacing times the least of the four
    // slownesses around the edge
    const float t1D1 = tv + dz * smin(S[i1 + jm * NZ + km * NXZ],
                                 smin(S[i1 + jm * NZ + kc * NXZ],
                                 smin(S[i1 + jc * NZ + km * NXZ], S[i1 + jc * NZ + kc * NXZ])));
    const float t1D2 = te + dx * smin(S[im + j1 * NZ + km * NXZ],
                                 smin(S[ic + j1 * NZ + km * NXZ],
                                 smin(S[im + j1 * NZ + kc * NXZ], S[ic + j1 * NZ + kc * NXZ])));
    const float t1D3 = tn + dy * smin(S[im + jm * NZ + k1 * NXZ],
                                 smin(S[im + jc * NZ + k1 * NXZ],
                                 smin(S[ic + jm * NZ + k1 * NXZ], S[ic + jc * NZ + k1 * NXZ])));
    const float t1D = smin(t1D1, smin(t1D2, t1D3));

    // 2D four-point operators in the xz, yz and xy planes
    float t2D1 = 1e6f, t2D2 = 1e6f, t2D3 = 1e6f;
    float Sref = smin(S[i1 + j1 * NZ + km * NXZ], S[i1 + j1 * NZ + kc * NXZ]);
    if ((tv < te + dx * Sref) && (te < tv + dz * Sref)) {
        const float ta = tev + te - tv;
        const float tb = tev - te + tv;
        t2D1 = ((tb * g.dz2i + ta * g.dx2i) +
                sqrtf(4.0f * Sref * Sref * (g.dz2i + g.dx2i) -
                      g.dz2i * g.dx2i * (ta - tb) * (ta - tb))) / (g.dz2i + g.dx2i);
    }
    Sref = smin(S[i1 + jm * NZ + k1 * NXZ], S[i1 + jc * NZ + k1 * NXZ]);
    if ((tv < tn + dy * Sref) && (tn < tv + dz * Sref)) {
        const float ta = tv - tn + tnv;
        const float tb = tn - tv + tnv;
        t2D2 = ((ta * g.dz2i + tb * g.dy2i) +
                sqrtf(4.0f * Sref * Sref * (g.dz2i + g.dy2i) -
                      g.dz2i * g.dy2i * (ta - tb) * (ta - tb))) / (g.dz2i + g.dy2i);
    }
    Sref = smin(S[im + j1 * NZ + k1 * NXZ], S[ic + j1 * NZ + k1 * NXZ]);
    if ((te < tn + dy * Sref) && (tn < te + dx * Sref)) {
        const float ta = te - tn + ten;
        const float tb = tn - te + ten;
        t2D3 = ((ta * g.dx2i + tb * g.dy2i) +
                sqrtf(4.0f * Sref * Sref * (g.dx2i + g.dy2i) -
                      g.dx2i * g.dy2i * (ta - tb) * (ta - tb))) / (g.dx2i + g.dy2i);
    }
    const float t2D = smin(t2D1, smin(t2D2, t2D3));

    // 3D operator on the upwind octant
    float t3D = 1e6f;
    Sref = S[i1 + j1 * NZ + k1 * NXZ];
    const float ta = te - 0.5f * tn + 0.5f * ten - 0.5f * tv + 0.5f * tev - tnv + tnve;
    const float tb = tv - 0.5f * tn + 0.5f * tnv - 0.5f * te + 0.5f * tev - ten + tnve;
    const float tc = tn - 0.5f * te + 0.5f * ten - 0.5f * tv + 0.5f * tnv - tev + tnve;
    if (smin(t1D, t2D) > smax(tv, smax(te, tn))) {
        const float t2 = 9.0f * Sref * Sref * g.dsum;
        const float t3 = g.dz2dx2 * (ta - tb) * (ta - tb) + g.dz2dy2 * (tb - tc) * (tb - tc) +
                         g.dx2dy2 * (ta - tc) * (ta - tc);
        if (t2 >= t3) {
            const float t1 = tb * g.dz2i + ta * g.dx2i + tc * g.dy2i;
            t3D = (t1 + sqrtf(t2 - t3)) / g.dsum;
        }
    }
    T[ijk] = smin(T[ijk], smin(t1D, smin(t2D, t3D)));
}

// Slowness padded by one cell, edge replicated: expand_model
// (accurate_FSM.cpp:8-60) copies the boundary plane outward along z, then
// x, then y, which is the model value at the clamped index.
__global__ void fsm_pad_model(const float *__restrict__ s, float *__restrict__ S, int nx, int ny,
                              int nz) {
    const int nzz = nz + 2, nxx = nx + 2, nyy = ny + 2;
    const long long n = (long long)nzz * nxx * nyy;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (long long)gridDim.x * blockDim.x) {
        const int z = (int)(q % nzz);
        const long long r = q / nzz;
        const int x = (int)(r % nxx), y = (int)(r / nxx);
        const int zs = min(max(z - 1, 0), nz - 1), xs = min(max(x - 1, 0), nx - 1),
                  ys = min(max(y - 1, 0), ny - 1);
        S[q] = s[zs + (long long)xs * nz + (long long)ys * nx * nz];
    }
}

// T = 1e6 everywhere (accurate_FSM.cpp:591-592), four cells per thread plus
// the n % 4 tail
__global__ void fsm_fill(float *__restrict__ T, long long n) {
    const long long n4 = n / 4;
    float4 *T4 = reinterpret_cast<float4 *>(T);
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n4;
         q += (long long)gridDim.x * blockDim.x)
        T4[q] = make_float4(1e6f, 1e6f, 1e6f, 1e6f);
    if (blockIdx.x == 0 && threadIdx.x < (unsigned)(n - 4 * n4)) T[4 * n4 + threadIdx.x] = 1e6f;
}

// The 27 analytic seeds around the source node (accurate_FSM.cpp:596-630):
// entry e writes the node at index offset (iz, ix, iy) with S[source] times
// the distance of the node at offset (dz, dx, dy).  The two xz diagonals of
// lines 608-609 pair an index offset with the mirrored node's distance; the
// table keeps that pairing so the results match.
__constant__ signed char kSeed[27][6] = {
    {0, 0, 0, 0, 0, 0},
    {1, 0, 0, 1, 0, 0},     {-1, 0, 0, -1, 0, 0},
    {0, 1, 0, 0, 1, 0},     {0, -1, 0, 0, -1, 0},
    {0, 0, 1, 0, 0, 1},     {0, 0, -1, 0, 0, -1},
    {1, 1, 0, 1, 1, 0},     {1, -1, 0, -1, 1, 0},
    {-1, 1, 0, 1, -1, 0},   {-1, -1, 0, -1, -1, 0},
    {1, 0, 1, 1, 0, 1},     {1, 0, -1, 1, 0, -1},
    {-1, 0, 1, -1, 0, 1},   {-1, 0, -1, -1, 0, -1},
    {0, 1, 1, 0, 1, 1},     {0, 1, -1, 0, 1, -1},
    {0, -1, 1, 0, -1, 1},   {0, -1, -1, 0, -1, -1},
    {1, 1, 1, 1, 1, 1},     {1, -1, 1, 1, -1, 1},
    {1, 1, -1, 1, 1, -1},   {1, -1, -1, 1, -1, -1},
    {-1, 1, 1, -1, 1, 1},   {-1, -1, 1, -1, -1, 1},
    {-1, 1, -1, -1, 1, -1}, {-1, -1, -1, -1, -1, -1},
};

// One thread per shot, the entries in table order (entries that name the
// same node are written in the reference's order: the last wins).
__global__ void fsm_seed(float *__restrict__ T, int NS, const float *__restrict__ S,
                         const float *__restrict__ src, const int *__restrict__ sid, int nshot,
                         FsmGeo g) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nshot) return;
    const int nb = 1;
    const float sx = src[3 * s], sy = src[3 * s + 1], sz = src[3 * s + 2];
    const int sidz = sid[3 * s], sidx = sid[3 * s + 1], sidy = sid[3 * s + 2];
    const long long NZ = g.nzz, NXZ = (long long)g.nxx * g.nzz;
    const long long sId = sidz + sidx * NZ + sidy * NXZ;
    float *Ts = T + s;
    const float S0 = S[sId];
    for (int e = 0; e < 27; e++) {
        const int iz = kSeed[e][0], ix = kSeed[e][1], iy = kSeed[e][2];
        const int oz = kSeed[e][3], ox = kSeed[e][4], oy = kSeed[e][5];
        const float ex = ((sidx - nb) + ox) * g.dx - sx;
        const float ey = ((sidy - nb) + oy) * g.dy - sy;
        const float ez = ((sidz - nb) + oz) * g.dz - sz;
        Ts[(sId + iz + ix * NZ + iy * NXZ) * NS] = S0 * sqrtf(ex * ex + ey * ey + ez * ez);
    }
}

// One plane p of one sweep for all shots.  The union box of the sweep
// starts at (zs, xs, ys) and runs along the sweep directions; cells of the
// plane are enumerated by (a, c) -- distances along z and y -- over the
// rectangle [a0, a0 + na) x [c0, c0 + nc), with b = p - a - c along x.
// Thread q: cell q / NS of the rectangle, shot q % NS (waves never straddle
// cells: NS and the block size are multiples of 64).
__global__ __launch_bounds__(256) void fsm_plane(float *__restrict__ T, int NS, int nshot,
                                                 const float *__restrict__ S,
                                                 const int *__restrict__ sid, FsmGeo g,
                                                 FsmSweep w, int zs, int xs, int ys, int xl, int p,
                                                 int a0, int na, int c0, int nc) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int cell = __builtin_amdgcn_readfirstlane((int)(q / NS));
    if (cell >= na * nc) return;
    const int a = a0 + cell % na, c = c0 + cell / na;
    const int b = p - a - c;
    if (b < 0 || b >= xl) return;
    const int s = (int)(q - (long long)cell * NS);
    if (s >= nshot) return;
    const int i = zs + w.tz * a, j = xs + w.tx * b, k = ys + w.ty * c;
    // this shot's own box along each axis
    int st, ln;
    sweep_axis(w.tz, w.init, g.nzz, sid[3 * s], st, ln);
    if (w.tz * (i - st) < 0 || w.tz * (i - st) >= ln) return;
    sweep_axis(w.tx, w.init, g.nxx, sid[3 * s + 1], st, ln);
    if (w.tx * (j - st) < 0 || w.tx * (j - st) >= ln) return;
    sweep_axis(w.ty, w.init, g.nyy, sid[3 * s + 2], st, ln);
    if (w.ty * (k - st) < 0 || w.ty * (k - st) >= ln) return;
    fsm_update(T + s, NS, S, g, i, j, k, w.tz, w.tx, w.ty);
}

// reduce_model (accurate_FSM.cpp:62-72): padded -> model extent, z fastest
__global__ void fsm_crop(const float *__restrict__ T, int NS, float *__restrict__ out, int nx,
                         int ny, int nz, int nshot) {
    const long long npts = (long long)nx * ny * nz;
    const long long n = npts * nshot;
    const int nzz = nz + 2, nxx = nx + 2;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (long long)gridDim.x * blockDim.x) {
        const long long s = q / npts, r = q - s * npts;
        const int z = (int)(r % nz);
        const long long r2 = r / nz;
        const int x = (int)(r2 % nx), y = (int)(r2 / nx);
        out[q] = T[((z + 1) + (long long)(x + 1) * nzz + (long long)(y + 1) * nxx * nzz) * NS + s];
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static unsigned grid_for(long long n) {
    return (unsigned)std::min<long long>(std::max<long long>((n + 255) / 256, 1), 1 << 20);
}

static FsmGeo fsm_geo(const st_ctx *ctx) {
    FsmGeo g;
    g.nxx = ctx->nx + 2; g.nyy = ctx->ny + 2; g.nzz = ctx->nz + 2;
    g.nB = (long long)g.nxx * g.nyy * g.nzz;
    g.dx = ctx->dx; g.dy = ctx->dy; g.dz = ctx->dz;
    g.dz2i = 1.0f / (g.dz * g.dz);
    g.dx2i = 1.0f / (g.dx * g.dx);
    g.dy2i = 1.0f / (g.dy * g.dy);
    g.dz2dx2 = g.dz2i * g.dx2i;
    g.dz2dy2 = g.dz2i * g.dy2i;
    g.dx2dy2 = g.dx2i * g.dy2i;
    g.dsum = g.dz2i + g.dx2i + g.dy2i;
    return g;
}

// The 16 sweeps of eikonal_equation: init_sweep then full_sweep, each in the
// reference's direction order (accurate_FSM.cpp:270-388, 393-511).
static const int kDirs[8][3] = {{1, 1, 1},  {-1, 1, 1},  {1, 1, -1},  {-1, 1, -1},
                                {1, -1, 1}, {-1, -1, 1}, {1, -1, -1}, {-1, -1, -1}};

// Shots per device pass are padded to a multiple of 64 (one wave = one cell
// for 64 shots).
int fsm_lanes(int nshot) { return (nshot + 63) / 64 * 64; }

// Solves nshot shots (device sources dsrc, host copies hsrc) into ctx->T
// (padded volume, shot fastest); returns the device time in ms.
double fsm_batch(st_ctx *ctx, const float *dsrc, const float *hsrc, int nshot, const float *S) {
    const FsmGeo g = fsm_geo(ctx);
    const int NS = fsm_lanes(nshot);
    std::vector<int> hsid(3 * nshot);
    for (int s = 0; s < nshot; s++) {
        hsid[3 * s] = (int)(hsrc[3 * s + 2] / ctx->dz) + 1;      // z
        hsid[3 * s + 1] = (int)(hsrc[3 * s] / ctx->dx) + 1;      // x
        hsid[3 * s + 2] = (int)(hsrc[3 * s + 1] / ctx->dy) + 1;  // y
    }
    ctx->fsm_sid.reserve(sizeof(int) * 3 * std::max(nshot, 1));
    int *dsid = ctx->fsm_sid.as<int>();
    ST_HIP(hipMemcpyAsync(dsid, hsid.data(), sizeof(int) * 3 * nshot, hipMemcpyHostToDevice,
                          ctx->stream));
    ST_HIP(hipEventRecord(ctx->ev[2], ctx->stream));
    const long long nT = (long long)NS * g.nB;
    float *T = ctx->T.as<float>();
    ST_REQUIRE(ctx->T.bytes >= sizeof(float) * (size_t)nT, ST_ERR_STATE, "travel-time buffer too small");
    hipLaunchKernelGGL(fsm_fill, dim3(grid_for(nT / 4)), dim3(256), 0, ctx->stream, T, nT);
    hipLaunchKernelGGL(fsm_seed, dim3((nshot + 63) / 64), dim3(64), 0, ctx->stream, T, NS, S, dsrc,
                       dsid, nshot, g);
    ST_HIP(hipGetLastError());
    for (int pass = 0; pass < 2; pass++) {
        for (int d = 0; d < 8; d++) {
            const FsmSweep w{kDirs[d][0], kDirs[d][1], kDirs[d][2], pass == 0 ? 1 : 0};
            // union of the shots' boxes: the earliest start along each direction
            int start[3], len[3];
            const int dir[3] = {w.tz, w.tx, w.ty}, n[3] = {g.nzz, g.nxx, g.nyy};
            for (int ax = 0; ax < 3; ax++) {
                int best = dir[ax] > 0 ? n[ax] : -1;
                for (int s = 0; s < nshot; s++) {
                    int st0, l;
                    sweep_axis(dir[ax], w.init, n[ax], hsid[3 * s + ax], st0, l);
                    best = dir[ax] > 0 ? std::min(best, st0) : std::max(best, st0);
                }
                start[ax] = best;
                len[ax] = dir[ax] > 0 ? n[ax] - best : best + 1;
            }
            const int zl = len[0], xl = len[1], yl = len[2];
            const int P = zl + xl + yl - 2;
            for (int p = 0; p < P; p++) {
                // a in [p - (xl-1) - (yl-1), p] cap [0, zl); c likewise
                const int a0 = std::max(0, p - (xl - 1) - (yl - 1)), a1 = std::min(zl - 1, p);
                const int c0 = std::max(0, p - (xl - 1) - (zl - 1)), c1 = std::min(yl - 1, p);
                if (a1 < a0 || c1 < c0) continue;
                const int na = a1 - a0 + 1, nc = c1 - c0 + 1;
                const long long nt = (long long)na * nc * NS;
                hipLaunchKernelGGL(fsm_plane, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0,
                                   ctx->stream, T, NS, nshot, S, dsid, g, w, start[0], start[1],
                                   start[2], xl, p, a0, na, c0, nc);
            }
            ST_HIP(hipGetLastError());
        }
    }
    ST_HIP(hipEventRecord(ctx->ev[3], ctx->stream));
    ST_HIP(hipEventSynchronize(ctx->ev[3]));
    float ms = 0.f;
    ST_HIP(hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]));
    return ms;
}

void fsm_pad(st_ctx *ctx, float *S) {
    const long long n = (long long)(ctx->nx + 2) * (ctx->ny + 2) * (ctx->nz + 2);
    hipLaunchKernelGGL(fsm_pad_model, dim3(grid_for(n)), dim3(256), 0, ctx->stream,
                       ctx->slow_zxy.as<float>(), S, ctx->nx, ctx->ny, ctx->nz);
    ST_HIP(hipGetLastError());
}

void fsm_picks(st_ctx *ctx, int nshot, const float *drcv, int nrcv, float *dpicks) {
    launch_picks_pad(ctx->stream, ctx->T.as<float>(), fsm_lanes(nshot), nshot, drcv, nrcv, dpicks,
                     ctx->nx, ctx->ny, ctx->nz, ctx->dx, ctx->dy, ctx->dz);
}

void fsm_crop_volumes(st_ctx *ctx, float *dout, int nshot) {
    const long long n = (long long)ctx->nx * ctx->ny * ctx->nz * nshot;
    hipLaunchKernelGGL(fsm_crop, dim3(grid_for(n)), dim3(256), 0, ctx->stream, ctx->T.as<float>(),
                       fsm_lanes(nshot), dout, ctx->nx, ctx->ny, ctx->nz, nshot);
    ST_HIP(hipGetLastError());
}

}  // namespace st
