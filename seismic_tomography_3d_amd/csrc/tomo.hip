// tomo.hip -- least-squares traveltime tomography on the device.
//
// Replaces, in src/tomography/least_squares/least_squares.cpp:
//   expand_fdm (:11-67)       -- not materialised: the tracer reads the FIM
//                                volume through a clamp-to-edge index map,
//                                which is what the one-cell replication pad is;
//   ray_tracing (:192-280)    -- one thread per ray, count -> scan -> fill CSR;
//   optimization (:321-356),
//   tk_reg_matrix (:358-422)  -- Tikhonov rows appended on the device;
//   sparse_cgls (:424-496)    -- CSR SpMV, SpMV^T from an explicit transposed
//                                matrix (deterministic, no atomics), FP64 dots;
//   converged (:301-319)      -- FP64 residual norm;
//   model_update (:498-570),
//   gaussian (src/utils/smoothing/gaussian.cpp:5-73).
//
// Compiled with -ffp-contract=off and correctly rounded division/sqrt: the
// tracer, the model update and the smoothing reproduce the CPU formulas bit
// for bit on the same input volume.
#include <algorithm>
#include <cmath>
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>

#include "st_common.hpp"

namespace st {

// ---------------------------------------------------------------------------
// state
// ---------------------------------------------------------------------------
struct TomoState {
    bool grid_set = false;
    int nxt = 0, nyt = 0, nzt = 0;
    float dxt = 0, dyt = 0, dzt = 0;
    int max_steps = 0, max_runs = 0;

    // sensitivity matrix G (CSR), grown batch by batch
    DevBuf rowptr, cols, vals;     // int64 rows+1, int32, float
    long long rows = 0, nnz = 0;
    long long cap_rows = 0, cap_nnz = 0;

    // per-ray scratch of the current batch
    DevBuf runs, nruns, fdist, rowlen, scanbuf;
    DevBuf err;                    // unsigned error word
    DevBuf illum;                  // nx*ny*nz
    bool illum_valid = false;

    // least squares
    int n_model = 0, n_data = 0, order = 0, N = 0, M = 0;
    long long annz = 0;
    DevBuf t_colptr, t_rows, t_vals, t_cursor;
    int n_long = 0;                // A^T columns reduced by a whole workgroup
    // A^T in row slices (large systems): per column kSlices + 1 entry bounds,
    // kSlices partial sums per column, the long pieces per slice
    DevBuf t_bounds, t_part, t_slong;
    bool sliced = false;
    int n_slong = 0;               // most long pieces in one slice
    DevBuf t_keys, t_idx0, t_idx1, t_sort;  // column-sort staging (radix sort pairs)
    DevBuf b, s, q, r, p, x, dots, scal;
    bool have_x = false;

    DevBuf dm, upd, kern;
    DevBuf gkern;                  // Gaussian weights read by the captured update
    hipGraphExec_t gexec = nullptr;   // CGLS + model update, captured
    std::vector<long long> gkey;      // what gexec was captured for
};

TomoState &tomo(st_ctx *ctx) {
    if (!ctx->tomo) ctx->tomo = new TomoState();
    return *static_cast<TomoState *>(ctx->tomo);
}

void tomo_release(st_ctx *ctx) {
    if (!ctx->tomo) return;
    TomoState &t = *static_cast<TomoState *>(ctx->tomo);
    for (DevBuf *b : {&t.rowptr, &t.cols, &t.vals, &t.runs, &t.nruns, &t.fdist, &t.rowlen,
                      &t.scanbuf, &t.err, &t.illum, &t.t_colptr, &t.t_rows, &t.t_vals,
                      &t.t_cursor, &t.t_bounds, &t.t_part, &t.t_slong, &t.t_keys, &t.t_idx0, &t.t_idx1, &t.t_sort, &t.b, &t.s, &t.q, &t.r, &t.p, &t.x, &t.dots, &t.scal, &t.dm,
                      &t.upd, &t.kern, &t.gkern})
        b->release();
    if (t.gexec) (void)hipGraphExecDestroy(t.gexec);
    delete &t;
    ctx->tomo = nullptr;
}

// grow a device buffer keeping its first `keep` bytes
static void grow_keep(DevBuf &b, size_t need, size_t keep, hipStream_t st) {
    if (need <= b.bytes) return;
    size_t nb = std::max(need, b.bytes + b.bytes / 2);
    void *np = nullptr;
    ST_HIP(hipMalloc(&np, nb));
    if (keep && b.ptr) ST_HIP(hipMemcpyAsync(np, b.ptr, keep, hipMemcpyDeviceToDevice, st));
    ST_HIP(hipStreamSynchronize(st));
    if (b.ptr) ST_HIP(hipFree(b.ptr));
    b.ptr = np;
    b.bytes = nb;
}

// ---------------------------------------------------------------------------
// ray tracing
// ---------------------------------------------------------------------------
struct ZxyLay {
    int nx, nz;
    __device__ long long operator()(int x, int y, int z) const {
        return z + (long long)x * nz + (long long)y * nx * nz;
    }
};
// the sweeping solvers' volumes, padded by one cell on every side (the
// reference's expand_fdm extent); cells are cs floats apart (Accurate_FSM
// keeps the shots of a batch fastest, cs = its shot lanes; Podvin cs = 1)
struct PadLay {
    int nxx, nzz, cs;
    __device__ long long operator()(int x, int y, int z) const {
        return ((long long)(z + 1) + (long long)(x + 1) * nzz + (long long)(y + 1) * nxx * nzz) *
               cs;
    }
};
// solver block layout for a power-of-two block edge 1 << lb and in-volume
// (non-negative) coordinates: shifts and masks instead of the three integer
// divisions of blk_addr, and 32-bit offsets (the caller checks that one
// shot's volume holds fewer than 2^31 cells) -- six lookups per ray step
struct BlkLay {
    int lb, nbx, nby;
    __device__ unsigned operator()(int x, int y, int z) const {
        const unsigned m = (1u << lb) - 1u;
        const unsigned b = (unsigned)(x >> lb) +
                           (unsigned)nbx * ((unsigned)(y >> lb) + (unsigned)nby * (unsigned)(z >> lb));
        return (b << (3 * lb)) + ((unsigned)x & m) + (((unsigned)y & m) << lb) +
               (((unsigned)z & m) << (2 * lb));
    }
};

// RN(a / b) for b > 0 from y = RN(1 / b) without a division: q = a y, one
// fma residual correction makes q faithful, and Markstein's final step (the
// residual a - b q is exact by fma, q + r y rounded once) returns the
// correctly rounded quotient.  Where intermediate results could leave the
// normal range (|a / b| < 2^-100), the quotient's integer part, the only use
// here, is 0 either way.  Checked against IEEE division on every float
// |a| in [2^-20, 2^20) for 16 spacings from 1e-3 to 9999: all bits equal.
__device__ __forceinline__ float div_rn(float a, float b, float y) {
    float q = a * y;
    float r = __builtin_fmaf(-q, b, a);
    q = __builtin_fmaf(r, y, q);
    r = __builtin_fmaf(-q, b, a);
    return __builtin_fmaf(r, y, q);
}

struct RayArgs {
    const float *T;
    long long vol;
    int nx, ny, nz;
    float dx, dy, dz;
    int nxt, nyt, nzt;
    float dxt, dyt, dzt;
    // correctly rounded reciprocals of the six spacings (host float division)
    float rdx, rdy, rdz, rdxt, rdyt, rdzt;
    const float *src;   // 3 per shot of this batch
    const float *rcv;   // 3 per receiver
    int nr, nshot;
    int max_steps, max_runs;
    int2 *runs;         // (cell, count) per ray, max_runs each
    int *nruns;         // merged entries per ray
    float *fdist;       // finalDist per ray
    float *illum;       // nullable
    unsigned *err;      // bit 0: step cap / lost gradient, bit 1: run capacity
};

__device__ __forceinline__ int iclampd(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Workgroups are dealt to the 8 XCDs round-robin; the launch is padded to
// 8 * per groups and group g takes the ray range (g % 8) * per + g / 8, so
// each XCD traces a contiguous run of rays (one or two shots) and its L2
// serves the cells those rays share near their sources.
constexpr int kXcd = 8;
constexpr int kRing = 16;   // runs / illumination additions a lane buffers in LDS

// SEG = false: one step per iteration, the next step's node loaded one step
// ahead; SEG = true: node segments (below), one memory round trip per node
// but lanes whose segments differ in length wait for the longest.  The
// second wins when many waves per SIMD share the latency, the first when a
// launch has too few rays to fill the SIMDs.
template <class L, bool SEG>
__global__ void ray_trace(RayArgs a, L lay, int per) {
    const long long grp = (long long)(blockIdx.x % kXcd) * per + blockIdx.x / kXcd;
    const long long ray = grp * blockDim.x + threadIdx.x;
    if (ray >= (long long)a.nshot * a.nr) return;
    const int s = (int)(ray / a.nr), r = (int)(ray - (long long)s * a.nr);
    const float *Ts = a.T + (long long)s * a.vol;
    const float sx = a.src[3 * s], sy = a.src[3 * s + 1], sz = a.src[3 * s + 2];
    const int sIdz = (int)div_rn(sz, a.dzt, a.rdzt), sIdx = (int)div_rn(sx, a.dxt, a.rdxt),
              sIdy = (int)div_rn(sy, a.dyt, a.rdyt);
    const int sId = sIdz + sIdx * a.nzt + sIdy * a.nxt * a.nzt;
    const float rayStep = 0.2f * a.dz;
    // expand_fdm pad: padded index p maps to model index clamp(p - 1)
    auto Tp = [&](int i, int j, int k) -> float {
        return Ts[lay(iclampd(j - 1, 0, a.nx - 1), iclampd(k - 1, 0, a.ny - 1),
                      iclampd(i - 1, 0, a.nz - 1))];
    };
    float zi = a.rcv[3 * r + 2], xi = a.rcv[3 * r], yi = a.rcv[3 * r + 1];
    int2 *my = a.runs + ray * a.max_runs;
    // Finished runs and illumination additions wait in per-lane LDS rings
    // and go to memory for all active lanes of the wave at once, when one
    // lane's ring is full: a store or atomic inside the step loop would
    // make the next step's loads wait for it to complete.
    __shared__ int2 run_ring[kRing][64];
    __shared__ int2 il_ring[kRing][64];
    const int lane = threadIdx.x;   // one wave per workgroup
    int nrun = 0, nb = 0, ni = 0;   // runs in memory, runs and additions buffered
    auto flush = [&]() {
        for (int u = 0; u < nb; u++) my[nrun + u] = run_ring[u][lane];
        nrun += nb;
        nb = 0;
        for (int u = 0; u < ni; u++) {
            const int2 e = il_ring[u][lane];
            atomicAdd(&a.illum[e.x], __int_as_float(e.y));
        }
        ni = 0;
    };
    int cur = -1, cnt = 0;
    long long il_cur = -1;
    float il_acc = 0.f;
    const long long npts = (long long)a.nx * a.ny * a.nz;
    bool ok = true;
    // padded node of a position, and the six neighbour values the central
    // differences at that node read
    auto node = [&](float z, float x, float y, int &i, int &j, int &k) {
        i = iclampd((int)div_rn(z, a.dz, a.rdz) + 1, 1, a.nz);
        j = iclampd((int)div_rn(x, a.dx, a.rdx) + 1, 1, a.nx);
        k = iclampd((int)div_rn(y, a.dy, a.rdy) + 1, 1, a.ny);
    };
    auto load6 = [&](int i, int j, int k, float (&t)[6]) {
        t[0] = Tp(i + 1, j, k);
        t[1] = Tp(i - 1, j, k);
        t[2] = Tp(i, j + 1, k);
        t[3] = Tp(i, j - 1, k);
        t[4] = Tp(i, j, k + 1);
        t[5] = Tp(i, j, k - 1);
    };
    if constexpr (!SEG) {
    // Software-pipelined by one step: the next node's loads are issued as
    // soon as the new position is known, and travel while this step's run
    // and illumination bookkeeping runs (a node past the last step is
    // clamped into the volume, so its loads stay in bounds).
    int i, j, k;
    float tv[6];
    node(zi, xi, yi, i, j, k);
    load6(i, j, k, tv);
    for (int step = 0;; step++) {
        if (step >= a.max_steps) { ok = false; atomicOr(a.err, 1u); break; }
        const float dTz = (tv[0] - tv[1]) / (2.0f * a.dz);
        const float dTx = (tv[2] - tv[3]) / (2.0f * a.dx);
        const float dTy = (tv[4] - tv[5]) / (2.0f * a.dy);
        const float norm = sqrtf(dTx * dTx + dTy * dTy + dTz * dTz);
        if (!(norm > 0.0f) || !isfinite(norm)) { ok = false; atomicOr(a.err, 1u); break; }
        zi -= rayStep * dTz / norm;
        xi -= rayStep * dTx / norm;
        yi -= rayStep * dTy / norm;
        int i2, j2, k2;
        float tn[6];
        node(zi, xi, yi, i2, j2, k2);
        load6(i2, j2, k2, tn);
        __builtin_amdgcn_sched_barrier(0);
        const int im = iclampd((int)div_rn(zi, a.dzt, a.rdzt), 0, a.nzt - 1);
        const int jm = iclampd((int)div_rn(xi, a.dxt, a.rdxt), 0, a.nxt - 1);
        const int km = iclampd((int)div_rn(yi, a.dyt, a.rdyt), 0, a.nyt - 1);
        const int cell = im + jm * a.nzt + km * a.nxt * a.nzt;
        if (cell != cur) {
            if (cnt) {
                if (nrun + nb >= a.max_runs) { ok = false; atomicOr(a.err, 2u); break; }
                run_ring[nb++][lane] = make_int2(cur, cnt);
            }
            cur = cell;
            cnt = 0;
        }
        cnt++;
        // illumination at the pre-step node with the reference's strides
        // (least_squares.cpp:234); writes outside the volume are dropped
        const long long il = i + (long long)j * a.nz + (long long)k * a.nx * a.nz;
        if (il != il_cur) {
            if (a.illum && il_cur >= 0 && il_cur < npts)
                il_ring[ni++][lane] = make_int2((int)il_cur, __float_as_int(il_acc));
            il_cur = il;
            il_acc = 0.f;
        }
        il_acc += rayStep;
        const bool done = cell == sId;
        if (__ballot(nb == kRing || ni == kRing)) flush();
        if (done) break;
        i = i2;
        j = j2;
        k = k2;
#pragma unroll
        for (int q = 0; q < 6; q++) tv[q] = tn[q];
    }
    } else {
    // Node segments.  While a ray stays in one padded node its six values,
    // hence its step (rayStep * dT / norm per axis), do not change, so the
    // positions of the whole segment follow from ALU work alone: a first
    // pass walks them to the step that enters the next node and issues that
    // node's six loads at once; a second pass replays the same steps (the
    // same FP32 operations, so the same positions) for the cell runs and
    // the illumination while the loads travel.  One memory round trip per
    // node instead of per step.
    int i, j, k;
    float tv[6];
    node(zi, xi, yi, i, j, k);
    load6(i, j, k, tv);
    int step = 0;
    for (;;) {
        const float dTz = (tv[0] - tv[1]) / (2.0f * a.dz);
        const float dTx = (tv[2] - tv[3]) / (2.0f * a.dx);
        const float dTy = (tv[4] - tv[5]) / (2.0f * a.dy);
        const float norm = sqrtf(dTx * dTx + dTy * dTy + dTz * dTz);
        if (!(norm > 0.0f) || !isfinite(norm)) { ok = false; atomicOr(a.err, 1u); break; }
        const float cz = rayStep * dTz / norm, cx = rayStep * dTx / norm, cy = rayStep * dTy / norm;
        // pass 1: the segment's length m (steps until the node changes)
        float qz = zi, qx = xi, qy = yi;
        int m = 0, i2 = i, j2 = j, k2 = k;
        bool capped = false;
        for (;;) {
            if (step + m >= a.max_steps) { capped = true; break; }
            qz -= cz;
            qx -= cx;
            qy -= cy;
            ++m;
            node(qz, qx, qy, i2, j2, k2);
            if (i2 != i || j2 != j || k2 != k) break;
        }
        // the next node (in bounds: nodes are clamped); past the step cap
        // the replay below may still reach the source cell first
        float tn[6];
        if (!capped) load6(i2, j2, k2, tn);
        __builtin_amdgcn_sched_barrier(0);
        // pass 2: the m steps again, with the bookkeeping of each
        bool done = false;
        const long long il = i + (long long)j * a.nz + (long long)k * a.nx * a.nz;
        for (int u = 0; u < m && !done; u++) {
            zi -= cz;
            xi -= cx;
            yi -= cy;
            const int im = iclampd((int)div_rn(zi, a.dzt, a.rdzt), 0, a.nzt - 1);
            const int jm = iclampd((int)div_rn(xi, a.dxt, a.rdxt), 0, a.nxt - 1);
            const int km = iclampd((int)div_rn(yi, a.dyt, a.rdyt), 0, a.nyt - 1);
            const int cell = im + jm * a.nzt + km * a.nxt * a.nzt;
            if (cell != cur) {
                if (cnt) {
                    if (nrun + nb >= a.max_runs) { ok = false; atomicOr(a.err, 2u); break; }
                    run_ring[nb++][lane] = make_int2(cur, cnt);
                }
                cur = cell;
                cnt = 0;
            }
            cnt++;
            // illumination at the pre-step node with the reference's strides
            // (least_squares.cpp:234); writes outside the volume are dropped
            if (il != il_cur) {
                if (a.illum && il_cur >= 0 && il_cur < npts)
                    il_ring[ni++][lane] = make_int2((int)il_cur, __float_as_int(il_acc));
                il_cur = il;
                il_acc = 0.f;
            }
            il_acc += rayStep;
            done = cell == sId;
            if (nb == kRing || ni == kRing) flush();
        }
        if (!ok) break;
        step += m;
        if (done) break;
        if (capped) { ok = false; atomicOr(a.err, 1u); break; }
        i = i2;
        j = j2;
        k = k2;
#pragma unroll
        for (int q = 0; q < 6; q++) tv[q] = tn[q];
    }
    }
    flush();
    if (a.illum && il_cur >= 0 && il_cur < npts) atomicAdd(&a.illum[il_cur], il_acc);
    if (!ok) { a.nruns[ray] = 0; a.fdist[ray] = 0.f; return; }
    if (nrun >= a.max_runs) { atomicOr(a.err, 2u); a.nruns[ray] = 0; return; }
    my[nrun++] = make_int2(cur, cnt);
    a.fdist[ray] = sqrtf((zi - sz) * (zi - sz) + (xi - sx) * (xi - sx) + (yi - sy) * (yi - sy));

    // sort the runs by cell (shell sort, Ciura gaps) and merge repeats: the
    // std::sort + run-length pass of least_squares.cpp:241-276
    const int gaps[8] = {701, 301, 132, 57, 23, 10, 4, 1};
    for (int gi = 0; gi < 8; gi++) {
        const int gap = gaps[gi];
        for (int u = gap; u < nrun; u++) {
            const int2 tmp = my[u];
            int w = u;
            while (w >= gap && my[w - gap].x > tmp.x) { my[w] = my[w - gap]; w -= gap; }
            my[w] = tmp;
        }
    }
    int m = 0;
    for (int u = 0; u < nrun; u++) {
        if (m > 0 && my[m - 1].x == my[u].x) my[m - 1].y += my[u].y;
        else my[m++] = my[u];
    }
    a.nruns[ray] = m;
}

// CSR values: first (smallest) cell gets rayStep + n additions, later cells
// rayStep + (n-1) additions (sequential FP32 sums), the source cell finalDist.
// One wave per ray, one lane per run: the runs are read and the row written
// as contiguous 64-entry spans; each lane's sum stays sequential.
constexpr int kFillRays = 4;  // rays (waves) per 256-thread workgroup

__global__ void __launch_bounds__(64 * kFillRays)
ray_fill(const int2 *__restrict__ runs, const int *__restrict__ nruns,
         const float *__restrict__ fdist, const long long *__restrict__ rowptr,
         int *__restrict__ cols, float *__restrict__ vals, long long nrays, int max_runs,
         const float *__restrict__ src, int nr, float dz, float dxt, float dyt, float dzt,
         int nxt, int nzt) {
    const long long ray = (long long)blockIdx.x * kFillRays + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (ray >= nrays) return;
    const int s = (int)(ray / nr);
    const float sx = src[3 * s], sy = src[3 * s + 1], sz = src[3 * s + 2];
    const int sId = (int)(sz / dzt) + (int)(sx / dxt) * nzt + (int)(sy / dyt) * nxt * nzt;
    const float rayStep = 0.2f * dz;
    const int2 *my = runs + ray * max_runs;
    const long long o = rowptr[ray];
    const int m = nruns[ray];
    const float fd = fdist[ray];
    for (int u = lane; u < m; u += 64) {
        const int2 r = my[u];
        float d = rayStep;
        const int adds = (u == 0) ? r.y : r.y - 1;
        for (int q = 0; q < adds; q++) d += rayStep;
        cols[o + u] = r.x;
        vals[o + u] = (r.x == sId) ? fd : d;
    }
}

// ---------------------------------------------------------------------------
// scan (exclusive, int -> int64), three phases
// ---------------------------------------------------------------------------
constexpr int kScanTile = 1024;

__global__ void scan_tiles(const int *__restrict__ in, long long *__restrict__ out,
                           long long *__restrict__ tile_sums, long long n) {
    __shared__ long long sh[kScanTile];
    const long long base = (long long)blockIdx.x * kScanTile;
    const int t = threadIdx.x;  // 256 threads, 4 items each
    long long v[4];
    long long acc = 0;
    for (int k = 0; k < 4; k++) {
        const long long i = base + t * 4 + k;
        v[k] = (i < n) ? in[i] : 0;
        acc += v[k];
    }
    sh[t] = acc;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        long long add = (t >= off) ? sh[t - off] : 0;
        __syncthreads();
        sh[t] += add;
        __syncthreads();
    }
    long long run = sh[t] - acc;
    for (int k = 0; k < 4; k++) {
        const long long i = base + t * 4 + k;
        if (i < n) out[i] = run;
        run += v[k];
    }
    if (t == 255) tile_sums[blockIdx.x] = sh[255];
}

__global__ void scan_sums(long long *tile_sums, int ntiles, long long *total) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    long long run = 0;
    for (int i = 0; i < ntiles; i++) {
        const long long v = tile_sums[i];
        tile_sums[i] = run;
        run += v;
    }
    *total = run;
}

__global__ void scan_add(long long *out, const long long *tile_sums, long long n, long long offset) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += tile_sums[i / kScanTile] + offset;
}

// exclusive scan of n ints into out[0..n) plus offset; returns the total.
static long long exclusive_scan(st_ctx *ctx, const int *in, long long *out, long long n,
                                long long offset, DevBuf &tmp) {
    const int ntiles = (int)((n + kScanTile - 1) / kScanTile);
    tmp.reserve(sizeof(long long) * (ntiles + 2));
    long long *sums = tmp.as<long long>();
    long long *total = sums + ntiles;
    hipLaunchKernelGGL(scan_tiles, dim3(ntiles), dim3(256), 0, ctx->stream, in, out, sums, n);
    hipLaunchKernelGGL(scan_sums, dim3(1), dim3(1), 0, ctx->stream, sums, ntiles, total);
    hipLaunchKernelGGL(scan_add, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, out,
                       sums, n, offset);
    ST_HIP(hipGetLastError());
    long long *h = static_cast<long long *>(ctx->host_pinned) + 64;
    ST_HIP(hipMemcpyAsync(h, total, sizeof(long long), hipMemcpyDeviceToHost, ctx->stream));
    ST_HIP(hipStreamSynchronize(ctx->stream));
    return *h;
}

// Traces nshot*nr rays on T (layout lay) and appends their rows to G.
template <class L>
static void trace_append(st_ctx *ctx, const float *T, long long vol, L lay, const float *dsrc,
                         int nshot, const float *drcv, int nr, bool with_illum) {
    TomoState &t = tomo(ctx);
    const long long nrays = (long long)nshot * nr;
    t.runs.reserve(sizeof(int2) * nrays * t.max_runs);
    t.nruns.reserve(sizeof(int) * nrays);
    t.fdist.reserve(sizeof(float) * nrays);
    t.err.reserve(64);
    ST_HIP(hipMemsetAsync(t.err.ptr, 0, 4, ctx->stream));
    RayArgs a;
    a.T = T; a.vol = vol;
    a.nx = ctx->nx; a.ny = ctx->ny; a.nz = ctx->nz;
    a.dx = ctx->dx; a.dy = ctx->dy; a.dz = ctx->dz;
    a.nxt = t.nxt; a.nyt = t.nyt; a.nzt = t.nzt;
    a.dxt = t.dxt; a.dyt = t.dyt; a.dzt = t.dzt;
    a.rdx = 1.0f / a.dx; a.rdy = 1.0f / a.dy; a.rdz = 1.0f / a.dz;
    a.rdxt = 1.0f / a.dxt; a.rdyt = 1.0f / a.dyt; a.rdzt = 1.0f / a.dzt;
    a.src = dsrc; a.rcv = drcv; a.nr = nr; a.nshot = nshot;
    a.max_steps = t.max_steps; a.max_runs = t.max_runs;
    a.runs = t.runs.as<int2>(); a.nruns = t.nruns.as<int>(); a.fdist = t.fdist.as<float>();
    a.illum = with_illum ? t.illum.as<float>() : nullptr;
    a.err = t.err.as<unsigned>();
    ST_REQUIRE((long long)ctx->nx * ctx->ny * ctx->nz < (1LL << 31), ST_ERR_CAPACITY,
               "model too large for ray tracing");
    const int per = (int)(((nrays + 63) / 64 + kXcd - 1) / kXcd);
    // node segments once the rays fill at least two waves per SIMD
    if ((nrays + 63) / 64 >= 2 * 1024) {
        hipLaunchKernelGGL((ray_trace<L, true>), dim3((unsigned)(per * kXcd)), dim3(64), 0,
                           ctx->stream, a, lay, per);
    } else {
        hipLaunchKernelGGL((ray_trace<L, false>), dim3((unsigned)(per * kXcd)), dim3(64), 0,
                           ctx->stream, a, lay, per);
    }
    ST_HIP(hipGetLastError());
    unsigned *herr = static_cast<unsigned *>(ctx->host_pinned) + 200;
    ST_HIP(hipMemcpyAsync(herr, t.err.ptr, 4, hipMemcpyDeviceToHost, ctx->stream));
    ST_HIP(hipStreamSynchronize(ctx->stream));
    ST_REQUIRE(!(*herr & 1u), ST_ERR_DIVERGED,
               "a ray exceeded its step cap or lost its gradient before reaching the source cell");
    ST_REQUIRE(!(*herr & 2u), ST_ERR_CAPACITY, "a ray crossed more cells than max_runs");

    // rows of this batch start at t.rows in the global rowptr
    grow_keep(t.rowptr, sizeof(long long) * (t.rows + nrays + 1), sizeof(long long) * (t.rows + 1),
              ctx->stream);
    long long *rp = t.rowptr.as<long long>() + t.rows;
    const long long add = exclusive_scan(ctx, t.nruns.as<int>(), rp, nrays, t.nnz, t.scanbuf);
    const long long new_nnz = t.nnz + add;
    ST_HIP(hipMemcpyAsync(rp + nrays, &new_nnz, sizeof(long long), hipMemcpyHostToDevice,
                          ctx->stream));
    grow_keep(t.cols, sizeof(int) * new_nnz, sizeof(int) * t.nnz, ctx->stream);
    grow_keep(t.vals, sizeof(float) * new_nnz, sizeof(float) * t.nnz, ctx->stream);
    ST_REQUIRE(nrays < (1LL << 26), ST_ERR_CAPACITY, "too many rays in one ray-tracing batch");
    hipLaunchKernelGGL(ray_fill, dim3((unsigned)((nrays + kFillRays - 1) / kFillRays)),
                       dim3(64 * kFillRays), 0, ctx->stream,
                       t.runs.as<int2>(), t.nruns.as<int>(), t.fdist.as<float>(), rp,
                       t.cols.as<int>(), t.vals.as<float>(), nrays, t.max_runs, dsrc, nr, ctx->dz,
                       t.dxt, t.dyt, t.dzt, t.nxt, t.nzt);
    ST_HIP(hipGetLastError());
    ST_HIP(hipStreamSynchronize(ctx->stream));
    t.rows += nrays;
    t.nnz = new_nnz;
}

void tomo_trace_blk(st_ctx *ctx, const float *T, long long vol, int B, int nbx, int nby,
                    const float *dsrc, int nshot, const float *drcv, int nr, bool with_illum) {
    ST_REQUIRE(B == 4 || B == 8, ST_ERR_ARG, "block edge must be 4 or 8");
    ST_REQUIRE(vol < (1LL << 31), ST_ERR_CAPACITY, "traveltime volume too large for ray tracing");
    trace_append(ctx, T, vol, BlkLay{B == 8 ? 3 : 2, nbx, nby}, dsrc, nshot, drcv, nr, with_illum);
}

// ---------------------------------------------------------------------------
// least squares: Tikhonov rows, transpose, CGLS
// ---------------------------------------------------------------------------

// rows n_data .. n_data + n_model - order - 1 (tk_reg_matrix, :397-417)
__global__ void tikhonov_rows(long long *rowptr, int *cols, float *vals, long long row0,
                              long long nnz0, int nrow, int order, const float *w) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrow) return;
    const int e = order + 1;
    const long long o = nnz0 + (long long)i * e;
    rowptr[row0 + i] = o;
    if (i == nrow - 1) rowptr[row0 + nrow] = o + e;
    for (int k = 0; k < e; k++) {
        cols[o + k] = i + k;
        vals[o + k] = w[k];
    }
}

__global__ void iota_int(int *__restrict__ v, long long n) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) v[e] = (int)e;
}

// Row of every entry of A: 16 lanes per row write its entries' row index
// (coalesced runs of the row's span).
__global__ void entry_rows(const long long *__restrict__ rowptr, int nrows, int *__restrict__ rowid) {
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long row = gid >> 4;
    if (row >= nrows) return;
    const int l = (int)(gid & 15);
    const long long e = rowptr[row + 1];
    for (long long k = rowptr[row] + l; k < e; k += 16) rowid[k] = (int)row;
}

// A^T gather: entry k of the column-sorted order is A's entry e = idx[k], of
// row rowid[e].  The sort is stable and A's entries are in row order, so rows
// ascend inside each column -- the summation order of the reference's
// sequential COO loop (least_squares.cpp:440, :479), which keeps SpMV^T
// deterministic.
__global__ void col_gather(const int *__restrict__ rowid, const float *__restrict__ vals,
                           const int *__restrict__ idx, long long nnz, int *__restrict__ trow,
                           float *__restrict__ tval) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    const long long e = idx[k];
    trow[k] = rowid[e];
    tval[k] = vals[e];
}

// Sum over each group of 16 lanes, the same in all of them: the butterfly of
// lane ^ 8, ^ 4, ^ 2, ^ 1, taken by DPP row rotations (after the step with
// 8 the values repeat every 8 lanes, so rotating by 4 pairs the same lanes
// as ^ 4, and so on; x + y = y + x keeps every lane's bits).
__device__ __forceinline__ float row16_sum(float v) {
    v += __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x128, 0xF, 0xF, false));
    v += __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x124, 0xF, 0xF, false));
    v += __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x122, 0xF, 0xF, false));
    v += __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x121, 0xF, 0xF, false));
    return v;
}

// acc += val[k] * x[idx[k]] for k = k0, k0 + S, ... below e, in that order,
// with four entries' loads (index, value, then the gathered x) in flight at
// a time instead of one: the same sequence of sums, fewer round trips.
template <int S>
__device__ __forceinline__ float strided_dot(const float *__restrict__ val,
                                             const int *__restrict__ idx,
                                             const float *__restrict__ x, long long k, long long e,
                                             float acc) {
    for (; k + 3 * S < e; k += 4 * S) {
        const int c0 = idx[k], c1 = idx[k + S], c2 = idx[k + 2 * S], c3 = idx[k + 3 * S];
        const float v0 = val[k], v1 = val[k + S], v2 = val[k + 2 * S], v3 = val[k + 3 * S];
        const float x0 = x[c0], x1 = x[c1], x2 = x[c2], x3 = x[c3];
        acc += v0 * x0;
        acc += v1 * x1;
        acc += v2 * x2;
        acc += v3 * x3;
    }
    for (; k < e; k += S) acc += val[k] * x[idx[k]];
    return acc;
}

// y[i] = sum_e A[i, col[e]] x[col[e]]: 16 lanes per row, fixed reduction tree
__global__ void spmv_csr(const long long *__restrict__ rowptr, const int *__restrict__ idx,
                         const float *__restrict__ val, const float *__restrict__ x,
                         float *__restrict__ y, int nrows) {
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long row = gid >> 4;
    const int l = threadIdx.x & 15;
    float acc = 0.f;
    if (row < nrows) {
        const long long b = rowptr[row], e = rowptr[row + 1];
        acc = strided_dot<16>(val, idx, x, b + l, e, acc);
    }
    acc = row16_sum(acc);
    if (row < nrows && l == 0) y[row] = acc;
}

// Column pointers of A^T from the column-sorted keys: colptr[c] = the first
// sorted position whose column is >= c (one write per c, no atomics).
__global__ void col_bounds(const unsigned *__restrict__ keys, long long nnz, int ncols,
                           long long *__restrict__ colptr) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > nnz) return;
    const int prev = i > 0 ? (int)keys[i - 1] : -1;
    const int cur = i < nnz ? (int)keys[i] : ncols;
    for (int c = prev + 1; c <= cur; c++) colptr[c] = i;
}

// A^T's columns longer than kLongCol entries (the cells every ray of a shot
// crosses near its source run to thousands): listed once per system, each
// then reduced by a whole workgroup instead of 16 lanes.
constexpr int kLongCol = 256;

__global__ void long_rows(const long long *__restrict__ rowptr, int nrows, int *__restrict__ count,
                          int *__restrict__ list) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row < nrows && rowptr[row + 1] - rowptr[row] > kLongCol) list[atomicAdd(count, 1)] = row;
}

// y = A x for a CSR matrix with a few long rows: blocks [0, nshort) give
// 16 lanes to each row of at most kLongCol entries (as spmv_csr), the
// blocks after them one workgroup to each listed long row (strided partial
// sums, fixed LDS tree) -- both orders are fixed, so results repeat.
__global__ __launch_bounds__(256) void spmv_split(const long long *__restrict__ rowptr,
                                                  const int *__restrict__ idx,
                                                  const float *__restrict__ val,
                                                  const float *__restrict__ x, float *__restrict__ y,
                                                  int nrows, int nshort,
                                                  const int *__restrict__ long_list) {
    if ((int)blockIdx.x < nshort) {
        const long long row = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 4;
        const int l = threadIdx.x & 15;
        float acc = 0.f;
        bool mine = false;
        if (row < nrows) {
            const long long b = rowptr[row], e = rowptr[row + 1];
            mine = e - b <= kLongCol;
            if (mine) acc = strided_dot<16>(val, idx, x, b + l, e, acc);
        }
        acc = row16_sum(acc);
        if (mine && l == 0) y[row] = acc;
        return;
    }
    __shared__ float sh[256];
    const int row = long_list[blockIdx.x - nshort];
    const long long b = rowptr[row], e = rowptr[row + 1];
    float acc = 0.f;
    acc = strided_dot<256>(val, idx, x, b + threadIdx.x, e, acc);
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) y[row] = sh[0];
}

// A^T products of large systems.  The input vector of A^T x has a value per
// ray; beyond an XCD's 4 MB L2 (a million rays) every gathered value costs a
// cache line from farther out.  Each column's entries (rows ascending) are
// cut into kSlices ranges of rows; the workgroups of slice s are those with
// w % kSlices == s, which the dispatcher deals to XCD s, so an XCD gathers
// from one eighth of the vector only.  Slices give partial sums per column,
// added in a fixed order afterwards: results repeat run to run.
constexpr int kSlices = 8;
constexpr long long kSliceRows = 1LL << 20;  // systems with at least this many rows

// bounds[c * (kSlices + 1) + s]: first entry of column c whose row is at
// least s * rps (s = 0 and kSlices: the column's ends); pieces longer than
// kLongCol entries are listed per slice for a whole workgroup each.
__global__ void slice_bounds(const long long *__restrict__ colptr, const int *__restrict__ rows,
                             int ncols, long long rps, long long *__restrict__ bounds,
                             int *__restrict__ slong_cnt, int *__restrict__ slong_list) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncols) return;
    const long long e = colptr[c + 1];
    long long prev = colptr[c];
    bounds[(long long)c * (kSlices + 1)] = prev;
    for (int s = 1; s <= kSlices; s++) {
        long long pos = e;
        if (s < kSlices) {
            const long long target = s * rps;
            long long lo = prev, hi = e;
            while (lo < hi) {
                const long long mid = (lo + hi) >> 1;
                if (rows[mid] < target) lo = mid + 1; else hi = mid;
            }
            pos = lo;
        }
        bounds[(long long)c * (kSlices + 1) + s] = pos;
        if (pos - prev > kLongCol)
            slong_list[(long long)(s - 1) * ncols + atomicAdd(&slong_cnt[s - 1], 1)] = c;
        prev = pos;
    }
}

// part[s * ncols + c] = the slice-s piece of column c times x: workgroups
// j < nshort give 16 lanes to each of 16 columns' pieces of at most
// kLongCol entries, the rest one workgroup to each listed long piece.
__global__ __launch_bounds__(256) void spmv_t_sliced(const long long *__restrict__ bounds,
                                                     const int *__restrict__ idx,
                                                     const float *__restrict__ val,
                                                     const float *__restrict__ x,
                                                     float *__restrict__ part, int ncols, int nshort,
                                                     const int *__restrict__ slong_cnt,
                                                     const int *__restrict__ slong_list) {
    const int s = (int)(blockIdx.x % kSlices);
    const int j = (int)(blockIdx.x / kSlices);
    if (j < nshort) {
        const int c = (j * 256 + (int)threadIdx.x) >> 4;
        const int l = threadIdx.x & 15;
        float acc = 0.f;
        bool mine = false;
        if (c < ncols) {
            const long long b = bounds[(long long)c * (kSlices + 1) + s];
            const long long e = bounds[(long long)c * (kSlices + 1) + s + 1];
            mine = e - b <= kLongCol;
            if (mine) acc = strided_dot<16>(val, idx, x, b + l, e, acc);
        }
        acc = row16_sum(acc);
        if (mine && l == 0) part[(long long)s * ncols + c] = acc;
        return;
    }
    if (j - nshort >= slong_cnt[s]) return;  // the whole workgroup
    __shared__ float sh[256];
    const int c = slong_list[(long long)s * ncols + (j - nshort)];
    const long long b = bounds[(long long)c * (kSlices + 1) + s];
    const long long e = bounds[(long long)c * (kSlices + 1) + s + 1];
    float acc = 0.f;
    acc = strided_dot<256>(val, idx, x, b + threadIdx.x, e, acc);
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(long long)s * ncols + c] = sh[0];
}

// y[c] = the kSlices partial sums of column c, pairwise in a fixed order
__global__ void sum_slices(const float *__restrict__ part, int ncols, float *__restrict__ y) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncols) return;
    float v[kSlices];
#pragma unroll
    for (int s = 0; s < kSlices; s++) v[s] = part[(long long)s * ncols + c];
#pragma unroll
    for (int w = 1; w < kSlices; w <<= 1)
#pragma unroll
        for (int s = 0; s < kSlices; s += 2 * w) v[s] += v[s + w];
    y[c] = v[0];
}

// FP64 dot products: per-block partials, then one ordered pass
__global__ void dot_partial(const float *__restrict__ a, const float *__restrict__ b, long long n,
                            double *__restrict__ part) {
    __shared__ double sh[256];
    double acc = 0.0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
        acc += (double)a[i] * (double)b[i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void dot_final(const double *__restrict__ part, int n, double *__restrict__ out) {
    __shared__ double sh[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) acc += part[i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

constexpr int kDotBlocks = 512;

static void dot(st_ctx *ctx, const float *a, const float *b, long long n, double *part, double *out) {
    hipLaunchKernelGGL(dot_partial, dim3(kDotBlocks), dim3(256), 0, ctx->stream, a, b, n, part);
    hipLaunchKernelGGL(dot_final, dim3(1), dim3(256), 0, ctx->stream, part, kDotBlocks, out);
}

// dot_final, then a CGLS step from the sum and a stored dot (one thread):
// step = sum / other (new_is_num) or other / sum, the FP64 ratio used as an
// FP32 step (a zero denominator -- an exactly solved system -- gives a zero
// step where the reference would produce NaN)
__global__ void dot_final_ratio(const double *__restrict__ part, int n, double *__restrict__ out,
                                const double *__restrict__ other, int new_is_num,
                                float *__restrict__ step) {
    __shared__ double sh[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) acc += part[i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        *out = sh[0];
        const double num = new_is_num ? sh[0] : *other, den = new_is_num ? *other : sh[0];
        *step = (den != 0.0) ? (float)(num / den) : 0.0f;
    }
}

// y1 += sign1 * alpha * x1 (n1 entries) and y2 += sign2 * alpha * x2 (n2
// entries) in one launch
__global__ void axpy2(float *__restrict__ y1, const float *__restrict__ x1, long long n1,
                      float sign1, float *__restrict__ y2, const float *__restrict__ x2,
                      long long n2, float sign2, const float *alpha) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n1) y1[i] += sign1 * (*alpha) * x1[i];
    if (i < n2) y2[i] += sign2 * (*alpha) * x2[i];
}

// p = r + beta * p
__global__ void xpby(float *__restrict__ p, const float *__restrict__ r, const float *beta, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = r[i] + (*beta) * p[i];
}

__global__ void sub_vec(const float *a, const float *b, float *out, long long n, long long ntot) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a[i] - b[i];
    else if (i < ntot) out[i] = 0.f;
}

static unsigned nblk256(long long n) { return (unsigned)((n + 255) / 256); }

// copies and fills as kernels (they may be captured into a graph)
__global__ void copy_vec(float *__restrict__ dst, const float *__restrict__ src, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}
__global__ void zero_vec(float *__restrict__ dst, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = 0.0f;
}

// Assembles A = [G; lambda*L] in place after G and builds A^T.
void tomo_cgls_setup(st_ctx *ctx, int n_model, int order, float lambda) {
    TomoState &t = tomo(ctx);
    const long long n_data = t.rows;
    // the stencil of order k holds the binomial coefficients C(k, j) (up to
    // C(16, 8) = 12870 at the cap, exact in int and in float); the
    // reference's parameter files use orders 0 to 2
    constexpr int kMaxTkOrder = 16;
    ST_REQUIRE(order >= 0 && order <= kMaxTkOrder && order < n_model, ST_ERR_ARG,
               "Tikhonov order must lie in 0..16 and below the model size");
    const int nL = n_model - order;
    const long long N = n_data + nL;
    const long long annz = t.nnz + (long long)(order + 1) * nL;
    // stencil (least_squares.cpp:369-410): order 0 -> [1], else repeated
    // differences of [-1, 1]
    std::vector<int> df(order + 2, 0), d1(order + 3, 0), d2(order + 3, 0);
    if (order == 0) {
        df[0] = 1;
    } else {
        df[0] = -1; df[1] = 1;
        for (int index = 1; index < order; index++)
            for (int k = 0; k < order + 1; k++) {
                d2[k] = df[k];
                d1[k + 1] = df[k];
                df[k] = d1[k] - d2[k];
            }
    }
    float *w = static_cast<float *>(ctx->host_pinned) + 300;
    for (int k = 0; k <= order; k++) w[k] = lambda * (float)df[k];

    grow_keep(t.rowptr, sizeof(long long) * (N + 1), sizeof(long long) * (n_data + 1), ctx->stream);
    grow_keep(t.cols, sizeof(int) * annz, sizeof(int) * t.nnz, ctx->stream);
    grow_keep(t.vals, sizeof(float) * annz, sizeof(float) * t.nnz, ctx->stream);
    t.kern.reserve(sizeof(float) * 64);
    ST_HIP(hipMemcpyAsync(t.kern.ptr, w, sizeof(float) * (order + 1), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(tikhonov_rows, dim3(nblk256(nL)), dim3(256), 0, ctx->stream,
                       t.rowptr.as<long long>(), t.cols.as<int>(), t.vals.as<float>(), n_data,
                       t.nnz, nL, order, t.kern.as<float>());

    // A^T (CSC of A), rows ascending inside each column: a stable radix sort
    // of the entry indices by column (the bits a column index needs), column
    // pointers from the sorted keys, then rows and values gathered in that
    // order
    t.t_colptr.reserve(sizeof(long long) * (n_model + 1));
    t.t_cursor.reserve(sizeof(int) * (n_model + 1));
    t.t_rows.reserve(sizeof(int) * annz);
    t.t_vals.reserve(sizeof(float) * annz);
    ST_REQUIRE(annz < (1LL << 31), ST_ERR_CAPACITY, "too many non-zeros for the transpose");
    t.t_keys.reserve(sizeof(unsigned) * annz);
    t.t_idx0.reserve(sizeof(int) * annz);
    t.t_idx1.reserve(sizeof(int) * annz);
    hipLaunchKernelGGL(iota_int, dim3(nblk256(annz)), dim3(256), 0, ctx->stream,
                       t.t_idx0.as<int>(), annz);
    unsigned bits = 1;
    while (bits < 32 && (1u << bits) < (unsigned)n_model) bits++;
    size_t tmp = 0;
    const unsigned *keys_in = reinterpret_cast<const unsigned *>(t.cols.as<int>());
    ST_HIP(rocprim::radix_sort_pairs(nullptr, tmp, keys_in, t.t_keys.as<unsigned>(),
                                     t.t_idx0.as<int>(), t.t_idx1.as<int>(), (size_t)annz, 0u, bits,
                                     ctx->stream));
    t.t_sort.reserve(std::max<size_t>(tmp, 16));
    ST_HIP(rocprim::radix_sort_pairs(t.t_sort.ptr, tmp, keys_in, t.t_keys.as<unsigned>(),
                                     t.t_idx0.as<int>(), t.t_idx1.as<int>(), (size_t)annz, 0u, bits,
                                     ctx->stream));
    hipLaunchKernelGGL(col_bounds, dim3(nblk256(annz + 1)), dim3(256), 0, ctx->stream,
                       t.t_keys.as<unsigned>(), annz, n_model, t.t_colptr.as<long long>());
    // the radix sort's input indices (t_idx0) are free again: each entry's row
    hipLaunchKernelGGL(entry_rows, dim3(nblk256(N * 16)), dim3(256), 0, ctx->stream,
                       t.rowptr.as<long long>(), (int)N, t.t_idx0.as<int>());
    hipLaunchKernelGGL(col_gather, dim3(nblk256(annz)), dim3(256), 0, ctx->stream,
                       t.t_idx0.as<int>(), t.vals.as<float>(), t.t_idx1.as<int>(), annz,
                       t.t_rows.as<int>(), t.t_vals.as<float>());
    // A^T's long columns: t_cursor[0] counts them, t_cursor[1..] lists them
    hipLaunchKernelGGL(zero_vec, dim3(1), dim3(64), 0, ctx->stream,
                       reinterpret_cast<float *>(t.t_cursor.ptr), 1LL);
    hipLaunchKernelGGL(long_rows, dim3(nblk256(n_model)), dim3(256), 0, ctx->stream,
                       t.t_colptr.as<long long>(), n_model, t.t_cursor.as<int>(),
                       t.t_cursor.as<int>() + 1);
    t.sliced = N >= kSliceRows;
    if (t.sliced) {
        t.t_bounds.reserve(sizeof(long long) * (kSlices + 1) * (size_t)n_model);
        t.t_part.reserve(sizeof(float) * kSlices * (size_t)n_model);
        t.t_slong.reserve(sizeof(int) * kSlices * ((size_t)n_model + 1));
        int *slong_cnt = t.t_slong.as<int>();
        hipLaunchKernelGGL(zero_vec, dim3(1), dim3(64), 0, ctx->stream,
                           reinterpret_cast<float *>(slong_cnt), (long long)kSlices);
        hipLaunchKernelGGL(slice_bounds, dim3(nblk256(n_model)), dim3(256), 0, ctx->stream,
                           t.t_colptr.as<long long>(), t.t_rows.as<int>(), n_model,
                           (N + kSlices - 1) / kSlices, t.t_bounds.as<long long>(), slong_cnt,
                           slong_cnt + kSlices);
    }
    ST_HIP(hipGetLastError());
    int *hcount = static_cast<int *>(ctx->host_pinned) + 64;
    ST_HIP(hipMemcpyAsync(hcount, t.t_cursor.ptr, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    if (t.sliced)
        ST_HIP(hipMemcpyAsync(hcount + 1, t.t_slong.ptr, sizeof(int) * kSlices,
                              hipMemcpyDeviceToHost, ctx->stream));
    ST_HIP(hipStreamSynchronize(ctx->stream));
    t.n_long = *hcount;
    t.n_slong = 0;
    if (t.sliced)
        for (int k = 0; k < kSlices; k++) t.n_slong = std::max(t.n_slong, hcount[1 + k]);

    t.N = (int)N; t.M = n_model; t.annz = annz; t.n_model = n_model; t.order = order;
    t.s.reserve(sizeof(float) * N);
    t.q.reserve(sizeof(float) * N);
    t.r.reserve(sizeof(float) * n_model);
    t.p.reserve(sizeof(float) * n_model);
    t.x.reserve(sizeof(float) * n_model);
    t.dots.reserve(sizeof(double) * (kDotBlocks + 8));
    t.scal.reserve(sizeof(float) * 8);
}

// The CGLS iterations on the assembled system (least_squares.cpp:424-496):
// kernels only (no copies or fills through the runtime), so the sequence
// can be captured into a graph.
void tomo_cgls_iterate(st_ctx *ctx, int iters) {
    TomoState &t = tomo(ctx);
    const long long N = t.N;
    const int n_model = t.n_model;
    double *part = t.dots.as<double>();
    // r.r is kept from the end of one iteration for the next (the reference
    // recomputes it, and again as rd before the update: the same sums)
    double *qTq = part + kDotBlocks, *rTr = qTq + 1, *rTr_next = qTq + 2;
    float *alpha = t.scal.as<float>(), *beta = alpha + 1;

    const long long *rp = t.rowptr.as<long long>();
    const long long *cp = t.t_colptr.as<long long>();
    auto spmv = [&](const float *in, float *out) {  // out = A in
        hipLaunchKernelGGL(spmv_csr, dim3((unsigned)((N * 16 + 255) / 256)), dim3(256), 0,
                           ctx->stream, rp, t.cols.as<int>(), t.vals.as<float>(), in, out, (int)N);
    };
    auto spmvt = [&](const float *in, float *out) {  // out = A^T in
        const int nshort = (int)(((long long)n_model * 16 + 255) / 256);
        if (t.sliced) {
            const int *slong = t.t_slong.as<int>();
            hipLaunchKernelGGL(spmv_t_sliced, dim3((unsigned)(kSlices * (nshort + t.n_slong))),
                               dim3(256), 0, ctx->stream, t.t_bounds.as<long long>(),
                               t.t_rows.as<int>(), t.t_vals.as<float>(), in, t.t_part.as<float>(),
                               n_model, nshort, slong, slong + kSlices);
            hipLaunchKernelGGL(sum_slices, dim3(nblk256(n_model)), dim3(256), 0, ctx->stream,
                               t.t_part.as<float>(), n_model, out);
            return;
        }
        hipLaunchKernelGGL(spmv_split, dim3((unsigned)(nshort + t.n_long)), dim3(256), 0,
                           ctx->stream, cp, t.t_rows.as<int>(), t.t_vals.as<float>(), in, out,
                           n_model, nshort, t.t_cursor.as<int>() + 1);
    };
    auto dot_step = [&](const float *a, long long n, double *out, const double *other,
                        int new_is_num, float *step) {
        hipLaunchKernelGGL(dot_partial, dim3(kDotBlocks), dim3(256), 0, ctx->stream, a, a, n, part);
        hipLaunchKernelGGL(dot_final_ratio, dim3(1), dim3(256), 0, ctx->stream, part, kDotBlocks,
                           out, other, new_is_num, step);
    };
    float *s = t.s.as<float>(), *q = t.q.as<float>(), *r = t.r.as<float>(), *p = t.p.as<float>(),
          *x = t.x.as<float>();
    // s = B (copied in by the caller into t.b), x = 0
    hipLaunchKernelGGL(copy_vec, dim3(nblk256(N)), dim3(256), 0, ctx->stream, s,
                       t.b.as<float>(), N);
    hipLaunchKernelGGL(zero_vec, dim3(nblk256(n_model)), dim3(256), 0, ctx->stream, x,
                       (long long)n_model);
    spmvt(s, r);
    hipLaunchKernelGGL(copy_vec, dim3(nblk256(n_model)), dim3(256), 0, ctx->stream, p, r,
                       (long long)n_model);
    spmv(p, q);
    if (iters > 0) dot(ctx, r, r, n_model, part, rTr);
    for (int it = 0; it < iters; it++) {
        dot_step(q, N, qTq, rTr, 0, alpha);                       // alpha = rTr / qTq
        hipLaunchKernelGGL(axpy2, dim3(nblk256(std::max<long long>(N, n_model))), dim3(256), 0,
                           ctx->stream, x, p, (long long)n_model, 1.0f, s, q, N, -1.0f, alpha);
        spmvt(s, r);
        dot_step(r, n_model, rTr_next, rTr, 1, beta);             // beta = rTr' / rTr
        hipLaunchKernelGGL(xpby, dim3(nblk256(n_model)), dim3(256), 0, ctx->stream, p, r, beta,
                           (long long)n_model);
        spmv(p, q);
        std::swap(rTr, rTr_next);
    }
    ST_HIP(hipGetLastError());
}

// Assembles the system and runs CGLS (eager).
void tomo_cgls(st_ctx *ctx, int n_model, int order, float lambda, int iters) {
    tomo_cgls_setup(ctx, n_model, order, lambda);
    tomo_cgls_iterate(ctx, iters);
    ST_HIP(hipStreamSynchronize(ctx->stream));
    tomo(ctx).have_x = true;
}

// ---------------------------------------------------------------------------
// model update
// ---------------------------------------------------------------------------
__device__ __forceinline__ float trilin(float c000, float c001, float c100, float c101, float c010,
                                        float c011, float c110, float c111, float x0, float x1,
                                        float y0, float y1, float z0, float z1, float x, float y,
                                        float z) {
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

// dm on the model grid by trilinear interpolation of xdm (:500-541)
__global__ void dm_trilinear(const float *__restrict__ xdm, float *__restrict__ dm, int nx, int ny,
                             int nz, float dx, float dy, float dz, int nxt, int nyt, int nzt,
                             float dxt, float dyt, float dzt) {
    const long long index = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (index >= (long long)nx * ny * nz) return;
    const int k = (int)(index / ((long long)nx * nz));
    const int j = (int)(index - (long long)k * nx * nz) / nz;
    const int i = (int)(index - (long long)j * nz - (long long)k * nx * nz);
    const float x = j * dx, y = k * dy, z = i * dz;
    const float x0 = floorf(x / dxt) * dxt, y0 = floorf(y / dyt) * dyt, z0 = floorf(z / dzt) * dzt;
    const float x1 = floorf(x / dxt) * dxt + dxt, y1 = floorf(y / dyt) * dyt + dyt,
                z1 = floorf(z / dzt) * dzt + dzt;
    float v = 0.0f;
    if ((i >= 11) && (i < nz - 2) && (j >= 15) && (j < nx - 16) && (k >= 15) && (k < ny - 16)) {
        const int idz = (int)(z / dzt), idx = (int)(x / dxt), idy = (int)(y / dyt);
        const long long oz = (idz + 1 < nzt) ? 1 : 0;
        const long long ox = (idx + 1 < nxt) ? nzt : 0;
        const long long oy = (idy + 1 < nyt) ? (long long)nxt * nzt : 0;
        const long long m = idz + (long long)idx * nzt + (long long)idy * nxt * nzt;
        v = trilin(xdm[m], xdm[m + oz], xdm[m + ox], xdm[m + oz + ox], xdm[m + oy], xdm[m + oz + oy],
                   xdm[m + ox + oy], xdm[m + oz + ox + oy], x0, x1, y0, y1, z0, z1, x, y, z);
    }
    dm[index] = v;
}

// gaussian (gaussian.cpp:65-90): interior cells get the weighted sum in
// the reference's (yk, xk, zk) order, edge cells keep the input
__global__ void gaussian3d(const float *__restrict__ in, float *__restrict__ out,
                           const float *__restrict__ kern, int nx, int ny, int nz, int samples) {
    const long long index = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (index >= (long long)nx * ny * nz) return;
    const int k = (int)(index / ((long long)nx * nz));
    const int j = (int)(index - (long long)k * nx * nz) / nz;
    const int i = (int)(index - (long long)j * nz - (long long)k * nx * nz);
    const int init = samples / 2;
    if (k < init || k >= ny - init || j < init || j >= nx - init || i < init || i >= nz - init) {
        out[index] = in[index];
        return;
    }
    float accum = 0.0f;
    for (int yk = 0; yk < samples; yk++)
        for (int xk = 0; xk < samples; xk++)
            for (int zk = 0; zk < samples; zk++) {
                const int w = zk + xk * samples + yk * samples * samples;
                const long long src = (i - init + zk) + (long long)(j - init + xk) * nz +
                                      (long long)(k - init + yk) * nx * nz;
                accum += in[src] * kern[w];
            }
    out[index] = accum;
}

// The same filter for a compile-time window S, tiled: a 256-thread
// workgroup owns a 32 (z) x 8 (x) column of outputs and marches along y
// through kGaussYC planes, keeping the S input planes the current output
// plane reads (with their S/2 halos, zero outside the volume) in an LDS
// ring, so each input value comes from memory about once per column
// instead of S^3 times.  Every interior cell adds its S^3 products in the
// reference's (yk, xk, zk) order, multiply then add, as gaussian3d does.
constexpr int kGaussTZ = 32, kGaussTX = 8, kGaussYC = 16;
template <int S>
__global__ __launch_bounds__(256) void gaussian3d_tiled(const float *__restrict__ in,
                                                        float *__restrict__ out,
                                                        const float *__restrict__ kern, int nx,
                                                        int ny, int nz) {
    constexpr int H = S / 2, PZ = kGaussTZ + S - 1, PX = kGaussTX + S - 1, PL = PZ * PX;
    __shared__ float ring[S][PL];
    const int tz = threadIdx.x % kGaussTZ, tx = threadIdx.x / kGaussTZ;
    const int z0 = blockIdx.x * kGaussTZ, x0 = blockIdx.y * kGaussTX, y0 = blockIdx.z * kGaussYC;
    const int y1 = min(y0 + kGaussYC, ny);
    const int i = z0 + tz, j = x0 + tx;
    const long long pstride = (long long)nx * nz;
    // input plane yp with its halo into ring slot `slot`
    auto load = [&](int yp, int slot) {
        for (int e = threadIdx.x; e < PL; e += 256) {
            const int zz = z0 - H + e % PZ, xx = x0 - H + e / PZ;
            float v = 0.0f;
            if (yp >= 0 && yp < ny && zz >= 0 && zz < nz && xx >= 0 && xx < nx)
                v = in[zz + (long long)xx * nz + (long long)yp * pstride];
            ring[slot][e] = v;
        }
    };
    // plane p sits in slot (p - y0 + H) mod S
    for (int q = 0; q < S - 1; q++) load(y0 - H + q, q);
    for (int y = y0; y < y1; y++) {
        load(y + H, (y - y0 + S - 1) % S);
        __syncthreads();
        if (i < nz && j < nx) {
            const long long o = i + (long long)j * nz + (long long)y * pstride;
            const int base = (y - y0) % S;  // slot of plane y - H
            if (y < H || y >= ny - H || j < H || j >= nx - H || i < H || i >= nz - H) {
                out[o] = ring[(base + H) % S][(tx + H) * PZ + tz + H];
            } else {
                float accum = 0.0f;
#pragma unroll
                for (int yk = 0; yk < S; yk++) {
                    const float *pl = ring[(base + yk) % S] + tx * PZ + tz;
#pragma unroll
                    for (int xk = 0; xk < S; xk++)
#pragma unroll
                        for (int zk = 0; zk < S; zk++)
                            accum += pl[xk * PZ + zk] * kern[zk + xk * S + yk * S * S];
                }
                out[o] = accum;
            }
        }
        __syncthreads();  // the next plane overwrites slot `base`
    }
}

static void launch_gaussian(hipStream_t st, const float *in, float *out, const float *kern, int nx,
                            int ny, int nz, int samples) {
    const long long n = (long long)nx * ny * nz;
    const dim3 g((nz + kGaussTZ - 1) / kGaussTZ, (nx + kGaussTX - 1) / kGaussTX,
                 (ny + kGaussYC - 1) / kGaussYC);
    ST_REQUIRE(g.y <= 65535u && g.z <= 65535u, ST_ERR_ARG, "model too large for the filter grid");
    switch (samples) {
    case 3: hipLaunchKernelGGL(gaussian3d_tiled<3>, g, dim3(256), 0, st, in, out, kern, nx, ny, nz); break;
    case 5: hipLaunchKernelGGL(gaussian3d_tiled<5>, g, dim3(256), 0, st, in, out, kern, nx, ny, nz); break;
    case 7: hipLaunchKernelGGL(gaussian3d_tiled<7>, g, dim3(256), 0, st, in, out, kern, nx, ny, nz); break;
    case 9: hipLaunchKernelGGL(gaussian3d_tiled<9>, g, dim3(256), 0, st, in, out, kern, nx, ny, nz); break;
    default:
        hipLaunchKernelGGL(gaussian3d, dim3(nblk256(n)), dim3(256), 0, st, in, out, kern, nx, ny, nz,
                           samples);
    }
}

// slowness += dm inside the margins (:553-565), velocity = 1 / slowness
__global__ void apply_update(float *__restrict__ slow, const float *__restrict__ dm,
                             float *__restrict__ vel, int nx, int ny, int nz) {
    const long long index = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (index >= (long long)nx * ny * nz) return;
    const int k = (int)(index / ((long long)nx * nz));
    const int j = (int)(index - (long long)k * nx * nz) / nz;
    const int i = (int)(index - (long long)j * nz - (long long)k * nx * nz);
    float s = slow[index];
    if ((i >= 11) && (i < nz - 2) && (j >= 15) && (j < nx - 15) && (k >= 15) && (k < ny - 15)) {
        s += dm[index];
        slow[index] = s;
    }
    vel[index] = 1.0f / s;
}

// Gaussian kernel weights as gaussian.cpp:7-43 builds them (host, FP32).
static void gaussian_weights(int samples, float stdv, std::vector<float> &kernel) {
    const int init = samples / 2;
    const int nK = samples * samples * samples;
    const float pi = 4.0f * atanf(1.0f);
    kernel.assign(nK, 0.0f);
    const int mid = samples / 2;
    kernel[mid + mid * samples + mid * samples * samples] = 1.0f;
    if (stdv != 0.0f) {
        float sum = 0.0f;
        for (int y = -init; y <= init; y++)
            for (int x = -init; x <= init; x++)
                for (int z = -init; z <= init; z++) {
                    const int index = (z + init) + (x + init) * samples + (y + init) * samples * samples;
                    kernel[index] = 1.0f / (2.0f * pi * stdv * stdv) *
                                    expf(-((x * x + y * y + z * z) / (2.0f * stdv * stdv)));
                    sum += kernel[index];
                }
        for (int i = 0; i < nK; i++) kernel[i] /= sum;
    }
}

// model_update kernels (least_squares.cpp:498-570): dm from the solution,
// optional Gaussian smoothing with weights already on the device, slowness
// update; the updated velocity lands in t.upd, dm in the returned buffer.
// Kernels only, so the sequence can be captured into a graph.
static float *update_kernels(st_ctx *ctx, TomoState &t, const float *dxdm, int smooth, int samples,
                             const float *kern) {
    const long long n = (long long)ctx->nx * ctx->ny * ctx->nz;
    hipLaunchKernelGGL(dm_trilinear, dim3(nblk256(n)), dim3(256), 0, ctx->stream, dxdm,
                       t.dm.as<float>(), ctx->nx, ctx->ny, ctx->nz, ctx->dx, ctx->dy, ctx->dz,
                       t.nxt, t.nyt, t.nzt, t.dxt, t.dyt, t.dzt);
    float *dm = t.dm.as<float>(), *tmp = t.upd.as<float>();
    if (smooth) {
        launch_gaussian(ctx->stream, dm, tmp, kern, ctx->nx, ctx->ny, ctx->nz, samples);
        std::swap(dm, tmp);
    }
    hipLaunchKernelGGL(apply_update, dim3(nblk256(n)), dim3(256), 0, ctx->stream,
                       ctx->slow_zxy.as<float>(), dm, tmp, ctx->nx, ctx->ny, ctx->nz);
    ST_HIP(hipGetLastError());
    return dm;
}

// CGLS and the model update as one captured graph (or eagerly): the system
// must be assembled (tomo_cgls_setup) and the right-hand side set.  The
// graph is captured on every call (sizes and buffers follow the current
// matrix) and the instantiated executable updated in place when its shape
// allows, re-instantiated otherwise.
void tomo_solve_update(st_ctx *ctx, int iters, int smooth, int samples, float stdv, bool graph,
                       float *x_out, float *dm_out, float *vel_out) {
    TomoState &t = tomo(ctx);
    const long long n = (long long)ctx->nx * ctx->ny * ctx->nz;
    t.dm.reserve(sizeof(float) * n);
    t.upd.reserve(sizeof(float) * n);
    if (smooth) {
        ST_REQUIRE(samples >= 1 && (samples & 1), ST_ERR_ARG, "filter_samples must be odd");
        std::vector<float> kern;
        gaussian_weights(samples, stdv, kern);
        t.gkern.reserve(sizeof(float) * kern.size());
        ST_HIP(hipMemcpyAsync(t.gkern.ptr, kern.data(), sizeof(float) * kern.size(),
                              hipMemcpyHostToDevice, ctx->stream));
        ST_HIP(hipStreamSynchronize(ctx->stream));   // before the host vector goes
    }
    const float *kern = t.gkern.as<float>();
    float *dm = nullptr;
    // every buffer address and size the captured kernels bake in: an equal
    // key means the instantiated graph can be launched as it is
    const std::vector<long long> key = {
        (long long)t.b.ptr, (long long)t.s.ptr, (long long)t.q.ptr, (long long)t.r.ptr,
        (long long)t.p.ptr, (long long)t.x.ptr, (long long)t.dots.ptr, (long long)t.scal.ptr,
        (long long)t.rowptr.ptr, (long long)t.cols.ptr, (long long)t.vals.ptr,
        (long long)t.t_colptr.ptr, (long long)t.t_rows.ptr, (long long)t.t_vals.ptr,
        (long long)t.t_cursor.ptr, t.n_long, (long long)t.t_bounds.ptr, (long long)t.t_part.ptr,
        (long long)t.t_slong.ptr, t.sliced ? 1LL + t.n_slong : 0LL,
        (long long)t.dm.ptr, (long long)t.upd.ptr, (long long)t.gkern.ptr,
        (long long)ctx->slow_zxy.ptr, t.N, t.n_model, t.annz, iters, smooth, samples, ctx->nx,
        ctx->ny, ctx->nz, t.nxt, t.nyt, t.nzt};
    if (graph && t.gexec && key == t.gkey) {
        ST_HIP(hipGraphLaunch(t.gexec, ctx->stream));
        dm = smooth ? t.upd.as<float>() : t.dm.as<float>();
    } else if (graph) {
        hipGraph_t g = nullptr;
        t.gkey.clear();
        ST_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        try {
            tomo_cgls_iterate(ctx, iters);
            dm = update_kernels(ctx, t, t.x.as<float>(), smooth, samples, kern);
        } catch (...) {
            // leave the stream usable: end the capture and drop its graph
            hipGraph_t partial = nullptr;
            (void)hipStreamEndCapture(ctx->stream, &partial);
            if (partial) (void)hipGraphDestroy(partial);
            (void)hipGetLastError();
            throw;
        }
        ST_HIP(hipStreamEndCapture(ctx->stream, &g));
        bool updated = false;
        if (t.gexec) {
            hipGraphExecUpdateResult res;
            hipGraphNode_t bad = nullptr;
            updated = hipGraphExecUpdate(t.gexec, g, &bad, &res) == hipSuccess &&
                      res == hipGraphExecUpdateSuccess;
            if (!updated) {
                (void)hipGetLastError();
                (void)hipGraphExecDestroy(t.gexec);
                t.gexec = nullptr;
            }
        }
        if (!updated) {
            const hipError_t e = hipGraphInstantiate(&t.gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) t.gexec = nullptr;
            ST_HIP(e);
        } else {
            (void)hipGraphDestroy(g);
        }
        t.gkey = key;
        ST_HIP(hipGraphLaunch(t.gexec, ctx->stream));
    } else {
        tomo_cgls_iterate(ctx, iters);
        dm = update_kernels(ctx, t, t.x.as<float>(), smooth, samples, kern);
    }
    t.have_x = true;
    if (x_out)
        ST_HIP(hipMemcpyAsync(x_out, t.x.ptr, sizeof(float) * t.n_model, hipMemcpyDeviceToHost,
                              ctx->stream));
    if (dm_out)
        ST_HIP(hipMemcpyAsync(dm_out, dm, sizeof(float) * n, hipMemcpyDeviceToHost, ctx->stream));
    if (vel_out) {
        // the updated velocity is in the buffer that did not receive dm
        const float *vel = dm == t.dm.as<float>() ? t.upd.as<float>() : t.dm.as<float>();
        ST_HIP(hipMemcpyAsync(vel_out, vel, sizeof(float) * n, hipMemcpyDeviceToHost,
                              ctx->stream));
    }
    ST_HIP(hipStreamSynchronize(ctx->stream));
    model_changed(ctx);  // derived copies of the slowness are rebuilt on the next solve
}

void tomo_model_update(st_ctx *ctx, const float *dxdm, int smooth, int samples, float stdv,
                       float *dm_out, float *vel_out) {
    TomoState &t = tomo(ctx);
    const long long n = (long long)ctx->nx * ctx->ny * ctx->nz;
    t.dm.reserve(sizeof(float) * n);
    t.upd.reserve(sizeof(float) * n);
    hipLaunchKernelGGL(dm_trilinear, dim3(nblk256(n)), dim3(256), 0, ctx->stream, dxdm,
                       t.dm.as<float>(), ctx->nx, ctx->ny, ctx->nz, ctx->dx, ctx->dy, ctx->dz,
                       t.nxt, t.nyt, t.nzt, t.dxt, t.dyt, t.dzt);
    float *dm = t.dm.as<float>();
    if (smooth) {
        ST_REQUIRE(samples >= 1 && (samples & 1), ST_ERR_ARG, "filter_samples must be odd");
        std::vector<float> kern;
        gaussian_weights(samples, stdv, kern);
        t.kern.reserve(sizeof(float) * kern.size());
        ST_HIP(hipMemcpyAsync(t.kern.ptr, kern.data(), sizeof(float) * kern.size(),
                              hipMemcpyHostToDevice, ctx->stream));
        launch_gaussian(ctx->stream, dm, t.upd.as<float>(), t.kern.as<float>(), ctx->nx, ctx->ny,
                        ctx->nz, samples);
        ST_HIP(hipStreamSynchronize(ctx->stream));  // kern is reused; keep host vector alive
        std::swap(t.dm, t.upd);
        dm = t.dm.as<float>();
    }
    hipLaunchKernelGGL(apply_update, dim3(nblk256(n)), dim3(256), 0, ctx->stream,
                       ctx->slow_zxy.as<float>(), dm, t.upd.as<float>(), ctx->nx, ctx->ny, ctx->nz);
    ST_HIP(hipGetLastError());
    if (dm_out)
        ST_HIP(hipMemcpyAsync(dm_out, dm, sizeof(float) * n, hipMemcpyDeviceToHost, ctx->stream));
    if (vel_out)
        ST_HIP(hipMemcpyAsync(vel_out, t.upd.ptr, sizeof(float) * n, hipMemcpyDeviceToHost,
                              ctx->stream));
    ST_HIP(hipStreamSynchronize(ctx->stream));
    model_changed(ctx);  // derived copies of the slowness are rebuilt on the next solve
}

void tomo_trace_pad(st_ctx *ctx, const float *T, long long shot_stride, int cell_stride,
                    const float *dsrc, int nshot, const float *drcv, int nr, bool with_illum) {
    trace_append(ctx, T, shot_stride, PadLay{ctx->nx + 2, ctx->nz + 2, cell_stride}, dsrc, nshot,
                 drcv, nr, with_illum);
}

void tomo_trace_zxy(st_ctx *ctx, const float *dT, const float *dsrc, const float *drcv, int nr,
                    bool with_illum) {
    trace_append(ctx, dT, (long long)ctx->nx * ctx->ny * ctx->nz, ZxyLay{ctx->nx, ctx->nz}, dsrc,
                 1, drcv, nr, with_illum);
}

// ---------------------------------------------------------------------------
// state helpers used by the C ABI
// ---------------------------------------------------------------------------
bool tomo_grid(st_ctx *ctx, int *n, float *d, int *max_steps, int *max_runs) {
    if (!ctx->tomo) return false;
    TomoState &t = tomo(ctx);
    if (!t.grid_set) return false;
    n[0] = t.nxt; n[1] = t.nyt; n[2] = t.nzt;
    d[0] = t.dxt; d[1] = t.dyt; d[2] = t.dzt;
    *max_steps = t.max_steps;
    *max_runs = t.max_runs;
    return true;
}

void tomo_set_grid(st_ctx *ctx, const int *n, const float *d, int max_steps, int max_runs) {
    TomoState &t = tomo(ctx);
    t.nxt = n[0]; t.nyt = n[1]; t.nzt = n[2];
    t.dxt = d[0]; t.dyt = d[1]; t.dzt = d[2];
    t.max_steps = max_steps;
    t.max_runs = max_runs;
    t.grid_set = true;
    t.have_x = false;
}

void tomo_reset(st_ctx *ctx, bool with_illum) {
    TomoState &t = tomo(ctx);
    t.rows = 0;
    t.nnz = 0;
    t.have_x = false;
    t.rowptr.reserve(sizeof(long long) * 1024);
    ST_HIP(hipMemsetAsync(t.rowptr.ptr, 0, sizeof(long long), ctx->stream));
    if (with_illum) {
        const long long n = (long long)ctx->nx * ctx->ny * ctx->nz;
        t.illum.reserve(sizeof(float) * n);
        ST_HIP(hipMemsetAsync(t.illum.ptr, 0, sizeof(float) * n, ctx->stream));
        t.illum_valid = true;
    }
}

void tomo_get_illum(st_ctx *ctx, float *out) {
    TomoState &t = tomo(ctx);
    ST_REQUIRE(t.illum_valid, ST_ERR_STATE, "no illumination");
    const long long n = (long long)ctx->nx * ctx->ny * ctx->nz;
    ST_HIP(hipMemcpyAsync(out, t.illum.ptr, sizeof(float) * n, hipMemcpyDeviceToHost, ctx->stream));
}

void tomo_csr(st_ctx *ctx, long long *rows, long long *nnz, const long long **rp, const int **c,
              const float **v) {
    TomoState &t = tomo(ctx);
    *rows = t.rows;
    *nnz = t.nnz;
    *rp = t.rowptr.as<long long>();
    *c = t.cols.as<int>();
    *v = t.vals.as<float>();
}

void tomo_load_csr(st_ctx *ctx, const long long *rowptr, const int *cols, const float *vals, int N,
                   long long nnz) {
    TomoState &t = tomo(ctx);
    t.rowptr.reserve(sizeof(long long) * (N + 1));
    t.cols.reserve(sizeof(int) * std::max(1LL, nnz));
    t.vals.reserve(sizeof(float) * std::max(1LL, nnz));
    ST_HIP(hipMemcpyAsync(t.rowptr.ptr, rowptr, sizeof(long long) * (N + 1), hipMemcpyHostToDevice,
                          ctx->stream));
    if (nnz) {
        ST_HIP(hipMemcpyAsync(t.cols.ptr, cols, sizeof(int) * nnz, hipMemcpyHostToDevice, ctx->stream));
        ST_HIP(hipMemcpyAsync(t.vals.ptr, vals, sizeof(float) * nnz, hipMemcpyHostToDevice, ctx->stream));
    }
    ST_HIP(hipStreamSynchronize(ctx->stream));
    t.rows = N;
    t.nnz = nnz;
    t.have_x = false;
}

// t.b = [dobs - dcal, 0 ...] of length Ntot (dcal nullable)
void tomo_set_rhs(st_ctx *ctx, const float *dobs, const float *dcal, int n, long long Ntot) {
    TomoState &t = tomo(ctx);
    t.b.reserve(sizeof(float) * Ntot);
    t.upd.reserve(sizeof(float) * 2 * n);
    float *d0 = t.upd.as<float>(), *d1 = d0 + n;
    ST_HIP(hipMemcpyAsync(d0, dobs, sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
    if (dcal)
        ST_HIP(hipMemcpyAsync(d1, dcal, sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
    else
        ST_HIP(hipMemsetAsync(d1, 0, sizeof(float) * n, ctx->stream));
    hipLaunchKernelGGL(sub_vec, dim3(nblk256(Ntot)), dim3(256), 0, ctx->stream, d0, d1,
                       t.b.as<float>(), (long long)n, Ntot);
    ST_HIP(hipGetLastError());
    ST_HIP(hipStreamSynchronize(ctx->stream));
}

const float *tomo_x(st_ctx *ctx) {
    if (!ctx->tomo) return nullptr;
    TomoState &t = tomo(ctx);
    return t.have_x ? t.x.as<float>() : nullptr;
}

double tomo_residual(st_ctx *ctx, const float *dobs, const float *dcal, int n) {
    TomoState &t = tomo(ctx);
    t.upd.reserve(sizeof(float) * 3 * n);
    t.dots.reserve(sizeof(double) * (kDotBlocks + 8));
    float *d0 = t.upd.as<float>(), *d1 = d0 + n, *df = d1 + n;
    ST_HIP(hipMemcpyAsync(d0, dobs, sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
    ST_HIP(hipMemcpyAsync(d1, dcal, sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(sub_vec, dim3(nblk256(n)), dim3(256), 0, ctx->stream, d0, d1, df,
                       (long long)n, (long long)n);
    double *part = t.dots.as<double>();
    dot(ctx, df, df, n, part, part + kDotBlocks);
    double *h = static_cast<double *>(ctx->host_pinned) + 400;
    ST_HIP(hipMemcpyAsync(h, part + kDotBlocks, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    ST_HIP(hipStreamSynchronize(ctx->stream));
    return std::sqrt(*h);
}

}  // namespace st
