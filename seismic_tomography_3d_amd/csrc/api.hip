// api.hip -- the C ABI of include/st_api.h: context, model, eikonal solve.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>

#include "st_common.hpp"

namespace st {
void ensure_block_model(st_ctx *ctx, int B);
void fim_batch(st_ctx *ctx, const float *dsrc, int nshot, const st_fim_opts &o,
               st_fim_stats *stats, bool warm_loaded, const float *keys, bool tentative);
void fim_batch_on(st_ctx *ctx, const FimBufs &bf, const float *dsrc, int nshot,
                  const st_fim_opts &o, st_fim_stats *stats, bool warm_loaded, const float *keys,
                  bool tentative);
int fim_waves(const st_fim_opts &o);
void pack_block_model(st_ctx *ctx, int B, float *S);
void warm_fill_fine(st_ctx *ctx, const float *dpred, const float *dsrc, int nshot, int B,
                    float delta, float r_cold, bool values, float *dkeys);
void judge_batch(st_ctx *ctx, const float *dsrc, int nshot, int B, float *res);
double fsm_batch(st_ctx *ctx, const float *dsrc, const float *hsrc, int nshot, const float *S);
void fsm_pad(st_ctx *ctx, float *S);
void fsm_picks(st_ctx *ctx, int nshot, const float *drcv, int nrcv, float *dpicks);
void fsm_crop_volumes(st_ctx *ctx, float *dout, int nshot);
int fsm_lanes(int nshot);
void launch_extract(st_ctx *ctx, float *dout, int nshot, int B);
void launch_picks_blk(hipStream_t st, const float *T, long long vol, int nshot, const float *rcv,
                      int nr, float *out, int nx, int ny, int nz, float dx, float dy, float dz,
                      int B, int nbx, int nby);
void launch_picks_zxy(hipStream_t st, const float *T, long long vol, int nshot, const float *rcv,
                      int nr, float *out, int nx, int ny, int nz, float dx, float dy, float dz);
// surrogate warm start (surrogate.hip); returns false when no network
bool surrogate_warm_fill(st_ctx *ctx, const float *dsrc, int nshot, int B, float delta,
                         float r_cold, bool values, float *dkeys);
float surrogate_cold_radius(st_ctx *ctx);
void tomo_release(st_ctx *ctx);
void surrogate_release(st_ctx *ctx);
void wave_release(st_ctx *ctx);
size_t podvin_bytes_per_shot(const st_ctx *ctx);
double podvin_batch(st_ctx *ctx, const float *dsrc, const float *hsrc, int nshot, const float *S,
                    float *T, int *stamp, int *list, float *newv, int *count, int *nit_dev,
                    DevBuf &temp);
void podvin_picks(st_ctx *ctx, const float *T, int nshot, const float *drcv, int nrcv,
                  float *dpicks);
void podvin_crop_volumes(st_ctx *ctx, const float *T, float *dout, int nshot);
void acoustic_dampers(float *d1, float *d2, float *d3, float factor, int nb);
void acoustic_ricker(float delay, float fmax, float dt, int nt, float *out);
void wave_set_model(st_ctx *ctx, const float *vel, int nx, int ny, int nz, float dx, float dy,
                    float dz, int nb, float factor, float dt);
void wave_shots(st_ctx *ctx, const float *wavelet, int nt, const float *src_xyz, int nshot,
                const float *rcv_xyz, int nrec, int group, float *seis_out, float *ufut_out,
                const float *pick_params, float *picks_out, double *device_ms);
void wave_pick(st_ctx *ctx, const float *seis, int ntraces, int nt, float dt, float window,
               float amp_cut, float shift, float *picks);
struct TomoState;
TomoState &tomo(st_ctx *ctx);
bool tomo_grid(st_ctx *ctx, int *n, float *d, int *max_steps, int *max_runs);
void tomo_set_grid(st_ctx *ctx, const int *n, const float *d, int max_steps, int max_runs);
void tomo_reset(st_ctx *ctx, bool with_illum);
void tomo_trace_blk(st_ctx *ctx, const float *T, long long vol, int B, int nbx, int nby,
                    const float *dsrc, int nshot, const float *drcv, int nr, bool with_illum);
void tomo_trace_zxy(st_ctx *ctx, const float *dT, const float *dsrc, const float *drcv, int nr,
                    bool with_illum);
void tomo_trace_pad(st_ctx *ctx, const float *T, long long shot_stride, int cell_stride,
                    const float *dsrc, int nshot, const float *drcv, int nr, bool with_illum);
void tomo_get_illum(st_ctx *ctx, float *out);
void tomo_csr(st_ctx *ctx, long long *rows, long long *nnz, const long long **rp, const int **c,
              const float **v);
void tomo_load_csr(st_ctx *ctx, const long long *rowptr, const int *cols, const float *vals, int N,
                   long long nnz);
void tomo_set_rhs(st_ctx *ctx, const float *dobs, const float *dcal, int n, long long Ntot);
void tomo_cgls(st_ctx *ctx, int n_model, int order, float lambda, int iters);
void tomo_cgls_setup(st_ctx *ctx, int n_model, int order, float lambda);
void tomo_solve_update(st_ctx *ctx, int iters, int smooth, int samples, float stdv, bool graph,
                       float *x_out, float *dm_out, float *vel_out);
const float *tomo_x(st_ctx *ctx);
double tomo_residual(st_ctx *ctx, const float *dobs, const float *dcal, int n);
void tomo_model_update(st_ctx *ctx, const float *dxdm, int smooth, int samples, float stdv,
                       float *dm_out, float *vel_out);

template <class F> int guarded(st_ctx *ctx, F &&f) {
    try {
        f();
        return ST_OK;
    } catch (const Error &e) {
        if (ctx) ctx->last_error = e.msg;
        return e.code;
    } catch (const std::exception &e) {
        if (ctx) ctx->last_error = e.what();
        return ST_ERR_HIP;
    }
}

void set_device(st_ctx *ctx) { ST_HIP(hipSetDevice(ctx->device)); }

void check_points(const float *p, int n, const st_ctx *ctx, const char *what) {
    const float xmax = (ctx->nx - 1) * ctx->dx, ymax = (ctx->ny - 1) * ctx->dy,
                zmax = (ctx->nz - 1) * ctx->dz;
    for (int i = 0; i < n; i++) {
        const float x = p[3 * i], y = p[3 * i + 1], z = p[3 * i + 2];
        const bool ok = std::isfinite(x) && std::isfinite(y) && std::isfinite(z) && x >= 0 &&
                        y >= 0 && z >= 0 && x <= xmax && y <= ymax && z <= zmax;
        ST_REQUIRE(ok, ST_ERR_ARG,
                   std::string(what) + " " + std::to_string(i) + " lies outside the model");
    }
}

// Smallest and largest slowness as IEEE bit patterns: for finite positive
// floats the unsigned order is the numeric order; +0 is the smallest
// pattern, while infinities, NaNs and negative values (sign bit set) are at
// or above 0x7F800000, so two integer extremes decide validity too.
__global__ void slowness_range(const unsigned *__restrict__ s, long long n,
                               unsigned *__restrict__ out) {
    unsigned lo = 0xFFFFFFFFu, hi = 0u;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const unsigned v = s[i];
        lo = min(lo, v);
        hi = max(hi, v);
    }
    for (int o = 32; o >= 1; o >>= 1) {
        lo = min(lo, (unsigned)__shfl_xor((int)lo, o));
        hi = max(hi, (unsigned)__shfl_xor((int)hi, o));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&out[0], lo);
        atomicMax(&out[1], hi);
    }
}

__global__ void range_init(unsigned *out) {
    out[0] = 0xFFFFFFFFu;
    out[1] = 0u;
}

void model_changed(st_ctx *ctx) {
    ctx->blk_valid[0] = ctx->blk_valid[1] = false;
    ctx->fsm_valid = false;
    ctx->range_valid = false;
}

// smin/smax of the device slowness through two words of device memory d.
static void ensure_range_on(st_ctx *ctx, unsigned *d) {
    if (ctx->range_valid) return;
    const long long n = (long long)ctx->nx * ctx->ny * ctx->nz;
    hipLaunchKernelGGL(range_init, dim3(1), dim3(1), 0, ctx->stream, d);
    const unsigned grid = (unsigned)std::min<long long>((n + 255) / 256, 2048);
    hipLaunchKernelGGL(slowness_range, dim3(grid), dim3(256), 0, ctx->stream,
                       ctx->slow_zxy.as<unsigned>(), n, d);
    ST_HIP(hipGetLastError());
    unsigned *h = static_cast<unsigned *>(ctx->host_pinned) + 240;
    ST_HIP(hipMemcpyAsync(h, d, 2 * sizeof(unsigned), hipMemcpyDeviceToHost, ctx->stream));
    ST_HIP(hipStreamSynchronize(ctx->stream));
    ST_REQUIRE(h[0] > 0u && h[1] < 0x7F800000u, ST_ERR_ARG,
               "slowness must be finite and positive everywhere (after a model update too)");
    ctx->smin = __builtin_bit_cast(float, h[0]);
    ctx->smax = __builtin_bit_cast(float, h[1]);
    ctx->range_valid = true;
}

void ensure_range(st_ctx *ctx) {
    if (ctx->range_valid) return;
    ctx->range.reserve(64);
    ensure_range_on(ctx, ctx->range.as<unsigned>());
}

// ---------------------------------------------------------------------------
// No-allocation FIM solve.  Every device buffer of st_fim_solve_noalloc --
// block-layout slowness, travel times, schedule flags, active lists,
// ordering keys, per-wave statistics, sources, receivers, picks and the
// slowness-range words -- is carved out of one file-scope device array.
// Each GPU the library runs on holds its own copy (kNoAllocBytes) from module
// load to process end.  Its address on a device is taken once, at context
// creation with that device current; host code passes that address to
// kernels and copies, never the array's name (a host-side stand-in).
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 256;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices];

// This device's scratch address (the device must be current).
static void *noalloc_scratch(int device) {
    ST_REQUIRE(device >= 0 && device < kMaxDevices, ST_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        ST_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        ST_REQUIRE(p != nullptr, ST_ERR_HIP, "no device address for the scratch array");
        g_scratch_ptr[device] = p;
    }
    return g_scratch_ptr[device];
}

// Bump allocator over the scratch bytes (256-byte aligned pieces).
struct ScratchCarve {
    unsigned char *base;
    size_t used = 0;
    template <class T> T *take(size_t count) {
        const size_t off = (used + 255) & ~size_t(255);
        used = off + sizeof(T) * count;
        return reinterpret_cast<T *>(base + off);
    }
};

}  // namespace st

using namespace st;

extern "C" {

void st_fim_default_opts(st_fim_opts *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->block = 8;
    o->batch = 256;
    o->max_inner = 0;
    o->eps = -1.0f;
    o->warm = 0;
    o->delta = 0.02f;
    o->judge = 1;
    o->grid = 0;
    o->rounds_per_sync = 0;
    o->use_graph = 0;
    o->band = 0.0f;
    o->cold_radius = 0.0f;
    o->warm_values = 0;
}

int st_ctx_create(int device, st_ctx **out) {
    if (!out) return ST_ERR_ARG;
    *out = nullptr;
    st_ctx *ctx = new st_ctx();
    ctx->device = device;
    int rc = guarded(ctx, [&] {
        int n = 0;
        ST_HIP(hipGetDeviceCount(&n));
        ST_REQUIRE(device >= 0 && device < n, ST_ERR_ARG, "no such HIP device");
        set_device(ctx);
        ST_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        ST_HIP(hipHostMalloc(&ctx->host_pinned, 16384, hipHostMallocDefault));
        for (auto &e : ctx->ev) ST_HIP(hipEventCreate(&e));
        ctx->counters.reserve(8192);
        ST_HIP(hipMemset(ctx->counters.ptr, 0, 8192));
        ctx->noalloc = noalloc_scratch(device);
    });
    if (rc != ST_OK) {
        st_ctx_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return ST_OK;
}

int st_ctx_destroy(st_ctx *ctx) {
    if (!ctx) return ST_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    tomo_release(ctx);
    surrogate_release(ctx);
    wave_release(ctx);
    if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
    if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
    for (auto *b : {&ctx->slow_zxy, &ctx->range, &ctx->slow_blk[0], &ctx->slow_blk[1], &ctx->T, &ctx->flags,
                    &ctx->list[0], &ctx->list[1], &ctx->counters, &ctx->wstats, &ctx->fsm_S,
                    &ctx->fsm_sid, &ctx->pod_T, &ctx->pod_stamp, &ctx->pod_list,
                    &ctx->pod_newv, &ctx->pod_counts, &ctx->pod_nit, &ctx->pod_temp,
                    &ctx->src, &ctx->rcv,
                    &ctx->picks, &ctx->scratch, &ctx->keys, &ctx->kmin, &ctx->judge, &ctx->pred,
                    &ctx->tkeys, &ctx->xfaces})
        b->release();
    for (auto &e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->host_pinned) (void)hipHostFree(ctx->host_pinned);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return ST_OK;
}

const char *st_last_error(const st_ctx *ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

int st_set_model(st_ctx *ctx, const float *slowness, int nx, int ny, int nz, float dx, float dy,
                 float dz) {
    if (!ctx) return ST_ERR_ARG;
    return guarded(ctx, [&] {
        ST_REQUIRE(slowness && nx >= 2 && ny >= 2 && nz >= 2, ST_ERR_ARG, "bad model extents");
        ST_REQUIRE(dx > 0 && dy > 0 && dz > 0, ST_ERR_ARG, "bad spacing");
        const long long n = (long long)nx * ny * nz;
        ST_REQUIRE(n < (1LL << 31), ST_ERR_ARG, "model too large");
        for (long long i = 0; i < n; i++)
            ST_REQUIRE(std::isfinite(slowness[i]) && slowness[i] > 0.f, ST_ERR_ARG,
                       "slowness must be finite and positive");
        set_device(ctx);
        ctx->nx = nx; ctx->ny = ny; ctx->nz = nz;
        ctx->dx = dx; ctx->dy = dy; ctx->dz = dz;
        ctx->slow_zxy.reserve(sizeof(float) * n);
        ST_HIP(hipMemcpyAsync(ctx->slow_zxy.ptr, slowness, sizeof(float) * n,
                              hipMemcpyHostToDevice, ctx->stream));
        model_changed(ctx);
        ctx->has_model = true;
        ctx->smin = *std::min_element(slowness, slowness + n);
        ctx->smax = *std::max_element(slowness, slowness + n);
        ctx->range_valid = true;
        ST_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int st_get_model(st_ctx *ctx, float *slowness) {
    if (!ctx) return ST_ERR_ARG;
    return guarded(ctx, [&] {
        ST_REQUIRE(ctx->has_model, ST_ERR_NO_MODEL, "no model");
        ST_REQUIRE(slowness, ST_ERR_ARG, "null output");
        set_device(ctx);
        const long long n = (long long)ctx->nx * ctx->ny * ctx->nz;
        ST_HIP(hipMemcpyAsync(slowness, ctx->slow_zxy.ptr, sizeof(float) * n,
                              hipMemcpyDeviceToHost, ctx->stream));
        ST_HIP(hipStreamSynchronize(ctx->stream));
    });
}

// Sources and receivers to the device (src: nshot triples, rcv: nrcv).
static void upload_points(st_ctx *ctx, const float *src_xyz, int nshot, const float *rcv_xyz,
                          int nrcv) {
    ctx->src.reserve(sizeof(float) * 3 * std::max(nshot, 1));
    ctx->rcv.reserve(sizeof(float) * 3 * std::max(nrcv, 1));
    ctx->picks.reserve(sizeof(float) * std::max<long long>(1, (long long)nshot * nrcv));
    ST_HIP(hipMemcpyAsync(ctx->src.ptr, src_xyz, sizeof(float) * 3 * nshot, hipMemcpyHostToDevice,
                          ctx->stream));
    if (nrcv)
        ST_HIP(hipMemcpyAsync(ctx->rcv.ptr, rcv_xyz, sizeof(float) * 3 * nrcv,
                              hipMemcpyHostToDevice, ctx->stream));
}

// The sweeping solver's padded slowness (rebuilt after a model change) and
// travel-time buffer for batches of shots; returns the batch size (batch <= 0:
// 256, halved until a batch fits in half of the free device memory).
static int fsm_setup(st_ctx *ctx, int nshot, int batch, bool with_T_out) {
    const long long nB = (long long)(ctx->nx + 2) * (ctx->ny + 2) * (ctx->nz + 2);
    const long long npts = (long long)ctx->nx * ctx->ny * ctx->nz;
    if (!ctx->fsm_valid) {
        ctx->fsm_S.reserve(sizeof(float) * nB);
        fsm_pad(ctx, ctx->fsm_S.as<float>());
        ctx->fsm_valid = true;
    }
    int b = batch > 0 ? std::min({batch, nshot, 1023}) : std::min(nshot, 256);
    size_t free_b = 0, total_b = 0;
    ST_HIP(hipMemGetInfo(&free_b, &total_b));
    const size_t per_shot = sizeof(float) * nB + (with_T_out ? sizeof(float) * npts : 0);
    const size_t have = free_b + ctx->T.bytes + ctx->scratch.bytes;
    while (b > 1 && per_shot * (size_t)fsm_lanes(b) > have / 2) b /= 2;
    ctx->T.reserve(sizeof(float) * nB * fsm_lanes(b));
    if (with_T_out) ctx->scratch.reserve(sizeof(float) * npts * b);
    return b;
}

// Podvin's padded slowness (the sweeping solver's one-cell replicated pad)
// and per-shot buffers; returns the batch size (batch <= 0: 16).
static int podvin_setup(st_ctx *ctx, int nshot, int batch, bool with_T_out) {
    const long long nB = (long long)(ctx->nx + 2) * (ctx->ny + 2) * (ctx->nz + 2);
    const long long npts = (long long)ctx->nx * ctx->ny * ctx->nz;
    if (!ctx->fsm_valid) {
        ctx->fsm_S.reserve(sizeof(float) * nB);
        fsm_pad(ctx, ctx->fsm_S.as<float>());
        ctx->fsm_valid = true;
    }
    int b = batch > 0 ? std::min(batch, nshot) : std::min(nshot, 16);
    // list items and counts are 32-bit: shots x padded cells below 2^31
    while (b > 1 && (long long)b * nB >= (1LL << 31)) b /= 2;
    size_t free_b = 0, total_b = 0;
    ST_HIP(hipMemGetInfo(&free_b, &total_b));
    const size_t per_shot = podvin_bytes_per_shot(ctx) + (with_T_out ? sizeof(float) * npts : 0);
    const size_t have = free_b + ctx->pod_T.bytes + ctx->pod_stamp.bytes + ctx->pod_list.bytes +
                        ctx->pod_newv.bytes + ctx->scratch.bytes;
    while (b > 1 && per_shot * (size_t)b > have / 2) b /= 2;
    ctx->pod_T.reserve(sizeof(float) * nB * b);
    ctx->pod_stamp.reserve(sizeof(int) * nB * b);
    ctx->pod_list.reserve(sizeof(int) * nB * b);
    ctx->pod_newv.reserve(sizeof(float) * nB * b);
    ctx->pod_counts.reserve(sizeof(int) * 64);
    ctx->pod_nit.reserve(sizeof(int) * b);
    if (with_T_out) ctx->scratch.reserve(sizeof(float) * npts * b);
    return b;
}

// Podvin over shots b0 .. b0 + nb - 1 of the uploaded sources into pod_T.
static double podvin_run(st_ctx *ctx, int b0, int nb, const float *src_xyz) {
    return podvin_batch(ctx, ctx->src.as<float>() + 3LL * b0, src_xyz + 3LL * b0, nb,
                        ctx->fsm_S.as<float>(), ctx->pod_T.as<float>(), ctx->pod_stamp.as<int>(),
                        ctx->pod_list.as<int>(), ctx->pod_newv.as<float>(),
                        ctx->pod_counts.as<int>(), ctx->pod_nit.as<int>(), ctx->pod_temp);
}

static void solve_impl(st_ctx *ctx, const float *src_xyz, int nshot, const float *rcv_xyz,
                       int nrcv, float *picks, float *T_out, const float *T_pred,
                       const st_fim_opts *opts, st_fim_stats *stats) {
    const auto t_start = std::chrono::steady_clock::now();
    ST_REQUIRE(ctx->has_model, ST_ERR_NO_MODEL, "st_set_model first");
    ST_REQUIRE(nshot >= 0 && nrcv >= 0, ST_ERR_ARG, "negative count");
    ST_REQUIRE(nshot == 0 || src_xyz, ST_ERR_ARG, "null sources");
    ST_REQUIRE(nrcv == 0 || (rcv_xyz && picks), ST_ERR_ARG, "null receivers or picks");
    st_fim_opts o;
    if (opts) o = *opts; else st_fim_default_opts(&o);
    ST_REQUIRE(o.block == 4 || o.block == 8, ST_ERR_ARG, "block must be 4 or 8");
    if (o.max_inner <= 0) o.max_inner = (o.block == 8) ? 24 : 16;
    if (o.eps < 0.f) o.eps = kFimEps;
    if (o.delta <= 0.f) o.delta = 0.02f;
    ST_REQUIRE(o.warm == 0 || o.warm == 1, ST_ERR_ARG, "warm must be 0 or 1");
    ST_REQUIRE(o.sweep == 0 || o.sweep == 1, ST_ERR_ARG, "sweep must be 0 or 1");
    ST_REQUIRE(o.order >= 0 && o.order <= 2, ST_ERR_ARG, "order must be 0, 1 or 2");
    const bool warm = (o.warm == 1) || T_pred;
    ST_REQUIRE(o.order != 2 || warm, ST_ERR_ARG, "order 2 needs a warm start's predictions");
    set_device(ctx);
    ensure_range(ctx);
    if (o.band == 0.f) o.band = o.block * ctx->dx * ctx->smax;  // one slow block crossing
    if (o.band < 0.f || (!warm && o.order == 0)) o.band = 1e30f;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (nshot == 0) return;
    check_points(src_xyz, nshot, ctx, "source");
    check_points(rcv_xyz, nrcv, ctx, "receiver");
    set_device(ctx);

    const int B = o.block;
    ensure_block_model(ctx, B);
    const BlockGeo g = make_geo(ctx->nx, ctx->ny, ctx->nz, B);
    ST_REQUIRE(g.nblk < (1LL << 22), ST_ERR_ARG, "model has too many blocks");
    const long long npts = (long long)ctx->nx * ctx->ny * ctx->nz;
    int batch = std::max(1, std::min({o.batch, nshot, 1023}));
    size_t free_b = 0, total_b = 0;
    ST_HIP(hipMemGetInfo(&free_b, &total_b));
    const size_t per_shot = sizeof(float) * g.vol + sizeof(unsigned) * g.nblk +
                            2 * fim_list_bytes(g.nblk, 1) +
                            (B == 8 ? sizeof(float) * 128 * g.nblk : 0) +
                            (warm ? sizeof(float) * g.nblk : 0) +
                            (T_pred ? sizeof(float) * npts : 0) + (T_out ? sizeof(float) * npts : 0);
    const size_t have = free_b + ctx->T.bytes + ctx->flags.bytes + ctx->list[0].bytes +
                        ctx->list[1].bytes + ctx->keys.bytes + ctx->pred.bytes + ctx->scratch.bytes +
                        ctx->xfaces.bytes;
    while (batch > 1 && per_shot * (size_t)batch > have / 2) batch /= 2;

    ctx->T.reserve(sizeof(float) * g.vol * batch);
    if (B == 8) ctx->xfaces.reserve(sizeof(float) * 128 * g.nblk * batch);
    ctx->flags.reserve(sizeof(unsigned) * g.nblk * batch);
    ctx->list[0].reserve(fim_list_bytes(g.nblk, batch));
    ctx->list[1].reserve(fim_list_bytes(g.nblk, batch));
    ctx->src.reserve(sizeof(float) * 3 * (nshot + batch));
    ctx->rcv.reserve(sizeof(float) * 3 * std::max(nrcv, 1));
    ctx->picks.reserve(sizeof(float) * std::max<long long>(1, (long long)nshot * nrcv));
    if (warm) ctx->keys.reserve(sizeof(float) * g.nblk * batch);
    if (T_pred) ctx->pred.reserve(sizeof(float) * npts * batch);
    if (T_out) ctx->scratch.reserve(sizeof(float) * npts * batch);
    ST_HIP(hipMemcpyAsync(ctx->src.ptr, src_xyz, sizeof(float) * 3 * nshot, hipMemcpyHostToDevice,
                          ctx->stream));
    if (nrcv)
        ST_HIP(hipMemcpyAsync(ctx->rcv.ptr, rcv_xyz, sizeof(float) * 3 * nrcv,
                              hipMemcpyHostToDevice, ctx->stream));
    float r_cold = o.cold_radius;
    if (r_cold <= 0.f)
        r_cold = (o.warm == 1) ? surrogate_cold_radius(ctx) : 2.0f * B * ctx->dx;

    std::vector<float> res(batch);
    for (int b0 = 0; b0 < nshot; b0 += batch) {
        const int nb = std::min(batch, nshot - b0);
        const float *dsrc = ctx->src.as<float>() + 3LL * b0;
        const float *keys = nullptr;
        if (warm) {
            if (T_pred) {
                ST_HIP(hipMemcpyAsync(ctx->pred.ptr, T_pred + (long long)b0 * npts,
                                      sizeof(float) * npts * nb, hipMemcpyHostToDevice, ctx->stream));
                warm_fill_fine(ctx, ctx->pred.as<float>(), dsrc, nb, B, o.delta, r_cold,
                               o.warm_values != 0, ctx->keys.as<float>());
            } else {
                const bool ok = surrogate_warm_fill(ctx, dsrc, nb, B, o.delta, r_cold,
                                                    o.warm_values != 0, ctx->keys.as<float>());
                ST_REQUIRE(ok, ST_ERR_NO_SURROGATE, "warm start requested without a network");
            }
            keys = ctx->keys.as<float>();
        }
        if (o.order == 1) keys = nullptr;  // the solve's own arrivals alone
        fim_batch(ctx, dsrc, nb, o, stats, warm, keys, o.order != 0);
        std::vector<int> bad;
        // the judge can only reject seeded values: an ordering-only warm start
        // keeps the cold initial state and reaches the cold fixed point
        if (warm && o.warm_values != 0 && o.judge) {
            judge_batch(ctx, dsrc, nb, B, res.data());
            for (int j = 0; j < nb; j++) {
                if (stats) stats->judge_max = std::max(stats->judge_max, (double)res[j]);
                if (res[j] > 2e-6f) bad.push_back(j);
            }
        }
        if (nrcv)
            launch_picks_blk(ctx->stream, ctx->T.as<float>(), g.vol, nb, ctx->rcv.as<float>(), nrcv,
                             ctx->picks.as<float>() + (long long)b0 * nrcv, ctx->nx, ctx->ny,
                             ctx->nz, ctx->dx, ctx->dy, ctx->dz, B, g.nbx, g.nby);
        if (T_out) {
            launch_extract(ctx, ctx->scratch.as<float>(), nb, B);
            ST_HIP(hipMemcpyAsync(T_out + (long long)b0 * npts, ctx->scratch.ptr,
                                  sizeof(float) * npts * nb, hipMemcpyDeviceToHost, ctx->stream));
            ST_HIP(hipStreamSynchronize(ctx->stream));
        }
        if (!bad.empty()) {
            // fallback: the rejected shots are solved again from a cold start
            const int k = (int)bad.size();
            float *gsrc = ctx->src.as<float>() + 3LL * nshot;
            std::vector<float> hs(3 * k);
            for (int q = 0; q < k; q++)
                for (int c = 0; c < 3; c++) hs[3 * q + c] = src_xyz[3LL * (b0 + bad[q]) + c];
            ST_HIP(hipMemcpyAsync(gsrc, hs.data(), sizeof(float) * 3 * k, hipMemcpyHostToDevice,
                                  ctx->stream));
            st_fim_opts oc = o;
            oc.warm = 0;
            oc.band = 1e30f;
            fim_batch(ctx, gsrc, k, oc, stats, false, nullptr, false);
            if (stats) stats->fallbacks += k;
            for (int q = 0; q < k; q++) {
                const long long row = b0 + bad[q];
                if (nrcv)
                    launch_picks_blk(ctx->stream, ctx->T.as<float>() + q * g.vol, g.vol, 1,
                                     ctx->rcv.as<float>(), nrcv,
                                     ctx->picks.as<float>() + row * nrcv, ctx->nx, ctx->ny,
                                     ctx->nz, ctx->dx, ctx->dy, ctx->dz, B, g.nbx, g.nby);
            }
            if (T_out) {
                launch_extract(ctx, ctx->scratch.as<float>(), k, B);
                for (int q = 0; q < k; q++)
                    ST_HIP(hipMemcpyAsync(T_out + (long long)(b0 + bad[q]) * npts,
                                          ctx->scratch.as<float>() + q * npts, sizeof(float) * npts,
                                          hipMemcpyDeviceToHost, ctx->stream));
                ST_HIP(hipStreamSynchronize(ctx->stream));
            }
        }
    }
    if (nrcv)
        ST_HIP(hipMemcpyAsync(picks, ctx->picks.ptr, sizeof(float) * (long long)nshot * nrcv,
                              hipMemcpyDeviceToHost, ctx->stream));
    ST_HIP(hipStreamSynchronize(ctx->stream));
    if (stats)
        stats->total_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start)
                .count();
}

int st_fim_solve(st_ctx *ctx, const float *src_xyz, int nshot, const float *rcv_xyz, int nrcv,
                 float *picks, float *T_out, const st_fim_opts *opts, st_fim_stats *stats) {
    if (!ctx) return ST_ERR_ARG;
    return guarded(ctx, [&] {
        solve_impl(ctx, src_xyz, nshot, rcv_xyz, nrcv, picks, T_out, nullptr, opts, stats);
    });
}

int st_fim_solve_noalloc(st_ctx *ctx, const float *src_xyz, int nshot, const float *rcv_xyz,
                         int nrcv, float *picks, const st_fim_opts *opts, st_fim_stats *stats) {
    if (!ctx) return ST_ERR_ARG;
    return guarded(ctx, [&] {
        const auto t_start = std::chrono::steady_clock::now();
        ST_REQUIRE(ctx->has_model, ST_ERR_NO_MODEL, "st_set_model first");
        ST_REQUIRE(nshot >= 0 && nrcv >= 0, ST_ERR_ARG, "negative count");
        ST_REQUIRE(nshot == 0 || src_xyz, ST_ERR_ARG, "null sources");
        ST_REQUIRE(nrcv == 0 || (rcv_xyz && picks), ST_ERR_ARG, "null receivers or picks");
        ST_REQUIRE(ctx->noalloc, ST_ERR_STATE, "context without a scratch address");
        st_fim_opts o;
        if (opts) o = *opts; else st_fim_default_opts(&o);
        ST_REQUIRE(o.block == 4 || o.block == 8, ST_ERR_ARG, "block must be 4 or 8");
        ST_REQUIRE(o.warm == 0, ST_ERR_ARG, "the no-allocation solve starts cold (warm = 0)");
        ST_REQUIRE(o.sweep == 0 || o.sweep == 1, ST_ERR_ARG, "sweep must be 0 or 1");
        ST_REQUIRE(o.order == 0 || o.order == 1, ST_ERR_ARG, "order must be 0 or 1");
        if (o.max_inner <= 0) o.max_inner = (o.block == 8) ? 24 : 16;
        if (o.eps < 0.f) o.eps = kFimEps;
        o.use_graph = 0;  // graph instantiation may allocate
        if (stats) std::memset(stats, 0, sizeof(*stats));
        if (nshot == 0) return;
        check_points(src_xyz, nshot, ctx, "source");
        check_points(rcv_xyz, nrcv, ctx, "receiver");
        set_device(ctx);

        const int B = o.block;
        const BlockGeo g = make_geo(ctx->nx, ctx->ny, ctx->nz, B);
        ST_REQUIRE(g.nblk < (1LL << 22), ST_ERR_ARG, "model has too many blocks");
        const bool tentative = o.order == 1;
        const int nwaves = fim_waves(o);
        // the layout for b shots per pass (each piece 256-byte aligned)
        auto layout = [&](int b, FimBufs *bf, float **S, float **src, float **rcv, float **pk,
                          unsigned **range) {
            ScratchCarve c{static_cast<unsigned char *>(ctx->noalloc)};
            float *s_ = c.take<float>(g.vol);
            unsigned *r_ = c.take<unsigned>(2);
            float *rc_ = c.take<float>(3 * (size_t)std::max(nrcv, 1));
            float *sr_ = c.take<float>(3 * (size_t)b);
            float *pk_ = c.take<float>((size_t)b * std::max(nrcv, 1));
            FimBufs f;
            f.S = s_;
            f.T = c.take<float>((size_t)b * g.vol);
            f.flags = c.take<unsigned>((size_t)b * g.nblk);
            f.list0 = c.take<unsigned>(fim_list_bytes(g.nblk, b) / sizeof(unsigned));
            f.list1 = c.take<unsigned>(fim_list_bytes(g.nblk, b) / sizeof(unsigned));
            f.kmin = tentative ? c.take<unsigned>(3 * (size_t)b) : nullptr;
            f.tk = tentative ? c.take<unsigned>(2 * (size_t)b * g.nblk) : nullptr;
            f.wstats = stats ? c.take<unsigned long long>(3 * (size_t)nwaves) : nullptr;
            f.xf = B == 8 ? c.take<float>(128 * (size_t)b * g.nblk) : nullptr;
            f.graphs = false;
            if (bf) {
                *bf = f; *S = s_; *src = sr_; *rcv = rc_; *pk = pk_; *range = r_;
            }
            return c.used;
        };
        int batch = std::max(1, std::min({o.batch, nshot, 1023}));
        while (batch > 1 && layout(batch, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr) >
                                kNoAllocBytes)
            batch = std::max(1, batch / 2);
        const size_t need = layout(batch, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
        ST_REQUIRE(need <= kNoAllocBytes, ST_ERR_ARG,
                   "model too large for the no-allocation solve: one shot needs " +
                       std::to_string(need >> 20) + " MiB of its " +
                       std::to_string(kNoAllocBytes >> 20) + " MiB scratch");
        FimBufs bf;
        float *S, *dsrc, *drcv, *dpk;
        unsigned *drange;
        layout(batch, &bf, &S, &dsrc, &drcv, &dpk, &drange);

        ensure_range_on(ctx, drange);
        if (o.band == 0.f) o.band = o.block * ctx->dx * ctx->smax;  // one slow block crossing
        if (o.band < 0.f || !tentative) o.band = 1e30f;
        pack_block_model(ctx, B, S);
        if (nrcv)
            ST_HIP(hipMemcpyAsync(drcv, rcv_xyz, sizeof(float) * 3 * nrcv, hipMemcpyHostToDevice,
                                  ctx->stream));
        for (int b0 = 0; b0 < nshot; b0 += batch) {
            const int nb = std::min(batch, nshot - b0);
            ST_HIP(hipMemcpyAsync(dsrc, src_xyz + 3LL * b0, sizeof(float) * 3 * nb,
                                  hipMemcpyHostToDevice, ctx->stream));
            fim_batch_on(ctx, bf, dsrc, nb, o, stats, false, nullptr, tentative);
            if (nrcv) {
                launch_picks_blk(ctx->stream, bf.T, g.vol, nb, drcv, nrcv, dpk, ctx->nx, ctx->ny,
                                 ctx->nz, ctx->dx, ctx->dy, ctx->dz, B, g.nbx, g.nby);
                ST_HIP(hipMemcpyAsync(picks + (long long)b0 * nrcv, dpk,
                                      sizeof(float) * (long long)nb * nrcv, hipMemcpyDeviceToHost,
                                      ctx->stream));
            }
            ST_HIP(hipStreamSynchronize(ctx->stream));
        }
        if (stats)
            stats->total_ms = std::chrono::duration<double, std::milli>(
                                  std::chrono::steady_clock::now() - t_start)
                                  .count();
    });
}

int st_fim_solve_warm(st_ctx *ctx, const float *src_xyz, int nshot, const float *rcv_xyz,
                      int nrcv, float *picks, float *T_out, const float *T_pred,
                      const st_fim_opts *opts, st_fim_stats *stats) {
    if (!ctx) return ST_ERR_ARG;
    return guarded(ctx, [&] {
        ST_REQUIRE(T_pred, ST_ERR_ARG, "null predictions");
        st_fim_opts o;
        if (opts) o = *opts; else st_fim_default_opts(&o);
        o.warm = 0;  // the predictions come from the caller
        solve_impl(ctx, src_xyz, nshot, rcv_xyz, nrcv, picks, T_out, T_pred, &o, stats);
    });
}

int st_fsm_solve(st_ctx *ctx, const float *src_xyz, int nshot, const float *rcv_xyz, int nrcv,
                 float *picks, float *T_out, int batch, double *device_ms) {
    if (!ctx) return ST_ERR_ARG;
    return guarded(ctx, [&] {
        ST_REQUIRE(ctx->has_model, ST_ERR_NO_MODEL, "st_set_model first");
        ST_REQUIRE(nshot >= 0 && nrcv >= 0, ST_ERR_ARG, "negative count");
        ST_REQUIRE(nshot == 0 || src_xyz, ST_ERR_ARG, "null sources");
        ST_REQUIRE(nrcv == 0 || (rcv_xyz && picks), ST_ERR_ARG, "null receivers or picks");
        if (device_ms) *device_ms = 0.0;
        if (nshot == 0) return;
        check_points(src_xyz, nshot, ctx, "source");
        check_points(rcv_xyz, nrcv, ctx, "receiver");
        set_device(ctx);
        ensure_range(ctx);
        const long long npts = (long long)ctx->nx * ctx->ny * ctx->nz;
        const int b = fsm_setup(ctx, nshot, batch, T_out != nullptr);
        upload_points(ctx, src_xyz, nshot, rcv_xyz, nrcv);
        double ms = 0.0;
        for (int b0 = 0; b0 < nshot; b0 += b) {
            const int nb = std::min(b, nshot - b0);
            ms += fsm_batch(ctx, ctx->src.as<float>() + 3LL * b0, src_xyz + 3LL * b0, nb,
                            ctx->fsm_S.as<float>());
            if (nrcv)
                fsm_picks(ctx, nb, ctx->rcv.as<float>(), nrcv,
                          ctx->picks.as<float>() + (long long)b0 * nrcv);
            if (T_out) {
                fsm_crop_volumes(ctx, ctx->scratch.as<float>(), nb);
                ST_HIP(hipMemcpyAsync(T_out + (long long)b0 * npts, ctx->scratch.ptr,
                                      sizeof(float) * npts * nb, hipMemcpyDeviceToHost,
                                      ctx->stream));
                ST_HIP(hipStreamSynchronize(ctx->stream));
            }
        }
        if (nrcv)
            ST_HIP(hipMemcpyAsync(picks, ctx->picks.ptr, sizeof(float) * (long long)nshot * nrcv,
                                  hipMemcpyDeviceToHost, ctx->stream));
        ST_HIP(hipStreamSynchronize(ctx->stream));
        if (device_ms) *device_ms = ms;
    });
}

int st_podvin_solve(st_ctx *ctx, const float *src_xyz, int nshot, const float *rcv_xyz, int nrcv,
                    float *picks, float *T_out, int batch, double *device_ms) {
    if (!ctx) return ST_ERR_ARG;
    return guarded(ctx, [&] {
        ST_REQUIRE(ctx->has_model, ST_ERR_NO_MODEL, "st_set_model first");
        ST_REQUIRE(nshot >= 0 && nrcv >= 0, ST_ERR_ARG, "negative count");
        ST_REQUIRE(nshot == 0 || src_xyz, ST_ERR_ARG, "null sources");
        ST_REQUIRE(nrcv == 0 || (rcv_xyz && picks), ST_ERR_ARG, "null receivers or picks");
        if (device_ms) *device_ms = 0.0;
        if (nshot == 0) return;
        check_points(src_xyz, nshot, ctx, "source");
        check_points(rcv_xyz, nrcv, ctx, "receiver");
        set_device(ctx);
        ensure_range(ctx);
        const long long npts = (long long)ctx->nx * ctx->ny * ctx->nz;
        const int b = podvin_setup(ctx, nshot, batch, T_out != nullptr);
        upload_points(ctx, src_xyz, nshot, rcv_xyz, nrcv);
        double ms = 0.0;
        for (int b0 = 0; b0 < nshot; b0 += b) {
            const int nb = std::min(b, nshot - b0);
            ms += podvin_run(ctx, b0, nb, src_xyz);
            if (nrcv)
                podvin_picks(ctx, ctx->pod_T.as<float>(), nb, ctx->rcv.as<float>(), nrcv,
                             ctx->picks.as<float>() + (long long)b0 * nrcv);
            if (T_out) {
                podvin_crop_volumes(ctx, ctx->pod_T.as<float>(), ctx->scratch.as<float>(), nb);
                ST_HIP(hipMemcpyAsync(T_out + (long long)b0 * npts, ctx->scratch.ptr,
                                      sizeof(float) * npts * nb, hipMemcpyDeviceToHost,
                                      ctx->stream));
                ST_HIP(hipStreamSynchronize(ctx->stream));
            }
        }
        if (nrcv)
            ST_HIP(hipMemcpyAsync(picks, ctx->picks.ptr, sizeof(float) * (long long)nshot * nrcv,
                                  hipMemcpyDeviceToHost, ctx->stream));
        ST_HIP(hipStreamSynchronize(ctx->stream));
        if (device_ms) *device_ms = ms;
    });
}

int st_picks_volume(st_ctx *ctx, const float *T, int nx, int ny, int nz, float dx, float dy,
                    float dz, const float *rcv_xyz, int nrcv, float *picks) {
    if (!ctx) return ST_ERR_ARG;
    return guarded(ctx, [&] {
        ST_REQUIRE(T && nx > 0 && ny > 0 && nz > 0 && dx > 0 && dy > 0 && dz > 0, ST_ERR_ARG,
                   "bad volume");
        ST_REQUIRE(nrcv >= 0 && (nrcv == 0 || (rcv_xyz && picks)), ST_ERR_ARG, "bad receivers");
        if (nrcv == 0) return;
        for (int i = 0; i < nrcv; i++) {
            const float x = rcv_xyz[3 * i], y = rcv_xyz[3 * i + 1], z = rcv_xyz[3 * i + 2];
            ST_REQUIRE(x >= 0 && y >= 0 && z >= 0 && x <= (nx - 1) * dx && y <= (ny - 1) * dy &&
                           z <= (nz - 1) * dz,
                       ST_ERR_ARG, "receiver outside the volume");
        }
        set_device(ctx);
        const long long n = (long long)nx * ny * nz;
        ctx->scratch.reserve(sizeof(float) * n);
        ctx->rcv.reserve(sizeof(float) * 3 * nrcv);
        ctx->picks.reserve(sizeof(float) * nrcv);
        ST_HIP(hipMemcpyAsync(ctx->scratch.ptr, T, sizeof(float) * n, hipMemcpyHostToDevice,
                              ctx->stream));
        ST_HIP(hipMemcpyAsync(ctx->rcv.ptr, rcv_xyz, sizeof(float) * 3 * nrcv,
                              hipMemcpyHostToDevice, ctx->stream));
        launch_picks_zxy(ctx->stream, ctx->scratch.as<float>(), n, 1, ctx->rcv.as<float>(), nrcv,
                         ctx->picks.as<float>(), nx, ny, nz, dx, dy, dz);
        ST_HIP(hipMemcpyAsync(picks, ctx->picks.ptr, sizeof(float) * nrcv, hipMemcpyDeviceToHost,
                              ctx->stream));
        ST_HIP(hipStreamSynchronize(ctx->stream));
    });
}


// ---- tomography --------------------------------------------------------

int st_tomo_set_grid(st_ctx *ctx, float dxt, float dyt, float dzt, int max_steps) {
    if (!ctx) return ST_ERR_ARG;
    return guarded(ctx, [&] {
        ST_REQUIRE(ctx->has_model, ST_ERR_NO_MODEL, "st_set_model first");
        ST_REQUIRE(dxt > 0 && dyt > 0 && dzt > 0, ST_ERR_ARG, "bad inversion spacing");
        int n[3];
        n[2] = (int)((ctx->nz - 1) * ctx->dz / dzt) + 1;
        n[0] = (int)((ctx->nx - 1) * ctx->dx / dxt) + 1;
        n[1] = (int)((ctx->ny - 1) * ctx->dy / dyt) + 1;
        const float d[3] = {dxt, dyt, dzt};
        const float step = 0.2f * ctx->dz;
        const float L = (ctx->nx + ctx->ny + ctx->nz) * std::max({ctx->dx, ctx->dy, ctx->dz});
        if (max_steps <= 0) max_steps = (int)std::min(1e8f, 8.0f * L / step) + 64;
        const int max_runs = 8 * (n[0] + n[1] + n[2]) + 64;
        tomo_set_grid(ctx, n, d, max_steps, max_runs);
    });
}

int st_tomo_dims(st_ctx *ctx, int *nxt, int *nyt, int *nzt) {
    if (!ctx) return ST_ERR_ARG;
    return guarded(ctx, [&] {
        int n[3], ms, mr;
        float d[3];
        ST_REQUIRE(tomo_grid(ctx, n, d, &ms, &mr), ST_ERR_STATE, "st_tomo_set_grid first");
        if (nxt) *nxt = n[0];
        if (nyt) *nyt = n[1];
        if (nzt) *nzt = n[2];
    });
}

int st_tomo_set_eikonal(st_ctx *ctx, int eikonal_type) {
    if (!ctx) return ST_ERR_ARG;
    return guarded(ctx, [&] {
        // least_squares.cpp:85-104: 0 Classic, 1 Block_FIM, 2 Accurate_FSM,
        // any other value Accurate_FSM (the switch's default branch)
        ctx->eikonal_type = (eikonal_type == 0 || eikonal_type == 1) ? eikonal_type : 2;
    });
}

static void forward_impl(st_ctx *ctx, const float *src_xyz, int nshot, const float *rcv_xyz,
                         int nrcv, float *dcal, float *illum, float *T_out,
                         const st_fim_opts *opts, st_fim_stats *stats) {
    const auto t_start = std::chrono::steady_clock::now();
    int n[3], ms, mr;
    float d[3];
    ST_REQUIRE(ctx->has_model, ST_ERR_NO_MODEL, "st_set_model first");
    ST_REQUIRE(tomo_grid(ctx, n, d, &ms, &mr), ST_ERR_STATE, "st_tomo_set_grid first");
    ST_REQUIRE(nshot > 0 && nrcv > 0 && src_xyz && rcv_xyz && dcal, ST_ERR_ARG,
               "forward needs sources, receivers and a dcal buffer");
    st_fim_opts o;
    if (opts) o = *opts; else st_fim_default_opts(&o);
    const int type = ctx->eikonal_type;
    if (type == 1) {
        ST_REQUIRE(o.block == 4 || o.block == 8, ST_ERR_ARG, "block must be 4 or 8");
        if (o.max_inner <= 0) o.max_inner = (o.block == 8) ? 24 : 16;
        if (o.eps < 0.f) o.eps = kFimEps;
        ST_REQUIRE(o.warm == 0 || o.warm == 1, ST_ERR_ARG, "warm must be 0 or 1");
        // a forward pass must reach the cold fixed point: a warm start here
        // only orders the block visits by the network's prediction
        ST_REQUIRE(o.warm == 0 || o.warm_values == 0, ST_ERR_ARG,
                   "forward warm starts order visits only (warm_values = 0)");
        ST_REQUIRE(o.sweep == 0 || o.sweep == 1, ST_ERR_ARG, "sweep must be 0 or 1");
        ST_REQUIRE(o.order >= 0 && o.order <= 2, ST_ERR_ARG, "order must be 0, 1 or 2");
        ST_REQUIRE(o.order != 2 || o.warm == 1, ST_ERR_ARG, "order 2 needs a warm start");
    } else {
        ST_REQUIRE(o.warm == 0, ST_ERR_ARG, "warm starts apply to the block FIM (eikonal_type 1)");
    }
    if (stats) std::memset(stats, 0, sizeof(*stats));
    check_points(src_xyz, nshot, ctx, "source");
    check_points(rcv_xyz, nrcv, ctx, "receiver");
    set_device(ctx);
    ensure_range(ctx);
    const long long npts = (long long)ctx->nx * ctx->ny * ctx->nz;
    upload_points(ctx, src_xyz, nshot, rcv_xyz, nrcv);
    tomo_reset(ctx, true);
    // one batch: solve, pick, trace its rays; the travel times to T_out
    auto to_host = [&](int b0, int nb) {
        ST_HIP(hipMemcpyAsync(T_out + (long long)b0 * npts, ctx->scratch.ptr,
                              sizeof(float) * npts * nb, hipMemcpyDeviceToHost, ctx->stream));
        ST_HIP(hipStreamSynchronize(ctx->stream));
    };
    const float *drcv = ctx->rcv.as<float>();
    if (type == 1) {
        const bool warm = o.warm == 1;
        if (o.band == 0.f) o.band = o.block * ctx->dx * ctx->smax;  // one slow block crossing
        if (o.band < 0.f || (!warm && o.order == 0)) o.band = 1e30f;
        const int B = o.block;
        ensure_block_model(ctx, B);
        const BlockGeo g = make_geo(ctx->nx, ctx->ny, ctx->nz, B);
        const int batch = std::max(1, std::min({o.batch, nshot, 1023}));
        ctx->T.reserve(sizeof(float) * g.vol * batch);
        if (B == 8) ctx->xfaces.reserve(sizeof(float) * 128 * g.nblk * batch);
        ctx->flags.reserve(sizeof(unsigned) * g.nblk * batch);
        ctx->list[0].reserve(fim_list_bytes(g.nblk, batch));
        ctx->list[1].reserve(fim_list_bytes(g.nblk, batch));
        if (warm) ctx->keys.reserve(sizeof(float) * g.nblk * batch);
        if (T_out) ctx->scratch.reserve(sizeof(float) * npts * batch);
        const float r_cold = o.cold_radius > 0.f ? o.cold_radius
                                                 : (warm ? surrogate_cold_radius(ctx) : 0.f);
        for (int b0 = 0; b0 < nshot; b0 += batch) {
            const int nb = std::min(batch, nshot - b0);
            const float *dsrc = ctx->src.as<float>() + 3LL * b0;
            const float *keys = nullptr;
            if (warm) {
                const bool ok = surrogate_warm_fill(ctx, dsrc, nb, B, o.delta, r_cold, false,
                                                    ctx->keys.as<float>());
                ST_REQUIRE(ok, ST_ERR_NO_SURROGATE, "warm start requested without a network");
                keys = ctx->keys.as<float>();
            }
            if (o.order == 1) keys = nullptr;  // the solve's own arrivals alone
            fim_batch(ctx, dsrc, nb, o, stats, warm, keys, o.order != 0);
            launch_picks_blk(ctx->stream, ctx->T.as<float>(), g.vol, nb, drcv, nrcv,
                             ctx->picks.as<float>() + (long long)b0 * nrcv, ctx->nx, ctx->ny,
                             ctx->nz, ctx->dx, ctx->dy, ctx->dz, B, g.nbx, g.nby);
            tomo_trace_blk(ctx, ctx->T.as<float>(), g.vol, B, g.nbx, g.nby, dsrc, nb, drcv, nrcv,
                           true);
            if (T_out) {
                launch_extract(ctx, ctx->scratch.as<float>(), nb, B);
                to_host(b0, nb);
            }
        }
    } else if (type == 2) {
        // Accurate_FSM: the padded, shot-fastest volumes are traced in place
        const int b = fsm_setup(ctx, nshot, o.batch, T_out != nullptr);
        for (int b0 = 0; b0 < nshot; b0 += b) {
            const int nb = std::min(b, nshot - b0);
            const float *dsrc = ctx->src.as<float>() + 3LL * b0;
            const double t = fsm_batch(ctx, dsrc, src_xyz + 3LL * b0, nb, ctx->fsm_S.as<float>());
            if (stats) { stats->solve_ms += t; stats->batches += 1; }
            fsm_picks(ctx, nb, drcv, nrcv, ctx->picks.as<float>() + (long long)b0 * nrcv);
            tomo_trace_pad(ctx, ctx->T.as<float>(), 1, fsm_lanes(nb), dsrc, nb, drcv, nrcv, true);
            if (T_out) {
                fsm_crop_volumes(ctx, ctx->scratch.as<float>(), nb);
                to_host(b0, nb);
            }
        }
    } else {
        // Podvin & Lecomte: shot-major padded volumes
        const int b = podvin_setup(ctx, nshot, 0, T_out != nullptr);
        const long long nB = (long long)(ctx->nx + 2) * (ctx->ny + 2) * (ctx->nz + 2);
        for (int b0 = 0; b0 < nshot; b0 += b) {
            const int nb = std::min(b, nshot - b0);
            const float *dsrc = ctx->src.as<float>() + 3LL * b0;
            const double t = podvin_run(ctx, b0, nb, src_xyz);
            if (stats) { stats->solve_ms += t; stats->batches += 1; }
            podvin_picks(ctx, ctx->pod_T.as<float>(), nb, drcv, nrcv,
                         ctx->picks.as<float>() + (long long)b0 * nrcv);
            tomo_trace_pad(ctx, ctx->pod_T.as<float>(), nB, 1, dsrc, nb, drcv, nrcv, true);
            if (T_out) {
                podvin_crop_volumes(ctx, ctx->pod_T.as<float>(), ctx->scratch.as<float>(), nb);
                to_host(b0, nb);
            }
        }
    }
    ST_HIP(hipMemcpyAsync(dcal, ctx->picks.ptr, sizeof(float) * (long long)nshot * nrcv,
                          hipMemcpyDeviceToHost, ctx->stream));
    if (illum) tomo_get_illum(ctx, illum);
    ST_HIP(hipStreamSynchronize(ctx->stream));
    if (stats)
        stats->total_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start)
                .count();
}

int st_forward(st_ctx *ctx, const float *src_xyz, int nshot, const float *rcv_xyz, int nrcv,
               float *dcal, float *illum, const st_fim_opts *opts, st_fim_stats *stats) {
    if (!ctx) return ST_ERR_ARG;
    return guarded(ctx, [&] {
        forward_impl(ctx, src_xyz, nshot, rcv_xyz, nrcv, dcal, illum, nullptr, opts, stats);
    });
}

int st_forward_times(st_ctx *ctx, const float *src_xyz, int nshot, const float *rcv_xyz,
                     int nrcv, float *dcal, float *illum, float *T_out, const st_fim_opts *opts,
                     st_fim_stats *stats) {
    if (!ctx) return ST_ERR_ARG;
    return guarded(ctx, [&] {
        ST_REQUIRE(T_out, ST_ERR_ARG, "null travel-time output");
        forward_impl(ctx, src_xyz, nshot, rcv_xyz, nrcv, dcal, illum, T_out, opts, stats);
    });
}

int st_csr_reset(st_ctx *ctx) {
    if (!ctx) return ST_ERR_ARG;
    return guarded(ctx, [&] {
        ST_REQUIRE(ctx->has_model, ST_ERR_NO_MODEL, "st_set_model first");
        set_device(ctx);
        tomo_reset(ctx, true);
    });
}

int st_raytrace_volume(st_ctx *ctx, const float *T, const float *src_xyz, const float *rcv_xyz,
                       int nrcv, float *illum) {
    if (!ctx) return ST_ERR_ARG;
    return guarded(ctx, [&] {
        int n[3], ms, mr;
        float d[3];
        ST_REQUIRE(ctx->has_model, ST_ERR_NO_MODEL, "st_set_model first");
        ST_REQUIRE(tomo_grid(ctx, n, d, &ms, &mr), ST_ERR_STATE, "st_tomo_set_grid first");
        ST_REQUIRE(T && src_xyz && rcv_xyz && nrcv > 0, ST_ERR_ARG, "bad arguments");
        check_points(src_xyz, 1, ctx, "source");
        check_points(rcv_xyz, nrcv, ctx, "receiver");
        set_device(ctx);
        const long long npts = (long long)ctx->nx * ctx->ny * ctx->nz;
        ctx->scratch.reserve(sizeof(float) * npts);
        ctx->src.reserve(sizeof(float) * 3);
        ctx->rcv.reserve(sizeof(float) * 3 * nrcv);
        ST_HIP(hipMemcpyAsync(ctx->scratch.ptr, T, sizeof(float) * npts, hipMemcpyHostToDevice,
                              ctx->stream));
        ST_HIP(hipMemcpyAsync(ctx->src.ptr, src_xyz, sizeof(float) * 3, hipMemcpyHostToDevice,
                              ctx->stream));
        ST_HIP(hipMemcpyAsync(ctx->rcv.ptr, rcv_xyz, sizeof(float) * 3 * nrcv,
                              hipMemcpyHostToDevice, ctx->stream));
        tomo_trace_zxy(ctx, ctx->scratch.as<float>(), ctx->src.as<float>(), ctx->rcv.as<float>(),
                       nrcv, true);
        if (illum) tomo_get_illum(ctx, illum);
        ST_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int st_csr_size(st_ctx *ctx, long long *nrows, long long *nnz) {
    if (!ctx) return ST_ERR_ARG;
    return guarded(ctx, [&] {
        const long long *rp;
        const int *c;
        const float *v;
        long long r = 0, z = 0;
        tomo_csr(ctx, &r, &z, &rp, &c, &v);
        if (nrows) *nrows = r;
        if (nnz) *nnz = z;
    });
}

int st_csr_copy(st_ctx *ctx, long long *rowptr, int *cols, float *vals) {
    if (!ctx) return ST_ERR_ARG;
    return guarded(ctx, [&] {
        const long long *rp;
        const int *c;
        const float *v;
        long long r = 0, z = 0;
        tomo_csr(ctx, &r, &z, &rp, &c, &v);
        set_device(ctx);
        if (rowptr && r >= 0)
            ST_HIP(hipMemcpyAsync(rowptr, rp, sizeof(long long) * (r + 1), hipMemcpyDeviceToHost,
                                  ctx->stream));
        if (cols && z)
            ST_HIP(hipMemcpyAsync(cols, c, sizeof(int) * z, hipMemcpyDeviceToHost, ctx->stream));
        if (vals && z)
            ST_HIP(hipMemcpyAsync(vals, v, sizeof(float) * z, hipMemcpyDeviceToHost, ctx->stream));
        ST_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int st_residual_norm(st_ctx *ctx, const float *dobs, const float *dcal, int n, double *norm) {
    if (!ctx) return ST_ERR_ARG;
    return guarded(ctx, [&] {
        ST_REQUIRE(dobs && dcal && norm && n > 0, ST_ERR_ARG, "bad arguments");
        set_device(ctx);
        *norm = tomo_residual(ctx, dobs, dcal, n);
    });
}

int st_optimize(st_ctx *ctx, const float *dobs, const float *dcal, int n, int tk_order,
                float lambda, int iters, float *xdm) {
    if (!ctx) return ST_ERR_ARG;
    return guarded(ctx, [&] {
        int nt[3], ms, mr;
        float d[3];
        ST_REQUIRE(tomo_grid(ctx, nt, d, &ms, &mr), ST_ERR_STATE, "st_tomo_set_grid first");
        long long rows = 0, nnz = 0;
        const long long *rp;
        const int *c;
        const float *v;
        tomo_csr(ctx, &rows, &nnz, &rp, &c, &v);
        ST_REQUIRE(rows > 0 && rows == n, ST_ERR_STATE,
                   "st_optimize needs the G of st_forward with n = shots*receivers rows");
        ST_REQUIRE(dobs && dcal && iters >= 0 && tk_order >= 0, ST_ERR_ARG, "bad arguments");
        set_device(ctx);
        const int n_model = nt[0] * nt[1] * nt[2];
        tomo_set_rhs(ctx, dobs, dcal, n, rows + n_model - tk_order);
        tomo_cgls(ctx, n_model, tk_order, lambda, iters);
        if (xdm) {
            ST_HIP(hipMemcpyAsync(xdm, tomo_x(ctx), sizeof(float) * n_model, hipMemcpyDeviceToHost,
                                  ctx->stream));
            ST_HIP(hipStreamSynchronize(ctx->stream));
        }
    });
}

int st_cgls_csr(st_ctx *ctx, const long long *rowptr, const int *cols, const float *vals, int N,
                int M, const float *b, int iters, float *x) {
    if (!ctx) return ST_ERR_ARG;
    return guarded(ctx, [&] {
        ST_REQUIRE(rowptr && cols && vals && b && x && N > 0 && M > 0 && iters >= 0, ST_ERR_ARG,
                   "bad arguments");
        const long long nnz = rowptr[N];
        ST_REQUIRE(rowptr[0] == 0 && nnz >= 0, ST_ERR_ARG, "bad rowptr");
        for (long long e = 0; e < nnz; e++)
            ST_REQUIRE(cols[e] >= 0 && cols[e] < M, ST_ERR_ARG, "column out of range");
        set_device(ctx);
        tomo_load_csr(ctx, rowptr, cols, vals, N, nnz);
        // lambda = 0 order-0 rows: zero rows that leave A^T A unchanged
        tomo_set_rhs(ctx, b, nullptr, N, (long long)N + M);
        tomo_cgls(ctx, M, 0, 0.0f, iters);
        ST_HIP(hipMemcpyAsync(x, tomo_x(ctx), sizeof(float) * M, hipMemcpyDeviceToHost,
                              ctx->stream));
        ST_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int st_optimize_update(st_ctx *ctx, const float *dobs, const float *dcal, int n, int tk_order,
                       float lambda, int iters, int smooth, int samples, float stdv, int use_graph,
                       float *xdm, float *dm, float *velocity) {
    if (!ctx) return ST_ERR_ARG;
    return guarded(ctx, [&] {
        int nt[3], ms, mr;
        float d[3];
        ST_REQUIRE(ctx->has_model, ST_ERR_NO_MODEL, "st_set_model first");
        ST_REQUIRE(tomo_grid(ctx, nt, d, &ms, &mr), ST_ERR_STATE, "st_tomo_set_grid first");
        long long rows = 0, nnz = 0;
        const long long *rp;
        const int *c;
        const float *v;
        tomo_csr(ctx, &rows, &nnz, &rp, &c, &v);
        ST_REQUIRE(rows > 0 && rows == n, ST_ERR_STATE,
                   "st_optimize_update needs the G of st_forward with n = shots*receivers rows");
        ST_REQUIRE(dobs && dcal && iters >= 0 && tk_order >= 0, ST_ERR_ARG, "bad arguments");
        set_device(ctx);
        const int n_model = nt[0] * nt[1] * nt[2];
        tomo_set_rhs(ctx, dobs, dcal, n, rows + n_model - tk_order);
        tomo_cgls_setup(ctx, n_model, tk_order, lambda);
        tomo_solve_update(ctx, iters, smooth, samples, stdv, use_graph != 0, xdm, dm, velocity);
    });
}

int st_model_update(st_ctx *ctx, const float *xdm, int smooth, int samples, float stdv, float *dm,
                    float *velocity) {
    if (!ctx) return ST_ERR_ARG;
    return guarded(ctx, [&] {
        int nt[3], ms, mr;
        float d[3];
        ST_REQUIRE(ctx->has_model, ST_ERR_NO_MODEL, "st_set_model first");
        ST_REQUIRE(tomo_grid(ctx, nt, d, &ms, &mr), ST_ERR_STATE, "st_tomo_set_grid first");
        set_device(ctx);
        const float *dx = nullptr;
        if (xdm) {
            const long long n_model = (long long)nt[0] * nt[1] * nt[2];
            ctx->scratch.reserve(sizeof(float) * n_model);
            ST_HIP(hipMemcpyAsync(ctx->scratch.ptr, xdm, sizeof(float) * n_model,
                                  hipMemcpyHostToDevice, ctx->stream));
            dx = ctx->scratch.as<float>();
        } else {
            dx = tomo_x(ctx);
            ST_REQUIRE(dx, ST_ERR_STATE, "no solution: call st_optimize or pass xdm");
        }
        tomo_model_update(ctx, dx, smooth, samples, stdv, dm, velocity);
    });
}


// ---- data synthesis ----------------------------------------------------

int st_acoustic_dampers(float *d1, float *d2, float *d3, float factor, int nb) {
    if (!d1 || !d2 || !d3 || nb <= 0) return ST_ERR_ARG;
    acoustic_dampers(d1, d2, d3, factor, nb);
    return ST_OK;
}

int st_ricker(float delay, float fmax, float dt, int nt, float *out) {
    if (!out || nt <= 0) return ST_ERR_ARG;
    acoustic_ricker(delay, fmax, dt, nt, out);
    return ST_OK;
}

int st_acoustic_set_model(st_ctx *ctx, const float *vel, int nx, int ny, int nz, float dx,
                          float dy, float dz, int nb, float factor, float dt) {
    if (!ctx) return ST_ERR_ARG;
    return guarded(ctx, [&] {
        set_device(ctx);
        wave_set_model(ctx, vel, nx, ny, nz, dx, dy, dz, nb, factor, dt);
    });
}

int st_acoustic_shots(st_ctx *ctx, const float *wavelet, int nt, const float *src_xyz, int nshot,
                      const float *rcv_xyz, int nrec, int group, float *seis, float *ufut_last,
                      double *device_ms) {
    if (!ctx) return ST_ERR_ARG;
    return guarded(ctx, [&] {
        set_device(ctx);
        wave_shots(ctx, wavelet, nt, src_xyz, nshot, rcv_xyz, nrec, group, seis, ufut_last,
                   nullptr, nullptr, device_ms);
    });
}

int st_acoustic_picks(st_ctx *ctx, const float *wavelet, int nt, const float *src_xyz, int nshot,
                      const float *rcv_xyz, int nrec, int group, float window, float amp_cut,
                      float shift, float *picks, double *device_ms) {
    if (!ctx) return ST_ERR_ARG;
    return guarded(ctx, [&] {
        ST_REQUIRE(picks, ST_ERR_ARG, "picks is null");
        set_device(ctx);
        const float pp[3] = {window, amp_cut, shift};
        wave_shots(ctx, wavelet, nt, src_xyz, nshot, rcv_xyz, nrec, group, nullptr, nullptr, pp,
                   picks, device_ms);
    });
}

int st_auto_pick(st_ctx *ctx, const float *seis, int ntraces, int nt, float dt, float window,
                 float amp_cut, float shift, float *picks) {
    if (!ctx) return ST_ERR_ARG;
    return guarded(ctx, [&] {
        set_device(ctx);
        wave_pick(ctx, seis, ntraces, nt, dt, window, amp_cut, shift, picks);
    });
}

}  // extern "C"
