// api.hip -- the C ABI of include/st_api.h: context, model, eikonal solve.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "st_common.hpp"

namespace st {
void ensure_block_model(st_ctx *ctx, int B);
void fim_batch(st_ctx *ctx, const float *dsrc, int nshot, const st_fim_opts &o,
               st_fim_stats *stats, bool warm_loaded);
void launch_extract(st_ctx *ctx, float *dout, int nshot, int B);
void launch_picks_blk(hipStream_t st, const float *T, long long vol, int nshot, const float *rcv,
                      int nr, float *out, int nx, int ny, int nz, float dx, float dy, float dz,
                      int B, int nbx, int nby);
void launch_picks_zxy(hipStream_t st, const float *T, long long vol, int nshot, const float *rcv,
                      int nr, float *out, int nx, int ny, int nz, float dx, float dy, float dz);
// surrogate warm start (surrogate.hip); returns false when no network
bool surrogate_warm_fill(st_ctx *ctx, const float *dsrc, int nshot, int B, float delta);
int surrogate_judge(st_ctx *ctx, const float *dsrc, int nshot, int B, float tol, int *bad);
void tomo_release(st_ctx *ctx);
void surrogate_release(st_ctx *ctx);

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

}  // namespace st

using namespace st;

extern "C" {

void st_fim_default_opts(st_fim_opts *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->block = 4;
    o->batch = 64;
    o->max_inner = 0;
    o->eps = -1.0f;
    o->warm = 0;
    o->delta = 0.02f;
    o->judge = 1;
    o->grid = 0;
    o->rounds_per_sync = 0;
    o->use_graph = 0;
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
        ST_HIP(hipHostMalloc(&ctx->host_pinned, 4096, hipHostMallocDefault));
        for (auto &e : ctx->ev) ST_HIP(hipEventCreate(&e));
        ctx->counters.reserve(256);
        ST_HIP(hipMemset(ctx->counters.ptr, 0, 256));
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
    if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
    if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
    for (auto *b : {&ctx->slow_zxy, &ctx->slow_blk[0], &ctx->slow_blk[1], &ctx->T, &ctx->flags,
                    &ctx->list[0], &ctx->list[1], &ctx->counters, &ctx->src, &ctx->rcv,
                    &ctx->picks, &ctx->scratch})
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
        ctx->blk_valid[0] = ctx->blk_valid[1] = false;
        ctx->has_model = true;
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

int st_fim_solve(st_ctx *ctx, const float *src_xyz, int nshot, const float *rcv_xyz, int nrcv,
                 float *picks, float *T_out, const st_fim_opts *opts, st_fim_stats *stats) {
    if (!ctx) return ST_ERR_ARG;
    return guarded(ctx, [&] {
        const auto t_start = std::chrono::steady_clock::now();
        ST_REQUIRE(ctx->has_model, ST_ERR_NO_MODEL, "st_set_model first");
        ST_REQUIRE(nshot >= 0 && nrcv >= 0, ST_ERR_ARG, "negative count");
        ST_REQUIRE(nshot == 0 || src_xyz, ST_ERR_ARG, "null sources");
        ST_REQUIRE(nrcv == 0 || (rcv_xyz && picks), ST_ERR_ARG, "null receivers or picks");
        st_fim_opts o;
        if (opts) o = *opts; else st_fim_default_opts(&o);
        ST_REQUIRE(o.block == 4 || o.block == 8, ST_ERR_ARG, "block must be 4 or 8");
        if (o.max_inner <= 0) o.max_inner = (o.block == 8) ? 8 : 16;
        if (o.eps < 0.f) o.eps = 0.f;
        if (o.delta <= 0.f) o.delta = 0.02f;
        ST_REQUIRE(o.warm == 0 || o.warm == 1, ST_ERR_ARG, "warm must be 0 or 1");
        if (stats) std::memset(stats, 0, sizeof(*stats));
        if (nshot == 0) return;
        check_points(src_xyz, nshot, ctx, "source");
        check_points(rcv_xyz, nrcv, ctx, "receiver");
        set_device(ctx);

        const int B = o.block;
        ensure_block_model(ctx, B);
        const BlockGeo g = make_geo(ctx->nx, ctx->ny, ctx->nz, B);
        ST_REQUIRE(g.nblk < (1LL << 22), ST_ERR_ARG, "model has too many blocks");
        int batch = std::max(1, std::min({o.batch, nshot, 1023}));
        // keep the batch's travel times within a quarter of free HBM
        size_t free_b = 0, total_b = 0;
        ST_HIP(hipMemGetInfo(&free_b, &total_b));
        const size_t per_shot = sizeof(float) * g.vol + 3 * sizeof(unsigned) * g.nblk;
        const size_t have = free_b + ctx->T.bytes + ctx->flags.bytes + ctx->list[0].bytes +
                            ctx->list[1].bytes;
        while (batch > 1 && per_shot * (size_t)batch > have / 2) batch /= 2;

        ctx->T.reserve(sizeof(float) * g.vol * batch);
        ctx->flags.reserve(sizeof(unsigned) * g.nblk * batch);
        ctx->list[0].reserve(sizeof(unsigned) * g.nblk * batch);
        ctx->list[1].reserve(sizeof(unsigned) * g.nblk * batch);
        ctx->src.reserve(sizeof(float) * 3 * nshot);
        ctx->rcv.reserve(sizeof(float) * 3 * std::max(nrcv, 1));
        ctx->picks.reserve(sizeof(float) * std::max<long long>(1, (long long)nshot * nrcv));
        ST_HIP(hipMemcpyAsync(ctx->src.ptr, src_xyz, sizeof(float) * 3 * nshot,
                              hipMemcpyHostToDevice, ctx->stream));
        if (nrcv)
            ST_HIP(hipMemcpyAsync(ctx->rcv.ptr, rcv_xyz, sizeof(float) * 3 * nrcv,
                                  hipMemcpyHostToDevice, ctx->stream));
        const long long npts = (long long)ctx->nx * ctx->ny * ctx->nz;
        if (T_out) ctx->scratch.reserve(sizeof(float) * npts * batch);

        for (int b0 = 0; b0 < nshot; b0 += batch) {
            const int nb = std::min(batch, nshot - b0);
            const float *dsrc = ctx->src.as<float>() + 3LL * b0;
            bool warm = false;
            if (o.warm == 1) {
                warm = surrogate_warm_fill(ctx, dsrc, nb, B, o.delta);
                ST_REQUIRE(warm, ST_ERR_NO_SURROGATE, "warm start requested without a network");
            }
            fim_batch(ctx, dsrc, nb, o, stats, warm);
            if (warm && o.judge) {
                // judge: a warm start that was not an upper bound leaves cells
                // below what their neighbours support; such batches are
                // re-solved from a cold start (the fallback)
                int bad[1024];
                const int nbad = surrogate_judge(ctx, dsrc, nb, B, 1e-5f, bad);
                if (nbad > 0) {
                    if (stats) stats->fallbacks += nbad;
                    fim_batch(ctx, dsrc, nb, o, stats, false);
                }
            }
            if (nrcv)
                launch_picks_blk(ctx->stream, ctx->T.as<float>(), g.vol, nb, ctx->rcv.as<float>(),
                                 nrcv, ctx->picks.as<float>() + (long long)b0 * nrcv, ctx->nx,
                                 ctx->ny, ctx->nz, ctx->dx, ctx->dy, ctx->dz, B, g.nbx, g.nby);
            if (T_out) {
                launch_extract(ctx, ctx->scratch.as<float>(), nb, B);
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
        if (stats)
            stats->total_ms = std::chrono::duration<double, std::milli>(
                                  std::chrono::steady_clock::now() - t_start)
                                  .count();
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

}  // extern "C"
