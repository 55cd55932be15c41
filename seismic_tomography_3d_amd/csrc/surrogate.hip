// surrogate.hip -- learned warm start for the batched FIM.
//
// A 3D U-Net (~0.54 M parameters, 3x3x3 convolutions, BF16 MFMA GEMMs with
// FP32 accumulation, nn.hip) maps, on an nc^3 node cube spanning the model,
//   inputs  [slowness * vref, (x - xs)/L, (y - ys)/L, (z - zs)/L, D/L,
//            vref * mean slowness on the straight segment source -> node]
//   output  y = T * vref / max(D, dmin)          (D = distance to the source)
// so T = y * max(D, dmin) / vref is trilinearly available everywhere.  The
// prediction orders the FIM's block visits (a block is visited once its
// predicted arrival is within a band of its shot's earliest scheduled
// block) and, inflated by (1 + delta), initialises the travel times; the
// judge in api.hip re-solves cold any shot the prediction did not bound.
//
// Training data are solved by the plain FIM of this library; training runs
// forward, backward and Adam with the kernels of nn.hip.  No counterpart in
// the reference (its block FIM always starts cold, block_FIM.cpp:295-369).
#include <cmath>
#include <cstring>
#include <fstream>
#include <random>

#include "nn.hpp"
#include "st_common.hpp"

namespace st {

void ensure_block_model(st_ctx *ctx, int B);
void fim_batch(st_ctx *ctx, const float *dsrc, int nshot, const st_fim_opts &o,
               st_fim_stats *stats, bool warm_loaded, const float *keys, bool tentative);
void warm_fill_coarse(st_ctx *ctx, const float *dQ, int nc, float vref, float dmin,
                      const float *dsrc, int nshot, int B, float delta, float r_cold, bool values,
                      float *dkeys);

constexpr float kVref = 3000.0f;
constexpr int kIn = 6;

// ---------------------------------------------------------------------------
// network definition
// ---------------------------------------------------------------------------
struct Conv {
    int cin, cout, ksz, res_div;  // res_div: cube edge = nc / res_div
    bool relu;
    long long ow, ob;             // offsets of W [taps*cin, cout] and b [cout]
    long long op;                 // offset of the packed BF16 weights (conv3d_fwd)
};

// conv index -> (cin, cout, ksz, resolution divisor, relu)
static const int kArch[10][5] = {
    {6, 16, 3, 1, 1},    // 0  enc 32^3
    {16, 16, 3, 1, 1},   // 1  enc 32^3  -> skip a2
    {16, 32, 3, 2, 1},   // 2  enc 16^3
    {32, 32, 3, 2, 1},   // 3  enc 16^3  -> skip a4
    {32, 64, 3, 4, 1},   // 4  enc 8^3   -> skip a5
    {64, 64, 3, 8, 1},   // 5  bottleneck 4^3
    {128, 64, 3, 4, 1},  // 6  dec 8^3   (up a7 | a5)
    {96, 32, 3, 2, 1},   // 7  dec 16^3  (up a8 | a4)
    {48, 16, 3, 1, 1},   // 8  dec 32^3  (up a9 | a2)
    {16, 1, 1, 1, 0},    // 9  head 1x1
};

struct Net {
    int nc = 0;
    std::vector<Conv> L;
    long long nparam = 0;
    DevBuf w, g, m, v, w16, wp;
    long long npacked = 0;
    unsigned wver = 1, pver = 0;  // weight version / version the packing holds
    int step = 0;
    bool trained = false;
    float dmin = 0.f, rmask = 0.f;  // metres, set from the training grid
    float under_weight = 1.f;       // loss weight of under-predictions
    // activation / gradient buffers for the current batch capacities
    int cap = 0, tcap = 0;
    DevBuf in0, a1, a2, p1, a3, a4, p2, a5, p3, a7, cat3, a8, cat2, a9, cat1, a10, out;
    DevBuf d_out, d_a10, d_cat1, d_a9, d_a2, d_t32, d_cat2, d_a8, d_a4, d_t16, d_cat3, d_a7, d_a5,
        d_t8, d_p3, d_p2, d_a3, d_p1, d_a1, d_in;
    DevBuf col, dcol, d16, lossacc;
    // training set
    DevBuf X, Y, M, bX, bY, bM, Q;
    long long nsamples = 0;
};

struct SurrState {
    Net net;
    std::vector<float> history;
};

static SurrState &surr(st_ctx *ctx) {
    if (!ctx->surr) ctx->surr = new SurrState();
    return *static_cast<SurrState *>(ctx->surr);
}

void surrogate_release(st_ctx *ctx) {
    if (!ctx->surr) return;
    SurrState &s = surr(ctx);
    Net &n = s.net;
    for (DevBuf *b : {&n.w, &n.g, &n.m, &n.v, &n.w16, &n.wp, &n.in0, &n.a1, &n.a2, &n.p1, &n.a3, &n.a4,
                      &n.p2, &n.a5, &n.p3, &n.a7, &n.cat3, &n.a8, &n.cat2, &n.a9, &n.cat1, &n.a10,
                      &n.out, &n.d_out, &n.d_a10, &n.d_cat1, &n.d_a9, &n.d_a2, &n.d_t32, &n.d_cat2,
                      &n.d_a8, &n.d_a4, &n.d_t16, &n.d_cat3, &n.d_a7, &n.d_a5, &n.d_t8, &n.d_p3,
                      &n.d_p2, &n.d_a3, &n.d_p1, &n.d_a1, &n.d_in, &n.col, &n.dcol, &n.d16,
                      &n.lossacc, &n.X, &n.Y, &n.M, &n.bX, &n.bY, &n.bM, &n.Q})
        b->release();
    delete &s;
    ctx->surr = nullptr;
}

static void net_define(Net &n, int nc) {
    ST_REQUIRE(nc >= 8 && (nc & (nc - 1)) == 0, ST_ERR_ARG, "node cube edge must be a power of 2 >= 8");
    n.nc = nc;
    n.L.clear();
    long long off = 0, poff = 0;
    for (auto &a : kArch) {
        Conv c{a[0], a[1], a[2], a[3], a[4] != 0, 0, 0, 0};
        const int taps = c.ksz * c.ksz * c.ksz;
        c.ow = off;
        off += (long long)taps * c.cin * c.cout;
        c.ob = off;
        off += c.cout;
        c.op = poff;
        poff += (nn::conv_packed_size(c.ksz, c.cin, c.cout) + 127) / 128 * 128;
        n.L.push_back(c);
    }
    n.nparam = off;
    n.npacked = poff;
}

static void net_alloc_params(st_ctx *ctx, Net &n, unsigned seed) {
    const size_t b = sizeof(float) * n.nparam;
    n.w.reserve(b); n.g.reserve(b); n.m.reserve(b); n.v.reserve(b);
    n.w16.reserve(sizeof(unsigned short) * n.nparam);
    n.wp.reserve(sizeof(unsigned short) * n.npacked);
    n.wver++;
    std::vector<float> h(n.nparam, 0.f);
    std::mt19937_64 rng(seed);
    for (auto &c : n.L) {
        const int fan = c.ksz * c.ksz * c.ksz * c.cin;
        std::normal_distribution<float> nd(0.f, std::sqrt(2.0f / fan) * (c.relu ? 1.f : 0.1f));
        for (long long i = 0; i < (long long)fan * c.cout; i++) h[c.ow + i] = nd(rng);
    }
    // the head's bias starts at a typical y (T*vref/D ~ 1)
    h[n.L.back().ob] = 1.0f;
    ST_HIP(hipMemcpyAsync(n.w.ptr, h.data(), b, hipMemcpyHostToDevice, ctx->stream));
    ST_HIP(hipMemsetAsync(n.m.ptr, 0, b, ctx->stream));
    ST_HIP(hipMemsetAsync(n.v.ptr, 0, b, ctx->stream));
    nn::bf16_copy(ctx->stream, n.w.as<float>(), n.w16.as<unsigned short>(), n.nparam);
    n.wver++;
    ST_HIP(hipStreamSynchronize(ctx->stream));
    n.step = 0;
}

__host__ __device__ inline long long cube(int e) { return (long long)e * e * e; }

// forward buffers for ns samples; with train also the gradient and im2col
// buffers of the backward pass
static void net_alloc_batch(Net &n, int ns, bool train) {
    const int e1 = n.nc, e2 = n.nc / 2, e4 = n.nc / 4, e8 = n.nc / 8;
    const long long v1 = cube(e1) * ns, v2 = cube(e2) * ns, v4 = cube(e4) * ns, v8 = cube(e8) * ns;
    auto F = [](DevBuf &b, long long n) { b.reserve(sizeof(float) * std::max(n, 1LL)); };
    if (ns > n.cap) {
        F(n.in0, v1 * kIn); F(n.a1, v1 * 16); F(n.a2, v1 * 16); F(n.p1, v2 * 16); F(n.a3, v2 * 32);
        F(n.a4, v2 * 32); F(n.p2, v4 * 32); F(n.a5, v4 * 64); F(n.p3, v8 * 64); F(n.a7, v8 * 64);
        F(n.cat3, v4 * 128); F(n.a8, v4 * 64); F(n.cat2, v2 * 96); F(n.a9, v2 * 32);
        F(n.cat1, v1 * 48); F(n.a10, v1 * 16); F(n.out, v1);
        n.cap = ns;
    }
    if (!train || ns <= n.tcap) return;
    F(n.d_out, v1); F(n.d_a10, v1 * 16); F(n.d_cat1, v1 * 48); F(n.d_a9, v2 * 32);
    F(n.d_a2, v1 * 16); F(n.d_t32, v1 * 16); F(n.d_cat2, v2 * 96); F(n.d_a8, v4 * 64);
    F(n.d_a4, v2 * 32); F(n.d_t16, v2 * 32); F(n.d_cat3, v4 * 128); F(n.d_a7, v8 * 64);
    F(n.d_a5, v4 * 64); F(n.d_t8, v4 * 64); F(n.d_p3, v8 * 64); F(n.d_p2, v4 * 32);
    F(n.d_a3, v2 * 32); F(n.d_p1, v2 * 16); F(n.d_a1, v1 * 16); F(n.d_in, v1 * kIn);
    long long colmax = 0, ymax = 0;
    for (auto &c : n.L) {
        const long long V = cube(n.nc / c.res_div) * ns;
        colmax = std::max(colmax, V * c.ksz * c.ksz * c.ksz * c.cin);
        ymax = std::max(ymax, V * c.cout);
    }
    n.col.reserve(sizeof(unsigned short) * colmax);
    n.dcol.reserve(sizeof(float) * colmax);
    n.d16.reserve(sizeof(unsigned short) * ymax);
    n.lossacc.reserve(64);
    n.tcap = ns;
}

// one convolution forward: Y = act(conv(X) + b), implicit GEMM on MFMA
static void conv_fwd(st_ctx *ctx, Net &n, int li, const float *X, float *Y, int ns) {
    const Conv &c = n.L[li];
    nn::conv_forward(ctx->stream, X, n.wp.as<unsigned short>() + c.op, n.w.as<float>() + c.ob, Y,
                     ns, n.nc / c.res_div, c.cin, c.cout, c.ksz, c.relu);
}

// one convolution of the inference pass on [X0 (nearest-upsampled from half
// resolution when up0) | X1], activations in FP32 or BF16
static void conv_any(st_ctx *ctx, Net &n, int li, const void *X0, int C0, bool up0, const void *X1,
                     int C1, bool in_bf16, void *Y, bool out_bf16, int ns) {
    const Conv &c = n.L[li];
    ST_REQUIRE(C0 + C1 == c.cin, ST_ERR_ARG, "input channels do not match the layer");
    nn::conv_forward_any(ctx->stream, X0, C0, up0, X1, C1, in_bf16, n.wp.as<unsigned short>() + c.op,
                         n.w.as<float>() + c.ob, Y, out_bf16, ns, n.nc / c.res_div, c.cout, c.ksz,
                         c.relu);
}

// re-packs the BF16 weights for conv3d_fwd after any change of n.w16
static void net_pack(st_ctx *ctx, Net &n) {
    if (n.pver == n.wver) return;
    for (auto &c : n.L)
        nn::pack_conv_weights(ctx->stream, n.w16.as<unsigned short>() + c.ow,
                              n.wp.as<unsigned short>() + c.op, c.ksz, c.cin, c.cout);
    n.pver = n.wver;
}

// one convolution backward: dY (in place masked) -> dW, db (+=), dX (written)
static void conv_bwd(st_ctx *ctx, Net &n, int li, const float *X, const float *Y, float *dY,
                     float *dX, int ns) {
    const Conv &c = n.L[li];
    const int e = n.nc / c.res_div;
    const int taps = c.ksz * c.ksz * c.ksz;
    const long long rows = cube(e) * ns;
    unsigned short *d16 = n.d16.as<unsigned short>();
    nn::relu_backward(ctx->stream, dY, Y, d16, rows * c.cout, c.relu);
    nn::bias_gradient(ctx->stream, dY, n.g.as<float>() + c.ob, rows, c.cout);
    nn::im2col(ctx->stream, X, n.col.as<unsigned short>(), ns, e, c.cin, c.ksz);
    // dW[taps*cin, cout] += col^T dY
    nn::GemmArgs gw{};
    gw.A = n.col.as<unsigned short>();
    gw.ta = 1; gw.lda = (long long)taps * c.cin;
    gw.B = d16; gw.tb = 0; gw.ldb = c.cout;
    gw.C = n.g.as<float>() + c.ow; gw.ldc = c.cout;
    gw.M = (long long)taps * c.cin; gw.N = c.cout; gw.K = (int)rows;
    gw.mode = 1;
    gw.ksplit = (int)std::min<long long>(512, std::max<long long>(1, rows / 2048));
    nn::gemm(ctx->stream, gw);
    if (dX) {
        // dcol[rows, taps*cin] = dY W^T
        nn::GemmArgs gd{};
        gd.A = d16; gd.ta = 0; gd.lda = c.cout;
        gd.B = n.w16.as<unsigned short>() + c.ow; gd.tb = 1; gd.ldb = c.cout;
        gd.C = n.dcol.as<float>(); gd.ldc = (long long)taps * c.cin;
        gd.M = rows; gd.N = taps * c.cin; gd.K = c.cout;
        gd.ksplit = 1;
        nn::gemm(ctx->stream, gd);
        nn::col2im(ctx->stream, n.dcol.as<float>(), dX, ns, e, c.cin, c.ksz);
    }
}

__global__ void add_into(float *__restrict__ y, const float *__restrict__ x, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        y[i] += x[i];
}

static void add(st_ctx *ctx, float *y, const float *x, long long n) {
    const long long b = std::min<long long>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(add_into, dim3((unsigned)b), dim3(256), 0, ctx->stream, y, x, n);
}

// Forward pass.  With fused, the decoder convolutions read the upsampled
// deeper activation and the skip activation directly (no concat buffers);
// the backward pass needs the materialised concats, so training and
// gradient evaluation run unfused.
// Inference pass: BF16 activations between layers (the convolutions stage
// their inputs as BF16 MFMA operands either way; only the 2x2x2 means are
// taken of BF16 rather than FP32 values), FP32 network input and output,
// and the decoder's upsampling and skip concatenation done while staging.
static void net_infer(st_ctx *ctx, Net &n, int ns) {
    hipStream_t st = ctx->stream;
    const int e1 = n.nc, e2 = e1 / 2, e4 = e1 / 4;
    net_pack(ctx, n);
    auto h = [](DevBuf &b) { return b.as<unsigned short>(); };
    conv_any(ctx, n, 0, n.in0.ptr, kIn, false, nullptr, 0, false, h(n.a1), true, ns);
    conv_any(ctx, n, 1, h(n.a1), 16, false, nullptr, 0, true, h(n.a2), true, ns);
    nn::pool_bf16(st, h(n.a2), h(n.p1), ns, e1, 16);
    conv_any(ctx, n, 2, h(n.p1), 16, false, nullptr, 0, true, h(n.a3), true, ns);
    conv_any(ctx, n, 3, h(n.a3), 32, false, nullptr, 0, true, h(n.a4), true, ns);
    nn::pool_bf16(st, h(n.a4), h(n.p2), ns, e2, 32);
    conv_any(ctx, n, 4, h(n.p2), 32, false, nullptr, 0, true, h(n.a5), true, ns);
    nn::pool_bf16(st, h(n.a5), h(n.p3), ns, e4, 64);
    conv_any(ctx, n, 5, h(n.p3), 64, false, nullptr, 0, true, h(n.a7), true, ns);
    conv_any(ctx, n, 6, h(n.a7), 64, true, h(n.a5), 64, true, h(n.a8), true, ns);
    conv_any(ctx, n, 7, h(n.a8), 64, true, h(n.a4), 32, true, h(n.a9), true, ns);
    conv_any(ctx, n, 8, h(n.a9), 32, true, h(n.a2), 16, true, h(n.a10), true, ns);
    conv_any(ctx, n, 9, h(n.a10), 16, false, nullptr, 0, true, n.out.ptr, false, ns);
    ST_HIP(hipGetLastError());
}

// inference (infer: net_infer) or the FP32 pass whose activations the
// backward pass reads
static void net_forward(st_ctx *ctx, Net &n, int ns, bool infer = false) {
    if (infer) {
        net_infer(ctx, n, ns);
        return;
    }
    hipStream_t st = ctx->stream;
    const int e1 = n.nc, e2 = e1 / 2, e4 = e1 / 4, e8 = e1 / 8;
    net_pack(ctx, n);
    conv_fwd(ctx, n, 0, n.in0.as<float>(), n.a1.as<float>(), ns);
    conv_fwd(ctx, n, 1, n.a1.as<float>(), n.a2.as<float>(), ns);
    nn::pool(st, n.a2.as<float>(), n.p1.as<float>(), ns, e1, 16);
    conv_fwd(ctx, n, 2, n.p1.as<float>(), n.a3.as<float>(), ns);
    conv_fwd(ctx, n, 3, n.a3.as<float>(), n.a4.as<float>(), ns);
    nn::pool(st, n.a4.as<float>(), n.p2.as<float>(), ns, e2, 32);
    conv_fwd(ctx, n, 4, n.p2.as<float>(), n.a5.as<float>(), ns);
    nn::pool(st, n.a5.as<float>(), n.p3.as<float>(), ns, e4, 64);
    conv_fwd(ctx, n, 5, n.p3.as<float>(), n.a7.as<float>(), ns);
    nn::upsample(st, n.a7.as<float>(), n.cat3.as<float>(), ns, e8, 64, 128, 0);
    nn::chancopy(st, n.a5.as<float>(), n.cat3.as<float>(), cube(e4) * ns, 64, 128, 64, false);
    conv_fwd(ctx, n, 6, n.cat3.as<float>(), n.a8.as<float>(), ns);
    nn::upsample(st, n.a8.as<float>(), n.cat2.as<float>(), ns, e4, 64, 96, 0);
    nn::chancopy(st, n.a4.as<float>(), n.cat2.as<float>(), cube(e2) * ns, 32, 96, 64, false);
    conv_fwd(ctx, n, 7, n.cat2.as<float>(), n.a9.as<float>(), ns);
    nn::upsample(st, n.a9.as<float>(), n.cat1.as<float>(), ns, e2, 32, 48, 0);
    nn::chancopy(st, n.a2.as<float>(), n.cat1.as<float>(), cube(e1) * ns, 16, 48, 32, false);
    conv_fwd(ctx, n, 8, n.cat1.as<float>(), n.a10.as<float>(), ns);
    conv_fwd(ctx, n, 9, n.a10.as<float>(), n.out.as<float>(), ns);
    ST_HIP(hipGetLastError());
}

// gradients of the weighted MSE into n.g (zeroed here); returns nothing,
// the loss sums stay in n.lossacc
static void net_backward(st_ctx *ctx, Net &n, int ns, const float *target, const float *mask,
                         bool want_dx) {
    hipStream_t st = ctx->stream;
    const int e1 = n.nc, e2 = e1 / 2, e4 = e1 / 4, e8 = e1 / 8;
    const long long V1 = cube(e1) * ns, V2 = cube(e2) * ns, V4 = cube(e4) * ns;
    ST_HIP(hipMemsetAsync(n.g.ptr, 0, sizeof(float) * n.nparam, st));
    ST_HIP(hipMemsetAsync(n.lossacc.ptr, 0, 2 * sizeof(double), st));
    nn::mse_loss(st, n.out.as<float>(), target, mask, n.d_out.as<float>(), n.lossacc.as<double>(), V1,
                 n.under_weight);
    conv_bwd(ctx, n, 9, n.a10.as<float>(), n.out.as<float>(), n.d_out.as<float>(), n.d_a10.as<float>(), ns);
    conv_bwd(ctx, n, 8, n.cat1.as<float>(), n.a10.as<float>(), n.d_a10.as<float>(), n.d_cat1.as<float>(), ns);
    nn::upsample_bwd(st, n.d_cat1.as<float>(), n.d_a9.as<float>(), ns, e2, 32, 48, 0);
    nn::chancopy(st, n.d_cat1.as<float>(), n.d_a2.as<float>(), V1, 16, 48, 32, true);
    conv_bwd(ctx, n, 7, n.cat2.as<float>(), n.a9.as<float>(), n.d_a9.as<float>(), n.d_cat2.as<float>(), ns);
    nn::upsample_bwd(st, n.d_cat2.as<float>(), n.d_a8.as<float>(), ns, e4, 64, 96, 0);
    nn::chancopy(st, n.d_cat2.as<float>(), n.d_a4.as<float>(), V2, 32, 96, 64, true);
    conv_bwd(ctx, n, 6, n.cat3.as<float>(), n.a8.as<float>(), n.d_a8.as<float>(), n.d_cat3.as<float>(), ns);
    nn::upsample_bwd(st, n.d_cat3.as<float>(), n.d_a7.as<float>(), ns, e8, 64, 128, 0);
    nn::chancopy(st, n.d_cat3.as<float>(), n.d_a5.as<float>(), V4, 64, 128, 64, true);
    conv_bwd(ctx, n, 5, n.p3.as<float>(), n.a7.as<float>(), n.d_a7.as<float>(), n.d_p3.as<float>(), ns);
    nn::pool_bwd(st, n.d_p3.as<float>(), n.d_t8.as<float>(), ns, e4, 64);
    add(ctx, n.d_a5.as<float>(), n.d_t8.as<float>(), V4 * 64);
    conv_bwd(ctx, n, 4, n.p2.as<float>(), n.a5.as<float>(), n.d_a5.as<float>(), n.d_p2.as<float>(), ns);
    nn::pool_bwd(st, n.d_p2.as<float>(), n.d_t16.as<float>(), ns, e2, 32);
    add(ctx, n.d_a4.as<float>(), n.d_t16.as<float>(), V2 * 32);
    conv_bwd(ctx, n, 3, n.a3.as<float>(), n.a4.as<float>(), n.d_a4.as<float>(), n.d_a3.as<float>(), ns);
    conv_bwd(ctx, n, 2, n.p1.as<float>(), n.a3.as<float>(), n.d_a3.as<float>(), n.d_p1.as<float>(), ns);
    nn::pool_bwd(st, n.d_p1.as<float>(), n.d_t32.as<float>(), ns, e1, 16);
    add(ctx, n.d_a2.as<float>(), n.d_t32.as<float>(), V1 * 16);
    conv_bwd(ctx, n, 1, n.a1.as<float>(), n.a2.as<float>(), n.d_a2.as<float>(), n.d_a1.as<float>(), ns);
    conv_bwd(ctx, n, 0, n.in0.as<float>(), n.a1.as<float>(), n.d_a1.as<float>(),
             want_dx ? n.d_in.as<float>() : nullptr, ns);
    ST_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// inputs / targets on the node cube
// ---------------------------------------------------------------------------
__device__ __forceinline__ float tri_fine(const float *V, int nx, int ny, int nz, float fx, float fy,
                                          float fz) {
    const int i0 = min((int)fx, nx - 2), j0 = min((int)fy, ny - 2), k0 = min((int)fz, nz - 2);
    const float u = fx - i0, v = fy - j0, w = fz - k0;
    // the two z neighbours of a corner are adjacent (z fastest, k0 + 1 < nz):
    // one 8-byte load each (dword alignment suffices for global loads)
    struct __attribute__((aligned(4))) Pair { float lo, hi; };
    auto at2 = [&](int i, int j) {
        return *reinterpret_cast<const Pair *>(V + k0 + (long long)i * nz + (long long)j * nx * nz);
    };
    const Pair p00 = at2(i0, j0), p10 = at2(i0 + 1, j0), p01 = at2(i0, j0 + 1),
               p11 = at2(i0 + 1, j0 + 1);
    const float c00 = p00.lo * (1 - u) + p10.lo * u;
    const float c10 = p01.lo * (1 - u) + p11.lo * u;
    const float c01 = p00.hi * (1 - u) + p10.hi * u;
    const float c11 = p01.hi * (1 - u) + p11.hi * u;
    return ((c00 * (1 - v) + c10 * v) * (1 - w) + (c01 * (1 - v) + c11 * v) * w);
}

struct CubeGeo {
    int nc, nx, ny, nz;
    float dx, dy, dz, L;
};

__device__ __forceinline__ void node_pos(const CubeGeo &g, int v, float &px, float &py, float &pz,
                                         float &fx, float &fy, float &fz) {
    const int i = v % g.nc, j = (v / g.nc) % g.nc, k = v / (g.nc * g.nc);
    fx = i * (float)(g.nx - 1) / (g.nc - 1);
    fy = j * (float)(g.ny - 1) / (g.nc - 1);
    fz = k * (float)(g.nz - 1) / (g.nc - 1);
    px = fx * g.dx; py = fy * g.dy; pz = fz * g.dz;
}

// X[s, v, 0..4] from the fine slowness (z fastest) and the sources.
// Threads walk the nodes z fastest (the node index v is x fastest), so the
// lanes of a wave sample neighbouring z columns of the fine volume, which is
// z fastest in memory: their gathers share cache lines.
__global__ void prep_inputs(const float *__restrict__ slow, const float *__restrict__ src, int ns,
                            CubeGeo g, float *__restrict__ X) {
    const long long V = cube(g.nc);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < ns * V;
         i += (long long)gridDim.x * blockDim.x) {
        const int s = (int)(i / V), w = (int)(i - s * V);
        const int wz = w % g.nc, wx = (w / g.nc) % g.nc, wy = w / (g.nc * g.nc);
        const int v = wx + g.nc * (wy + g.nc * wz);
        float px, py, pz, fx, fy, fz;
        node_pos(g, v, px, py, pz, fx, fy, fz);
        const float ex = px - src[3 * s], ey = py - src[3 * s + 1], ez = pz - src[3 * s + 2];
        const float D = sqrtf(ex * ex + ey * ey + ez * ez);
        float *o = X + ((long long)s * V + v) * kIn;
        const float sn = tri_fine(slow, g.nx, g.ny, g.nz, fx, fy, fz);
        // straight-ray mean slowness (first arrivals never exceed it)
        const float gx = src[3 * s] / g.dx, gy = src[3 * s + 1] / g.dy, gz = src[3 * s + 2] / g.dz;
        float acc = 0.f;
        constexpr int K = 16;
        for (int k = 0; k < K; k++) {
            const float t = (k + 0.5f) / K;
            acc += tri_fine(slow, g.nx, g.ny, g.nz, gx + t * (fx - gx), gy + t * (fy - gy),
                            gz + t * (fz - gz));
        }
        o[0] = sn * kVref;
        o[1] = ex / g.L;
        o[2] = ey / g.L;
        o[3] = ez / g.L;
        o[4] = D / g.L;
        o[5] = acc / K * kVref;
    }
}

// targets y = T * vref / max(D, dmin) from the block-layout FIM volume, and
// the loss mask (nodes farther than rmask from the source)
__global__ void prep_targets(const float *__restrict__ T, long long vol, int B, int nbx, int nby,
                             const float *__restrict__ src, int ns, CubeGeo g, float dmin,
                             float rmask, float *__restrict__ Y, float *__restrict__ M) {
    const long long V = cube(g.nc);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < ns * V;
         i += (long long)gridDim.x * blockDim.x) {
        const int s = (int)(i / V), v = (int)(i - s * V);
        float px, py, pz, fx, fy, fz;
        node_pos(g, v, px, py, pz, fx, fy, fz);
        const float ex = px - src[3 * s], ey = py - src[3 * s + 1], ez = pz - src[3 * s + 2];
        const float D = sqrtf(ex * ex + ey * ey + ez * ez);
        const float *Ts = T + s * vol;
        const int i0 = min((int)fx, g.nx - 2), j0 = min((int)fy, g.ny - 2), k0 = min((int)fz, g.nz - 2);
        const float u = fx - i0, w1 = fy - j0, w2 = fz - k0;
        auto at = [&](int a, int b, int c) { return Ts[blk_addr(a, b, c, B, nbx, nby)]; };
        const float c00 = at(i0, j0, k0) * (1 - u) + at(i0 + 1, j0, k0) * u;
        const float c10 = at(i0, j0 + 1, k0) * (1 - u) + at(i0 + 1, j0 + 1, k0) * u;
        const float c01 = at(i0, j0, k0 + 1) * (1 - u) + at(i0 + 1, j0, k0 + 1) * u;
        const float c11 = at(i0, j0 + 1, k0 + 1) * (1 - u) + at(i0 + 1, j0 + 1, k0 + 1) * u;
        const float t = (c00 * (1 - w1) + c10 * w1) * (1 - w2) + (c01 * (1 - w1) + c11 * w1) * w2;
        Y[i] = t * kVref / fmaxf(D, dmin);
        M[i] = D >= rmask ? 1.f : 0.f;
    }
}

// synthetic training model: layered (kind 0) or linear gradient (kind 1)
// velocity times the Gaussian anomaly factors; written as slowness
struct ModelParams {
    int kind, nl, na;
    float tops[8], vels[8];
    float v0, v1;
    float ax[4], ay[4], az[4], asig[4], aamp[4];
};

__global__ void synth_model(ModelParams P, float *__restrict__ slow, int nx, int ny, int nz,
                            float dx, float dy, float dz) {
    const long long n = (long long)nx * ny * nz;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const int z = (int)(i % nz);
        const long long r = i / nz;
        const int x = (int)(r % nx), y = (int)(r / nx);
        const float pz = z * dz, px = x * dx, py = y * dy;
        float v;
        if (P.kind == 0) {
            v = P.vels[0];
            for (int l = 0; l < P.nl - 1; l++)
                if (pz >= P.tops[l]) v = P.vels[l + 1];
        } else {
            v = P.v0 + (P.v1 - P.v0) * pz / ((nz - 1) * dz);
        }
        for (int a = 0; a < P.na; a++) {
            const float ex = px - P.ax[a], ey = py - P.ay[a], ez = pz - P.az[a];
            v *= 1.f + P.aamp[a] * expf(-(ex * ex + ey * ey + ez * ez) / (2.f * P.asig[a] * P.asig[a]));
        }
        slow[i] = 1.f / v;
    }
}

__global__ void gather_samples(const float *__restrict__ X, const float *__restrict__ Y,
                               const float *__restrict__ M, const int *__restrict__ idx, int nb,
                               long long V, float *__restrict__ bX, float *__restrict__ bY,
                               float *__restrict__ bM) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nb * V;
         i += (long long)gridDim.x * blockDim.x) {
        const int b = (int)(i / V);
        const long long v = i - b * V;
        const long long src = (long long)idx[b] * V + v;
        bY[i] = Y[src];
        bM[i] = M[src];
        for (int c = 0; c < kIn; c++) bX[i * kIn + c] = X[src * kIn + c];
    }
}

static unsigned nblocks(long long n) { return (unsigned)std::min<long long>((n + 255) / 256, 65536); }

static CubeGeo cube_geo(st_ctx *ctx, int nc) {
    CubeGeo g;
    g.nc = nc; g.nx = ctx->nx; g.ny = ctx->ny; g.nz = ctx->nz;
    g.dx = ctx->dx; g.dy = ctx->dy; g.dz = ctx->dz;
    g.L = std::max({(ctx->nx - 1) * ctx->dx, (ctx->ny - 1) * ctx->dy, (ctx->nz - 1) * ctx->dz});
    return g;
}

static float coarse_spacing(st_ctx *ctx, int nc) {
    return std::max({(ctx->nx - 1) * ctx->dx, (ctx->ny - 1) * ctx->dy, (ctx->nz - 1) * ctx->dz}) /
           (nc - 1);
}

// ---------------------------------------------------------------------------
// entry points used by the solver
// ---------------------------------------------------------------------------
// Cells closer than five node spacings to the source start cold: near the
// source the coarse prediction, trilinearly upsampled, misses the curvature
// of the arrival cone by more than the (1 + delta) margin (at 201x201x101,
// node spacing 161 m, the cells furthest below the solution lay 320-400 m
// from the source; with the network trained towards an upper bound the
// judge rejects 2 of 256 shots at delta = 0.02 with five spacings, against
// 105 of 256 with two spacings for one trained with a lighter weight).
float surrogate_cold_radius(st_ctx *ctx) {
    const int nc = ctx->surr ? surr(ctx).net.nc : 32;
    return 5.0f * coarse_spacing(ctx, nc > 0 ? nc : 32);
}

bool surrogate_warm_fill(st_ctx *ctx, const float *dsrc, int nshot, int B, float delta, float r_cold,
                         bool values, float *dkeys) {
    if (!ctx->surr) return false;
    Net &n = surr(ctx).net;
    if (!n.trained) return false;
    const long long V = cube(n.nc);
    const CubeGeo g = cube_geo(ctx, n.nc);
    n.Q.reserve(sizeof(float) * V * nshot);
    const int chunk = 256;
    net_alloc_batch(n, std::min(chunk, nshot), false);
    for (int s0 = 0; s0 < nshot; s0 += chunk) {
        const int ns = std::min(chunk, nshot - s0);
        hipLaunchKernelGGL(prep_inputs, dim3(nblocks(V * ns)), dim3(256), 0, ctx->stream,
                           ctx->slow_zxy.as<float>(), dsrc + 3LL * s0, ns, g, n.in0.as<float>());
        net_forward(ctx, n, ns, true);
        ST_HIP(hipMemcpyAsync(n.Q.as<float>() + s0 * V, n.out.ptr, sizeof(float) * V * ns,
                              hipMemcpyDeviceToDevice, ctx->stream));
    }
    warm_fill_coarse(ctx, n.Q.as<float>(), n.nc, kVref, n.dmin, dsrc, nshot, B, delta, r_cold,
                     values, dkeys);
    return true;
}

}  // namespace st

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
using namespace st;

namespace {
template <class F> int run(st_ctx *ctx, F &&f) {
    try {
        f();
        return ST_OK;
    } catch (const st::Error &e) {
        if (ctx) ctx->last_error = e.msg;
        return e.code;
    } catch (const std::exception &e) {
        if (ctx) ctx->last_error = e.what();
        return ST_ERR_HIP;
    }
}
}  // namespace

extern "C" {

int st_surrogate_init(st_ctx *ctx, int nc, unsigned seed) {
    if (!ctx) return ST_ERR_ARG;
    return run(ctx, [&] {
        ST_HIP(hipSetDevice(ctx->device));
        Net &n = surr(ctx).net;
        net_define(n, nc);
        net_alloc_params(ctx, n, seed);
        n.trained = false;
        n.under_weight = 1.f;  // a fresh network trains on the plain MSE
        n.cap = 0;
        n.tcap = 0;
    });
}

int st_surrogate_nparams(st_ctx *ctx, long long *n) {
    if (!ctx || !n) return ST_ERR_ARG;
    return run(ctx, [&] {
        ST_REQUIRE(ctx->surr && surr(ctx).net.nc, ST_ERR_NO_SURROGATE, "st_surrogate_init first");
        *n = surr(ctx).net.nparam;
    });
}

int st_surrogate_params(st_ctx *ctx, float *get, const float *set) {
    if (!ctx) return ST_ERR_ARG;
    return run(ctx, [&] {
        ST_REQUIRE(ctx->surr && surr(ctx).net.nc, ST_ERR_NO_SURROGATE, "st_surrogate_init first");
        ST_HIP(hipSetDevice(ctx->device));
        Net &n = surr(ctx).net;
        const size_t b = sizeof(float) * n.nparam;
        if (set) {
            ST_HIP(hipMemcpyAsync(n.w.ptr, set, b, hipMemcpyHostToDevice, ctx->stream));
            nn::bf16_copy(ctx->stream, n.w.as<float>(), n.w16.as<unsigned short>(), n.nparam);
            n.wver++;
        }
        if (get) ST_HIP(hipMemcpyAsync(get, n.w.ptr, b, hipMemcpyDeviceToHost, ctx->stream));
        ST_HIP(hipStreamSynchronize(ctx->stream));
    });
}

// forward on caller inputs X[ns, nc^3, 5] -> Y[ns, nc^3]; with target and
// mask also the weighted-MSE loss and its parameter gradient
int st_surrogate_eval(st_ctx *ctx, const float *X, int ns, const float *target, const float *mask,
                      float *Y, double *loss, float *grad) {
    if (!ctx) return ST_ERR_ARG;
    return run(ctx, [&] {
        ST_REQUIRE(ctx->surr && surr(ctx).net.nc, ST_ERR_NO_SURROGATE, "st_surrogate_init first");
        ST_REQUIRE(X && ns > 0, ST_ERR_ARG, "bad inputs");
        ST_HIP(hipSetDevice(ctx->device));
        Net &n = surr(ctx).net;
        net_alloc_batch(n, ns, target && mask);
        const long long V = cube(n.nc) * ns;
        ST_HIP(hipMemcpyAsync(n.in0.ptr, X, sizeof(float) * V * kIn, hipMemcpyHostToDevice, ctx->stream));
        net_forward(ctx, n, ns, !(target && mask));
        if (Y) ST_HIP(hipMemcpyAsync(Y, n.out.ptr, sizeof(float) * V, hipMemcpyDeviceToHost, ctx->stream));
        if (target && mask) {
            n.bY.reserve(sizeof(float) * V);
            n.bM.reserve(sizeof(float) * V);
            ST_HIP(hipMemcpyAsync(n.bY.ptr, target, sizeof(float) * V, hipMemcpyHostToDevice, ctx->stream));
            ST_HIP(hipMemcpyAsync(n.bM.ptr, mask, sizeof(float) * V, hipMemcpyHostToDevice, ctx->stream));
            net_backward(ctx, n, ns, n.bY.as<float>(), n.bM.as<float>(), false);
            double h[2];
            ST_HIP(hipMemcpyAsync(h, n.lossacc.ptr, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
            ST_HIP(hipStreamSynchronize(ctx->stream));
            if (loss) *loss = h[1] > 0 ? h[0] / h[1] : 0.0;
            if (grad) ST_HIP(hipMemcpyAsync(grad, n.g.ptr, sizeof(float) * n.nparam, hipMemcpyDeviceToHost, ctx->stream));
        }
        ST_HIP(hipStreamSynchronize(ctx->stream));
    });
}

// Training data on the current model grid: n_models seeded synthetic
// models (layered, 3-7 layers in [1500, 4500] m/s, or linear gradient, each
// with 1-4 Gaussian anomalies of up to +-10%), spm sources each (x, y
// uniform, depth uniform in [0, zmax_src]), solved by the plain FIM.  The
// caller's model is restored afterwards.
int st_surrogate_make_data(st_ctx *ctx, int n_models, int spm, float zmax_src, unsigned seed,
                           float grad_fraction) {
    if (!ctx) return ST_ERR_ARG;
    return run(ctx, [&] {
        ST_REQUIRE(ctx->has_model, ST_ERR_NO_MODEL, "st_set_model first (sets the grid)");
        ST_REQUIRE(ctx->surr && surr(ctx).net.nc, ST_ERR_NO_SURROGATE, "st_surrogate_init first");
        ST_REQUIRE(n_models > 0 && spm > 0 && spm <= 256, ST_ERR_ARG, "bad counts");
        ST_HIP(hipSetDevice(ctx->device));
        Net &n = surr(ctx).net;
        const long long npts = (long long)ctx->nx * ctx->ny * ctx->nz;
        const long long V = cube(n.nc);
        const long long total = (long long)n_models * spm;
        n.X.reserve(sizeof(float) * total * V * kIn);
        n.Y.reserve(sizeof(float) * total * V);
        n.M.reserve(sizeof(float) * total * V);
        n.nsamples = total;
        n.dmin = 0.5f * coarse_spacing(ctx, n.nc);
        n.rmask = 2.0f * coarse_spacing(ctx, n.nc);
        // keep the caller's model
        DevBuf saved;
        saved.reserve(sizeof(float) * npts);
        ST_HIP(hipMemcpyAsync(saved.ptr, ctx->slow_zxy.ptr, sizeof(float) * npts,
                              hipMemcpyDeviceToDevice, ctx->stream));
        const float smin_saved = ctx->smin, smax_saved = ctx->smax;
        const bool range_saved = ctx->range_valid;
        st_fim_opts o;
        st_fim_default_opts(&o);
        o.block = 8;
        o.max_inner = 16;
        o.eps = 0.f;
        o.band = 1e30f;
        const BlockGeo bg = make_geo(ctx->nx, ctx->ny, ctx->nz, o.block);
        ctx->T.reserve(sizeof(float) * bg.vol * spm);
        ctx->flags.reserve(sizeof(unsigned) * bg.nblk * spm);
        ctx->list[0].reserve(fim_list_bytes(bg.nblk, spm));
        ctx->list[1].reserve(fim_list_bytes(bg.nblk, spm));
        ctx->src.reserve(sizeof(float) * 3 * spm);
        const CubeGeo cg = cube_geo(ctx, n.nc);
        std::mt19937_64 rng(seed);
        std::uniform_real_distribution<float> U(0.f, 1.f);
        const float Lx = (ctx->nx - 1) * ctx->dx, Ly = (ctx->ny - 1) * ctx->dy, Lz = (ctx->nz - 1) * ctx->dz;
        std::vector<float> hs(3 * spm);
        for (int mdl = 0; mdl < n_models; mdl++) {
            ModelParams P{};
            P.kind = U(rng) < grad_fraction ? 1 : 0;
            P.nl = 3 + (int)(U(rng) * 5);
            if (P.nl > 7) P.nl = 7;
            std::vector<float> vv(P.nl), tt(P.nl - 1);
            for (auto &x : vv) x = 1500.f + 3000.f * U(rng);
            for (auto &x : tt) x = (0.1f + 0.8f * U(rng)) * Lz;
            std::sort(vv.begin(), vv.end());
            std::sort(tt.begin(), tt.end());
            for (int l = 0; l < P.nl; l++) P.vels[l] = vv[l];
            for (int l = 0; l < P.nl - 1; l++) P.tops[l] = tt[l];
            P.v0 = 1500.f + 1000.f * U(rng);
            P.v1 = P.v0 + 500.f + 2000.f * U(rng);
            P.na = 1 + (int)(U(rng) * 4);
            if (P.na > 4) P.na = 4;
            for (int a = 0; a < P.na; a++) {
                P.ax[a] = (0.2f + 0.6f * U(rng)) * Lx;
                P.ay[a] = (0.2f + 0.6f * U(rng)) * Ly;
                P.az[a] = (0.2f + 0.6f * U(rng)) * Lz;
                P.asig[a] = (0.05f + 0.1f * U(rng)) * std::max(Lx, Ly);
                P.aamp[a] = (U(rng) < 0.5f ? -1.f : 1.f) * (0.04f + 0.06f * U(rng));
            }
            hipLaunchKernelGGL(synth_model, dim3(nblocks(npts)), dim3(256), 0, ctx->stream, P,
                               ctx->slow_zxy.as<float>(), ctx->nx, ctx->ny, ctx->nz, ctx->dx,
                               ctx->dy, ctx->dz);
            model_changed(ctx);
            ctx->smax = 1.f / 1400.f;
            ensure_block_model(ctx, o.block);
            for (int k = 0; k < spm; k++) {
                hs[3 * k] = U(rng) * Lx;
                hs[3 * k + 1] = U(rng) * Ly;
                hs[3 * k + 2] = U(rng) * std::min(zmax_src, Lz);
            }
            ST_HIP(hipMemcpyAsync(ctx->src.ptr, hs.data(), sizeof(float) * 3 * spm,
                                  hipMemcpyHostToDevice, ctx->stream));
            fim_batch(ctx, ctx->src.as<float>(), spm, o, nullptr, false, nullptr, false);
            const long long s0 = (long long)mdl * spm;
            hipLaunchKernelGGL(prep_inputs, dim3(nblocks(V * spm)), dim3(256), 0, ctx->stream,
                               ctx->slow_zxy.as<float>(), ctx->src.as<float>(), spm, cg,
                               n.X.as<float>() + s0 * V * kIn);
            hipLaunchKernelGGL(prep_targets, dim3(nblocks(V * spm)), dim3(256), 0, ctx->stream,
                               ctx->T.as<float>(), bg.vol, o.block, bg.nbx, bg.nby,
                               ctx->src.as<float>(), spm, cg, n.dmin, n.rmask,
                               n.Y.as<float>() + s0 * V, n.M.as<float>() + s0 * V);
            ST_HIP(hipGetLastError());
        }
        ST_HIP(hipMemcpyAsync(ctx->slow_zxy.ptr, saved.ptr, sizeof(float) * npts,
                              hipMemcpyDeviceToDevice, ctx->stream));
        ST_HIP(hipStreamSynchronize(ctx->stream));
        saved.release();
        model_changed(ctx);
        ctx->smin = smin_saved;
        ctx->smax = smax_saved;
        ctx->range_valid = range_saved;
    });
}

// Adam on the generated data; the last n_val samples are held out.
// history (nullable, epochs*2): per epoch mean training loss and the
// validation loss.
int st_surrogate_loss(st_ctx *ctx, float under_weight) {
    if (!ctx) return ST_ERR_ARG;
    return run(ctx, [&] {
        ST_REQUIRE(under_weight >= 1.f && under_weight <= 1e4f, ST_ERR_ARG,
                   "under-prediction weight must lie in [1, 1e4]");
        surr(ctx).net.under_weight = under_weight;
    });
}

int st_surrogate_train(st_ctx *ctx, int epochs, int batch, float lr, int n_val, unsigned seed,
                       float *history) {
    if (!ctx) return ST_ERR_ARG;
    return run(ctx, [&] {
        ST_REQUIRE(ctx->surr && surr(ctx).net.nc, ST_ERR_NO_SURROGATE, "st_surrogate_init first");
        Net &n = surr(ctx).net;
        ST_REQUIRE(n.nsamples > n_val && n_val >= 0 && batch > 0, ST_ERR_STATE,
                   "st_surrogate_make_data first");
        ST_HIP(hipSetDevice(ctx->device));
        const long long V = cube(n.nc);
        const long long ntrain = n.nsamples - n_val;
        net_alloc_batch(n, batch, true);
        n.bX.reserve(sizeof(float) * V * kIn * batch);
        n.bY.reserve(sizeof(float) * V * batch);
        n.bM.reserve(sizeof(float) * V * batch);
        DevBuf didx;
        didx.reserve(sizeof(int) * batch);
        std::mt19937_64 rng(seed);
        std::vector<int> perm(ntrain);
        for (long long i = 0; i < ntrain; i++) perm[i] = (int)i;
        auto run_batch = [&](const int *hidx, int nb, bool train) -> double {
            ST_HIP(hipMemcpyAsync(didx.ptr, hidx, sizeof(int) * nb, hipMemcpyHostToDevice, ctx->stream));
            hipLaunchKernelGGL(gather_samples, dim3(nblocks(V * nb)), dim3(256), 0, ctx->stream,
                               n.X.as<float>(), n.Y.as<float>(), n.M.as<float>(), didx.as<int>(), nb,
                               V, n.in0.as<float>(), n.bY.as<float>(), n.bM.as<float>());
            net_forward(ctx, n, nb);
            net_backward(ctx, n, nb, n.bY.as<float>(), n.bM.as<float>(), false);
            if (train) {
                n.step++;
                nn::adam_step(ctx->stream, n.w.as<float>(), n.g.as<float>(), n.m.as<float>(),
                              n.v.as<float>(), n.w16.as<unsigned short>(), n.nparam, lr, n.step);
                n.wver++;
            }
            double h[2];
            ST_HIP(hipMemcpyAsync(h, n.lossacc.ptr, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
            ST_HIP(hipStreamSynchronize(ctx->stream));
            return h[1] > 0 ? h[0] / h[1] : 0.0;
        };
        const float lr0 = lr;
        for (int ep = 0; ep < epochs; ep++) {
            // cosine decay from lr0 to lr0/20 over the epochs
            lr = lr0 * (0.05f + 0.95f * 0.5f * (1.f + std::cos(3.14159265f * ep / std::max(1, epochs))));
            std::shuffle(perm.begin(), perm.end(), rng);
            double acc = 0.0;
            int nbat = 0;
            for (long long b0 = 0; b0 + batch <= ntrain; b0 += batch, nbat++)
                acc += run_batch(perm.data() + b0, batch, true);
            double vl = 0.0;
            int nv = 0;
            std::vector<int> vidx(batch);
            for (long long b0 = 0; b0 < n_val; b0 += batch, nv++) {
                const int nb = (int)std::min<long long>(batch, n_val - b0);
                for (int k = 0; k < nb; k++) vidx[k] = (int)(ntrain + b0 + k);
                vl += run_batch(vidx.data(), nb, false);
            }
            if (history) {
                history[2 * ep] = (float)(nbat ? acc / nbat : 0.0);
                history[2 * ep + 1] = (float)(nv ? vl / nv : 0.0);
            }
        }
        didx.release();
        n.trained = true;
    });
}

int st_surrogate_save(st_ctx *ctx, const char *path) {
    if (!ctx || !path) return ST_ERR_ARG;
    return run(ctx, [&] {
        ST_REQUIRE(ctx->surr && surr(ctx).net.trained, ST_ERR_NO_SURROGATE, "no trained network");
        Net &n = surr(ctx).net;
        std::vector<float> h(n.nparam);
        ST_HIP(hipMemcpy(h.data(), n.w.ptr, sizeof(float) * n.nparam, hipMemcpyDeviceToHost));
        std::ofstream f(path, std::ios::binary);
        ST_REQUIRE(f.good(), ST_ERR_ARG, "cannot open weight file");
        const char magic[8] = {'S', 'T', 'W', 'A', 'R', 'M', '0', '1'};
        f.write(magic, 8);
        const int hdr[2] = {n.nc, (int)n.nparam};
        f.write(reinterpret_cast<const char *>(hdr), sizeof(hdr));
        const float geo[2] = {n.dmin, n.rmask};
        f.write(reinterpret_cast<const char *>(geo), sizeof(geo));
        f.write(reinterpret_cast<const char *>(h.data()), sizeof(float) * n.nparam);
    });
}

int st_surrogate_load(st_ctx *ctx, const char *path) {
    if (!ctx || !path) return ST_ERR_ARG;
    return run(ctx, [&] {
        ST_HIP(hipSetDevice(ctx->device));
        std::ifstream f(path, std::ios::binary);
        ST_REQUIRE(f.good(), ST_ERR_ARG, "cannot open weight file");
        char magic[8];
        int hdr[2];
        float geo[2];
        f.read(magic, 8);
        f.read(reinterpret_cast<char *>(hdr), sizeof(hdr));
        f.read(reinterpret_cast<char *>(geo), sizeof(geo));
        ST_REQUIRE(f.good() && std::memcmp(magic, "STWARM01", 8) == 0, ST_ERR_ARG, "not a weight file");
        Net &n = surr(ctx).net;
        net_define(n, hdr[0]);
        ST_REQUIRE(n.nparam == hdr[1], ST_ERR_ARG, "weight file does not match the network");
        std::vector<float> h(n.nparam);
        f.read(reinterpret_cast<char *>(h.data()), sizeof(float) * n.nparam);
        ST_REQUIRE(f.good(), ST_ERR_ARG, "truncated weight file");
        net_alloc_params(ctx, n, 0);
        ST_HIP(hipMemcpy(n.w.ptr, h.data(), sizeof(float) * n.nparam, hipMemcpyHostToDevice));
        nn::bf16_copy(ctx->stream, n.w.as<float>(), n.w16.as<unsigned short>(), n.nparam);
        n.wver++;
        ST_HIP(hipStreamSynchronize(ctx->stream));
        n.dmin = geo[0];
        n.rmask = geo[1];
        n.trained = true;
        n.cap = 0;
        n.tcap = 0;
    });
}

// Network prediction of the travel time at the model's nodes for the
// given sources: T_pred[nshot, nc^3] on the node cube (diagnostics).
int st_surrogate_predict(st_ctx *ctx, const float *src_xyz, int nshot, float *y_out) {
    if (!ctx) return ST_ERR_ARG;
    return run(ctx, [&] {
        ST_REQUIRE(ctx->has_model, ST_ERR_NO_MODEL, "st_set_model first");
        ST_REQUIRE(ctx->surr && surr(ctx).net.nc, ST_ERR_NO_SURROGATE, "no network");
        ST_REQUIRE(src_xyz && y_out && nshot > 0 && nshot <= 1024, ST_ERR_ARG, "bad arguments");
        ST_HIP(hipSetDevice(ctx->device));
        Net &n = surr(ctx).net;
        const long long V = cube(n.nc);
        net_alloc_batch(n, nshot, false);
        ctx->src.reserve(sizeof(float) * 3 * nshot);
        ST_HIP(hipMemcpyAsync(ctx->src.ptr, src_xyz, sizeof(float) * 3 * nshot, hipMemcpyHostToDevice,
                              ctx->stream));
        hipLaunchKernelGGL(prep_inputs, dim3(nblocks(V * nshot)), dim3(256), 0, ctx->stream,
                           ctx->slow_zxy.as<float>(), ctx->src.as<float>(), nshot,
                           cube_geo(ctx, n.nc), n.in0.as<float>());
        net_forward(ctx, n, nshot, true);
        ST_HIP(hipMemcpyAsync(y_out, n.out.ptr, sizeof(float) * V * nshot, hipMemcpyDeviceToHost,
                              ctx->stream));
        ST_HIP(hipStreamSynchronize(ctx->stream));
    });
}

// Diagnostic: C = A B with the network's BF16 MFMA GEMM (A [M,K], B [K,N]
// row-major FP32 on the host, rounded to BF16 on the device).
int st_nn_gemm(st_ctx *ctx, const float *A, const float *B, float *C, int M, int N, int K) {
    if (!ctx) return ST_ERR_ARG;
    return run(ctx, [&] {
        ST_REQUIRE(A && B && C && M > 0 && N > 0 && K > 0, ST_ERR_ARG, "bad arguments");
        ST_HIP(hipSetDevice(ctx->device));
        DevBuf a32, b32, a16, b16, c;
        a32.reserve(sizeof(float) * M * K); b32.reserve(sizeof(float) * K * N);
        a16.reserve(2LL * M * K); b16.reserve(2LL * K * N); c.reserve(sizeof(float) * M * N);
        ST_HIP(hipMemcpyAsync(a32.ptr, A, sizeof(float) * M * K, hipMemcpyHostToDevice, ctx->stream));
        ST_HIP(hipMemcpyAsync(b32.ptr, B, sizeof(float) * K * N, hipMemcpyHostToDevice, ctx->stream));
        nn::bf16_copy(ctx->stream, a32.as<float>(), a16.as<unsigned short>(), (long long)M * K);
        nn::bf16_copy(ctx->stream, b32.as<float>(), b16.as<unsigned short>(), (long long)K * N);
        nn::GemmArgs g{};
        g.A = a16.as<unsigned short>(); g.lda = K;
        g.B = b16.as<unsigned short>(); g.ldb = N;
        g.C = c.as<float>(); g.ldc = N;
        g.M = M; g.N = N; g.K = K; g.ksplit = 1;
        nn::gemm(ctx->stream, g);
        ST_HIP(hipMemcpyAsync(C, c.ptr, sizeof(float) * M * N, hipMemcpyDeviceToHost, ctx->stream));
        ST_HIP(hipStreamSynchronize(ctx->stream));
        a32.release(); b32.release(); a16.release(); b16.release(); c.release();
    });
}

}  // extern "C"
