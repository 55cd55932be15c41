// st_common.hpp -- shared definitions of the HIP implementation behind
// include/st_api.h.  gfx950 (CDNA4) only: wave64, one wave per block visit.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "st_api.h"

namespace st {

constexpr float kInf = 1e6f;  // "not reached" (cuda_kernel.cuh:7)
// default FIM change threshold in seconds: a cell counts as changed (and
// its block as unconverged) when it decreased by more than this -- the
// reference's block-FIM tolerance EPS (cuda_kernel.cuh:8, used at
// cuda_kernel.cu:414)
constexpr float kFimEps = 1e-6f;

// Thrown inside the library, turned into a status at the C boundary.
struct Error {
    int code;
    std::string msg;
};

#define ST_HIP(call)                                                          \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess)                                                 \
            throw ::st::Error{ST_ERR_HIP, std::string(#call) + ": " +         \
                                              hipGetErrorString(e_)};         \
    } while (0)

#define ST_REQUIRE(cond, code, text)                                          \
    do {                                                                      \
        if (!(cond)) throw ::st::Error{(code), (text)};                       \
    } while (0)

// Growable device buffer owned by the context (allocated once per size,
// reused across shots and iterations).
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    void reserve(size_t n) {
        if (n <= bytes) return;
        if (ptr) ST_HIP(hipFree(ptr));
        ptr = nullptr;
        ST_HIP(hipMalloc(&ptr, n));
        bytes = n;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    template <class T> T *as() const { return static_cast<T *>(ptr); }
};

// Active lists of the block FIM are split into kListSegments append
// segments, each with its own counter on its own 128-byte line, so that the
// per-visit append atomics of a round spread over as many addresses.  A
// segment can receive every block of the batch, so a list holds
// kListSegments * nblk * nshot items.
constexpr int kListSegments = 16;
inline size_t fim_list_bytes(long long nblk, long long nshot) {
    return sizeof(unsigned) * (size_t)kListSegments * (size_t)nblk * (size_t)std::max(nshot, 1LL);
}

// Workgroups of a launch over n items, tpb per group, for kernels that walk
// their items with a grid stride.  A dispatch counts its size in work-items
// with 32 bits: grids of 2^32 work-items or more wrap around and leave items
// unvisited, so large launches are capped (2^20 groups) and loop instead.
inline unsigned stride_grid(long long n, int tpb) {
    const long long g = (n + tpb - 1) / tpb;
    return (unsigned)std::max(1LL, std::min(g, 1LL << 20));
}

// Block-major volume geometry.  The FIM domain is the reference's
// pad-to-4 extent (n4*); blocks of edge B tile the pad-to-B extent, cells
// beyond n4* are outside the domain and stay at kInf.
struct BlockGeo {
    int nx, ny, nz;        // model extent
    int n4x, n4y, n4z;     // FIM domain (pad to 4, block_FIM.cpp:144-150)
    int B;                 // block edge
    int nbx, nby, nbz;     // blocks per axis
    long long nblk;        // blocks
    long long vol;         // cells per shot in block layout = nblk * B^3
};

inline BlockGeo make_geo(int nx, int ny, int nz, int B) {
    auto pad = [](int n, int b) { return n + (b - n % b) % b; };
    BlockGeo g;
    g.nx = nx; g.ny = ny; g.nz = nz;
    g.n4x = pad(nx, 4); g.n4y = pad(ny, 4); g.n4z = pad(nz, 4);
    g.B = B;
    g.nbx = pad(g.n4x, B) / B; g.nby = pad(g.n4y, B) / B; g.nbz = pad(g.n4z, B) / B;
    g.nblk = (long long)g.nbx * g.nby * g.nbz;
    g.vol = g.nblk * B * B * B;
    return g;
}

// Address of cell (x, y, z) in block layout: block b = bx + nbx*(by +
// nby*bz) (x fastest, as block_FIM.cpp:271-292), in-block t = tx + B*ty +
// B*B*tz.
__host__ __device__ inline long long blk_addr(int x, int y, int z, int B, int nbx, int nby) {
    int bx = x / B, by = y / B, bz = z / B;
    int tx = x - bx * B, ty = y - by * B, tz = z - bz * B;
    long long b = bx + (long long)nbx * (by + (long long)nby * bz);
    return b * (B * B * B) + tx + B * (ty + B * tz);
}

}  // namespace st

struct st_ctx;
namespace st {
// The device slowness (slow_zxy) changed: a new model, a model update or the
// training-set generator's temporary models.  Drops every copy derived from
// it (block layouts, the sweeping solvers' padded copy, the slowness range).
void model_changed(st_ctx *ctx);
// Recomputes smin/smax on the device when they are stale and requires a
// finite, positive slowness everywhere (ST_ERR_ARG otherwise).
void ensure_range(st_ctx *ctx);
}  // namespace st

namespace st {
// Device buffers of one batched FIM solve (sizes as fim_batch reserves them
// for nshot shots: T nshot*vol, flags nshot*nblk, each list
// fim_list_bytes(nblk, nshot), kmin 3*nshot, tk 2*nshot*nblk, wstats
// 3*fim_waves(opts) counters).
struct FimBufs {
    const float *S = nullptr;      // block-layout slowness
    float *T = nullptr;
    unsigned *flags = nullptr, *list0 = nullptr, *list1 = nullptr;
    unsigned *kmin = nullptr, *tk = nullptr;
    unsigned long long *wstats = nullptr;
    float *xf = nullptr;           // 8^3 blocks: copies of each block's two x faces
                                   // (nshot * nblk * 128 floats; nullable)
    bool graphs = false;           // the context's round-graph cache may be used
};
}  // namespace st

// The context.
struct st_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string last_error;

    // model
    bool has_model = false;
    int nx = 0, ny = 0, nz = 0;
    float dx = 0, dy = 0, dz = 0;
    float smin = 0, smax = 0;          // slowness range of the model
    bool range_valid = false;          // smin/smax describe slow_zxy
    st::DevBuf slow_zxy;               // slowness, z fastest (for updates)
    st::DevBuf range;                  // device min/max of the slowness bits
    st::DevBuf slow_blk[2];            // block layouts for B = 4, 8
    bool blk_valid[2] = {false, false};

    // FIM state
    st::DevBuf T;                      // batch x vol travel times
    st::DevBuf flags;                  // batch x nblk schedule bits
    st::DevBuf list[2];                // active lists (by round parity)
    st::DevBuf counters;               // list counts, stats
    st::DevBuf wstats;                 // per-wave FIM statistics slots
    st::DevBuf fsm_S, fsm_sid;         // sweeping solver: padded slowness, source nodes
    st::DevBuf pod_T, pod_stamp, pod_list, pod_newv, pod_counts, pod_nit, pod_temp;   // Podvin
    bool fsm_valid = false;            // fsm_S matches the model
    st::DevBuf src, rcv, picks, scratch;
    st::DevBuf keys, kmin, judge, pred;   // warm-start state
    st::DevBuf tkeys;                  // arrival-ordering keys (2 x batch x nblk)
    st::DevBuf xfaces;                 // 8^3 blocks: x-face copies (batch x nblk x 128)
    void *host_pinned = nullptr;       // small pinned staging area
    void *noalloc = nullptr;           // this device's no-allocation scratch array
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};

    // graph cache: captured rounds for one (geometry, batch) shape
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    long long graph_key = -1;

    // the eikonal solver of the tomography's forward pass (eikonal_type of
    // Least_squares::set_parameters: 0 Podvin, 1 block FIM, 2 Accurate_FSM)
    int eikonal_type = 1;

    // opaque extension state (tomography, surrogate, acoustic modelling)
    void *tomo = nullptr;
    void *surr = nullptr;
    void *wave = nullptr;
};
