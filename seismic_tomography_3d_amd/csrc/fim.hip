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

constexpr int kSeg = kListSegments;
constexpr int kCntStride = 32;                       // one 128-byte line per counter
constexpr int kRoundsIdx = 3 * kSeg * kCntStride;    // rounds with a non-empty list
constexpr int kStatsIdx = kRoundsIdx + kCntStride;   // 3 x u64 summed statistics
constexpr int kCntWords = kStatsIdx + 8;

struct FimArgs {
    const float *S;            // slowness, block layout (shared by shots)
    float *T;                  // travel times, shot-major block layout
    unsigned *flags;           // per (shot, block): bit p = scheduled for parity p
    unsigned *list0, *list1;   // active lists by round parity, kSeg segments of cap
    long long cap;             // items per segment (nblk * nshot)
    unsigned *cnt;             // segment lengths: cnt[(slot * kSeg + seg) * kCntStride]
                               // for rotating slots 0..2; rounds run at cnt[kRoundsIdx]
    unsigned long long *stats; // per wave gw: [3 gw] visits, [3 gw + 1] inner
                               // iterations, [3 gw + 2] deferrals
    int nbx, nby, nbz;
    int n4x, n4y, n4z;
    long long vol, nblk;
    float h, eps;
    int max_inner;
    int round;                 // round index modulo 6
    // predicted-arrival scheduling (nullable keys = plain FIM): a block whose
    // key exceeds (smallest key of its shot's current list + band) is
    // deferred to the next round instead of visited
    const float *keys;         // per (shot, block) predicted first arrival
    unsigned *kmin;            // [3][nshot] per-shot smallest key of each list (float bits)
    int nshot;
    float band;
    const float *src;          // the batch's sources, 3 per shot (x, y, z in m)
    // arrival ordering (nullable; replaces keys when set): per list parity
    // and (shot, block), the smallest face value (float bits) through which
    // a visit scheduled the block -- a lower bound on the block's arrival
    // known from the data itself.  [2][nshot * nblk]
    unsigned *tk;
    // 8^3 blocks (nullable): a copy of each block's two x faces, [shot][block]
    // [face: 0 tx = 0, 1 tx = 7][ty + 8 tz], written with the block, so a
    // neighbour's x halo is 64 contiguous floats instead of one float in each
    // 32-byte row of the block
    float *xf;
    // 8^3 cold starts without the volume fills: the first wave to schedule
    // a block claims it (flag bit 2), fills its cells and x faces with kInf
    // and stamps the flag word (bits 3 and up) with epoch + 2 of its launch
    // (1: the source blocks, filled before the first launch).  A block is
    // first visited in a later launch than its claim, so every visited block
    // is filled; a visit trusts a neighbour stamped by an earlier launch
    // (its fill and writes are visible) and reads any other neighbour as
    // unreached (kInf) -- unclaimed, or claimed in this launch and so not
    // yet visited: what a cold fill would hold.  Blocks never scheduled are
    // filled after the solve.
    int lazy;
    int epoch;                 // launch index in the call (lazy only)
};

constexpr unsigned kInfBits = 0x7F800000u;

constexpr int kShotShift = 22;
constexpr unsigned kBlkMask = (1u << kShotShift) - 1u;

__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// First-order Godunov upwind update from the smallest neighbour of each axis
// (a, b, c), hs = h*s: the 1D candidate lo + hs while it does not exceed the
// middle neighbour, else the 2D solution while it does not exceed the
// largest, else the 3D solution -- the branch structure of get_time_eikonal
// (cuda_kernel.cu:217-248), evaluated without branches (all three
// candidates, then two selects) so that a wave issues no exec-mask
// sequences.  The 3D discriminant is 3(hs)^2 - [(a-b)^2 + (b-c)^2 + (c-a)^2]
// in sorted differences d = mid - lo, e1 = hi - mid, e3 = hi - lo = d + e1:
// the bracket equals 2 (d e3 + e1^2), with no cancellation at large T (the
// reference's expanded 2(a(b-a)+b(c-b)+c(a-c)) + 3(hs)^2 has it).  No square
// root ever sees a negative argument, so no NaN is formed: the 3D one takes
// max(., 0), the 2D one the absolute value (exact whenever it is selected).  Unreached neighbours are the
// finite sentinel kInf; a cell whose neighbours are all kInf gets a value
// above kInf, which min(T, .) discards.
__device__ __forceinline__ float godunov(float a, float b, float c, float hs, float hs2) {
    const float lo = fminf(fminf(a, b), c);
    const float mid = __builtin_amdgcn_fmed3f(a, b, c);
    const float hi = fmaxf(fmaxf(a, b), c);
    const float t1 = lo + hs;
    const float d = mid - lo;
    const float sm = mid + lo;
    // |.| (a free source modifier of v_sqrt): whenever t2 is selected, t1 =
    // lo + hs > mid, so d < hs (up to rounding) and the argument exceeds hs^2
    const float t2 = 0.5f * (sm + fsqrt(fabsf(2.0f * hs2 - d * d)));
    const float e1 = hi - mid, e3 = hi - lo;
    const float u = d * e3 + e1 * e1;
    const float t3 = ((hi + sm) + fsqrt(fmaxf(3.0f * hs2 - 2.0f * u, 0.0f))) * (1.0f / 3.0f);
    const float r = t1 > mid ? t2 : t1;
    return r > hi ? t3 : r;
}

// Schedules the candidates held by lanes 0..6 for the next round and
// appends the newly scheduled ones to the next list.
__device__ __forceinline__ void schedule(const FimArgs &A, unsigned shot, unsigned *fl,
                                         unsigned target, unsigned bit_nxt, unsigned *nxt,
                                         unsigned *nxt_cnt, int lane, float key) {
    bool fresh = false, claim = false;
    if (target != 0xFFFFFFFFu) {
        const unsigned old = atomicOr(&fl[target], bit_nxt | (A.lazy ? 4u : 0u));
        fresh = !(old & bit_nxt);
        claim = A.lazy && !(old & 4u);
        // the claim's stamp (one writer, so an OR)
        if (claim) atomicOr(&fl[target], ((unsigned)A.epoch + 2u) << 3);
        if (A.tk) {
            // every scheduler lowers the next list's key (the item itself is
            // appended once, by the first); with predictions too, the block
            // counts as due no earlier than its predicted arrival
            atomicMin(&A.tk[((long long)((A.round + 1) & 1) * A.nshot + shot) * A.nblk + target],
                      __float_as_uint(key));
            const float eff = A.keys ? fmaxf(key, A.keys[(long long)shot * A.nblk + target]) : key;
            atomicMin(&A.kmin[((A.round + 1) % 3) * A.nshot + shot], __float_as_uint(eff));
        } else if (fresh && A.keys) {
            atomicMin(&A.kmin[((A.round + 1) % 3) * A.nshot + shot],
                      __float_as_uint(A.keys[(long long)shot * A.nblk + target]));
        }
    }
    const unsigned long long fm = __ballot(fresh);
    if (fm) {
        unsigned base = 0;
        if (lane == 0) base = atomicAdd(nxt_cnt, (unsigned)__popcll(fm));
        base = (unsigned)__builtin_amdgcn_readlane((int)base, 0);
        if (fresh) {
            const unsigned rank = (unsigned)__popcll(fm & ((1ull << lane) - 1ull));
            nxt[base + rank] = (shot << kShotShift) | target;
        }
    }
    // claimed blocks' cold fill, by the whole wave: 512 cells, 128 face values
    unsigned long long cm = __ballot(claim);
    while (cm) {
        const int j = __builtin_ctzll(cm);
        cm &= cm - 1ull;
        const long long tb = (long long)(unsigned)__builtin_amdgcn_readlane((int)target, j);
        float4 *p = reinterpret_cast<float4 *>(A.T + (long long)shot * A.vol + tb * 512);
        p[lane] = make_float4(kInf, kInf, kInf, kInf);
        p[lane + 64] = make_float4(kInf, kInf, kInf, kInf);
        if (A.xf)
            reinterpret_cast<float2 *>(A.xf + ((long long)shot * A.nblk + tb) * 128)[lane] =
                make_float2(kInf, kInf);
    }
}

// Minimum over the lanes whose index differs from this lane's only in the
// bits of mask (a butterfly over those bits).
// Cross-lane minima in the VALU (DPP), without LDS permutes: a DPP move
// reads a lane of the same 16-lane row (every lane active where these run).
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
    return __int_as_float(
        __builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), CTRL, 0xF, 0xF, false));
}
__device__ __forceinline__ float lane_val(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
// Minima over lanes of values that are never negative (travel times and
// +inf): their bit patterns order as unsigned integers, and an unsigned min
// whose DPP source falls back to its identity 0xFFFFFFFF compiles to one
// v_min_u32_dpp per step (a float min takes a copy, a DPP move and the min).
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ unsigned dpp_umin(unsigned v) {
    return min(v, (unsigned)__builtin_amdgcn_update_dpp((int)0xFFFFFFFFu, (int)v, CTRL, ROWS, 0xF,
                                                         false));
}
// minimum over each group of 8 lanes (quad swaps 1 and 2, then the half-row
// mirror pairs the two quads)
__device__ __forceinline__ float min8(float f) {
    unsigned v = __float_as_uint(f);
    v = dpp_umin<0xB1>(v);    // quad_perm [1, 0, 3, 2]
    v = dpp_umin<0x4E>(v);    // quad_perm [2, 3, 0, 1]
    v = dpp_umin<0x141>(v);   // row_half_mirror
    return __uint_as_float(v);
}
// minimum over the whole wave (values never negative), the same in every
// lane: rows by DPP, then row_bcast:15 combines rows 1 and 3 with the last
// lane of the row before and row_bcast:31 rows 2 and 3 with lane 31, so lane
// 63 ends with the whole wave's minimum (the other rows keep theirs)
__device__ __forceinline__ float wave_min_all(float f) {
    unsigned v = __float_as_uint(min8(f));
    v = dpp_umin<0x140>(v);         // row_mirror: each row's 16 lanes
    v = dpp_umin<0x142, 0xA>(v);    // row_bcast:15 into rows 1 and 3
    v = dpp_umin<0x143, 0xC>(v);    // row_bcast:31 into rows 2 and 3
    return __uint_as_float((unsigned)__builtin_amdgcn_readlane((int)v, 63));
}

// maximum over the whole wave, the same in every lane (any sign)
__device__ __forceinline__ float wave_max_all(float v) {
    v = fmaxf(v, dpp_mov<0xB1>(v));
    v = fmaxf(v, dpp_mov<0x4E>(v));
    v = fmaxf(v, dpp_mov<0x141>(v));
    v = fmaxf(v, dpp_mov<0x140>(v));
    return fmaxf(fmaxf(lane_val(v, 0), lane_val(v, 16)), fmaxf(lane_val(v, 32), lane_val(v, 48)));
}

template <int MASK>
__device__ __forceinline__ float lane_min(float v) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1)
        if (MASK & o) v = fminf(v, __shfl_xor(v, o));
    return v;
}

constexpr float kKeyInf = __builtin_inff();  // "no key" for the arrival ordering

// Round kernels run kWaves independent waves per workgroup (each with its
// own LDS tile and item stream): workgroup slots per CU, not registers, cap
// single-wave launches at four waves per SIMD.  Waves exchange nothing, so
// they order their own LDS traffic with wave-scope fences instead of
// s_barrier (which would couple waves with different trip counts).
constexpr int kWaves = 2;

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Segment lengths of this round's list: lane k < kSeg gets the inclusive
// prefix over segments 0..k in incl, every lane the total in count.
__device__ __forceinline__ void round_counts(const FimArgs &A, int lane, unsigned &incl,
                                             unsigned &count) {
    const int slot = A.round % 3;
    unsigned v = lane < kSeg ? A.cnt[(slot * kSeg + lane) * kCntStride] : 0u;
    // inclusive scan inside the first 16-lane row by DPP row shifts (a lane
    // with no source lane takes 0)
    static_assert(kSeg == 16, "one DPP row of segments");
    v += (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false);  // row_shr:1
    v += (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, false);  // row_shr:2
    v += (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false);  // row_shr:4
    v += (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false);  // row_shr:8
    incl = v;
    count = (unsigned)__builtin_amdgcn_readlane((int)v, kSeg - 1);
}

// Walks this wave's items of the current list (i = gw + j * G for global
// wave gw of G, the same interleaving for every round) 64 at a time: lane l
// reads item j0 + l.  With arrival ordering, the items whose block is not
// due yet (predicted arrival above its shot's earliest scheduled key plus
// the band) move to the next list right here, all lanes at once with one
// append; the others are visited one after another by the whole wave.
template <class Visit>
__device__ __forceinline__ void for_items(const FimArgs &A, unsigned count, unsigned incl,
                                          const unsigned *cur, unsigned *nxt, unsigned *nxt_cnt,
                                          unsigned bit_cur, unsigned bit_nxt, int lane,
                                          unsigned gw, unsigned G, unsigned long long &my_defer,
                                          Visit visit) {
    for (unsigned long long j0 = 0; gw + j0 * G < count; j0 += 64) {
        const unsigned long long i = gw + (j0 + lane) * G;
        const bool in = i < count;
        // flat index -> (segment, offset): incl holds, in lane k < kSeg, the
        // inclusive prefix of the segment lengths
        unsigned seg = 0, before = 0;
#pragma unroll
        for (int k = 0; k < kSeg - 1; k++) {
            const unsigned ik = (unsigned)__builtin_amdgcn_readlane((int)incl, k);
            if (i >= ik) { seg = k + 1; before = ik; }
        }
        const unsigned item = in ? cur[seg * A.cap + (i - before)] : 0u;
        bool run = in;
        if (A.keys || A.tk) {
            bool fresh = false;
            unsigned shot = 0;
            if (in) {
                shot = item >> kShotShift;
                const unsigned b = item & kBlkMask;
                const float tau =
                    __uint_as_float(A.kmin[(A.round % 3) * A.nshot + shot]) + A.band;
                float key, kt = 0.0f;
                if (A.tk) {
                    // this list's key is consumed here (visited or carried
                    // over); only the next list's keys are written this round
                    unsigned *kc = A.tk + ((long long)(A.round & 1) * A.nshot + shot) * A.nblk + b;
                    kt = __uint_as_float(*kc);
                    *kc = kInfBits;
                    key = A.keys ? fmaxf(kt, A.keys[(long long)shot * A.nblk + b]) : kt;
                } else {
                    key = A.keys[(long long)shot * A.nblk + b];
                }
                if (key > tau) {  // not its time yet: carried to the next round
                    run = false;
                    unsigned *fl = A.flags + (long long)shot * A.nblk;
                    atomicAnd(&fl[b], ~bit_cur);
                    fresh = !(atomicOr(&fl[b], bit_nxt) & bit_nxt);
                    if (A.tk) {
                        atomicMin(&A.tk[((long long)((A.round + 1) & 1) * A.nshot + shot) * A.nblk + b],
                                  __float_as_uint(kt));
                        atomicMin(&A.kmin[((A.round + 1) % 3) * A.nshot + shot], __float_as_uint(key));
                    } else if (fresh) {
                        atomicMin(&A.kmin[((A.round + 1) % 3) * A.nshot + shot],
                                  __float_as_uint(key));
                    }
                }
            }
            const unsigned long long fm = __ballot(fresh);
            if (fm) {
                unsigned base = 0;
                if (lane == 0) base = atomicAdd(nxt_cnt, (unsigned)__popcll(fm));
                base = (unsigned)__builtin_amdgcn_readlane((int)base, 0);
                if (fresh) nxt[base + (unsigned)__popcll(fm & ((1ull << lane) - 1ull))] = item;
            }
            my_defer += (unsigned long long)__popcll(__ballot(in && !run));
        }
        unsigned long long m = __ballot(run);
        while (m) {
            const int j = __builtin_ctzll(m);
            m &= m - 1ull;
            visit((unsigned)__builtin_amdgcn_readlane((int)item, j));
        }
    }
}

// ---------------------------------------------------------------------------
// 4^3 blocks: lane = tx + 4 ty + 16 tz, one cell per lane, Jacobi sweeps.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64 * kWaves) void fim_round4(FimArgs A) {
    __shared__ float tiles[kWaves][216];  // (4+2)^3, index x + 6 y + 36 z
    const int lane = threadIdx.x & 63;
    float *tile = tiles[threadIdx.x >> 6];
    const unsigned gw = blockIdx.x * kWaves + (threadIdx.x >> 6), G = gridDim.x * kWaves;
    const int p = A.round & 1;
    unsigned incl = 0, count = 0;
    round_counts(A, lane, incl, count);
    if (gw == 0) {
        if (lane < kSeg) A.cnt[(((A.round + 2) % 3) * kSeg + lane) * kCntStride] = 0;
        if (lane == 0 && count) atomicAdd(&A.cnt[kRoundsIdx], 1u);
        if (A.keys || A.tk)
            for (int j = lane; j < A.nshot; j += 64) A.kmin[((A.round + 2) % 3) * A.nshot + j] = kInfBits;
    }
    if (count == 0) return;
    const unsigned *cur = p ? A.list1 : A.list0;
    // this wave appends to segment gw mod kSeg of the next list
    const int myseg = (int)(gw & (kSeg - 1));
    unsigned *nxt = (p ? A.list0 : A.list1) + myseg * A.cap;
    unsigned *nxt_cnt = &A.cnt[(((A.round + 1) % 3) * kSeg + myseg) * kCntStride];
    const unsigned bit_cur = 1u << p, bit_nxt = 1u << (p ^ 1);

    const int tx = lane & 3, ty = (lane >> 2) & 3, tz = lane >> 4;
    const int tc = (tx + 1) + 6 * (ty + 1) + 36 * (tz + 1);
    const int nbxy = A.nbx * A.nby;
    unsigned long long my_visits = 0, my_iters = 0, my_defer = 0;

    for_items(A, count, incl, cur, nxt, nxt_cnt, bit_cur, bit_nxt, lane, gw, G, my_defer, [&](const unsigned item) {
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
        wave_sync();

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
            wave_sync();
            tile[tc] = nv;
            v = nv;
            ++it;
            more = __ballot(ch) != 0ull;
            wave_sync();
            if (!more) break;
        }

        const float dl = v0 - v;
        if (__ballot(dl > 0.0f)) Tsh[base + lane] = v;
        // a changed face cell can lower the neighbour only if it is now
        // below the neighbour's adjacent cell (held in the halo since the
        // visit began; the neighbour has only decreased since, so the test
        // never misses a needed visit)
        const bool ch = dl > A.eps;
        const bool xm = __ballot(ch && tx == 0 && v < tile[tc - 1]) != 0ull;
        const bool xp = __ballot(ch && tx == 3 && v < tile[tc + 1]) != 0ull;
        const bool ym = __ballot(ch && ty == 0 && v < tile[tc - 6]) != 0ull;
        const bool yp = __ballot(ch && ty == 3 && v < tile[tc + 6]) != 0ull;
        const bool zm = __ballot(ch && tz == 0 && v < tile[tc - 36]) != 0ull;
        const bool zp = __ballot(ch && tz == 3 && v < tile[tc + 36]) != 0ull;
        unsigned target = 0xFFFFFFFFu;
        switch (lane) {
        case 0: if (more) target = b; break;
        case 1: if (xm && bx > 0) target = b - 1; break;
        case 2: if (xp && bx < A.nbx - 1) target = b + 1; break;
        case 3: if (ym && by > 0) target = b - A.nbx; break;
        case 4: if (yp && by < A.nby - 1) target = b + A.nbx; break;
        case 5: if (zm && bz > 0) target = b - nbxy; break;
        case 6: if (zp && bz < A.nbz - 1) target = b + nbxy; break;
        default: break;
        }
        float key = kKeyInf;
        if (A.tk) {
            // per face, the smallest new value among the cells that activate
            // the neighbour; for a self-reschedule, among the changed cells
            auto face = [&](bool c) { return lane_min<63>(c ? v : kKeyInf); };
            const float k0 = face(ch), k1 = face(ch && tx == 0 && v < tile[tc - 1]),
                        k2 = face(ch && tx == 3 && v < tile[tc + 1]),
                        k3 = face(ch && ty == 0 && v < tile[tc - 6]),
                        k4 = face(ch && ty == 3 && v < tile[tc + 6]),
                        k5 = face(ch && tz == 0 && v < tile[tc - 36]),
                        k6 = face(ch && tz == 3 && v < tile[tc + 36]);
            key = lane == 0 ? k0 : lane == 1 ? k1 : lane == 2 ? k2 : lane == 3 ? k3
                : lane == 4 ? k4 : lane == 5 ? k5 : k6;
        }
        schedule(A, shot, fl, target, bit_nxt, nxt, nxt_cnt, lane, key);
        my_visits += 1;
        my_iters += it;
        wave_sync();
    });
    if (lane == 0 && A.stats && (my_visits | my_defer)) {
        // each wave owns its slot: no atomics on shared counters
        unsigned long long *w = A.stats + 3ull * gw;
        w[0] += my_visits;
        w[1] += my_iters;
        w[2] += my_defer;
    }
}

// ---------------------------------------------------------------------------
// 8^3 blocks: lane = tx + 8 ty owns the z column tz = 0..7 in registers.
// x/y neighbours come through the LDS tile, z neighbours from registers
// (Gauss-Seidel along z, alternating direction each sweep).
// HYP: each visit first makes one full Gauss-Seidel pass in the octant
// pointing away from the shot's source -- cells taken by diagonal planes
// x' + y' + z' = d (coordinates flipped into the octant), 22 planes of at
// most 64 cells, one lane per column; a cell's upwind neighbours lie on
// plane d - 1, updated one step before -- so a front crossing the block
// in that octant is carried through it in one pass; the z sweeps then
// confirm or finish the fixed point.
// ---------------------------------------------------------------------------
// Six waves per SIMD (80 VGPRs, a few spilled values outside the plane
// loops) measured 10 % faster than five without spills.
//
// Tile layout: cell (x, y, z) of the (8+2)^3 tile (halo at 0 and 9) at word
// x + 104 y + 10 z - 11.  gfx950's LDS serves a wave64 access from 64 banks:
// the z sweeps (all lanes on one z: own cell and its x / y neighbours) then
// touch 64 distinct banks, since x + 104 y for x, y in 1..8 covers each bank
// once (104 = 40 mod 64), and the diagonal pass, whose lanes sit on
// different z, averages 0.55 extra cycles per access.  With x + 10 y + 100 z
// every z-sweep access cost one extra cycle; measured, the layout makes the
// 8^3 solve 7 % faster.  Only the interior and the six face halos are ever
// addressed (never an edge or corner of the tile): on them the map is
// one-to-one and spans words 0..1013.
constexpr int kTX = 1, kTY = 104, kTZ = 10, kTBase = 11;
__host__ __device__ __forceinline__ constexpr int tix(int x, int y, int z) {
    return kTX * x + kTY * y + kTZ * z - kTBase;
}
constexpr int kTile8 = tix(8, 9, 8) + 1;
static_assert(tix(1, 0, 1) == 0 && tix(0, 1, 1) > 0 && tix(1, 1, 0) > 0, "tile base");
static_assert(tix(9, 8, 8) < kTile8 && tix(8, 8, 9) < kTile8 && kTile8 == 1014, "tile span");
template <bool HYP>
__global__ __launch_bounds__(64 * kWaves, 6) void fim_round8(FimArgs A) {
    __shared__ float tiles[kWaves][kTile8];  // (8+2)^3 at tix(x, y, z)
    __shared__ float hss[HYP ? kWaves : 1][512];  // h * slowness, tx + 8 ty + 64 z
    const int lane = threadIdx.x & 63;
    float *tile = tiles[threadIdx.x >> 6];
    float *hsl = hss[HYP ? (threadIdx.x >> 6) : 0];
    const unsigned gw = blockIdx.x * kWaves + (threadIdx.x >> 6), G = gridDim.x * kWaves;
    const int p = A.round & 1;
    unsigned incl = 0, count = 0;
    round_counts(A, lane, incl, count);
    if (gw == 0) {
        if (lane < kSeg) A.cnt[(((A.round + 2) % 3) * kSeg + lane) * kCntStride] = 0;
        if (lane == 0 && count) atomicAdd(&A.cnt[kRoundsIdx], 1u);
        if (A.keys || A.tk)
            for (int j = lane; j < A.nshot; j += 64) A.kmin[((A.round + 2) % 3) * A.nshot + j] = kInfBits;
    }
    if (count == 0) return;
    const unsigned *cur = p ? A.list1 : A.list0;
    // this wave appends to segment gw mod kSeg of the next list
    const int myseg = (int)(gw & (kSeg - 1));
    unsigned *nxt = (p ? A.list0 : A.list1) + myseg * A.cap;
    unsigned *nxt_cnt = &A.cnt[(((A.round + 1) % 3) * kSeg + myseg) * kCntStride];
    const unsigned bit_cur = 1u << p, bit_nxt = 1u << (p ^ 1);

    const int tx = lane & 7, ty = lane >> 3;
    const int tcol = tix(tx + 1, ty + 1, 0);  // tile index of the column's z halo below
    const int nbxy = A.nbx * A.nby;
    unsigned long long my_visits = 0, my_iters = 0, my_defer = 0;

    for_items(A, count, incl, cur, nxt, nxt_cnt, bit_cur, bit_nxt, lane, gw, G, my_defer, [&](const unsigned item) {
        const unsigned shot = item >> kShotShift, b = item & kBlkMask;
        float *Tsh = A.T + (long long)shot * A.vol;
        unsigned *fl = A.flags + (long long)shot * A.nblk;
        const int bx = b % A.nbx, rest = b / A.nbx, by = rest % A.nby, bz = rest / A.nby;
        if (lane == 0) atomicAnd(&fl[b], ~bit_cur);
        // lazy starts: lanes 1..6 load the flag words of the -x, +x, -y, +y,
        // -z, +z neighbours (in flight with the block's own values); okm bit
        // i: neighbour i's values are trusted (bit 0, the block itself, always)
        unsigned fw = 0u, okm = 0x7Fu;
        if (A.lazy) {
            long long nb = -1;
            switch (lane) {
            case 1: if (bx > 0) nb = (long long)b - 1; break;
            case 2: if (bx < A.nbx - 1) nb = (long long)b + 1; break;
            case 3: if (by > 0) nb = (long long)b - A.nbx; break;
            case 4: if (by < A.nby - 1) nb = (long long)b + A.nbx; break;
            case 5: if (bz > 0) nb = (long long)b - nbxy; break;
            case 6: if (bz < A.nbz - 1) nb = (long long)b + nbxy; break;
            default: break;
            }
            if (nb >= 0) fw = fl[nb];
        }
        const long long base = (long long)b * 512;
        const bool vxy = (bx * 8 + tx < A.n4x) && (by * 8 + ty < A.n4y);

        // vz[0] / vz[9]: z halos below / above the column; vz[1..8]: the cells
        float vz[10], v0[8], hs[8];
        #pragma unroll
        for (int k = 0; k < 8; k++) {
            vz[k + 1] = Tsh[base + lane + 64 * k];
            hs[k] = A.h * A.S[base + lane + 64 * k];
        }
        // the halos are loaded whenever the neighbour exists, in flight with
        // the flag words, and kept if trusted: a neighbour's trust cannot
        // change within a launch, so values read from an untrusted block are
        // never used
        const float zlo = (bz > 0) ? Tsh[base - 512LL * nbxy + lane + 64 * 7] : kInf;
        const float zhi = (bz < A.nbz - 1) ? Tsh[base + 512LL * nbxy + lane] : kInf;
        // x / y face halos: face f = 0..3 (-x, +x, -y, +y), lane = u + 8 w
        // (u = in-face coordinate, w = z)
        const int u = lane & 7, w = lane >> 3;
        float hv[4];
        hv[0] = bx > 0 ? (A.xf ? A.xf[((long long)shot * A.nblk + b - 1) * 128 + 64 + lane]
                               : Tsh[base - 512 + 7 + 8 * u + 64 * w]) : kInf;
        hv[1] = bx < A.nbx - 1 ? (A.xf ? A.xf[((long long)shot * A.nblk + b + 1) * 128 + lane]
                                       : Tsh[base + 512 + 8 * u + 64 * w]) : kInf;
        hv[2] = by > 0 ? Tsh[base - 512LL * A.nbx + u + 56 + 64 * w] : kInf;
        hv[3] = by < A.nby - 1 ? Tsh[base + 512LL * A.nbx + u + 64 * w] : kInf;
        if (A.lazy) {
            const unsigned st = fw >> 3;
            okm = (unsigned)__ballot(st != 0u && st <= (unsigned)A.epoch + 1u) | 1u;
        }
        #pragma unroll
        for (int k = 0; k < 8; k++) {
            v0[k] = vz[k + 1];
            tile[tcol + kTZ * (k + 1)] = vz[k + 1];
        }
        vz[0] = (okm & 32u) ? zlo : kInf;
        vz[9] = (okm & 64u) ? zhi : kInf;
        #pragma unroll
        for (int f = 0; f < 4; f++) hv[f] = (okm & (2u << f)) ? hv[f] : kInf;
        tile[tix(0, u + 1, w + 1)] = hv[0];
        tile[tix(9, u + 1, w + 1)] = hv[1];
        tile[tix(u + 1, 0, w + 1)] = hv[2];
        tile[tix(u + 1, 9, w + 1)] = hv[3];
        int it = 0, zflip = 0;
        if constexpr (HYP) {
            tile[tcol] = vz[0];
            tile[tcol + 9 * kTZ] = vz[9];
#pragma unroll
            for (int k = 0; k < 8; k++) hsl[lane + 64 * k] = hs[k];
            wave_sync();
            // octant: along each axis the front enters through the face
            // holding the earlier arrivals (diving waves reach shallow blocks
            // from below); on a tie, away from the source
            // (the face halos from the registers that staged them)
            auto wmin = [](float v) { return wave_min_all(v); };
            const float mxm = wmin(hv[0]), mxp = wmin(hv[1]);
            const float mym = wmin(hv[2]), myp = wmin(hv[3]);
            const float mzm = wmin(vz[0]), mzp = wmin(vz[9]);
            const float *sp = A.src + 3 * shot;
            const bool fx = mxm != mxp ? mxm < mxp : (bx * 8 + 3.5f) * A.h >= sp[0];
            const bool fy = mym != myp ? mym < myp : (by * 8 + 3.5f) * A.h >= sp[1];
            const bool fz = mzm != mzp ? mzm < mzp : (bz * 8 + 3.5f) * A.h >= sp[2];
            const int s0 = (fx ? tx : 7 - tx) + (fy ? ty : 7 - ty);
            // the lane's cell on plane d is z' = d - s0, i.e. column level
            // k(d) = kb + ks * d: tile and h*s addresses are linear in d.  It
            // updates on the planes d in [dlo, dlo + span]: 0 <= z' < 8 and
            // the cell inside the model (k < zmax; the column inside, vxy).
            const int zmax = A.n4z - bz * 8;  // >= 1: the block holds model cells
            const int ks = fz ? 1 : -1, kb = fz ? -s0 : 7 + s0;
            const int zlo = fz ? 0 : max(0, 8 - zmax), zhi = fz ? min(7, zmax - 1) : 7;
            const int dlo = vxy ? s0 + zlo : 64;
            const unsigned span = (unsigned)(zhi - zlo);
            const int i0 = tcol + kTZ * (kb + 1), h0 = lane + 64 * kb;
            // (indexed from the LDS arrays themselves: 32-bit addresses)
            const int w = threadIdx.x >> 6;
#pragma unroll
            for (int d = 0; d < 22; d++) {
                if ((unsigned)(d - dlo) <= span) {
                    const int i = i0 + kTZ * ks * d;
                    const float own = tiles[w][i];
                    const float a = fminf(tiles[w][i - kTX], tiles[w][i + kTX]);
                    const float bb = fminf(tiles[w][i - kTY], tiles[w][i + kTY]);
                    const float c = fminf(tiles[w][i - kTZ], tiles[w][i + kTZ]);
                    const float hk = hss[w][h0 + 64 * ks * d];
                    const float t = godunov(a, bb, c, hk, hk * hk);
                    tiles[w][i] = fminf(own, t);
                }
                wave_sync();
            }
            // the z sweeps start in the pass's z direction
            zflip = fz ? 1 : 0;
#pragma unroll
            for (int k = 0; k < 8; k++) vz[k + 1] = tile[tcol + kTZ * (k + 1)];
            it = 1;
        }
        wave_sync();

        bool more = true;
        // dirty: z planes whose inputs changed since their last update.  A
        // plane reads its own plane (x/y neighbours, Jacobi) and planes k - 1
        // and k + 1 (z neighbours); the halos are fixed for the visit.  A
        // plane none of whose inputs changed reproduces its values, so it is
        // skipped (wave-uniform), and the visit ends when no plane is dirty.
        unsigned dirty = 0xFFu;
        while (it < A.max_inner) {
            // one plane of the column: x/y neighbours from the tile (Jacobi
            // across lanes), z neighbours from registers (Gauss-Seidel)
            auto plane = [&](const int k) {
                if (!((dirty >> k) & 1u)) return;
                dirty &= ~(1u << k);
                const int ti = tcol + kTZ * (k + 1);
                const float a = fminf(tile[ti - kTX], tile[ti + kTX]);
                const float bb = fminf(tile[ti - kTY], tile[ti + kTY]);
                const float c = fminf(vz[k], vz[k + 2]);
                // HYP keeps h*s in LDS (the pass reads it there): fewer live registers
                const float hk = HYP ? hsl[lane + 64 * k] : hs[k];
                const float t = godunov(a, bb, c, hk, hk * hk);
                const bool valid = vxy && (bz * 8 + k < A.n4z);
                const float nv = valid ? fminf(vz[k + 1], t) : vz[k + 1];
                const bool ch = (vz[k + 1] - nv) > A.eps;
                vz[k + 1] = nv;
                wave_sync();
                tile[ti] = nv;
                wave_sync();
                if (__ballot(ch)) dirty |= ((7u << k) >> 1) & 0xFFu;
            };
            if (((it + zflip) & 1) == 0) {
                #pragma unroll
                for (int k = 0; k < 8; k++) plane(k);
            } else {
                #pragma unroll
                for (int k = 7; k >= 0; k--) plane(k);
            }
            ++it;
            more = dirty != 0u;
            if (!more) break;
        }
        float v[8];
        #pragma unroll
        for (int k = 0; k < 8; k++) v[k] = vz[k + 1];


        // face activation as in the 4^3 kernel: a changed face cell must be
        // below the neighbour's adjacent cell (tile halo / z halo registers)
        bool anydec = false, axl = false, ayl = false;
        // arrival-ordering keys: smallest activating value on the x faces
        // (tx = 0 / 7 lanes), the y faces (ty = 0 / 7) and among the changed
        // cells of the column
        float kx = kKeyInf, ky = kKeyInf, kc = kKeyInf;
        // a face lane compares with the one halo beyond its own face
        const bool xface = tx == 0 || tx == 7, yface = ty == 0 || ty == 7;
        const int ox = tx == 0 ? -kTX : kTX, oy = ty == 0 ? -kTY : kTY;
        #pragma unroll
        for (int k = 0; k < 8; k++) {
            const float dl = v0[k] - v[k];
            anydec |= dl > 0.0f;
            const bool c = dl > A.eps;
            const int ti = tcol + kTZ * (k + 1);
            const bool ax = c && xface && v[k] < tile[ti + ox];
            const bool ay = c && yface && v[k] < tile[ti + oy];
            axl |= ax;
            ayl |= ay;
            if (A.tk) {
                kx = ax ? fminf(kx, v[k]) : kx;
                ky = ay ? fminf(ky, v[k]) : ky;
                kc = c ? fminf(kc, v[k]) : kc;
            }
        }
        const bool fzm = (v0[0] - v[0] > A.eps) && v[0] < vz[0];
        const bool fzp = (v0[7] - v[7] > A.eps) && v[7] < vz[9];
        if (__ballot(anydec)) {
            #pragma unroll
            for (int k = 0; k < 8; k++) Tsh[base + lane + 64 * k] = v[k];
            if (A.xf) {
                // the face copies from the tile (it holds the final values):
                // lane l writes entries 2l, 2l + 1 of the 128, one float2
                const int f = lane >> 5, q = 2 * (lane & 31);  // q = ty + 8 z
                const int ti = tix(f ? 8 : 1, (q & 7) + 1, (q >> 3) + 1);
                *reinterpret_cast<float2 *>(A.xf + ((long long)shot * A.nblk + b) * 128 + 64 * f + q) =
                    make_float2(tile[ti], tile[ti + kTY]);
            }
        }
        // faces by lane: tx = 0 are lanes 8 ty, tx = 7 lanes 8 ty + 7, ty = 0
        // lanes 0..7, ty = 7 lanes 56..63
        const unsigned long long bx_ = __ballot(axl), by_ = __ballot(ayl);
        const bool xm = (bx_ & 0x0101010101010101ull) != 0ull, xp = (bx_ & 0x8080808080808080ull) != 0ull;
        const bool ym = (by_ & 0x00000000000000FFull) != 0ull, yp = (by_ & 0xFF00000000000000ull) != 0ull;
        const bool zm = __ballot(fzm) != 0ull, zp = __ballot(fzp) != 0ull;
        unsigned target = 0xFFFFFFFFu;
        switch (lane) {
        case 0: if (more) target = b; break;
        case 1: if (xm && bx > 0) target = b - 1; break;
        case 2: if (xp && bx < A.nbx - 1) target = b + 1; break;
        case 3: if (ym && by > 0) target = b - A.nbx; break;
        case 4: if (yp && by < A.nby - 1) target = b + A.nbx; break;
        case 5: if (zm && bz > 0) target = b - nbxy; break;
        case 6: if (zp && bz < A.nbz - 1) target = b + nbxy; break;
        default: break;
        }
        float key = kKeyInf;
        if (A.tk) {
            // x faces: min over ty of the tx = 0 / 7 lanes (lanes tx + 8 ty;
            // row_ror 8 pairs ty with ty ^ 1 inside a row, readlane the rest);
            // y faces: min over tx of the ty = 0 / 7 rows of 8 lanes.  Only
            // for the faces that schedule a neighbour (the same in every lane)
            float k0 = kKeyInf, k1 = kKeyInf, k2 = kKeyInf, k3 = kKeyInf, k4 = kKeyInf,
                  k5 = kKeyInf, k6 = kKeyInf;
            if (xm || xp) {
                const float mx = __uint_as_float(dpp_umin<0x128>(__float_as_uint(kx)));  // row_ror:8
                k1 = fminf(fminf(lane_val(mx, 0), lane_val(mx, 16)),
                           fminf(lane_val(mx, 32), lane_val(mx, 48)));
                k2 = fminf(fminf(lane_val(mx, 7), lane_val(mx, 23)),
                           fminf(lane_val(mx, 39), lane_val(mx, 55)));
            }
            if (ym || yp) {
                const float my = min8(ky);
                k3 = lane_val(my, 0);
                k4 = lane_val(my, 56);
            }
            if (zm) k5 = wave_min_all(fzm ? v[0] : kKeyInf);
            if (zp) k6 = wave_min_all(fzp ? v[7] : kKeyInf);
            if (more) k0 = wave_min_all(kc);
            key = lane == 0 ? k0 : lane == 1 ? k1 : lane == 2 ? k2 : lane == 3 ? k3
                : lane == 4 ? k4 : lane == 5 ? k5 : k6;
        }
        schedule(A, shot, fl, target, bit_nxt, nxt, nxt_cnt, lane, key);
        my_visits += 1;
        my_iters += it;
        wave_sync();
    });
    if (lane == 0 && A.stats && (my_visits | my_defer)) {
        // each wave owns its slot: no atomics on shared counters
        unsigned long long *w = A.stats + 3ull * gw;
        w[0] += my_visits;
        w[1] += my_iters;
        w[2] += my_defer;
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
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
         i += (long long)gridDim.x * blockDim.x)
        T[i] = make_float4(kInf, kInf, kInf, kInf);
}

__global__ void fill_u32(unsigned *__restrict__ p, long long n, unsigned v) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

// After a lazy 8^3 solve: the (shot, block) items never claimed (flag bit 2
// clear: never scheduled, never visited) get their cold fill, 64 items per
// wave and pass.
__global__ void fill_unwritten8(const unsigned *__restrict__ flags, float *__restrict__ T,
                                long long nitems, long long nblk, long long vol) {
    const int lane = threadIdx.x & 63;
    const long long w0 = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long W = ((long long)gridDim.x * blockDim.x) >> 6;
    const float4 inf4 = make_float4(kInf, kInf, kInf, kInf);
    for (long long i0 = w0 * 64; i0 < nitems; i0 += W * 64) {
        const long long i = i0 + lane;
        unsigned long long m = __ballot(i < nitems && !(flags[i] & 4u));
        while (m) {
            const long long it = i0 + __builtin_ctzll(m);
            m &= m - 1ull;
            const long long shot = it / nblk, b = it - shot * nblk;
            float4 *p = reinterpret_cast<float4 *>(T + shot * vol + b * 512);
            p[lane] = inf4;
            p[lane + 64] = inf4;
        }
    }
}

// apply_source_time (block_FIM.cpp:295-369): the source's grid node gets
// S*distance, its block is the only active one.
__global__ void seed_sources(const float *__restrict__ src, int nshot, const float *__restrict__ S,
                             float *T, unsigned *flags, unsigned *list0, unsigned *cnt,
                             float dx, float dy, float dz, int B, int nbx, int nby,
                             long long vol, long long nblk, const float *keys, unsigned *kmin,
                             unsigned *tk, float *xf, int lazy) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j == 0) {  // the sources form segment 0 of the first list
        for (int k = 0; k < 3 * kSeg; k++) cnt[k * kCntStride] = 0;
        cnt[0] = nshot;
        cnt[kRoundsIdx] = 0;
    }
    if (j >= nshot) return;
    const float sx = src[3 * j], sy = src[3 * j + 1], sz = src[3 * j + 2];
    const int ix = (int)(sx / dx), iy = (int)(sy / dy), iz = (int)(sz / dz);
    const long long a = blk_addr(ix, iy, iz, B, nbx, nby);
    const float ex = (float)(ix * dx) - sx, ey = (float)(iy * dy) - sy, ez = (float)(iz * dz) - sz;
    const float t0 = S[a] * sqrtf(ex * ex + ey * ey + ez * ez);
    const unsigned b = (unsigned)(a / (B * B * B));
    if (lazy) {  // the source block's cold fill (8^3 blocks; the others are claimed)
        const float4 inf4 = make_float4(kInf, kInf, kInf, kInf);
        float4 *p = reinterpret_cast<float4 *>(T + (long long)j * vol + (long long)b * 512);
        for (int k = 0; k < 128; k++) p[k] = inf4;
        if (xf) {
            float4 *q = reinterpret_cast<float4 *>(xf + ((long long)j * nblk + b) * 128);
            for (int k = 0; k < 32; k++) q[k] = inf4;
        }
    }
    T[(long long)j * vol + a] = t0;
    if (xf && B == 8) {  // the source cell's x-face copy
        const int tx = ix & 7, ty = iy & 7, tz = iz & 7;
        if (tx == 0 || tx == 7)
            xf[((long long)j * nblk + b) * 128 + (tx == 7 ? 64 : 0) + ty + 8 * tz] = t0;
    }
    flags[(long long)j * nblk + b] = lazy ? 1u | 4u | (1u << 3) : 1u;
    list0[j] = ((unsigned)j << kShotShift) | b;
    if (tk) {  // the source block opens the first list with key 0
        tk[(long long)j * nblk + b] = 0u;
        kmin[j] = keys ? __float_as_uint(keys[(long long)j * nblk + b]) : 0u;
        kmin[nshot + j] = kInfBits;
        kmin[2 * nshot + j] = kInfBits;
    } else if (keys) {
        kmin[j] = __float_as_uint(keys[(long long)j * nblk + b]);
        kmin[nshot + j] = kInfBits;
        kmin[2 * nshot + j] = kInfBits;
    }
}

// block layout -> z-fastest model extent (extract_solution + reduce_model,
// block_FIM.cpp:371-401 and :59-69)
__global__ void extract_volume(const float *__restrict__ T, float *__restrict__ out, int nx, int ny,
                               int nz, int B, int nbx, int nby, long long vol, int nshot) {
    const long long npts = (long long)nx * ny * nz;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < npts * nshot;
         i += (long long)gridDim.x * blockDim.x) {
        const long long sh = i / npts, c = i - sh * npts;
        const int y = (int)(c / ((long long)nx * nz));
        const long long r = c - (long long)y * nx * nz;
        const int x = (int)(r / nz), z = (int)(r - (long long)x * nz);
        out[i] = T[sh * vol + blk_addr(x, y, z, B, nbx, nby)];
    }
}


// ---------------------------------------------------------------------------
// warm start: T0 = (1 + delta) * prediction, cold (kInf) near the source and
// on the pad cells past the model; per-block key = smallest prediction in
// the block (the block's predicted first arrival).  One wave per block.
// ---------------------------------------------------------------------------
// Predictors of the warm fill: begin() runs once per (shot, block) with the
// block's first cell and may stage what the block's cells read; operator()
// gives the predicted travel time of one cell.
struct FinePred {  // caller volumes, z fastest, one per shot
    const float *P;
    int nx, nz;
    long long npts;
    struct Blk {};
    __device__ Blk begin(int, int, int, int, int, float, float, float, float *) const { return {}; }
    __device__ Blk direct(int) const { return {}; }
    __device__ float operator()(const Blk &, int s, int x, int y, int z, float, float, float,
                                float) const {
        return P[s * npts + z + (long long)x * nz + (long long)y * nx * nz];
    }
};

struct CoarsePred {  // network output y on an nc^3 node cube spanning the model
    const float *Q;
    int nc;
    float ux, uy, uz;    // (nc - 1) / extent per axis
    float inv_vref, dmin;
    struct Blk {};
    __device__ Blk begin(int, int, int, int, int, float, float, float, float *) const { return {}; }
    __device__ Blk direct(int) const { return {}; }
    __device__ float operator()(const Blk &, int s, int, int, int, float px, float py, float pz,
                                float dist) const {
        const float u = px * ux, v = py * uy, w = pz * uz;
        const int i0 = min((int)u, nc - 2), j0 = min((int)v, nc - 2), k0 = min((int)w, nc - 2);
        const float fu = u - i0, fv = v - j0, fw = w - k0;
        const float *q = Q + (long long)s * nc * nc * nc;
        auto at = [&](int i, int j, int k) { return q[i + nc * (j + nc * k)]; };
        const float c00 = at(i0, j0, k0) * (1 - fu) + at(i0 + 1, j0, k0) * fu;
        const float c10 = at(i0, j0 + 1, k0) * (1 - fu) + at(i0 + 1, j0 + 1, k0) * fu;
        const float c01 = at(i0, j0, k0 + 1) * (1 - fu) + at(i0 + 1, j0, k0 + 1) * fu;
        const float c11 = at(i0, j0 + 1, k0 + 1) * (1 - fu) + at(i0 + 1, j0 + 1, k0 + 1) * fu;
        const float c0 = c00 * (1 - fv) + c10 * fv, c1 = c01 * (1 - fv) + c11 * fv;
        return (c0 * (1 - fw) + c1 * fw) * fmaxf(dist, dmin) * inv_vref;
    }
};

__device__ __forceinline__ float wave_min(float v) { return wave_min_all(v); }

template <class Pred>
__global__ __launch_bounds__(64) void warm_fill(float *__restrict__ T, float *__restrict__ keys,
                                                const float *__restrict__ src, int B, int nbx,
                                                int nby, long long nblk, long long vol, int nx,
                                                int ny, int nz, float dx, float dy, float dz,
                                                float scale, float r_cold, int write_values,
                                                int nshot, float *__restrict__ xf, Pred pred) {
    // one (shot, block) per workgroup and step of the grid stride
    for (long long w = blockIdx.x; w < (long long)nshot * nblk; w += gridDim.x) {
    const int s = (int)(w / nblk);
    const long long b = w - (long long)s * nblk;
    const int lane = threadIdx.x;
    const int B3 = B * B * B;
    const int bx = (int)(b % nbx);
    const long long rest = b / nbx;
    const int by = (int)(rest % nby), bz = (int)(rest / nby);
    const float sx = src[3 * s], sy = src[3 * s + 1], sz = src[3 * s + 2];
    const auto blk = pred.begin(s, bx * B, by * B, bz * B, lane, dx, dy, dz, nullptr);
    float kmin = kInf;
    for (int t = lane; t < B3; t += 64) {
        const int tx = t % B, ty = (t / B) % B, tz = t / (B * B);
        const int x = bx * B + tx, y = by * B + ty, z = bz * B + tz;
        float v = kInf;
        if (x < nx && y < ny && z < nz) {
            const float px = x * dx, py = y * dy, pz = z * dz;
            const float ex = px - sx, ey = py - sy, ez = pz - sz;
            const float dist = sqrtf(ex * ex + ey * ey + ez * ez);
            float p = pred(blk, s, x, y, z, px, py, pz, dist);
            // a non-finite prediction (exponent all ones) counts as unreached;
            // tested on the bits, which survive no-NaN code generation
            if ((__float_as_uint(p) & 0x7F800000u) == 0x7F800000u) p = kInf;
            kmin = fminf(kmin, p);
            if (dist >= r_cold) v = fminf(scale * p, kInf);
        }
        if (write_values) {
            T[s * vol + b * B3 + t] = v;
            if (xf && B == 8) {  // the block's x-face copy
                const int tx = t & 7;
                if (tx == 0 || tx == 7)
                    xf[(s * nblk + b) * 128 + (tx == 7 ? 64 : 0) + ((t >> 3) & 7) + 8 * (t >> 6)] = v;
            }
        }
    }
    kmin = wave_min(kmin);
    // no usable prediction in the block: schedule it as early as possible
    if (lane == 0) keys[s * nblk + b] = kmin < kInf ? fmaxf(kmin, 0.0f) : 0.0f;
    }
}

// warm_fill for the network's node cube on 8^3 blocks whose cells span under
// three node spacings along z.  A lane's eight cells form one z column: the
// x / y interpolation (the c00..c11 and c0 / c1 terms of CoarsePred, the same
// expressions) is made once per column at the block's five lowest node
// levels, and each cell only interpolates between two of them in z.
__global__ __launch_bounds__(64) void warm_fill_coarse8(float *__restrict__ T,
                                                        float *__restrict__ keys,
                                                        const float *__restrict__ src, int nbx,
                                                        int nby, long long nblk, long long vol,
                                                        int nx, int ny, int nz, float dx, float dy,
                                                        float dz, float scale, float r_cold,
                                                        int write_values, int nshot,
                                                        float *__restrict__ xf, CoarsePred cp) {
    const int nc = cp.nc;
    for (long long w = blockIdx.x; w < (long long)nshot * nblk; w += gridDim.x) {
    const int s = (int)(w / nblk);
    const long long b = w - (long long)s * nblk;
    const int lane = threadIdx.x, tx = lane & 7, ty = lane >> 3;
    const int bx = (int)(b % nbx);
    const long long rest = b / nbx;
    const int by = (int)(rest % nby), bz = (int)(rest / nby);
    const int x = bx * 8 + tx, y = by * 8 + ty;
    const float sx = src[3 * s], sy = src[3 * s + 1], sz = src[3 * s + 2];
    const float *q = cp.Q + (long long)s * nc * nc * nc;
    const float px = x * dx, py = y * dy;
    const float u = px * cp.ux, v = py * cp.uy;
    const int i0 = min((int)u, nc - 2), j0 = min((int)v, nc - 2);
    const float fu = u - i0, fv = v - j0;
    // lowest node level of the block's cells (w grows with z)
    const int kb = min((int)((bz * 8 * dz) * cp.uz), nc - 2);
    auto at = [&](int i, int j, int k) { return q[i + nc * (j + nc * k)]; };
    float cl[5];
#pragma unroll
    for (int l = 0; l < 5; l++) {
        const int k = min(kb + l, nc - 1);
        const float c0 = at(i0, j0, k) * (1 - fu) + at(i0 + 1, j0, k) * fu;
        const float c1 = at(i0, j0 + 1, k) * (1 - fu) + at(i0 + 1, j0 + 1, k) * fu;
        cl[l] = c0 * (1 - fv) + c1 * fv;
    }
    const bool inxy = x < nx && y < ny;
    const float ex = px - sx, ey = py - sy;
    float kmin = kInf;
#pragma unroll
    for (int tz = 0; tz < 8; tz++) {
        const int z = bz * 8 + tz;
        float val = kInf;
        if (inxy && z < nz) {
            const float pz = z * dz;
            const float ez = pz - sz;
            const float dist = sqrtf(ex * ex + ey * ey + ez * ez);
            const float ww = pz * cp.uz;
            const int k0 = min((int)ww, nc - 2);
            const float fw = ww - k0;
            const int l = k0 - kb;  // 0 .. 3
            const float c0 = l == 0 ? cl[0] : l == 1 ? cl[1] : l == 2 ? cl[2] : cl[3];
            const float c1 = l == 0 ? cl[1] : l == 1 ? cl[2] : l == 2 ? cl[3] : cl[4];
            float p = (c0 * (1 - fw) + c1 * fw) * fmaxf(dist, cp.dmin) * cp.inv_vref;
            if ((__float_as_uint(p) & 0x7F800000u) == 0x7F800000u) p = kInf;
            kmin = fminf(kmin, p);
            if (dist >= r_cold) val = fminf(scale * p, kInf);
        }
        if (write_values) {
            T[s * vol + b * 512 + lane + 64 * tz] = val;
            if (xf && (tx == 0 || tx == 7))
                xf[(s * nblk + b) * 128 + (tx == 7 ? 64 : 0) + ty + 8 * tz] = val;
        }
    }
    kmin = wave_min(kmin);
    if (lane == 0) keys[s * nblk + b] = kmin < kInf ? fmaxf(kmin, 0.0f) : 0.0f;
    }
}

// Keys only (ordering-only warm starts): one thread per (shot, block), the
// block's predicted first arrival taken as the smallest prediction on the
// 3 x 3 x 3 lattice of its corners, edge and face midpoints and centre
// (clipped to the model).  An ordering key needs no more: the schedule stays
// exact whatever the keys are.
template <class Pred>
__global__ void warm_keys(float *__restrict__ keys, const float *__restrict__ src, int nshot, int B,
                          int nbx, int nby, long long nblk, int nx, int ny, int nz, float dx,
                          float dy, float dz, Pred pred) {
    for (long long w = (long long)blockIdx.x * blockDim.x + threadIdx.x; w < (long long)nshot * nblk;
         w += (long long)gridDim.x * blockDim.x) {
    const int s = (int)(w / nblk);
    const long long b = w - (long long)s * nblk;
    const int bx = (int)(b % nbx);
    const long long rest = b / nbx;
    const int by = (int)(rest % nby), bz = (int)(rest / nby);
    const int x0 = bx * B, y0 = by * B, z0 = bz * B;
    const int x1 = min(x0 + B - 1, nx - 1), y1 = min(y0 + B - 1, ny - 1), z1 = min(z0 + B - 1, nz - 1);
    const float sx = src[3 * s], sy = src[3 * s + 1], sz = src[3 * s + 2];
    float kmin = kInf;
    for (int k = 0; k < 3; k++) {
        const int z = k == 0 ? z0 : (k == 1 ? (z0 + z1) / 2 : z1);
        for (int j = 0; j < 3; j++) {
            const int y = j == 0 ? y0 : (j == 1 ? (y0 + y1) / 2 : y1);
            for (int i = 0; i < 3; i++) {
                const int x = i == 0 ? x0 : (i == 1 ? (x0 + x1) / 2 : x1);
                const float px = x * dx, py = y * dy, pz = z * dz;
                const float ex = px - sx, ey = py - sy, ez = pz - sz;
                float p = pred(pred.direct(s), s, x, y, z, px, py, pz,
                               sqrtf(ex * ex + ey * ey + ez * ez));
                if ((__float_as_uint(p) & 0x7F800000u) == 0x7F800000u) p = kInf;
                kmin = fminf(kmin, p);
            }
        }
    }
    keys[w] = kmin < kInf ? fmaxf(kmin, 0.0f) : 0.0f;
    }
}

// Judge: largest residual G(T) - T over the domain cells of each shot
// (source node excluded).  A converged solve from an upper bound has none:
// every cell holds the update of its final neighbours.  A positive residual
// marks cells whose warm value was below what their neighbours support.
// One thread per padded cell; vol is a multiple of 64 (B^3 cells per block),
// so every wave lies in one shot and a wave-level max leaves one atomic per
// wave instead of one per cell.
__global__ void judge_residual(const float *__restrict__ T, const float *__restrict__ S,
                               const float *__restrict__ src, int nshot, int B, int nbx, int nby,
                               int n4x, int n4y, int n4z, long long vol, float dx, float dy,
                               float dz, unsigned *__restrict__ resmax) {
    const long long n = (long long)nshot * vol;
    // the launch covers whole waves: every step of the stride keeps them whole
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i - (threadIdx.x & 63) < n;
         i += (long long)gridDim.x * blockDim.x) {
    const long long iw = i < n ? i : n - 1;  // tail lanes repeat the last cell
    const int s = (int)(iw / vol);
    const long long a = iw - (long long)s * vol;
    const int B3 = B * B * B;
    const long long b = a / B3;
    const int t = (int)(a - b * B3);
    const int x = (int)(b % nbx) * B + t % B;
    const long long rest = b / nbx;
    const int y = (int)(rest % nby) * B + (t / B) % B;
    const int z = (int)(rest / nby) * B + t / (B * B);
    const int ix = (int)(src[3 * s] / dx), iy = (int)(src[3 * s + 1] / dy), iz = (int)(src[3 * s + 2] / dz);
    float r = 0.0f;
    if (x < n4x && y < n4y && z < n4z && !(x == ix && y == iy && z == iz)) {
        const float *Ts = T + (long long)s * vol;
        auto at = [&](int xx, int yy, int zz) -> float {
            if (xx < 0 || yy < 0 || zz < 0 || xx >= n4x || yy >= n4y || zz >= n4z) return kInf;
            return Ts[blk_addr(xx, yy, zz, B, nbx, nby)];
        };
        const float v = Ts[a];
        const float hs = dx * S[a];
        const float g = godunov(fminf(at(x - 1, y, z), at(x + 1, y, z)),
                                fminf(at(x, y - 1, z), at(x, y + 1, z)),
                                fminf(at(x, y, z - 1), at(x, y, z + 1)), hs, hs * hs);
        r = fmaxf(g - v, 0.0f);
    }
    r = wave_max_all(r);
    if ((threadIdx.x & 63) == 0 && r > 0.0f) atomicMax(&resmax[s], __float_as_uint(r));
    }
}
// The same judge for 8^3 blocks, one wave per (shot, block): the block and
// its six face halos are staged as a visit stages them (z column per lane,
// x / y neighbours through an LDS tile), so each cell's neighbours come
// from registers or LDS instead of six address computations and scattered
// loads per cell.
__global__ __launch_bounds__(64 * kWaves) void judge_blocks8(
    const float *__restrict__ T, const float *__restrict__ S, const float *__restrict__ src,
    int nshot, int nbx, int nby, int nbz, long long nblk, int n4x, int n4y, int n4z,
    long long vol, float dx, float dy, float dz, const float *__restrict__ xf,
    unsigned *__restrict__ resmax) {
    __shared__ float tiles[kWaves][kTile8];  // the FIM kernel's tile layout
    const int lane = threadIdx.x & 63;
    float *tile = tiles[threadIdx.x >> 6];
    // block-major: consecutive items are one block of consecutive shots, so
    // the block's slowness is read from L2 after the first
    for (long long item = (long long)blockIdx.x * kWaves + (threadIdx.x >> 6);
         item < (long long)nshot * nblk; item += (long long)gridDim.x * kWaves) {  // wave-uniform
    const long long b = item / nshot;
    const int shot = (int)(item - b * nshot);
    const int bx = (int)(b % nbx);
    const long long rest = b / nbx;
    const int by = (int)(rest % nby), bz = (int)(rest / nby);
    const float *Ts = T + (long long)shot * vol;
    const long long base = b * 512, nbxy = (long long)nbx * nby;
    const int tx = lane & 7, ty = lane >> 3;
    const int tcol = tix(tx + 1, ty + 1, 0);
    float v[10];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        v[k + 1] = Ts[base + lane + 64 * k];
        tile[tcol + kTZ * (k + 1)] = v[k + 1];
    }
    v[0] = (bz > 0) ? Ts[base - 512 * nbxy + lane + 448] : kInf;
    v[9] = (bz < nbz - 1) ? Ts[base + 512 * nbxy + lane] : kInf;
    for (int hid = lane; hid < 256; hid += 64) {
        const int f = hid >> 6, k = hid & 63, u = k & 7, w = k >> 3;
        int ti;
        float hv = kInf;
        switch (f) {
        case 0: ti = tix(0, u + 1, w + 1);
                if (bx > 0)
                    hv = xf ? xf[((long long)shot * nblk + b - 1) * 128 + 64 + k]
                            : Ts[base - 512 + 7 + 8 * u + 64 * w];
                break;
        case 1: ti = tix(9, u + 1, w + 1);
                if (bx < nbx - 1)
                    hv = xf ? xf[((long long)shot * nblk + b + 1) * 128 + k]
                            : Ts[base + 512 + 8 * u + 64 * w];
                break;
        case 2: ti = tix(u + 1, 0, w + 1);
                if (by > 0) hv = Ts[base - 512LL * nbx + u + 56 + 64 * w]; break;
        default: ti = tix(u + 1, 9, w + 1);
                if (by < nby - 1) hv = Ts[base + 512LL * nbx + u + 64 * w]; break;
        }
        tile[ti] = hv;
    }
    wave_sync();
    const int ix = (int)(src[3 * shot] / dx), iy = (int)(src[3 * shot + 1] / dy),
              iz = (int)(src[3 * shot + 2] / dz);
    const int x = bx * 8 + tx, y = by * 8 + ty;
    float r = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int z = bz * 8 + k;
        if (x < n4x && y < n4y && z < n4z && !(x == ix && y == iy && z == iz)) {
            const int ti = tcol + kTZ * (k + 1);
            const float a = fminf(tile[ti - kTX], tile[ti + kTX]);
            const float bb = fminf(tile[ti - kTY], tile[ti + kTY]);
            const float c = fminf(v[k], v[k + 2]);
            const float hs = dx * S[base + lane + 64 * k];
            r = fmaxf(r, godunov(a, bb, c, hs, hs * hs) - v[k + 1]);
        }
    }
    r = wave_max_all(r);
    if (lane == 0 && r > 0.0f) atomicMax(&resmax[shot], __float_as_uint(r));
    wave_sync();  // the tile is restaged for the next item
    }
}
}  // namespace st

// Sums the per-wave statistics slots into out[0..2] (one workgroup of 1,024
// threads, four slots per thread in flight: the sum is latency-bound).
__global__ __launch_bounds__(1024) void sum_wave_stats(const unsigned long long *__restrict__ w,
                                                       int nw, unsigned long long *__restrict__ out) {
    __shared__ unsigned long long part[3][1024];
    unsigned long long a = 0, b = 0, c = 0;
    const int t = threadIdx.x;
    for (int i = t; i < nw; i += 4 * 1024) {
        unsigned long long v[4][3];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int j = i + u * 1024;
#pragma unroll
            for (int k = 0; k < 3; k++) v[u][k] = j < nw ? w[3LL * j + k] : 0ull;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            a += v[u][0];
            b += v[u][1];
            c += v[u][2];
        }
    }
    part[0][t] = a;
    part[1][t] = b;
    part[2][t] = c;
    __syncthreads();
    for (int h = 512; h > 0; h >>= 1) {
        if (t < h)
            for (int k = 0; k < 3; k++) part[k][t] += part[k][t + h];
        __syncthreads();
    }
    if (t < 3) out[t] = part[t][0];
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace st {

static int blk_index(int B) { return B == 8 ? 1 : 0; }

// The model's slowness in the block layout of edge B, into S (vol floats).
void pack_block_model(st_ctx *ctx, int B, float *S) {
    const BlockGeo g = make_geo(ctx->nx, ctx->ny, ctx->nz, B);
    const int tpb = 256;
    hipLaunchKernelGGL(pack_model, dim3((unsigned)((g.vol + tpb - 1) / tpb)), dim3(tpb), 0,
                       ctx->stream, ctx->slow_zxy.as<float>(), S, ctx->nx, ctx->ny, ctx->nz, B,
                       g.nbx, g.nby, g.vol);
    ST_HIP(hipGetLastError());
}

void ensure_block_model(st_ctx *ctx, int B) {
    const int bi = blk_index(B);
    if (ctx->blk_valid[bi]) return;
    const BlockGeo g = make_geo(ctx->nx, ctx->ny, ctx->nz, B);
    ctx->slow_blk[bi].reserve(sizeof(float) * g.vol);
    pack_block_model(ctx, B, ctx->slow_blk[bi].as<float>());
    ctx->blk_valid[bi] = true;
}

const float *block_model(st_ctx *ctx, int B) { return ctx->slow_blk[blk_index(B)].as<float>(); }

void fim_batch_on(st_ctx *ctx, const FimBufs &bf, const float *dsrc, int nshot,
                  const st_fim_opts &o, st_fim_stats *stats, bool warm_loaded, const float *keys,
                  bool tentative);

int fim_waves(const st_fim_opts &o) {
    // o.grid counts waves; a launch has o.grid / kWaves workgroups.  By
    // default 192 waves per CU: more than are resident (24 per CU), so a
    // CU's slots work through several waves' shares of a round's list,
    // which evens out the round's tail (a 401x401x201 forward of 1,024 shots:
    // 0.34 s at 32 waves per CU, 0.28 s at 192; the launch is capped by the
    // batch's blocks, so small solves launch fewer)
    const int grid = o.grid > 0 ? o.grid : 256 * 192;
    return std::max(1, (grid + kWaves - 1) / kWaves) * kWaves;
}

// Device-resident FIM over one batch of shots already uploaded to dsrc, on
// the context's buffers.  Leaves the travel times in ctx->T (shot-major
// block layout).
void fim_batch(st_ctx *ctx, const float *dsrc, int nshot, const st_fim_opts &o,
               st_fim_stats *stats, bool warm_loaded, const float *keys, bool tentative) {
    const BlockGeo g = make_geo(ctx->nx, ctx->ny, ctx->nz, o.block);
    ctx->list[0].reserve(fim_list_bytes(g.nblk, nshot));
    ctx->list[1].reserve(fim_list_bytes(g.nblk, nshot));
    if (keys || tentative) ctx->kmin.reserve(sizeof(unsigned) * 3 * std::max(nshot, 1));
    if (tentative) ctx->tkeys.reserve(sizeof(unsigned) * 2 * nshot * g.nblk);
    if (stats) ctx->wstats.reserve(sizeof(unsigned long long) * 3 * fim_waves(o));
    if (o.block == 8) ctx->xfaces.reserve(sizeof(float) * 128 * g.nblk * nshot);
    FimBufs bf;
    bf.S = block_model(ctx, o.block);
    bf.T = ctx->T.as<float>();
    bf.flags = ctx->flags.as<unsigned>();
    bf.list0 = ctx->list[0].as<unsigned>();
    bf.list1 = ctx->list[1].as<unsigned>();
    bf.kmin = ctx->kmin.as<unsigned>();
    bf.tk = ctx->tkeys.as<unsigned>();
    bf.wstats = ctx->wstats.as<unsigned long long>();
    bf.xf = o.block == 8 ? ctx->xfaces.as<float>() : nullptr;
    bf.graphs = true;
    fim_batch_on(ctx, bf, dsrc, nshot, o, stats, warm_loaded, keys, tentative);
}

// The same on caller-provided device buffers (sized as fim_batch reserves
// them); graph replay only when bf.graphs.
void fim_batch_on(st_ctx *ctx, const FimBufs &bf, const float *dsrc, int nshot,
                  const st_fim_opts &o, st_fim_stats *stats, bool warm_loaded, const float *keys,
                  bool tentative) {
    const int B = o.block;
    const BlockGeo g = make_geo(ctx->nx, ctx->ny, ctx->nz, B);
    unsigned *cnt = ctx->counters.as<unsigned>();
    unsigned long long *dstats = reinterpret_cast<unsigned long long *>(cnt + kStatsIdx);
    const int tpb = 256;

    unsigned *tk = nullptr;
    if (tentative) {
        const long long n = 2LL * nshot * g.nblk;
        tk = bf.tk;
        hipLaunchKernelGGL(fill_u32, dim3(stride_grid(n, tpb)), dim3(tpb), 0, ctx->stream, tk, n,
                           kInfBits);
    }
    float *xf = B == 8 ? bf.xf : nullptr;
    // cold 8^3 starts skip the volume fills (stamped blocks, see FimArgs);
    // graph replays, whose launches cannot carry a launch index, fill
    const bool lazy = B == 8 && !warm_loaded && !(o.use_graph && bf.graphs);
    if (!warm_loaded && !lazy) {
        const long long n4 = (long long)nshot * g.vol / 4;
        hipLaunchKernelGGL(fill_inf, dim3(stride_grid(n4, tpb)), dim3(tpb), 0, ctx->stream,
                           reinterpret_cast<float4 *>(bf.T), n4);
        if (xf) {  // the face copies of cold blocks (a warm fill writes its own)
            const long long f4 = (long long)nshot * g.nblk * 32;
            hipLaunchKernelGGL(fill_inf, dim3(stride_grid(f4, tpb)), dim3(tpb), 0, ctx->stream,
                               reinterpret_cast<float4 *>(xf), f4);
        }
    }
    ST_HIP(hipMemsetAsync(bf.flags, 0, sizeof(unsigned) * g.nblk * nshot, ctx->stream));
    hipLaunchKernelGGL(seed_sources, dim3((nshot + 63) / 64), dim3(64), 0, ctx->stream, dsrc,
                       nshot, bf.S, bf.T, bf.flags, bf.list0, cnt, ctx->dx, ctx->dy, ctx->dz, B,
                       g.nbx, g.nby, g.vol, g.nblk, keys, bf.kmin, tk, xf, lazy ? 1 : 0);
    ST_HIP(hipGetLastError());

    FimArgs A;
    A.S = bf.S;
    A.T = bf.T;
    A.flags = bf.flags;
    A.cap = g.nblk * (long long)nshot;
    A.list0 = bf.list0;
    A.list1 = bf.list1;
    A.cnt = cnt;
    A.stats = nullptr;
    A.nbx = g.nbx; A.nby = g.nby; A.nbz = g.nbz;
    A.n4x = g.n4x; A.n4y = g.n4y; A.n4z = g.n4z;
    A.vol = g.vol; A.nblk = g.nblk;
    A.h = ctx->dx;
    A.eps = o.eps;
    A.max_inner = o.max_inner;
    A.keys = keys;
    A.kmin = bf.kmin;
    A.nshot = nshot;
    A.band = o.band;
    A.src = dsrc;
    A.tk = tk;
    A.xf = xf;
    A.lazy = lazy ? 1 : 0;
    A.epoch = 0;

    // no more waves than an eighth of the batch's blocks (at least 1,024): a
    // small batch, such as the cold re-solve of rejected warm starts, keeps
    // its rounds from launching waves that find no item
    const long long wcap = std::max<long long>(1024, (long long)nshot * g.nblk / 8);
    const int nwaves = (int)std::min<long long>(fim_waves(o), (wcap + kWaves - 1) / kWaves * kWaves);
    const int grid = nwaves / kWaves;
    if (stats) {
        ST_HIP(hipMemsetAsync(bf.wstats, 0, sizeof(unsigned long long) * 3 * nwaves, ctx->stream));
        A.stats = bf.wstats;
    }
    const int K = o.rounds_per_sync > 0 ? o.rounds_per_sync : 6;
    unsigned *hcnt = static_cast<unsigned *>(ctx->host_pinned);
    // two snapshots of a slot's segment counters (polled one chunk behind)
    unsigned *hseg[2] = {static_cast<unsigned *>(ctx->host_pinned) + 2048,
                         static_cast<unsigned *>(ctx->host_pinned) + 2048 + kSeg * kCntStride};
    auto seg_total = [&](const unsigned *h) {
        unsigned long long t = 0;
        for (int k = 0; k < kSeg; k++) t += h[k * kCntStride];
        return t;
    };
    const long long max_rounds = 4LL * (g.nbx + g.nby + g.nbz) * 64 + 4096;
    const bool hyp = o.sweep != 1;
    auto launch_round = [&](const FimArgs &a) {
        if (B == 8 && hyp)
            hipLaunchKernelGGL(fim_round8<true>, dim3(grid), dim3(64 * kWaves), 0, ctx->stream, a);
        else if (B == 8)
            hipLaunchKernelGGL(fim_round8<false>, dim3(grid), dim3(64 * kWaves), 0, ctx->stream, a);
        else
            hipLaunchKernelGGL(fim_round4, dim3(grid), dim3(64 * kWaves), 0, ctx->stream, a);
    };

    ST_HIP(hipEventRecord(ctx->ev[2], ctx->stream));
    long long r = 0;
    int slot = 0, pending = -1;
    // Graph replay: six rounds (one period of the list parity and counter
    // rotation) captured once per launch configuration and replayed; the
    // round kernels clear their own counters, so the graph holds kernel nodes
    // only.
    hipGraphExec_t exec = nullptr;
    if (o.use_graph && bf.graphs) {
        long long key = 1469598103934665603LL;
        auto mix = [&](long long v) { key = (key ^ v) * 1099511628211LL; };
        mix((long long)(uintptr_t)A.T); mix((long long)(uintptr_t)A.S); mix((long long)(uintptr_t)A.keys);
        mix((long long)(uintptr_t)A.flags); mix((long long)(uintptr_t)A.list0);
        mix((long long)(uintptr_t)A.list1); mix((long long)(uintptr_t)A.kmin);
        mix((long long)(uintptr_t)A.stats);
        mix(nshot); mix(B); mix(grid); mix(A.nblk); mix(A.max_inner);
        // every geometry field the kernels read: two models with the same
        // block count can differ in the pad-to-4 extent or the block grid
        mix(A.nbx); mix(A.nby); mix(A.nbz); mix(A.n4x); mix(A.n4y); mix(A.n4z);
        mix(A.vol); mix(A.cap); mix(A.nshot);
        mix((long long)__builtin_bit_cast(unsigned, A.band)); mix((long long)__builtin_bit_cast(unsigned, A.eps));
        mix((long long)__builtin_bit_cast(unsigned, A.h));
        mix((long long)(uintptr_t)A.src); mix(hyp ? 1 : 0); mix((long long)(uintptr_t)A.tk);
        mix((long long)(uintptr_t)A.xf); mix(A.lazy);
        if (!ctx->graph_exec || ctx->graph_key != key) {
            if (ctx->graph_exec) ST_HIP(hipGraphExecDestroy(ctx->graph_exec));
            if (ctx->graph) ST_HIP(hipGraphDestroy(ctx->graph));
            ctx->graph_exec = nullptr;
            ctx->graph = nullptr;
            ST_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            for (int k = 0; k < 6; k++) {
                A.round = k;
                launch_round(A);
            }
            ST_HIP(hipStreamEndCapture(ctx->stream, &ctx->graph));
            ST_HIP(hipGraphInstantiate(&ctx->graph_exec, ctx->graph, nullptr, nullptr, 0));
            ctx->graph_key = key;
        }
        exec = ctx->graph_exec;
    }
    const int Kg = exec ? 6 * std::max(1, K / 6) : K;
    for (;;) {
        if (exec) {
            for (int k = 0; k < Kg; k += 6, r += 6) ST_HIP(hipGraphLaunch(exec, ctx->stream));
        } else {
            for (int k = 0; k < Kg; k++, r++) {
                A.round = (int)(r % 6);
                A.epoch = (int)r;
                launch_round(A);
            }
        }
        ST_HIP(hipGetLastError());
        ST_HIP(hipMemcpyAsync(hseg[slot], &cnt[(r % 3) * kSeg * kCntStride],
                              sizeof(unsigned) * kSeg * kCntStride, hipMemcpyDeviceToHost,
                              ctx->stream));
        ST_HIP(hipEventRecord(ctx->ev[slot], ctx->stream));
        if (pending >= 0) {
            ST_HIP(hipEventSynchronize(ctx->ev[pending]));
            if (seg_total(hseg[pending]) == 0) break;
        }
        pending = slot;
        slot ^= 1;
        ST_REQUIRE(r < max_rounds, ST_ERR_HIP, "FIM did not terminate within the round cap");
    }
    ST_HIP(hipEventRecord(ctx->ev[3], ctx->stream));
    if (lazy) {
        const long long n = (long long)nshot * g.nblk;
        hipLaunchKernelGGL(fill_unwritten8, dim3(stride_grid(n, 256)), dim3(256), 0, ctx->stream,
                           bf.flags, bf.T, n, g.nblk, g.vol);
        ST_HIP(hipGetLastError());
    }
    if (stats) {
        hipLaunchKernelGGL(sum_wave_stats, dim3(1), dim3(1024), 0, ctx->stream, bf.wstats, nwaves,
                           dstats);
        ST_HIP(hipGetLastError());
        ST_HIP(hipMemcpyAsync(hcnt + 4, cnt + kRoundsIdx, sizeof(unsigned), hipMemcpyDeviceToHost,
                              ctx->stream));
        ST_HIP(hipMemcpyAsync(hcnt + 8, dstats, 3 * sizeof(unsigned long long),
                              hipMemcpyDeviceToHost, ctx->stream));
        ST_HIP(hipStreamSynchronize(ctx->stream));
        const unsigned long long *hs = reinterpret_cast<const unsigned long long *>(hcnt + 8);
        stats->rounds += hcnt[4];
        stats->launches += r;
        stats->visits += (long long)hs[0];
        stats->inner_iters += (long long)hs[1];
        stats->deferred += (long long)hs[2];
        stats->batches += 1;
        float ms = 0.f;
        ST_HIP(hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]));
        stats->solve_ms += ms;
    }
}

// The context's x-face copies for a warm fill of 8^3 blocks (null otherwise).
static float *warm_faces(st_ctx *ctx, const BlockGeo &g, int nshot, int B) {
    if (B != 8) return nullptr;
    ctx->xfaces.reserve(sizeof(float) * 128 * g.nblk * nshot);
    return ctx->xfaces.as<float>();
}

// Warm values + keys from caller volumes (device, nshot x npts, z fastest).
void warm_fill_fine(st_ctx *ctx, const float *dpred, const float *dsrc, int nshot, int B,
                    float delta, float r_cold, bool values, float *dkeys) {
    const BlockGeo g = make_geo(ctx->nx, ctx->ny, ctx->nz, B);
    FinePred fp{dpred, ctx->nx, ctx->nz, (long long)ctx->nx * ctx->ny * ctx->nz};
    float *xf = warm_faces(ctx, g, nshot, B);
    if (!values) {  // cold values, keys from a sampled lattice
        const long long n4 = (long long)nshot * g.vol / 4;
        hipLaunchKernelGGL(fill_inf, dim3(stride_grid(n4, 256)), dim3(256), 0, ctx->stream,
                           ctx->T.as<float4>(), n4);
        if (xf) {
            const long long f4 = (long long)nshot * g.nblk * 32;
            hipLaunchKernelGGL(fill_inf, dim3(stride_grid(f4, 256)), dim3(256), 0, ctx->stream,
                               reinterpret_cast<float4 *>(xf), f4);
        }
        const long long nk = (long long)nshot * g.nblk;
        hipLaunchKernelGGL(warm_keys<FinePred>, dim3(stride_grid(nk, 256)), dim3(256), 0,
                           ctx->stream, dkeys, dsrc, nshot, B, g.nbx, g.nby, g.nblk, ctx->nx,
                           ctx->ny, ctx->nz, ctx->dx, ctx->dy, ctx->dz, fp);
        ST_HIP(hipGetLastError());
        return;
    }
    hipLaunchKernelGGL(warm_fill<FinePred>, dim3(stride_grid(g.nblk * nshot, 1)), dim3(64), 0,
                       ctx->stream, ctx->T.as<float>(), dkeys, dsrc, B, g.nbx, g.nby, g.nblk,
                       g.vol, ctx->nx, ctx->ny, ctx->nz, ctx->dx, ctx->dy, ctx->dz, 1.0f + delta,
                       r_cold, values ? 1 : 0, nshot, xf, fp);
    ST_HIP(hipGetLastError());
}

// Warm values + keys from the network's node-cube output (device,
// nshot x nc^3): T = y * max(D, dmin) / vref.
void warm_fill_coarse(st_ctx *ctx, const float *dQ, int nc, float vref, float dmin,
                      const float *dsrc, int nshot, int B, float delta, float r_cold, bool values,
                      float *dkeys) {
    const BlockGeo g = make_geo(ctx->nx, ctx->ny, ctx->nz, B);
    CoarsePred cp{dQ, nc, (nc - 1) / ((ctx->nx - 1) * ctx->dx), (nc - 1) / ((ctx->ny - 1) * ctx->dy),
                  (nc - 1) / ((ctx->nz - 1) * ctx->dz), 1.0f / vref, dmin};
    // a block's cells span 7 spacings along z: under three node spacings,
    // their lower node levels take at most four values (five levels with the
    // upper ones)
    const bool columns = B == 8 && 7.0f * ctx->dz * cp.uz < 2.99f;
    float *xf = warm_faces(ctx, g, nshot, B);
    if (!values) {  // cold values, keys from a sampled lattice
        const long long n4 = (long long)nshot * g.vol / 4;
        hipLaunchKernelGGL(fill_inf, dim3(stride_grid(n4, 256)), dim3(256), 0, ctx->stream,
                           ctx->T.as<float4>(), n4);
        if (xf) {
            const long long f4 = (long long)nshot * g.nblk * 32;
            hipLaunchKernelGGL(fill_inf, dim3(stride_grid(f4, 256)), dim3(256), 0, ctx->stream,
                               reinterpret_cast<float4 *>(xf), f4);
        }
        const long long nk = (long long)nshot * g.nblk;
        hipLaunchKernelGGL(warm_keys<CoarsePred>, dim3(stride_grid(nk, 256)), dim3(256), 0,
                           ctx->stream, dkeys, dsrc, nshot, B, g.nbx, g.nby, g.nblk, ctx->nx,
                           ctx->ny, ctx->nz, ctx->dx, ctx->dy, ctx->dz, cp);
        ST_HIP(hipGetLastError());
        return;
    }
    if (columns) {
        hipLaunchKernelGGL(warm_fill_coarse8, dim3(stride_grid(g.nblk * nshot, 1)), dim3(64), 0,
                           ctx->stream, ctx->T.as<float>(), dkeys, dsrc, g.nbx, g.nby, g.nblk,
                           g.vol, ctx->nx, ctx->ny, ctx->nz, ctx->dx, ctx->dy, ctx->dz,
                           1.0f + delta, r_cold, 1, nshot, xf, cp);
        ST_HIP(hipGetLastError());
        return;
    }
    hipLaunchKernelGGL(warm_fill<CoarsePred>, dim3(stride_grid(g.nblk * nshot, 1)), dim3(64), 0,
                       ctx->stream, ctx->T.as<float>(), dkeys, dsrc, B, g.nbx, g.nby, g.nblk,
                       g.vol, ctx->nx, ctx->ny, ctx->nz, ctx->dx, ctx->dy, ctx->dz, 1.0f + delta,
                       r_cold, values ? 1 : 0, nshot, xf, cp);
    ST_HIP(hipGetLastError());
}

// Per-shot judge residuals (host array of nshot floats).
void judge_batch(st_ctx *ctx, const float *dsrc, int nshot, int B, float *res) {
    const BlockGeo g = make_geo(ctx->nx, ctx->ny, ctx->nz, B);
    ctx->judge.reserve(sizeof(unsigned) * std::max(nshot, 1));
    ST_HIP(hipMemsetAsync(ctx->judge.ptr, 0, sizeof(unsigned) * nshot, ctx->stream));
    if (B == 8) {
        const long long items = (long long)nshot * g.nblk;
        // the solve leaves the x-face copies equal to the volumes' faces
        const float *xf = ctx->xfaces.bytes >= sizeof(float) * 128 * (size_t)items
                              ? ctx->xfaces.as<float>() : nullptr;
        hipLaunchKernelGGL(judge_blocks8, dim3(stride_grid(items, kWaves)),
                           dim3(64 * kWaves), 0, ctx->stream, ctx->T.as<float>(), block_model(ctx, B),
                           dsrc, nshot, g.nbx, g.nby, g.nbz, g.nblk, g.n4x, g.n4y, g.n4z, g.vol,
                           ctx->dx, ctx->dy, ctx->dz, xf, ctx->judge.as<unsigned>());
        ST_HIP(hipGetLastError());
        ST_HIP(hipMemcpyAsync(res, ctx->judge.ptr, sizeof(float) * nshot, hipMemcpyDeviceToHost,
                              ctx->stream));
        ST_HIP(hipStreamSynchronize(ctx->stream));
        return;
    }
    const long long n = (long long)nshot * g.vol;
    hipLaunchKernelGGL(judge_residual, dim3(stride_grid(n, 256)), dim3(256), 0, ctx->stream,
                       ctx->T.as<float>(), block_model(ctx, B), dsrc, nshot, B, g.nbx, g.nby, g.n4x,
                       g.n4y, g.n4z, g.vol, ctx->dx, ctx->dy, ctx->dz, ctx->judge.as<unsigned>());
    ST_HIP(hipGetLastError());
    ST_HIP(hipMemcpyAsync(res, ctx->judge.ptr, sizeof(float) * nshot, hipMemcpyDeviceToHost,
                          ctx->stream));
    ST_HIP(hipStreamSynchronize(ctx->stream));
}

void launch_extract(st_ctx *ctx, float *dout, int nshot, int B) {
    const BlockGeo g = make_geo(ctx->nx, ctx->ny, ctx->nz, B);
    const long long n = (long long)ctx->nx * ctx->ny * ctx->nz * nshot;
    hipLaunchKernelGGL(extract_volume, dim3(stride_grid(n, 256)), dim3(256), 0, ctx->stream,
                       ctx->T.as<float>(), dout, ctx->nx, ctx->ny, ctx->nz, B, g.nbx, g.nby, g.vol,
                       nshot);
    ST_HIP(hipGetLastError());
}

}  // namespace st
