// nn.hip -- hand-written training and inference kernels of the warm-start
// network: BF16 MFMA GEMM with FP32 accumulation, 3D im2col / col2im,
// average pooling, nearest upsampling, channel concat, bias/ReLU, MSE loss
// and Adam.  Activations are channels-last [sample][z][y][x][c], voxel
// v = x + n*(y + n*z) on an n^3 cube.
#include <type_traits>

#include "nn.hpp"

namespace st {
namespace nn {

// ---------------------------------------------------------------------------
// GEMM  C[M,N] = A[M,K] * B[K,N]  (BF16 operands, FP32 accumulate)
//   ta = 0: A(m,k) = A[m*lda + k]     ta = 1: A(m,k) = A[k*lda + m]
//   tb = 0: B(k,n) = B[k*ldb + n]     tb = 1: B(k,n) = B[n*ldb + k]
//   mode 0: C = act(acc + bias) (and a BF16 copy when C16 is set)
//   mode 1: C += acc with atomics (split-K over blockIdx.z)
// 256 threads = 4 waves; 64x64 tile per workgroup, 32x32 per wave as 2x2
// v_mfma_f32_16x16x16_bf16 tiles; K staged through LDS in steps of 32.
// ---------------------------------------------------------------------------
typedef short s4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef __bf16 b8v __attribute__((ext_vector_type(8)));

__global__ __launch_bounds__(256) void gemm_bf16(GemmArgs g) {
    __shared__ unsigned short As[64][40];  // [m][k]
    __shared__ unsigned short Bs[64][40];  // [n][k]
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const long long m0 = (long long)blockIdx.y * 64;
    const int n0 = blockIdx.x * 64;
    const int kchunk = (((g.K + g.ksplit - 1) / g.ksplit) + 31) / 32 * 32;
    const int kbeg = blockIdx.z * kchunk;
    const int kend = min(g.K, kbeg + kchunk);
    f4v acc[2][2];
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) acc[i][j] = f4v{0.f, 0.f, 0.f, 0.f};
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    for (int k0 = kbeg; k0 < kend; k0 += 32) {
        for (int e = tid; e < 2048; e += 256) {
            int mm, kk;
            if (!g.ta) { mm = e >> 5; kk = e & 31; } else { kk = e >> 6; mm = e & 63; }
            const long long m = m0 + mm;
            const int k = k0 + kk;
            unsigned short v = 0;
            if (m < g.M && k < kend)
                v = g.ta ? g.A[(long long)k * g.lda + m] : g.A[m * g.lda + k];
            As[mm][kk] = v;
        }
        for (int e = tid; e < 2048; e += 256) {
            int nn, kk;
            if (!g.tb) { kk = e >> 6; nn = e & 63; } else { nn = e >> 5; kk = e & 31; }
            const int n = n0 + nn, k = k0 + kk;
            unsigned short v = 0;
            if (n < g.N && k < kend)
                v = g.tb ? g.B[(long long)n * g.ldb + k] : g.B[(long long)k * g.ldb + n];
            Bs[nn][kk] = v;
        }
        __syncthreads();
        for (int ks = 0; ks < 32; ks += 16) {
            const int kq = ks + 4 * (lane >> 4);
            s4v a[2], b[2];
            for (int i = 0; i < 2; i++) {
                const int r = wm + i * 16 + (lane & 15);
                a[i] = s4v{(short)As[r][kq], (short)As[r][kq + 1], (short)As[r][kq + 2],
                            (short)As[r][kq + 3]};
            }
            for (int j = 0; j < 2; j++) {
                const int c = wn + j * 16 + (lane & 15);
                b[j] = s4v{(short)Bs[c][kq], (short)Bs[c][kq + 1], (short)Bs[c][kq + 2],
                           (short)Bs[c][kq + 3]};
            }
            for (int i = 0; i < 2; i++)
                for (int j = 0; j < 2; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
            for (int r = 0; r < 4; r++) {
                const long long m = m0 + wm + i * 16 + 4 * (lane >> 4) + r;
                const int n = n0 + wn + j * 16 + (lane & 15);
                if (m >= g.M || n >= g.N) continue;
                float v = acc[i][j][r];
                if (g.mode == 1) {
                    atomicAdd(&g.C[m * g.ldc + n], v);
                } else {
                    if (g.bias) v += g.bias[n];
                    if (g.relu) v = fmaxf(v, 0.f);
                    if (g.C) g.C[m * g.ldc + n] = v;
                    if (g.C16) g.C16[m * g.ldc + n] = f2bf(v);
                }
            }
}

void gemm(hipStream_t st, const GemmArgs &g) {
    dim3 grid((g.N + 63) / 64, (unsigned)((g.M + 63) / 64), g.ksplit);
    hipLaunchKernelGGL(gemm_bf16, grid, dim3(256), 0, st, g);
    ST_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Implicit-GEMM 3D convolution, forward (stride 1, zero "same" padding):
//   Y[s, v, n] = act(b[n] + sum_{tap, c} X[s, v + off(tap), c] * W[tap*cin + c, n])
// where X is the channel concatenation of X0 (C0 channels; at half
// resolution, nearest-upsampled, when up0) and X1 (C1 channels, may be
// empty) -- the decoder's upsample + skip concat happens while staging.
//
// One workgroup (4 waves) per BX x 4 x 4 output brick of one sample.  The
// (BX + KS - 1) x (4 + KS - 1)^2 input halo, all channels, is staged once in
// LDS as BF16.  M-tile t holds the 16 voxels x = 4 (t >> 2) + (m & 3) and
// z = m >> 2, y = t & 3 (3^3 layers) or y = m >> 2, z = t & 3 (1^3 layers).
// K = tap * CINP + c in steps of 32
// (v_mfma_f32_16x16x32_bf16, lane m + 16 q holding k = 8 q .. 8 q + 7):
// CINP is a multiple of 8, so a lane's eight k share one tap and its A
// fragment is one 16-byte LDS read, its B fragment one 16-byte load of the
// fragment-ordered weights (a wave reads 1 KB contiguous; n padded to 16 NT,
// zero weights past the last tap).  The B fragments of a wave's next PD
// steps are in flight while the current one computes.
//
// LDS banks (64 x 4 bytes; a 16-byte read is served 16 lanes at a time): the
// 16 lanes of one q read one channel chunk of the 16 voxels of an m-tile,
// x = m & 3 plus a row index m >> 2 along z (3^3 layers) or y (1^3 layers).
// With a voxel stride of an odd number of 16-byte chunks and a row stride of
// 4 or 12 mod 16 voxels, the 16 voxels fall on 16 distinct 4-bank groups.
// ---------------------------------------------------------------------------
template <int KS, int CINP, int NT, int BX, bool VEC, bool IBF, bool OBF>
__global__ __launch_bounds__(256) void conv3d_fwd(const void *__restrict__ X0v, int C0, int up0,
                                                  const void *__restrict__ X1v, int C1,
                                                  const unsigned short *__restrict__ Wp,
                                                  const float *__restrict__ bias,
                                                  void *__restrict__ Yv, int e, int cout, int relu) {
    // activations in FP32, or in BF16 (IBF: inputs, OBF: output)
    using TI = typename std::conditional<IBF, unsigned short, float>::type;
    const TI *__restrict__ X0 = static_cast<const TI *>(X0v);
    const TI *__restrict__ X1 = static_cast<const TI *>(X1v);
    constexpr int HF = KS / 2, HX = BX + KS - 1, H = 4 + KS - 1;
    // ZROW: an m-tile's four voxel rows run along z (halo stride HX H: 36,
    // 60 or 108 for the 3^3 layers, all 4 or 12 mod 16); otherwise along y
    // with the halo's x extent padded to 4 or 12 mod 16 (1^3 layers: 16 -> 20)
    constexpr bool ZROW = KS == 3;
    constexpr int HXP = ZROW ? HX : (HX + 3) / 4 * 4 + (((HX + 3) / 4) % 2 == 0 ? 4 : 0);
    constexpr int NH = HXP * H * H;
    // voxel stride in halves: a multiple of 8 with an odd quotient
    constexpr int TAPS = KS * KS * KS, CST = (CINP / 8) % 2 == 1 ? CINP : CINP + 8;
    constexpr int KTOT = TAPS * CINP, NS = (KTOT + 31) / 32;  // 32-wide k steps
    static_assert(CINP % 8 == 0 && (CST / 8) % 2 == 1, "halo layout");
    // KS = 3 with 4 x 4 x 4 bricks (the wide-input layers): the k steps are
    // dealt to the 4 waves (s = w, w + 4, ...), each wave computes all 4
    // m-tiles over its steps (every weight fragment is read by one wave of
    // the workgroup), and the partial sums meet in LDS in a fixed order.
    // Otherwise wave w owns m-tiles w and w + 4 over all steps (measured
    // faster for the 8 x 4 x 4 bricks: 1.03 against 1.24 ms for the
    // 48-input layer at 256 x 32^3).
    constexpr bool KSPLIT = KS > 1 && BX == 4;
    constexpr int MT = KSPLIT ? BX : BX / 4, GSTEP = KSPLIT ? 4 : 1;
    constexpr int NHALO = NH * CST, NPART = KSPLIT ? 1024 * MT * NT : 0;  // halves
    __shared__ __align__(16) unsigned short halo[NHALO > NPART ? NHALO : NPART];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int nbx = (e + BX - 1) / BX, nb4 = (e + 3) >> 2;
    int bid = blockIdx.x;
    const int bx = bid % nbx; bid /= nbx;
    const int by = bid % nb4; bid /= nb4;
    const int bz = bid % nb4;
    const int s = bid / nb4;
    const long long V = (long long)e * e * e;
    const int eh = e >> 1;
    const long long Vh = (long long)eh * eh * eh;
    const int ox = BX * bx - HF, oy = 4 * by - HF, oz = 4 * bz - HF;
    const int cin = C0 + C1;

    // stage the halo, 4 channels per item, SU items per thread in flight:
    // every load of a group is issued before the first one is converted.
    // VEC: both inputs have a multiple of 4 channels (one 16-byte load per
    // item); otherwise the channels are read one by one.
    constexpr int C4 = CINP / 4;
    // BF16 halos of the 16-output-channel layers (few registers in use)
    // are loaded in one batch where 24 items per thread cover them (two
    // registers an item); the others in batches of 8
    constexpr int NI = HX * H * H * C4, NPT = (NI + 255) / 256;
    constexpr int SU = (IBF && NT == 1) ? (NPT < 24 ? NPT : 24) : 8;
    for (int base = 0; base < NI; base += 256 * SU) {
        float f[SU][4];      // FP32 inputs
        uint2 raw[SU];       // BF16 inputs, 4 channels as loaded
#pragma unroll
        for (int u = 0; u < SU; u++) {
            const int i = base + u * 256 + tid;
            f[u][0] = f[u][1] = f[u][2] = f[u][3] = 0.f;
            raw[u] = make_uint2(0u, 0u);
            if (i >= NI) continue;
            const int hv = i / C4, c = (i - hv * C4) * 4;
            const int hx = hv % HX, hy = (hv / HX) % H, hz = hv / (HX * H);
            const int xx = ox + hx, yy = oy + hy, zz = oz + hz;
            if (c < cin && xx >= 0 && yy >= 0 && zz >= 0 && xx < e && yy < e && zz < e) {
                const TI *p;
                int cc, nav;
                if (c < C0) {
                    p = up0 ? X0 + ((long long)s * Vh + (xx >> 1) + (long long)eh * ((yy >> 1) + (long long)eh * (zz >> 1))) * C0
                            : X0 + ((long long)s * V + xx + (long long)e * (yy + (long long)e * zz)) * C0;
                    cc = c;
                    nav = C0 - c;
                } else {
                    p = X1 + ((long long)s * V + xx + (long long)e * (yy + (long long)e * zz)) * C1;
                    cc = c - C0;
                    nav = C1 - cc;
                }
                if constexpr (IBF) {
                    if constexpr (VEC) {
                        raw[u] = *reinterpret_cast<const uint2 *>(p + cc);
                    } else {
                        unsigned short hb4[4] = {0, 0, 0, 0};
#pragma unroll
                        for (int k = 0; k < 4; k++)
                            if (k < nav) hb4[k] = p[cc + k];
                        raw[u].x = (unsigned)hb4[0] | ((unsigned)hb4[1] << 16);
                        raw[u].y = (unsigned)hb4[2] | ((unsigned)hb4[3] << 16);
                    }
                } else if constexpr (VEC) {
                    const float4 v = *reinterpret_cast<const float4 *>(p + cc);
                    f[u][0] = v.x; f[u][1] = v.y; f[u][2] = v.z; f[u][3] = v.w;
                } else {
#pragma unroll
                    for (int k = 0; k < 4; k++)
                        if (k < nav) f[u][k] = p[cc + k];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < SU; u++) {
            const int i = base + u * 256 + tid;
            if (i >= NI) continue;
            const int hv = i / C4, c = (i - hv * C4) * 4;
            const int hx = hv % HX, hr = hv / HX;  // hr = hy + H hz
            const int lv = hx + HXP * hr;           // LDS voxel
            uint2 w2;
            if constexpr (IBF) {
                w2 = raw[u];
            } else {
                w2.x = (unsigned)f2bf(f[u][0]) | ((unsigned)f2bf(f[u][1]) << 16);
                w2.y = (unsigned)f2bf(f[u][2]) | ((unsigned)f2bf(f[u][3]) << 16);
            }
            *reinterpret_cast<uint2 *>(&halo[lv * CST + c]) = w2;
        }
    }
    __syncthreads();

    const int m = lane & 15, q = lane >> 4;
    // halo index of this lane's voxel in each of its m-tiles (tap offset 0)
    int hb[MT];
#pragma unroll
    for (int t = 0; t < MT; t++) {
        const int mt = KSPLIT ? t : w + 4 * t;
        const int vx = 4 * (mt >> 2) + (m & 3);
        const int vy = ZROW ? mt & 3 : m >> 2, vz = ZROW ? m >> 2 : mt & 3;
        hb[t] = vx + HXP * (vy + H * vz);
    }

    f4v acc[MT][NT];
#pragma unroll
    for (int t = 0; t < MT; t++)
#pragma unroll
        for (int j = 0; j < NT; j++) acc[t][j] = f4v{0.f, 0.f, 0.f, 0.f};

    // B fragments of the next PD steps in flight: a ring indexed at compile
    // time (the step loop is unrolled PD times), so no register copy waits
    // for a load issued in the same iteration
    constexpr int PDR = 32 / (4 * NT);
    constexpr int PD = PDR < 2 ? 2 : (PDR > 8 ? 8 : PDR);
    b8v bq[PD][NT];
    auto load_b = [&](int st, b8v (&b)[NT]) {
#pragma unroll
        for (int j = 0; j < NT; j++)
            b[j] = *reinterpret_cast<const b8v *>(Wp + (((long long)st * NT + j) * 64 + lane) * 8);
    };
    // DBUF: A fragments double-buffered, a step's LDS reads issued before the
    // previous step's MFMAs so that their latency overlaps them.  It pays
    // where one A fragment feeds one MFMA (16 output channels) from 16 or more
    // inputs (48-input layer 1,030 -> 886 us per 256-shot inference); with
    // more output channels, or the k split, the extra registers cost more
    // than the latency they hide, so those layers read each step's fragments
    // just before its MFMAs.
    constexpr bool DBUF = NT == 1 && CINP >= 16 && KS == 3;
    int hc[MT];
#pragma unroll
    for (int t = 0; t < MT; t++) hc[t] = hb[t] * CST;
    auto load_a = [&](int st, b8v (&a)[MT]) {
        const int kb = 32 * st + 8 * q;
        int tap = kb / CINP;
        const int c = kb - tap * CINP;
        if (tap >= TAPS) tap = 0;  // past the last tap: the weights are zero
        const int tx = tap % KS, ty = (tap / KS) % KS, tz = tap / (KS * KS);
        const int off = (tx + HXP * (ty + H * tz)) * CST + c;
#pragma unroll
        for (int t = 0; t < MT; t++) a[t] = *reinterpret_cast<const b8v *>(&halo[hc[t] + off]);
    };
    auto mfmas = [&](const b8v (&a)[MT], const b8v (&b)[NT]) {
#pragma unroll
        for (int t = 0; t < MT; t++)
#pragma unroll
            for (int j = 0; j < NT; j++)
                acc[t][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[t], b[j], acc[t][j], 0, 0, 0);
    };
    const int s0 = KSPLIT ? w : 0;
#pragma unroll
    for (int d = 0; d < PD; d++)
        if (s0 + d * GSTEP < NS) load_b(s0 + d * GSTEP, bq[d]);
    static_assert(PD % 2 == 0, "the A buffers alternate within an unrolled group");
    b8v a0[MT], a1[MT];
    if (DBUF && s0 < NS) load_a(s0, a0);
    for (int sb = s0; sb < NS; sb += PD * GSTEP) {
#pragma unroll
        for (int d = 0; d < PD; d++) {
            const int st = sb + d * GSTEP;
            if (st >= NS) break;
            if constexpr (!DBUF) {
                load_a(st, a0);
                mfmas(a0, bq[d]);
            } else if (d % 2 == 0) {
                if (st + GSTEP < NS) load_a(st + GSTEP, a1);
                mfmas(a0, bq[d]);
            } else {
                if (st + GSTEP < NS) load_a(st + GSTEP, a0);
                mfmas(a1, bq[d]);
            }
            if (st + PD * GSTEP < NS) load_b(st + PD * GSTEP, bq[d]);
        }
    }

    if constexpr (KSPLIT) {
        // partial sums: waves 2, 3 -> waves 0, 1, then wave 1 -> wave 0
        float *part = reinterpret_cast<float *>(halo);
        constexpr int PT = MT * NT * 4;
        auto put = [&](int slot) {
#pragma unroll
            for (int t = 0; t < MT; t++)
#pragma unroll
                for (int j = 0; j < NT; j++)
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        part[(slot * PT + (t * NT + j) * 4 + r) * 64 + lane] = acc[t][j][r];
        };
        auto add = [&](int slot) {
#pragma unroll
            for (int t = 0; t < MT; t++)
#pragma unroll
                for (int j = 0; j < NT; j++)
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        acc[t][j][r] += part[(slot * PT + (t * NT + j) * 4 + r) * 64 + lane];
        };
        __syncthreads();  // every wave is done with the halo
        if (w >= 2) put(w - 2);
        __syncthreads();
        if (w < 2) add(w);
        __syncthreads();
        if (w == 1) put(0);
        __syncthreads();
        if (w != 0) return;
        add(0);
    }
#pragma unroll
    for (int t = 0; t < MT; t++) {
        const int mt = KSPLIT ? t : w + 4 * t;
#pragma unroll
        for (int j = 0; j < NT; j++) {
            const int n = j * 16 + m;
            if (n >= cout) continue;
            const float bn = bias ? bias[n] : 0.f;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int mo = 4 * q + r;
                const int x = BX * bx + 4 * (mt >> 2) + (mo & 3),
                          y = 4 * by + (ZROW ? mt & 3 : mo >> 2),
                          z = 4 * bz + (ZROW ? mo >> 2 : mt & 3);
                if (x >= e || y >= e || z >= e) continue;
                float v = acc[t][j][r] + bn;
                if (relu) v = fmaxf(v, 0.f);
                const long long o = ((long long)s * V + x + (long long)e * (y + (long long)e * z)) * cout + n;
                if constexpr (OBF)
                    static_cast<unsigned short *>(Yv)[o] = f2bf(v);
                else
                    static_cast<float *>(Yv)[o] = v;
            }
        }
    }
}

// Wp in MFMA fragment order: fragment (s, j) is 64 lanes x 8 BF16, lane =
// m + 16 q holding W[k][n = 16 j + m] for k = 32 s + 8 q .. + 7, where
// k = tap * cinp + c; zero where tap >= taps, n >= cout or c >= cin.  A
// wave's B fragment is one contiguous 1 KB read.  W16 is [tap*cin + c][cout].
__global__ void pack_conv(const unsigned short *__restrict__ W16, unsigned short *__restrict__ Wp,
                          int taps, int cin, int cout, int cinp, int coutp) {
    const int ns = (taps * cinp + 31) / 32, nt = coutp / 16;
    const long long tot = (long long)ns * nt * 512;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < tot;
         i += (long long)gridDim.x * blockDim.x) {
        const int e = (int)(i & 7), lane = (int)((i >> 3) & 63);
        const long long f = i >> 9;
        const int j = (int)(f % nt), s = (int)(f / nt);
        const int m = lane & 15, q = lane >> 4, n = 16 * j + m;
        const int k = 32 * s + 8 * q + e;
        const int tap = k / cinp, c = k - tap * cinp;
        unsigned short v = 0;
        if (tap < taps && n < cout && c < cin) v = W16[((long long)tap * cin + c) * cout + n];
        Wp[i] = v;
    }
}

// ---------------------------------------------------------------------------
// elementwise / layout kernels
// ---------------------------------------------------------------------------
#define GRID_STRIDE(i, n) \
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (n); \
         i += (long long)gridDim.x * blockDim.x)

// col[(s,v), tap*C + c] = X[s, v + off(tap), c] (zero outside), BF16
__global__ void im2col3(const float *__restrict__ X, unsigned short *__restrict__ col, int ns,
                        int n, int C, int ksz) {
    const long long V = (long long)n * n * n;
    const int taps = ksz * ksz * ksz, h = ksz / 2;
    const long long total = (long long)ns * V * taps * C;
    GRID_STRIDE(i, total) {
        const int c = (int)(i % C);
        long long r = i / C;
        const int tap = (int)(r % taps);
        const long long sv = r / taps;
        const long long s = sv / V;
        const int v = (int)(sv - s * V);
        const int x = v % n, y = (v / n) % n, z = v / (n * n);
        const int tx = tap % ksz - h, ty = (tap / ksz) % ksz - h, tz = tap / (ksz * ksz) - h;
        const int xx = x + tx, yy = y + ty, zz = z + tz;
        float val = 0.f;
        if (xx >= 0 && yy >= 0 && zz >= 0 && xx < n && yy < n && zz < n)
            val = X[(s * V + xx + (long long)n * (yy + (long long)n * zz)) * C + c];
        col[i] = f2bf(val);
    }
}

// dX[s, v, c] = sum_tap dcol[(s, v - off(tap)), tap*C + c]
__global__ void col2im3(const float *__restrict__ dcol, float *__restrict__ dX, int ns, int n,
                        int C, int ksz) {
    const long long V = (long long)n * n * n;
    const int taps = ksz * ksz * ksz, h = ksz / 2;
    const long long total = (long long)ns * V * C;
    GRID_STRIDE(i, total) {
        const int c = (int)(i % C);
        const long long sv = i / C;
        const long long s = sv / V;
        const int v = (int)(sv - s * V);
        const int x = v % n, y = (v / n) % n, z = v / (n * n);
        float acc = 0.f;
        for (int tap = 0; tap < taps; tap++) {
            const int tx = tap % ksz - h, ty = (tap / ksz) % ksz - h, tz = tap / (ksz * ksz) - h;
            const int xx = x - tx, yy = y - ty, zz = z - tz;
            if (xx < 0 || yy < 0 || zz < 0 || xx >= n || yy >= n || zz >= n) continue;
            const long long src = s * V + xx + (long long)n * (yy + (long long)n * zz);
            acc += dcol[(src * taps + tap) * C + c];
        }
        dX[i] = acc;
    }
}

// 2x average pooling: out n/2 cube
__global__ void avgpool2(const float *__restrict__ X, float *__restrict__ Y, int ns, int n, int C) {
    const int m = n / 2;
    const long long V2 = (long long)m * m * m, V = (long long)n * n * n;
    GRID_STRIDE(i, (long long)ns * V2 * C) {
        const int c = (int)(i % C);
        const long long sv = i / C;
        const long long s = sv / V2;
        const int v = (int)(sv - s * V2);
        const int x = v % m, y = (v / m) % m, z = v / (m * m);
        float acc = 0.f;
        for (int d = 0; d < 8; d++) {
            const int xx = 2 * x + (d & 1), yy = 2 * y + ((d >> 1) & 1), zz = 2 * z + (d >> 2);
            acc += X[(s * V + xx + (long long)n * (yy + (long long)n * zz)) * C + c];
        }
        Y[i] = 0.125f * acc;
    }
}

// the same on BF16 activations (FP32 sum, rounded once)
__global__ void avgpool2_bf16(const unsigned short *__restrict__ X, unsigned short *__restrict__ Y,
                              int ns, int n, int C) {
    const int m = n / 2;
    const long long V2 = (long long)m * m * m, V = (long long)n * n * n;
    GRID_STRIDE(i, (long long)ns * V2 * C) {
        const int c = (int)(i % C);
        const long long sv = i / C;
        const long long s = sv / V2;
        const int v = (int)(sv - s * V2);
        const int x = v % m, y = (v / m) % m, z = v / (m * m);
        float acc = 0.f;
        for (int d = 0; d < 8; d++) {
            const int xx = 2 * x + (d & 1), yy = 2 * y + ((d >> 1) & 1), zz = 2 * z + (d >> 2);
            acc += bf2f(X[(s * V + xx + (long long)n * (yy + (long long)n * zz)) * C + c]);
        }
        Y[i] = f2bf(0.125f * acc);
    }
}

__global__ void avgpool2_bwd(const float *__restrict__ dY, float *__restrict__ dX, int ns, int n,
                             int C) {
    const int m = n / 2;
    const long long V2 = (long long)m * m * m, V = (long long)n * n * n;
    GRID_STRIDE(i, (long long)ns * V * C) {
        const int c = (int)(i % C);
        const long long sv = i / C;
        const long long s = sv / V;
        const int v = (int)(sv - s * V);
        const int x = v % n, y = (v / n) % n, z = v / (n * n);
        dX[i] = 0.125f * dY[(s * V2 + x / 2 + (long long)m * (y / 2 + (long long)m * (z / 2))) * C + c];
    }
}

// nearest 2x upsampling of X (m cube) into channels [off, off+C) of Y (2m
// cube, Ctot channels)
__global__ void upsample2(const float *__restrict__ X, float *__restrict__ Y, int ns, int m, int C,
                          int Ctot, int off) {
    const int n = 2 * m;
    const long long V = (long long)n * n * n, V2 = (long long)m * m * m;
    GRID_STRIDE(i, (long long)ns * V * C) {
        const int c = (int)(i % C);
        const long long sv = i / C;
        const long long s = sv / V;
        const int v = (int)(sv - s * V);
        const int x = v % n, y = (v / n) % n, z = v / (n * n);
        Y[sv * Ctot + off + c] = X[(s * V2 + x / 2 + (long long)m * (y / 2 + (long long)m * (z / 2))) * C + c];
    }
}

__global__ void upsample2_bwd(const float *__restrict__ dY, float *__restrict__ dX, int ns, int m,
                              int C, int Ctot, int off) {
    const int n = 2 * m;
    const long long V = (long long)n * n * n, V2 = (long long)m * m * m;
    GRID_STRIDE(i, (long long)ns * V2 * C) {
        const int c = (int)(i % C);
        const long long sv = i / C;
        const long long s = sv / V2;
        const int v = (int)(sv - s * V2);
        const int x = v % m, y = (v / m) % m, z = v / (m * m);
        float acc = 0.f;
        for (int d = 0; d < 8; d++) {
            const int xx = 2 * x + (d & 1), yy = 2 * y + ((d >> 1) & 1), zz = 2 * z + (d >> 2);
            acc += dY[(s * V + xx + (long long)n * (yy + (long long)n * zz)) * Ctot + off + c];
        }
        dX[i] = acc;
    }
}

// copy X (C channels) into channels [off, off+C) of Y (Ctot), or back
__global__ void chan_copy(const float *__restrict__ X, float *__restrict__ Y, long long rows, int C,
                          int Ctot, int off, int backward) {
    GRID_STRIDE(i, rows * C) {
        const long long r = i / C;
        const int c = (int)(i - r * C);
        if (!backward) Y[r * Ctot + off + c] = X[i];
        else Y[i] = X[r * Ctot + off + c];  // here X is the wide gradient
    }
}

// dY *= (Y > 0), and a BF16 copy
__global__ void relu_bwd(float *__restrict__ dY, const float *__restrict__ Y,
                         unsigned short *__restrict__ d16, long long n, int relu) {
    GRID_STRIDE(i, n) {
        float d = dY[i];
        if (relu && !(Y[i] > 0.f)) d = 0.f;
        dY[i] = d;
        d16[i] = f2bf(d);
    }
}

// db[c] += sum_r dY[r, c]
__global__ void bias_grad(const float *__restrict__ dY, float *__restrict__ db, long long rows, int C) {
    __shared__ float sh[256];
    const int c = blockIdx.y;
    float acc = 0.f;
    for (long long r = (long long)blockIdx.x * 256 + threadIdx.x; r < rows; r += (long long)gridDim.x * 256)
        acc += dY[r * C + c];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(&db[c], sh[0]);
}

// weighted MSE with under-predictions weighted wu: w = m (p < t ? wu : 1),
// loss = sum w (p - t)^2 / sum m; grad = 2 w (p - t) / sum m.  wu > 1 pulls
// the prediction towards an upper bound of the target (an expectile).
__global__ void mse(const float *__restrict__ p, const float *__restrict__ t,
                    const float *__restrict__ m, float *__restrict__ grad, double *__restrict__ acc,
                    long long n, float wu) {
    __shared__ double sl[256], sw[256];
    double l = 0.0, ws = 0.0;
    GRID_STRIDE(i, n) {
        const float d = p[i] - t[i];
        const float w = d < 0.f ? m[i] * wu : m[i];
        l += (double)w * d * d;
        ws += m[i];
        grad[i] = 2.f * w * d;  // scaled by 1/sum m afterwards
    }
    sl[threadIdx.x] = l;
    sw[threadIdx.x] = ws;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) { sl[threadIdx.x] += sl[threadIdx.x + o]; sw[threadIdx.x] += sw[threadIdx.x + o]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { atomicAdd(&acc[0], sl[0]); atomicAdd(&acc[1], sw[0]); }
}

__global__ void scale_by_inv(float *__restrict__ g, const double *__restrict__ acc, long long n) {
    const float s = acc[1] > 0 ? (float)(1.0 / acc[1]) : 0.f;
    GRID_STRIDE(i, n) g[i] *= s;
}

__global__ void adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                     float *__restrict__ v, unsigned short *__restrict__ w16, long long n, float lr,
                     float b1, float b2, float c1, float c2, float eps) {
    GRID_STRIDE(i, n) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        w[i] = wi;
        if (w16) w16[i] = f2bf(wi);
    }
}

__global__ void to_bf16(const float *__restrict__ x, unsigned short *__restrict__ y, long long n) {
    GRID_STRIDE(i, n) y[i] = f2bf(x[i]);
}

static unsigned blocks_for(long long n) {
    long long b = (n + 255) / 256;
    return (unsigned)(b < 65536 ? b : 65536);
}

// ---------------------------------------------------------------------------
// host wrappers
// ---------------------------------------------------------------------------
void im2col(hipStream_t st, const float *X, unsigned short *col, int ns, int n, int C, int ksz) {
    const long long tot = (long long)ns * n * n * n * ksz * ksz * ksz * C;
    hipLaunchKernelGGL(im2col3, dim3(blocks_for(tot)), dim3(256), 0, st, X, col, ns, n, C, ksz);
}
void col2im(hipStream_t st, const float *dcol, float *dX, int ns, int n, int C, int ksz) {
    const long long tot = (long long)ns * n * n * n * C;
    hipLaunchKernelGGL(col2im3, dim3(blocks_for(tot)), dim3(256), 0, st, dcol, dX, ns, n, C, ksz);
}
void pool(hipStream_t st, const float *X, float *Y, int ns, int n, int C) {
    hipLaunchKernelGGL(avgpool2, dim3(blocks_for((long long)ns * n * n * n / 8 * C)), dim3(256), 0,
                       st, X, Y, ns, n, C);
}
void pool_bf16(hipStream_t st, const unsigned short *X, unsigned short *Y, int ns, int n, int C) {
    hipLaunchKernelGGL(avgpool2_bf16, dim3(blocks_for((long long)ns * n * n * n / 8 * C)), dim3(256),
                       0, st, X, Y, ns, n, C);
}
void pool_bwd(hipStream_t st, const float *dY, float *dX, int ns, int n, int C) {
    hipLaunchKernelGGL(avgpool2_bwd, dim3(blocks_for((long long)ns * n * n * n * C)), dim3(256), 0,
                       st, dY, dX, ns, n, C);
}
void upsample(hipStream_t st, const float *X, float *Y, int ns, int m, int C, int Ctot, int off) {
    hipLaunchKernelGGL(upsample2, dim3(blocks_for((long long)ns * 8 * m * m * m * C)), dim3(256), 0,
                       st, X, Y, ns, m, C, Ctot, off);
}
void upsample_bwd(hipStream_t st, const float *dY, float *dX, int ns, int m, int C, int Ctot, int off) {
    hipLaunchKernelGGL(upsample2_bwd, dim3(blocks_for((long long)ns * m * m * m * C)), dim3(256), 0,
                       st, dY, dX, ns, m, C, Ctot, off);
}
void chancopy(hipStream_t st, const float *X, float *Y, long long rows, int C, int Ctot, int off,
              bool backward) {
    hipLaunchKernelGGL(chan_copy, dim3(blocks_for(rows * C)), dim3(256), 0, st, X, Y, rows, C, Ctot,
                       off, backward ? 1 : 0);
}
void relu_backward(hipStream_t st, float *dY, const float *Y, unsigned short *d16, long long n,
                   bool relu) {
    hipLaunchKernelGGL(relu_bwd, dim3(blocks_for(n)), dim3(256), 0, st, dY, Y, d16, n, relu ? 1 : 0);
}
void bias_gradient(hipStream_t st, const float *dY, float *db, long long rows, int C) {
    hipLaunchKernelGGL(bias_grad, dim3(64, C), dim3(256), 0, st, dY, db, rows, C);
}
void mse_loss(hipStream_t st, const float *p, const float *t, const float *m, float *grad,
              double *acc, long long n, float under_weight) {
    hipLaunchKernelGGL(mse, dim3(blocks_for(n) < 1024 ? blocks_for(n) : 1024), dim3(256), 0, st, p,
                       t, m, grad, acc, n, under_weight);
    hipLaunchKernelGGL(scale_by_inv, dim3(blocks_for(n)), dim3(256), 0, st, grad, acc, n);
}
void adam_step(hipStream_t st, float *w, const float *g, float *m, float *v, unsigned short *w16,
               long long n, float lr, int step) {
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(adam, dim3(blocks_for(n)), dim3(256), 0, st, w, g, m, v, w16, n, lr, b1, b2,
                       c1, c2, 1e-8f);
}
void bf16_copy(hipStream_t st, const float *x, unsigned short *y, long long n) {
    hipLaunchKernelGGL(to_bf16, dim3(blocks_for(n)), dim3(256), 0, st, x, y, n);
}


int conv_cinp(int cin) { return (cin + 7) / 8 * 8; }
int conv_coutp(int cout) { return cout <= 16 ? 16 : (cout <= 32 ? 32 : 64); }
long long conv_packed_size(int ksz, int cin, int cout) {
    // 32-wide k steps over taps x padded channels
    return (long long)((ksz * ksz * ksz * conv_cinp(cin) + 31) / 32) * 32 * conv_coutp(cout);
}

void pack_conv_weights(hipStream_t st, const unsigned short *W16, unsigned short *Wp, int ksz,
                       int cin, int cout) {
    const int taps = ksz * ksz * ksz;
    const long long tot = conv_packed_size(ksz, cin, cout);
    hipLaunchKernelGGL(pack_conv, dim3(blocks_for(tot)), dim3(256), 0, st, W16, Wp, taps, cin,
                       cout, conv_cinp(cin), conv_coutp(cout));
    ST_HIP(hipGetLastError());
}

void conv_forward_any(hipStream_t st, const void *X0, int C0, bool up0, const void *X1, int C1,
                      bool in_bf16, const unsigned short *Wp, const float *bias, void *Y,
                      bool out_bf16, int ns, int e, int cout, int ksz, bool relu) {
    const int cin = C0 + C1;
    ST_REQUIRE(ksz == 1 || ksz == 3, ST_ERR_ARG, "conv kernel size must be 1 or 3");
    ST_REQUIRE(C0 >= 1 && C1 >= 0 && (C1 == 0 || X1) && cout >= 1 && cout <= 64 && e >= 1 &&
                   (!up0 || e % 2 == 0),
               ST_ERR_ARG, "conv shape out of range");
    ST_REQUIRE(C1 == 0 || C0 % 4 == 0, ST_ERR_ARG, "concatenated inputs need C0 % 4 == 0");
    const int cinp = conv_cinp(cin), coutp = conv_coutp(cout);
    // bricks: 16 wide for 16 output channels from at most 16 inputs (four
    // m-tiles per wave share each weight fragment), 8 up to 48 inputs (a
    // 16-wide 48-channel halo fills 67 KB of LDS and ran slower), else 4
    const int bx = (coutp == 16 && cinp <= 16) ? 16 : (cinp <= 48 ? 8 : 4);
    const long long nwg = (long long)ns * ((e + bx - 1) / bx) * ((e + 3) / 4) * ((e + 3) / 4);
    ST_REQUIRE(nwg < (1LL << 31), ST_ERR_ARG, "conv grid too large");
    if (nwg == 0) return;
    const dim3 g((unsigned)nwg), b(256);
    const int u = up0 ? 1 : 0, r = relu ? 1 : 0;
    const bool vec = ((C0 | C1) & 3) == 0;  // 16-byte (8-byte for BF16) channel groups
    const int mode = (vec ? 4 : 0) | (in_bf16 ? 2 : 0) | (out_bf16 ? 1 : 0);
#define ST_CONV_MODE(KS, CP, NT, BXV, M)                                                          \
    case M:                                                                                       \
        hipLaunchKernelGGL((conv3d_fwd<KS, CP, NT, BXV, ((M) & 4) != 0, ((M) & 2) != 0,           \
                                       ((M) & 1) != 0>),                                          \
                           g, b, 0, st, X0, C0, u, X1, C1, Wp, bias, Y, e, cout, r);              \
        break;
#define ST_CONV(KS, CP, NT, BXV)                                                                  \
    if (ksz == KS && cinp == CP && coutp == 16 * NT) {                                            \
        ST_REQUIRE(bx == BXV, ST_ERR_ARG, "conv brick");                                          \
        switch (mode) {                                                                           \
            ST_CONV_MODE(KS, CP, NT, BXV, 0) ST_CONV_MODE(KS, CP, NT, BXV, 1)                     \
            ST_CONV_MODE(KS, CP, NT, BXV, 2) ST_CONV_MODE(KS, CP, NT, BXV, 3)                     \
            ST_CONV_MODE(KS, CP, NT, BXV, 4) ST_CONV_MODE(KS, CP, NT, BXV, 5)                     \
            ST_CONV_MODE(KS, CP, NT, BXV, 6) ST_CONV_MODE(KS, CP, NT, BXV, 7)                     \
        }                                                                                         \
        ST_HIP(hipGetLastError());                                                                \
        return;                                                                                   \
    }
    ST_CONV(3, 8, 1, 16)
    ST_CONV(3, 16, 1, 16)
    ST_CONV(3, 48, 1, 8)
    ST_CONV(3, 16, 2, 8)
    ST_CONV(3, 32, 2, 8)
    ST_CONV(3, 96, 2, 4)
    ST_CONV(3, 32, 4, 8)
    ST_CONV(3, 64, 4, 4)
    ST_CONV(3, 128, 4, 4)
    ST_CONV(1, 16, 1, 16)
#undef ST_CONV
#undef ST_CONV_MODE
    ST_REQUIRE(false, ST_ERR_ARG, "no convolution kernel for this channel configuration");
}

void conv_forward2(hipStream_t st, const float *X0, int C0, bool up0, const float *X1, int C1,
                   const unsigned short *Wp, const float *bias, float *Y, int ns, int e, int cout,
                   int ksz, bool relu) {
    conv_forward_any(st, X0, C0, up0, X1, C1, false, Wp, bias, Y, false, ns, e, cout, ksz, relu);
}

void conv_forward(hipStream_t st, const float *X, const unsigned short *Wp, const float *bias,
                  float *Y, int ns, int e, int cin, int cout, int ksz, bool relu) {
    conv_forward2(st, X, cin, false, nullptr, 0, Wp, bias, Y, ns, e, cout, ksz, relu);
}

}  // namespace nn
}  // namespace st
