/*
 * oracle.c -- CPU restatement of the first-arrival tomography path of
 * phbastosa/seismic_tomography_3D.  TEST INFRASTRUCTURE ONLY: it is the
 * checker for the HIP path, never the thing measured or shipped.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Each function cites the reference file:line it restates.  Build flags
 * (oracle/Makefile) pin FP semantics: -O2 -ffp-contract=off, SSE math, so
 * float expressions evaluate in float with the reference's association.
 *
 * Pinning (which tests check what) is summarised per function below.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* std::min / std::max semantics (return the first argument on ties) */
static inline float fmin_std(float a, float b) { return (b < a) ? b : a; }
static inline float fmax_std(float a, float b) { return (a < b) ? b : a; }
static inline int imin(int a, int b) { return (b < a) ? b : a; }
static inline int imax(int a, int b) { return (a < b) ? b : a; }

/* ------------------------------------------------------------------------ */
/* padding: block_FIM.cpp:144-150 (extent) and :8-57 (replication).         */
/* Pinned: test_oracle_eikonal.py::test_pad_* (replication order against a  */
/* literal loop-by-loop transcription on random volumes).                  */
/* ------------------------------------------------------------------------ */
int or_pad_extent(int n, int B) { return n + (B - n % B) % B; }

void or_pad_replicate(const float *s, int nx, int ny, int nz,
                      float *S, int nxx, int nyy, int nzz)
{
    /* The reference fills the interior, then the z pad from plane nz-1,
     * then the x pad from column nx-1 (all z), then the y pad from row
     * ny-1 (all z, x).  The composition is clamp-to-edge replication. */
    for (int y = 0; y < nyy; y++)
        for (int x = 0; x < nxx; x++)
            for (int z = 0; z < nzz; z++)
                S[z + (long)x * nzz + (long)y * nxx * nzz] =
                    s[imin(z, nz - 1) + (long)imin(x, nx - 1) * nz +
                      (long)imin(y, ny - 1) * nx * nz];
}

/* ------------------------------------------------------------------------ */
/* get_time_eikonal: cuda_kernel.cu:217-248.                                */
/* Pinned: 1D/2D/3D closed forms (test_godunov_*), homogeneous axis times,  */
/* FP64 fixed point vs brute-force Dijkstra-free invariants.                */
/* ------------------------------------------------------------------------ */
float or_godunov_f32(float a, float b, float c, float h, float s)
{
    float ret, tmp;
    if (a < b) { tmp = a; a = b; b = tmp; }
    if (b < c) { tmp = b; b = c; c = tmp; }
    if (a < b) { tmp = a; a = b; b = tmp; }
    ret = OR_INF;
    if (c < OR_INF) {
        ret = c + h * s;
        if (ret > b) {
            tmp = ((b + c) + sqrtf(2.0f * s * s * h * h - (b - c) * (b - c))) * 0.5f;
            if (tmp > b) ret = tmp;
            if (ret > a) {
                tmp = (a + b + c) / 3.0f +
                      sqrtf(2.0f * (a * (b - a) + b * (c - b) + c * (a - c)) +
                            3.0f * s * s * h * h) / 3.0f;
                if (tmp > a) ret = tmp;
            }
        }
    }
    return ret;
}

double or_godunov_f64(double a, double b, double c, double h, double s)
{
    double ret, tmp;
    const double INF = (double)OR_INF;
    if (a < b) { tmp = a; a = b; b = tmp; }
    if (b < c) { tmp = b; b = c; c = tmp; }
    if (a < b) { tmp = a; a = b; b = tmp; }
    ret = INF;
    if (c < INF) {
        ret = c + h * s;
        if (ret > b) {
            tmp = ((b + c) + sqrt(2.0 * s * s * h * h - (b - c) * (b - c))) * 0.5;
            if (tmp > b) ret = tmp;
            if (ret > a) {
                tmp = (a + b + c) / 3.0 +
                      sqrt(2.0 * (a * (b - a) + b * (c - b) + c * (a - c)) +
                           3.0 * s * s * h * h) / 3.0;
                if (tmp > a) ret = tmp;
            }
        }
    }
    return ret;
}

/* apply_source_time: block_FIM.cpp:305-317 */
void or_source_seed(const float *S, int nxx, int nyy, int nzz,
                    float dx, float dy, float dz,
                    float sx, float sy, float sz, long *sid, float *t0)
{
    (void)nyy;
    int sidx = (int)(sx / dx);
    int sidy = (int)(sy / dy);
    int sidz = (int)(sz / dz);
    long id = sidz + (long)sidx * nzz + (long)sidy * nxx * nzz;
    *sid = id;
    *t0 = S[id] * sqrtf(powf((float)(sidx * dx) - sx, 2.0f) +
                        powf((float)(sidy * dy) - sy, 2.0f) +
                        powf((float)(sidz * dz) - sz, 2.0f));
}

/* ------------------------------------------------------------------------ */
/* FP64 fixed point of the same discretisation (what the 1e-4 s gate is     */
/* measured against).  Domain = padded extent, slowness replicated, missing */
/* neighbours at the domain edge ignored (equivalent to the reference's     */
/* "halo = own value", cuda_kernel.cu:294-297, because the update is        */
/* min(old, G)).                                                            */
/* Pinned: exact n*h*s on axes of a homogeneous medium, exact sum of h*s    */
/* down a vertical line of a layered medium, order independence (forward vs */
/* reverse sweep start), agreement with or_block_fim_ref.                   */
/* ------------------------------------------------------------------------ */
int or_eikonal_fp64(const float *S, int nxx, int nyy, int nzz, double h,
                    long sid, double t0, double tol, int max_pass, double *T)
{
    const double INF = (double)OR_INF;
    const long nxz = (long)nxx * nzz;
    const long n = nxz * nyy;
    for (long i = 0; i < n; i++) T[i] = INF;
    T[sid] = t0;

    for (int pass = 0; pass < max_pass; pass++) {
        double maxchg = 0.0;
        for (int dir = 0; dir < 8; dir++) {
            int ys = (dir & 4) ? -1 : 1, xs = (dir & 2) ? -1 : 1, zs = (dir & 1) ? -1 : 1;
            for (int yy = 0; yy < nyy; yy++) {
                int y = ys > 0 ? yy : nyy - 1 - yy;
                for (int xx = 0; xx < nxx; xx++) {
                    int x = xs > 0 ? xx : nxx - 1 - xx;
                    long base = (long)x * nzz + (long)y * nxz;
                    for (int zz = 0; zz < nzz; zz++) {
                        int z = zs > 0 ? zz : nzz - 1 - zz;
                        long c = base + z;
                        double xm = x > 0 ? T[c - nzz] : INF;
                        double xp = x < nxx - 1 ? T[c + nzz] : INF;
                        double ym = y > 0 ? T[c - nxz] : INF;
                        double yp = y < nyy - 1 ? T[c + nxz] : INF;
                        double zm = z > 0 ? T[c - 1] : INF;
                        double zp = z < nzz - 1 ? T[c + 1] : INF;
                        double a = xm < xp ? xm : xp;
                        double b = ym < yp ? ym : yp;
                        double cc = zm < zp ? zm : zp;
                        double t = or_godunov_f64(a, b, cc, h, (double)S[c]);
                        if (t < T[c]) {
                            double d = T[c] - t;
                            if (d > maxchg) maxchg = d;
                            T[c] = t;
                        }
                    }
                }
            }
        }
        if (maxchg < tol) return pass + 1;
    }
    return -1;
}

/* ------------------------------------------------------------------------ */
/* The reference's host-driven block FIM, restated on the CPU in FP32:      */
/* block_FIM_solver (cuda_kernel.cu:17-215), run_solver (:250-417),         */
/* run_reduction (:419-457), run_check_neighbor (:459-619).  The volume is  */
/* kept z-fastest here; blocks are 4^3 with the reference's block order     */
/* b = bx + by*nBx + bz*nBx*nBy where the reference's "x" is the volume's   */
/* x (block_FIM.cpp:271-292).  Inside a visit the update is Jacobi (all    */
/* reads, then all writes): the reference's missing barrier after the      */
/* write (cuda_kernel.cu:408) makes it a Jacobi/Gauss-Seidel mix with the   */
/* same fixed point.                                                       */
/* Pinned: agrees with or_eikonal_fp64 to < 1e-5 s on seeded models         */
/* (test_oracle_eikonal.py::test_ref_fim_vs_fp64).                          */
/* ------------------------------------------------------------------------ */
#define RB 4

typedef struct {
    int nxx, nyy, nzz, nbx, nby, nbz;
    long nblk;
} fim_geo;

/* global cell index of in-block (tx,ty,tz) of block b */
static inline long fim_cell(const fim_geo *g, long b, int tx, int ty, int tz)
{
    long bx = b % g->nbx;
    long rest = b / g->nbx;
    long by = rest % g->nby;
    long bz = rest / g->nby;
    long x = bx * RB + tx, y = by * RB + ty, z = bz * RB + tz;
    return z + x * g->nzz + y * (long)g->nxx * g->nzz;
}

/* Stage block b from sol (centre + 6 face halos, domain edge = own value)
 * into tile[6][6][6] indexed [x][y][z] as the reference's SOL(i,j,k). */
static void fim_stage(const fim_geo *g, const float *sol, long b,
                      float tile[RB + 2][RB + 2][RB + 2])
{
    long bx = b % g->nbx;
    long rest = b / g->nbx;
    long by = rest % g->nby;
    long bz = rest / g->nby;
    for (int tz = 0; tz < RB; tz++)
        for (int ty = 0; ty < RB; ty++)
            for (int tx = 0; tx < RB; tx++) {
                long c = fim_cell(g, b, tx, ty, tz);
                tile[tx + 1][ty + 1][tz + 1] = sol[c];
                if (tx == 0)
                    tile[0][ty + 1][tz + 1] = (bx == 0) ? sol[c] : sol[c - g->nzz];
                if (tx == RB - 1)
                    tile[RB + 1][ty + 1][tz + 1] = (bx == g->nbx - 1) ? sol[c] : sol[c + g->nzz];
                if (ty == 0)
                    tile[tx + 1][0][tz + 1] = (by == 0) ? sol[c] : sol[c - (long)g->nxx * g->nzz];
                if (ty == RB - 1)
                    tile[tx + 1][RB + 1][tz + 1] = (by == g->nby - 1) ? sol[c] : sol[c + (long)g->nxx * g->nzz];
                if (tz == 0)
                    tile[tx + 1][ty + 1][0] = (bz == 0) ? sol[c] : sol[c - 1];
                if (tz == RB - 1)
                    tile[tx + 1][ty + 1][RB + 1] = (bz == g->nbz - 1) ? sol[c] : sol[c + 1];
            }
}

int or_block_fim_ref(const float *S, int nxx, int nyy, int nzz, float h,
                     long sid, float t0, int n_iter, float *T, long *stats)
{
    const float EPS = 1e-6f;
    fim_geo g;
    g.nxx = nxx; g.nyy = nyy; g.nzz = nzz;
    g.nbx = nxx / RB; g.nby = nyy / RB; g.nbz = nzz / RB;
    g.nblk = (long)g.nbx * g.nby * g.nbz;
    const long n = (long)nxx * nyy * nzz;

    float *d_sol = (float *)malloc(sizeof(float) * n);
    float *t_sol = (float *)malloc(sizeof(float) * n);
    unsigned char *con = (unsigned char *)malloc(n);
    unsigned char *listVol = (unsigned char *)calloc(g.nblk, 1);
    unsigned char *listed = (unsigned char *)calloc(g.nblk, 1);
    long *list = (long *)malloc(sizeof(long) * g.nblk);
    float (*tiles)[RB + 2][RB + 2][RB + 2] = malloc(sizeof(*tiles));

    for (long i = 0; i < n; i++) { d_sol[i] = OR_INF; con[i] = 1; }
    d_sol[sid] = t0;
    memcpy(t_sol, d_sol, sizeof(float) * n);

    /* seed block: apply_source_time (block_FIM.cpp:319-366) */
    {
        long z = sid % nzz, x = (sid / nzz) % nxx, y = sid / ((long)nxx * nzz);
        long b = (x / RB) + (y / RB) * g.nbx + (z / RB) * (long)g.nbx * g.nby;
        list[0] = b; listed[b] = 1; listVol[b] = 1;
    }
    long nActive = 1, rounds = 0, visits = 0;
    const long nBlkXY = (long)g.nbx * g.nby;

    while (nActive > 0) {
        rounds++;
        visits += nActive;
        /* 1. run_solver: sol_in = d_sol, sol_out = t_sol */
        for (long li = 0; li < nActive; li++) {
            long b = list[li];
            fim_stage(&g, d_sol, b, *tiles);
            float oldT[RB][RB][RB], newT[RB][RB][RB], Fb[RB][RB][RB];
            for (int tz = 0; tz < RB; tz++)
                for (int ty = 0; ty < RB; ty++)
                    for (int tx = 0; tx < RB; tx++)
                        Fb[tx][ty][tz] = S[fim_cell(&g, b, tx, ty, tz)];
            for (int it = 0; it < n_iter; it++) {
                for (int tz = 0; tz < RB; tz++)
                    for (int ty = 0; ty < RB; ty++)
                        for (int tx = 0; tx < RB; tx++) {
                            float (*s)[RB + 2][RB + 2] = *tiles;
                            float o = s[tx + 1][ty + 1][tz + 1];
                            float a = fmin_std(s[tx][ty + 1][tz + 1], s[tx + 2][ty + 1][tz + 1]);
                            float bb = fmin_std(s[tx + 1][ty][tz + 1], s[tx + 1][ty + 2][tz + 1]);
                            float c = fmin_std(s[tx + 1][ty + 1][tz], s[tx + 1][ty + 1][tz + 2]);
                            float t = or_godunov_f32(a, bb, c, h, Fb[tx][ty][tz]);
                            oldT[tx][ty][tz] = o;
                            newT[tx][ty][tz] = fmin_std(t, o);
                        }
                for (int tz = 0; tz < RB; tz++)
                    for (int ty = 0; ty < RB; ty++)
                        for (int tx = 0; tx < RB; tx++)
                            (*tiles)[tx + 1][ty + 1][tz + 1] = newT[tx][ty][tz];
            }
            for (int tz = 0; tz < RB; tz++)
                for (int ty = 0; ty < RB; ty++)
                    for (int tx = 0; tx < RB; tx++) {
                        long c = fim_cell(&g, b, tx, ty, tz);
                        float res = oldT[tx][ty][tz] - newT[tx][ty][tz];
                        con[c] = (res < EPS) ? 1 : 0;
                        t_sol[c] = newT[tx][ty][tz];
                    }
        }
        /* 2. run_reduction */
        for (long li = 0; li < nActive; li++) {
            long b = list[li];
            int all = 1;
            for (int t = 0; t < RB * RB * RB && all; t++)
                all = con[fim_cell(&g, b, t % RB, (t / RB) % RB, t / (RB * RB))];
            listVol[b] = !all;
        }
        /* 3. add the 6 neighbours of converged blocks (cuda_kernel.cu:128-154),
         * with the reference's index clamps verbatim */
        long nOld = nActive;
        for (long li = 0; li < nOld; li++) {
            long cur = list[li];
            if (!listVol[cur]) {
                long nb[6];
                nb[0] = (cur < nBlkXY) ? cur : cur - nBlkXY;
                nb[1] = (cur + nBlkXY >= g.nblk) ? cur : cur + nBlkXY;
                nb[2] = (cur < g.nbx) ? cur : cur - g.nbx;
                nb[3] = (cur + g.nbx >= g.nblk) ? cur : cur + g.nbx;
                nb[4] = (cur % g.nbx == 0) ? cur : cur - 1;
                nb[5] = ((cur + 1) % g.nbx == 0) ? cur : cur + 1;
                for (int k = 0; k < 6; k++)
                    if (!listed[nb[k]]) { listed[nb[k]] = 1; list[nActive++] = nb[k]; }
            }
        }
        /* 4. run_check_neighbor: sol_in = t_sol, sol_out = d_sol */
        for (long li = 0; li < nActive; li++) {
            long b = list[li];
            if (li < nOld) {
                for (int t = 0; t < RB * RB * RB; t++) {
                    long c = fim_cell(&g, b, t % RB, (t / RB) % RB, t / (RB * RB));
                    d_sol[c] = t_sol[c];
                }
            } else {
                fim_stage(&g, t_sol, b, *tiles);
                float (*s)[RB + 2][RB + 2] = *tiles;
                for (int tz = 0; tz < RB; tz++)
                    for (int ty = 0; ty < RB; ty++)
                        for (int tx = 0; tx < RB; tx++) {
                            long c = fim_cell(&g, b, tx, ty, tz);
                            float o = s[tx + 1][ty + 1][tz + 1];
                            float a = fmin_std(s[tx][ty + 1][tz + 1], s[tx + 2][ty + 1][tz + 1]);
                            float bb = fmin_std(s[tx + 1][ty][tz + 1], s[tx + 1][ty + 2][tz + 1]);
                            float cc = fmin_std(s[tx + 1][ty + 1][tz], s[tx + 1][ty + 1][tz + 2]);
                            float t = or_godunov_f32(a, bb, cc, h, S[c]);
                            float nw = fmin_std(t, o);
                            d_sol[c] = nw;
                            con[c] = (o - nw < EPS) ? 1 : 0;
                        }
            }
        }
        /* 5. run_reduction over the grown list */
        for (long li = 0; li < nActive; li++) {
            long b = list[li];
            int all = 1;
            for (int t = 0; t < RB * RB * RB && all; t++)
                all = con[fim_cell(&g, b, t % RB, (t / RB) % RB, t / (RB * RB))];
            listVol[b] = !all;
        }
        /* 6. rebuild the list (cuda_kernel.cu:183-197) */
        nActive = 0;
        for (long b = 0; b < g.nblk; b++) {
            if (listVol[b]) { listed[b] = 1; list[nActive++] = b; }
            else listed[b] = 0;
        }
    }
    memcpy(T, d_sol, sizeof(float) * n);
    if (stats) { stats[0] = rounds; stats[1] = visits; }
    free(d_sol); free(t_sol); free(con); free(listVol); free(listed); free(list); free(tiles);
    return 0;
}

/* reduce_model (block_FIM.cpp:59-69) */
void or_crop_f32(const float *Tp, int nxx, int nyy, int nzz,
                 float *T, int nx, int ny, int nz)
{
    (void)nyy;
    for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++)
            for (int z = 0; z < nz; z++)
                T[z + (long)x * nz + (long)y * nx * nz] =
                    Tp[z + (long)x * nzz + (long)y * nxx * nzz];
}

void or_crop_f64(const double *Tp, int nxx, int nyy, int nzz,
                 double *T, int nx, int ny, int nz)
{
    (void)nyy;
    for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++)
            for (int z = 0; z < nz; z++)
                T[z + (long)x * nz + (long)y * nx * nz] =
                    Tp[z + (long)x * nzz + (long)y * nxx * nzz];
}

/* ------------------------------------------------------------------------ */
/* trilinear.cpp:3-20 and the receiver pick of block_FIM.cpp:228-262.       */
/* Pinned: bit-exact against the reference trilinear() built in oracle/_ref */
/* and against on-node exactness (tests/test_oracle_picks.py).              */
/* ------------------------------------------------------------------------ */
float or_trilinear(float c000, float c001, float c100, float c101,
                   float c010, float c011, float c110, float c111,
                   float x0, float x1, float y0, float y1, float z0, float z1,
                   float x, float y, float z)
{
    float xd = (x - x0) / (x1 - x0);
    float yd = (y - y0) / (y1 - y0);
    float zd = (z - z0) / (z1 - z0);

    float c00 = c000 * (1 - xd) + c100 * xd;
    float c01 = c001 * (1 - xd) + c101 * xd;
    float c10 = c010 * (1 - xd) + c110 * xd;
    float c11 = c011 * (1 - xd) + c111 * xd;

    float c0 = c00 * (1 - yd) + c10 * yd;
    float c1 = c01 * (1 - yd) + c11 * yd;

    return (c0 * (1 - zd) + c1 * zd);
}

void or_picks(const float *T, int nx, int ny, int nz,
              float dx, float dy, float dz,
              const float *rx, const float *ry, const float *rz, int nr,
              float *out)
{
    for (int r = 0; r < nr; r++) {
        float x = rx[r], y = ry[r], z = rz[r];
        float x0 = floorf(x / dx) * dx;
        float y0 = floorf(y / dy) * dy;
        float z0 = floorf(z / dz) * dz;
        float x1 = floorf(x / dx) * dx + dx;
        float y1 = floorf(y / dy) * dy + dy;
        float z1 = floorf(z / dz) * dz + dz;
        int iz = (int)(z / dz), ix = (int)(x / dx), iy = (int)(y / dy);
        /* the reference's +1 corners; clamped where it would read past the
         * volume (those corners carry weight 0 for on-node receivers) */
        long oz = (iz + 1 < nz) ? 1 : 0;
        long ox = (ix + 1 < nx) ? nz : 0;
        long oy = (iy + 1 < ny) ? (long)nx * nz : 0;
        long id = iz + (long)ix * nz + (long)iy * nx * nz;
        float c000 = T[id];
        float c001 = T[id + oz];
        float c100 = T[id + ox];
        float c101 = T[id + oz + ox];
        float c010 = T[id + oy];
        float c011 = T[id + oz + oy];
        float c110 = T[id + ox + oy];
        float c111 = T[id + oz + ox + oy];
        out[r] = or_trilinear(c000, c001, c100, c101, c010, c011, c110, c111,
                              x0, x1, y0, y1, z0, z1, x, y, z);
    }
}

/* ------------------------------------------------------------------------ */
/* Accurate_FSM (Noble, Gesret & Belayouni 2014), accurate_FSM.cpp:8-650,   */
/* restated with the same operation order.  Pinned: bit-exact against the   */
/* reference's own accurate_FSM.cpp compiled in oracle/_ref and against the */
/* committed fixtures tests/golden/fsm_*.npz.                               */
/* ------------------------------------------------------------------------ */
typedef struct {
    int nx, ny, nz, nxx, nyy, nzz;
    float dx, dy, dz;
    int i, j, k;
    int sgntz, sgntx, sgnty, sgnvz, sgnvx, sgnvy;
    float dsum, dz2i, dx2i, dy2i, dz2dx2, dz2dy2, dx2dy2;
    int sidx, sidy, sidz;
    float *S, *T;
} fsm_state;

/* expand_model (accurate_FSM.cpp:8-60) */
static void fsm_expand(fsm_state *f, const float *slowness)
{
    int nxx = f->nxx, nyy = f->nyy, nzz = f->nzz, nx = f->nx, nz = f->nz;
    float *S = f->S;
    for (int z = 1; z < nzz - 1; z++)
        for (int y = 1; y < nyy - 1; y++)
            for (int x = 1; x < nxx - 1; x++)
                S[z + (long)x * nzz + (long)y * nxx * nzz] =
                    slowness[(z - 1) + (long)(x - 1) * nz + (long)(y - 1) * nx * nz];
    for (int y = 1; y < nyy - 1; y++)
        for (int x = 1; x < nxx - 1; x++) {
            S[0 + (long)x * nzz + (long)y * nxx * nzz] =
                slowness[0 + (long)(x - 1) * nz + (long)(y - 1) * nx * nz];
            S[(nzz - 1) + (long)x * nzz + (long)y * nxx * nzz] =
                slowness[(nz - 1) + (long)(x - 1) * nz + (long)(y - 1) * nx * nz];
        }
    for (int z = 0; z < nzz; z++)
        for (int y = 1; y < nyy - 1; y++) {
            S[z + 0 + (long)y * nxx * nzz] = S[z + 1L * nzz + (long)y * nxx * nzz];
            S[z + (long)(nxx - 1) * nzz + (long)y * nxx * nzz] =
                S[z + (long)(nxx - 2) * nzz + (long)y * nxx * nzz];
        }
    for (int z = 0; z < nzz; z++)
        for (int x = 0; x < nxx; x++) {
            S[z + (long)x * nzz + 0] = S[z + (long)x * nzz + 1L * nxx * nzz];
            S[z + (long)x * nzz + (long)(nyy - 1) * nxx * nzz] =
                S[z + (long)x * nzz + (long)(nyy - 2) * nxx * nzz];
        }
}

/* inner_sweep (accurate_FSM.cpp:159-265) */
static void fsm_inner(fsm_state *f)
{
    const int i = f->i, j = f->j, k = f->k;
    const int nxx = f->nxx, nyy = f->nyy, nzz = f->nzz;
    const long NZ = nzz, NXZ = (long)nxx * nzz;
    const float *S = f->S;
    float *T = f->T;
    const float dx = f->dx, dy = f->dy, dz = f->dz;
    float ta, tb, tc, t1, t2, t3, Sref;
    float t1D1, t1D2, t1D3, t1D, t2D1, t2D2, t2D3, t2D, t3D;

    int i1 = i - f->sgnvz;
    int j1 = j - f->sgnvx;
    int k1 = k - f->sgnvy;

    float tv = T[(i - f->sgntz) + j * NZ + k * NXZ];
    float te = T[i + (j - f->sgntx) * NZ + k * NXZ];
    float tn = T[i + j * NZ + (k - f->sgnty) * NXZ];
    float tev = T[(i - f->sgntz) + (j - f->sgntx) * NZ + k * NXZ];
    float ten = T[i + (j - f->sgntx) * NZ + (k - f->sgnty) * NXZ];
    float tnv = T[(i - f->sgntz) + j * NZ + (k - f->sgnty) * NXZ];
    float tnve = T[(i - f->sgntz) + (j - f->sgntx) * NZ + (k - f->sgnty) * NXZ];

    long ijk = i + j * NZ + k * NXZ;

    t1D1 = tv + dz * fmin_std(S[i1 + imax(j - 1, 1) * NZ + imax(k - 1, 1) * NXZ],
                     fmin_std(S[i1 + imax(j - 1, 1) * NZ + imin(k, nyy - 1) * NXZ],
                     fmin_std(S[i1 + imin(j, nxx - 1) * NZ + imax(k - 1, 1) * NXZ],
                              S[i1 + imin(j, nxx - 1) * NZ + imin(k, nyy - 1) * NXZ])));

    t1D2 = te + dx * fmin_std(S[imax(i - 1, 1) + j1 * NZ + imax(k - 1, 1) * NXZ],
                     fmin_std(S[imin(i, nzz - 1) + j1 * NZ + imax(k - 1, 1) * NXZ],
                     fmin_std(S[imax(i - 1, 1) + j1 * NZ + imin(k, nyy - 1) * NXZ],
                              S[imin(i, nzz - 1) + j1 * NZ + imin(k, nyy - 1) * NXZ])));

    t1D3 = tn + dy * fmin_std(S[imax(i - 1, 1) + imax(j - 1, 1) * NZ + k1 * NXZ],
                     fmin_std(S[imax(i - 1, 1) + imin(j, nxx - 1) * NZ + k1 * NXZ],
                     fmin_std(S[imin(i, nzz - 1) + imax(j - 1, 1) * NZ + k1 * NXZ],
                              S[imin(i, nzz - 1) + imin(j, nxx - 1) * NZ + k1 * NXZ])));

    t1D = fmin_std(t1D1, fmin_std(t1D2, t1D3));

    t2D1 = 1e6f; t2D2 = 1e6f; t2D3 = 1e6f;

    Sref = fmin_std(S[i1 + j1 * NZ + imax(k - 1, 1) * NXZ], S[i1 + j1 * NZ + imin(k, nyy - 1) * NXZ]);
    if ((tv < te + dx * Sref) && (te < tv + dz * Sref)) {
        ta = tev + te - tv;
        tb = tev - te + tv;
        t2D1 = ((tb * f->dz2i + ta * f->dx2i) +
                sqrtf(4.0f * Sref * Sref * (f->dz2i + f->dx2i) -
                      f->dz2i * f->dx2i * (ta - tb) * (ta - tb))) / (f->dz2i + f->dx2i);
    }

    Sref = fmin_std(S[i1 + imax(j - 1, 1) * NZ + k1 * NXZ], S[i1 + imin(j, nxx - 1) * NZ + k1 * NXZ]);
    if ((tv < tn + dy * Sref) && (tn < tv + dz * Sref)) {
        ta = tv - tn + tnv;
        tb = tn - tv + tnv;
        t2D2 = ((ta * f->dz2i + tb * f->dy2i) +
                sqrtf(4.0f * Sref * Sref * (f->dz2i + f->dy2i) -
                      f->dz2i * f->dy2i * (ta - tb) * (ta - tb))) / (f->dz2i + f->dy2i);
    }

    Sref = fmin_std(S[imax(i - 1, 1) + j1 * NZ + k1 * NXZ], S[imin(i, nzz - 1) + j1 * NZ + k1 * NXZ]);
    if ((te < tn + dy * Sref) && (tn < te + dx * Sref)) {
        ta = te - tn + ten;
        tb = tn - te + ten;
        t2D3 = ((ta * f->dx2i + tb * f->dy2i) +
                sqrtf(4.0f * Sref * Sref * (f->dx2i + f->dy2i) -
                      f->dx2i * f->dy2i * (ta - tb) * (ta - tb))) / (f->dx2i + f->dy2i);
    }

    t2D = fmin_std(t2D1, fmin_std(t2D2, t2D3));

    t3D = 1e6f;
    Sref = S[i1 + j1 * NZ + k1 * NXZ];

    ta = te - 0.5f * tn + 0.5f * ten - 0.5f * tv + 0.5f * tev - tnv + tnve;
    tb = tv - 0.5f * tn + 0.5f * tnv - 0.5f * te + 0.5f * tev - ten + tnve;
    tc = tn - 0.5f * te + 0.5f * ten - 0.5f * tv + 0.5f * tnv - tev + tnve;

    if (fmin_std(t1D, t2D) > fmax_std(tv, fmax_std(te, tn))) {
        t2 = 9.0f * Sref * Sref * f->dsum;
        t3 = f->dz2dx2 * (ta - tb) * (ta - tb) + f->dz2dy2 * (tb - tc) * (tb - tc) +
             f->dx2dy2 * (ta - tc) * (ta - tc);
        if (t2 >= t3) {
            t1 = tb * f->dz2i + ta * f->dx2i + tc * f->dy2i;
            t3D = (t1 + sqrtf(t2 - t3)) / f->dsum;
        }
    }

    T[ijk] = fmin_std(T[ijk], fmin_std(t1D, fmin_std(t2D, t3D)));
}

#define FSM_SET(f, tz, tx, ty, vz, vx, vy) \
    do { (f)->sgntz = tz; (f)->sgntx = tx; (f)->sgnty = ty; \
         (f)->sgnvz = vz; (f)->sgnvx = vx; (f)->sgnvy = vy; } while (0)

/* init_sweep (accurate_FSM.cpp:267-388).  The reverse loops start at
 * sid+1; for a source on the last node of an axis the reference then reads
 * one cell past the padded array (T[i - sgnt] with i = n+1).  The start is
 * clamped to n here (identical results wherever the reference is defined). */
static void fsm_init_sweep(fsm_state *f)
{
    int nxx = f->nxx, nyy = f->nyy, nzz = f->nzz;
    int sidx = imin(f->sidx, nxx - 3), sidy = imin(f->sidy, nyy - 3), sidz = imin(f->sidz, nzz - 3);
    int fwdx = f->sidx, fwdy = f->sidy, fwdz = f->sidz;

    FSM_SET(f, 1, 1, 1, 1, 1, 1);
    for (f->k = imax(1, fwdy); f->k < nyy; f->k++)
        for (f->j = imax(1, fwdx); f->j < nxx; f->j++)
            for (f->i = imax(1, fwdz); f->i < nzz; f->i++) fsm_inner(f);

    FSM_SET(f, -1, 1, 1, 0, 1, 1);
    for (f->k = imax(1, fwdy); f->k < nyy; f->k++)
        for (f->j = imax(1, fwdx); f->j < nxx; f->j++)
            for (f->i = sidz + 1; f->i >= 0; f->i--) fsm_inner(f);

    FSM_SET(f, 1, 1, -1, 1, 1, 0);
    for (f->k = sidy + 1; f->k >= 0; f->k--)
        for (f->j = imax(1, fwdx); f->j < nxx; f->j++)
            for (f->i = imax(1, fwdz); f->i < nzz; f->i++) fsm_inner(f);

    FSM_SET(f, -1, 1, -1, 0, 1, 0);
    for (f->k = sidy + 1; f->k >= 0; f->k--)
        for (f->j = imax(1, fwdx); f->j < nxx; f->j++)
            for (f->i = sidz + 1; f->i >= 0; f->i--) fsm_inner(f);

    FSM_SET(f, 1, -1, 1, 1, 0, 1);
    for (f->k = imax(1, fwdy); f->k < nyy; f->k++)
        for (f->j = sidx + 1; f->j >= 0; f->j--)
            for (f->i = imax(1, fwdz); f->i < nzz; f->i++) fsm_inner(f);

    FSM_SET(f, -1, -1, 1, 0, 0, 1);
    for (f->k = imax(1, fwdy); f->k < nyy; f->k++)
        for (f->j = sidx + 1; f->j >= 0; f->j--)
            for (f->i = sidz + 1; f->i >= 0; f->i--) fsm_inner(f);

    FSM_SET(f, 1, -1, -1, 1, 0, 0);
    for (f->k = sidy + 1; f->k >= 0; f->k--)
        for (f->j = sidx + 1; f->j >= 0; f->j--)
            for (f->i = imax(1, fwdz); f->i < nzz; f->i++) fsm_inner(f);

    FSM_SET(f, -1, -1, -1, 0, 0, 0);
    for (f->k = sidy + 1; f->k >= 0; f->k--)
        for (f->j = sidx + 1; f->j >= 0; f->j--)
            for (f->i = sidz + 1; f->i >= 0; f->i--) fsm_inner(f);
}

/* full_sweep (accurate_FSM.cpp:390-511) */
static void fsm_full_sweep(fsm_state *f)
{
    int nxx = f->nxx, nyy = f->nyy, nzz = f->nzz;

    FSM_SET(f, 1, 1, 1, 1, 1, 1);
    for (f->k = 1; f->k < nyy; f->k++)
        for (f->j = 1; f->j < nxx; f->j++)
            for (f->i = 1; f->i < nzz; f->i++) fsm_inner(f);

    FSM_SET(f, -1, 1, 1, 0, 1, 1);
    for (f->k = 1; f->k < nyy; f->k++)
        for (f->j = 1; f->j < nxx; f->j++)
            for (f->i = nzz - 2; f->i >= 0; f->i--) fsm_inner(f);

    FSM_SET(f, 1, 1, -1, 1, 1, 0);
    for (f->k = nyy - 2; f->k >= 0; f->k--)
        for (f->j = 1; f->j < nxx; f->j++)
            for (f->i = 1; f->i < nzz; f->i++) fsm_inner(f);

    FSM_SET(f, -1, 1, -1, 0, 1, 0);
    for (f->k = nyy - 2; f->k >= 0; f->k--)
        for (f->j = 1; f->j < nxx; f->j++)
            for (f->i = nzz - 2; f->i >= 0; f->i--) fsm_inner(f);

    FSM_SET(f, 1, -1, 1, 1, 0, 1);
    for (f->k = 1; f->k < nyy; f->k++)
        for (f->j = nxx - 2; f->j >= 0; f->j--)
            for (f->i = 1; f->i < nzz; f->i++) fsm_inner(f);

    FSM_SET(f, -1, -1, 1, 0, 0, 1);
    for (f->k = 1; f->k < nyy; f->k++)
        for (f->j = nxx - 2; f->j >= 0; f->j--)
            for (f->i = nzz - 2; f->i >= 0; f->i--) fsm_inner(f);

    FSM_SET(f, 1, -1, -1, 1, 0, 0);
    for (f->k = nyy - 2; f->k >= 0; f->k--)
        for (f->j = nxx - 2; f->j >= 0; f->j--)
            for (f->i = 1; f->i < nzz; f->i++) fsm_inner(f);

    FSM_SET(f, -1, -1, -1, 0, 0, 0);
    for (f->k = nyy - 2; f->k >= 0; f->k--)
        for (f->j = nxx - 2; f->j >= 0; f->j--)
            for (f->i = nzz - 2; f->i >= 0; f->i--) fsm_inner(f);
}

/* eikonal_equation (accurate_FSM.cpp:577-650) on prepared S, T */
static void fsm_solve(fsm_state *f, float sx, float sy, float sz, float *Tout)
{
    const int nb = 1;
    const float dx = f->dx, dy = f->dy, dz = f->dz;
    const int nxx = f->nxx, nzz = f->nzz;
    const long NZ = nzz, NXZ = (long)nxx * nzz;
    const long nPointsB = (long)f->nxx * f->nyy * f->nzz;
    float *T = f->T;
    const float *S = f->S;

    f->sidx = (int)(sx / dx) + nb;
    f->sidy = (int)(sy / dy) + nb;
    f->sidz = (int)(sz / dz) + nb;
    const int sidx = f->sidx, sidy = f->sidy, sidz = f->sidz;
    long sId = sidz + sidx * NZ + sidy * NXZ;

    for (long index = 0; index < nPointsB; index++) T[index] = 1e6f;

    /* 27-node analytic seed (accurate_FSM.cpp:596-630), transcribed line by
     * line: node at index offset (iz, ix, iy) receives S[sId] times the
     * distance of the node at offset (dz_, dx_, dy_).  The two xz diagonals
     * of lines 608-609 pair an index with the mirrored node's distance; that
     * quirk is kept. */
#define SEED(iz, ix, iy, oz, ox, oy) \
    T[sId + (iz) + (ix) * NZ + (iy) * NXZ] = S[sId] * sqrtf( \
        powf(((sidx - nb) + (ox)) * dx - sx, 2.0f) + \
        powf(((sidy - nb) + (oy)) * dy - sy, 2.0f) + \
        powf(((sidz - nb) + (oz)) * dz - sz, 2.0f))
    SEED(0, 0, 0, 0, 0, 0);                                  /* :596 */
    SEED(1, 0, 0, 1, 0, 0);     SEED(-1, 0, 0, -1, 0, 0);    /* :598-599 */
    SEED(0, 1, 0, 0, 1, 0);     SEED(0, -1, 0, 0, -1, 0);    /* :601-602 */
    SEED(0, 0, 1, 0, 0, 1);     SEED(0, 0, -1, 0, 0, -1);    /* :604-605 */
    SEED(1, 1, 0, 1, 1, 0);                                  /* :607 */
    SEED(1, -1, 0, -1, 1, 0);                                /* :608 */
    SEED(-1, 1, 0, 1, -1, 0);                                /* :609 */
    SEED(-1, -1, 0, -1, -1, 0);                              /* :610 */
    SEED(1, 0, 1, 1, 0, 1);     SEED(1, 0, -1, 1, 0, -1);    /* :612-613 */
    SEED(-1, 0, 1, -1, 0, 1);   SEED(-1, 0, -1, -1, 0, -1);  /* :614-615 */
    SEED(0, 1, 1, 0, 1, 1);     SEED(0, 1, -1, 0, 1, -1);    /* :617-618 */
    SEED(0, -1, 1, 0, -1, 1);   SEED(0, -1, -1, 0, -1, -1);  /* :619-620 */
    SEED(1, 1, 1, 1, 1, 1);     SEED(1, -1, 1, 1, -1, 1);    /* :622-623 */
    SEED(1, 1, -1, 1, 1, -1);   SEED(1, -1, -1, 1, -1, -1);  /* :624-625 */
    SEED(-1, 1, 1, -1, 1, 1);   SEED(-1, -1, 1, -1, -1, 1);  /* :627-628 */
    SEED(-1, 1, -1, -1, 1, -1); SEED(-1, -1, -1, -1, -1, -1);/* :629-630 */
#undef SEED

    f->dz2i = 1.0f / (dz * dz);
    f->dx2i = 1.0f / (dx * dx);
    f->dy2i = 1.0f / (dy * dy);
    f->dz2dx2 = f->dz2i * f->dx2i;
    f->dz2dy2 = f->dz2i * f->dy2i;
    f->dx2dy2 = f->dx2i * f->dy2i;
    f->dsum = f->dz2i + f->dx2i + f->dy2i;

    fsm_init_sweep(f);
    fsm_full_sweep(f);

    /* reduce_model (accurate_FSM.cpp:62-72) */
    for (int y = 0; y < f->ny; y++)
        for (int x = 0; x < f->nx; x++)
            for (int z = 0; z < f->nz; z++)
                Tout[z + (long)x * f->nz + (long)y * f->nx * f->nz] =
                    T[(z + 1) + (long)(x + 1) * NZ + (long)(y + 1) * NXZ];
}

static void fsm_prepare(fsm_state *f, const float *slowness, int nx, int ny,
                        int nz, float dx, float dy, float dz)
{
    memset(f, 0, sizeof(*f));
    f->nx = nx; f->ny = ny; f->nz = nz;
    f->dx = dx; f->dy = dy; f->dz = dz;
    f->nxx = nx + 2; f->nyy = ny + 2; f->nzz = nz + 2;
    long nB = (long)f->nxx * f->nyy * f->nzz;
    f->S = (float *)calloc(nB, sizeof(float));
    f->T = (float *)calloc(nB, sizeof(float));
    fsm_expand(f, slowness);
}

void or_fsm(const float *slowness, int nx, int ny, int nz,
            float dx, float dy, float dz, float sx, float sy, float sz,
            float *T)
{
    fsm_state f;
    fsm_prepare(&f, slowness, nx, ny, nz, dx, dy, dz);
    fsm_solve(&f, sx, sy, sz, T);
    free(f.S); free(f.T);
}

void or_fsm_batch(const float *slowness, int nx, int ny, int nz,
                  float dx, float dy, float dz, const float *sxyz, int nshot,
                  const float *rx, const float *ry, const float *rz, int nr,
                  float *picks, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
    {
        fsm_state f;
        fsm_prepare(&f, slowness, nx, ny, nz, dx, dy, dz);
        float *T = (float *)malloc(sizeof(float) * (long)nx * ny * nz);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int s = 0; s < nshot; s++) {
            fsm_solve(&f, sxyz[3 * s], sxyz[3 * s + 1], sxyz[3 * s + 2], T);
            or_picks(T, nx, ny, nz, dx, dy, dz, rx, ry, rz, nr, picks + (long)s * nr);
        }
        free(T); free(f.S); free(f.T);
    }
}

/* ------------------------------------------------------------------------ */
/* Least squares tomography (least_squares.cpp)                            */
/* ------------------------------------------------------------------------ */

/* expand_fdm (least_squares.cpp:11-67) */
void or_expand_fdm(const float *T, int nx, int ny, int nz, float *Tp)
{
    int nxx = nx + 2, nyy = ny + 2, nzz = nz + 2;
    const long NZ = nzz, NXZ = (long)nxx * nzz;
    for (int z = 1; z < nzz - 1; z++)
        for (int y = 1; y < nyy - 1; y++)
            for (int x = 1; x < nxx - 1; x++)
                Tp[z + x * NZ + y * NXZ] = T[(z - 1) + (long)(x - 1) * nz + (long)(y - 1) * nx * nz];
    for (int y = 1; y < nyy - 1; y++)
        for (int x = 1; x < nxx - 1; x++) {
            Tp[0 + x * NZ + y * NXZ] = T[0 + (long)(x - 1) * nz + (long)(y - 1) * nx * nz];
            Tp[(nzz - 1) + x * NZ + y * NXZ] = T[(nz - 1) + (long)(x - 1) * nz + (long)(y - 1) * nx * nz];
        }
    for (int z = 0; z < nzz; z++)
        for (int y = 1; y < nyy - 1; y++) {
            Tp[z + 0 * NZ + y * NXZ] = Tp[z + 1 * NZ + y * NXZ];
            Tp[z + (nxx - 1) * NZ + y * NXZ] = Tp[z + (nxx - 2) * NZ + y * NXZ];
        }
    for (int z = 0; z < nzz; z++)
        for (int x = 0; x < nxx; x++) {
            Tp[z + x * NZ + 0] = Tp[z + x * NZ + 1 * NXZ];
            Tp[z + x * NZ + (nyy - 1) * NXZ] = Tp[z + x * NZ + (nyy - 2) * NXZ];
        }
}

static int cmp_int(const void *a, const void *b)
{
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

static inline int iclamp(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* ray_tracing (least_squares.cpp:192-280) for one shot.  Where the
 * reference reads outside the padded T (ray leaving the model) the node
 * index is clamped into [1, n]; where its tomography cell would be outside
 * the inversion grid the cell coordinates are clamped into it.  A ray whose
 * gradient vanishes (norm not > 0) or that exceeds max_steps is reported
 * as diverged (the reference loops forever there). */
long or_ray_trace_shot(const float *Tp, int nx, int ny, int nz,
                       float dx, float dy, float dz,
                       int nxt, int nyt, int nzt,
                       float dxt, float dyt, float dzt,
                       float sx, float sy, float sz,
                       const float *rx, const float *ry, const float *rz,
                       int nr, int max_steps,
                       int *row_nnz, int *cols, float *vals, long cap,
                       float *illum)
{
    const int nxx = nx + 2, nzz = nz + 2;
    const long NZ = nzz, NXZ = (long)nxx * nzz;
    const long npts = (long)nx * ny * nz;
    int sIdz = (int)(sz / dzt);
    int sIdx = (int)(sx / dxt);
    int sIdy = (int)(sy / dyt);
    int sId = sIdz + sIdx * nzt + sIdy * nxt * nzt;
    float rayStep = 0.2f * dz;

    int *ray = (int *)malloc(sizeof(int) * (size_t)(max_steps > 0 ? max_steps : 1));
    long nnz = 0;

    for (int r = 0; r < nr; r++) {
        float zi = rz[r], xi = rx[r], yi = ry[r];
        int nsteps = 0;
        for (;;) {
            if (nsteps >= max_steps) { free(ray); return -1; }
            int i = iclamp((int)(zi / dz) + 1, 1, nz);
            int j = iclamp((int)(xi / dx) + 1, 1, nx);
            int k = iclamp((int)(yi / dy) + 1, 1, ny);

            float dTz = (Tp[(i + 1) + j * NZ + k * NXZ] - Tp[(i - 1) + j * NZ + k * NXZ]) / (2.0f * dz);
            float dTx = (Tp[i + (j + 1) * NZ + k * NXZ] - Tp[i + (j - 1) * NZ + k * NXZ]) / (2.0f * dx);
            float dTy = (Tp[i + j * NZ + (k + 1) * NXZ] - Tp[i + j * NZ + (k - 1) * NXZ]) / (2.0f * dy);

            float norm = sqrtf(dTx * dTx + dTy * dTy + dTz * dTz);
            if (!(norm > 0.0f) || !isfinite(norm)) { free(ray); return -1; }

            zi -= rayStep * dTz / norm;
            xi -= rayStep * dTx / norm;
            yi -= rayStep * dTy / norm;

            int im = iclamp((int)(zi / dzt), 0, nzt - 1);
            int jm = iclamp((int)(xi / dxt), 0, nxt - 1);
            int km = iclamp((int)(yi / dyt), 0, nyt - 1);

            int cell = im + jm * nzt + km * nxt * nzt;
            ray[nsteps++] = cell;
            long il = i + (long)j * nz + (long)k * nx * nz;
            if (illum && il >= 0 && il < npts) illum[il] += rayStep;
            if (cell == sId) break;
        }
        float finalDist = sqrtf(powf(zi - sz, 2.0f) + powf(xi - sx, 2.0f) + powf(yi - sy, 2.0f));

        qsort(ray, nsteps, sizeof(int), cmp_int);

        int cnt = 0;
        int current = ray[0];
        float distance = rayStep;
        for (int n = 0; n < nsteps; n++) {
            if (ray[n] == current) {
                distance += rayStep;
            } else {
                if (nnz >= cap) { free(ray); return -2; }
                cols[nnz] = current;
                vals[nnz] = (current == sId) ? finalDist : distance;
                nnz++; cnt++;
                distance = rayStep;
                current = ray[n];
            }
        }
        if (nnz >= cap) { free(ray); return -2; }
        cols[nnz] = current;
        vals[nnz] = (current == sId) ? finalDist : distance;
        nnz++; cnt++;
        row_nnz[r] = cnt;
    }
    free(ray);
    return nnz;
}

/* tk_reg_matrix stencil (least_squares.cpp:369-410), restated verbatim:
 * order 0 gives [1]; otherwise the repeated difference of [-1, 1]. */
void or_tk_stencil(int order, int *df)
{
    int elements = order + 1;
    if (order == 0) { df[0] = 1; return; }
    int *d = (int *)calloc(elements, sizeof(int));
    int *d1 = (int *)calloc(elements + 1, sizeof(int));
    int *d2 = (int *)calloc(elements + 1, sizeof(int));
    d[0] = -1; d[1] = 1;
    for (int index = 1; index < order; index++)
        for (int k = 0; k < elements; k++) {
            d2[k] = d[k];
            d1[k + 1] = d[k];
            d[k] = d1[k] - d2[k];
        }
    memcpy(df, d, sizeof(int) * elements);
    free(d); free(d1); free(d2);
}

/* sparse_cgls (least_squares.cpp:424-496), FP64 throughout.
 * Pinned: equals the normal-equation least-squares solution when iters
 * >= M on small full-rank systems (test_oracle_cgls.py). */
void or_cgls_f64(const long *rowptr, const int *col, const float *val,
                 int N, int M, const float *b, int iters, double *x)
{
    double *s = (double *)calloc(N, sizeof(double));
    double *q = (double *)calloc(N, sizeof(double));
    double *r = (double *)calloc(M, sizeof(double));
    double *p = (double *)calloc(M, sizeof(double));
    for (int i = 0; i < N; i++) s[i] = b[i];
    for (int i = 0; i < N; i++)
        for (long e = rowptr[i]; e < rowptr[i + 1]; e++) r[col[e]] += val[e] * s[i];
    for (int i = 0; i < M; i++) { p[i] = r[i]; x[i] = 0.0; }
    for (int i = 0; i < N; i++) {
        double acc = 0.0;
        for (long e = rowptr[i]; e < rowptr[i + 1]; e++) acc += val[e] * p[col[e]];
        q[i] = acc;
    }
    for (int it = 0; it < iters; it++) {
        double qTq = 0.0, rTr = 0.0, rd = 0.0;
        for (int k = 0; k < N; k++) qTq += q[k] * q[k];
        for (int k = 0; k < M; k++) rTr += r[k] * r[k];
        double a = rTr / qTq;
        for (int k = 0; k < M; k++) x[k] += a * p[k];
        for (int k = 0; k < N; k++) s[k] -= a * q[k];
        for (int k = 0; k < M; k++) rd += r[k] * r[k];
        for (int k = 0; k < M; k++) r[k] = 0.0;
        for (int i = 0; i < N; i++)
            for (long e = rowptr[i]; e < rowptr[i + 1]; e++) r[col[e]] += val[e] * s[i];
        rTr = 0.0;
        for (int k = 0; k < M; k++) rTr += r[k] * r[k];
        double bb = rTr / rd;
        for (int k = 0; k < M; k++) p[k] = r[k] + bb * p[k];
        for (int i = 0; i < N; i++) {
            double acc = 0.0;
            for (long e = rowptr[i]; e < rowptr[i + 1]; e++) acc += val[e] * p[col[e]];
            q[i] = acc;
        }
    }
    free(s); free(q); free(r); free(p);
}

/* The reference's FP32 version with COO scatter order (COO order equals
 * CSR order for the assembled system, least_squares.cpp:338-343,412-417). */
void or_cgls_ref_f32(const long *rowptr, const int *col, const float *val,
                     int N, int M, const float *b, int iters, float *x)
{
    float *s = (float *)calloc(N, sizeof(float));
    float *q = (float *)calloc(N, sizeof(float));
    float *r = (float *)calloc(M, sizeof(float));
    float *p = (float *)calloc(M, sizeof(float));
    float a, bb, qTq, rTr, rd;
    for (int i = 0; i < N; i++) s[i] = b[i];
    for (int i = 0; i < N; i++)
        for (long e = rowptr[i]; e < rowptr[i + 1]; e++) r[col[e]] += val[e] * s[i];
    for (int i = 0; i < M; i++) { p[i] = r[i]; x[i] = 0.0f; }
    for (int i = 0; i < N; i++)
        for (long e = rowptr[i]; e < rowptr[i + 1]; e++) q[i] += val[e] * p[col[e]];
    for (int it = 0; it < iters; it++) {
        qTq = 0.0f;
        for (int k = 0; k < N; k++) qTq += q[k] * q[k];
        rTr = 0.0f;
        for (int k = 0; k < M; k++) rTr += r[k] * r[k];
        a = rTr / qTq;
        for (int k = 0; k < M; k++) x[k] += a * p[k];
        for (int k = 0; k < N; k++) s[k] -= a * q[k];
        rd = 0.0f;
        for (int k = 0; k < M; k++) rd += r[k] * r[k];
        for (int k = 0; k < M; k++) r[k] = 0.0f;
        for (int i = 0; i < N; i++)
            for (long e = rowptr[i]; e < rowptr[i + 1]; e++) r[col[e]] += val[e] * s[i];
        rTr = 0.0f;
        for (int k = 0; k < M; k++) rTr += r[k] * r[k];
        bb = rTr / rd;
        for (int k = 0; k < M; k++) p[k] = r[k] + bb * p[k];
        for (int k = 0; k < N; k++) q[k] = 0.0f;
        for (int i = 0; i < N; i++)
            for (long e = rowptr[i]; e < rowptr[i + 1]; e++) q[i] += val[e] * p[col[e]];
    }
    free(s); free(q); free(r); free(p);
}

/* converged() (least_squares.cpp:301-307) */
float or_residual_norm(const float *dobs, const float *dcal, int n)
{
    float r = 0.0f;
    for (int i = 0; i < n; i++) r += powf(dobs[i] - dcal[i], 2.0f);
    return sqrtf(r);
}

/* gaussian kernel weights (gaussian.cpp:7-41) */
void or_gaussian_kernel(int samples, float stdv, float *kernel)
{
    int init = samples / 2;
    int nKernel = samples * samples * samples;
    float pi = 4.0f * atanf(1.0f);
    for (int i = 0; i < nKernel; i++) kernel[i] = 0.0f;
    int mid = (int)(samples / 2);
    kernel[mid + mid * samples + mid * samples * samples] = 1.0f;
    if (stdv != 0.0f) {
        float sum = 0.0f;
        for (int y = -init; y <= init; y++)
            for (int x = -init; x <= init; x++)
                for (int z = -init; z <= init; z++) {
                    int index = (z + init) + (x + init) * samples + (y + init) * samples * samples;
                    kernel[index] = 1.0f / (2.0f * pi * stdv * stdv) *
                                    expf(-((x * x + y * y + z * z) / (2.0f * stdv * stdv)));
                    sum += kernel[index];
                }
        for (int i = 0; i < nKernel; i++) kernel[i] /= sum;
    }
}

/* gaussian (gaussian.cpp:5-73); samples must be odd as the reference
 * parameter file requires. */
void or_gaussian(const float *in, float *out, int nx, int ny, int nz,
                 int samples, float stdv)
{
    int init = samples / 2;
    long nPoints = (long)nx * ny * nz;
    float *kernel = (float *)malloc(sizeof(float) * samples * samples * samples);
    or_gaussian_kernel(samples, stdv, kernel);
    for (long i = 0; i < nPoints; i++) out[i] = in[i];
    for (int k = init; k < ny - init; k++)
        for (int j = init; j < nx - init; j++)
            for (int i = init; i < nz - init; i++) {
                float accum = 0.0f;
                for (int yk = 0; yk < samples; yk++)
                    for (int xk = 0; xk < samples; xk++)
                        for (int zk = 0; zk < samples; zk++) {
                            int index = zk + xk * samples + yk * samples * samples;
                            long partial = (i - init + zk) + (long)(j - init + xk) * nz +
                                           (long)(k - init + yk) * nx * nz;
                            accum += in[partial] * kernel[index];
                        }
                out[i + (long)j * nz + (long)k * nx * nz] = accum;
            }
    free(kernel);
}

/* model_update (least_squares.cpp:498-570).  The +1 corners of the coarse
 * lookup are clamped into the inversion grid where the reference would
 * read past it. */
void or_model_update(float *slowness, const float *xdm,
                     int nx, int ny, int nz, float dx, float dy, float dz,
                     int nxt, int nyt, int nzt,
                     float dxt, float dyt, float dzt,
                     int smooth, int samples, float stdv,
                     float *dm, float *velocity)
{
    long nPoints = (long)nx * ny * nz;
    for (long index = 0; index < nPoints; index++) {
        int k = (int)(index / ((long)nx * nz));
        int j = (int)(index - (long)k * nx * nz) / nz;
        int i = (int)(index - (long)j * nz - (long)k * nx * nz);

        float x = j * dx;
        float y = k * dy;
        float z = i * dz;

        float x0 = floorf(x / dxt) * dxt;
        float y0 = floorf(y / dyt) * dyt;
        float z0 = floorf(z / dzt) * dzt;
        float x1 = floorf(x / dxt) * dxt + dxt;
        float y1 = floorf(y / dyt) * dyt + dyt;
        float z1 = floorf(z / dzt) * dzt + dzt;

        dm[index] = 0.0f;
        if ((i >= 11) && (i < nz - 2) && (j >= 15) && (j < nx - 16) && (k >= 15) && (k < ny - 16)) {
            int idz = (int)(z / dzt);
            int idx = (int)(x / dxt);
            int idy = (int)(y / dyt);
            long oz = (idz + 1 < nzt) ? 1 : 0;
            long ox = (idx + 1 < nxt) ? nzt : 0;
            long oy = (idy + 1 < nyt) ? (long)nxt * nzt : 0;
            long ind_m = idz + (long)idx * nzt + (long)idy * nxt * nzt;
            float c000 = xdm[ind_m];
            float c001 = xdm[ind_m + oz];
            float c100 = xdm[ind_m + ox];
            float c101 = xdm[ind_m + oz + ox];
            float c010 = xdm[ind_m + oy];
            float c011 = xdm[ind_m + oz + oy];
            float c110 = xdm[ind_m + ox + oy];
            float c111 = xdm[ind_m + oz + ox + oy];
            dm[i + (long)j * nz + (long)k * nx * nz] =
                or_trilinear(c000, c001, c100, c101, c010, c011, c110, c111,
                             x0, x1, y0, y1, z0, z1, x, y, z);
        }
    }
    float *update = (float *)calloc(nPoints, sizeof(float));
    if (smooth) {
        or_gaussian(dm, update, nx, ny, nz, samples, stdv);
        for (long index = 0; index < nPoints; index++) dm[index] = update[index];
    }
    for (long index = 0; index < nPoints; index++) {
        int k = (int)(index / ((long)nx * nz));
        int j = (int)(index - (long)k * nx * nz) / nz;
        int i = (int)(index - (long)j * nz - (long)k * nx * nz);
        if ((i >= 11) && (i < nz - 2) && (j >= 15) && (j < nx - 15) && (k >= 15) && (k < ny - 15))
            slowness[index] += dm[index];
        update[index] = 1.0f / slowness[index];
    }
    memcpy(velocity, update, sizeof(float) * nPoints);
    free(update);
}


/* ====================================================================== */
/* data synthesis                                                          */
/* ====================================================================== */

/* dampers (acoustic.cpp:100-137): d1[i] = exp(-(factor (nb - i))^2);
 * d2[a + b nb] = d1[a] + d1[b] - 1; d3[i + j nb + k nb^2] = d2(i,j) +
 * d2(j,k) + d2(i,k) before the -1 (summed in that order) - 5. */
void or_acoustic_dampers(float *d1, float *d2, float *d3, float factor, int nb)
{
    memset(d1, 0, sizeof(float) * nb);
    memset(d2, 0, sizeof(float) * nb * nb);
    memset(d3, 0, sizeof(float) * (size_t)nb * nb * nb);
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
                float *c = d3 + i + j * nb + (size_t)k * nb * nb;
                *c += d2[i + j * nb];
                *c += d2[j + k * nb];
                *c += d2[i + k * nb];
            }
    for (int q = 0; q < nb * nb; q++) d2[q] -= 1.0f;
    for (long q = 0; q < (long)nb * nb * nb; q++) d3[q] -= 5.0f;
}

/* depth of index i into the absorbing layer of an axis of length n (0 at
 * the outer face side ... nb-1 next to the model), -1 outside the layer */
static int layer_index(int i, int n, int nb)
{
    if (i < nb) return i;
    if (i >= n - nb) return nb - (i - (n - nb)) - 1;
    return -1;
}

/* the damper of cell (z i, x j, y k): 1 in the model, the 1D, 2D or 3D
 * profile where one, two or three axes are inside the layer
 * (the case list of acoustic.cpp:192-302) */
static float damper_at(int i, int j, int k, int nb, int nxx, int nyy, int nzz,
                       const float *d1, const float *d2, const float *d3)
{
    const int lz = layer_index(i, nzz, nb), lx = layer_index(j, nxx, nb), ly = layer_index(k, nyy, nb);
    const int n = (lz >= 0) + (lx >= 0) + (ly >= 0);
    if (n == 0) return 1.0f;
    if (n == 1) return d1[lz >= 0 ? lz : (lx >= 0 ? lx : ly)];
    if (n == 2) {
        if (lz < 0) return d2[lx + ly * nb];
        if (lx < 0) return d2[lz + ly * nb];
        return d2[lz + lx * nb];
    }
    return d3[lz + lx * nb + (size_t)ly * nb * nb];
}

void or_acoustic_shot(const float *vel, int nx, int ny, int nz, float dx, float dy, float dz,
                      int nb, float factor, const float *wavelet, int nt, float dt,
                      float sx, float sy, float sz, const float *rx, const float *ry,
                      const float *rz, int nrec, float *seismogram, float *ufut)
{
    const int nxx = nx + 2 * nb, nyy = ny + 2 * nb, nzz = nz + 2 * nb;
    const long N = (long)nxx * nyy * nzz, NZ = nzz, NXZ = (long)nxx * nzz;
    float *V = (float *)malloc(sizeof(float) * N);
    float *P0 = (float *)calloc(N, sizeof(float));   /* U_pas */
    float *P1 = (float *)calloc(N, sizeof(float));   /* U_pre */
    float *P2 = (float *)calloc(N, sizeof(float));   /* U_fut */
    float *d1 = (float *)malloc(sizeof(float) * nb), *d2 = (float *)malloc(sizeof(float) * nb * nb);
    float *d3 = (float *)malloc(sizeof(float) * (size_t)nb * nb * nb);
    /* expand (acoustic.cpp:16-85): edge replication = clamped index */
    for (int k = 0; k < nyy; k++)
        for (int j = 0; j < nxx; j++)
            for (int i = 0; i < nzz; i++) {
                const int zi = imin(imax(i - nb, 0), nz - 1), xi = imin(imax(j - nb, 0), nx - 1),
                          yi = imin(imax(k - nb, 0), ny - 1);
                V[i + j * NZ + k * NXZ] = vel[zi + (long)xi * nz + (long)yi * nx * nz];
            }
    or_acoustic_dampers(d1, d2, d3, factor, nb);
    const long sId = ((int)(sz / dz) + nb) + (long)((int)(sx / dx) + nb) * NZ +
                     (long)((int)(sy / dy) + nb) * NXZ;
    const float cz = 5040.0f * (dz * dz), cx = 5040.0f * (dx * dx), cy = 5040.0f * (dy * dy);
    for (int t = 0; t < nt; t++) {
        P1[sId] += wavelet[t];                                   /* applyWavelet */
        for (int k = 4; k < nyy - 4; k++)                        /* wavePropagation */
            for (int j = 4; j < nxx - 4; j++)
                for (int i = 4; i < nzz - 4; i++) {
                    const long c = i + j * NZ + k * NXZ;
#define ST8(s) (-9.0f * (P1[c - 4 * (s)] + P1[c + 4 * (s)]) + 128.0f * (P1[c - 3 * (s)] + P1[c + 3 * (s)]) \
                - 1008.0f * (P1[c - 2 * (s)] + P1[c + 2 * (s)]) + 8064.0f * (P1[c - (s)] + P1[c + (s)]) \
                - 14350.0f * P1[c])
                    const float px = ST8(NZ) / cx;
                    const float py = ST8(NXZ) / cy;
                    const float pz = ST8(1) / cz;
#undef ST8
                    const float d = damper_at(i, j, k, nb, nxx, nyy, nzz, d1, d2, d3);
                    const float lap = px + py + pz;
                    P2[c] = V[c] * V[c] * (dt * dt) * lap + 2.0f * P1[c] - P0[c];
                    P2[c] *= d;
                    P1[c] *= d;
                    P0[c] *= d;
                }
        float *tmp = P0; P0 = P1; P1 = P2; P2 = tmp;             /* wavefieldUpdate */
        memcpy(P2, P1, sizeof(float) * N);                       /* U_fut keeps its value */
        for (int r = 0; r < nrec; r++) {                         /* buildSeismogram */
            const long c = ((int)(rz[r] / dz) + nb) + (long)((int)(rx[r] / dx) + nb) * NZ +
                           (long)((int)(ry[r] / dy) + nb) * NXZ;
            seismogram[t + (long)r * nt] = P1[c];
        }
    }
    if (ufut) memcpy(ufut, P1, sizeof(float) * N);
    free(V); free(P0); free(P1); free(P2); free(d1); free(d2); free(d3);
}

void or_auto_pick(const float *seis, int ntraces, int nt, float dt, float window,
                  float amp_cut, float shift, float *picks)
{
    const int iw = (int)(window / dt) + 1;
    const long n = (long)ntraces * nt;
    float mx = seis[0];
    for (long q = 1; q < n; q++) if (seis[q] > mx) mx = seis[q];
    const float inv = 1.0f / mx;
    float *x = (float *)malloc(sizeof(float) * nt);
    float *S = (float *)malloc(sizeof(float) * nt);
    for (int tr = 0; tr < ntraces; tr++) {
        for (int q = 0; q < nt; q++) { x[q] = seis[(long)tr * nt + q] * inv; S[q] = 0.0f; }
        picks[tr] = 0.0f;
        for (int time = iw; time < nt - iw; time++) {
            float A = 0.0f, B = 0.0f;
            for (int k = 0; k < iw; k++) {
                A += x[time - k] + 1e-15f;
                B += x[time + k] + 1e-15f;
            }
            /* powf(iw + time, 2.0f): the square, as optimised builds fold it */
            const float w = (float)(iw + time);
            S[time] = fabsf(B / A * (B - A));
            S[time] *= (B + A) * (B - A) / (w * w);
        }
        float am = S[0];
        for (int q = 1; q < nt; q++) if (S[q] > am) am = S[q];
        const float ia = 1.0f / am;
        for (int q = 0; q < nt; q++) {
            if (S[q] * ia > amp_cut) { picks[tr] = q * dt + shift; break; }
        }
    }
    free(x); free(S);
}


/* ====================================================================== */
/* eikonal: Podvin & Lecomte (classic.cpp)                                 */
/* ====================================================================== */

/* Offsets (dz, dx, dy) of the operators of classic.cpp:301-1773, in the
 * reference's order.  S cells are the voxels whose lower corner is the
 * given node. */
typedef struct { signed char t[3], s[4][3]; } PodHead;          /* :301-341 */
typedef struct { signed char t[3], s[2][3]; } PodDiff;          /* :343-395 */
typedef struct { signed char t[3], s[3]; } PodCorner;           /* :397-429 */
typedef struct { signed char s[2][3], tt[2][2][3]; } PodQuad;   /* :431-765: Sref, (T1, T2) x 2 */
typedef struct { signed char s[3], f[3][4][3]; } PodOct;        /* :767-1773: Sref, faces (M,N,P,Q) */

static const PodHead POD_HEAD[6] = {
    {{0,-1,0}, {{0,-1,0}, {-1,-1,0}, {0,-1,-1}, {-1,-1,-1}}},
    {{0,1,0}, {{0,0,0}, {-1,0,0}, {0,0,-1}, {-1,0,-1}}},
    {{0,0,-1}, {{0,0,-1}, {0,-1,-1}, {-1,0,-1}, {-1,-1,-1}}},
    {{0,0,1}, {{0,0,0}, {-1,0,0}, {0,-1,0}, {-1,-1,0}}},
    {{-1,0,0}, {{-1,0,0}, {-1,-1,0}, {-1,0,-1}, {-1,-1,-1}}},
    {{1,0,0}, {{0,0,0}, {0,-1,0}, {0,0,-1}, {0,-1,-1}}},
};
static const PodDiff POD_DIFF[12] = {
    {{-1,-1,0}, {{-1,-1,0}, {-1,-1,-1}}}, {{-1,1,0}, {{-1,0,0}, {-1,0,-1}}},
    {{1,-1,0}, {{0,-1,0}, {0,-1,-1}}},    {{1,1,0}, {{0,0,0}, {0,0,-1}}},
    {{-1,0,-1}, {{-1,0,-1}, {-1,-1,-1}}}, {{-1,0,1}, {{-1,0,0}, {-1,-1,0}}},
    {{1,0,-1}, {{0,0,-1}, {0,-1,-1}}},    {{1,0,1}, {{0,0,0}, {0,-1,0}}},
    {{0,-1,-1}, {{0,-1,-1}, {-1,-1,-1}}}, {{0,-1,1}, {{0,-1,0}, {-1,-1,0}}},
    {{0,1,-1}, {{0,0,-1}, {-1,0,-1}}},    {{0,1,1}, {{0,0,0}, {-1,0,0}}},
};
static const PodCorner POD_CORNER[8] = {
    {{-1,-1,-1}, {-1,-1,-1}}, {{-1,-1,1}, {-1,-1,0}}, {{1,-1,-1}, {0,-1,-1}},
    {{1,-1,1}, {0,-1,0}},     {{-1,1,-1}, {-1,0,-1}}, {{-1,1,1}, {-1,0,0}},
    {{1,1,-1}, {0,0,-1}},     {{1,1,1}, {0,0,0}},
};
static const PodQuad POD_QUAD[12] = {
    {{{-1,-1,0}, {-1,-1,-1}}, {{{0,-1,0}, {-1,-1,0}}, {{-1,0,0}, {-1,-1,0}}}},
    {{{0,-1,0}, {0,-1,-1}}, {{{0,-1,0}, {1,-1,0}}, {{1,0,0}, {1,-1,0}}}},
    {{{0,0,0}, {0,0,-1}}, {{{1,0,0}, {1,1,0}}, {{0,1,0}, {1,1,0}}}},
    {{{-1,0,0}, {-1,0,-1}}, {{{0,1,0}, {-1,1,0}}, {{-1,0,0}, {-1,1,0}}}},
    {{{-1,0,-1}, {-1,-1,-1}}, {{{0,0,-1}, {-1,0,-1}}, {{-1,0,0}, {-1,0,-1}}}},
    {{{0,0,-1}, {0,-1,-1}}, {{{0,0,-1}, {1,0,-1}}, {{1,0,0}, {1,0,-1}}}},
    {{{0,0,0}, {0,-1,0}}, {{{1,0,0}, {1,0,1}}, {{0,0,1}, {1,0,1}}}},
    {{{-1,0,0}, {-1,-1,0}}, {{{0,0,1}, {-1,0,1}}, {{-1,0,0}, {-1,0,1}}}},
    {{{0,-1,-1}, {-1,-1,-1}}, {{{0,-1,0}, {0,-1,-1}}, {{0,0,-1}, {0,-1,-1}}}},
    {{{0,-1,0}, {-1,-1,0}}, {{{0,-1,0}, {0,-1,1}}, {{0,0,1}, {0,-1,1}}}},
    {{{0,0,0}, {-1,0,0}}, {{{0,0,1}, {0,1,1}}, {{0,1,0}, {0,1,1}}}},
    {{{0,0,-1}, {-1,0,-1}}, {{{0,1,0}, {0,1,-1}}, {{0,0,-1}, {0,1,-1}}}},
};
static const PodOct POD_OCT[8] = {
    {{-1,-1,-1}, {{{-1,-1,-1}, {-1,-1,0}, {-1,0,-1}, {-1,0,0}}, {{-1,-1,-1}, {-1,-1,0}, {0,-1,-1}, {0,-1,0}}, {{-1,-1,-1}, {0,-1,-1}, {-1,0,-1}, {0,0,-1}}}},
    {{-1,0,-1}, {{{-1,1,-1}, {-1,0,-1}, {-1,1,0}, {-1,0,0}}, {{-1,1,-1}, {-1,1,0}, {0,1,-1}, {0,1,0}}, {{-1,1,-1}, {-1,0,-1}, {0,1,-1}, {0,0,-1}}}},
    {{-1,0,0}, {{{-1,1,1}, {-1,1,0}, {-1,0,1}, {-1,0,0}}, {{-1,1,1}, {0,1,1}, {-1,1,0}, {0,1,0}}, {{-1,1,1}, {-1,0,1}, {0,1,1}, {0,0,1}}}},
    {{-1,-1,0}, {{{-1,-1,1}, {-1,0,1}, {-1,-1,0}, {-1,0,0}}, {{-1,-1,1}, {-1,-1,0}, {0,-1,1}, {0,-1,0}}, {{-1,-1,1}, {0,-1,1}, {-1,0,1}, {0,0,1}}}},
    {{0,-1,-1}, {{{1,-1,-1}, {1,0,-1}, {1,-1,0}, {1,0,0}}, {{1,-1,-1}, {1,-1,0}, {0,-1,-1}, {0,-1,0}}, {{1,-1,-1}, {0,-1,-1}, {1,0,-1}, {0,0,-1}}}},
    {{0,0,-1}, {{{1,1,-1}, {1,1,0}, {1,0,-1}, {1,0,0}}, {{1,1,-1}, {0,1,-1}, {1,1,0}, {0,1,0}}, {{1,1,-1}, {1,0,-1}, {0,1,-1}, {0,0,-1}}}},
    {{0,-1,0}, {{{1,-1,1}, {1,-1,0}, {1,0,1}, {1,0,0}}, {{1,-1,1}, {0,-1,1}, {1,-1,0}, {0,-1,0}}, {{1,-1,1}, {1,0,1}, {0,-1,1}, {0,0,1}}}},
    {{0,0,0}, {{{1,1,1}, {1,0,1}, {1,1,0}, {1,0,0}}, {{1,1,1}, {1,1,0}, {0,1,1}, {0,1,0}}, {{1,1,1}, {0,1,1}, {1,0,1}, {0,0,1}}}},
};

static inline float pod_min(float a, float b) { return b < a ? b : a; }   /* std::min */

/* the lowest of T and every operator's arrival at cell c (classic.cpp:297-1773) */
static float pod_update(const float *T, const float *S, long c, long NZ, long NXZ, float h,
                        float sqrt2, float sqrt3)
{
#define OFF(o) ((long)(o)[0] + (long)(o)[1] * NZ + (long)(o)[2] * NXZ)
    float lowest = T[c], Tijk;
    for (int q = 0; q < 6; q++) {          /* 1D head waves */
        const PodHead *o = &POD_HEAD[q];
        const float m = pod_min(S[c + OFF(o->s[0])], pod_min(S[c + OFF(o->s[1])],
                        pod_min(S[c + OFF(o->s[2])], S[c + OFF(o->s[3])])));
        Tijk = T[c + OFF(o->t)] + h * m;
        if (Tijk < lowest) lowest = Tijk;
    }
    for (int q = 0; q < 12; q++) {         /* edge diffractions */
        const PodDiff *o = &POD_DIFF[q];
        Tijk = T[c + OFF(o->t)] + h * sqrt2 * pod_min(S[c + OFF(o->s[0])], S[c + OFF(o->s[1])]);
        if (Tijk < lowest) lowest = Tijk;
    }
    for (int q = 0; q < 8; q++) {          /* corner diffractions */
        const PodCorner *o = &POD_CORNER[q];
        Tijk = T[c + OFF(o->t)] + h * sqrt3 * S[c + OFF(o->s)];
        if (Tijk < lowest) lowest = Tijk;
    }
    for (int q = 0; q < 12; q++) {         /* 2D operators */
        const PodQuad *o = &POD_QUAD[q];
        const float Sref = pod_min(S[c + OFF(o->s[0])], S[c + OFF(o->s[1])]);
        for (int e = 0; e < 2; e++) {
            const float T1 = T[c + OFF(o->tt[e][0])], T2 = T[c + OFF(o->tt[e][1])];
            if ((T1 - T2) > 0.0f && (T1 - T2) < h * Sref / sqrt2) {
                Tijk = T1 + sqrtf(h * h * Sref * Sref - (T1 - T2) * (T1 - T2));
                if (Tijk < lowest) lowest = Tijk;
            }
        }
    }
    for (int q = 0; q < 8; q++) {          /* 3D operators */
        const PodOct *o = &POD_OCT[q];
        const float Sref = S[c + OFF(o->s)];
        const float hs2 = h * h * Sref * Sref;
        for (int f = 0; f < 3; f++) {
            const float M = T[c + OFF(o->f[f][0])], N = T[c + OFF(o->f[f][1])],
                        P = T[c + OFF(o->f[f][2])], Q = T[c + OFF(o->f[f][3])];
            if ((M <= N) && (M <= P) &&
                ((2.0f * (P - M) * (P - M) + (N - M) * (N - M)) <= hs2) &&
                ((2.0f * (N - M) * (N - M) + (P - M) * (P - M)) <= hs2) &&
                ((N - M) * (N - M) + (P - M) * (P - M) + (N - M) * (P - M) >= 0.5f * hs2)) {
                Tijk = N + P - M + sqrtf(hs2 - (N - M) * (N - M) - (P - M) * (P - M));
                if (Tijk < lowest) lowest = Tijk;
            }
            if ((N <= Q) && (P <= Q) &&
                ((Q - N) * (Q - N) + (Q - P) * (Q - P) + (Q - N) * (Q - P) <= 0.5f * hs2)) {
                Tijk = Q + sqrtf(hs2 - (Q - N) * (Q - N) - (Q - P) * (Q - P));
                if (Tijk < lowest) lowest = Tijk;
            }
            if ((N - M >= 0) && (N - M <= Q - N) &&
                (2 * (Q - N) * (Q - N) + (N - M) * (N - M) <= hs2)) {
                Tijk = Q + sqrtf(hs2 - (Q - N) * (Q - N) - (N - M) * (N - M));
                if (Tijk < lowest) lowest = Tijk;
            }
            if ((P - M >= 0) && (P - M <= Q - P) &&
                (2 * (Q - P) * (Q - P) + (P - M) * (P - M) <= hs2)) {
                Tijk = Q + sqrtf(hs2 - (Q - P) * (Q - P) - (P - M) * (P - M));
                if (Tijk < lowest) lowest = Tijk;
            }
        }
    }
#undef OFF
    return lowest;
}

void or_podvin(const float *slowness, int nx, int ny, int nz, float dx, float dy, float dz,
               float sx, float sy, float sz, float *Tout)
{
    const int nxx = nx + 2, nyy = ny + 2, nzz = nz + 2;
    const long NZ = nzz, NXZ = (long)nxx * nzz, N = NXZ * nyy;
    float *S = (float *)malloc(sizeof(float) * N), *T = (float *)malloc(sizeof(float) * N);
    float *nT = (float *)malloc(sizeof(float) * N);
    unsigned char *K = (unsigned char *)calloc(N, 1), *nK = (unsigned char *)calloc(N, 1);
    /* expand_model (classic.cpp:8-60): replicate the edge cells outwards */
    for (int k = 0; k < nyy; k++)
        for (int j = 0; j < nxx; j++)
            for (int i = 0; i < nzz; i++)
                S[i + j * NZ + k * NXZ] = slowness[imin(imax(i - 1, 0), nz - 1) +
                                                   (long)imin(imax(j - 1, 0), nx - 1) * nz +
                                                   (long)imin(imax(k - 1, 0), ny - 1) * nx * nz];
    /* seeding (classic.cpp:186-246) */
    const int sidx = (int)(sx / dx) + 1, sidy = (int)(sy / dy) + 1, sidz = (int)(sz / dz) + 1;
    const long sId = sidz + sidx * NZ + sidy * NXZ;
    for (long q = 0; q < N; q++) T[q] = nT[q] = 1e6f;
    {
        const float gx = floorf(sx / dx) * dx, gy = floorf(sy / dy) * dy, gz = floorf(sz / dz) * dz;
        const float ex = sx - gx, ey = sy - gy, ez = sz - gz;
        const float dist = sqrtf(ex * ex + ey * ey + ez * ez);
        T[sId] = dist * S[sId];
        nT[sId] = T[sId];
    }
    for (int a = -1; a <= 1; a++)
        for (int b = -1; b <= 1; b++)
            for (int e = -1; e <= 1; e++)
                if (a || b || e) K[sId + a + b * NZ + e * NXZ] = 1;
    /* iteration count: the largest source-to-corner distance in nodes (:248-273) */
    int nIt = 0;
    {
        const float fx = (float)sidx, fy = (float)sidy, fz = (float)sidz;
        const float gx = (float)(nxx - sidx), gy = (float)(nyy - sidy), gz = (float)(nzz - sidz);
        const float c[8][3] = {{fx, fy, fz}, {gx, fy, fz}, {fx, gy, fz}, {fx, fy, gz},
                               {fx, gy, gz}, {gx, fy, gz}, {gx, gy, fz}, {gx, gy, gz}};
        for (int q = 0; q < 8; q++) {
            const int aux = (int)sqrtf(c[q][0] * c[q][0] + c[q][1] * c[q][1] + c[q][2] * c[q][2]);
            if (aux > nIt) nIt = aux;
        }
    }
    const float h = dx, sqrt2 = sqrtf(2.0f), sqrt3 = sqrtf(3.0f);
    (void)dy; (void)dz;
    for (int it = 0; it < nIt; it++) {
        for (int k = 1; k < nyy - 1; k++)
            for (int j = 1; j < nxx - 1; j++)
                for (int i = 1; i < nzz - 1; i++) {
                    const long c = i + j * NZ + k * NXZ;
                    if (!K[c]) continue;
                    const float lowest = pod_update(T, S, c, NZ, NXZ, h, sqrt2, sqrt3);
                    if (lowest == T[c]) K[c] = 0;
                    nT[c] = lowest;
                }
        memset(nK, 0, N);
        for (int k = 1; k < nyy - 1; k++)
            for (int j = 1; j < nxx - 1; j++)
                for (int i = 1; i < nzz - 1; i++) {
                    const long c = i + j * NZ + k * NXZ;
                    if (!K[c]) continue;
                    for (int a = -1; a <= 1; a++)
                        for (int b = -1; b <= 1; b++)
                            for (int e = -1; e <= 1; e++)
                                if (a || b || e) nK[c + a + b * NZ + e * NXZ] = 1;
                }
        memcpy(T, nT, sizeof(float) * N);
        memcpy(K, nK, N);
    }
    /* reduce_model (:62-72) */
    for (int k = 0; k < ny; k++)
        for (int j = 0; j < nx; j++)
            for (int i = 0; i < nz; i++)
                Tout[i + (long)j * nz + (long)k * nx * nz] = T[(i + 1) + (j + 1) * NZ + (k + 1) * NXZ];
    free(S); free(T); free(nT); free(K); free(nK);
}
