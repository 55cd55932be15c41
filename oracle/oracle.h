/*
 * oracle.h -- CPU restatement of the first-arrival tomography path of
 * phbastosa/seismic_tomography_3D, used ONLY as test infrastructure.
 *
 * This library is the checker: tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  No product code links it.
 *
 * Volumes are FP32, z fastest: idx = z + x*nz + y*nx*nz (the reference's
 * layout, e.g. src/eikonal/block_FIM/block_FIM.cpp:17).
 *
 * Pinning status of each function is listed in its comment; the tests that
 * pin it live in tests/test_oracle_*.py.
 */
#ifndef ST_ORACLE_H
#define ST_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define OR_INF 1e6f

/* ---- eikonal: block-FIM discretisation (cuda_kernel.cu, block_FIM.cpp) ---- */

/* pad extent to a multiple of B (block_FIM.cpp:144-150) */
int or_pad_extent(int n, int B);

/* expand_model (block_FIM.cpp:8-57): replicate last plane into the pad.
 * S has extents (nzz, nxx, nyy), z fastest. */
void or_pad_replicate(const float *s, int nx, int ny, int nz,
                      float *S, int nxx, int nyy, int nzz);

/* get_time_eikonal (cuda_kernel.cu:217-248), FP32 and FP64 */
float  or_godunov_f32(float a, float b, float c, float h, float s);
double or_godunov_f64(double a, double b, double c, double h, double s);

/* apply_source_time (block_FIM.cpp:305-317): seed node and its time */
void or_source_seed(const float *S, int nxx, int nyy, int nzz,
                    float dx, float dy, float dz,
                    float sx, float sy, float sz, long *sid, float *t0);

/* FP64 fixed point of the block-FIM discretisation on the padded domain
 * (Gauss-Seidel sweeps in 8 orders until max change < tol).  T has the
 * padded extents.  Returns the number of 8-sweep passes, or -1 if
 * max_pass was reached. */
int or_eikonal_fp64(const float *S, int nxx, int nyy, int nzz, double h,
                    long sid, double t0, double tol, int max_pass, double *T);

/* Restatement of the reference's host-driven block FIM
 * (cuda_kernel.cu:17-215 + kernels :250-619), FP32, 4^3 blocks, Jacobi
 * semantics inside a block, ping-pong between rounds.  T has padded extents
 * (z fastest).  stats[0]=rounds, stats[1]=block visits.  Returns 0. */
int or_block_fim_ref(const float *S, int nxx, int nyy, int nzz, float h,
                     long sid, float t0, int n_iter, float *T, long *stats);

/* crop padded (nzz,nxx,nyy) -> (nz,nx,ny) (block_FIM.cpp:59-69) */
void or_crop_f32(const float *Tp, int nxx, int nyy, int nzz,
                 float *T, int nx, int ny, int nz);
void or_crop_f64(const double *Tp, int nxx, int nyy, int nzz,
                 double *T, int nx, int ny, int nz);

/* trilinear (trilinear.cpp:3-20) */
float or_trilinear(float c000, float c001, float c100, float c101,
                   float c010, float c011, float c110, float c111,
                   float x0, float x1, float y0, float y1, float z0, float z1,
                   float x, float y, float z);

/* write_first_arrival picks (block_FIM.cpp:228-262).  Where the reference
 * reads past the last node (weight exactly 0), the +1 corner index is
 * clamped into the volume. */
void or_picks(const float *T, int nx, int ny, int nz,
              float dx, float dy, float dz,
              const float *rx, const float *ry, const float *rz, int nr,
              float *out);

/* ---- eikonal: Noble-Gesret-Belayouni FSM (accurate_FSM.cpp) ---- */
/* Full Accurate_FSM::eikonal_equation (accurate_FSM.cpp:577-650):
 * pad by 1, 27-node seed, init_sweep, full_sweep, crop.  Output (nz,nx,ny). */
void or_fsm(const float *slowness, int nx, int ny, int nz,
            float dx, float dy, float dz, float sx, float sy, float sz,
            float *T);
/* Many shots on nthreads host threads (CPU baseline timing). */
void or_fsm_batch(const float *slowness, int nx, int ny, int nz,
                  float dx, float dy, float dz, const float *sxyz, int nshot,
                  const float *rx, const float *ry, const float *rz, int nr,
                  float *picks, int nthreads);

/* ---- eikonal: Podvin & Lecomte (classic.cpp) ---- */
/* Classic::eikonal_equation (classic.cpp:184-1842) for one shot: pad by one
 * with replication (expand_model :8-60), seed the source node and activate
 * its 26 neighbours, nItEikonal (:248-273) Jacobi iterations of the 1D head
 * wave, edge and corner diffraction, 2D and 3D Huygens operators over the
 * active cells, crop (reduce_model :62-72).  h = dx only, as the
 * reference.  Output (nz,nx,ny), z fastest. */
void or_podvin(const float *slowness, int nx, int ny, int nz, float dx, float dy, float dz,
               float sx, float sy, float sz, float *T);

/* ---- tomography: least squares (least_squares.cpp) ---- */

/* expand_fdm (least_squares.cpp:11-67): pad T by one with replication */
void or_expand_fdm(const float *T, int nx, int ny, int nz, float *Tp);

/* ray_tracing for one shot (least_squares.cpp:192-280).  Rows are
 * row_base + receiver.  Writes per-row counts in row_nnz[nr], columns and
 * values in row order into cols/vals (capacity cap).  illum (nx*ny*nz) gets
 * rayStep per step at the reference's index (writes outside the volume
 * are dropped).  Returns total nnz, -1 if a ray exceeded max_steps,
 * -2 if cap is too small. */
long or_ray_trace_shot(const float *Tp, int nx, int ny, int nz,
                       float dx, float dy, float dz,
                       int nxt, int nyt, int nzt,
                       float dxt, float dyt, float dzt,
                       float sx, float sy, float sz,
                       const float *rx, const float *ry, const float *rz,
                       int nr, int max_steps,
                       int *row_nnz, int *cols, float *vals, long cap,
                       float *illum);

/* tk_reg_matrix difference stencil (least_squares.cpp:369-410) */
void or_tk_stencil(int order, int *df);

/* CGLS restatements.  A is CSR (rowptr has N+1 entries). */
/* FP64 recurrence of least_squares.cpp:424-496 (x0 = 0, fixed iters) */
void or_cgls_f64(const long *rowptr, const int *col, const float *val,
                 int N, int M, const float *b, int iters, double *x);
/* FP32 with the reference's COO scatter order and sequential sums */
void or_cgls_ref_f32(const long *rowptr, const int *col, const float *val,
                     int N, int M, const float *b, int iters, float *x);

/* converged() residual norm (least_squares.cpp:301-307) */
float or_residual_norm(const float *dobs, const float *dcal, int n);

/* gaussian (gaussian.cpp:5-73) */
void or_gaussian(const float *in, float *out, int nx, int ny, int nz,
                 int samples, float stdv);
/* kernel weights as the reference builds them (gaussian.cpp:11-43) */
void or_gaussian_kernel(int samples, float stdv, float *kernel);

/* model_update (least_squares.cpp:498-570): slowness updated in place,
 * dm and velocity written out. */
void or_model_update(float *slowness, const float *xdm,
                     int nx, int ny, int nz, float dx, float dy, float dz,
                     int nxt, int nyt, int nzt,
                     float dxt, float dyt, float dzt,
                     int smooth, int samples, float stdv,
                     float *dm, float *velocity);

/* ---- data synthesis (src/utils/wave, src/utils/auto_picking) --------- */

/* dampers (acoustic.cpp:100-137): Cerjan profiles d1[nb], d2[nb*nb],
 * d3[nb*nb*nb] (zero-initialised here) */
void or_acoustic_dampers(float *d1, float *d2, float *d3, float factor, int nb);

/* One shot of the acoustic modelling loop (wave_main.cpp:129-166 with
 * acoustic.cpp: expand, applyWavelet, wavePropagation, wavefieldUpdate,
 * buildSeismogram), serial; seismogram[nrec][nt]; ufut (nullable): the
 * final U_fut, padded nyy x nxx x nzz, z fastest. */
void or_acoustic_shot(const float *vel, int nx, int ny, int nz, float dx, float dy, float dz,
                      int nb, float factor, const float *wavelet, int nt, float dt,
                      float sx, float sy, float sz, const float *rx, const float *ry,
                      const float *rz, int nrec, float *seismogram, float *ufut);

/* The auto-picker of auto_picking.cpp:28-82 on one file's traces
 * (seis[ntraces][nt], normalised by its largest sample); picks[ntraces]
 * (0 where no sample crosses amp_cut). */
void or_auto_pick(const float *seis, int ntraces, int nt, float dt, float window,
                  float amp_cut, float shift, float *picks);

#ifdef __cplusplus
}
#endif
#endif
