/*
 * st_api.h -- C ABI of the MI355X-native first-arrival tomography path.
 *
 * Drop-in boundary for the eikonal and least-squares tomography path of
 * phbastosa/seismic_tomography_3D.  Plain pointers and sizes only; every
 * call returns an st_status (0 = ST_OK).  Host arrays are owned by the
 * caller; the context owns all device memory and reuses it across calls.
 * A context is not thread-safe; calls are synchronous at return.
 *
 * Volumes are FP32, z fastest: idx = z + x*nz + y*nx*nz (the reference's
 * layout, src/eikonal/block_FIM/block_FIM.cpp:17).  Point triples
 * (sources, receivers) are interleaved x, y, z in metres.
 */
#ifndef ST_API_H
#define ST_API_H

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    ST_OK = 0,
    ST_ERR_ARG = 1,          /* bad argument / size / pointer             */
    ST_ERR_HIP = 2,          /* HIP runtime error (message in st_last_error) */
    ST_ERR_NO_MODEL = 3,     /* st_set_model not called                   */
    ST_ERR_DIVERGED = 4,     /* ray exceeded its step cap or lost its gradient */
    ST_ERR_NOT_UPPER_BOUND = 5, /* warm start violated the upper-bound check */
    ST_ERR_CAPACITY = 6,     /* an output buffer is too small             */
    ST_ERR_NO_SURROGATE = 7, /* warm start requested without a network    */
    ST_ERR_STATE = 8         /* call order (e.g. cgls before forward)     */
} st_status;

typedef struct st_ctx st_ctx;

/* Eikonal solver options (block Fast Iterative Method). */
typedef struct {
    int block;          /* block edge: 4 or 8 (default 8)                      */
    int batch;          /* shots per persistent solve (default 256); 1 = one
                           shot at a time                                      */
    int max_inner;      /* inner iteration cap per block visit (0 = 24 for
                           8^3 blocks, 16 for 4^3)                             */
    float eps;          /* change threshold in seconds (<0 = default 1e-6, the
                           reference's EPS; 0 = iterate to the exact fixed point) */
    int warm;           /* 0 cold start, 1 surrogate warm start                */
    float delta;        /* warm start inflation (1+delta) (0 = 0.02)           */
    int judge;          /* 1: verify shots started from seeded values
                           (warm_values = 1), re-solve cold on failure         */
    int grid;           /* waves per round launch (0 = 192 per CU)             */
    int rounds_per_sync;/* rounds enqueued between host checks (0 = auto)      */
    int use_graph;      /* 1: replay rounds from a captured hipGraph           */
    float band;         /* warm starts: visit a block only once its predicted
                           arrival is within `band` s of its shot's earliest
                           scheduled block (0 = default, <0 = off)             */
    float cold_radius;  /* warm starts: cells closer than this to the source
                           start cold (m, 0 = default)                         */
    int warm_values;    /* warm starts: 1 initialise T = (1+delta)*prediction
                           (an upper bound only if the prediction errs by less
                           than delta; the judge re-solves the others cold),
                           0 (default) keep T cold and use the prediction only
                           to order block visits (exact by construction)       */
    int sweep;          /* 8^3 blocks: 0 (default) = each visit opens with one
                           Gauss-Seidel pass over diagonal planes in the octant
                           facing away from the source, then z sweeps; 1 = z
                           sweeps only                                          */
    int order;          /* block visit ordering (rounds visit, per shot, the
                           scheduled blocks whose key lies within `band` s of
                           the shot's smallest key; the others wait; every
                           mode reaches the same fixed point): 0 (default) =
                           keys from the network's predicted arrival for warm
                           starts, no ordering for cold ones; 1 = each
                           block's tentative arrival, the smallest face value
                           through which a visit scheduled it (no network
                           needed); 2 = the larger of the two (warm starts)  */
} st_fim_opts;

typedef struct {
    long long rounds;       /* FIM rounds executed (summed over batches)  */
    long long visits;       /* block visits                               */
    long long inner_iters;  /* inner iterations summed over visits        */
    long long fallbacks;    /* warm-started shots re-solved cold          */
    long long batches;      /* persistent solves launched                 */
    long long deferred;     /* block visits postponed by arrival ordering */
    double judge_max;       /* largest warm-start residual seen (s)       */
    double solve_ms;        /* device time of the FIM rounds              */
    double total_ms;        /* wall time of the call                      */
    long long launches;     /* round kernels launched (rounds + the empty
                               ones enqueued before the host saw the end) */
} st_fim_stats;

/* Fills opts with defaults. */
void st_fim_default_opts(st_fim_opts *opts);

/* Context lifetime.  Replaces the per-shot cudaMalloc/cudaFree of
 * block_FIM_solver (src/eikonal/block_FIM/cuda_kernel.cu:54-68,208-214). */
int st_ctx_create(int device, st_ctx **out);
int st_ctx_destroy(st_ctx *ctx);
const char *st_last_error(const st_ctx *ctx);

/* Slowness model (s/m), z fastest, nx*ny*nz.  Replaces
 * Block_FIM::set_parameters' model read (block_FIM.cpp:85-91),
 * prepare_volumes/expand_model (:142-188, :8-57) and apply_model_mask
 * (:264-293): the padded, block-major copy lives on the device.  The FIM
 * uses dx as the isotropic spacing, as the reference does (:173). */
int st_set_model(st_ctx *ctx, const float *slowness, int nx, int ny, int nz,
                 float dx, float dy, float dz);

/* Current device slowness back to the host (z fastest, nx*ny*nz). */
int st_get_model(st_ctx *ctx, float *slowness);

/* Solve nshot point sources and pick nrcv receivers per shot.  Replaces
 * the shot loop of eikonal_main.cpp:31-39 over
 * Block_FIM::eikonal_equation (block_FIM.cpp:212-220: apply_source_time,
 * block_FIM_solver, extract_solution) and write_first_arrival (:228-262).
 * picks: nshot*nrcv (shot-major).  T_out (nullable): nshot*nx*ny*nz.
 * opts/stats nullable. */
int st_fim_solve(st_ctx *ctx, const float *src_xyz, int nshot,
                 const float *rcv_xyz, int nrcv, float *picks, float *T_out,
                 const st_fim_opts *opts, st_fim_stats *stats);

/* As st_fim_solve (cold start, picks only), but the call allocates no
 * device memory: the block-layout slowness, travel times, their x-face
 * copies, schedule flags, active lists, ordering keys, statistics, sources,
 * receivers and picks all
 * live in one static device array of 64 MiB per GPU, held from library load
 * to process end.  No counterpart in the reference; it is for callers that
 * cannot allocate device memory (serving loops, harnesses that forbid
 * hipMalloc).  Limit: the slowness plus one shot's state must fit in the
 * 64 MiB (201x201x101 with 8^3 blocks: 41 MiB; shots are solved in passes
 * of as many as fit); larger models return ST_ERR_ARG.  opts: warm must be
 * 0, order 0 or 1; use_graph is ignored.  Every call on a GPU shares that
 * GPU's array: calls that overlap on one GPU (two threads, two contexts on
 * the same device) race. */
int st_fim_solve_noalloc(st_ctx *ctx, const float *src_xyz, int nshot,
                         const float *rcv_xyz, int nrcv, float *picks,
                         const st_fim_opts *opts, st_fim_stats *stats);

/* As st_fim_solve, warm-started from caller predictions T_pred (nshot
 * volumes, z fastest; e.g. the previous tomography iteration's times).
 * Each shot starts from (1+delta)*T_pred (cold near the source), block
 * visits are ordered by predicted arrival, and the judge re-solves from a
 * cold start every shot whose result is not a fixed point of the upwind
 * update (a prediction that was not an upper bound).  No counterpart in the
 * reference (its FIM always starts cold, block_FIM.cpp:295-369). */
int st_fim_solve_warm(st_ctx *ctx, const float *src_xyz, int nshot,
                      const float *rcv_xyz, int nrcv, float *picks, float *T_out,
                      const float *T_pred, const st_fim_opts *opts,
                      st_fim_stats *stats);

/* The reference's CPU eikonal, Accurate_FSM (Noble-Gesret-Belayouni
 * operators: accurate_FSM.cpp:577-650 eikonal_equation with
 * write_first_arrival :541-575), on the device for all shots of the call,
 * bit-identical to the CPU code on the same model (isotropic or not).
 * Replaces Accurate_FSM::eikonal_equation + write_first_arrival per shot
 * (selected by eikonal_type, eikonal_main.cpp:15-22).  picks[nshot x nrcv];
 * T_out (nullable) [nshot x ny x nx x nz], z fastest; batch = shots per
 * device pass (0 = 256); device_ms (nullable) = device time of the sweeps. */
int st_fsm_solve(st_ctx *ctx, const float *src_xyz, int nshot,
                 const float *rcv_xyz, int nrcv, float *picks, float *T_out,
                 int batch, double *device_ms);

/* Receiver picks on a caller volume: write_first_arrival's trilinear
 * (block_FIM.cpp:228-262, trilinear.cpp:3-20) on the device. */
int st_picks_volume(st_ctx *ctx, const float *T, int nx, int ny, int nz,
                    float dx, float dy, float dz, const float *rcv_xyz,
                    int nrcv, float *picks);


/* ---- least-squares tomography (src/tomography/least_squares) ---------- */

/* Inversion grid: n*_tomo = (int)((n-1)*d/d_tomo) + 1
 * (least_squares.cpp:124-126).  max_steps caps each ray (0 = default);
 * the reference's ray loop has none (:213-237). */
int st_tomo_set_grid(st_ctx *ctx, float dx_tomo, float dy_tomo, float dz_tomo,
                     int max_steps);
int st_tomo_dims(st_ctx *ctx, int *nx_tomo, int *ny_tomo, int *nz_tomo);

/* Least_squares::forward_modeling (least_squares.cpp:163-190): for every
 * shot the eikonal solve, the receiver picks (dcal, nshot*nrcv) and the
 * gradient ray tracing (:192-280) into the device-resident sensitivity
 * matrix G (CSR, row = shot*nrcv + receiver, columns ascending).  illum
 * (nullable, nx*ny*nz) receives the ray illumination summed over shots.
 * opts->warm = 1 orders the block visits by the warm-start network's
 * prediction (warm_values must stay 0: the result is the cold fixed point). */
int st_forward(st_ctx *ctx, const float *src_xyz, int nshot,
               const float *rcv_xyz, int nrcv, float *dcal, float *illum,
               const st_fim_opts *opts, st_fim_stats *stats);

/* Least_squares::set_parameters' choice of the forward eikonal
 * (least_squares.cpp:85-104): 0 Podvin & Lecomte (Classic), 1 block FIM,
 * 2 Accurate_FSM; any other value selects Accurate_FSM, the switch's default
 * branch.  st_forward then solves with that method and traces the rays on
 * its travel times (Accurate_FSM and Podvin bit-identical to the CPU code,
 * so picks and G equal the reference's).  A new context uses 1.  The
 * st_fim_opts of st_forward apply to the block FIM only (a warm start with
 * another method is ST_ERR_ARG). */
int st_tomo_set_eikonal(st_ctx *ctx, int eikonal_type);

/* As st_forward, and T_out (nshot*nx*ny*nz, z fastest) receives each shot's
 * travel-time volume, the one its rays were traced on: the reference's
 * eikonal->travel_time that expand_fdm pads for ray_tracing
 * (least_squares.cpp:175-181). */
int st_forward_times(st_ctx *ctx, const float *src_xyz, int nshot,
                     const float *rcv_xyz, int nrcv, float *dcal, float *illum,
                     float *T_out, const st_fim_opts *opts, st_fim_stats *stats);

/* Ray tracing on a caller volume T (z fastest, nx*ny*nz, the current model
 * grid) for one source; rows are appended to G.  st_csr_reset clears G. */
int st_csr_reset(st_ctx *ctx);
int st_raytrace_volume(st_ctx *ctx, const float *T, const float *src_xyz,
                       const float *rcv_xyz, int nrcv, float *illum);
int st_csr_size(st_ctx *ctx, long long *nrows, long long *nnz);
int st_csr_copy(st_ctx *ctx, long long *rowptr, int *cols, float *vals);

/* converged() residual (least_squares.cpp:301-307), FP64 on the device */
int st_residual_norm(st_ctx *ctx, const float *dobs, const float *dcal, int n,
                     double *norm);

/* Least_squares::optimization (least_squares.cpp:321-356): B = dobs - dcal
 * padded with zeros, Tikhonov rows lambda*stencil (tk_reg_matrix,
 * :358-422), CGLS (sparse_cgls, :424-496) for `iters` iterations with FP32
 * vectors and FP64 dot products.  The solution stays on the device for
 * st_model_update; xdm (nullable, n_model) receives a copy. */
int st_optimize(st_ctx *ctx, const float *dobs, const float *dcal, int n,
                int tk_order, float lambda, int iters, float *xdm);

/* CGLS on a caller CSR matrix (N rows, M columns, rowptr has N+1 entries),
 * x0 = 0: the sparse_cgls recurrence without regularisation rows. */
int st_cgls_csr(st_ctx *ctx, const long long *rowptr, const int *cols,
                const float *vals, int N, int M, const float *b, int iters,
                float *x);

/* Least_squares::model_update (least_squares.cpp:498-570) with the optional
 * gaussian (src/utils/smoothing/gaussian.cpp:5-73): trilinear coarse-to-fine
 * dm, smoothing, slowness += dm inside the margins.  xdm = NULL uses the
 * device solution of st_optimize.  The device model is updated in place;
 * dm and velocity (nullable, nx*ny*nz) receive copies. */
int st_model_update(st_ctx *ctx, const float *xdm, int smooth, int samples,
                    float stdv, float *dm, float *velocity);

/* ---- learned warm start (no counterpart in the reference) ------------- */
/* A 3D U-Net on an nc^3 node cube (nc a power of two, 32 by default)
 * predicting y = T*3000/max(D, dmin) from the slowness, the source
 * position and the straight-ray mean slowness; st_fim_solve with opts.warm = 1 uses it to order block visits
 * and to initialise (1+delta)*T.  Training and inference run on the device
 * with BF16 MFMA GEMMs (FP32 accumulation). */
int st_surrogate_init(st_ctx *ctx, int nc, unsigned seed);
int st_surrogate_nparams(st_ctx *ctx, long long *n);
/* copy parameters out (get) and/or in (set); either may be NULL */
int st_surrogate_params(st_ctx *ctx, float *get, const float *set);
/* forward on X[ns, nc^3, 6] -> Y[ns, nc^3]; with target and mask also the
 * weighted MSE loss and its parameter gradient (grad: nparams) */
int st_surrogate_eval(st_ctx *ctx, const float *X, int ns, const float *target,
                      const float *mask, float *Y, double *loss, float *grad);
/* training set on the current model grid, solved by the plain FIM */
int st_surrogate_make_data(st_ctx *ctx, int n_models, int shots_per_model,
                           float zmax_src, unsigned seed, float grad_fraction);
/* Weight of under-predictions (prediction below the target) in the
 * training and evaluation loss, >= 1 (1 = plain MSE, the default; reset by
 * st_surrogate_init).  A large
 * weight trains the network towards an upper bound of the travel times,
 * which the (1+delta) value warm start needs. */
int st_surrogate_loss(st_ctx *ctx, float under_weight);
/* Adam; history (nullable): per epoch [train loss, validation loss] */
int st_surrogate_train(st_ctx *ctx, int epochs, int batch, float lr, int n_val,
                       unsigned seed, float *history);
int st_surrogate_save(st_ctx *ctx, const char *path);
int st_surrogate_load(st_ctx *ctx, const char *path);
/* network output y on the node cube for the given sources (diagnostics) */
int st_surrogate_predict(st_ctx *ctx, const float *src_xyz, int nshot, float *y);
/* diagnostic: C = A B with the BF16 MFMA GEMM of the network */
int st_nn_gemm(st_ctx *ctx, const float *A, const float *B, float *C, int M,
               int N, int K);

/* One least-squares step after st_forward: the right-hand side, Tikhonov
 * rows and A^T are assembled, then the CGLS iterations and the model
 * update (with Gaussian smoothing when smooth != 0) run as one captured
 * HIP graph (use_graph != 0) or eagerly; the same kernels in the same
 * order, so both give the same result.  Replaces Least_squares::
 * optimization followed by model_update (least_squares.cpp:321-570);
 * xdm (n_model), dm and velocity (nx*ny*nz) are optional outputs as for
 * st_optimize and st_model_update. */
int st_optimize_update(st_ctx *ctx, const float *dobs, const float *dcal, int n,
                       int tk_order, float lambda, int iters, int smooth,
                       int samples, float stdv, int use_graph, float *xdm,
                       float *dm, float *velocity);

/* ---- data synthesis --------------------------------------------------
 * The acoustic modelling of src/utils/wave (8th-order Laplacian, Cerjan
 * absorbing layers) and the first-break auto-picker of
 * src/utils/auto_picking, on the device, bit-identical to the reference's
 * serial CPU code (its in-place damping order included). */

/* Absorbing-layer profiles on the host: d1[nb], d2[nb*nb], d3[nb^3].
 * Replaces dampers (acoustic.cpp:100-137). */
int st_acoustic_dampers(float *d1, float *d2, float *d3, float factor, int nb);

/* The reference's Podvin & Lecomte (1991) eikonal (classic.cpp:184-1842,
 * Classic::eikonal_equation with write_first_arrival :1850-1884;
 * eikonal_type 0) on the device for all shots of the call, bit-identical to
 * the CPU code.  Same arguments as st_fsm_solve; batch = shots per device
 * pass (0 = 16).  Replaces Classic::eikonal_equation per shot. */
int st_podvin_solve(st_ctx *ctx, const float *src_xyz, int nshot,
                    const float *rcv_xyz, int nrcv, float *picks, float *T_out,
                    int batch, double *device_ms);

/* Ricker wavelet on the host: out[t] for t*dt - delay, peak frequency
 * parameter fmax.  Replaces rickerGeneration (acoustic.cpp:81-98). */
int st_ricker(float delay, float fmax, float dt, int nt, float *out);

/* Velocity model vel[ny][nx][nz] (z fastest, m/s), spacings, nb absorbing
 * cells per side (>= 4), damping factor and time step.  Replaces
 * wave_main.cpp:10-48 (expand, dampers). */
int st_acoustic_set_model(st_ctx *ctx, const float *vel, int nx, int ny, int nz,
                          float dx, float dy, float dz, int nb, float factor,
                          float dt);

/* Model nshot shots with one wavelet[nt] and record nrec receivers per
 * shot.  Replaces the shot loop of wave_main.cpp:118-166 (setWaveField,
 * then per step applyWavelet, wavePropagation, wavefieldUpdate,
 * buildSeismogram).  Sources and receivers sit at cell (int)(p / d) + nb
 * per axis and must lie in the model.  group = shots advanced together
 * (0 = automatic: enough to fill the device; each holds three wavefields).
 * seis (nullable):
 * [nshot][nrec][nt] (the reference's seismogram[t + r nt] per shot);
 * ufut_last (nullable): the last shot's final U_fut (nyy*nxx*nzz padded
 * cells, z fastest); device_ms (nullable): device time of the call. */
int st_acoustic_shots(st_ctx *ctx, const float *wavelet, int nt,
                      const float *src_xyz, int nshot, const float *rcv_xyz,
                      int nrec, int group, float *seis, float *ufut_last,
                      double *device_ms);

/* As st_acoustic_shots, and pick every trace on the device without
 * returning the seismograms: each shot's seismogram is one file of the
 * auto-picker (normalised by its own largest sample).  picks[nshot][nrec]. */
int st_acoustic_picks(st_ctx *ctx, const float *wavelet, int nt,
                      const float *src_xyz, int nshot, const float *rcv_xyz,
                      int nrec, int group, float window, float amp_cut,
                      float shift, float *picks, double *device_ms);

/* The auto-picker over one file of traces seis[ntraces][nt] (nt <= 8192):
 * normalise by the file's largest sample, energy ratio of the iw =
 * (int)(window / dt) + 1 samples after and before each time, first time
 * whose ratio over its maximum exceeds amp_cut; picks[ntraces] = that time
 * * dt + shift, 0 without a crossing.  Replaces auto_picking.cpp:26-83
 * (shift 0.045 s and amp_cut 1e-15 there). */
int st_auto_pick(st_ctx *ctx, const float *seis, int ntraces, int nt, float dt,
                 float window, float amp_cut, float shift, float *picks);

#ifdef __cplusplus
}
#endif
#endif
