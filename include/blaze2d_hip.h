/*
 * blaze2d_hip: MI355X (gfx950) backend for the Blaze2D band-solve path.
 *
 * C ABI.  Plain pointers and sizes only; every call returns 0 on success
 * and a negative status on failure, with the message in bh_last_error().
 * Complex arrays are interleaved doubles (re, im).  Precision argument:
 * 0 = FP64 storage, 1 = FP32 storage with FP64 accumulation.
 *
 * The entry points mirror the reference's boundary (Rust, relative to the
 * reference repository root):
 *   SpectralBackend / SpectralBuffer   crates/core/src/backend.rs:43-165
 *   CpuBackend (what this replaces)    crates/backend-cpu/src/lib.rs:28-452
 *   LinearOperator (Theta, mass)       crates/core/src/operators.rs:91-165
 *   OperatorPreconditioner             crates/core/src/preconditioners.rs:102-113
 *   Eigensolver::solve + run_core      crates/core/src/eigensolver.rs:1684-2192,
 *                                      crates/core/src/drivers/bandstructure.rs:381-885
 *   BulkDriver::benchmark job loop     crates/bulk-driver/src/driver.rs:1106-1257
 */
#ifndef BLAZE2D_HIP_H
#define BLAZE2D_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define BH_MAX_ATOMS 64

typedef struct bh_context bh_context;
typedef struct bh_buffer bh_buffer;
typedef struct bh_theta bh_theta;
typedef struct bh_cnn bh_cnn;

/* Library identification. */
const char *bh_version(void);
/* Message of the last failing call on this thread ("" if none). */
const char *bh_last_error(void);
/* Number of visible HIP devices (0 when none). */
int bh_device_count(void);

/* Context on one device (one HIP stream).  Replaces CpuBackend::new
 * (backend-cpu/src/lib.rs:40-90). */
int bh_context_create(int device, bh_context **out);
void bh_context_destroy(bh_context *ctx);
int bh_synchronize(bh_context *ctx);

/* ---- SpectralBackend level (backend.rs:73-165) ------------------------- */

/* alloc_field: `count` zeroed fields of nx*ny complex values. */
int bh_buffer_alloc(bh_context *ctx, int nx, int ny, int precision, int count, bh_buffer **out);
void bh_buffer_free(bh_buffer *buf);
/* Host (FP64 interleaved) <-> device (storage precision) for fields
 * [first, first + count). */
int bh_buffer_upload(bh_buffer *buf, int first, int count, const double *host);
int bh_buffer_download(const bh_buffer *buf, int first, int count, double *host);
/* forward_fft_2d (direction -1, unnormalised e^{-2 pi i jk/n}) or
 * inverse_fft_2d (direction +1, scaled by 1/(nx ny)) of fields
 * [first, first + count), in place (backend-cpu/src/lib.rs:94-275). */
int bh_fft2d(bh_context *ctx, bh_buffer *buf, int first, int count, int direction);
/* scale: x <- alpha x (backend-cpu/src/lib.rs:414-419). */
int bh_scale(bh_context *ctx, bh_buffer *x, int ix, double alpha_re, double alpha_im);
/* axpy: y <- alpha x + y (backend-cpu/src/lib.rs:421-426). */
int bh_axpy(bh_context *ctx, double alpha_re, double alpha_im, const bh_buffer *x, int ix, bh_buffer *y, int iy);
/* dot: sum conj(x) y accumulated in FP64 (backend-cpu/src/lib.rs:428-443). */
int bh_dot(bh_context *ctx, const bh_buffer *x, int ix, const bh_buffer *y, int iy, double out[2]);
/* gram_matrix: out[i*q + j] = <x_{xi0+i}, y_{yi0+j}>, row-major p x q
 * (backend.rs:105-128). */
int bh_gram(bh_context *ctx, const bh_buffer *x, int xi0, int p, const bh_buffer *y, int yi0, int q,
            double *out);
/* bh_gram with no allocation of any kind: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, limits (p, q <= 48) and result as bh_gram, bit for bit.  Its
 * partial sums live in a fixed 2.4 MB device array per GPU, set up by
 * bh_context_create and held until the process ends; calls that overlap on
 * one GPU share that array and race. */
int bh_gram_noalloc(bh_context *ctx, const bh_buffer *x, int xi0, int p, const bh_buffer *y, int yi0, int q,
                    double *out);
/* linear_combinations: out_j = sum_i q_{qi0+i} c[i + j*r], c col-major r x m
 * (backend.rs:130-150). */
int bh_lincomb(bh_context *ctx, const bh_buffer *q, int qi0, int r, const double *coeffs, int m, bh_buffer *out,
               int oi0);

/* Dense Hermitian eigendecomposition of `count` independent n x n matrices
 * (n <= 64) on the device, FP64: a is col-major n*n per matrix (lower
 * triangle read), w receives ascending eigenvalues (n per matrix) and v the
 * col-major eigenvectors.  dense::solve_hermitian_eigen (eigensolver/dense.rs:71-134). */
int bh_hermitian_eigh(bh_context *ctx, int n, int count, const double *a, double *w, double *v);

/* Rayleigh-Ritz of `count` subspaces Z of width n (n <= 96) from their Grams,
 * FP64 on the device.  g1 = Z^H B Z and g2 = Z^H A Z, col-major n x n complex
 * per problem (lower triangles read).  The block is B-orthonormalised as
 * svqb_orthonormalize does (normalization.rs:281-544: unit B-norm columns,
 * directions below 1e-12 of the largest Gram eigenvalue dropped; Cholesky QR
 * is taken only when it provably drops nothing), the operator projected
 * (eigensolver.rs:1232-1316) and diagonalised in ascending order
 * (dense.rs:71-134).  Outputs per problem: lam[n] (rank ascending Ritz values,
 * then zeros); coef[n x n] col-major complex whose first cnt_x + cnt_w columns
 * are the coefficients over Z's columns of the new X (cnt_x = min(rank,
 * nsplit)) and W (cnt_w = min(2 nsplit, rank) - cnt_x) Ritz vectors
 * (eigensolver.rs:1341-1637), the rest zero; info[4] = {rank, SVQB path taken
 * (0 or 1), cnt_x, cnt_w}. */
int bh_rayleigh_ritz(bh_context *ctx, int n, int count, int nsplit, const double *g1, const double *g2, double *lam,
                     double *coef, int *info);

/* ---- LinearOperator / OperatorPreconditioner level -------------------- */

typedef struct {
    int nx, ny;
    int pol;                 /* 0 = TM, 1 = TE */
    double b1[2], b2[2];     /* reciprocal lattice vectors (lattice.rs:170-180) */
    double bloch[2];         /* Cartesian Bloch wavevector */
    const double *eps;       /* N values: smoothed eps (TM mass, TP preconditioner) */
    const double *tensor;    /* N*4 values [xx, xy, yx, yy] inverse-eps tensor (TE); may be NULL -> 1/eps */
    int precond;             /* -1 auto (TM -> fourier-diagonal, TE -> transverse projection), 0 none, 1 FD, 2 TP */
    int precision;
} bh_theta_desc;

/* ThetaOperator::new (maxwell.rs:126-176) plus the adaptive-shift
 * preconditioner (maxwell.rs:218-379). */
int bh_theta_create(bh_context *ctx, const bh_theta_desc *desc, bh_theta **out);
void bh_theta_destroy(bh_theta *op);
/* batch_apply: out_j = Theta in_j (maxwell.rs:647-713,1532-1542). */
int bh_theta_apply(bh_theta *op, const bh_buffer *in, int ii0, bh_buffer *out, int oi0, int count);
/* batch_apply_mass: out_j = B in_j (TM: eps, TE: identity) (maxwell.rs:1522-1530). */
int bh_mass_apply(bh_theta *op, const bh_buffer *in, int ii0, bh_buffer *out, int oi0, int count);
/* OperatorPreconditioner::batch_apply in place (fourier_diagonal.rs:310-342,
 * transverse_projection.rs:167-231). */
int bh_precond_apply(bh_theta *op, bh_buffer *buf, int first, int count);

/* ---- Dielectric setup (dielectric.rs:146-210,473-574) ------------------ */

typedef struct {
    double a1[2], a2[2];     /* lattice vectors */
    double eps_bg;
    int n_atoms;
    double atoms[BH_MAX_ATOMS][4]; /* fractional x, y, radius (units of |a1|), eps_inside */
} bh_geometry;

/* Smoothed eps, inverse eps and inverse-eps tensor on an nx x ny node grid. */
int bh_dielectric(const bh_geometry *geom, int nx, int ny, int smoothing, double *eps, double *inv_eps,
                  double *tensor);
/* The same tables for ngeom geometries on one grid, computed on the GPU (one
 * thread per node; replaces Dielectric2D::from_geometry, dielectric.rs:146-210,
 * and the analytic smoothing of analytic_geometry.rs:183-541, for whole
 * sweeps and large grids).  Host outputs laid out [geometry][node] (tensor
 * [geometry][node][4]); any may be NULL.  gpu_ms (may be NULL) receives the
 * kernel's event-timed duration. */
int bh_dielectric_batch(bh_context *ctx, const bh_geometry *geoms, int ngeom, int nx, int ny, int smoothing,
                        double *eps, double *inv_eps, double *tensor, double *gpu_ms);

/* ---- Band structures (Eigensolver::solve driven by run_core) ----------- */

typedef struct {
    bh_geometry geom;
    int nx, ny;
    int pol;                 /* 0 TM, 1 TE */
    int n_bands;
    int max_iter;            /* default 200 */
    double tol;              /* relative eigenvalue change (eigensolver.rs:345-400) */
    int block_size;          /* 0 -> n_bands + 2 */
    int smoothing;           /* 1 analytic smoothing (default) */
    int precond;             /* -1 auto */
    int subspace_prediction; /* 1 (default) */
    int extrapolation;       /* 1 (default) */
    int band_tracking;       /* 1 (default) */
    int reuse_gamma;         /* 1 (default) */
    int exit_rule;           /* 0: Eigensolver::solve() accounting; 1: soft-locked bands only */
    int nk;
    const double *kpath;     /* nk x 2 fractional k-points */
    /* Learned start (optional; initialize_block's warm vectors, initialization.rs:248-309):
     * warm_count vectors for the job's first k-point given by their coefficients on
     * the plane waves G = n1 b1 + n2 b2, |n1|, |n2| <= pw_half (0..8), complex,
     * index (n2 + pw_half)(2 pw_half + 1) + n1 + pw_half per vector; the field
     * is sum_G c_G exp(2 pi i (n1 ix / nx + n2 iy / ny)).  NULL: default start. */
    const double *warm_pw;
    int warm_count;
    int pw_half;
} bh_job;

typedef struct {
    int slots;               /* concurrent band structures on the GPU (0 -> 64) */
    int precision;           /* 0 FP64; 1 FP32 fields (FP64 accumulation and dense algebra);
                                2 FP64 fields with the preconditioners (FD, TP) computed in FP32 */
    int profile;             /* 1: time operator kernels with HIP events (runs one slot group) */
    int groups;              /* slot groups pipelined on separate HIP streams (0 -> 4) */
    int timed_kernel;        /* kernel class timed with HIP events on its stream:
                                0 none, 1 Theta operator, 2 dual Rayleigh-Ritz Gram,
                                3 Rayleigh-Ritz solve (Cholesky/SVQB + projection + Jacobi),
                                4 Ritz-update linear combinations,
                                5 residual and direction deflation,
                                6 k-point transitions (tracking, prediction, start block) */
    int engines;             /* engines run side by side, each with its share of the slots, its own
                                streams and host thread (0 -> 1): the jobs of each grid size form one
                                engine's share, the largest shares are split further to reach this count */
    int gram_mfma_min;       /* FP64 Gram products at least this many columns wide run on the matrix
                                cores (0 -> 33); a tuning knob (blaze2d_amd.autotune) */
} bh_run_options;

typedef struct {
    long steps;
    long iterations;         /* LOBPCG iterations summed over solves */
    long solves;             /* k-point solves (closing Gamma copies excluded) */
    double seconds;          /* wall time of the solve loop */
    long eigh_calls[2];      /* [SVQB, Rayleigh-Ritz] dense problems (profile = 1) */
    long eigh_sweeps[2];     /* Jacobi sweeps summed over them */
    long eigh_max_sweeps[2]; /* largest sweep count of a single problem */
    long eigh_launch_max_sum[2]; /* per-launch maximum sweep count, summed over launches */
    double timed_ms;             /* summed duration of the timed_kernel launches */
    long timed_launches;
    double timed_bytes;          /* their algorithmic HBM bytes */
    double timed_flops;          /* their algorithmic FLOPs */
    double host_wait_seconds;    /* host time blocked on group completion events */
    double host_plan_seconds;    /* host time spent planning and enqueueing steps */
    long rr_problems;            /* Rayleigh-Ritz problems solved */
    long svqb_fallbacks;         /* of them, B-orthonormalised on the SVQB path */
    long rank_drops;             /* of them, with a direction dropped (rank below the subspace width) */
    double timed_flops_executed; /* the FLOPs the timed kernels ran (differs from timed_flops for the
                                    Rayleigh-Ritz class: its Jacobi sweeps as run, not a method-independent count) */
    double alg_bytes;            /* algorithmic HBM bytes of every launch of the run (each column read or
                                    written once per launch, tables once): the whole path's HBM traffic floor */
} bh_run_stats;

/* One-time operator diagnostics of a job (drivers/bandstructure.rs:437-498):
 * at the job's first non-Gamma k-point (|k_frac|^2 >= 1e-8), the Theta
 * operator's self-adjointness error and power-iteration condition estimates
 * of Theta and of the preconditioned operator (maxwell.rs:382-582). */
typedef struct {
    int k_index;             /* k-point used; -1 when the path has only Gamma (then all zero) */
    int precond;             /* preconditioner of the pm estimate: 0 none, 1 FD, 2 TP */
    double lambda_max;       /* Rayleigh quotient of the last power iteration on Theta */
    double lambda_min;       /* smallest clamped |k+G|^2 (k_plus_g_sq_min) */
    double kappa;            /* lambda_max / lambda_min, +inf when lambda_min <= 1e-15 */
    double pm_lambda_max;    /* the same on M^-1 Theta; 0 when precond = 0 */
    double pm_kappa;         /* pm_lambda_max / 1 (the reference's lambda_min estimate) */
    double self_adjoint_err; /* |<Ax, y> - <x, Ay>| / (|Ax| |y|) */
} bh_diagnostics;

/* The diagnostics of njobs jobs (sharing nx, ny) on the GPU: dielectric
 * tables, start vectors and power_iterations steps (the reference uses 15)
 * batched over the jobs.  gpu_ms (may be NULL) receives the event-timed device
 * time.  Replaces ThetaOperator::{check_self_adjointness,
 * estimate_condition_number, estimate_preconditioned_condition_number}
 * (maxwell.rs:382-582) as run once per job by run_core. */
int bh_operator_diagnostics(bh_context *ctx, const bh_job *jobs, int njobs, int power_iterations,
                            bh_diagnostics *out, double *gpu_ms);

/* Runs every job to completion (jobs may use different grids; see
 * bh_run_options.engines).  Outputs per
 * job j at offsets: omegas[(sum of nk*n_bands over jobs < j) ...] as
 * omega (not omega / 2 pi), iters / converged per k-point. */
int bh_bandstructure_run(bh_context *ctx, const bh_job *jobs, int njobs, const bh_run_options *opts,
                         double *omegas, int *iters, int *converged, bh_run_stats *stats);

/* Optional outputs of bh_bandstructure_run_ex; any pointer may be NULL. */
typedef struct {
    /* The residual diagnostics the reference's solve() computes on its log
     * iterations (1-5, 10, 20, 50, 100, then every 50; eigensolver.rs:251-259,
     * 1729-1737): per k-point the relative residual B-norms
     * ||r_i||_B / max(|lambda_i|, 1e-10) of the first n_bands block entries
     * (eigensolver.rs:744-779; block order, soft-locked bands 0) at the solve's
     * last log iteration, at the same offsets as omegas, and that iteration
     * (1-based) at the same offsets as iters. */
    double *rel_residuals;
    int *rel_residual_iters;
    /* pw_half > 0 (up to 8): per k-point the n_bands converged eigenvectors
     * (all_eigenvectors order, eigensolver.rs:2404-2437: ascending eigenvalue,
     * B-normalised) projected on the plane waves |n1|, |n2| <= pw_half,
     * c_G = (1/N) sum_r x(r) exp(-2 pi i (n1 ix / nx + n2 iy / ny)), layout as
     * bh_job::warm_pw, at offset (sum of nk * n_bands over earlier jobs + k
     * n_bands + band) (2 pw_half + 1)^2 complex values.  The training data of
     * the learned start. */
    int pw_half;
    double *pw_coefs;
    /* Per k-point, at the same offsets as omegas: the B-norms ||M^-1 r_i||_B
     * of the preconditioned residuals of the solve's first iteration (the
     * start block after its Rayleigh-Ritz; M = I without a preconditioner).
     * No counterpart in the reference: the learned start's fallback test
     * compares them between a predicted and the default start. */
    double *prec_residuals;
} bh_run_outputs;

/* bh_bandstructure_run with the optional outputs above. */
int bh_bandstructure_run_ex(bh_context *ctx, const bh_job *jobs, int njobs, const bh_run_options *opts,
                            double *omegas, int *iters, int *converged, const bh_run_outputs *extra,
                            bh_run_stats *stats);

/* ---- learned start ------------------------------------------------------
 * The reference has no learned start: its warm vectors come from the
 * previous k-point (subspace_prediction.rs:276-380) or none.  This network
 * predicts warm vectors for a solve from the eps grid and the k-point; they
 * enter the solver through bh_job::warm_pw, the warm-start ingestion of
 * initialize_block (initialization.rs:248-309).
 *
 * Network on a height x width input (multiples of 8), C = 32 channels:
 *   e   = W_e2 relu(W_e1 [k1, k2, pol, lattice] + b_e1) + b_e2   (4 C)
 *   a_1 = relu(conv3x3(x, stride 1) + b_1 + e[0:C])
 *   a_l = relu(conv3x3(a_{l-1}, stride 2) + b_l + e[(l-1) C : l C]), l = 2..4
 *   y   = W_h a_4 + b_h   (n_vectors x (2 pw_half + 1)^2 complex, bh_job::warm_pw layout)
 * with periodic 3 x 3 convolutions.  Parameters (FP32 master copy, used as
 * BF16 with FP32 accumulation), in this order in the flat array:
 *   per layer l: w_l [C][c_in][3][3] (c_in = 2 for l = 1, else C), b_l [C];
 *   W_e1 [C][4], b_e1 [C]; W_e2 [4C][C], b_e2 [4C];
 *   W_h [2 n_vectors (2 pw_half + 1)^2][C (height/8) (width/8)], b_h.
 * Inputs per sample: grid [2][height][width] floats (eps / 16, 1 / eps),
 * embed [4] floats (k1, k2 fractional; pol 0 TM / 1 TE; lattice 0 square /
 * 1 hexagonal).  Training loss per sample, invariant to the basis of either
 * block: L = 1 - tr(P_Y P_U) / n_vectors, P the orthogonal projector on the
 * columns (predicted Y, target U), averaged over the batch. */
int bh_cnn_create(bh_context *ctx, int height, int width, int n_vectors, int pw_half, int max_batch,
                  unsigned long long seed, bh_cnn **out);
void bh_cnn_destroy(bh_cnn *net);
long long bh_cnn_param_count(const bh_cnn *net);
int bh_cnn_get_params(bh_cnn *net, float *host);
/* Replaces the master weights (and their BF16 copy); restarts Adam. */
int bh_cnn_set_params(bh_cnn *net, const float *host);
/* Gradients of the last bh_cnn_train_step's loss w.r.t. the parameters. */
int bh_cnn_get_grads(bh_cnn *net, float *host);
/* coefs: [batch][n_vectors][(2 pw_half + 1)^2] complex (interleaved doubles). */
int bh_cnn_predict(bh_cnn *net, int batch, const float *grid, const float *embed, double *coefs);
/* bh_cnn_predict with the grid input computed on the device from the
 * geometries: the smoothed (smoothing = 1) or plain eps of each on the
 * network's height x width grid (as bh_dielectric with nx = width, ny =
 * height), then eps / 16 and 1 / eps. */
int bh_cnn_predict_geometries(bh_cnn *net, int batch, const bh_geometry *geoms, int smoothing, const float *embed,
                              double *coefs);
/* Forward, loss, backward and (learning_rate > 0) one Adam step (beta 0.9,
 * 0.999, eps 1e-8, bias-corrected).  targets as coefs.  loss may be NULL. */
int bh_cnn_train_step(bh_cnn *net, int batch, const float *grid, const float *embed, const double *targets,
                      float learning_rate, double *loss);
/* No reference counterpart (the reference has no network).  enable = 1 (the
 * default): bh_cnn_train_step replays its device work (forward, loss,
 * backward, Adam; about thirty kernels) from a HIP graph captured on the
 * first step of each batch size, the inputs, targets and the step's learning
 * rate and bias corrections copied in before each replay; enable = 0: the
 * same kernels launched one by one.  Both give identical results. */
int bh_cnn_set_graphs(bh_cnn *net, int enable);

#ifdef __cplusplus
}
#endif
#endif
