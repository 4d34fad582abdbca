/*
 * FP64 CPU oracle for the Blaze2D band-solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  The HIP implementation's results are
 * compared against this library's.  Only tests/, the smoke() entry and
 * bench.py's cpu_baseline leg may load it; the product library never links
 * or calls it.
 *
 * It restates, in plain C++ with complex<double> arithmetic, the reference
 * CPU path (crates/core + crates/backend-cpu, Rust).  Every function in
 * oracle.cpp cites the reference file:line it follows.
 *
 * Complex arrays cross this ABI as interleaved doubles (re, im).
 */
#ifndef BLAZE2D_ORACLE_H
#define BLAZE2D_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_MAX_ATOMS 64

typedef struct {
    double a1[2], a2[2];          /* lattice vectors (crates/core/src/lattice.rs) */
    double eps_bg;                 /* background permittivity */
    int n_atoms;
    double atoms[ORACLE_MAX_ATOMS][4]; /* frac x, frac y, radius (units of |a1|), eps_inside */
    int nx, ny;
    int pol;                       /* 0 = TM, 1 = TE */
    int n_bands;
    int max_iter;
    double tol;
    int block_size;                /* 0 = n_bands + 2 */
    int smoothing;                 /* 1 = analytic smoothing (reference default), 0 = raw sampling */
    int precond;                   /* -1 auto, 0 none, 1 fourier-diagonal, 2 transverse-projection */
    int subspace_prediction;       /* default 1 */
    int extrapolation;             /* default 1 */
    int band_tracking;             /* default 1 */
    int reuse_gamma;               /* default 1 */
    int exit_rule;                 /* 0 = Eigensolver::solve() exit accounting, 1 = soft-locked only */
    int nk;
    const double *kpath;           /* nk x 2 fractional k-points */
} OracleJob;

/* Full band structure (bandstructure::run_core with the plain solve()).
 * omegas: nk*n_bands (omega, not omega/2pi); iters, converged: nk each.
 * Returns total LOBPCG iterations, or -1 on error. */
int oracle_bandstructure(const OracleJob *job, double *omegas, int *iters, int *converged);
/* The same, also returning per k-point the relative residual B-norms
 * ||r_i||_B / max(|lambda_i|, 1e-10) of the first n_bands block entries (in
 * the solver's block order) at the last log iteration of the solve
 * (eigensolver.rs:251-259,744-779,1729-1737) and that iteration (1-based). */
/* oracle_bandstructure with warm vectors for the first k-point (ingested by
 * initialize_block, initialization.rs:248-309): warm_count fields given by
 * plane-wave coefficients on |n1|, |n2| <= pw_half (index (n2 + h)(2h + 1) +
 * n1 + h, interleaved complex), x(r) = sum_G c_G exp(+2 pi i (n1 ix / nx +
 * n2 iy / ny)).  rel_residuals / rel_residual_iters (may be NULL) as in
 * oracle_bandstructure_residuals. */
int oracle_bandstructure_warm(const OracleJob *job, const double *warm_pw, int warm_count, int pw_half,
                              double *omegas, int *iters, int *converged, double *rel_residuals,
                              int *rel_residual_iters);
int oracle_bandstructure_residuals(const OracleJob *job, double *omegas, int *iters, int *converged,
                                   double *rel_residuals, int *rel_residual_iters);

/* Dielectric sampling / analytic smoothing (dielectric.rs).  Outputs N values
 * (eps, inv_eps) and N*4 tensor entries [xx, xy, yx, yy]. */
int oracle_dielectric(const OracleJob *job, double *eps, double *inv_eps, double *tensor);

/* Unnormalised forward 2D DFT (dir=-1) or inverse scaled by 1/(nx ny)
 * (dir=+1) of `count` row-major nx*ny fields, in place. */
void oracle_fft2d(int nx, int ny, int count, double *data, int dir);

/* Theta operator / mass / preconditioner applied to `count` fields at one
 * fractional k-point. which: 0 = A (Theta), 1 = B (mass), 2 = preconditioner
 * (job->precond, adaptive shift). */
int oracle_operator_apply(const OracleJob *job, const double kfrac[2], int which,
                          int count, const double *in, double *out);

/* Hermitian eigendecomposition (ascending) of an n x n column-major matrix,
 * lower triangle read.  w: n, v: n*n column-major eigenvectors. */
void oracle_eigh(int n, const double *a, double *w, double *v);

/* SVQB on p vectors of length len with mass vectors (normalization.rs:281-544).
 * Overwrites x and bx with the B-orthonormal output; returns the rank. */
int oracle_svqb(int len, int p, double *x, double *bx);

/* Polar factor U = W V^H of an m x m column-major matrix; returns sigma_min. */
double oracle_polar(int m, const double *o, double *u);

/* Hungarian maximum-weight assignment on an n x n row-major weight matrix:
 * perm[row] = column. */
void oracle_hungarian(int n, const double *w, int *perm);

/* LOBPCG on the diagonal test operator A = diag(diag), B = I, bloch (0.1,0.1)
 * (crates/core/src/_tests_eigensolver.rs:17-119).  Writes n_bands
 * eigenvalues; returns iterations (negative if not converged). */
int oracle_lobpcg_diagonal(int len, const double *diag, int n_bands, int max_iter, double tol,
                           int use_precond, double *eigenvalues);

#ifdef __cplusplus
}
#endif
#endif
