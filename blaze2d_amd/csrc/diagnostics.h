// One-time operator diagnostics of a band-structure job on the GPU: the
// self-adjointness check and the power-iteration condition estimates of the
// Theta operator and of the preconditioned operator (maxwell.rs:382-582, as
// drivers/bandstructure.rs:437-498 runs them), batched over many jobs.
#pragma once

#include "common.h"

namespace bh {

// One job's columns and outputs for diag_reduce.  mode 0 (power-iteration
// step): out = Re(v^H av), then av /= |av| when |av| > 1e-15.  mode 1
// (self-adjointness): out = |<Ax, y> - <x, Ay>| / (|Ax| |y|), 0 when that
// scale is <= 1e-15; v = x, av = Ax, y, ay = Ay.  mode 2: v /= |v| when
// |v| > 1e-15.
struct DiagItem {
    cplx<double> *v;
    cplx<double> *av;
    const cplx<double> *y;
    const cplx<double> *ay;
    double *out;
};

// x and y start vectors of the reference (maxwell.rs:386-391,529-538) in the
// columns x0 + j * stride and y0 + j * stride, j < njobs.
void launch_diag_init(cplx<double> *x0, cplx<double> *y0, long stride, int njobs, long n, hipStream_t st);
void launch_diag_reduce(const DiagItem *d_items, int nitems, int mode, long n, hipStream_t st);
// eps_eff and mass_floor of each job's preconditioner from its eps table
// (arithmetic mean for TM, harmonic mean for TE, as fourier_diagonal.rs:138-189)
void launch_diag_params(OpParams *d_params, const double *d_eps, int njobs, long n, hipStream_t st);

}  // namespace bh
