// Elementwise and reduction helpers for the band-solve engine and the
// trait-level buffer API.
#pragma once

#include "common.h"

namespace bh {

// r_c = A x_{act[c]} - lambda_{act[c]} B x_{act[c]}   (eigensolver.rs:707-738)
// With rel set, also the relative residual B-norm of each column
// (eigensolver.rs:744-779): rel[act[c]] = sqrt(Re r_c^H B r_c) / max(|lambda|, 1e-10)
template <typename T>
struct ResidualWork {
    const cplx<T> *ax;
    const cplx<T> *bx;    // B x, or x itself when bw is set
    const T *bw;          // optional: B x = bw * x (TM mass: bw = eps)
    cplx<T> *out;
    const int *act;       // band index of each active column
    const double *lam;
    double *rel;          // optional: residual norms, indexed by band
    long stride;          // column stride for ax, bx, out
    int count, pad;
};

// lambda_c = Re(x^H A x) / Re(x^H B x)   (initialization.rs:52-57)
template <typename T>
struct RayleighWork {
    const cplx<T> *x;
    const cplx<T> *ax;
    const cplx<T> *bx;    // B x, or x itself when bw is set
    const T *bw;          // optional: B x = bw * x
    double *lam;
    long stride;
    int count, pad;
};

// Relative residual B-norms of columns already formed (the row-chunked
// deflation path): rel[act[c]] = sqrt(Re r_c^H B r_c) / max(|lambda|, 1e-10)
template <typename T>
struct NormWork {
    const cplx<T> *r;     // count columns, stride apart
    const T *bw;          // optional B weight (TM: eps); null: B = I
    const int *act;       // band index of each column
    const double *lam;    // null: the plain B-norm
    double *rel;          // indexed by band
    long stride;
    int count, pad;
};

// Low plane waves G = n1 b1 + n2 b2, |n1|, |n2| <= h, index
// g = (n2 + h)(2h + 1) + n1 + h, of fields on an nx x ny node grid.
// Projection: c_g = (1/N) sum_r x(r) exp(-2 pi i (n1 ix / nx + n2 iy / ny));
// synthesis:  x(r) = sum_g c_g exp(+2 pi i (n1 ix / nx + n2 iy / ny)).
template <typename T>
struct PwWork {
    const cplx<T> *x;     // projection: count columns, stride apart
    double *c;            // projection output: count x (2h+1)^2 complex
    const double *cin;    // synthesis input: count x (2h+1)^2 complex
    cplx<T> *out;         // synthesis output: count columns, stride apart
    long stride;
    int count, pad;
};
template <typename T>
void launch_pw_project(const PwWork<T> *dwork, int nwork, int maxcount, int nx, int ny, int h, hipStream_t st);
template <typename T>
void launch_pw_synth(const PwWork<T> *dwork, int nwork, int maxcount, int nx, int ny, int h, hipStream_t st);

template <typename T>
void launch_residual(const ResidualWork<T> *dwork, int nwork, int maxcount, long rows, hipStream_t st);
template <typename T>
void launch_bnorm(const NormWork<T> *dwork, int nwork, int maxcount, long rows, hipStream_t st);
template <typename T>
void launch_rayleigh(const RayleighWork<T> *dwork, int nwork, int maxcount, long rows, hipStream_t st);
// xorshift start vectors, phase = first_phase + column (initialization.rs:113-125)
template <typename T>
void launch_random(cplx<T> *out, int ncols, long rows, double first_phase, hipStream_t st);
// out = value (complex), optional bout = value * w
template <typename T>
void launch_fill(cplx<T> *out, cplx<T> *bout, const T *w, double re, double im, long rows, hipStream_t st);

// trait-level ops (backend-cpu/src/lib.rs:414-443)
template <typename T>
void launch_scale(cplx<T> *x, long n, double re, double im, hipStream_t st);
template <typename T>
void launch_axpy(const cplx<T> *x, cplx<T> *y, long n, double re, double im, hipStream_t st);
template <typename T>
void launch_dot(const cplx<T> *x, const cplx<T> *y, long n, double *out2, hipStream_t st);

}  // namespace bh
