// Batched tall-skinny dense kernels (Gram, linear combination, residual) and
// the batched small Hermitian eigensolver.
#pragma once

#include "common.h"

namespace bh {

// G(i, j) = sum_r conj(A_i[r]) w[r] B_j[r], partial sums per row chunk.
template <typename T>
struct GramWork {
    const cplx<T> *A;
    const cplx<T> *B;
    const T *weight;        // optional real weight per row (eps for B-inner products)
    const cplx<T> *const *a_ptrs;  // optional per-column pointers for A (else A + c*a_stride)
    const cplx<T> *const *b_ptrs;  // optional per-column pointers for B
    const int *dev_na;      // optional device-side counts
    const int *dev_nb;
    double *partial;        // [nchunks][ld][ld] complex (interleaved re, im)
    long a_stride, b_stride;
    int na, nb, ld;
    int lower_only;         // Hermitian products: only i >= j (16 x 16 tiles on or below the diagonal)
};

// out_j = sum_i A_i C(i, j)            (mode 0)
// out_j = base_j - sum_i A_i C(i, j)   (mode 1; base may alias out)
// optional out2_j = w * out_j          (mass output, TM)
template <typename T>
struct LincombWork {
    const cplx<T> *A;
    const cplx<T> *const *a_ptrs;  // optional per-column pointers for A
    const int *dev_na;
    const int *dev_nq;
    const double *C;        // complex coefficients, C[i + j*ldc]
    const cplx<T> *base;
    cplx<T> *out;
    cplx<T> *out2;
    const T *w2;
    long a_stride, base_stride, out_stride;
    int na, nq, ldc, mode;
};

// Batched Hermitian eigenproblems on row-major partial Gram sums.
struct EighWork {
    const double *partial;  // [nchunks][ld][ld] complex
    double *w;              // eigenvalues (ascending for RR, kept descending for SVQB)
    double *V;              // RR: eigenvectors col-major (leading dim ldv); SVQB: T col-major n x rank
    double *Vw;             // RR: copy of columns [cnt_x, cnt_x + cnt_w) (history directions W)
    int *rank_out;          // SVQB: output rank; RR: n
    int *cnt_x;             // RR: min(nsplit, n)
    int *cnt_w;             // RR: min(2 nsplit, n) - cnt_x
    const int *dev_n;
    int n, nchunks, ld, mode;  // mode 0 = Rayleigh-Ritz, 1 = SVQB
    int ldv, nsplit;
    int *sweeps_out;        // optional: Jacobi sweeps used (diagnostics)
};

// Both Rayleigh-Ritz Grams in one pass over the rows of the subspace Z:
// G1 = Z^H (B Z) and G2 = Z^H (A Z), lower triangles, partial sums per row
// chunk (row-major [chunk][ld][ld] complex, entry i*ld + j for i >= j).
template <typename T>
struct Gram2Work {
    const cplx<T> *Z, *AZ;  // column j at base + j*stride
    const T *bweight;       // B Z = bweight * Z (TM mass); null: B = I
    double *partial1, *partial2;
    long stride;
    int n, ld;
};

// Rayleigh-Ritz on the dual Gram: B-orthonormalising transform T of Z from the
// scaled G1 (Cholesky QR, or SVQB when a pivot shows G1 ill-conditioned),
// projected operator M = T^H G2 T, its eigendecomposition M = Y diag(theta) Y^H,
// and the coefficients C = T Y of the new X (first cnt_x Ritz vectors) and W
// (next cnt_w) as combinations of the columns of Z.
struct RRWork {
    const double *partial1, *partial2;  // from the dual Gram
    double *lam;            // ascending Ritz values (rank of them)
    double *C;              // col-major n x (cnt_x + cnt_w), leading dim ldc, complex interleaved
    int *cnt;               // [cnt_x, cnt_w, cnt_x + cnt_w, Jacobi sweeps run, SVQB path taken (0/1), rank]
    int n, nchunks, ld, ldc, nsplit, pad;
    int *sweeps_out;        // optional diagnostics (see rr_kernel)
    double *ws;             // orders above 48: device workspace of rr_wide_bytes(n) (T, eigenvectors, G2 T)
};

constexpr int GRAM_MAXC = 48;   // columns of the one-workgroup subspace kernels (up to 16 bands)
constexpr int GRAM_WIDE = 128;  // columns of the wide (tiled / multi-workgroup) paths
constexpr int EIGH_MAXN = 64;

// Gram products of up to GRAM_WIDE columns; in FP64, products wider than
// mfma_min - 1 columns run on the matrix cores (gram_mfma_kernel)
constexpr int GRAM_MFMA_MIN_DEFAULT = 33;
template <typename T>
void launch_gram(const GramWork<T> *dwork, int nwork, int nchunks, long rows, int maxa, int maxb,
                 hipStream_t st, int mfma_min = GRAM_MFMA_MIN_DEFAULT);
void launch_gram_reduce(const double *partial, double *out, int nwork, int nchunks, int ld, long part_stride,
                        long out_stride, hipStream_t st);
// Batched reduce with per-item partial/output pointers and an optional
// uint8 mask (col-major ld x ld) zeroing selected coefficients.
struct ReduceWork {
    const double *partial;
    double *out;
    const unsigned char *mask;
    int nchunks, pad;
};
void launch_reduce(const ReduceWork *dwork, int nwork, int ld, hipStream_t st);

// Block deflation of a columns against nd vectors (deflation.rs:318-332,
// eigensolver.rs:859-865,978-1008):
//   r_c <- r_c - sum_i d_i <B d_i, r_c>,  <B d, r> = sum conj(d) w r,
// every coefficient taken from the undeflated r_c; a pair (i, c) whose
// mask[i + c*ld] is 0 is skipped.  One launch replaces Gram + reduce + lincomb.
template <typename T>
struct DeflateWork {
    const cplx<T> *const *d_ptrs;  // nd vectors
    const T *weight;               // optional B weight per row (TM: eps)
    cplx<T> *R;                    // a columns, stride apart, updated in place
    const unsigned char *mask;     // optional, col-major ld x ld
    // optional relative residual B-norms of the deflated columns
    // (eigensolver.rs:744-779): rel[act[c]] = ||r_c||_B / max(|lam[act[c]]|, 1e-10)
    const int *act;
    const double *lam;
    double *rel;
    long stride;
    int nd, a, ld, pad;
};
constexpr int DEFL_MAXD = 24;  // most deflation vectors per item
constexpr long DEFL_MAXROWS = 4096;  // longest column the one-launch deflation holds in registers
template <typename T>
void launch_deflate(const DeflateWork<T> *dwork, int nwork, long rows, int maxa, int maxd, hipStream_t st);
// maxna: most inputs of an item (above 48 selects the wide variant)
template <typename T>
void launch_lincomb(const LincombWork<T> *dwork, int nwork, long rows, int maxq, hipStream_t st,
                    int maxna = GRAM_MAXC);
void launch_eigh(const EighWork *dwork, int nwork, int maxn, hipStream_t st);
// maxn: largest n in the launch (selects the 36- or 48-column kernel)
template <typename T>
void launch_gram2(const Gram2Work<T> *dwork, int nwork, int nchunks, long rows, int maxn, hipStream_t st);
// Sums the dual Gram's row-chunk partials (lower triangles) into chunk 0 of
// each product, in chunk order from 0.0, so a reader loads one block.
template <typename T>
void launch_gram2_reduce(const Gram2Work<T> *dwork, int nwork, int nchunks, int maxn, hipStream_t st);
// The same for a Hermitian Gram of launch_gram (na = nb, lower triangle).
template <typename T>
void launch_gram_lower_reduce(const GramWork<T> *dwork, int nwork, int nchunks, int maxn, hipStream_t st);
// orders up to 48: one 512-thread workgroup per problem, all in LDS; above
// 48 (up to 128): one 1024-thread workgroup per problem with the projected
// matrix in LDS and T, the eigenvectors and G2 T in the item's workspace
void launch_rr(const RRWork *dwork, int nwork, int maxn, hipStream_t st);
size_t rr_wide_bytes(int n);
size_t eigh_shared_bytes(int maxn);

}  // namespace bh
