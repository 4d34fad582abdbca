// Parallel cyclic Jacobi for small Hermitian matrices held in LDS (or any
// memory one workgroup addresses), shared by the batched eigensolver, the
// Rayleigh-Ritz kernels and the k-point transition kernel.
#pragma once

#include "common.h"

namespace bh {

// One plane rotation of a Jacobi round, packed so that a reader needs two
// 16-byte LDS loads: cosine c, sigma = s conj(h / |h|), and p | q << 8 | flag << 16
// with flag 0 identity, 1 rotate, 2 negligible (zero the pair).
struct alignas(16) JacobiRot {
    double c;
    int pk, pad;
    double2 s;
};

// Parallel cyclic Jacobi for the Hermitian matrix A held in LDS (col-major,
// leading dimension ld, order ne even, padding rows/columns zero), with the
// rotations accumulated into V.  All NT threads of the workgroup call it
// (NT >= 64, ne / 2 <= 32); rot holds 2 * ne / 2 entries (two consecutive
// rounds), red one double.
//
// Each round applies ne/2 disjoint rotations (round-robin pairing).  The matrix
// is partitioned into np x np 2x2 blocks (row pair, column pair); the thread
// owning a block applies both the column rotation of its column pair and the
// row rotation of its row pair in registers.  The rotations of round r are
// computed by the first np threads while the other threads apply round r-1's
// rotations to V; a round needs two barriers.  Each rotating thread keeps the
// largest relative rotation |a_pq|^2 / |a_pp a_qq| of the sweep in a register;
// one wave reduction per sweep ends the iteration when it is below
// stop_ratio^2 (quadratic convergence leaves the off-diagonal near stop_ratio^2
// relative).  An odd leading dimension keeps the 16-byte LDS accesses free of
// bank conflicts.  Returns the number of sweeps (at most 40).
template <int NT>
__device__ __forceinline__ int jacobi_lds(double2 *A, double2 *V, int ld, int ne, JacobiRot *rot, double *red,
                          double stop_ratio = 1e-9) {
    const int tid = threadIdx.x;
    const int np = ne / 2;
    const int nblk = np * np;
    const int nv = np * ne;  // eigenvector items (pair, row)
    const double stop2 = stop_ratio * stop_ratio;
    // V items of one round, rotations from parity slot `par`; threads [first, NT) share them
    auto rotate_v = [&](int par, int first) {
        const int stride = NT - first;
        int e = tid - first;
        if (e < 0) return;
        int t = e / ne, k = e - t * ne;
        const int dt = stride / ne, dk = stride - dt * ne;
        for (; e < nv; e += stride) {
            const JacobiRot R = rot[par * np + t];
            if (((R.pk >> 16) & 3) == 1) {
                const int p = R.pk & 255, q = (R.pk >> 8) & 255;
                const double c = R.c;
                const double2 sg = R.s;
                const double2 vkp = V[k + p * ld], vkq = V[k + q * ld];
                V[k + p * ld] = make_double2(c * vkp.x + sg.x * vkq.x - sg.y * vkq.y,
                                             c * vkp.y + sg.x * vkq.y + sg.y * vkq.x);
                V[k + q * ld] = make_double2(-(sg.x * vkp.x + sg.y * vkp.y) + c * vkq.x,
                                             -(sg.x * vkp.y - sg.y * vkp.x) + c * vkq.y);
            }
            t += dt;
            k += dk;
            if (k >= ne) { k -= ne; ++t; }
        }
    };
    int sweep = 0;
    for (; sweep < 40; ++sweep) {
        // rotating threads: 1 once a rotation of this sweep had relative size
        // |a_pq|^2 / |a_pp a_qq| >= stop2 (a_pp a_qq = 0 counts as large)
        double my_max2 = 0.0;
        for (int r = 0; r < ne - 1; ++r) {
            const int par = r & 1;
            if (tid < np) {
                // round-robin pairing, pair slot t: slot 0 (r, ne-1); slot t ((r+t) mod (ne-1), (r-t) mod (ne-1))
                // (both sums lie in [0, 2(ne-1)), so one conditional subtraction reduces them)
                int p, q;
                if (tid == 0) { p = r; q = ne - 1; }
                else {
                    p = r + tid;
                    q = r - tid + ne - 1;
                    if (p >= ne - 1) p -= ne - 1;
                    if (q >= ne - 1) q -= ne - 1;
                }
                if (p > q) { int t = p; p = q; q = t; }
                const double2 h = A[p + q * ld];
                const double hh = h.x * h.x + h.y * h.y;
                const double app = A[p + p * ld].x, aqq = A[q + q * ld].x;
                const double papq = fabs(app * aqq);
                JacobiRot R;
                R.c = 1.0;
                R.s = make_double2(0.0, 0.0);
                R.pad = 0;
                int flag = 0;
                if (hh != 0.0 && hh > 4.84e-32 * papq) {  // |h| > 2.2e-16 sqrt(|a_pp a_qq|)
                    flag = 1;
                    // t = sgn / u with u = |tau| + sqrt(1 + tau^2) >= 1, so
                    // c = 1 / sqrt(1 + t^2) = u w and |s| = |t| c = w with
                    // w = 1 / sqrt(u^2 + 1): no division, c^2 + |s|^2 = 1 to rounding
                    const double ri = rsqrt(hh);  // 1 / |h|
                    const double tau = (aqq - app) * 0.5 * ri;
                    const double at = fmin(fabs(tau), 1e100);  // u^2 stays finite
                    const double u = at + sqrt(1.0 + at * at);
                    const double wv = rsqrt(u * u + 1.0);
                    const double sr = (tau >= 0 ? -wv : wv) * ri;
                    R.c = u * wv;
                    R.s = make_double2(sr * h.x, -sr * h.y);  // sigma = s conj(h/|h|)
                    if (hh >= stop2 * papq) my_max2 = 1.0;
                } else if (hh != 0.0) {
                    flag = 2;
                }
                R.pk = p | (q << 8) | (flag << 16);
                rot[par * np + tid] = R;
            } else if (r > 0) {
                rotate_v(1 - par, np);  // previous round's rotations
            }
            __syncthreads();
            for (int b = tid; b < nblk; b += NT) {
                const int bt1 = b / np, bt2 = b - bt1 * np;
                const JacobiRot R1 = rot[par * np + bt1], R2 = rot[par * np + bt2];  // row pair, column pair
                const int f1 = (R1.pk >> 16) & 3, f2 = (R2.pk >> 16) & 3;
                const int p1 = R1.pk & 255, q1 = (R1.pk >> 8) & 255, p2 = R2.pk & 255, q2 = (R2.pk >> 8) & 255;
                if (f1 != 1 && f2 != 1) {
                    if (bt1 == bt2 && f1 == 2) {
                        A[p1 + q1 * ld] = make_double2(0.0, 0.0);
                        A[q1 + p1 * ld] = make_double2(0.0, 0.0);
                    }
                    continue;
                }
                double2 a11 = A[p1 + p2 * ld], a12 = A[p1 + q2 * ld];
                double2 a21 = A[q1 + p2 * ld], a22 = A[q1 + q2 * ld];
                if (f2 == 1) {  // columns: [:,p] = c [:,p] + s [:,q]; [:,q] = -conj(s) [:,p] + c [:,q]
                    const double c = R2.c;
                    const double2 s = R2.s;
                    double2 n11 = make_double2(c * a11.x + s.x * a12.x - s.y * a12.y, c * a11.y + s.x * a12.y + s.y * a12.x);
                    double2 n12 = make_double2(-(s.x * a11.x + s.y * a11.y) + c * a12.x, -(s.x * a11.y - s.y * a11.x) + c * a12.y);
                    double2 n21 = make_double2(c * a21.x + s.x * a22.x - s.y * a22.y, c * a21.y + s.x * a22.y + s.y * a22.x);
                    double2 n22 = make_double2(-(s.x * a21.x + s.y * a21.y) + c * a22.x, -(s.x * a21.y - s.y * a21.x) + c * a22.y);
                    a11 = n11; a12 = n12; a21 = n21; a22 = n22;
                }
                if (f1 == 1) {  // rows: [p,:] = c [p,:] + conj(s) [q,:]; [q,:] = -s [p,:] + c [q,:]
                    const double c = R1.c;
                    const double2 s = R1.s;
                    double2 n11 = make_double2(c * a11.x + s.x * a21.x + s.y * a21.y, c * a11.y + s.x * a21.y - s.y * a21.x);
                    double2 n21 = make_double2(-(s.x * a11.x - s.y * a11.y) + c * a21.x, -(s.x * a11.y + s.y * a11.x) + c * a21.y);
                    double2 n12 = make_double2(c * a12.x + s.x * a22.x + s.y * a22.y, c * a12.y + s.x * a22.y - s.y * a22.x);
                    double2 n22 = make_double2(-(s.x * a12.x - s.y * a12.y) + c * a22.x, -(s.x * a12.y + s.y * a12.x) + c * a22.y);
                    a11 = n11; a12 = n12; a21 = n21; a22 = n22;
                }
                if (bt1 == bt2) {  // the rotated pair: exact zero off-diagonal, real diagonal
                    a12 = make_double2(0.0, 0.0);
                    a21 = make_double2(0.0, 0.0);
                    a11.y = 0.0;
                    a22.y = 0.0;
                }
                A[p1 + p2 * ld] = a11;
                A[p1 + q2 * ld] = a12;
                A[q1 + p2 * ld] = a21;
                A[q1 + q2 * ld] = a22;
            }
            __syncthreads();
        }
        rotate_v((ne - 2) & 1, 0);  // the sweep's last round
        if (tid < 64) {  // the rotating threads are all in wave 0
            double m = my_max2;
            for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
            if (tid == 0) red[0] = m;
        }
        __syncthreads();
        const bool done = red[0] < stop2;  // also true when the sweep needed no rotation
        __syncthreads();  // every thread has read red[0] before the next sweep writes it
        if (done) break;
    }
    return sweep < 40 ? sweep + 1 : 40;
}

}  // namespace bh
