// k-point transitions of the band-solve engine on the device: band tracking,
// subspace-history prediction and block initialisation, in the coefficient
// space of the transition basis K = [E | P | R | y0 | Wc] over its B-weighted
// Gram M = K^H B K (one workgroup of one wave per slot).
#pragma once

#include "common.h"

namespace bh {

enum TransFlags {
    TR_TRACK = 1,    // band tracking between the previous k-point's vectors (P) and E
    TR_PREDICT = 2,  // predict_with_rotation from the history (else the tracked E)
    TR_EXTRAP = 4,   // extrapolation allowed and the path does not turn a corner
    TR_GAMMA = 8,    // the new k-point is Gamma: deflate y0 from the block
};

// One slot's transition.  K's columns: E (nE finished eigenvectors, ascending
// eigenvalue), P (nP vectors of the current history), R (nR start vectors),
// y0 (the Gamma mode), Wc (nWc learned-start vectors, first k-point only).
struct TransWork {
    const double *gram;     // Gram partial sums: M(i, j), i >= j, is the sum over c < nchunks of the
                            // complex at 2 (c ld ld + i ld + j)
    const double *om;       // nE frequencies of the finished solve (ascending)
    const double *prev_om;  // nprev tracked frequencies of the previous k-point
    double *cx;             // out: kd x m0 block coefficients over K, ld kd
    double *hist;           // out: nE x nE history-copy coefficients (H_new_i = E_perm[i]), ld nE
    int *out;               // out: [0] block vectors built (-1 start vectors ran out, -2 shape error),
                            //      [1] 1 when the tracking overlap had full rank, [2, 2 + nE) tracking permutation
    char *scratch;          // per-item workspace of transition_bytes() (null: dynamic LDS)
    double alpha;           // extrapolation factor |dk_next| / |dk_prev|
    int ld, nchunks, kd, nE, nP, nR, nWc, nb;
    int m0;                 // block vectors to build (0: tracking only)
    int nprev, flags;
};

// Capacities of one launch: KD coefficient rows, NBW warm vectors, M0 block
// vectors, MM order of the polar problems (at most 32), NC candidates of the
// block (warm vectors and start vectors).
size_t transition_bytes(int KD, int NBW, int M0, int MM, int NC);
// scratch: items bring their own workspace (else dynamic LDS of transition_bytes)
void launch_transition(const TransWork *dwork, int nwork, int KD, int NBW, int M0, int MM, int NC, bool scratch,
                       hipStream_t st);

}  // namespace bh
