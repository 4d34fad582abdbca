// The learned start: a small CNN over the eps(r) grid plus an MLP embedding of
// (k1, k2, polarisation, lattice) that predicts the initial LOBPCG block as
// coefficients on the low plane waves G = n1 b1 + n2 b2, |n1|, |n2| <= h.
//
// Network (input H x W, both multiples of 8; C = 32 channels):
//   x0  [2][H][W]   eps / 16 and 1 / eps on the node grid
//   e   = W_e2 relu(W_e1 [k1, k2, pol, hex] + b_e1) + b_e2         (4 C: one C-vector per layer)
//   a1  = relu(conv3x3(x0, stride 1) + b_1 + e_1)                   [C][H][W]
//   a_l = relu(conv3x3(a_{l-1}, stride 2) + b_l + e_l), l = 2..4     [C][H/2^(l-1)][W/2^(l-1)]
//   y   = W_h flatten(a4) + b_h                                     [nb][(2h+1)^2] complex
// conv3x3 with stride s: out[co][oy][ox] = sum_{ci,dy,dx} w[co][ci][dy][dx]
// in[ci][(s oy + dy - 1) mod Hin][(s ox + dx - 1) mod Win], periodic because
// the grid is one lattice cell.  Weights are held in FP32 (master copy, Adam
// state) and used as BF16 by the forward and backward passes, which
// accumulate in FP32.
//
// Loss, per sample (gauge- and degeneracy-invariant): the predicted block Y
// and the target U (the converged eigenvectors on the same plane waves, ng x
// nb) define the projectors P_Y = Y (Y^H Y)^-1 Y^H and P_U;
//   L = 1 - tr(P_Y P_U) / nb  in [0, 1],
// zero exactly when Y spans the span of U, whatever basis either uses.  With
// M = (I - P_Y) P_U Y (Y^H Y)^-1, dL/dRe Y = -(2/nb) Re M and dL/dIm Y =
// -(2/nb) Im M.  The batch loss is the mean over samples.  A ridge of 1e-10
// times the mean diagonal is added to Y^H Y and U^H U before inversion.
#include "cnn.h"

#include <algorithm>
#include <cmath>
#include <random>
#include <stdexcept>
#include <vector>

namespace bh {

namespace {

constexpr int CH = Cnn::CH;
constexpr int EMB_IN = Cnn::EMB_IN;
constexpr int EMB_OUT = Cnn::LAYERS * CH;
constexpr int NT = 256;
constexpr int LOSS_NT = 256;
constexpr int LOSS_MAXNB = 32;
constexpr size_t LOSS_MAX_LDS = 160 * 1024;
constexpr int WGRAD_NT = 320;

__device__ inline float bf(const __hip_bfloat16 &v) { return __bfloat162float(v); }
__device__ inline int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// pre-activation + relu of one conv layer, one output element per thread
__global__ void conv_fwd_kernel(const float *__restrict__ in, const __hip_bfloat16 *__restrict__ w,
                                const float *__restrict__ bias, const float *__restrict__ emb, int emb_ld,
                                float *__restrict__ out, int B, int cin, int hin, int win, int stride) {
    const int hout = hin / stride, wout = win / stride;
    const long total = (long)B * CH * hout * wout;
    for (long e = (long)blockIdx.x * NT + threadIdx.x; e < total; e += (long)gridDim.x * NT) {
        const int ox = (int)(e % wout);
        const int oy = (int)((e / wout) % hout);
        const int co = (int)((e / ((long)wout * hout)) % CH);
        const int b = (int)(e / ((long)wout * hout * CH));
        float acc = bias[co] + emb[(long)b * emb_ld + co];
        int iy[3], ix[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            iy[d] = wrap(oy * stride + d - 1, hin);
            ix[d] = wrap(ox * stride + d - 1, win);
        }
        for (int ci = 0; ci < cin; ++ci) {
            const float *ip = in + ((long)b * cin + ci) * hin * win;
            const __hip_bfloat16 *wp = w + ((long)co * cin + ci) * 9;
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                for (int dx = 0; dx < 3; ++dx) acc = fmaf(bf(wp[dy * 3 + dx]), ip[(long)iy[dy] * win + ix[dx]], acc);
        }
        out[e] = acc > 0.0f ? acc : 0.0f;
    }
}

// g <- g * (act > 0): the gradient at the pre-activation, in place
__global__ void relu_bwd_kernel(float *__restrict__ g, const float *__restrict__ act, long n) {
    for (long e = (long)blockIdx.x * NT + threadIdx.x; e < n; e += (long)gridDim.x * NT)
        if (!(act[e] > 0.0f)) g[e] = 0.0f;
}

// gin[b][ci][iy][ix] = sum over (co, dy, dx) of gpre[b][co][oy][ox] w[co][ci][dy][dx]
// for the output positions with s oy + dy - 1 = iy, s ox + dx - 1 = ix (mod the grid)
__global__ void conv_bwd_data_kernel(const float *__restrict__ gpre, const __hip_bfloat16 *__restrict__ w,
                                     float *__restrict__ gin, int B, int cin, int hin, int win, int stride) {
    const int hout = hin / stride, wout = win / stride;
    const long total = (long)B * cin * hin * win;
    for (long e = (long)blockIdx.x * NT + threadIdx.x; e < total; e += (long)gridDim.x * NT) {
        const int ix = (int)(e % win);
        const int iy = (int)((e / win) % hin);
        const int ci = (int)((e / ((long)win * hin)) % cin);
        const int b = (int)(e / ((long)win * hin * cin));
        float acc = 0.0f;
        for (int dy = 0; dy < 3; ++dy) {
            const int ty = wrap(iy - dy + 1, hin);  // = s oy
            if (ty % stride) continue;
            const int oy = ty / stride;
            for (int dx = 0; dx < 3; ++dx) {
                const int tx = wrap(ix - dx + 1, win);
                if (tx % stride) continue;
                const int ox = tx / stride;
                const float *gp = gpre + ((long)b * CH * hout + oy) * wout + ox;
                const __hip_bfloat16 *wp = w + (long)ci * 9 + dy * 3 + dx;
                for (int co = 0; co < CH; ++co)
                    acc = fmaf(bf(wp[(long)co * cin * 9]), gp[(long)co * hout * wout], acc);
            }
        }
        gin[e] = acc;
    }
}

// Weight and bias gradients as sums over (sample, output position):
//   dw[co][ci][dy][dx] = sum gpre[b][co][oy][ox] in[b][ci][s oy + dy - 1][s ox + dx - 1],
//   db[co] = sum gpre[b][co][oy][ox].
// A chunk is R output rows of one sample.  Each workgroup stages a chunk's
// gradient rows ([position][channel] in LDS) and the input rows they read
// (s (R - 1) + 3 rows per input channel), and moves on to chunk + gridDim.x.
// Thread t takes tap j = t / NGRP (j = 9 ci + 3 dy + dx, or j = 9 CIN for the
// bias) and COPT = CH / NGRP output channels: per position one input read and
// COPT gradient reads (the same for the wave's lanes of a group: broadcast)
// feed COPT FMAs.  Its sums go to one row of `partial`, which
// wgrad_reduce_kernel adds up in a fixed order.
template <int CIN, int COPT>
__global__ void __launch_bounds__(WGRAD_NT) wgrad_partial_kernel(const float *__restrict__ in,
                                                                 const float *__restrict__ gpre,
                                                                 float *__restrict__ partial, int B, int hin, int win,
                                                                 int stride, int R) {
    constexpr int PER = CIN * 9 + 1;
    constexpr int NGRP = CH / COPT;
    static_assert(PER * NGRP <= WGRAD_NT, "taps x channel groups exceed the workgroup");
    static_assert(COPT % 4 == 0 && CH % COPT == 0, "channels per thread: a multiple of 4 dividing CH");
    extern __shared__ __attribute__((aligned(16))) float wl[];
    const int hout = hin / stride, wout = win / stride;
    const int rows_in = stride * (R - 1) + 3;
    const int npos = R * wout;
    float *g_l = wl;                 // [npos][CH]
    float *i_l = wl + npos * CH;     // [CIN][rows_in][win]
    const int tid = threadIdx.x;
    const int tap = tid / NGRP, cg = tid - tap * NGRP;
    const bool active = tap < PER;
    const bool bias = tap == PER - 1;
    const int ci = active && !bias ? tap / 9 : 0, dy = active && !bias ? (tap % 9) / 3 : 0,
              dx = active && !bias ? tap % 3 : 0;
    float acc[COPT];
#pragma unroll
    for (int c = 0; c < COPT; ++c) acc[c] = 0.0f;
    const int per_b = hout / R, nchunks = B * per_b;
    for (int ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
        const int b = ch / per_b, oy0 = (ch - b * per_b) * R;
        __syncthreads();  // the previous chunk's reads are done
        for (int e = tid; e < npos * CH; e += WGRAD_NT) {
            const int co = e / npos, p = e - co * npos;
            g_l[p * CH + co] = gpre[((long)b * CH + co) * hout * wout + (long)oy0 * wout + p];
        }
        for (int e = tid; e < CIN * rows_in * win; e += WGRAD_NT) {
            const int c = e / (rows_in * win), r = (e / win) % rows_in, x = e % win;
            const int iy = wrap(stride * oy0 - 1 + r, hin);
            i_l[e] = in[(((long)b * CIN + c) * hin + iy) * win + x];
        }
        __syncthreads();
        if (active) {
            const float *irow = i_l + (ci * rows_in + dy) * win;
            for (int oy = 0; oy < R; ++oy) {
                const float *ir = irow + stride * oy * win;
                const float *gr = g_l + oy * wout * CH + cg * COPT;
                for (int ox = 0; ox < wout; ++ox) {
                    const float iv = bias ? 1.0f : ir[wrap(stride * ox + dx - 1, win)];
                    const float *gp = gr + ox * CH;
#pragma unroll
                    for (int c = 0; c < COPT; c += 4) {
                        const float4 g4 = *reinterpret_cast<const float4 *>(gp + c);
                        acc[c] = fmaf(g4.x, iv, acc[c]);
                        acc[c + 1] = fmaf(g4.y, iv, acc[c + 1]);
                        acc[c + 2] = fmaf(g4.z, iv, acc[c + 2]);
                        acc[c + 3] = fmaf(g4.w, iv, acc[c + 3]);
                    }
                }
            }
        }
    }
    if (active) {
        // row layout e = co (9 CIN + 1) + tap
#pragma unroll
        for (int c = 0; c < COPT; ++c) partial[(long)blockIdx.x * CH * PER + (long)(cg * COPT + c) * PER + tap] = acc[c];
    }
}

// gw / gb from the partial rows: e = co (9 cin + 1) + j, j = 9 cin the bias.
// A workgroup takes 64 outputs; its four waves add every fourth row each
// (consecutive lanes, consecutive outputs), then combine in a fixed order.
__global__ void __launch_bounds__(NT) wgrad_reduce_kernel(const float *__restrict__ partial, int nrows, int cin,
                                                          float *__restrict__ gw, float *__restrict__ gb) {
    __shared__ float red[NT];
    const int per = cin * 9 + 1, nout = CH * per;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int e = blockIdx.x * 64 + lane;
    float t = 0.0f;
    if (e < nout)
        for (int r = wv; r < nrows; r += NT / 64) t += partial[(long)r * nout + e];
    red[threadIdx.x] = t;
    __syncthreads();
    if (wv == 0 && e < nout) {
        const float v = ((red[lane] + red[64 + lane]) + red[128 + lane]) + red[192 + lane];
        const int co = e / per, j = e - co * per;
        if (j == per - 1) gb[co] = v;
        else gw[(long)co * (per - 1) + j] = v;
    }
}

// block reduction of one float per thread (NT threads); result valid in thread 0
__device__ inline float block_sum(float acc) {
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    __shared__ float part[NT / 64];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    float t = 0.0f;
    if (threadIdx.x == 0)
        for (int v = 0; v < NT / 64; ++v) t += part[v];
    return t;
}

// out[b out_ld + c] = sum over the plane of g[b][c] (block b CH + c)
__global__ void __launch_bounds__(NT) channel_sum_kernel(const float *__restrict__ g, float *__restrict__ out,
                                                         long plane, int out_ld) {
    const float *p = g + (long)blockIdx.x * plane;
    float acc = 0.0f;
    for (long i = threadIdx.x; i < plane; i += NT) acc += p[i];
    const float t = block_sum(acc);
    if (threadIdx.x == 0) out[(long)(blockIdx.x / CH) * out_ld + blockIdx.x % CH] = t;
}

// y[b][m] = sum_k W[m][k] x[b][k] + bias[m] (relu optional); one wavefront per
// (b, m), lanes split k so the weight row is read contiguously
__global__ void __launch_bounds__(NT) dense_fwd_kernel(const float *__restrict__ x,
                                                       const __hip_bfloat16 *__restrict__ W,
                                                       const float *__restrict__ bias, float *__restrict__ y, int B,
                                                       int K, int M, int relu) {
    const long nw = (long)B * M;
    const int lane = threadIdx.x & 63;
    for (long r = (long)blockIdx.x * (NT / 64) + (threadIdx.x >> 6); r < nw; r += (long)gridDim.x * (NT / 64)) {
        const int b = (int)(r / M), m = (int)(r - (long)b * M);
        const float *xp = x + (long)b * K;
        const __hip_bfloat16 *wp = W + (long)m * K;
        float acc = 0.0f;
        for (int k = lane; k < K; k += 64) acc = fmaf(bf(wp[k]), xp[k], acc);
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
        if (lane == 0) {
            acc += bias[m];
            y[r] = relu && !(acc > 0.0f) ? 0.0f : acc;
        }
    }
}

// dx[b][k] = sum_m dy[b][m] W[m][k] as S partial sums over slices of m.
// Workgroup (kt, s): k = NT kt + thread, m in slice s, samples in blocks of
// DBB with that block's dy[:, slice] staged in LDS ([m][b], read as float4
// broadcasts), so each weight read feeds DBB FMAs.  part[s][b][k].
constexpr int DBB = 64;
constexpr int MSUB = 128;  // m rows staged at a time (32 KB of LDS)
__global__ void __launch_bounds__(NT) dense_bwd_data_partial_kernel(const float *__restrict__ dy,
                                                                    const __hip_bfloat16 *__restrict__ W,
                                                                    float *__restrict__ part, int B, int K, int M,
                                                                    int ms) {
    __shared__ __attribute__((aligned(16))) float dl[MSUB * DBB];  // [m][b]
    const int k = blockIdx.x * NT + threadIdx.x;
    const int s = blockIdx.y;
    const int m0 = s * ms, m1 = min(M, m0 + ms);
    for (int b0 = 0; b0 < B; b0 += DBB) {
        const int nbb = min(DBB, B - b0);
        float acc[DBB];
#pragma unroll
        for (int i = 0; i < DBB; ++i) acc[i] = 0.0f;
        for (int mc = m0; mc < m1; mc += MSUB) {
            const int nm = min(MSUB, m1 - mc);
            __syncthreads();  // the previous sub-slice's reads are done
            for (int e = threadIdx.x; e < nm * DBB; e += NT) {
                const int mm = e / DBB, i = e - mm * DBB;
                dl[e] = i < nbb ? dy[(long)(b0 + i) * M + mc + mm] : 0.0f;
            }
            __syncthreads();
            if (k < K)
                for (int mm = 0; mm < nm; ++mm) {
                    const float w = bf(W[(long)(mc + mm) * K + k]);
                    const float *dp = dl + mm * DBB;
#pragma unroll
                    for (int i = 0; i < DBB; i += 4) {
                        const float4 d4 = *reinterpret_cast<const float4 *>(dp + i);
                        acc[i] = fmaf(d4.x, w, acc[i]);
                        acc[i + 1] = fmaf(d4.y, w, acc[i + 1]);
                        acc[i + 2] = fmaf(d4.z, w, acc[i + 2]);
                        acc[i + 3] = fmaf(d4.w, w, acc[i + 3]);
                    }
                }
        }
        if (k < K) {
#pragma unroll
            for (int i = 0; i < DBB; ++i)
                if (i < nbb) part[((long)s * B + b0 + i) * K + k] = acc[i];
        }
    }
}

// dx[b][k] = sum_s part[s][b][k]
__global__ void dense_bwd_data_reduce_kernel(const float *__restrict__ part, float *__restrict__ dx, int S, long n) {
    for (long e = (long)blockIdx.x * NT + threadIdx.x; e < n; e += (long)gridDim.x * NT) {
        float t = 0.0f;
        for (int s = 0; s < S; ++s) t += part[(long)s * n + e];
        dx[e] = t;
    }
}

// dW[m][k] = sum_b dy[b][m] x[b][k]; k = K gives the bias dbias[m] = sum_b dy[b][m]
__global__ void dense_bwd_weight_kernel(const float *__restrict__ dy, const float *__restrict__ x,
                                        float *__restrict__ dW, float *__restrict__ db, int B, int K, int M) {
    const long total = (long)M * (K + 1);
    for (long e = (long)blockIdx.x * NT + threadIdx.x; e < total; e += (long)gridDim.x * NT) {
        const int m = (int)(e / (K + 1)), k = (int)(e - (long)m * (K + 1));
        float acc = 0.0f;
        if (k == K) {
            for (int b = 0; b < B; ++b) acc += dy[(long)b * M + m];
            db[m] = acc;
        } else {
            for (int b = 0; b < B; ++b) acc = fmaf(dy[(long)b * M + m], x[(long)b * K + k], acc);
            dW[(long)m * K + k] = acc;
        }
    }
}

// In-place inverse of an n x n complex matrix (row-major, n <= LOSS_MAXNB)
// by Gauss-Jordan with partial pivoting, the whole workgroup cooperating.
// I: n x n workspace, F: n entries.
__device__ void small_inverse(double2 *A, double2 *I, double2 *F, int n) {
    const int tid = threadIdx.x;
    for (int e = tid; e < n * n; e += LOSS_NT) I[e] = make_double2((e / n) == (e % n) ? 1.0 : 0.0, 0.0);
    __syncthreads();
    for (int c = 0; c < n; ++c) {
        if (tid == 0) {
            int p = c;
            double best = -1.0;
            for (int r = c; r < n; ++r) {
                const double m = A[r * n + c].x * A[r * n + c].x + A[r * n + c].y * A[r * n + c].y;
                if (m > best) { best = m; p = r; }
            }
            if (p != c)
                for (int j = 0; j < n; ++j) {
                    double2 t = A[c * n + j]; A[c * n + j] = A[p * n + j]; A[p * n + j] = t;
                    t = I[c * n + j]; I[c * n + j] = I[p * n + j]; I[p * n + j] = t;
                }
        }
        __syncthreads();
        const double2 d = A[c * n + c];
        __syncthreads();  // every thread holds the pivot before row c changes
        const double dd = d.x * d.x + d.y * d.y;
        const double2 inv = make_double2(d.x / dd, -d.y / dd);
        for (int j = tid; j < 2 * n; j += LOSS_NT) {
            double2 *x = j < n ? &A[c * n + j] : &I[c * n + j - n];
            const double2 a = *x;
            *x = make_double2(a.x * inv.x - a.y * inv.y, a.x * inv.y + a.y * inv.x);
        }
        for (int r = tid; r < n; r += LOSS_NT) F[r] = A[r * n + c];
        __syncthreads();
        for (int e = tid; e < n * 2 * n; e += LOSS_NT) {
            const int r = e / (2 * n), j = e - r * 2 * n;
            if (r == c) continue;
            const double2 f = F[r];
            double2 *x = j < n ? &A[r * n + j] : &I[r * n + j - n];
            const double2 b = j < n ? A[c * n + j] : I[c * n + j - n];
            x->x -= f.x * b.x - f.y * b.y;
            x->y -= f.x * b.y + f.y * b.x;
        }
        __syncthreads();
    }
    for (int e = tid; e < n * n; e += LOSS_NT) A[e] = I[e];
    __syncthreads();
}

// One workgroup per sample.  y: [B][nb][ng][2] float (the head output), u:
// [B][nb][ng][2] double (targets, held as FP32 in LDS like y); every product
// accumulates in FP64.  Writes loss[b] and gy = d(mean loss)/dy.
__global__ void __launch_bounds__(LOSS_NT) projector_loss_kernel(const float *__restrict__ y,
                                                                 const double *__restrict__ u,
                                                                 double *__restrict__ loss, float *__restrict__ gy,
                                                                 int B, int nb, int ng) {
    const int b = blockIdx.x;
    const int tid = threadIdx.x;
    extern __shared__ __attribute__((aligned(16))) double lsm[];
    double2 *R = reinterpret_cast<double2 *>(lsm);  // [ng][nb]: P_U Y, then (I - P_Y) P_U Y
    double2 *Gy = R + (size_t)ng * nb;              // Y^H Y, then its inverse
    double2 *Gu = Gy + nb * nb;                     // U^H U, then its inverse
    double2 *A = Gu + nb * nb;
    double2 *T = A + nb * nb;                       // also the inversions' workspace
    double2 *Fv = T + nb * nb;
    float2 *Y = reinterpret_cast<float2 *>(Fv + nb);  // [ng][nb]
    float2 *U = Y + (size_t)ng * nb;                  // [ng][nb]
    const float *yb = y + (size_t)b * nb * ng * 2;
    const double *ub = u + (size_t)b * nb * ng * 2;
    for (int e = tid; e < ng * nb; e += LOSS_NT) {
        const int g = e / nb, j = e - g * nb;
        Y[e] = make_float2(yb[2 * (j * ng + g)], yb[2 * (j * ng + g) + 1]);
        U[e] = make_float2((float)ub[2 * (j * ng + g)], (float)ub[2 * (j * ng + g) + 1]);
    }
    __syncthreads();
    auto herm_prod = [&](const auto *P, const auto *Q, double2 *out) {  // P^H Q, nb x nb
        for (int e = tid; e < nb * nb; e += LOSS_NT) {
            const int i = e / nb, j = e - i * nb;
            double re = 0.0, im = 0.0;
            for (int g = 0; g < ng; ++g)
                cmacc(re, im, P[g * nb + i].x, P[g * nb + i].y, Q[g * nb + j].x, Q[g * nb + j].y);
            out[e] = make_double2(re, im);
        }
    };
    auto small_prod = [&](const double2 *P, const double2 *Q, double2 *out) {  // P Q, nb x nb
        for (int e = tid; e < nb * nb; e += LOSS_NT) {
            const int i = e / nb, j = e - i * nb;
            double re = 0.0, im = 0.0;
            for (int k = 0; k < nb; ++k)
                cmac(re, im, P[i * nb + k].x, P[i * nb + k].y, Q[k * nb + j].x, Q[k * nb + j].y);
            out[e] = make_double2(re, im);
        }
    };
    herm_prod(Y, Y, Gy);
    herm_prod(U, U, Gu);
    herm_prod(U, Y, A);
    __syncthreads();
    if (tid == 0) {
        double ty = 0.0, tu = 0.0;
        for (int i = 0; i < nb; ++i) { ty += Gy[i * nb + i].x; tu += Gu[i * nb + i].x; }
        for (int i = 0; i < nb; ++i) { Gy[i * nb + i].x += 1e-10 * ty / nb; Gu[i * nb + i].x += 1e-10 * tu / nb; }
    }
    __syncthreads();
    small_inverse(Gy, T, Fv, nb);
    small_inverse(Gu, T, Fv, nb);
    small_prod(Gu, A, T);  // (U^H U)^-1 U^H Y
    __syncthreads();
    for (int e = tid; e < ng * nb; e += LOSS_NT) {  // R = U T = P_U Y
        const int g = e / nb, j = e - g * nb;
        double re = 0.0, im = 0.0;
        for (int k = 0; k < nb; ++k) cmac(re, im, U[g * nb + k].x, U[g * nb + k].y, T[k * nb + j].x, T[k * nb + j].y);
        R[e] = make_double2(re, im);
    }
    __syncthreads();
    herm_prod(Y, R, A);  // Y^H P_U Y
    __syncthreads();
    small_prod(Gy, A, T);  // (Y^H Y)^-1 Y^H P_U Y; tr = tr(P_Y P_U)
    __syncthreads();
    if (tid == 0) {
        double tr = 0.0;
        for (int i = 0; i < nb; ++i) tr += T[i * nb + i].x;
        loss[b] = 1.0 - tr / nb;
    }
    // R <- R - Y T = (I - P_Y) P_U Y (each thread rewrites only its own entries)
    for (int e = tid; e < ng * nb; e += LOSS_NT) {
        const int g = e / nb, j = e - g * nb;
        double re = R[e].x, im = R[e].y;
        for (int k = 0; k < nb; ++k)
            cmac(re, im, -(double)Y[g * nb + k].x, -(double)Y[g * nb + k].y, T[k * nb + j].x, T[k * nb + j].y);
        R[e] = make_double2(re, im);
    }
    __syncthreads();
    const double sc = -2.0 / ((double)nb * B);
    float *gb = gy + (size_t)b * nb * ng * 2;
    for (int e = tid; e < ng * nb; e += LOSS_NT) {
        const int g = e / nb, j = e - g * nb;
        double re = 0.0, im = 0.0;
        for (int k = 0; k < nb; ++k) cmac(re, im, R[g * nb + k].x, R[g * nb + k].y, Gy[k * nb + j].x, Gy[k * nb + j].y);
        gb[2 * (j * ng + g)] = (float)(sc * re);
        gb[2 * (j * ng + g) + 1] = (float)(sc * im);
    }
}

// Adam with bias correction on the FP32 master weights, then their BF16 copy
// hyper = [learning rate, 1 - b1^t, 1 - b2^t] of step t, from device memory
__global__ void adam_kernel(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, __hip_bfloat16 *__restrict__ pb, long n,
                            const float *__restrict__ hyper, float b1, float b2, float eps) {
    const float lr = hyper[0], c1 = hyper[1], c2 = hyper[2];
    for (long e = (long)blockIdx.x * NT + threadIdx.x; e < n; e += (long)gridDim.x * NT) {
        const float ge = g[e];
        const float me = b1 * m[e] + (1.0f - b1) * ge;
        const float ve = b2 * v[e] + (1.0f - b2) * ge * ge;
        m[e] = me;
        v[e] = ve;
        const float pe = p[e] - lr * (me / c1) / (sqrtf(ve / c2) + eps);
        p[e] = pe;
        pb[e] = __float2bfloat16(pe);
    }
}

__global__ void to_bf16_kernel(const float *__restrict__ p, __hip_bfloat16 *__restrict__ pb, long n) {
    for (long e = (long)blockIdx.x * NT + threadIdx.x; e < n; e += (long)gridDim.x * NT) pb[e] = __float2bfloat16(p[e]);
}

// x0[b][0][i] = eps / 16, x0[b][1][i] = 1 / eps from eps[b][i] (FP64)
__global__ void eps_input_kernel(const double *__restrict__ eps, float *__restrict__ x0, int B, long plane) {
    const long total = (long)B * plane;
    for (long e = (long)blockIdx.x * NT + threadIdx.x; e < total; e += (long)gridDim.x * NT) {
        const long b = e / plane, i = e - b * plane;
        const double v = eps[e];
        x0[(2 * b) * plane + i] = (float)(v / 16.0);
        x0[(2 * b + 1) * plane + i] = (float)(1.0 / v);
    }
}

__global__ void fill_kernel(float *__restrict__ p, long n) {
    for (long e = (long)blockIdx.x * NT + threadIdx.x; e < n; e += (long)gridDim.x * NT) p[e] = 0.0f;
}

constexpr int WGRAD_MAX_WG = 512;
constexpr size_t WGRAD_MAX_LDS = 64 * 1024;

size_t wgrad_lds_bytes(int cin, int R, int win, int wout, int stride) {
    return sizeof(float) * ((size_t)R * wout * CH + (size_t)cin * (stride * (R - 1) + 3) * win);
}

// rows per chunk: the largest divisor of the output height whose staging fits
// in 64 KB of LDS (two workgroups per CU)
int wgrad_rows(int cin, int hin, int win, int stride) {
    const int hout = hin / stride, wout = win / stride;
    for (int R = hout; R >= 1; --R)
        if (hout % R == 0 && wgrad_lds_bytes(cin, R, win, wout, stride) <= WGRAD_MAX_LDS) return R;
    throw std::runtime_error("CNN layer too wide for the weight-gradient staging");
}

unsigned grid_for(long n) { return (unsigned)std::max<long>(1, std::min<long>((n + NT - 1) / NT, 16384)); }

size_t loss_lds_bytes(int nb, int ng) {
    return sizeof(double2) * ((size_t)ng * nb + 4 * (size_t)nb * nb + nb) + sizeof(float2) * 2 * (size_t)ng * nb;
}

}  // namespace

Cnn::Cnn(int h_, int w_, int nb_, int pw_half_, int max_batch_, unsigned long long seed, hipStream_t st_)
    : H(h_), W(w_), nb(nb_), pw_half(pw_half_), max_batch(max_batch_), st(st_) {
    if (H % 8 || W % 8 || H < 8 || W < 8 || H > 1024 || W > 1024)
        throw std::runtime_error("CNN input sides must be multiples of 8 in [8, 1024]");
    if (nb < 1 || nb > LOSS_MAXNB) throw std::runtime_error("CNN predicts 1..32 vectors");
    if (pw_half < 0 || pw_half > 8) throw std::runtime_error("CNN plane-wave half width must be 0..8");
    if (max_batch < 1 || max_batch > 4096) throw std::runtime_error("CNN batch must be 1..4096");
    ng = (2 * pw_half + 1) * (2 * pw_half + 1);
    if (ng < nb) throw std::runtime_error("CNN predicts more vectors than plane waves");
    F = CH * (H / 8) * (W / 8);
    OUT = nb * ng * 2;
    loss_lds = loss_lds_bytes(nb, ng);
    if (loss_lds > LOSS_MAX_LDS) throw std::runtime_error("projector loss: vectors x plane waves exceed the LDS");
    long o = 0;
    auto take = [&](long n) { long r = o; o += n; return r; };
    for (int l = 0; l < LAYERS; ++l) {
        off_c[l] = take((long)CH * (l == 0 ? 2 : CH) * 9);
        off_cb[l] = take(CH);
    }
    off_e1 = take((long)CH * EMB_IN);
    off_e1b = take(CH);
    off_e2 = take((long)EMB_OUT * CH);
    off_e2b = take(EMB_OUT);
    off_h = take((long)OUT * F);
    off_hb = take(OUT);
    P = o;
    // He initialisation (normal, variance 2 / fan-in; 1 / fan-in for the linear
    // layers), seeded; biases zero
    std::vector<float> hp(P, 0.0f);
    std::mt19937_64 rng(seed);
    auto init = [&](long off, long n, int fan_in, float gain) {
        std::normal_distribution<float> nd(0.0f, std::sqrt(gain / (float)fan_in));
        for (long i = 0; i < n; ++i) hp[off + i] = nd(rng);
    };
    for (int l = 0; l < LAYERS; ++l) init(off_c[l], (long)CH * (l == 0 ? 2 : CH) * 9, (l == 0 ? 2 : CH) * 9, 2.0f);
    init(off_e1, (long)CH * EMB_IN, EMB_IN, 2.0f);
    init(off_e2, (long)EMB_OUT * CH, CH, 1.0f);
    init(off_h, (long)OUT * F, F, 1.0f);
    auto dal = [&](size_t bytes) {
        void *p = nullptr;
        BH_HIP_CHECK(hipMalloc(&p, std::max<size_t>(bytes, 256)));
        bufs.push_back(p);
        return p;
    };
    params = (float *)dal(sizeof(float) * P);
    grads = (float *)dal(sizeof(float) * P);
    adam_m = (float *)dal(sizeof(float) * P);
    adam_v = (float *)dal(sizeof(float) * P);
    pbf = (__hip_bfloat16 *)dal(sizeof(__hip_bfloat16) * P);
    const long Bm = max_batch;
    x0 = (float *)dal(sizeof(float) * Bm * 2 * H * W);
    e0 = (float *)dal(sizeof(float) * Bm * EMB_IN);
    eh = (float *)dal(sizeof(float) * Bm * CH);
    ev = (float *)dal(sizeof(float) * Bm * EMB_OUT);
    for (int l = 0; l < LAYERS; ++l) {
        const long hw = (long)(H >> l) * (W >> l);
        act[l] = (float *)dal(sizeof(float) * Bm * CH * hw);
        gact[l] = (float *)dal(sizeof(float) * Bm * CH * hw);
    }
    yout = (float *)dal(sizeof(float) * Bm * OUT);
    gy = (float *)dal(sizeof(float) * Bm * OUT);
    geh = (float *)dal(sizeof(float) * Bm * CH);
    gev = (float *)dal(sizeof(float) * Bm * EMB_OUT);
    tgt = (double *)dal(sizeof(double) * Bm * OUT);
    wpart = (float *)dal(sizeof(float) * WGRAD_MAX_WG * CH * (CH * 9 + 1));
    // dense backward partials: at least one slice per sample block for either layer
    dpart_floats = std::max<size_t>((size_t)64 << 20 >> 2, (size_t)Bm * std::max(F, CH));
    dpart = (float *)dal(sizeof(float) * dpart_floats);
    for (int l = 0; l < LAYERS; ++l) {  // the staging of every layer must fit
        const int cin = l == 0 ? 2 : CH, hin = l == 0 ? H : (H >> (l - 1)), win = l == 0 ? W : (W >> (l - 1));
        (void)wgrad_rows(cin, hin, win, l == 0 ? 1 : 2);
    }
    lossb = (double *)dal(sizeof(double) * Bm);
    d_hyper = (float *)dal(sizeof(float) * 4);
    BH_HIP_CHECK(hipHostMalloc((void **)&h_hyper, sizeof(float) * 4, hipHostMallocDefault));
    BH_HIP_CHECK(hipFuncSetAttribute((const void *)projector_loss_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)LOSS_MAX_LDS));
    BH_HIP_CHECK(hipFuncSetAttribute((const void *)wgrad_partial_kernel<2, 4>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)WGRAD_MAX_LDS));
    BH_HIP_CHECK(hipFuncSetAttribute((const void *)wgrad_partial_kernel<CH, CH>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)WGRAD_MAX_LDS));
    set_params(hp.data());
}

Cnn::~Cnn() {
    (void)hipStreamSynchronize(st);
    for (auto &g : graphs) (void)hipGraphExecDestroy(g.second);
    if (h_hyper) (void)hipHostFree(h_hyper);
    for (void *p : bufs) (void)hipFree(p);
}

void Cnn::upload_inputs(int B, const float *x, const float *emb) {
    if (B < 1 || B > max_batch) throw std::runtime_error("CNN batch outside 1..max_batch");
    BH_HIP_CHECK(hipMemcpyAsync(x0, x, sizeof(float) * B * 2 * H * W, hipMemcpyHostToDevice, st));
    BH_HIP_CHECK(hipMemcpyAsync(e0, emb, sizeof(float) * B * EMB_IN, hipMemcpyHostToDevice, st));
}

void Cnn::inputs_from_geometries(int B, const DielGeom *geoms, bool smoothing, const float *emb) {
    if (B < 1 || B > max_batch) throw std::runtime_error("CNN batch outside 1..max_batch");
    if (!d_eps) {
        BH_HIP_CHECK(hipMalloc((void **)&d_eps, sizeof(double) * max_batch * H * W));
        bufs.push_back(d_eps);
        BH_HIP_CHECK(hipMalloc(&d_geo, dielectric_geom_bytes(max_batch)));
        bufs.push_back(d_geo);
    }
    launch_dielectric(geoms, B, W, H, smoothing, d_eps, nullptr, nullptr, d_geo, st);
    const long plane = (long)H * W;
    hipLaunchKernelGGL(eps_input_kernel, dim3(grid_for(B * plane)), dim3(NT), 0, st, d_eps, x0, B, plane);
    BH_HIP_CHECK(hipGetLastError());
    BH_HIP_CHECK(hipMemcpyAsync(e0, emb, sizeof(float) * B * EMB_IN, hipMemcpyHostToDevice, st));
}

void Cnn::forward(int B) {
    if (B < 1 || B > max_batch) throw std::runtime_error("CNN batch outside 1..max_batch");
    const __hip_bfloat16 *wb = pbf;
    hipLaunchKernelGGL(dense_fwd_kernel, dim3(grid_for((long)B * CH * 64)), dim3(NT), 0, st, e0, wb + off_e1,
                       params + off_e1b, eh, B, EMB_IN, CH, 1);
    hipLaunchKernelGGL(dense_fwd_kernel, dim3(grid_for((long)B * EMB_OUT * 64)), dim3(NT), 0, st, eh, wb + off_e2,
                       params + off_e2b, ev, B, CH, EMB_OUT, 0);
    const float *in = x0;
    for (int l = 0; l < LAYERS; ++l) {
        const int cin = l == 0 ? 2 : CH, hin = l == 0 ? H : (H >> (l - 1)), win = l == 0 ? W : (W >> (l - 1));
        const int stride = l == 0 ? 1 : 2;
        const long n = (long)B * CH * (hin / stride) * (win / stride);
        hipLaunchKernelGGL(conv_fwd_kernel, dim3(grid_for(n)), dim3(NT), 0, st, in, wb + off_c[l],
                           params + off_cb[l], ev + l * CH, EMB_OUT, act[l], B, cin, hin, win, stride);
        in = act[l];
    }
    hipLaunchKernelGGL(dense_fwd_kernel, dim3(grid_for((long)B * OUT * 64)), dim3(NT), 0, st, act[LAYERS - 1],
                       wb + off_h, params + off_hb, yout, B, F, OUT, 0);
    BH_HIP_CHECK(hipGetLastError());
}

// dx = dy W (dy: B x M, W: M x K): partial sums over S slices of M, enough
// workgroups to fill the device, within the partial buffer
void Cnn::dense_bwd_data(const float *dy, const __hip_bfloat16 *Wm, float *dx, int B, int K, int M) {
    const int kt = (K + NT - 1) / NT;
    long S = std::max(1, 1024 / kt);
    S = std::min<long>(S, M);
    S = std::min<long>(S, (long)(dpart_floats / ((size_t)B * K)));
    if (S < 1) throw std::runtime_error("CNN dense backward: partial buffer too small");
    const int ms = (int)((M + S - 1) / S);
    S = (M + ms - 1) / ms;
    hipLaunchKernelGGL(dense_bwd_data_partial_kernel, dim3(kt, (unsigned)S), dim3(NT), 0, st, dy, Wm, dpart, B, K, M,
                       ms);
    const long n = (long)B * K;
    hipLaunchKernelGGL(dense_bwd_data_reduce_kernel, dim3(grid_for(n)), dim3(NT), 0, st, dpart, dx, (int)S, n);
    BH_HIP_CHECK(hipGetLastError());
}

void Cnn::step_launches(int B, bool update) {
    forward(B);
    hipLaunchKernelGGL(projector_loss_kernel, dim3(B), dim3(LOSS_NT), loss_lds, st, yout, tgt, lossb, gy, B, nb, ng);
    const __hip_bfloat16 *wb = pbf;
    hipLaunchKernelGGL(dense_bwd_weight_kernel, dim3(grid_for((long)OUT * (F + 1))), dim3(NT), 0, st, gy,
                       act[LAYERS - 1], grads + off_h, grads + off_hb, B, F, OUT);
    dense_bwd_data(gy, wb + off_h, gact[LAYERS - 1], B, F, OUT);
    for (int l = LAYERS - 1; l >= 0; --l) {
        const int cin = l == 0 ? 2 : CH, hin = l == 0 ? H : (H >> (l - 1)), win = l == 0 ? W : (W >> (l - 1));
        const int stride = l == 0 ? 1 : 2;
        const long plane = (long)(hin / stride) * (win / stride);
        const long n = (long)B * CH * plane;
        hipLaunchKernelGGL(relu_bwd_kernel, dim3(grid_for(n)), dim3(NT), 0, st, gact[l], act[l], n);
        hipLaunchKernelGGL(channel_sum_kernel, dim3(B * CH), dim3(NT), 0, st, gact[l], gev + l * CH, plane, EMB_OUT);
        const float *in = l == 0 ? x0 : act[l - 1];
        {
            const int hout = hin / stride, wout = win / stride;
            const int R = wgrad_rows(cin, hin, win, stride);
            const int nchunks = B * (hout / R);
            const int nwg = std::min(nchunks, WGRAD_MAX_WG);
            const size_t lds = wgrad_lds_bytes(cin, R, win, wout, stride);
            if (cin == 2)
                hipLaunchKernelGGL((wgrad_partial_kernel<2, 4>), dim3(nwg), dim3(WGRAD_NT), lds, st, in, gact[l],
                                   wpart, B, hin, win, stride, R);
            else
                hipLaunchKernelGGL((wgrad_partial_kernel<CH, CH>), dim3(nwg), dim3(WGRAD_NT), lds, st, in, gact[l],
                                   wpart, B, hin, win, stride, R);
            const int nout = CH * (cin * 9 + 1);
            hipLaunchKernelGGL(wgrad_reduce_kernel, dim3((nout + 63) / 64), dim3(NT), 0, st, wpart, nwg, cin,
                               grads + off_c[l], grads + off_cb[l]);
        }
        if (l > 0)
            hipLaunchKernelGGL(conv_bwd_data_kernel, dim3(grid_for((long)B * cin * hin * win)), dim3(NT), 0, st,
                               gact[l], wb + off_c[l], gact[l - 1], B, cin, hin, win, stride);
    }
    hipLaunchKernelGGL(dense_bwd_weight_kernel, dim3(grid_for((long)EMB_OUT * (CH + 1))), dim3(NT), 0, st, gev, eh,
                       grads + off_e2, grads + off_e2b, B, CH, EMB_OUT);
    dense_bwd_data(gev, wb + off_e2, geh, B, CH, EMB_OUT);
    hipLaunchKernelGGL(relu_bwd_kernel, dim3(grid_for((long)B * CH)), dim3(NT), 0, st, geh, eh, (long)B * CH);
    hipLaunchKernelGGL(dense_bwd_weight_kernel, dim3(grid_for((long)CH * (EMB_IN + 1))), dim3(NT), 0, st, geh, e0,
                       grads + off_e1, grads + off_e1b, B, EMB_IN, CH);
    BH_HIP_CHECK(hipGetLastError());
    if (update) {
        hipLaunchKernelGGL(adam_kernel, dim3(grid_for(P)), dim3(NT), 0, st, params, grads, adam_m, adam_v, pbf, P,
                           d_hyper, 0.9f, 0.999f, 1e-8f);
        BH_HIP_CHECK(hipGetLastError());
    }
}

double Cnn::train_step(int B, const double *targets, float lr) {
    if (B < 1 || B > max_batch) throw std::runtime_error("CNN batch outside 1..max_batch");
    BH_HIP_CHECK(hipMemcpyAsync(tgt, targets, sizeof(double) * B * OUT, hipMemcpyHostToDevice, st));
    const bool update = lr > 0.0f;
    if (update) {
        ++step;
        h_hyper[0] = lr;
        h_hyper[1] = 1.0f - std::pow(0.9f, (float)step);
        h_hyper[2] = 1.0f - std::pow(0.999f, (float)step);
        BH_HIP_CHECK(hipMemcpyAsync(d_hyper, h_hyper, sizeof(float) * 3, hipMemcpyHostToDevice, st));
    }
    if (use_graphs) {
        const int key = 2 * B + (update ? 1 : 0);
        auto it = graphs.find(key);
        if (it == graphs.end()) {
            BH_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            hipGraph_t g = nullptr;
            try {
                step_launches(B, update);
            } catch (...) {
                (void)hipStreamEndCapture(st, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            BH_HIP_CHECK(hipStreamEndCapture(st, &g));
            hipGraphExec_t ge = nullptr;
            const hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            BH_HIP_CHECK(ie);
            it = graphs.emplace(key, ge).first;
        }
        BH_HIP_CHECK(hipGraphLaunch(it->second, st));
    } else {
        step_launches(B, update);
    }
    std::vector<double> lb(B);
    BH_HIP_CHECK(hipMemcpyAsync(lb.data(), lossb, sizeof(double) * B, hipMemcpyDeviceToHost, st));
    BH_HIP_CHECK(hipStreamSynchronize(st));
    double s = 0.0;
    for (double v : lb) s += v;
    return s / B;
}

void Cnn::get_output(int B, double *host) const {
    if (B < 1 || B > max_batch) throw std::runtime_error("CNN batch outside 1..max_batch");
    std::vector<float> y((size_t)B * OUT);
    BH_HIP_CHECK(hipMemcpyAsync(y.data(), yout, sizeof(float) * B * OUT, hipMemcpyDeviceToHost, st));
    BH_HIP_CHECK(hipStreamSynchronize(st));
    for (size_t i = 0; i < y.size(); ++i) host[i] = y[i];
}

void Cnn::get_params(float *host) const {
    BH_HIP_CHECK(hipMemcpyAsync(host, params, sizeof(float) * P, hipMemcpyDeviceToHost, st));
    BH_HIP_CHECK(hipStreamSynchronize(st));
}

// new master weights (and their BF16 copy); the Adam state restarts
void Cnn::set_params(const float *host) {
    BH_HIP_CHECK(hipMemcpyAsync(params, host, sizeof(float) * P, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(to_bf16_kernel, dim3(grid_for(P)), dim3(NT), 0, st, params, pbf, P);
    hipLaunchKernelGGL(fill_kernel, dim3(grid_for(P)), dim3(NT), 0, st, adam_m, P);
    hipLaunchKernelGGL(fill_kernel, dim3(grid_for(P)), dim3(NT), 0, st, adam_v, P);
    hipLaunchKernelGGL(fill_kernel, dim3(grid_for(P)), dim3(NT), 0, st, grads, P);
    BH_HIP_CHECK(hipGetLastError());
    BH_HIP_CHECK(hipStreamSynchronize(st));
    step = 0;
}

void Cnn::get_grads(float *host) const {
    BH_HIP_CHECK(hipMemcpyAsync(host, grads, sizeof(float) * P, hipMemcpyDeviceToHost, st));
    BH_HIP_CHECK(hipStreamSynchronize(st));
}

}  // namespace bh
