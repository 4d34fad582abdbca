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
constexpr int LOSS_NT = 64;
constexpr int LOSS_MAXNB = 16;
constexpr size_t LOSS_MAX_LDS = 160 * 1024;

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

// one workgroup per weight w[co][ci][dy][dx] (j < 9 cin) or per bias of co (j = 9 cin):
// the sum over the batch and the output positions
__global__ void __launch_bounds__(NT) conv_bwd_weight_kernel(const float *__restrict__ in,
                                                             const float *__restrict__ gpre, float *__restrict__ gw,
                                                             float *__restrict__ gb, int B, int cin, int hin, int win,
                                                             int stride) {
    const int hout = hin / stride, wout = win / stride;
    const int per = cin * 9 + 1;
    const int co = blockIdx.x / per, j = blockIdx.x - co * per;
    const bool bias = j == cin * 9;
    const int ci = bias ? 0 : j / 9, dy = bias ? 0 : (j % 9) / 3, dx = bias ? 0 : j % 3;
    const long plane = (long)hout * wout;
    const long npos = (long)B * plane;
    float acc = 0.0f;
    for (long p = threadIdx.x; p < npos; p += NT) {
        const int b = (int)(p / plane);
        const long q = p - (long)b * plane;
        const int oy = (int)(q / wout), ox = (int)(q - (long)oy * wout);
        const float g = gpre[((long)b * CH + co) * plane + q];
        if (bias) {
            acc += g;
        } else {
            const int iy = wrap(oy * stride + dy - 1, hin), ix = wrap(ox * stride + dx - 1, win);
            acc = fmaf(g, in[(((long)b * cin + ci) * hin + iy) * win + ix], acc);
        }
    }
    const float t = block_sum(acc);
    if (threadIdx.x == 0) {
        if (bias) gb[co] = t;
        else gw[((long)co * cin + ci) * 9 + dy * 3 + dx] = t;
    }
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

// dx[b][k] = sum_m dy[b][m] W[m][k]; consecutive threads take consecutive k
__global__ void dense_bwd_data_kernel(const float *__restrict__ dy, const __hip_bfloat16 *__restrict__ W,
                                      float *__restrict__ dx, int B, int K, int M) {
    const long total = (long)B * K;
    for (long e = (long)blockIdx.x * NT + threadIdx.x; e < total; e += (long)gridDim.x * NT) {
        const int b = (int)(e / K), k = (int)(e - (long)b * K);
        const float *dp = dy + (long)b * M;
        float acc = 0.0f;
        for (int m = 0; m < M; ++m) acc = fmaf(dp[m], bf(W[(long)m * K + k]), acc);
        dx[e] = acc;
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

// In-place inverse of a Hermitian positive definite n x n matrix (row-major,
// n <= LOSS_MAXNB) by Gauss-Jordan with partial pivoting, on thread 0.
__device__ void small_inverse(double2 *A, double2 *I, int n) {
    for (int i = 0; i < n * n; ++i) I[i] = make_double2((i / n) == (i % n) ? 1.0 : 0.0, 0.0);
    for (int c = 0; c < n; ++c) {
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
        const double2 d = A[c * n + c];
        const double dd = d.x * d.x + d.y * d.y;
        const double2 inv = make_double2(d.x / dd, -d.y / dd);
        for (int j = 0; j < n; ++j) {
            const double2 a = A[c * n + j], b = I[c * n + j];
            A[c * n + j] = make_double2(a.x * inv.x - a.y * inv.y, a.x * inv.y + a.y * inv.x);
            I[c * n + j] = make_double2(b.x * inv.x - b.y * inv.y, b.x * inv.y + b.y * inv.x);
        }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            const double2 f = A[r * n + c];
            for (int j = 0; j < n; ++j) {
                const double2 a = A[c * n + j], b = I[c * n + j];
                A[r * n + j].x -= f.x * a.x - f.y * a.y;
                A[r * n + j].y -= f.x * a.y + f.y * a.x;
                I[r * n + j].x -= f.x * b.x - f.y * b.y;
                I[r * n + j].y -= f.x * b.y + f.y * b.x;
            }
        }
    }
    for (int i = 0; i < n * n; ++i) A[i] = I[i];
}

// One workgroup per sample, FP64.  y: [B][nb][ng][2] float (the head output),
// u: [B][nb][ng][2] double (targets); writes loss[b] and gy = d(mean loss)/dy.
__global__ void __launch_bounds__(LOSS_NT) projector_loss_kernel(const float *__restrict__ y,
                                                                 const double *__restrict__ u,
                                                                 double *__restrict__ loss, float *__restrict__ gy,
                                                                 int B, int nb, int ng) {
    const int b = blockIdx.x;
    const int tid = threadIdx.x;
    extern __shared__ __attribute__((aligned(16))) double lsm[];
    double2 *Y = reinterpret_cast<double2 *>(lsm);  // [ng][nb]
    double2 *U = Y + (size_t)ng * nb;               // [ng][nb]
    double2 *R = U + (size_t)ng * nb;               // P_U Y, then (I - P_Y) P_U Y
    double2 *Gy = R + (size_t)ng * nb;              // Y^H Y, then its inverse
    double2 *Gu = Gy + nb * nb;                     // U^H U, then its inverse
    double2 *A = Gu + nb * nb;
    double2 *T = A + nb * nb;
    double2 *S = T + nb * nb;  // inversion workspace
    const float *yb = y + (size_t)b * nb * ng * 2;
    const double *ub = u + (size_t)b * nb * ng * 2;
    for (int e = tid; e < ng * nb; e += LOSS_NT) {
        const int g = e / nb, j = e - g * nb;
        Y[e] = make_double2(yb[2 * (j * ng + g)], yb[2 * (j * ng + g) + 1]);
        U[e] = make_double2(ub[2 * (j * ng + g)], ub[2 * (j * ng + g) + 1]);
    }
    __syncthreads();
    auto herm_prod = [&](const double2 *P, const double2 *Q, double2 *out) {  // P^H Q, nb x nb
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
        small_inverse(Gy, S, nb);
        small_inverse(Gu, S, nb);
    }
    __syncthreads();
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
        for (int k = 0; k < nb; ++k) cmac(re, im, -Y[g * nb + k].x, -Y[g * nb + k].y, T[k * nb + j].x, T[k * nb + j].y);
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
__global__ void adam_kernel(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, __hip_bfloat16 *__restrict__ pb, long n, float lr, float b1,
                            float b2, float c1, float c2, float eps) {
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

unsigned grid_for(long n) { return (unsigned)std::max<long>(1, std::min<long>((n + NT - 1) / NT, 16384)); }

size_t loss_lds_bytes(int nb, int ng) { return sizeof(double2) * (3 * (size_t)ng * nb + 5 * (size_t)nb * nb); }

}  // namespace

Cnn::Cnn(int h_, int w_, int nb_, int pw_half_, int max_batch_, unsigned long long seed, hipStream_t st_)
    : H(h_), W(w_), nb(nb_), pw_half(pw_half_), max_batch(max_batch_), st(st_) {
    if (H % 8 || W % 8 || H < 8 || W < 8 || H > 1024 || W > 1024)
        throw std::runtime_error("CNN input sides must be multiples of 8 in [8, 1024]");
    if (nb < 1 || nb > LOSS_MAXNB) throw std::runtime_error("CNN predicts 1..16 vectors");
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
    lossb = (double *)dal(sizeof(double) * Bm);
    BH_HIP_CHECK(hipFuncSetAttribute((const void *)projector_loss_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)LOSS_MAX_LDS));
    set_params(hp.data());
}

Cnn::~Cnn() {
    (void)hipStreamSynchronize(st);
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

double Cnn::loss_and_backward(int B, const double *targets) {
    if (B < 1 || B > max_batch) throw std::runtime_error("CNN batch outside 1..max_batch");
    BH_HIP_CHECK(hipMemcpyAsync(tgt, targets, sizeof(double) * B * OUT, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(projector_loss_kernel, dim3(B), dim3(LOSS_NT), loss_lds, st, yout, tgt, lossb, gy, B, nb, ng);
    const __hip_bfloat16 *wb = pbf;
    hipLaunchKernelGGL(dense_bwd_weight_kernel, dim3(grid_for((long)OUT * (F + 1))), dim3(NT), 0, st, gy,
                       act[LAYERS - 1], grads + off_h, grads + off_hb, B, F, OUT);
    hipLaunchKernelGGL(dense_bwd_data_kernel, dim3(grid_for((long)B * F)), dim3(NT), 0, st, gy, wb + off_h,
                       gact[LAYERS - 1], B, F, OUT);
    for (int l = LAYERS - 1; l >= 0; --l) {
        const int cin = l == 0 ? 2 : CH, hin = l == 0 ? H : (H >> (l - 1)), win = l == 0 ? W : (W >> (l - 1));
        const int stride = l == 0 ? 1 : 2;
        const long plane = (long)(hin / stride) * (win / stride);
        const long n = (long)B * CH * plane;
        hipLaunchKernelGGL(relu_bwd_kernel, dim3(grid_for(n)), dim3(NT), 0, st, gact[l], act[l], n);
        hipLaunchKernelGGL(channel_sum_kernel, dim3(B * CH), dim3(NT), 0, st, gact[l], gev + l * CH, plane, EMB_OUT);
        const float *in = l == 0 ? x0 : act[l - 1];
        hipLaunchKernelGGL(conv_bwd_weight_kernel, dim3(CH * (cin * 9 + 1)), dim3(NT), 0, st, in, gact[l],
                           grads + off_c[l], grads + off_cb[l], B, cin, hin, win, stride);
        if (l > 0)
            hipLaunchKernelGGL(conv_bwd_data_kernel, dim3(grid_for((long)B * cin * hin * win)), dim3(NT), 0, st,
                               gact[l], wb + off_c[l], gact[l - 1], B, cin, hin, win, stride);
    }
    hipLaunchKernelGGL(dense_bwd_weight_kernel, dim3(grid_for((long)EMB_OUT * (CH + 1))), dim3(NT), 0, st, gev, eh,
                       grads + off_e2, grads + off_e2b, B, CH, EMB_OUT);
    hipLaunchKernelGGL(dense_bwd_data_kernel, dim3(grid_for((long)B * CH)), dim3(NT), 0, st, gev, wb + off_e2, geh,
                       B, CH, EMB_OUT);
    hipLaunchKernelGGL(relu_bwd_kernel, dim3(grid_for((long)B * CH)), dim3(NT), 0, st, geh, eh, (long)B * CH);
    hipLaunchKernelGGL(dense_bwd_weight_kernel, dim3(grid_for((long)CH * (EMB_IN + 1))), dim3(NT), 0, st, geh, e0,
                       grads + off_e1, grads + off_e1b, B, EMB_IN, CH);
    BH_HIP_CHECK(hipGetLastError());
    std::vector<double> lb(B);
    BH_HIP_CHECK(hipMemcpyAsync(lb.data(), lossb, sizeof(double) * B, hipMemcpyDeviceToHost, st));
    BH_HIP_CHECK(hipStreamSynchronize(st));
    double s = 0.0;
    for (double v : lb) s += v;
    return s / B;
}

void Cnn::adam(float lr) {
    ++step;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)step), c2 = 1.0f - std::pow(b2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3(grid_for(P)), dim3(NT), 0, st, params, grads, adam_m, adam_v, pbf, P, lr, b1,
                       b2, c1, c2, 1e-8f);
    BH_HIP_CHECK(hipGetLastError());
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
