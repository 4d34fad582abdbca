// Elementwise and reduction helpers (residuals, Rayleigh quotients, start
// vectors, trait-level scale/axpy/dot).
#include "aux_kernels.h"

namespace bh {

// One workgroup per (column, work item), so that the column's B-norm is
// reduced inside the workgroup (no atomics; independent of launch order).
template <typename T>
__global__ void __launch_bounds__(256) residual_kernel(const ResidualWork<T> *__restrict__ work, long rows) {
    const ResidualWork<T> &w = work[blockIdx.y];
    const int c = blockIdx.x;
    if (c >= w.count) return;
    const int b = w.act[c];
    const double lam = w.lam[b];
    // backend axpy with alpha = (-lambda, 0) in storage precision
    const T al = (T)(-lam);
    const cplx<T> *ax = w.ax + (size_t)b * w.stride, *bx = w.bx + (size_t)b * w.stride;
    cplx<T> *out = w.out + (size_t)c * w.stride;
    double acc = 0.0;
#pragma unroll 4
    for (long r = threadIdx.x; r < rows; r += 256) {
        const cplx<T> a = gload(ax, r);
        cplx<T> m = gload(bx, r);
        const T s = w.bw ? gmem(w.bw)[r] : (T)1;
        if (w.bw) m = {m.x * s, m.y * s};
        const cplx<T> o = {a.x + al * m.x, a.y + al * m.y};
        gstore(out, r, o);
        // Re(conj(r) (B r)), B r in storage precision as a mass apply forms it
        acc += (double)o.x * (double)(o.x * s) + (double)o.y * (double)(o.y * s);
    }
    if (!w.rel) return;
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    __shared__ double part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0)
        w.rel[b] = sqrt(fmax(part[0] + part[1] + part[2] + part[3], 0.0)) / fmax(fabs(lam), 1e-10);
}

template <typename T>
void launch_residual(const ResidualWork<T> *dwork, int nwork, int maxcount, long rows, hipStream_t st) {
    if (nwork <= 0 || maxcount <= 0) return;
    hipLaunchKernelGGL((residual_kernel<T>), dim3(maxcount, nwork), dim3(256), 0, st, dwork, rows);
    BH_HIP_CHECK(hipGetLastError());
}

// grid = (columns, items, 2h + 1 values of n2); each thread accumulates the
// 2h + 1 coefficients of its n2 over a stride of rows, then a workgroup reduction
constexpr int PW_MAXH = 8;
template <typename T>
__global__ void __launch_bounds__(256) pw_project_kernel(const PwWork<T> *__restrict__ work, int nx, int ny,
                                                         int h) {
    const PwWork<T> &w = work[blockIdx.y];
    const int c = blockIdx.x;
    if (c >= w.count) return;
    const int n2 = (int)blockIdx.z - h, nw = 2 * h + 1;
    const cplx<T> *x = w.x + (size_t)c * w.stride;
    double ar[2 * PW_MAXH + 1], ai[2 * PW_MAXH + 1];
    for (int j = 0; j < nw; ++j) ar[j] = ai[j] = 0.0;
    const long n = (long)nx * ny;
    for (long r = threadIdx.x; r < n; r += 256) {
        const int iy = (int)(r / nx), ix = (int)(r - (long)iy * nx);
        const cplx<T> v = gload(x, r);
        // v exp(-2 pi i n2 iy / ny), then times exp(-2 pi i n1 ix / nx) for n1 = -h..h
        double sy, cy, s1, c1, s0, c0;
        sincospi(-2.0 * n2 * iy / ny, &sy, &cy);
        sincospi(-2.0 * ix / nx, &s1, &c1);       // step exp(-2 pi i ix / nx)
        sincospi(2.0 * h * ix / nx, &s0, &c0);    // start exp(+2 pi i h ix / nx) (n1 = -h)
        double br = (double)v.x * cy - (double)v.y * sy, bi = (double)v.x * sy + (double)v.y * cy;
        double pr = c0, pi = s0;
        for (int j = 0; j < nw; ++j) {
            cmac(ar[j], ai[j], br, bi, pr, pi);
            const double t = pr * c1 - pi * s1;
            pi = pr * s1 + pi * c1;
            pr = t;
        }
    }
    __shared__ double red[4][2 * (2 * PW_MAXH + 1)];
    for (int j = 0; j < nw; ++j) {
        double vr = ar[j], vi = ai[j];
        for (int o = 32; o > 0; o >>= 1) {
            vr += __shfl_xor(vr, o);
            vi += __shfl_xor(vi, o);
        }
        if ((threadIdx.x & 63) == 0) {
            red[threadIdx.x >> 6][2 * j] = vr;
            red[threadIdx.x >> 6][2 * j + 1] = vi;
        }
    }
    __syncthreads();
    if ((int)threadIdx.x < 2 * nw) {
        const double t = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
        const int j = threadIdx.x >> 1;
        w.c[2 * ((size_t)c * nw * nw + (size_t)(n2 + h) * nw + j) + (threadIdx.x & 1)] = t / (double)n;
    }
}

template <typename T>
__global__ void __launch_bounds__(256) pw_synth_kernel(const PwWork<T> *__restrict__ work, int nx, int ny, int h) {
    const PwWork<T> &w = work[blockIdx.z];
    const int c = blockIdx.y;
    if (c >= w.count) return;
    const long n = (long)nx * ny;
    const long r = (long)blockIdx.x * 256 + threadIdx.x;
    if (r >= n) return;
    const int iy = (int)(r / nx), ix = (int)(r - (long)iy * nx);
    const int nw = 2 * h + 1;
    const double *cf = w.cin + 2 * (size_t)c * nw * nw;
    double sr = 0.0, si = 0.0;
    double s1, c1, s2, c2;
    sincospi(2.0 * ix / nx, &s1, &c1);
    for (int b = 0; b < nw; ++b) {
        const int n2 = b - h;
        double sy, cy, s0, c0;
        sincospi(2.0 * n2 * iy / ny, &sy, &cy);
        sincospi(-2.0 * h * ix / nx, &s0, &c0);
        // e = exp(+2 pi i (n1 ix / nx + n2 iy / ny)), n1 from -h
        double er = c0 * cy - s0 * sy, ei = c0 * sy + s0 * cy;
        for (int a = 0; a < nw; ++a) {
            const double gr = cf[2 * (b * nw + a)], gi = cf[2 * (b * nw + a) + 1];
            cmac(sr, si, gr, gi, er, ei);
            const double t = er * c1 - ei * s1;
            ei = er * s1 + ei * c1;
            er = t;
        }
    }
    (void)s2;
    (void)c2;
    gstore(w.out + (size_t)c * w.stride, r, cplx<T>{(T)sr, (T)si});
}

template <typename T>
void launch_pw_project(const PwWork<T> *dwork, int nwork, int maxcount, int nx, int ny, int h, hipStream_t st) {
    if (nwork <= 0 || maxcount <= 0) return;
    if (h < 0 || h > PW_MAXH) throw std::runtime_error("plane-wave half width must be 0..8");
    hipLaunchKernelGGL((pw_project_kernel<T>), dim3(maxcount, nwork, 2 * h + 1), dim3(256), 0, st, dwork, nx, ny, h);
    BH_HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_pw_synth(const PwWork<T> *dwork, int nwork, int maxcount, int nx, int ny, int h, hipStream_t st) {
    if (nwork <= 0 || maxcount <= 0) return;
    if (h < 0 || h > PW_MAXH) throw std::runtime_error("plane-wave half width must be 0..8");
    const long n = (long)nx * ny;
    hipLaunchKernelGGL((pw_synth_kernel<T>), dim3((unsigned)((n + 255) / 256), maxcount, nwork), dim3(256), 0, st,
                       dwork, nx, ny, h);
    BH_HIP_CHECK(hipGetLastError());
}

// one workgroup per (column, work item), the column's sum reduced in the workgroup
template <typename T>
__global__ void __launch_bounds__(256) bnorm_kernel(const NormWork<T> *__restrict__ work, long rows) {
    const NormWork<T> &w = work[blockIdx.y];
    const int c = blockIdx.x;
    if (c >= w.count) return;
    const cplx<T> *r = w.r + (size_t)c * w.stride;
    double acc = 0.0;
#pragma unroll 4
    for (long i = threadIdx.x; i < rows; i += 256) {
        const cplx<T> v = gload(r, i);
        const T s = w.bw ? gmem(w.bw)[i] : (T)1;
        acc += (double)v.x * (double)(v.x * s) + (double)v.y * (double)(v.y * s);
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    __shared__ double part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int b = w.act[c];
        const double den = w.lam ? fmax(fabs(w.lam[b]), 1e-10) : 1.0;
        w.rel[b] = sqrt(fmax(part[0] + part[1] + part[2] + part[3], 0.0)) / den;
    }
}

template <typename T>
void launch_bnorm(const NormWork<T> *dwork, int nwork, int maxcount, long rows, hipStream_t st) {
    if (nwork <= 0 || maxcount <= 0) return;
    hipLaunchKernelGGL((bnorm_kernel<T>), dim3(maxcount, nwork), dim3(256), 0, st, dwork, rows);
    BH_HIP_CHECK(hipGetLastError());
}

template <typename T>
__global__ void __launch_bounds__(256) rayleigh_kernel(const RayleighWork<T> *__restrict__ work, long rows) {
    const RayleighWork<T> &w = work[blockIdx.y];
    const int c = blockIdx.x;
    if (c >= w.count) return;
    const cplx<T> *x = w.x + c * w.stride, *ax = w.ax + c * w.stride, *bx = w.bx + c * w.stride;
    double num = 0, den = 0;
    for (long r = threadIdx.x; r < rows; r += 256) {
        cplx<T> xv = x[r], av = ax[r], bv = bx[r];
        if (w.bw) {
            const T s = w.bw[r];
            bv = {bv.x * s, bv.y * s};
        }
        num += (double)xv.x * av.x + (double)xv.y * av.y;
        den += (double)xv.x * bv.x + (double)xv.y * bv.y;
    }
    __shared__ double sn[256], sd[256];
    sn[threadIdx.x] = num;
    sd[threadIdx.x] = den;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            sn[threadIdx.x] += sn[threadIdx.x + s];
            sd[threadIdx.x] += sd[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) w.lam[c] = fabs(sd[0]) > 1e-15 ? sn[0] / sd[0] : 0.0;
}

template <typename T>
void launch_rayleigh(const RayleighWork<T> *dwork, int nwork, int maxcount, long rows, hipStream_t st) {
    if (nwork <= 0 || maxcount <= 0) return;
    hipLaunchKernelGGL((rayleigh_kernel<T>), dim3(maxcount, nwork), dim3(256), 0, st, dwork, rows);
    BH_HIP_CHECK(hipGetLastError());
}

// state = bits(phase) * 0x9E3779B97F4A7C15 ; xorshift 13/7/17 ; (s >> 12) / 2^52 * 2 - 1
template <typename T>
__global__ void random_kernel(cplx<T> *out, int ncols, long rows, double first_phase) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncols) return;
    double phase = first_phase + (double)c;
    unsigned long long s = (unsigned long long)__double_as_longlong(phase) * 0x9E3779B97F4A7C15ULL;
    if (s == 0) s = 0xDEADBEEFCAFEBABEULL;
    cplx<T> *o = out + (long)c * rows;
    for (long r = 0; r < rows; ++r) {
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        double re = (double)(s >> 12) / 4503599627370496.0 * 2.0 - 1.0;
        o[r] = {(T)re, (T)0};
    }
}

template <typename T>
void launch_random(cplx<T> *out, int ncols, long rows, double first_phase, hipStream_t st) {
    if (ncols <= 0) return;
    hipLaunchKernelGGL((random_kernel<T>), dim3((ncols + 63) / 64), dim3(64), 0, st, out, ncols, rows, first_phase);
    BH_HIP_CHECK(hipGetLastError());
}

template <typename T>
__global__ void fill_kernel(cplx<T> *out, cplx<T> *bout, const T *w, double re, double im, long rows) {
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    cplx<T> v = {(T)re, (T)im};
    out[r] = v;
    if (bout) bout[r] = {v.x * w[r], v.y * w[r]};
}

template <typename T>
void launch_fill(cplx<T> *out, cplx<T> *bout, const T *w, double re, double im, long rows, hipStream_t st) {
    hipLaunchKernelGGL((fill_kernel<T>), dim3((rows + 255) / 256), dim3(256), 0, st, out, bout, w, re, im, rows);
    BH_HIP_CHECK(hipGetLastError());
}

template <typename T>
__global__ void scale_kernel(cplx<T> *x, long n, double re, double im) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    cplx<T> a = {(T)re, (T)im};
    x[i] = cmul(x[i], a);
}
template <typename T>
__global__ void axpy_kernel(const cplx<T> *x, cplx<T> *y, long n, double re, double im) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    cplx<T> a = {(T)re, (T)im};
    y[i] = cadd(y[i], cmul(a, x[i]));
}
template <typename T>
__global__ void __launch_bounds__(256) dot_kernel(const cplx<T> *x, const cplx<T> *y, long n, double *out) {
    double re = 0, im = 0;
    for (long i = threadIdx.x; i < n; i += 256) {
        double a = x[i].x, b = x[i].y, c = y[i].x, d = y[i].y;
        re += a * c + b * d;
        im += a * d - b * c;
    }
    __shared__ double sr[256], si[256];
    sr[threadIdx.x] = re;
    si[threadIdx.x] = im;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            sr[threadIdx.x] += sr[threadIdx.x + s];
            si[threadIdx.x] += si[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { out[0] = sr[0]; out[1] = si[0]; }
}

template <typename T>
void launch_scale(cplx<T> *x, long n, double re, double im, hipStream_t st) {
    hipLaunchKernelGGL((scale_kernel<T>), dim3((n + 255) / 256), dim3(256), 0, st, x, n, re, im);
    BH_HIP_CHECK(hipGetLastError());
}
template <typename T>
void launch_axpy(const cplx<T> *x, cplx<T> *y, long n, double re, double im, hipStream_t st) {
    hipLaunchKernelGGL((axpy_kernel<T>), dim3((n + 255) / 256), dim3(256), 0, st, x, y, n, re, im);
    BH_HIP_CHECK(hipGetLastError());
}
template <typename T>
void launch_dot(const cplx<T> *x, const cplx<T> *y, long n, double *out2, hipStream_t st) {
    hipLaunchKernelGGL((dot_kernel<T>), dim3(1), dim3(256), 0, st, x, y, n, out2);
    BH_HIP_CHECK(hipGetLastError());
}

#define BH_INST(T)                                                                                         \
    template void launch_residual<T>(const ResidualWork<T> *, int, int, long, hipStream_t);                 \
    template void launch_rayleigh<T>(const RayleighWork<T> *, int, int, long, hipStream_t);                 \
    template void launch_bnorm<T>(const NormWork<T> *, int, int, long, hipStream_t);                        \
    template void launch_pw_project<T>(const PwWork<T> *, int, int, int, int, int, hipStream_t);             \
    template void launch_pw_synth<T>(const PwWork<T> *, int, int, int, int, int, hipStream_t);               \
    template void launch_random<T>(cplx<T> *, int, long, double, hipStream_t);                              \
    template void launch_fill<T>(cplx<T> *, cplx<T> *, const T *, double, double, long, hipStream_t);       \
    template void launch_scale<T>(cplx<T> *, long, double, double, hipStream_t);                            \
    template void launch_axpy<T>(const cplx<T> *, cplx<T> *, long, double, double, hipStream_t);            \
    template void launch_dot<T>(const cplx<T> *, const cplx<T> *, long, double *, hipStream_t);
BH_INST(double)
BH_INST(float)
#undef BH_INST

}  // namespace bh
