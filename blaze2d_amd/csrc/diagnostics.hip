// Operator diagnostics kernels: start vectors, per-job reductions with
// normalisation, and the per-job preconditioner means.
#include "diagnostics.h"

#include <cmath>

namespace bh {

namespace {

constexpr int DG_NT = 256;

// sum over the workgroup of NV doubles; every thread receives the sums
template <int NV>
__device__ void block_sum(double (&v)[NV], double *sh) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k)
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) sh[wv * NV + k] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double s = 0.0;
        for (int w = 0; w < DG_NT / 64; ++w) s += sh[w * NV + k];
        v[k] = s;
    }
}

__global__ void __launch_bounds__(DG_NT) diag_init_kernel(cplx<double> *x0, cplx<double> *y0, long stride, long n) {
    const long i = (long)blockIdx.x * DG_NT + threadIdx.x;
    if (i >= n) return;
    const double t = (double)i;
    cplx<double> *x = x0 + (size_t)blockIdx.y * stride;
    cplx<double> *y = y0 + (size_t)blockIdx.y * stride;
    x[i] = cplx<double>{sin(t * 0.618033988749895), cos(t * 0.414213562373095)};
    y[i] = cplx<double>{cos(t * 1.4142135623730951), sin(t * 1.7320508075688772)};
}

__global__ void __launch_bounds__(DG_NT) diag_reduce_kernel(const DiagItem *__restrict__ items, int mode, long n) {
    const DiagItem &it = items[blockIdx.x];
    __shared__ double sh[(DG_NT / 64) * 6];
    if (mode == 2) {
        double v[1] = {0.0};
        for (long i = threadIdx.x; i < n; i += DG_NT) v[0] += it.v[i].x * it.v[i].x + it.v[i].y * it.v[i].y;
        block_sum<1>(v, sh);
        const double nrm = sqrt(v[0]);
        if (nrm > 1e-15)
            for (long i = threadIdx.x; i < n; i += DG_NT) it.v[i] = cplx<double>{it.v[i].x / nrm, it.v[i].y / nrm};
    } else if (mode == 0) {
        double v[2] = {0.0, 0.0};  // Re(v^H av), |av|^2
        for (long i = threadIdx.x; i < n; i += DG_NT) {
            const cplx<double> a = it.v[i], b = it.av[i];
            v[0] += a.x * b.x + a.y * b.y;
            v[1] += b.x * b.x + b.y * b.y;
        }
        block_sum<2>(v, sh);
        if (threadIdx.x == 0) *it.out = v[0];
        const double nrm = sqrt(v[1]);
        if (nrm > 1e-15)
            for (long i = threadIdx.x; i < n; i += DG_NT) it.av[i] = cplx<double>{it.av[i].x / nrm, it.av[i].y / nrm};
    } else {
        double v[6] = {0, 0, 0, 0, 0, 0};  // <Ax, y>, <x, Ay> (re, im), |Ax|^2, |y|^2
        for (long i = threadIdx.x; i < n; i += DG_NT) {
            const cplx<double> ax = it.av[i], y = it.y[i], x = it.v[i], ay = it.ay[i];
            v[0] += ax.x * y.x + ax.y * y.y;
            v[1] += ax.x * y.y - ax.y * y.x;
            v[2] += x.x * ay.x + x.y * ay.y;
            v[3] += x.x * ay.y - x.y * ay.x;
            v[4] += ax.x * ax.x + ax.y * ax.y;
            v[5] += y.x * y.x + y.y * y.y;
        }
        block_sum<6>(v, sh);
        if (threadIdx.x == 0) {
            const double dr = v[0] - v[2], di = v[1] - v[3];
            const double scale = sqrt(v[4]) * sqrt(v[5]);
            *it.out = scale > 1e-15 ? sqrt(dr * dr + di * di) / scale : 0.0;
        }
    }
}

__global__ void __launch_bounds__(DG_NT) diag_params_kernel(OpParams *params, const double *eps, long n) {
    OpParams &P = params[blockIdx.x];
    const double *e = eps + (size_t)blockIdx.x * n;
    __shared__ double sh[(DG_NT / 64) * 2];
    double v[2] = {0.0, 0.0};  // sum eps, sum 1 / eps
    for (long i = threadIdx.x; i < n; i += DG_NT) {
        v[0] += e[i];
        v[1] += 1.0 / e[i];
    }
    block_sum<2>(v, sh);
    if (threadIdx.x != 0) return;
    if (P.pol == 0) {
        P.eps_eff = v[0] / (double)n;
        P.mass_floor = (isfinite(P.eps_eff) && P.eps_eff > 0) ? P.eps_eff * 1e-2 : 0.0;
    } else {
        P.eps_eff = v[1] > 0 ? 1.0 / (v[1] / (double)n) : 1.0;
        P.mass_floor = 0.0;
    }
}

}  // namespace

void launch_diag_init(cplx<double> *x0, cplx<double> *y0, long stride, int njobs, long n, hipStream_t st) {
    if (njobs <= 0) return;
    hipLaunchKernelGGL(diag_init_kernel, dim3((unsigned)((n + DG_NT - 1) / DG_NT), (unsigned)njobs), dim3(DG_NT), 0,
                       st, x0, y0, stride, n);
    BH_HIP_CHECK(hipGetLastError());
}

void launch_diag_reduce(const DiagItem *d_items, int nitems, int mode, long n, hipStream_t st) {
    if (nitems <= 0) return;
    hipLaunchKernelGGL(diag_reduce_kernel, dim3((unsigned)nitems), dim3(DG_NT), 0, st, d_items, mode, n);
    BH_HIP_CHECK(hipGetLastError());
}

void launch_diag_params(OpParams *d_params, const double *d_eps, int njobs, long n, hipStream_t st) {
    if (njobs <= 0) return;
    hipLaunchKernelGGL(diag_params_kernel, dim3((unsigned)njobs), dim3(DG_NT), 0, st, d_params, d_eps, n);
    BH_HIP_CHECK(hipGetLastError());
}

}  // namespace bh
