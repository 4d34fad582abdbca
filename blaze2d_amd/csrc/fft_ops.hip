// Fused LDS-resident 2D FFT operators for gfx950.
//
// One workgroup owns one NX x NY field (one column of one slot) for the whole
// operator: the field is read from HBM once, every FFT pass runs in LDS, and
// the result is written once.  The forward transform is an in-place mixed
// radix decimation-in-frequency (natural order in, digit-reversed out); the
// inverse is the matching decimation-in-time (digit-reversed in, natural
// out), so no reorder pass is ever needed: Fourier-space factors are
// evaluated at the digit-reversed frequency of each LDS position, with k+G
// generated from (b1, b2, bloch) on the fly instead of streamed tables.
//
// Operators (reference citations):
//   THETA_TM   y = F^-1 |k+G|^2 F x                         maxwell.rs:686-713
//   THETA_TE   y = -F^-1 i(k+G).F eta F^-1 i(k+G) F x       maxwell.rs:647-684,1745-1839
//   PREC_FD    y = F^-1 eps_eff/(|k+G|^2+m+s eps_eff) F x   fourier_diagonal.rs:310-342
//   PREC_TP    transverse projection, 6 FFTs            transverse_projection.rs:167-231
// FFT convention (backend-cpu/src/lib.rs:94-177): forward unnormalised
// e^{-2 pi i jk/n}, inverse e^{+...} scaled by 1/(nx ny).
#include "common.h"
#include "fft_core.h"
#include "fft_ops.h"

#include <type_traits>

namespace bh {

// Threads per workgroup of the 64 x 64 operator kernels, for one field (TM,
// FD, plain FFT) and two fields (TE, TP) per workgroup.  Measured on MI355X
// (FP64, 64 x 64): one field runs 20-25% faster with 512 threads (two or
// more workgroups per CU overlap their load and FFT phases); two fields need
// 1024, since at 512 threads the 128-VGPR budget spills and they run 12%
// slower.
#ifndef OP_NT_64_F1
#define OP_NT_64_F1 512
#endif
#ifndef OP_NT_64_F2
#define OP_NT_64_F2 1024
#endif

// ---------------------------------------------------------------------------
// The fused operator kernel.  grid = (max columns, work items).  TS is the
// storage precision of the columns and tables, T the precision of the FFTs
// and pointwise steps (T = float with TS = double: a preconditioner applied in
// FP32 to FP64 columns).
// ---------------------------------------------------------------------------
template <typename TS, int NX, int NY, int MODE, int NT, typename T = TS>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) op_kernel(const OpWork<TS> *__restrict__ work, const cplx<TS> *__restrict__ gtwx,
                                                const cplx<TS> *__restrict__ gtwy) {
    constexpr int N = NX * NY;
    constexpr int NF = (MODE == MODE_THETA_TM || MODE == MODE_PREC_FD) ? 1 : 2;
    extern __shared__ __attribute__((aligned(16))) char smem_raw[];
    cplx<T> *F = reinterpret_cast<cplx<T> *>(smem_raw);
    const OpWork<TS> &w = work[blockIdx.y];
    const int count = w.dev_count ? *w.dev_count : w.count;
    const int col = blockIdx.x;
    if (col >= count) return;
    const OpParams P = *w.params;
    const cplx<TS> *in = w.in + (size_t)col * w.in_stride;
    cplx<TS> *out = w.out + (size_t)col * w.out_stride;
    const T invN = (T)(1.0 / (double)N);
    // twiddle tables staged behind the fields
    cplx<T> *tlx = F + NF * N, *tly = tlx + NX;
    for (int i = threadIdx.x; i < NX; i += NT) tlx[i] = {(T)gtwx[i].x, (T)gtwx[i].y};
    for (int i = threadIdx.x; i < NY; i += NT) tly[i] = {(T)gtwy[i].x, (T)gtwy[i].y};
    const cplx<T> *twx = tlx, *twy = tly;
    __syncthreads();

    if constexpr (MODE == MODE_THETA_TM || MODE == MODE_PREC_FD) {
        fft2d_forward_from<T, NX, NY, NT, TS>(F, in, twx, twy);
#pragma unroll 2
        for (int i = threadIdx.x; i < N; i += NT) {
            int py = i / NX, px = i - py * NX;
            double kx, ky, ksq;
            bool cl;
            kplusg<NX, NY>(P, px, py, kx, ky, ksq, cl);
            double fac;
            if (MODE == MODE_THETA_TM) {
                fac = ksq;
            } else {  // inverse_scale (maxwell.rs:1692-1711) with Gamma DC mask
                double e = P.eps_eff;
                if (!(e > 0.0) || !isfinite(e) || !isfinite(ksq)) fac = 0.0;
                else {
                    double mf = (isfinite(P.mass_floor) && P.mass_floor > 0.0) ? P.mass_floor : 0.0;
                    fac = e / (fmax(ksq, 0.0) + mf + P.shift * fmax(e, 1e-12));
                }
                if (P.gamma && cl) fac = 0.0;
            }
            F[lds_pos<NX>(i)] = scale_c(F[lds_pos<NX>(i)], (T)fac * invN);
        }
        __syncthreads();
        PassIO io;
        io.gout = out;
        if (MODE == MODE_PREC_FD && w.bout) {
            io.geps = w.eps;
            io.gbout = w.bout + (size_t)col * w.out_stride;
        }
        fft2d_inverse_to<T, NX, NY, NT, TS>(F, io, twx, twy);
    } else {
        cplx<T> *G2 = F + N;
        constexpr int KX = Plan<NX>::K, KY = Plan<NY>::K;
        auto at = [&](int f, int i) -> cplx<T> & { return F[f * N + lds_pos<NX>(i)]; };
        // forward transform of the field; its last column pass applies the
        // gradient step on the way out: THETA_TE g = i(k+G) u_hat, PREC_TP
        // g = -i(k+G) u_hat / (|k+G|^2 + s)
        {
            PassIO io;
            io.gin = in;
            fft_pass<T, NX, NY, true, 0, false, 1, NT, N, NX, true, false, TS>(F, twx, io);
            __syncthreads();
            fft_passes_fwd<T, NX, NY, true, false, 1, NT, N, NX, 1>(F, twx);
            fft_passes_fwd<T, NY, NX, false, false, 1, NT, N, NX, 0, KY - 1>(F, twy);
            auto grad = [&](int, int i, cplx<T> u) {
                const int py = i / NX, px = i - py * NX;
                double kx, ky, ksq;
                bool cl;
                kplusg<NX, NY>(P, px, py, kx, ky, ksq, cl);
                if (MODE == MODE_THETA_TE) {
                    const T skx = (T)kx * invN, sky = (T)ky * invN;
                    at(0, i) = {-u.y * skx, u.x * skx};  // u * (0 + i kx)
                    at(1, i) = {-u.y * sky, u.x * sky};
                } else {
                    double ik = 0.0;
                    if (!cl) { double den = ksq + P.shift; ik = den > 1e-15 ? 1.0 / den : 0.0; }
                    const T fx = (T)(kx * ik) * invN, fy = (T)(ky * ik) * invN;
                    at(0, i) = {u.y * fx, -u.x * fx};  // u * (0 - i kx ik)
                    at(1, i) = {u.y * fy, -u.x * fy};
                }
            };
            fft_pass_fn<T, NY, NX, false, KY - 1, false, 1, NT>(twy, at, grad);
            __syncthreads();
        }
        // inverse transform of both gradient components, then the real-space
        // step: THETA_TE eta tensor [xx, xy, yx, yy] (mixes the two fields, so
        // it runs as its own step); PREC_TP eps, applied by the last row pass
        fft_passes_inv<T, NY, NX, false, true, 2, NT, N, NX, KY - 1>(F, twy);
        if constexpr (MODE == MODE_THETA_TE) {
            fft_passes_inv<T, NX, NY, true, true, 2, NT, N, NX, KX - 1>(F, twx);
#pragma unroll 2
            for (int i = threadIdx.x; i < N; i += NT) {
                cplx<T> gx = F[lds_pos<NX>(i)], gy = G2[lds_pos<NX>(i)];
                const auto *t = gmem(w.tensor + 4 * (size_t)i);
                T t0 = (T)t[0], t1 = (T)t[1], t2 = (T)t[2], t3 = (T)t[3];
                F[lds_pos<NX>(i)] = {gx.x * t0 + gy.x * t1, gx.y * t0 + gy.y * t1};
                G2[lds_pos<NX>(i)] = {gx.x * t2 + gy.x * t3, gx.y * t2 + gy.y * t3};
            }
        } else {
            fft_passes_inv<T, NX, NY, true, true, 2, NT, N, NX, KX - 1, 1>(F, twx);
            auto scale_eps = [&](int f, int i, cplx<T> g) {
                const T e = (T)gmem(w.eps)[i];
                at(f, i) = {g.x * e, g.y * e};
            };
            fft_pass_fn<T, NX, NY, true, 0, true, 2, NT>(twx, at, scale_eps);
        }
        __syncthreads();
        fft2d_forward<T, NX, NY, 2, NT>(F, twx, twy);
        // inverse transform of the divergence; its first column pass forms
        // out_hat = -(i kx gx + i ky gy) [ * 1/(|k+G|^2+s) for TP ] as it loads
        // (it writes only field 0 at the positions it reads)
        {
            auto div = [&](int, int i) -> cplx<T> {
                const int py = i / NX, px = i - py * NX;
                double kx, ky, ksq;
                bool cl;
                kplusg<NX, NY>(P, px, py, kx, ky, ksq, cl);
                double sc = 1.0;
                if (MODE == MODE_PREC_TP) {
                    double ik = 0.0;
                    if (!cl) { double den = ksq + P.shift; ik = den > 1e-15 ? 1.0 / den : 0.0; }
                    sc = ik;
                }
                const T skx = (T)kx, sky = (T)ky, s = (T)sc * invN;
                const cplx<T> gx = at(0, i), gy = at(1, i);
                // div = i kx gx + i ky gy ; out = -div
                const T dre = -(skx * gx.y + sky * gy.y);
                const T dim = skx * gx.x + sky * gy.x;
                return cplx<T>{-dre * s, -dim * s};
            };
            auto put = [&](int, int i, cplx<T> v) { at(0, i) = v; };
            if constexpr (MODE == MODE_THETA_TE) {
                fft_pass_fn<T, NY, NX, false, KY - 1, true, 1, NT>(twy, div, put);
            } else {
                // TP: the division would spill registers inside the pass; the
                // divergence runs as its own step before the first pass
#pragma unroll 2
                for (int i = threadIdx.x; i < N; i += NT) at(0, i) = div(0, i);
                __syncthreads();
                fft_pass<T, NY, NX, false, KY - 1, true, 1, NT, N, NX>(F, twy);
            }
            __syncthreads();
            fft_passes_inv<T, NY, NX, false, true, 1, NT, N, NX, KY - 2>(F, twy);
            fft_passes_inv<T, NX, NY, true, true, 1, NT, N, NX, KX - 1, 1>(F, twx);
            PassIO io;
            io.gout = out;
            fft_pass<T, NX, NY, true, 0, true, 1, NT, N, NX, false, true, TS>(F, twx, io);
        }
    }
}

// Plain 2D FFT of columns (trait-level SpectralBackend::forward/inverse_fft_2d).
template <typename T, int NX, int NY, int NT>
__global__ void __launch_bounds__(NT) fft_kernel(const OpWork<T> *__restrict__ work, const cplx<T> *__restrict__ twx,
                                                 const cplx<T> *__restrict__ twy, int inverse) {
    constexpr int N = NX * NY;
    extern __shared__ __attribute__((aligned(16))) char smem_raw[];
    cplx<T> *F = reinterpret_cast<cplx<T> *>(smem_raw);
    const OpWork<T> &w = work[blockIdx.y];
    const int count = w.dev_count ? *w.dev_count : w.count;
    const int col = blockIdx.x;
    if (col >= count) return;
    const cplx<T> *in = w.in + (size_t)col * w.in_stride;
    cplx<T> *out = w.out + (size_t)col * w.out_stride;
    cplx<T> *tlx = F + N, *tly = tlx + NX;  // twiddle tables staged behind the field
    for (int i = threadIdx.x; i < NX; i += NT) tlx[i] = twx[i];
    for (int i = threadIdx.x; i < NY; i += NT) tly[i] = twy[i];
    twx = tlx;
    twy = tly;
    if (!inverse) {
        for (int i = threadIdx.x; i < N; i += NT) F[lds_pos<NX>(i)] = in[i];
        __syncthreads();
        fft2d_forward<T, NX, NY, 1, NT>(F, twx, twy);
        // digit-reversed -> natural on the way out
#pragma unroll 2
        for (int i = threadIdx.x; i < N; i += NT) {
            int py = i / NX, px = i - py * NX;
            out[rev_index<NY>(py) * NX + rev_index<NX>(px)] = F[lds_pos<NX>(i)];
        }
    } else {
        const T invN = (T)(1.0 / (double)N);
#pragma unroll 2
        for (int i = threadIdx.x; i < N; i += NT) {
            int py = i / NX, px = i - py * NX;
            F[lds_pos<NX>(i)] = scale_c(in[rev_index<NY>(py) * NX + rev_index<NX>(px)], invN);
        }
        __syncthreads();
        fft2d_inverse<T, NX, NY, 1, NT>(F, twx, twy);
        for (int i = threadIdx.x; i < N; i += NT) out[i] = F[lds_pos<NX>(i)];
    }
}

// ---------------------------------------------------------------------------
// Host-side dispatch
// ---------------------------------------------------------------------------
// Threads per workgroup for an N-point field and NF fields per workgroup.
constexpr int op_threads(int N, int NF) {
    if (N >= 4096) return NF == 1 ? OP_NT_64_F1 : OP_NT_64_F2;
    return N >= 1024 ? 256 : 128;
}

// FP32 preconditioners on FP64 columns: FP32 fields need half the LDS, so the
// two-field TP kernel fits two 512-thread workgroups per CU
template <int NX, int NY, int MODE>
static void launch_prec_fp32(const OpWork<double> *dwork, int nwork, int maxcols, const cplx<double> *twx,
                             const cplx<double> *twy, hipStream_t st) {
    constexpr int N = NX * NY;
    constexpr int NF = MODE == MODE_PREC_FD ? 1 : 2;
    constexpr int NT = N >= 4096 ? 512 : (N >= 1024 ? 256 : 128);
    const size_t sh = sizeof(cplx<float>) * (NF * N + NX + NY);
    static const bool attr_set = [&] {
        BH_HIP_CHECK(hipFuncSetAttribute((const void *)op_kernel<double, NX, NY, MODE, NT, float>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
        return true;
    }();
    (void)attr_set;
    hipLaunchKernelGGL((op_kernel<double, NX, NY, MODE, NT, float>), dim3(maxcols, nwork), dim3(NT), sh, st, dwork,
                       twx, twy);
    BH_HIP_CHECK(hipGetLastError());
}

template <typename T, int NX, int NY>
static void launch_sized(int mode, const OpWork<T> *dwork, int nwork, int maxcols, const cplx<T> *twx,
                         const cplx<T> *twy, hipStream_t st, bool prec_fp32) {
    if constexpr (std::is_same<T, double>::value) {
        if (prec_fp32 && mode == MODE_PREC_FD) {
            launch_prec_fp32<NX, NY, MODE_PREC_FD>(dwork, nwork, maxcols, twx, twy, st);
            return;
        }
        if (prec_fp32 && mode == MODE_PREC_TP) {
            launch_prec_fp32<NX, NY, MODE_PREC_TP>(dwork, nwork, maxcols, twx, twy, st);
            return;
        }
    }
    constexpr int N = NX * NY;
    constexpr int NT1 = op_threads(N, 1), NT = op_threads(N, 2);
    dim3 grid(maxcols, nwork);
    // one or two fields, then the two twiddle tables
    const size_t sh1 = sizeof(cplx<T>) * (N + NX + NY), sh2 = sizeof(cplx<T>) * (2 * N + NX + NY);
    // kernels using more than the default 64 KB of dynamic LDS must opt in
    static const bool attrs_set = [&] {
        BH_HIP_CHECK(hipFuncSetAttribute((const void *)op_kernel<T, NX, NY, MODE_THETA_TE, NT>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh2));
        BH_HIP_CHECK(hipFuncSetAttribute((const void *)op_kernel<T, NX, NY, MODE_PREC_TP, NT>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh2));
        BH_HIP_CHECK(hipFuncSetAttribute((const void *)op_kernel<T, NX, NY, MODE_THETA_TM, NT1>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh1));
        BH_HIP_CHECK(hipFuncSetAttribute((const void *)op_kernel<T, NX, NY, MODE_PREC_FD, NT1>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh1));
        BH_HIP_CHECK(hipFuncSetAttribute((const void *)fft_kernel<T, NX, NY, NT1>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh1));
        return true;
    }();
    (void)attrs_set;
    switch (mode) {
        case MODE_THETA_TM:
            hipLaunchKernelGGL((op_kernel<T, NX, NY, MODE_THETA_TM, NT1>), grid, dim3(NT1), sh1, st, dwork, twx, twy);
            break;
        case MODE_THETA_TE:
            hipLaunchKernelGGL((op_kernel<T, NX, NY, MODE_THETA_TE, NT>), grid, dim3(NT), sh2, st, dwork, twx, twy);
            break;
        case MODE_PREC_FD:
            hipLaunchKernelGGL((op_kernel<T, NX, NY, MODE_PREC_FD, NT1>), grid, dim3(NT1), sh1, st, dwork, twx, twy);
            break;
        case MODE_PREC_TP:
            hipLaunchKernelGGL((op_kernel<T, NX, NY, MODE_PREC_TP, NT>), grid, dim3(NT), sh2, st, dwork, twx, twy);
            break;
        case MODE_FFT_FWD:
            hipLaunchKernelGGL((fft_kernel<T, NX, NY, NT1>), grid, dim3(NT1), sh1, st, dwork, twx, twy, 0);
            break;
        case MODE_FFT_INV:
            hipLaunchKernelGGL((fft_kernel<T, NX, NY, NT1>), grid, dim3(NT1), sh1, st, dwork, twx, twy, 1);
            break;
        default:
            throw std::runtime_error("unknown operator mode");
    }
    BH_HIP_CHECK(hipGetLastError());
}

static bool lds_grid(int nx, int ny) {
    return nx == ny && (nx == 8 || nx == 16 || nx == 24 || nx == 32 || nx == 48 || nx == 64);
}

static bool line_supported(int n) {
    switch (n) {
        case 8: case 16: case 24: case 32: case 48: case 64: case 96: case 128: case 192: case 256: case 384:
        case 512: return true;
        default: return false;
    }
}

bool grid_supported(int nx, int ny) { return lds_grid(nx, ny) || (line_supported(nx) && line_supported(ny)); }
bool grid_multipass(int nx, int ny) { return grid_supported(nx, ny) && !lds_grid(nx, ny); }

template <int N>
static std::vector<int> digit_reversal_n() {
    std::vector<int> r(N);
    for (int p = 0; p < N; ++p) {  // rev_index on the host
        int k = 0, mult = 1, rem = p, L = N;
        for (int s = 0; s < Plan<N>::K; ++s) {
            const int R = Plan<N>::r(s);
            L /= R;
            const int d = rem / L;
            rem -= d * L;
            k += d * mult;
            mult *= R;
        }
        r[p] = k;
    }
    return r;
}

std::vector<int> digit_reversal(int n) {
    switch (n) {
        case 8: return digit_reversal_n<8>();
        case 16: return digit_reversal_n<16>();
        case 24: return digit_reversal_n<24>();
        case 32: return digit_reversal_n<32>();
        case 48: return digit_reversal_n<48>();
        case 64: return digit_reversal_n<64>();
        case 96: return digit_reversal_n<96>();
        case 128: return digit_reversal_n<128>();
        case 192: return digit_reversal_n<192>();
        case 256: return digit_reversal_n<256>();
        case 384: return digit_reversal_n<384>();
        case 512: return digit_reversal_n<512>();
        default: throw std::runtime_error("FFT length " + std::to_string(n) + " has no radix plan");
    }
}

template <typename T>
void launch_op(int mode, int nx, int ny, const OpWork<T> *dwork, int nwork, int maxcols, const cplx<T> *twx,
               const cplx<T> *twy, hipStream_t st, bool prec_fp32, const int *revx, const int *revy) {
    if (nwork <= 0 || maxcols <= 0) return;
    if (!lds_grid(nx, ny)) {
        if (!grid_supported(nx, ny))
            throw std::runtime_error("grid " + std::to_string(nx) + " x " + std::to_string(ny) + " not supported");
        if (prec_fp32)
            throw std::runtime_error("FP32 preconditioners are instantiated for square grids up to 64 x 64");
        if (!revx || !revy) throw std::runtime_error("multi-pass operator without digit-reversal tables");
        launch_op_big<T>(mode, nx, ny, dwork, nwork, maxcols, twx, twy, revx, revy, st);
        return;
    }
    switch (nx) {
        case 8: launch_sized<T, 8, 8>(mode, dwork, nwork, maxcols, twx, twy, st, prec_fp32); break;
        case 16: launch_sized<T, 16, 16>(mode, dwork, nwork, maxcols, twx, twy, st, prec_fp32); break;
        case 24: launch_sized<T, 24, 24>(mode, dwork, nwork, maxcols, twx, twy, st, prec_fp32); break;
        case 32: launch_sized<T, 32, 32>(mode, dwork, nwork, maxcols, twx, twy, st, prec_fp32); break;
        case 48: launch_sized<T, 48, 48>(mode, dwork, nwork, maxcols, twx, twy, st, prec_fp32); break;
        case 64: launch_sized<T, 64, 64>(mode, dwork, nwork, maxcols, twx, twy, st, prec_fp32); break;
        default: throw std::runtime_error("grid size " + std::to_string(nx) + " not instantiated");
    }
}

template void launch_op<double>(int, int, int, const OpWork<double> *, int, int, const cplx<double> *,
                                const cplx<double> *, hipStream_t, bool, const int *, const int *);
template void launch_op<float>(int, int, int, const OpWork<float> *, int, int, const cplx<float> *,
                               const cplx<float> *, hipStream_t, bool, const int *, const int *);

}  // namespace bh
