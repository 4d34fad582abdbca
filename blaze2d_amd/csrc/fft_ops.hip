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
// Radix plans: N = R0 * R1 * ... (largest radix first).  Radix 8 halves the
// number of LDS round trips of a 64-point line compared with radix 4.
// ---------------------------------------------------------------------------
template <int N>
struct Plan;
#define BH_PLAN(NN, KK, A, B, C, D)                                                      \
    template <>                                                                          \
    struct Plan<NN> {                                                                    \
        static constexpr int K = KK;                                                     \
        __host__ __device__ static constexpr int r(int s) {                              \
            return s == 0 ? A : s == 1 ? B : s == 2 ? C : D;                             \
        }                                                                                \
    };
BH_PLAN(8, 2, 4, 2, 1, 1)
BH_PLAN(16, 2, 4, 4, 1, 1)
BH_PLAN(24, 3, 4, 3, 2, 1)
BH_PLAN(32, 2, 8, 4, 1, 1)
BH_PLAN(48, 3, 4, 4, 3, 1)
BH_PLAN(64, 2, 8, 8, 1, 1)
#undef BH_PLAN

template <int N>
__host__ __device__ constexpr int plan_len(int s) {  // transform length at pass s
    int L = N;
    for (int i = 0; i < s; ++i) L /= Plan<N>::r(i);
    return L;
}

// frequency index held at digit-reversed storage position p
template <int N>
__device__ inline int rev_index(int p) {
    int k = 0, mult = 1, rem = p, L = N;
#pragma unroll
    for (int s = 0; s < Plan<N>::K; ++s) {
        const int R = Plan<N>::r(s);
        L /= R;
        int d = rem / L;
        rem -= d * L;
        k += d * mult;
        mult *= R;
    }
    return k;
}

// ---------------------------------------------------------------------------
// Small DFT butterflies (in registers).  INV selects e^{+2 pi i/R}.
// ---------------------------------------------------------------------------
template <typename T>
__device__ inline cplx<T> mul_mi(cplx<T> a) { return {a.y, -a.x}; }  // * (-i)
template <typename T>
__device__ inline cplx<T> mul_pi(cplx<T> a) { return {-a.y, a.x}; }  // * (+i)

template <typename T, int R, bool INV>
__device__ inline void small_dft(cplx<T> *v) {
    if constexpr (R == 2) {
        cplx<T> a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 3) {
        const T s = (T)0.86602540378443864676;
        cplx<T> t1 = cadd(v[1], v[2]);
        cplx<T> t2 = {v[0].x - (T)0.5 * t1.x, v[0].y - (T)0.5 * t1.y};
        cplx<T> t3 = {s * (v[1].x - v[2].x), s * (v[1].y - v[2].y)};
        v[0] = cadd(v[0], t1);
        if (!INV) {
            v[1] = cadd(t2, mul_mi(t3));
            v[2] = cadd(t2, mul_pi(t3));
        } else {
            v[1] = cadd(t2, mul_pi(t3));
            v[2] = cadd(t2, mul_mi(t3));
        }
    } else if constexpr (R == 4) {
        cplx<T> s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        cplx<T> s13 = cadd(v[1], v[3]), d13 = csub(v[1], v[3]);
        v[0] = cadd(s02, s13);
        v[2] = csub(s02, s13);
        if (!INV) {
            v[1] = cadd(d02, mul_mi(d13));
            v[3] = cadd(d02, mul_pi(d13));
        } else {
            v[1] = cadd(d02, mul_pi(d13));
            v[3] = cadd(d02, mul_mi(d13));
        }
    } else if constexpr (R == 8) {
        // radix 2 x 4: length-4 DFTs of the even and odd samples, odd half
        // twiddled by W8^k (W8 = e^{-+ i pi/4}), then one radix-2 stage
        cplx<T> e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
        small_dft<T, 4, INV>(e);
        small_dft<T, 4, INV>(o);
        const T h = (T)0.70710678118654752440;
        if (!INV) {
            o[1] = {h * (o[1].x + o[1].y), h * (o[1].y - o[1].x)};   // * (1 - i)/sqrt2
            o[2] = mul_mi(o[2]);                                      // * -i
            o[3] = {h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y)};  // * (-1 - i)/sqrt2
        } else {
            o[1] = {h * (o[1].x - o[1].y), h * (o[1].y + o[1].x)};   // * (1 + i)/sqrt2
            o[2] = mul_pi(o[2]);                                      // * +i
            o[3] = {-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y)};  // * (-1 + i)/sqrt2
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            v[k] = cadd(e[k], o[k]);
            v[k + 4] = csub(e[k], o[k]);
        }
    }
}

// LDS position of the field element with flat index i (row-major, rows RW
// long).  For 64-wide rows the low three bits of the column are XORed with the
// next three: a radix-8 pass over rows whose eight lanes read elements 8 apart
// (128 bytes, one bank cycle) then hits eight different bank groups instead of
// one, and the accesses that are contiguous across lanes stay within the same
// 128-byte block.  It is a bijection on every 64-element row.
template <int RW>
__device__ inline int lds_pos(int i) {
    if constexpr (RW == 64) return i ^ ((i >> 3) & 7);
    else return i;
}

// Where a pass takes its input or puts its output: the LDS field, or (for
// the first forward and the last inverse pass of a one-field operator) the
// field in HBM, which saves one LDS round trip and one barrier at each end.
struct PassIO {
    const void *gin = nullptr;   // global input (natural order)
    void *gout = nullptr;        // global output (natural order)
    const void *geps = nullptr;  // optional: gbout = geps * output
    void *gbout = nullptr;
};

// One radix pass over NLINES independent length-N transforms stored in LDS.
// ROWS: elements contiguous (stride 1), lines NX apart; else elements NX
// apart, lines contiguous.  NF fields are processed together.  GIN / GOUT:
// read the input from / write the output to HBM (io) instead of LDS.
template <typename T, int N, int NLINES, bool ROWS, int S, bool INV, int NF, int NT, int FSTRIDE, int RW,
          bool GIN = false, bool GOUT = false, typename TS = T>
__device__ inline void fft_pass(cplx<T> *buf, const cplx<T> *__restrict__ tw, const PassIO &io = PassIO()) {
    constexpr int R = Plan<N>::r(S);
    constexpr int L = plan_len<N>(S);
    constexpr int M = L / R;
    constexpr int NBF = N / R;
    constexpr int TOTAL = NBF * NLINES * NF;
    constexpr int TWS = N / L;
    constexpr int ES = ROWS ? 1 : NLINES;  // element stride
    constexpr int LS = ROWS ? N : 1;       // line stride
#pragma unroll 1
    for (int id = threadIdx.x; id < TOTAL; id += NT) {
        int f = id / (NBF * NLINES);
        int rem = id - f * (NBF * NLINES);
        int line, bf;
        if (ROWS) { line = rem / NBF; bf = rem - line * NBF; }
        else { bf = rem / NLINES; line = rem - bf * NLINES; }
        int b = bf / M, j2 = bf - b * M;
        const int e0 = f * FSTRIDE + line * LS + (b * L + j2) * ES;
        cplx<T> v[R];
#pragma unroll
        for (int j = 0; j < R; ++j) {
            if constexpr (GIN) {  // HBM in storage precision TS
                const cplx<TS> g = gload(static_cast<const cplx<TS> *>(io.gin), e0 + j * M * ES);
                v[j] = {(T)g.x, (T)g.y};
            }
            else v[j] = buf[lds_pos<RW>(e0 + j * M * ES)];
        }
        if (!INV) {
            small_dft<T, R, false>(v);
            if (M > 1) {
#pragma unroll
                for (int k1 = 1; k1 < R; ++k1) v[k1] = cmul(v[k1], tw[j2 * k1 * TWS]);
            }
        } else {
            if (M > 1) {
#pragma unroll
                for (int k1 = 1; k1 < R; ++k1) v[k1] = cmulc(v[k1], tw[j2 * k1 * TWS]);
            }
            small_dft<T, R, true>(v);
        }
#pragma unroll
        for (int j = 0; j < R; ++j) {
            if constexpr (GOUT) {
                const int e = e0 + j * M * ES;
                const cplx<TS> o = {(TS)v[j].x, (TS)v[j].y};
                gstore(static_cast<cplx<TS> *>(io.gout), e, o);
                if (io.gbout) {
                    const TS w = gmem(static_cast<const TS *>(io.geps))[e];
                    gstore(static_cast<cplx<TS> *>(io.gbout), e, cplx<TS>{o.x * w, o.y * w});
                }
            } else {
                buf[lds_pos<RW>(e0 + j * M * ES)] = v[j];
            }
        }
    }
}

template <typename T, int N, int NLINES, bool ROWS, bool INV, int NF, int NT, int FSTRIDE, int RW, int S>
__device__ inline void fft_passes_fwd(cplx<T> *buf, const cplx<T> *tw) {
    if constexpr (S < Plan<N>::K) {
        fft_pass<T, N, NLINES, ROWS, S, INV, NF, NT, FSTRIDE, RW>(buf, tw);
        __syncthreads();
        fft_passes_fwd<T, N, NLINES, ROWS, INV, NF, NT, FSTRIDE, RW, S + 1>(buf, tw);
    }
}
template <typename T, int N, int NLINES, bool ROWS, bool INV, int NF, int NT, int FSTRIDE, int RW, int S,
          int SEND = 0>
__device__ inline void fft_passes_inv(cplx<T> *buf, const cplx<T> *tw) {
    if constexpr (S >= SEND) {
        fft_pass<T, N, NLINES, ROWS, S, INV, NF, NT, FSTRIDE, RW>(buf, tw);
        __syncthreads();
        fft_passes_inv<T, N, NLINES, ROWS, INV, NF, NT, FSTRIDE, RW, S - 1, SEND>(buf, tw);
    }
}

// natural -> digit-reversed (rows then columns)
template <typename T, int NX, int NY, int NF, int NT>
__device__ inline void fft2d_forward(cplx<T> *buf, const cplx<T> *twx, const cplx<T> *twy) {
    fft_passes_fwd<T, NX, NY, true, false, NF, NT, NX * NY, NX, 0>(buf, twx);
    fft_passes_fwd<T, NY, NX, false, false, NF, NT, NX * NY, NX, 0>(buf, twy);
}
// digit-reversed -> natural, unscaled
template <typename T, int NX, int NY, int NF, int NT>
__device__ inline void fft2d_inverse(cplx<T> *buf, const cplx<T> *twx, const cplx<T> *twy) {
    fft_passes_inv<T, NY, NX, false, true, NF, NT, NX * NY, NX, Plan<NY>::K - 1>(buf, twy);
    fft_passes_inv<T, NX, NY, true, true, NF, NT, NX * NY, NX, Plan<NX>::K - 1>(buf, twx);
}
// one field, natural order in HBM -> digit-reversed in LDS (the first row
// pass reads HBM; twiddles must be visible in LDS before the call)
template <typename T, int NX, int NY, int NT, typename TS = T>
__device__ inline void fft2d_forward_from(cplx<T> *buf, const cplx<TS> *gin, const cplx<T> *twx,
                                          const cplx<T> *twy) {
    PassIO io;
    io.gin = gin;
    fft_pass<T, NX, NY, true, 0, false, 1, NT, NX * NY, NX, true, false, TS>(buf, twx, io);
    __syncthreads();
    fft_passes_fwd<T, NX, NY, true, false, 1, NT, NX * NY, NX, 1>(buf, twx);
    fft_passes_fwd<T, NY, NX, false, false, 1, NT, NX * NY, NX, 0>(buf, twy);
}
// one field, digit-reversed in LDS -> natural order in HBM (the last row pass
// writes gout, and gbout = geps * gout when gbout is set), unscaled
template <typename T, int NX, int NY, int NT, typename TS = T>
__device__ inline void fft2d_inverse_to(cplx<T> *buf, const PassIO &io, const cplx<T> *twx, const cplx<T> *twy) {
    fft_passes_inv<T, NY, NX, false, true, 1, NT, NX * NY, NX, Plan<NY>::K - 1>(buf, twy);
    fft_passes_inv<T, NX, NY, true, true, 1, NT, NX * NY, NX, Plan<NX>::K - 1, 1>(buf, twx);
    fft_pass<T, NX, NY, true, 0, true, 1, NT, NX * NY, NX, false, true, TS>(buf, twx, io);
}

// k+G at the frequency stored at LDS position (px, py), with the reference's
// |k+G|^2 <= 1e-9 clamp (maxwell.rs:1609-1690).
template <int NX, int NY>
__device__ inline void kplusg(const OpParams &P, int px, int py, double &kx, double &ky, double &ksq,
                              bool &clamped) {
    int kxi = rev_index<NX>(px), kyi = rev_index<NY>(py);
    double n1 = kxi <= NX / 2 ? kxi : kxi - NX;
    double n2 = kyi <= NY / 2 ? kyi : kyi - NY;
    double gx = n1 * P.b1[0] + n2 * P.b2[0];
    double gy = n1 * P.b1[1] + n2 * P.b2[1];
    double rx = gx + P.bloch[0], ry = gy + P.bloch[1];
    double raw = rx * rx + ry * ry;
    clamped = raw <= KG_FLOOR;
    if (raw >= KG_FLOOR) { kx = rx; ky = ry; }
    else if (raw == 0.0) { kx = sqrt(KG_FLOOR); ky = 0.0; }
    else { double s = sqrt(KG_FLOOR / raw); kx = rx * s; ky = ry * s; }
    ksq = kx * kx + ky * ky;
}

template <typename T>
__device__ inline cplx<T> scale_c(cplx<T> a, T s) { return {a.x * s, a.y * s}; }

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
    fft2d_forward_from<T, NX, NY, NT, TS>(F, in, twx, twy);

    if constexpr (MODE == MODE_THETA_TM || MODE == MODE_PREC_FD) {
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
        // gradient step: THETA_TE  g = i(k+G) u_hat ; PREC_TP  g = -i(k+G) u_hat / (|k+G|^2 + s)
#pragma unroll 2
        for (int i = threadIdx.x; i < N; i += NT) {
            int py = i / NX, px = i - py * NX;
            double kx, ky, ksq;
            bool cl;
            kplusg<NX, NY>(P, px, py, kx, ky, ksq, cl);
            cplx<T> u = F[lds_pos<NX>(i)];
            if (MODE == MODE_THETA_TE) {
                T skx = (T)kx * invN, sky = (T)ky * invN;
                F[lds_pos<NX>(i)] = {-u.y * skx, u.x * skx};   // u * (0 + i kx)
                G2[lds_pos<NX>(i)] = {-u.y * sky, u.x * sky};
            } else {
                double ik = 0.0;
                if (!cl) { double den = ksq + P.shift; ik = den > 1e-15 ? 1.0 / den : 0.0; }
                T fx = (T)(kx * ik) * invN, fy = (T)(ky * ik) * invN;
                F[lds_pos<NX>(i)] = {u.y * fx, -u.x * fx};     // u * (0 - i kx ik)
                G2[lds_pos<NX>(i)] = {u.y * fy, -u.x * fy};
            }
        }
        __syncthreads();
        fft2d_inverse<T, NX, NY, 2, NT>(F, twx, twy);
        // real-space step: THETA_TE eta tensor [xx, xy, yx, yy]; PREC_TP eps
#pragma unroll 2
        for (int i = threadIdx.x; i < N; i += NT) {
            cplx<T> gx = F[lds_pos<NX>(i)], gy = G2[lds_pos<NX>(i)];
            if (MODE == MODE_THETA_TE) {
                const auto *t = gmem(w.tensor + 4 * (size_t)i);
                T t0 = (T)t[0], t1 = (T)t[1], t2 = (T)t[2], t3 = (T)t[3];
                F[lds_pos<NX>(i)] = {gx.x * t0 + gy.x * t1, gx.y * t0 + gy.y * t1};
                G2[lds_pos<NX>(i)] = {gx.x * t2 + gy.x * t3, gx.y * t2 + gy.y * t3};
            } else {
                T e = (T)gmem(w.eps)[i];
                F[lds_pos<NX>(i)] = {gx.x * e, gx.y * e};
                G2[lds_pos<NX>(i)] = {gy.x * e, gy.y * e};
            }
        }
        __syncthreads();
        fft2d_forward<T, NX, NY, 2, NT>(F, twx, twy);
        // divergence: out_hat = -(i kx gx + i ky gy) [ * 1/(|k+G|^2+s) for TP ]
#pragma unroll 2
        for (int i = threadIdx.x; i < N; i += NT) {
            int py = i / NX, px = i - py * NX;
            double kx, ky, ksq;
            bool cl;
            kplusg<NX, NY>(P, px, py, kx, ky, ksq, cl);
            double sc = 1.0;
            if (MODE == MODE_PREC_TP) {
                double ik = 0.0;
                if (!cl) { double den = ksq + P.shift; ik = den > 1e-15 ? 1.0 / den : 0.0; }
                sc = ik;
            }
            T skx = (T)kx, sky = (T)ky, s = (T)sc * invN;
            cplx<T> gx = F[lds_pos<NX>(i)], gy = G2[lds_pos<NX>(i)];
            // div = i kx gx + i ky gy ; out = -div
            T dre = -(skx * gx.y + sky * gy.y);
            T dim = skx * gx.x + sky * gy.x;
            F[lds_pos<NX>(i)] = {-dre * s, -dim * s};
        }
        __syncthreads();
        PassIO io;
        io.gout = out;
        fft2d_inverse_to<T, NX, NY, NT, TS>(F, io, twx, twy);
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

bool grid_supported(int nx, int ny) {
    return nx == ny && (nx == 8 || nx == 16 || nx == 24 || nx == 32 || nx == 48 || nx == 64);
}

template <typename T>
void launch_op(int mode, int nx, int ny, const OpWork<T> *dwork, int nwork, int maxcols, const cplx<T> *twx,
               const cplx<T> *twy, hipStream_t st, bool prec_fp32) {
    if (nwork <= 0 || maxcols <= 0) return;
    if (nx != ny) throw std::runtime_error("only square grids are instantiated");
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
                                const cplx<double> *, hipStream_t, bool);
template void launch_op<float>(int, int, int, const OpWork<float> *, int, int, const cplx<float> *,
                               const cplx<float> *, hipStream_t, bool);

}  // namespace bh
