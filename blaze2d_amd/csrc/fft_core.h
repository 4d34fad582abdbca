// Device building blocks of the FFT operators: radix plans, in-register
// butterflies, LDS radix passes and the in-kernel k+G, shared by the
// LDS-resident single-workgroup operators (fft_ops.hip) and the multi-pass
// operators for large and rectangular grids (fft_big.hip).
#pragma once

#include "common.h"

namespace bh {


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
BH_PLAN(96, 3, 8, 4, 3, 1)
BH_PLAN(128, 3, 8, 4, 4, 1)
BH_PLAN(192, 3, 8, 8, 3, 1)
BH_PLAN(256, 3, 8, 8, 4, 1)
BH_PLAN(384, 4, 8, 4, 4, 3)
BH_PLAN(512, 3, 8, 8, 8, 1)
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

// One radix pass like fft_pass, with the element loads and stores done by the
// caller: ld(f, i) returns element i (flat row-major position) of field f and
// st(f, i, v) stores it.  This folds a pointwise step into a pass that reads or
// writes the data anyway.  Each element of a field is loaded and stored by one
// thread, so ld and st may also touch other fields at position i, as long as
// no thread of the same pass writes what they read.
template <typename T, int N, int NLINES, bool ROWS, int S, bool INV, int NF, int NT, class LD, class ST>
__device__ inline void fft_pass_fn(const cplx<T> *__restrict__ tw, LD ld, ST st) {
    constexpr int R = Plan<N>::r(S);
    constexpr int L = plan_len<N>(S);
    constexpr int M = L / R;
    constexpr int NBF = N / R;
    constexpr int TOTAL = NBF * NLINES * NF;
    constexpr int TWS = N / L;
    constexpr int ES = ROWS ? 1 : NLINES;
    constexpr int LS = ROWS ? N : 1;
#pragma unroll 1
    for (int id = threadIdx.x; id < TOTAL; id += NT) {
        const int f = id / (NBF * NLINES);
        const int rem = id - f * (NBF * NLINES);
        int line, bf;
        if (ROWS) { line = rem / NBF; bf = rem - line * NBF; }
        else { bf = rem / NLINES; line = rem - bf * NLINES; }
        const int b = bf / M, j2 = bf - b * M;
        const int e0 = line * LS + (b * L + j2) * ES;
        cplx<T> v[R];
#pragma unroll
        for (int j = 0; j < R; ++j) v[j] = ld(f, e0 + j * M * ES);
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
        for (int j = 0; j < R; ++j) st(f, e0 + j * M * ES, v[j]);
    }
}

// passes S .. SEND - 1 of a forward transform, a barrier after each
template <typename T, int N, int NLINES, bool ROWS, bool INV, int NF, int NT, int FSTRIDE, int RW, int S,
          int SEND = Plan<N>::K>
__device__ inline void fft_passes_fwd(cplx<T> *buf, const cplx<T> *tw) {
    if constexpr (S < SEND) {
        fft_pass<T, N, NLINES, ROWS, S, INV, NF, NT, FSTRIDE, RW>(buf, tw);
        __syncthreads();
        fft_passes_fwd<T, N, NLINES, ROWS, INV, NF, NT, FSTRIDE, RW, S + 1, SEND>(buf, tw);
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
// k+G at the FFT frequency indices (kxi, kyi) of an nx x ny grid.
__device__ inline void kplusg_at(const OpParams &P, int kxi, int kyi, int nx, int ny, double &kx, double &ky,
                                 double &ksq, bool &clamped) {
    double n1 = kxi <= nx / 2 ? kxi : kxi - nx;
    double n2 = kyi <= ny / 2 ? kyi : kyi - ny;
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

}  // namespace bh
