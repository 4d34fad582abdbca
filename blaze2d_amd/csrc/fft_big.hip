// Multi-pass FFT operators for grids whose field does not fit one workgroup's
// LDS (256 x 256, 512 x 512) and for rectangular grids.
//
// A 2D transform is split into column passes and row passes.  A column pass
// owns a tile of BIG_TC adjacent columns over all ny rows (each row segment a
// contiguous 128-byte read), runs the y transform of every column of the tile
// in LDS and writes the tile back.  A row pass owns whole rows, runs the x
// transform, the Fourier-space step of the operator and the inverse x
// transform in LDS.  Between passes the field goes through HBM (and L2 /
// Infinity Cache, which hold a 1 MB field).  As in the LDS-resident kernels,
// forward transforms are decimation in frequency (natural order in,
// digit-reversed out) and inverse ones decimation in time, so the operators
// never reorder: the Fourier-space step reads the frequency of each position
// from the digit-reversal tables (rows: compile-time, columns: revy).
//
// Operators (reference citations as in fft_ops.hip):
//   THETA_TM  colF -> row(|k+G|^2) -> colI                       maxwell.rs:686-713
//   PREC_FD   colF -> row(FD diagonal) -> colI                    fourier_diagonal.rs:310-342
//   THETA_TE  colF -> row(grad) -> col(eta, both fields) -> row(div) -> colI
//                                                                 maxwell.rs:647-684,1745-1839
//   PREC_TP   colF -> row(grad / (k^2+s)) -> col(eps) -> row(div / (k^2+s)) -> colI
//                                                                 transverse_projection.rs:167-231
//   FFT_FWD / FFT_INV: natural order in and out (backend-cpu/src/lib.rs:94-275)
// TE and TP carry their second gradient component in OpWork::tmp.
#include "fft_core.h"
#include "fft_ops.h"

#include <string>

namespace bh {

constexpr int BIG_NT = 256;  // threads per workgroup
constexpr int BIG_TC = 8;    // columns per column-pass tile (8 x 16 B = one 128-byte segment per row in FP64)
// rows per row-pass workgroup (ny is a multiple of 8 for every supported length)
template <int L>
constexpr int big_rows() {
    return L <= 256 ? 8 : 4;
}

enum ColStage { COL_FWD = 0, COL_INV = 1, COL_MID_TE = 2, COL_MID_TP = 3, COL_FWD_NAT = 4, COL_INV_NAT = 5 };
enum RowStage {
    ROW_TM = 0,
    ROW_FD = 1,
    ROW_TE1 = 2,
    ROW_TP1 = 3,
    ROW_TE2 = 4,
    ROW_TP2 = 5,
    ROW_FWD_NAT = 6,
    ROW_INV_NAT = 7
};

struct BigGrid {
    int nx, ny;
    const int *revy;  // frequency index at each digit-reversed row position
};

template <int STAGE>
constexpr int col_fields() {
    return (STAGE == COL_MID_TE || STAGE == COL_MID_TP) ? 2 : 1;
}
template <int STAGE>
constexpr int row_fields() {
    return (STAGE >= ROW_TE1 && STAGE <= ROW_TP2) ? 2 : 1;
}

template <typename T, int L, int STAGE>
__global__ void __launch_bounds__(BIG_NT) col_kernel(const OpWork<T> *__restrict__ work,
                                                     const cplx<T> *__restrict__ gtw, BigGrid g) {
    constexpr int TC = BIG_TC, NF = col_fields<STAGE>(), TILE = L * TC;
    extern __shared__ __attribute__((aligned(16))) char smem_raw[];
    cplx<T> *F = reinterpret_cast<cplx<T> *>(smem_raw);  // NF tiles, [y][c]
    cplx<T> *tw = F + NF * TILE;
    const OpWork<T> &w = work[blockIdx.z];
    const int count = w.dev_count ? *w.dev_count : w.count;
    const int col = blockIdx.y;
    if (col >= count) return;
    const int nx = g.nx;
    const long x0 = (long)blockIdx.x * TC;
    // the first pass of an operator reads the input; later passes work in place on the output
    constexpr bool FIRST = STAGE == COL_FWD || STAGE == COL_FWD_NAT || STAGE == COL_INV_NAT;
    const cplx<T> *src = FIRST ? w.in + (size_t)col * w.in_stride : w.out + (size_t)col * w.out_stride;
    cplx<T> *dst = w.out + (size_t)col * w.out_stride;
    cplx<T> *tmp = NF == 2 ? w.tmp + (size_t)col * w.tmp_stride : nullptr;
    for (int i = threadIdx.x; i < L; i += BIG_NT) tw[i] = gtw[i];
    for (int e = threadIdx.x; e < TILE; e += BIG_NT) {
        const int y = e / TC, c = e - y * TC;
        // natural-order input of an inverse transform: position y holds frequency revy[y]
        const int ys = STAGE == COL_INV_NAT ? g.revy[y] : y;
        F[e] = gload(src, (long)ys * nx + x0 + c);
        if constexpr (NF == 2) F[TILE + e] = gload(tmp, (long)y * nx + x0 + c);
    }
    __syncthreads();
    if constexpr (STAGE == COL_FWD || STAGE == COL_FWD_NAT) {
        fft_passes_fwd<T, L, TC, false, false, 1, BIG_NT, TILE, 0, 0>(F, tw);
    } else if constexpr (STAGE == COL_INV || STAGE == COL_INV_NAT) {
        fft_passes_inv<T, L, TC, false, true, 1, BIG_NT, TILE, 0, Plan<L>::K - 1>(F, tw);
    } else {
        // both gradient components back to real space along y (x already is),
        // the real-space step at natural (x, y), then forward along y again
        fft_passes_inv<T, L, TC, false, true, 2, BIG_NT, TILE, 0, Plan<L>::K - 1>(F, tw);
        for (int e = threadIdx.x; e < TILE; e += BIG_NT) {
            const int y = e / TC, c = e - y * TC;
            const long i = (long)y * nx + x0 + c;
            const cplx<T> gx = F[e], gy = F[TILE + e];
            if constexpr (STAGE == COL_MID_TE) {  // eta tensor [xx, xy, yx, yy]
                const auto *t = gmem(w.tensor + 4 * i);
                const T t0 = t[0], t1 = t[1], t2 = t[2], t3 = t[3];
                F[e] = {gx.x * t0 + gy.x * t1, gx.y * t0 + gy.y * t1};
                F[TILE + e] = {gx.x * t2 + gy.x * t3, gx.y * t2 + gy.y * t3};
            } else {  // eps
                const T ev = gmem(w.eps)[i];
                F[e] = {gx.x * ev, gx.y * ev};
                F[TILE + e] = {gy.x * ev, gy.y * ev};
            }
        }
        __syncthreads();
        fft_passes_fwd<T, L, TC, false, false, 2, BIG_NT, TILE, 0, 0>(F, tw);
    }
    for (int e = threadIdx.x; e < TILE; e += BIG_NT) {
        const int y = e / TC, c = e - y * TC;
        // natural-order output of a forward transform: frequency revy[y] to row revy[y]
        const int yd = STAGE == COL_FWD_NAT ? g.revy[y] : y;
        gstore(dst, (long)yd * nx + x0 + c, F[e]);
        if constexpr (NF == 2) gstore(tmp, (long)y * nx + x0 + c, F[TILE + e]);
    }
}

template <typename T, int L, int STAGE>
__global__ void __launch_bounds__(BIG_NT) row_kernel(const OpWork<T> *__restrict__ work,
                                                     const cplx<T> *__restrict__ gtw, BigGrid g) {
    constexpr int R = big_rows<L>(), NF = row_fields<STAGE>(), RL = R * L;
    constexpr bool TWO_IN = STAGE == ROW_TE2 || STAGE == ROW_TP2;
    constexpr bool TWO_OUT = STAGE == ROW_TE1 || STAGE == ROW_TP1;
    extern __shared__ __attribute__((aligned(16))) char smem_raw[];
    cplx<T> *F = reinterpret_cast<cplx<T> *>(smem_raw);  // NF blocks of R rows
    cplx<T> *G2 = F + RL;
    cplx<T> *tw = F + NF * RL;
    const OpWork<T> &w = work[blockIdx.z];
    const int count = w.dev_count ? *w.dev_count : w.count;
    const int col = blockIdx.y;
    if (col >= count) return;
    const int ny = g.ny;
    const int y0 = blockIdx.x * R;
    cplx<T> *buf = w.out + (size_t)col * w.out_stride;
    cplx<T> *tmp = NF == 2 ? w.tmp + (size_t)col * w.tmp_stride : nullptr;
    const T invN = (T)(1.0 / ((double)L * ny));
    for (int i = threadIdx.x; i < L; i += BIG_NT) tw[i] = gtw[i];
    for (int e = threadIdx.x; e < RL; e += BIG_NT) {
        const int r = e / L, x = e - r * L;
        const int xs = STAGE == ROW_INV_NAT ? rev_index<L>(x) : x;
        F[e] = gload(buf, (long)(y0 + r) * L + xs);
        if constexpr (TWO_IN) G2[e] = gload(tmp, (long)(y0 + r) * L + x);
    }
    __syncthreads();
    if constexpr (STAGE == ROW_INV_NAT) {
        fft_passes_inv<T, L, R, true, true, 1, BIG_NT, RL, 0, Plan<L>::K - 1>(F, tw);
        for (int e = threadIdx.x; e < RL; e += BIG_NT) {
            const int r = e / L, x = e - r * L;
            gstore(buf, (long)(y0 + r) * L + x, scale_c(F[e], invN));
        }
        return;
    }
    fft_passes_fwd<T, L, R, true, false, TWO_IN ? 2 : 1, BIG_NT, RL, 0, 0>(F, tw);
    if constexpr (STAGE == ROW_FWD_NAT) {
        for (int e = threadIdx.x; e < RL; e += BIG_NT) {
            const int r = e / L, x = e - r * L;
            gstore(buf, (long)(y0 + r) * L + rev_index<L>(x), F[e]);
        }
        return;
    } else {
        const OpParams P = *w.params;
        for (int e = threadIdx.x; e < RL; e += BIG_NT) {
            const int r = e / L, px = e - r * L;
            double kx, ky, ksq;
            bool cl;
            kplusg_at(P, rev_index<L>(px), g.revy[y0 + r], L, ny, kx, ky, ksq, cl);
            if constexpr (STAGE == ROW_TM || STAGE == ROW_FD) {
                double fac;
                if (STAGE == ROW_TM) {
                    fac = ksq;
                } else {  // inverse_scale (maxwell.rs:1692-1711) with the Gamma DC mask
                    const double ee = P.eps_eff;
                    if (!(ee > 0.0) || !isfinite(ee) || !isfinite(ksq)) fac = 0.0;
                    else {
                        const double mf = (isfinite(P.mass_floor) && P.mass_floor > 0.0) ? P.mass_floor : 0.0;
                        fac = ee / (fmax(ksq, 0.0) + mf + P.shift * fmax(ee, 1e-12));
                    }
                    if (P.gamma && cl) fac = 0.0;
                }
                F[e] = scale_c(F[e], (T)fac * invN);
            } else if constexpr (STAGE == ROW_TE1) {  // u (i kx), u (i ky)
                const cplx<T> u = F[e];
                const T skx = (T)kx * invN, sky = (T)ky * invN;
                F[e] = {-u.y * skx, u.x * skx};
                G2[e] = {-u.y * sky, u.x * sky};
            } else if constexpr (STAGE == ROW_TP1) {  // u (-i kx) / (k^2 + s), u (-i ky) / (k^2 + s)
                const cplx<T> u = F[e];
                double ik = 0.0;
                if (!cl) {
                    const double den = ksq + P.shift;
                    ik = den > 1e-15 ? 1.0 / den : 0.0;
                }
                const T fx = (T)(kx * ik) * invN, fy = (T)(ky * ik) * invN;
                F[e] = {u.y * fx, -u.x * fx};
                G2[e] = {u.y * fy, -u.x * fy};
            } else {  // divergence: -(i kx gx + i ky gy) [ / (k^2 + s) for TP ]
                double sc = 1.0;
                if (STAGE == ROW_TP2) {
                    double ik = 0.0;
                    if (!cl) {
                        const double den = ksq + P.shift;
                        ik = den > 1e-15 ? 1.0 / den : 0.0;
                    }
                    sc = ik;
                }
                const T skx = (T)kx, sky = (T)ky, s = (T)sc * invN;
                const cplx<T> gx = F[e], gy = G2[e];
                const T dre = -(skx * gx.y + sky * gy.y);
                const T dim = skx * gx.x + sky * gy.x;
                F[e] = {-dre * s, -dim * s};
            }
        }
        __syncthreads();
        fft_passes_inv<T, L, R, true, true, TWO_OUT ? 2 : 1, BIG_NT, RL, 0, Plan<L>::K - 1>(F, tw);
        for (int e = threadIdx.x; e < RL; e += BIG_NT) {
            const int r = e / L, x = e - r * L;
            gstore(buf, (long)(y0 + r) * L + x, F[e]);
            if constexpr (TWO_OUT) gstore(tmp, (long)(y0 + r) * L + x, G2[e]);
        }
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
template <typename T, int L, int STAGE>
static void launch_col(const OpWork<T> *dwork, int nwork, int maxcols, const cplx<T> *tw, const BigGrid &g,
                       hipStream_t st) {
    const size_t sh = sizeof(cplx<T>) * ((size_t)col_fields<STAGE>() * L * BIG_TC + L);
    static const bool attr = [sh] {
        BH_HIP_CHECK(hipFuncSetAttribute((const void *)col_kernel<T, L, STAGE>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
        return true;
    }();
    (void)attr;
    hipLaunchKernelGGL((col_kernel<T, L, STAGE>), dim3(g.nx / BIG_TC, maxcols, nwork), dim3(BIG_NT), sh, st, dwork,
                       tw, g);
    BH_HIP_CHECK(hipGetLastError());
}

template <typename T, int L, int STAGE>
static void launch_row(const OpWork<T> *dwork, int nwork, int maxcols, const cplx<T> *tw, const BigGrid &g,
                       hipStream_t st) {
    const size_t sh = sizeof(cplx<T>) * ((size_t)row_fields<STAGE>() * big_rows<L>() * L + L);
    static const bool attr = [sh] {
        BH_HIP_CHECK(hipFuncSetAttribute((const void *)row_kernel<T, L, STAGE>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
        return true;
    }();
    (void)attr;
    hipLaunchKernelGGL((row_kernel<T, L, STAGE>), dim3(g.ny / big_rows<L>(), maxcols, nwork), dim3(BIG_NT), sh, st,
                       dwork, tw, g);
    BH_HIP_CHECK(hipGetLastError());
}

#define BH_BIG_LENGTHS(X) X(8) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512)

template <typename T, int STAGE>
static void col_pass(const OpWork<T> *dwork, int nwork, int maxcols, const cplx<T> *twy, const BigGrid &g,
                     hipStream_t st) {
    switch (g.ny) {
#define BH_CASE(n) \
    case n: launch_col<T, n, STAGE>(dwork, nwork, maxcols, twy, g, st); return;
        BH_BIG_LENGTHS(BH_CASE)
#undef BH_CASE
        default: throw std::runtime_error("no column pass for ny = " + std::to_string(g.ny));
    }
}

template <typename T, int STAGE>
static void row_pass(const OpWork<T> *dwork, int nwork, int maxcols, const cplx<T> *twx, const BigGrid &g,
                     hipStream_t st) {
    switch (g.nx) {
#define BH_CASE(n) \
    case n: launch_row<T, n, STAGE>(dwork, nwork, maxcols, twx, g, st); return;
        BH_BIG_LENGTHS(BH_CASE)
#undef BH_CASE
        default: throw std::runtime_error("no row pass for nx = " + std::to_string(g.nx));
    }
}

template <typename T>
void launch_op_big(int mode, int nx, int ny, const OpWork<T> *dwork, int nwork, int maxcols, const cplx<T> *twx,
                   const cplx<T> *twy, const int *revx, const int *revy, hipStream_t st) {
    (void)revx;  // the row passes take the x digit reversal from the compile-time plan
    if (nwork <= 0 || maxcols <= 0) return;
    if (nx % BIG_TC || ny % 8) throw std::runtime_error("multi-pass grid sides must be multiples of 8");
    const BigGrid g{nx, ny, revy};
    switch (mode) {
        case MODE_THETA_TM:
            col_pass<T, COL_FWD>(dwork, nwork, maxcols, twy, g, st);
            row_pass<T, ROW_TM>(dwork, nwork, maxcols, twx, g, st);
            col_pass<T, COL_INV>(dwork, nwork, maxcols, twy, g, st);
            break;
        case MODE_PREC_FD:
            col_pass<T, COL_FWD>(dwork, nwork, maxcols, twy, g, st);
            row_pass<T, ROW_FD>(dwork, nwork, maxcols, twx, g, st);
            col_pass<T, COL_INV>(dwork, nwork, maxcols, twy, g, st);
            break;
        case MODE_THETA_TE:
            col_pass<T, COL_FWD>(dwork, nwork, maxcols, twy, g, st);
            row_pass<T, ROW_TE1>(dwork, nwork, maxcols, twx, g, st);
            col_pass<T, COL_MID_TE>(dwork, nwork, maxcols, twy, g, st);
            row_pass<T, ROW_TE2>(dwork, nwork, maxcols, twx, g, st);
            col_pass<T, COL_INV>(dwork, nwork, maxcols, twy, g, st);
            break;
        case MODE_PREC_TP:
            col_pass<T, COL_FWD>(dwork, nwork, maxcols, twy, g, st);
            row_pass<T, ROW_TP1>(dwork, nwork, maxcols, twx, g, st);
            col_pass<T, COL_MID_TP>(dwork, nwork, maxcols, twy, g, st);
            row_pass<T, ROW_TP2>(dwork, nwork, maxcols, twx, g, st);
            col_pass<T, COL_INV>(dwork, nwork, maxcols, twy, g, st);
            break;
        case MODE_FFT_FWD:
            col_pass<T, COL_FWD_NAT>(dwork, nwork, maxcols, twy, g, st);
            row_pass<T, ROW_FWD_NAT>(dwork, nwork, maxcols, twx, g, st);
            break;
        case MODE_FFT_INV:
            col_pass<T, COL_INV_NAT>(dwork, nwork, maxcols, twy, g, st);
            row_pass<T, ROW_INV_NAT>(dwork, nwork, maxcols, twx, g, st);
            break;
        default:
            throw std::runtime_error("unknown operator mode");
    }
}

template void launch_op_big<double>(int, int, int, const OpWork<double> *, int, int, const cplx<double> *,
                                    const cplx<double> *, const int *, const int *, hipStream_t);
template void launch_op_big<float>(int, int, int, const OpWork<float> *, int, int, const cplx<float> *,
                                   const cplx<float> *, const int *, const int *, hipStream_t);

}  // namespace bh
