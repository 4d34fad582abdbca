// Batched tall-skinny kernels for the LOBPCG subspace algebra on gfx950.
//
// Gram products (normalization.rs:313-345 SVQB Gram, eigensolver.rs:1232-1316
// Q^H A Q, deflation/overlap products) and linear combinations
// (normalization.rs:433-531 X T, eigensolver.rs:1341-1637 Ritz/W updates)
// for many independent instances in one launch.  Accumulation is always FP64.
#include "blas.h"

namespace bh {

// ---------------------------------------------------------------------------
// Gram: grid = (nchunks, nwork); 256 threads as 16 x 16, each owning a 3 x 3
// block of the (up to 48 x 48) output; rows staged through LDS 32 at a time.
// ---------------------------------------------------------------------------
constexpr int GRAM_SR = 32;

template <typename T>
__global__ void __launch_bounds__(256) gram_kernel(const GramWork<T> *__restrict__ work, long rows_per_chunk,
                                                   long nrows) {
    const GramWork<T> &w = work[blockIdx.y];
    const int na = w.dev_na ? *w.dev_na : w.na;
    const int nb = w.dev_nb ? *w.dev_nb : w.nb;
    if (na <= 0 || nb <= 0) return;
    __shared__ cplx<T> As[GRAM_SR][GRAM_MAXC + 1];
    __shared__ cplx<T> Bs[GRAM_SR][GRAM_MAXC + 1];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const bool lower = w.lower_only != 0;
    double ar[3][3] = {}, ai[3][3] = {};
    const long r0 = (long)blockIdx.x * rows_per_chunk;
    const long r1 = r0 + rows_per_chunk < nrows ? r0 + rows_per_chunk : nrows;
    for (long rb = r0; rb < r1; rb += GRAM_SR) {
        const int nr = (int)(r1 - rb < GRAM_SR ? r1 - rb : GRAM_SR);
        for (int e = threadIdx.x; e < GRAM_SR * na; e += 256) {
            int c = e / GRAM_SR, r = e - c * GRAM_SR;
            const cplx<T> *col = w.a_ptrs ? w.a_ptrs[c] : w.A + (long)c * w.a_stride;
            As[r][c] = r < nr ? col[rb + r] : cplx<T>{(T)0, (T)0};
        }
        for (int e = threadIdx.x; e < GRAM_SR * nb; e += 256) {
            int c = e / GRAM_SR, r = e - c * GRAM_SR;
            const cplx<T> *col = w.b_ptrs ? w.b_ptrs[c] : w.B + (long)c * w.b_stride;
            cplx<T> v = r < nr ? col[rb + r] : cplx<T>{(T)0, (T)0};
            if (w.weight && r < nr) {
                T s = w.weight[rb + r];
                v.x *= s;
                v.y *= s;
            }
            Bs[r][c] = v;
        }
        __syncthreads();
        for (int r = 0; r < nr; ++r) {
            double axr[3], axi[3], bxr[3], bxi[3];
#pragma unroll
            for (int u = 0; u < 3; ++u) {
                int i = ty + 16 * u;
                cplx<T> a = i < na ? As[r][i] : cplx<T>{(T)0, (T)0};
                axr[u] = (double)a.x;
                axi[u] = (double)a.y;
                int j = tx + 16 * u;
                cplx<T> b = j < nb ? Bs[r][j] : cplx<T>{(T)0, (T)0};
                bxr[u] = (double)b.x;
                bxi[u] = (double)b.y;
            }
#pragma unroll
            for (int u = 0; u < 3; ++u)
#pragma unroll
                for (int v = 0; v < 3; ++v) {
                    if (lower && v > u) continue;
                    ar[u][v] += axr[u] * bxr[v] + axi[u] * bxi[v];
                    ai[u][v] += axr[u] * bxi[v] - axi[u] * bxr[v];
                }
        }
        __syncthreads();
    }
    double *out = w.partial + 2 * (size_t)blockIdx.x * w.ld * w.ld;
#pragma unroll
    for (int u = 0; u < 3; ++u)
#pragma unroll
        for (int v = 0; v < 3; ++v) {
            int i = ty + 16 * u, j = tx + 16 * v;
            if (i < na && j < nb) {
                out[2 * (i * w.ld + j)] = ar[u][v];
                out[2 * (i * w.ld + j) + 1] = ai[u][v];
            }
        }
}

template <typename T>
void launch_gram(const GramWork<T> *dwork, int nwork, int nchunks, long rows, int maxa, int maxb, hipStream_t st) {
    if (nwork <= 0) return;
    if (maxa > GRAM_MAXC || maxb > GRAM_MAXC) throw std::runtime_error("gram operand wider than 48 columns");
    long per = (rows + nchunks - 1) / nchunks;
    per = (per + GRAM_SR - 1) / GRAM_SR * GRAM_SR;
    hipLaunchKernelGGL((gram_kernel<T>), dim3(nchunks, nwork), dim3(256), 0, st, dwork, per, rows);
    BH_HIP_CHECK(hipGetLastError());
}

// out(i, j) col-major [i + j*ld] = sum_c partial[c][i*ld + j]
__global__ void gram_reduce_kernel(const double *__restrict__ partial, double *__restrict__ out, int nchunks, int ld,
                                   long part_stride, long out_stride) {
    const double *p = partial + (size_t)blockIdx.y * part_stride;
    double *o = out + (size_t)blockIdx.y * out_stride;
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ld * ld) return;
    int i = e / ld, j = e - i * ld;
    double re = 0, im = 0;
    for (int c = 0; c < nchunks; ++c) {
        re += p[2 * ((size_t)c * ld * ld + e)];
        im += p[2 * ((size_t)c * ld * ld + e) + 1];
    }
    o[2 * (i + (size_t)j * ld)] = re;
    o[2 * (i + (size_t)j * ld) + 1] = im;
}

void launch_gram_reduce(const double *partial, double *out, int nwork, int nchunks, int ld, long part_stride,
                        long out_stride, hipStream_t st) {
    if (nwork <= 0) return;
    int n = ld * ld;
    hipLaunchKernelGGL(gram_reduce_kernel, dim3((n + 255) / 256, nwork), dim3(256), 0, st, partial, out, nchunks, ld,
                       part_stride, out_stride);
    BH_HIP_CHECK(hipGetLastError());
}

__global__ void reduce_kernel(const ReduceWork *__restrict__ work, int ld) {
    const ReduceWork &w = work[blockIdx.y];
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ld * ld) return;
    int i = e / ld, j = e - i * ld;
    double re = 0, im = 0;
    for (int c = 0; c < w.nchunks; ++c) {
        re += w.partial[2 * ((size_t)c * ld * ld + e)];
        im += w.partial[2 * ((size_t)c * ld * ld + e) + 1];
    }
    size_t o = (size_t)i + (size_t)j * ld;
    if (w.mask && !w.mask[o]) { re = 0; im = 0; }
    w.out[2 * o] = re;
    w.out[2 * o + 1] = im;
}

void launch_reduce(const ReduceWork *dwork, int nwork, int ld, hipStream_t st) {
    if (nwork <= 0) return;
    hipLaunchKernelGGL(reduce_kernel, dim3((ld * ld + 255) / 256, nwork), dim3(256), 0, st, dwork, ld);
    BH_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Linear combination: grid = (row blocks, nwork, output groups of QG).
// ---------------------------------------------------------------------------
constexpr int LC_QG = 16;
constexpr int LC_NT = 256;

template <typename T>
__global__ void __launch_bounds__(LC_NT) lincomb_kernel(const LincombWork<T> *__restrict__ work, long nrows) {
    const LincombWork<T> &w = work[blockIdx.y];
    const int na = w.dev_na ? *w.dev_na : w.na;
    const int nq = w.dev_nq ? *w.dev_nq : w.nq;
    const int q0 = blockIdx.z * LC_QG;
    if (q0 >= nq) return;
    const int qn = nq - q0 < LC_QG ? nq - q0 : LC_QG;
    __shared__ double Cs[2 * GRAM_MAXC * LC_QG];
    for (int e = threadIdx.x; e < na * LC_QG; e += LC_NT) {
        int i = e % na, j = e / na;
        double cr = 0, ci = 0;
        if (j < qn) {
            cr = w.C[2 * (i + (size_t)(q0 + j) * w.ldc)];
            ci = w.C[2 * (i + (size_t)(q0 + j) * w.ldc) + 1];
        }
        Cs[2 * (j * GRAM_MAXC + i)] = cr;
        Cs[2 * (j * GRAM_MAXC + i) + 1] = ci;
    }
    __syncthreads();
    const long row = (long)blockIdx.x * LC_NT + threadIdx.x;
    if (row >= nrows) return;
    double accr[LC_QG], acci[LC_QG];
#pragma unroll
    for (int j = 0; j < LC_QG; ++j) { accr[j] = 0; acci[j] = 0; }
    for (int i = 0; i < na; ++i) {
        const cplx<T> *col = w.a_ptrs ? w.a_ptrs[i] : w.A + (long)i * w.a_stride;
        cplx<T> a = col[row];
        double xr = (double)a.x, xi = (double)a.y;
#pragma unroll
        for (int j = 0; j < LC_QG; ++j) {
            double cr = Cs[2 * (j * GRAM_MAXC + i)], ci = Cs[2 * (j * GRAM_MAXC + i) + 1];
            accr[j] += xr * cr - xi * ci;
            acci[j] += xr * ci + xi * cr;
        }
    }
#pragma unroll
    for (int j = 0; j < LC_QG; ++j) {
        if (j >= qn) break;
        const long oc = q0 + j;
        double vr = accr[j], vi = acci[j];
        if (w.mode == 1) {
            cplx<T> b = w.base[oc * w.base_stride + row];
            vr = (double)b.x - vr;
            vi = (double)b.y - vi;
        }
        cplx<T> o = {(T)vr, (T)vi};
        w.out[oc * w.out_stride + row] = o;
        if (w.out2) {
            T s = w.w2[row];
            w.out2[oc * w.out_stride + row] = {o.x * s, o.y * s};
        }
    }
}

template <typename T>
void launch_lincomb(const LincombWork<T> *dwork, int nwork, long rows, int maxq, hipStream_t st) {
    if (nwork <= 0 || maxq <= 0) return;
    dim3 grid((unsigned)((rows + LC_NT - 1) / LC_NT), nwork, (maxq + LC_QG - 1) / LC_QG);
    hipLaunchKernelGGL((lincomb_kernel<T>), grid, dim3(LC_NT), 0, st, dwork, rows);
    BH_HIP_CHECK(hipGetLastError());
}

template void launch_gram<double>(const GramWork<double> *, int, int, long, int, int, hipStream_t);
template void launch_gram<float>(const GramWork<float> *, int, int, long, int, int, hipStream_t);
template void launch_lincomb<double>(const LincombWork<double> *, int, long, int, hipStream_t);
template void launch_lincomb<float>(const LincombWork<float> *, int, long, int, hipStream_t);

// ---------------------------------------------------------------------------
// Batched Hermitian eigensolver: one workgroup per matrix, parallel cyclic
// Jacobi (round-robin pairing, n/2 disjoint rotations per round) in LDS, FP64.
// Stands in for faer self_adjoint_eigen (dense.rs:71-134) in Rayleigh-Ritz
// and for hermitian_eigendecomposition in SVQB (normalization.rs:631-690);
// lower triangle read, stable ascending order.
// SVQB mode (normalization.rs:281-544): unit-B-norm column scaling D folded
// into the Gram, eigenpairs in descending order, drop lambda < 1e-12 lambda_max
// or <= 0, T = D Q_kept Lambda_kept^{-1/2}.
// ---------------------------------------------------------------------------
constexpr int EIGH_NT = 1024;

size_t eigh_shared_bytes(int maxn) {
    size_t ne = maxn + (maxn & 1), np = ne / 2;
    // A, V (complex), D (ne), rc (np), rs (complex np), rp/rq/rf (int np)
    return sizeof(double) * (4 * ne * ne + ne + np + 2 * np) + sizeof(int) * 3 * np;
}

__global__ void __launch_bounds__(1024) eigh_kernel(const EighWork *__restrict__ work) {
    const EighWork &w = work[blockIdx.x];
    const int n = w.dev_n ? *w.dev_n : w.n;
    const int tid = threadIdx.x;
    if (n <= 0) {
        if (tid == 0) {
            if (w.rank_out) *w.rank_out = 0;
            if (w.cnt_x) *w.cnt_x = 0;
            if (w.cnt_w) *w.cnt_w = 0;
        }
        return;
    }
    const int ne = n + (n & 1);
    const int np = ne / 2;
    extern __shared__ double sm[];
    double2 *A = reinterpret_cast<double2 *>(sm);  // col-major [i + j*ne]
    double2 *V = A + ne * ne;
    double2 *rs = V + ne * ne;                      // 16-byte aligned block first
    double *D = reinterpret_cast<double *>(rs + np);
    double *rc = D + ne;
    int *rp = reinterpret_cast<int *>(rc + np);
    int *rq = rp + np;
    int *rf = rq + np;  // 0 identity, 1 rotate, 2 negligible (zero the pair)
    __shared__ int nrot;

    const int ld = w.ld;
    auto G = [&](int a, int b) -> double2 {
        double re = 0, im = 0;
        for (int c = 0; c < w.nchunks; ++c) {
            const double *p = w.partial + 2 * ((size_t)c * ld * ld + (size_t)a * ld + b);
            re += p[0];
            im += p[1];
        }
        return make_double2(re, im);
    };
    for (int e = tid; e < ne * ne; e += EIGH_NT) {
        int i = e % ne, j = e / ne;
        double2 v = make_double2(0, 0);
        if (i < n && j < n) {
            if (i > j) v = G(i, j);
            else if (i < j) { double2 g = G(j, i); v = make_double2(g.x, -g.y); }
            else v = make_double2(G(i, i).x, 0.0);
        }
        A[e] = v;
        V[e] = make_double2(i == j ? 1.0 : 0.0, 0.0);
    }
    __syncthreads();
    if (w.mode >= 1) {
        for (int i = tid; i < ne; i += EIGH_NT) {
            double nrm = i < n ? sqrt(fmax(A[i + i * ne].x, 0.0)) : 0.0;
            D[i] = nrm > 1e-30 ? 1.0 / nrm : 1.0;
        }
        __syncthreads();
        for (int e = tid; e < ne * ne; e += EIGH_NT) {
            int i = e % ne, j = e / ne;
            double s = D[i] * D[j];
            A[e].x *= s;
            A[e].y *= s;
        }
        __syncthreads();
    }
    if (w.mode == 2) {
        // Cholesky QR of the scaled Gram G = L L^H, T = D L^{-H}: the same span
        // as SVQB when no direction would be dropped.  Falls back to the SVQB
        // eigendecomposition below when a pivot shows G is ill-conditioned.
        double2 *L = V;  // V is free until the Jacobi path starts
        __shared__ int chol_ok;
        if (tid == 0) chol_ok = 1;
        for (int e = tid; e < ne * ne; e += EIGH_NT) L[e] = A[e];
        __syncthreads();
        for (int j = 0; j < n; ++j) {
            if (tid == 0) {
                double d = L[j + j * ne].x;
                for (int k = 0; k < j; ++k) {
                    double2 l = L[j + k * ne];
                    d -= l.x * l.x + l.y * l.y;
                }
                if (!(d > 1e-8)) chol_ok = 0;  // relative pivot: columns have unit B-norm
                L[j + j * ne] = make_double2(d > 0.0 ? sqrt(d) : 1.0, 0.0);
            }
            __syncthreads();
            if (!chol_ok) break;
            const double ljj = L[j + j * ne].x;
            for (int i = j + 1 + tid; i < n; i += EIGH_NT) {
                double2 a = L[i + j * ne];
                for (int k = 0; k < j; ++k) {  // a -= L[i,k] conj(L[j,k])
                    double2 x = L[i + k * ne], y = L[j + k * ne];
                    a.x -= x.x * y.x + x.y * y.y;
                    a.y -= x.y * y.x - x.x * y.y;
                }
                L[i + j * ne] = make_double2(a.x / ljj, a.y / ljj);
            }
            __syncthreads();
        }
        if (chol_ok) {
            // M = L^{-1} by forward substitution, one column per thread; T = D M^H
            for (int c = tid; c < n; c += EIGH_NT) {
                // M[i][c] for i >= c; kept in the A buffer's column c (A no longer needed)
                for (int i = 0; i < n; ++i) A[i + c * ne] = make_double2(0.0, 0.0);
                for (int i = c; i < n; ++i) {
                    double2 acc = make_double2(i == c ? 1.0 : 0.0, 0.0);
                    for (int k = c; k < i; ++k) {
                        double2 l = L[i + k * ne], m = A[k + c * ne];
                        acc.x -= l.x * m.x - l.y * m.y;
                        acc.y -= l.x * m.y + l.y * m.x;
                    }
                    double lii = L[i + i * ne].x;
                    A[i + c * ne] = make_double2(acc.x / lii, acc.y / lii);
                }
            }
            __syncthreads();
            for (int e = tid; e < n * n; e += EIGH_NT) {
                int r = e % n, c = e / n;  // T[r][c] = D[r] conj(M[c][r]), zero below the diagonal
                double2 m = r <= c ? A[c + r * ne] : make_double2(0.0, 0.0);
                size_t o = 2 * ((size_t)r + (size_t)c * w.ldv);
                w.V[o] = D[r] * m.x;
                w.V[o + 1] = -D[r] * m.y;
            }
            if (tid == 0) {
                if (w.rank_out) *w.rank_out = n;
                if (w.sweeps_out) atomicAdd(w.sweeps_out + 1, 1);  // Cholesky path taken
            }
            return;
        }
        // fallback: restore V = I and run SVQB's eigendecomposition
        for (int e = tid; e < ne * ne; e += EIGH_NT) {
            int i = e % ne, j = e / ne;
            V[e] = make_double2(i == j ? 1.0 : 0.0, 0.0);
        }
        __syncthreads();
    }
    // Each round applies np disjoint rotations.  The matrix is partitioned
    // into np x np 2x2 blocks (row pair, column pair); a thread owning a block
    // applies both the column rotation of its column pair and the row
    // rotation of its row pair in registers, so a round needs two barriers.
    __shared__ unsigned long long maxratio_bits;  // non-negative doubles order like their bit patterns
    // fixed per-thread work for every round: one 2x2 block (np*np <= 1024) and
    // up to two eigenvector items (np*ne <= 2048)
    const bool has_blk = tid < np * np;
    const int bt1 = has_blk ? tid / np : 0, bt2 = has_blk ? tid - bt1 * np : 0;
    const int nv = np * ne;
    const int vt0 = tid < nv ? tid / ne : 0, vk0 = tid < nv ? tid - vt0 * ne : 0;
    const int vt1 = tid + EIGH_NT < nv ? (tid + EIGH_NT) / ne : 0;
    const int vk1 = tid + EIGH_NT < nv ? (tid + EIGH_NT) - vt1 * ne : 0;
    // round-robin pairing for this thread's pair slot (tid < np), per round r:
    // slot 0: (r, ne-1); slot t: ((r+t) mod (ne-1), (r-t) mod (ne-1))
    for (int sweep = 0; sweep < 40; ++sweep) {
        if (tid == 0) { nrot = 0; maxratio_bits = 0ull; }
        __syncthreads();
        for (int r = 0; r < ne - 1; ++r) {
            if (tid < np) {
                int p, q;
                if (tid == 0) { p = r; q = ne - 1; }
                else { p = (r + tid) % (ne - 1); q = (r - tid + ne - 1) % (ne - 1); }
                if (p > q) { int t = p; p = q; q = t; }
                double2 h = A[p + q * ne];
                double habs = sqrt(h.x * h.x + h.y * h.y);
                double app = A[p + p * ne].x, aqq = A[q + q * ne].x;
                double c = 1.0;
                int flag = 0;
                double2 sg = make_double2(0.0, 0.0);
                double scale = sqrt(fabs(app) * fabs(aqq));
                if (habs != 0.0 && habs > 2.2e-16 * scale) {
                    flag = 1;
                    double tau = (aqq - app) / (2.0 * habs);
                    double t = (tau >= 0 ? -1.0 : 1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                    c = 1.0 / sqrt(1.0 + t * t);
                    double s = t * c;
                    sg = make_double2(s * h.x / habs, -s * h.y / habs);  // sigma = s conj(h/|h|)
                    atomicAdd(&nrot, 1);
                    double ratio = scale > 0.0 ? habs / scale : 1.0;
                    atomicMax(&maxratio_bits, (unsigned long long)__double_as_longlong(ratio));
                } else if (habs != 0.0) {
                    flag = 2;
                }
                rf[tid] = flag;
                rp[tid] = p;
                rq[tid] = q;
                rc[tid] = c;
                rs[tid] = sg;
            }
            __syncthreads();
            do {
                if (!has_blk) break;
                const int t1 = bt1, t2 = bt2;  // row pair, column pair
                const int f1 = rf[t1], f2 = rf[t2];
                if (f1 != 1 && f2 != 1) {
                    if (t1 == t2 && f1 == 2) {
                        int p = rp[t1], q = rq[t1];
                        A[p + q * ne] = make_double2(0.0, 0.0);
                        A[q + p * ne] = make_double2(0.0, 0.0);
                    }
                    break;
                }
                const int p1 = rp[t1], q1 = rq[t1], p2 = rp[t2], q2 = rq[t2];
                double2 a11 = A[p1 + p2 * ne], a12 = A[p1 + q2 * ne];
                double2 a21 = A[q1 + p2 * ne], a22 = A[q1 + q2 * ne];
                if (f2 == 1) {  // columns: [:,p] = c [:,p] + s [:,q]; [:,q] = -conj(s) [:,p] + c [:,q]
                    const double c = rc[t2];
                    const double2 s = rs[t2];
                    double2 n11 = make_double2(c * a11.x + s.x * a12.x - s.y * a12.y, c * a11.y + s.x * a12.y + s.y * a12.x);
                    double2 n12 = make_double2(-(s.x * a11.x + s.y * a11.y) + c * a12.x, -(s.x * a11.y - s.y * a11.x) + c * a12.y);
                    double2 n21 = make_double2(c * a21.x + s.x * a22.x - s.y * a22.y, c * a21.y + s.x * a22.y + s.y * a22.x);
                    double2 n22 = make_double2(-(s.x * a21.x + s.y * a21.y) + c * a22.x, -(s.x * a21.y - s.y * a21.x) + c * a22.y);
                    a11 = n11; a12 = n12; a21 = n21; a22 = n22;
                }
                if (f1 == 1) {  // rows: [p,:] = c [p,:] + conj(s) [q,:]; [q,:] = -s [p,:] + c [q,:]
                    const double c = rc[t1];
                    const double2 s = rs[t1];
                    double2 n11 = make_double2(c * a11.x + s.x * a21.x + s.y * a21.y, c * a11.y + s.x * a21.y - s.y * a21.x);
                    double2 n21 = make_double2(-(s.x * a11.x - s.y * a11.y) + c * a21.x, -(s.x * a11.y + s.y * a11.x) + c * a21.y);
                    double2 n12 = make_double2(c * a12.x + s.x * a22.x + s.y * a22.y, c * a12.y + s.x * a22.y - s.y * a22.x);
                    double2 n22 = make_double2(-(s.x * a12.x - s.y * a12.y) + c * a22.x, -(s.x * a12.y + s.y * a12.x) + c * a22.y);
                    a11 = n11; a12 = n12; a21 = n21; a22 = n22;
                }
                if (t1 == t2) {  // the rotated pair: exact zero off-diagonal, real diagonal
                    a12 = make_double2(0.0, 0.0);
                    a21 = make_double2(0.0, 0.0);
                    a11.y = 0.0;
                    a22.y = 0.0;
                }
                A[p1 + p2 * ne] = a11;
                A[p1 + q2 * ne] = a12;
                A[q1 + p2 * ne] = a21;
                A[q1 + q2 * ne] = a22;
            } while (false);
            for (int e = tid, it = 0; e < nv; e += EIGH_NT, ++it) {  // eigenvectors: column rotations only
                const int t = it == 0 ? vt0 : vt1, k = it == 0 ? vk0 : vk1;
                if (rf[t] != 1) continue;
                double c = rc[t];
                double2 sg = rs[t];
                int p = rp[t], q = rq[t];
                double2 vkp = V[k + p * ne], vkq = V[k + q * ne];
                V[k + p * ne] = make_double2(c * vkp.x + sg.x * vkq.x - sg.y * vkq.y,
                                             c * vkp.y + sg.x * vkq.y + sg.y * vkq.x);
                V[k + q * ne] = make_double2(-(sg.x * vkp.x + sg.y * vkp.y) + c * vkq.x,
                                             -(sg.x * vkp.y - sg.y * vkp.x) + c * vkq.y);
            }
            __syncthreads();
        }
        // quadratic convergence: rotations of relative size < 1e-9 leave
        // off-diagonal terms below 1e-18 relative, so no further sweep is needed
        const bool done = nrot == 0 || __longlong_as_double((long long)maxratio_bits) < 1e-9;
        if (done || sweep == 39) {
            if (tid == 0 && w.sweeps_out) {
                atomicAdd(w.sweeps_out, sweep + 1);
                atomicMax(w.sweeps_out + 2, sweep + 1);  // maximum over the launch
            }
            break;
        }
        __syncthreads();
    }
    // stable ascending ranks over the first n diagonal entries
    __shared__ double wsh[EIGH_MAXN];
    __shared__ int rank_of[EIGH_MAXN];
    for (int i = tid; i < n; i += EIGH_NT) wsh[i] = A[i + i * ne].x;
    __syncthreads();
    for (int i = tid; i < n; i += EIGH_NT) {
        double wi = wsh[i];
        int rk = 0;
        for (int j = 0; j < n; ++j) rk += (wsh[j] < wi || (wsh[j] == wi && j < i)) ? 1 : 0;
        rank_of[i] = rk;
    }
    __syncthreads();
    if (w.mode == 0) {
        const int cx = n < w.nsplit ? n : w.nsplit;
        const int cw = (2 * w.nsplit < n ? 2 * w.nsplit : n) - cx;
        for (int i = tid; i < n; i += EIGH_NT) w.w[rank_of[i]] = wsh[i];
        for (int e = tid; e < n * n; e += EIGH_NT) {
            int row = e % n, i = e / n;
            double2 v = V[row + i * ne];
            int col = rank_of[i];
            size_t o = 2 * ((size_t)row + (size_t)col * w.ldv);
            w.V[o] = v.x;
            w.V[o + 1] = v.y;
            if (w.Vw && col >= cx && col < cx + cw) {
                size_t ow = 2 * ((size_t)row + (size_t)(col - cx) * w.ldv);
                w.Vw[ow] = v.x;
                w.Vw[ow + 1] = v.y;
            }
        }
        if (tid == 0) {
            if (w.rank_out) *w.rank_out = n;
            if (w.cnt_x) *w.cnt_x = cx;
            if (w.cnt_w) *w.cnt_w = cw;
        }
    } else {
        // descending position = n - 1 - ascending rank
        __shared__ double mx;
        if (tid == 0) {
            double m = 0.0;
            for (int i = 0; i < n; ++i) m = fmax(m, wsh[i]);
            mx = m;
        }
        __syncthreads();
        const double thr = 1e-12 * mx;
        // kept = prefix of the descending order with ev >= thr && ev > 0
        __shared__ int rank;
        if (tid == 0) {
            int r = 0;
            for (int i = 0; i < n; ++i)
                if (wsh[i] >= thr && wsh[i] > 0.0) ++r;
            rank = r;
            if (w.rank_out) *w.rank_out = r;
        }
        __syncthreads();
        for (int e = tid; e < n * n; e += EIGH_NT) {
            int row = e % n, i = e / n;
            int pos = n - 1 - rank_of[i];
            if (pos >= rank) continue;
            double lam = wsh[i];
            double sc = D[row] / sqrt(fmax(lam, 1e-30));
            double2 v = V[row + i * ne];
            size_t o = 2 * ((size_t)row + (size_t)pos * w.ldv);
            w.V[o] = v.x * sc;
            w.V[o + 1] = v.y * sc;
        }
        for (int i = tid; i < n; i += EIGH_NT) {
            int pos = n - 1 - rank_of[i];
            w.w[pos] = wsh[i];
        }
    }
}

void launch_eigh(const EighWork *dwork, int nwork, int maxn, hipStream_t st) {
    if (nwork <= 0) return;
    if (maxn > EIGH_MAXN) throw std::runtime_error("eigh dimension above 64");
    int ne = maxn + (maxn & 1);
    int np = ne / 2;
    size_t sh = eigh_shared_bytes(maxn);
    static const bool attr_set = [] {
        BH_HIP_CHECK(hipFuncSetAttribute((const void *)eigh_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)eigh_shared_bytes(EIGH_MAXN)));
        return true;
    }();
    (void)attr_set;
    hipLaunchKernelGGL(eigh_kernel, dim3(nwork), dim3(EIGH_NT), sh, st, dwork);
    BH_HIP_CHECK(hipGetLastError());
}

}  // namespace bh
