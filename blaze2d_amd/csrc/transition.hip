// k-point transitions in coefficient space, one wave per slot.
//
// A slot that finished the solve at one k-point and moves to the next needs
// band tracking (band_tracking.rs:294-700), the history update and
// predict_with_rotation (subspace_prediction.rs:248-380), and the start block
// of the next solve (initialize_block, initialization.rs:248-367, with the
// Gamma deflation of :166-224).  All of it is linear algebra inside
// span K = [E | P | R | y0 | Wc], so with the B-weighted Gram M = K^H B K the
// inner products <a, b>_B become a^H M b over coefficient vectors of length
// kd <= 99.  The kernel reads M where the Gram reduction left it and writes
// the coefficients that the engine's linear combinations turn into the new
// history and the new block, so a k-point change needs no readback: the
// tracking permutation is read back with the step's other results, for the
// record of the finished k-point only.
//
// Layout: every coefficient vector is spread over the lanes (element i on
// lane i mod 64), so dot products are wave reductions (DPP) and updates touch
// only the lane's own elements; independent inner products (Gram-Schmidt
// against one vector) take one lane each.  M is kept in full as a
// Hermitian matrix.  The assignment problem runs column-per-lane; only the short
// block scans of the tracking rules run on lane 0.
#include "transition.h"

#include "jacobi.h"

namespace bh {

namespace {

constexpr int TR_NT = 64;  // one wave

inline __host__ __device__ size_t tr_take(size_t &o, size_t bytes) {
    const size_t r = o;
    o += (bytes + 15) & ~size_t(15);
    return r;
}

// Byte offsets of the kernel's arrays in its workspace (LDS or per-item memory).
struct TransLayout {
    size_t mp, warm, mc, gr, tv, tmv, col, cb, A, V, rot, red, Wt, Ut, Up, fw, mag, hu, sig, ord, perm, tperm, ci,
        lp, amb, blk, rr, tk, dk, total;
    __host__ __device__ TransLayout(int KD, int NBW, int M0, int MM, int NC) {
        size_t o = 0;
        const size_t z = 16;  // complex FP64
        const size_t ne = MM + (MM & 1);
        const int NE = NBW > M0 ? NBW : M0;  // entries of one Gram-Schmidt
        mp = tr_take(o, z * KD * KD);
        // candidates and their images M c; the block's entries and their
        // images replace them in place
        warm = tr_take(o, z * NC * KD);
        mc = tr_take(o, z * NC * KD);
        gr = tr_take(o, z * NC * NC);  // the candidates' Gram
        tv = tr_take(o, z * KD);
        tmv = tr_take(o, z * KD);
        col = tr_take(o, z * MM);
        cb = tr_take(o, z * NC);
        // the polar and tracking arrays, free once the warm vectors are built
        const size_t p0 = o;
        A = tr_take(o, z * ne * ne);
        V = tr_take(o, z * ne * ne);
        rot = tr_take(o, sizeof(JacobiRot) * ne);  // two rounds of ne / 2 rotations
        red = tr_take(o, 8);
        Wt = tr_take(o, z * MM * MM);
        Ut = tr_take(o, z * MM * MM);
        Up = tr_take(o, z * MM * MM);
        fw = tr_take(o, 8 * MM * MM);
        mag = tr_take(o, 8 * MM * MM);
        // the Gram-Schmidt coefficients reuse them when they fit
        const size_t rsz = z * NE * NC;
        rr = rsz <= o - p0 ? p0 : tr_take(o, rsz);
        hu = tr_take(o, 8 * (MM + 1));
        sig = tr_take(o, 8 * MM);
        ord = tr_take(o, 4 * MM);
        perm = tr_take(o, 4 * KD);
        tperm = tr_take(o, 4 * MM);
        ci = tr_take(o, 4 * MM);
        lp = tr_take(o, 4 * MM);
        amb = tr_take(o, MM);
        blk = tr_take(o, 4 * 8 * (MM + 1));
        tk = tr_take(o, 4 * NE);
        dk = tr_take(o, 8 * NE);  // 1 / d_k of the entries
        total = o;
    }
};

// Wave-wide reductions of a double through DPP row operations (quad
// permutations, half-row and row mirrors, then the broadcasts of lanes 15 and
// 31 into the rows above), read from lane 63: no LDS round trips, unlike a
// shuffle through ds_bpermute.  Every lane of the wave must be active.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ double dpp_d(double old, double v) {
    const long long b = __double_as_longlong(v), o = __double_as_longlong(old);
    const int lo = __builtin_amdgcn_update_dpp((int)(o & 0xffffffff), (int)(b & 0xffffffff), CTRL, ROW_MASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(b >> 32), CTRL, ROW_MASK, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double lane63_d(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), 63);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ void wave_sum2(double &a, double &b) {
    a += dpp_d<0xb1>(0.0, a);          // quad_perm [1,0,3,2]
    b += dpp_d<0xb1>(0.0, b);
    a += dpp_d<0x4e>(0.0, a);          // quad_perm [2,3,0,1]
    b += dpp_d<0x4e>(0.0, b);
    a += dpp_d<0x141>(0.0, a);         // row_half_mirror: sums of 8
    b += dpp_d<0x141>(0.0, b);
    a += dpp_d<0x140>(0.0, a);         // row_mirror: sums of 16
    b += dpp_d<0x140>(0.0, b);
    a += dpp_d<0x142, 0xa>(0.0, a);    // row_bcast15 into rows 1 and 3
    b += dpp_d<0x142, 0xa>(0.0, b);
    a += dpp_d<0x143, 0xc>(0.0, a);    // row_bcast31 into rows 2 and 3: lane 63 holds the total
    b += dpp_d<0x143, 0xc>(0.0, b);
    a = lane63_d(a);
    b = lane63_d(b);
}
__device__ __forceinline__ double wave_min(double v) {
    v = fmin(v, dpp_d<0xb1>(v, v));
    v = fmin(v, dpp_d<0x4e>(v, v));
    v = fmin(v, dpp_d<0x141>(v, v));
    v = fmin(v, dpp_d<0x140>(v, v));
    v = fmin(v, dpp_d<0x142, 0xa>(v, v));
    v = fmin(v, dpp_d<0x143, 0xc>(v, v));
    return lane63_d(v);
}
__device__ __forceinline__ double wave_max(double v) { return -wave_min(-v); }
// sums over groups of S adjacent lanes (S = 1, 2, 4, 8 or 16, uniform), in every lane of the group
__device__ __forceinline__ void group_sum2(double &a, double &b, int S) {
    if (S >= 2) {
        a += dpp_d<0xb1>(0.0, a);   // quad_perm [1,0,3,2]: pairs
        b += dpp_d<0xb1>(0.0, b);
    }
    if (S >= 4) {
        a += dpp_d<0x4e>(0.0, a);   // quad_perm [2,3,0,1]: quads
        b += dpp_d<0x4e>(0.0, b);
    }
    if (S >= 8) {
        a += dpp_d<0x141>(0.0, a);  // row_half_mirror: eights
        b += dpp_d<0x141>(0.0, b);
    }
    if (S >= 16) {
        a += dpp_d<0x140>(0.0, a);  // row_mirror: sixteens
        b += dpp_d<0x140>(0.0, b);
    }
}
// lanes per item for n items of length len over one wave: the smallest power of
// two S <= 16 with n * S >= 64 spread, so that short items use the whole wave
__device__ __forceinline__ int lanes_per_item(int n, int len) {
    int S = 1;
    while (S < 16 && n * S * 2 <= TR_NT && len >= 2 * S) S *= 2;
    return S;
}

// Maximum-weight assignment of an n x n weight matrix (w[i + j n] = weight of
// row i to column j, n < 64) by the shortest-augmenting-path Hungarian method
// on the costs max(w) - w, as band_tracking.rs's hungarian_algorithm; perm[i]
// is the column of row i.  All lanes of the wave: column j's potential,
// matching, path and slack live on lane j, the row potentials u in memory,
// and each pass's smallest slack is a wave minimum taken at the lowest such
// column (the order of the sequential scan).  Returns false (perm =
// identity) when the weights are not finite.
__device__ __forceinline__ bool hungarian_wave(int n, const double *w, int *perm, double *u) {
    const int lane = threadIdx.x;
    for (int i = lane; i < n; i += TR_NT) perm[i] = i;
    __syncthreads();
    if (n <= 1) return true;
    double mx = -INFINITY;
    for (int e = lane; e < n * n; e += TR_NT) mx = fmax(mx, w[e]);
    mx = wave_max(mx);
    const bool col = lane >= 1 && lane <= n;  // lane j holds column j (lane 0: the virtual column)
    double v = 0.0;
    int p = 0, way = 0;
    if (lane <= n) u[lane] = 0.0;
    __syncthreads();
    for (int i = 1; i <= n; ++i) {
        if (lane == 0) p = i;
        int j0 = 0;
        double minv = INFINITY;
        bool used = false;
        for (int pass = 0; pass <= n + 1; ++pass) {
            if (lane == j0) used = true;
            const int i0 = __builtin_amdgcn_readlane(p, j0);
            const double ui0 = u[i0];
            double cand = INFINITY;
            if (col && !used) {
                const double cur = (mx - w[(i0 - 1) + (lane - 1) * n]) - ui0 - v;
                if (cur < minv) {
                    minv = cur;
                    way = j0;
                }
                cand = minv;
            }
            const double delta = wave_min(cand);
            if (!(delta < INFINITY)) {  // no finite reduced cost
                __syncthreads();
                for (int k = lane; k < n; k += TR_NT) perm[k] = k;
                __syncthreads();
                return false;
            }
            const int j1 = __ffsll((unsigned long long)__ballot(col && !used && cand == delta)) - 1;
            __syncthreads();
            if (lane <= n) {
                if (used) {
                    u[p] += delta;  // the used columns' rows are distinct
                    v -= delta;
                } else {
                    minv -= delta;
                }
            }
            __syncthreads();
            j0 = j1;
            if (__builtin_amdgcn_readlane(p, j0) == 0) break;
        }
        for (int guard = 0; guard <= n && j0 != 0; ++guard) {  // augment along the path
            const int j1 = __builtin_amdgcn_readlane(way, j0);
            const int pj1 = __builtin_amdgcn_readlane(p, j1);
            if (lane == j0) p = pj1;
            j0 = j1;
        }
    }
    __syncthreads();
    if (col && p >= 1 && p <= n) perm[p - 1] = lane - 1;
    __syncthreads();
    return true;
}

// band_tracking.rs:294-700: assignment of the previous k-point's bands (rows
// of the polar factor U of the overlap, m x m) to the current ones by |U|,
// with the reference's refinements: rows whose rotation is ambiguous (largest
// |U| below 0.7 or second largest above half of it) that form blocks of
// near-equal previous frequencies are reassigned by frequency, and so are
// blocks of consecutive bands whose assigned frequencies jump.  All lanes;
// the block rules are short scans on lane 0.
struct TrackScratch {
    double *mag, *fw, *hu;
    int *ci, *lp, *blk;
    char *amb;
};

__device__ __forceinline__ void track_assign(int m, const double2 *U, const double *pw, int npw, const double *cw, int ncw, int *perm,
                             const TrackScratch &s) {
    const int lane = threadIdx.x;
    // |U| <= 1: no scaling needed
    for (int e = lane; e < m * m; e += TR_NT) s.mag[e] = sqrt(U[e].x * U[e].x + U[e].y * U[e].y);
    __syncthreads();
    // detect_degenerate_bands_from_rotation: the ambiguous rows; each row's
    // largest entry's column is kept for the shortcut below
    int jmax = lane;
    bool clear = true;  // this lane's row is not ambiguous
    if (m >= 2 && lane < m) {
        double r0 = -INFINITY, r1 = -INFINITY;
        for (int j = 0; j < m; ++j) {
            const double x = s.mag[lane + j * m];
            if (x > r0) {
                r1 = r0;
                r0 = x;
                jmax = j;
            } else if (x > r1) {
                r1 = x;
            }
        }
        const double ratio = r0 > 1e-10 ? r1 / r0 : 0.0;
        clear = !(r0 < 0.7 || ratio > 0.5);
        s.amb[lane] = clear ? 0 : 1;
        s.ci[lane] = jmax;
    }
    __syncthreads();
    // When no row is ambiguous (each row's largest |U| above 0.7, the next at
    // most half of it) and the rows' largest entries lie in distinct columns,
    // that assignment is the unique maximum-weight one: any other permutation
    // takes at most half of the maximum in every row where it differs.  The
    // assignment problem and the ambiguous-block refinement are then skipped.
    bool distinct = true;
    if (m >= 2 && lane < m)
        for (int i = 0; i < lane; ++i) distinct = distinct && s.ci[i] != jmax;
    const bool shortcut = m >= 2 && !__any(!(lane >= m || (clear && distinct)));
    __syncthreads();
    // blk: [0] rb count, [1] db count, [2] mb count, then (start, size) pairs
    int *const cnt = s.blk;
    int *const rb = s.blk + 3, *const db = rb + 2 * (m + 1), *const mb = db + 2 * (m + 1);
    auto near = [](double a, double b) {
        const double gap = fabs(a - b);
        if (gap <= 5e-3) return true;
        return gap / fmax(0.5 * (fabs(a) + fabs(b)), 1e-6) <= 0.015;
    };
    if (lane == 0) {
        int nrb = 0, ndb = 0;
        if (m >= 2) {  // runs of ambiguous rows longer than one
            int bs = -1;
            for (int i = 0; i < m; ++i) {
                if (s.amb[i]) {
                    if (bs < 0) bs = i;
                } else if (bs >= 0) {
                    if (i - bs > 1) { rb[2 * nrb] = bs; rb[2 * nrb + 1] = i - bs; ++nrb; }
                    bs = -1;
                }
            }
            if (bs >= 0 && m - bs > 1) { rb[2 * nrb] = bs; rb[2 * nrb + 1] = m - bs; ++nrb; }
        }
        // split_blocks_by_frequency: within each block, runs of near-equal previous frequencies
        for (int k = 0; k < nrb; ++k) {
            const int st = rb[2 * k], sz = rb[2 * k + 1];
            if (sz < 2 || st >= npw) continue;
            const int en = min(npw, st + sz);
            if (en - st < 2) continue;
            int cs = st, cl = 1;
            for (int i = st + 1; i < en; ++i) {
                if (near(pw[i - 1], pw[i])) {
                    ++cl;
                } else {
                    if (cl > 1) { db[2 * ndb] = cs; db[2 * ndb + 1] = cl; ++ndb; }
                    cs = i;
                    cl = 1;
                }
            }
            if (cl > 1) { db[2 * ndb] = cs; db[2 * ndb + 1] = cl; ++ndb; }
        }
        cnt[0] = nrb;
        cnt[1] = ndb;
    }
    if (shortcut) {
        if (lane < m) perm[lane] = jmax;
        if (lane == 0) cnt[1] = 0;
        __syncthreads();
    } else {
        hungarian_wave(m, s.mag, perm, s.hu);  // begins with a barrier
    }
    // reassign a block of consecutive previous bands by frequency distance
    auto refine = [&](int st, int sz) {
        const int en = st + sz;
        if (en > m || en > npw) return;
        for (int k = lane; k < sz; k += TR_NT) s.ci[k] = perm[st + k];
        __syncthreads();
        for (int k = 0; k < sz; ++k)
            if (s.ci[k] >= ncw) return;
        for (int e = lane; e < sz * sz; e += TR_NT) {
            const int a = e % sz, b = e / sz;
            s.fw[e] = -fabs(pw[st + a] - cw[s.ci[b]]);
        }
        hungarian_wave(sz, s.fw, s.lp, s.hu);
        for (int li = lane; li < sz; li += TR_NT)
            if (s.lp[li] < sz) perm[st + li] = s.ci[s.lp[li]];
        __syncthreads();
    };
    const int ndb = cnt[1];
    for (int k = 0; k < ndb; ++k) refine(db[2 * k], db[2 * k + 1]);
    // detect_frequency_mismatch_blocks: runs of consecutive bands whose
    // assigned frequency moved by more than 5e-3 and 2 % relative
    if (lane == 0) {
        int nmb = 0, bs = -1;
        for (int i = 0; i < m; ++i) {
            const int c = perm[i];
            if (i >= npw || c >= ncw) {
                if (bs >= 0) {
                    if (i - bs > 1) { mb[2 * nmb] = bs; mb[2 * nmb + 1] = i - bs; ++nmb; }
                    bs = -1;
                }
                continue;
            }
            const double mis = fabs(pw[i] - cw[c]);
            const double sc = fmax(fmax(fabs(pw[i]), fabs(cw[c])), 1e-3);
            const bool bad = mis > 5e-3 && mis / sc > 0.02;
            if (bad) {
                if (bs < 0) bs = i;
            } else if (bs >= 0) {
                if (i - bs > 1) { mb[2 * nmb] = bs; mb[2 * nmb + 1] = i - bs; ++nmb; }
                bs = -1;
            }
        }
        if (bs >= 0 && m - bs > 1) { mb[2 * nmb] = bs; mb[2 * nmb + 1] = m - bs; ++nmb; }
        cnt[2] = nmb;
    }
    __syncthreads();
    const int nmb = cnt[2];
    for (int k = 0; k < nmb; ++k) refine(mb[2 * k], mb[2 * k + 1]);
}

// SCR: the workspace is the item's memory (wide subspaces), else dynamic LDS;
// a template parameter, so that LDS accesses compile to ds_* instructions
template <bool SCR>
__global__ void __launch_bounds__(TR_NT) transition_kernel(const TransWork *__restrict__ work, int KD, int NBW,
                                                           int M0, int MM, int NC) {
    const TransWork &w = work[blockIdx.x];
    extern __shared__ __attribute__((aligned(16))) double tsm[];
    char *const base = SCR ? w.scratch : reinterpret_cast<char *>(tsm);
    const TransLayout L(KD, NBW, M0, MM, NC);
    double2 *const Mp = reinterpret_cast<double2 *>(base + L.mp);
    double2 *const warm = reinterpret_cast<double2 *>(base + L.warm);
    double2 *const MC = reinterpret_cast<double2 *>(base + L.mc);
    double2 *const Gc = reinterpret_cast<double2 *>(base + L.gr);
    double2 *const Rr = reinterpret_cast<double2 *>(base + L.rr);
    int *const tk = reinterpret_cast<int *>(base + L.tk);
    double *const dk = reinterpret_cast<double *>(base + L.dk);
    double2 *const ent = warm;  // the block's entries replace the candidates
    double2 *const ment = MC;
    double2 *const tv = reinterpret_cast<double2 *>(base + L.tv);
    double2 *const tmv = reinterpret_cast<double2 *>(base + L.tmv);
    double2 *const col = reinterpret_cast<double2 *>(base + L.col);
    double2 *const A = reinterpret_cast<double2 *>(base + L.A);
    double2 *const V = reinterpret_cast<double2 *>(base + L.V);
    JacobiRot *const rot = reinterpret_cast<JacobiRot *>(base + L.rot);
    double *const red = reinterpret_cast<double *>(base + L.red);
    double2 *const Wt = reinterpret_cast<double2 *>(base + L.Wt);
    double2 *const Ut = reinterpret_cast<double2 *>(base + L.Ut);
    double2 *const Up = reinterpret_cast<double2 *>(base + L.Up);
    double *const sig = reinterpret_cast<double *>(base + L.sig);
    int *const ord = reinterpret_cast<int *>(base + L.ord);
    int *const perm = reinterpret_cast<int *>(base + L.perm);
    int *const tperm = reinterpret_cast<int *>(base + L.tperm);
    const TrackScratch ts{reinterpret_cast<double *>(base + L.mag), reinterpret_cast<double *>(base + L.fw),
                          reinterpret_cast<double *>(base + L.hu),  reinterpret_cast<int *>(base + L.ci),
                          reinterpret_cast<int *>(base + L.lp),     reinterpret_cast<int *>(base + L.blk),
                          reinterpret_cast<char *>(base + L.amb)};
    double2 *const cb = reinterpret_cast<double2 *>(base + L.cb);  // batched inner products

    const int tid = threadIdx.x;
    const int kd = w.kd, nE = w.nE, nP = w.nP, nR = w.nR;
    const int iP = nE, iR = nE + nP, iY = iR + nR, iW = iY + 1;
    const int mt = min(nE, nP);        // order of the tracking overlap
    const int nbp = min(nE, w.nb);     // eigenvectors carried into the history
    const int mp2 = min(nbp, nP);      // order of the prediction overlap
    const bool shape_ok = kd >= 1 && kd <= KD && kd <= w.ld && kd == iW + (nE == 0 ? w.nWc : 0) && mt <= MM &&
                          (w.m0 == 0 || max(nbp, w.nWc) + nR <= NC) &&
                          mp2 <= MM && w.m0 >= 0 && w.m0 <= M0 && nbp <= NBW && w.nWc <= NBW && nE <= KD &&
                          nE >= 0 && nP >= 0 && nR >= 0 && w.nprev >= 0;
    if (!shape_ok) {
        if (tid == 0) w.out[0] = -2;
        return;
    }

    // M (col-major, leading dimension KD) from the lower triangle of the
    // Gram's row-chunk partial sums (summed in chunk order), diagonal real
    for (int e = tid; e < kd * kd; e += TR_NT) {
        const int i = e % kd, j = e / kd;
        if (i < j) continue;
        const __attribute__((address_space(1))) double *g = gmem(w.gram) + 2 * ((size_t)i * w.ld + j);
        double gx = 0.0, gy = 0.0;
        const size_t cs = 2 * (size_t)w.ld * w.ld;
        int c = 0;
        for (; c + 8 <= w.nchunks; c += 8) {  // eight partials in flight, summed in chunk order
            double px[8], py[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                px[u] = g[(c + u) * cs];
                py[u] = g[(c + u) * cs + 1];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                gx += px[u];
                gy += py[u];
            }
        }
        for (; c < w.nchunks; ++c) {
            gx += g[c * cs];
            gy += g[c * cs + 1];
        }
        if (i == j) gy = 0.0;
        Mp[i + j * KD] = make_double2(gx, gy);
        Mp[j + i * KD] = make_double2(gx, -gy);
    }
    for (int i = tid; i < nE; i += TR_NT) perm[i] = i;
    __syncthreads();
    auto Mij = [&](int i, int j) __attribute__((always_inline)) -> double2 { return Mp[i + j * KD]; };
    // <a, b> = a^H b over the lanes' elements, reduced over the wave
    auto cdot = [&](const double2 *a, const double2 *b) __attribute__((always_inline)) -> double2 {
        double re = 0.0, im = 0.0;
        for (int i = tid; i < kd; i += TR_NT) {
            const double2 x = a[i], y = b[i];
            re += x.x * y.x + x.y * y.y;
            im += x.x * y.y - x.y * y.x;
        }
        wave_sum2(re, im);
        return make_double2(re, im);
    };
    // out = M v (v complete in memory: a barrier precedes every call); a row
    // per group of S adjacent lanes, each lane a segment of the row
    auto matvec = [&](const double2 *v, double2 *out) __attribute__((always_inline)) {
        const int S = lanes_per_item(kd, kd), R = TR_NT / S, L = (kd + S - 1) / S;
        const int sg = tid % S, j0 = sg * L, j1 = min(kd, j0 + L);
        for (int i0 = 0; i0 < kd; i0 += R) {
            const int i = i0 + tid / S;
            double re = 0.0, im = 0.0;
            if (i < kd) {
#pragma unroll 4
                for (int j = j0; j < j1; ++j) {
                    const double2 g = Mp[i + j * KD], x = v[j];
                    re = fma(g.x, x.x, re);
                    re = fma(-g.y, x.y, re);
                    im = fma(g.x, x.y, im);
                    im = fma(g.y, x.x, im);
                }
            }
            group_sum2(re, im, S);
            if (i < kd && sg == 0) out[i] = make_double2(re, im);
        }
        __syncthreads();  // out is read by other lanes
    };
    // cb[j] = a^H B_j for the nb vectors B_j = B + j KD (a and B complete in
    // memory): a vector per group of S adjacent lanes, each lane a segment
    auto dots = [&](const double2 *a, const double2 *B, int nb) __attribute__((always_inline)) {
        if (nb <= 0) return;
        const int S = lanes_per_item(nb, kd), R = TR_NT / S, L = (kd + S - 1) / S;
        const int sg = tid % S, q0 = sg * L, q1 = min(kd, q0 + L);
        for (int jb = 0; jb < nb; jb += R) {
            const int j = jb + tid / S;
            double re = 0.0, im = 0.0;
            if (j < nb) {
                const double2 *bj = B + (size_t)j * KD;
#pragma unroll 4
                for (int q = q0; q < q1; ++q) {
                    const double2 x = a[q], y = bj[q];
                    re += x.x * y.x + x.y * y.y;
                    im += x.x * y.y - x.y * y.x;
                }
            }
            group_sum2(re, im, S);
            if (j < nb && sg == 0) cb[j] = make_double2(re, im);
        }
        __syncthreads();  // cb is read by every lane
    };
    auto unit = [&](double2 *v, int idx) __attribute__((always_inline)) {
        for (int i = tid; i < kd; i += TR_NT) v[i] = make_double2(i == idx ? 1.0 : 0.0, 0.0);
    };

    // Gram-Schmidt in the B inner product of the nc candidates Cc (stride KD),
    // in order, as the right-looking Cholesky factorisation of their Gram
    // G = Cc^H M Cc: with the earlier entries projected out, candidate t has
    // squared B-norm G_tt, and projecting entry k out of a later candidate u
    // takes the coefficient R_ku = G_tu / d_k; G's trailing block then loses
    // conj(R_ku) R_ku'.  In exact arithmetic this is the modified Gram-Schmidt
    // of orthonormalize_fields and of initialize_block's project_out, with the
    // same norms and thresholds; each step is a few parallel passes over small
    // matrices instead of a chain of inner products.  A candidate whose norm
    // is not above tol is rejected (reject), counting as an attempt from
    // index nfree on (at most cap * 8), or kept unnormalised (!reject, the
    // 1e-14 rule of orthonormalize_fields).  MC holds M c_t for the candidates
    // (computed here for the first ndense).  Entry k goes to E + k KD and M e_k to
    // ME + k KD; E and ME may be Cc and MC themselves when nothing is
    // rejected.  Returns the number of entries (at most cap).
    auto orth = [&](double2 *Cc, int nc, int ndense, int cap, double tol, bool reject, int nfree, double2 *E,
                    double2 *ME) __attribute__((always_inline)) -> int {
        if (ndense > 0) {
            for (int e = tid; e < kd * ndense; e += TR_NT) {  // MC = M Cc for the leading dense candidates
                const int i = e % kd, t = e / kd;
                const double2 *v = Cc + (size_t)t * KD;
                double re = 0.0, im = 0.0;
                int j = 0;
                for (; j + 4 <= kd; j += 4) {
                    double2 g[4], x[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        g[u] = Mp[i + (j + u) * KD];
                        x[u] = v[j + u];
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        re = fma(g[u].x, x[u].x, re);
                        re = fma(-g[u].y, x[u].y, re);
                        im = fma(g[u].x, x[u].y, im);
                        im = fma(g[u].y, x[u].x, im);
                    }
                }
                for (; j < kd; ++j) {
                    const double2 g = Mp[i + j * KD], x = v[j];
                    re = fma(g.x, x.x, re);
                    re = fma(-g.y, x.y, re);
                    im = fma(g.x, x.y, im);
                    im = fma(g.y, x.x, im);
                }
                MC[(size_t)t * KD + i] = make_double2(re, im);
            }
            __syncthreads();
        }
        for (int e = tid; e < nc * nc; e += TR_NT) {  // G = Cc^H (M Cc), Hermitian
            const int t = e % nc, u = e / nc;
            if (t > u) continue;
            const double2 *a = Cc + (size_t)t * KD, *b = MC + (size_t)u * KD;
            double re = 0.0, im = 0.0;
            int q = 0;
            for (; q + 4 <= kd; q += 4) {
                double2 x[4], y[4];
#pragma unroll
                for (int w4 = 0; w4 < 4; ++w4) {
                    x[w4] = a[q + w4];
                    y[w4] = b[q + w4];
                }
#pragma unroll
                for (int w4 = 0; w4 < 4; ++w4) {
                    re += x[w4].x * y[w4].x + x[w4].y * y[w4].y;
                    im += x[w4].x * y[w4].y - x[w4].y * y[w4].x;
                }
            }
            for (; q < kd; ++q) {
                const double2 x = a[q], y = b[q];
                re += x.x * y.x + x.y * y.y;
                im += x.x * y.y - x.y * y.x;
            }
            Gc[t + u * nc] = make_double2(re, im);
            if (t != u) Gc[u + t * nc] = make_double2(re, -im);
        }
        __syncthreads();
        int n = 0, attempts = 0;
        for (int t = 0; t < nc && n < cap; ++t) {
            if (t >= nfree && attempts >= cap * 8) break;
            const double piv = Gc[t + t * nc].x;
            const double d = sqrt(fmax(piv, 0.0));
            double dd = d;
            if (!(d > tol)) {
                if (reject) {
                    if (t >= nfree) ++attempts;
                    continue;
                }
                dd = 1.0;
            }
            const double rd = 1.0 / dd;
            const double f = 2.0 - piv * rd * rd;  // 1 for a normalised entry
            // the coefficients R_nu = G_tu / d (kept for the entries) and the
            // trailing update, each lane forming the ones it needs
            double2 *rn = Rr + (size_t)n * NC;
            const int nt = nc - t - 1;
            for (int u = t + 1 + tid; u < nc; u += TR_NT) {
                const double2 g = Gc[t + u * nc];
                rn[u] = make_double2(g.x * rd, g.y * rd);
            }
            for (int e = tid; e < nt * nt; e += TR_NT) {
                const int u = t + 1 + e % nt, v = t + 1 + e / nt;
                const double2 gu = Gc[t + u * nc], gv = Gc[t + v * nc];
                const double fr = f * rd * rd;
                double2 &g = Gc[u + v * nc];
                g.x -= fr * (gu.x * gv.x + gu.y * gv.y);
                g.y -= fr * (gu.x * gv.y - gu.y * gv.x);
            }
            if (tid == 0) {
                tk[n] = t;
                dk[n] = rd;
            }
            __syncthreads();
            ++n;
        }
        // e_k = (c_{t_k} - sum_{l<k} R_{l t_k} e_l) / d_k and the same for M e_k,
        // each lane its own elements, both in one pass
        for (int k = 0; k < n; ++k) {
            const int t = tk[k];
            const double rd = dk[k];  // 1 / d_k
            for (int i = tid; i < kd; i += TR_NT) {
                double2 x = Cc[(size_t)t * KD + i], mx = MC[(size_t)t * KD + i];
                int l = 0;
                for (; l + 4 <= k; l += 4) {
                    double2 c[4], el[4], ml[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        c[u] = Rr[(size_t)(l + u) * NC + t];
                        el[u] = E[(size_t)(l + u) * KD + i];
                        ml[u] = ME[(size_t)(l + u) * KD + i];
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        x.x -= c[u].x * el[u].x - c[u].y * el[u].y;
                        x.y -= c[u].x * el[u].y + c[u].y * el[u].x;
                        mx.x -= c[u].x * ml[u].x - c[u].y * ml[u].y;
                        mx.y -= c[u].x * ml[u].y + c[u].y * ml[u].x;
                    }
                }
                for (; l < k; ++l) {
                    const double2 c = Rr[(size_t)l * NC + t], el = E[(size_t)l * KD + i], ml = ME[(size_t)l * KD + i];
                    x.x -= c.x * el.x - c.y * el.y;
                    x.y -= c.x * el.y + c.y * el.x;
                    mx.x -= c.x * ml.x - c.y * ml.y;
                    mx.y -= c.x * ml.y + c.y * ml.x;
                }
                E[(size_t)k * KD + i] = make_double2(x.x * rd, x.y * rd);
                ME[(size_t)k * KD + i] = make_double2(mx.x * rd, mx.y * rd);
            }
        }
        __syncthreads();
        return n;
    };

    // Polar factor U of the m x m overlap O(i, j) = M(ro + i, cmap[j]) from
    // the eigendecomposition O^H O = V S^2 V^H: U = O V S^-1 V^H, columns of
    // vanishing singular value completed by Gram-Schmidt on unit vectors
    // (band_tracking.rs polar_decomposition); returns the smallest singular
    // value, full = 0 when a column had to be completed.
    auto polar = [&](int m, int ro, const int *cmap, double2 *U, int &full) __attribute__((always_inline)) -> double {
        const int ne = m + (m & 1);
        auto O = [&](int i, int j) { return Mij(ro + i, cmap ? cmap[j] : j); };
        for (int e = tid; e < ne * ne; e += TR_NT) {
            const int i = e % ne, j = e / ne;
            double re = 0.0, im = 0.0;
            if (i < m && j < m)
#pragma unroll 4
                for (int k = 0; k < m; ++k) {
                    const double2 a = O(k, i), b = O(k, j);
                    re += a.x * b.x + a.y * b.y;
                    im += a.x * b.y - a.y * b.x;
                }
            A[e] = make_double2(re, im);
            V[e] = make_double2(i == j ? 1.0 : 0.0, 0.0);
        }
        __syncthreads();
        jacobi_lds<TR_NT>(A, V, ne, ne, rot, red);
        __syncthreads();
        // ascending order, ties by index; a NaN sorts last, so the ranks are a permutation
        auto key = [&](int i) {
            const double d = A[i + i * ne].x;
            return isnan(d) ? INFINITY : d;
        };
        for (int i = tid; i < m; i += TR_NT) {
            const double di = key(i);
            int r = 0;
            for (int j = 0; j < m; ++j) {
                const double dj = key(j);
                r += (dj < di || (dj == di && j < i)) ? 1 : 0;
            }
            ord[r] = i;
        }
        __syncthreads();
        for (int j = tid; j < m; j += TR_NT) sig[j] = sqrt(fmax(A[ord[j] + ord[j] * ne].x, 0.0));
        __syncthreads();
        double smin = INFINITY, smax = 0.0;
        for (int j = 0; j < m; ++j) {
            smin = fmin(smin, sig[j]);
            smax = fmax(smax, sig[j]);
        }
        const double thr = 1e-13 * fmax(smax, 1e-300);
        full = 1;
        for (int j = 0; j < m; ++j)
            if (!(sig[j] > thr)) full = 0;
        for (int e = tid; e < m * m; e += TR_NT) {
            const int i = e % m, j = e / m;
            if (!(sig[j] > thr)) continue;
            double re = 0.0, im = 0.0;
#pragma unroll 4
            for (int k = 0; k < m; ++k) {
                const double2 a = O(i, k), v = V[k + ord[j] * ne];
                re += a.x * v.x - a.y * v.y;
                im += a.x * v.y + a.y * v.x;
            }
            Wt[i + j * m] = make_double2(re / sig[j], im / sig[j]);
        }
        __syncthreads();
        if (!full && tid == 0) {
            for (int j = m - 1; j >= 0; --j) {
                if (sig[j] > thr) continue;
                for (int e = 0; e < m; ++e) {
                    for (int i = 0; i < m; ++i) col[i] = make_double2(i == e ? 1.0 : 0.0, 0.0);
                    for (int jj = j + 1; jj < m; ++jj) {
                        double cr = 0.0, cim = 0.0;
                        for (int i = 0; i < m; ++i) {
                            const double2 a = Wt[i + jj * m], b = col[i];
                            cr += a.x * b.x + a.y * b.y;
                            cim += a.x * b.y - a.y * b.x;
                        }
                        for (int i = 0; i < m; ++i) {
                            const double2 a = Wt[i + jj * m];
                            col[i].x -= cr * a.x - cim * a.y;
                            col[i].y -= cr * a.y + cim * a.x;
                        }
                    }
                    double nn = 0.0;
                    for (int i = 0; i < m; ++i) nn += col[i].x * col[i].x + col[i].y * col[i].y;
                    if (nn > 0.25) {
                        nn = sqrt(nn);
                        for (int i = 0; i < m; ++i) col[i] = make_double2(col[i].x / nn, col[i].y / nn);
                        break;
                    }
                }
                for (int i = 0; i < m; ++i) Wt[i + j * m] = col[i];
            }
        }
        __syncthreads();
        for (int e = tid; e < m * m; e += TR_NT) {
            const int i = e % m, j = e / m;
            double re = 0.0, im = 0.0;
#pragma unroll 4
            for (int k = 0; k < m; ++k) {
                const double2 a = Wt[i + k * m], v = V[j + ord[k] * ne];
                re += a.x * v.x + a.y * v.y;
                im += a.y * v.x - a.x * v.y;
            }
            U[i + j * m] = make_double2(re, im);
        }
        __syncthreads();
        return smin;
    };

    // ---- band tracking ----
    const bool track = (w.flags & TR_TRACK) && mt > 0;
    int full_t = 0;
    double smin_t = 0.0;
    if (track) {
        smin_t = polar(mt, iP, nullptr, Ut, full_t);
        track_assign(mt, Ut, w.prev_om, w.nprev, w.om, nE, tperm, ts);
        for (int i = tid; i < mt; i += TR_NT) perm[i] = tperm[i];
        __syncthreads();
    }

    if (w.m0 > 0) {
        // ---- warm vectors: the tracked eigenvectors or their prediction ----
        int nw = 0;
        if (nE > 0 && !(w.flags & TR_PREDICT)) {
            nw = nbp;  // the unit vectors e_perm[v], set with the start vectors below
        } else if (nE > 0) {
            // predict_with_rotation: O = prev^H B curr with curr the tracked E;
            // a full-rank tracking overlap's polar factor permuted by the
            // tracking is this overlap's factor
            const int m = mp2;
            double smin;
            if (track && full_t && mt == m) {
                for (int e = tid; e < m * m; e += TR_NT) {
                    const int i = e % m, j = e / m;
                    Up[i + j * m] = Ut[i + perm[j] * m];
                }
                smin = smin_t;
                __syncthreads();
            } else {
                int f;
                smin = polar(m, iP, perm, Up, f);
            }
            const bool extrap = (w.flags & TR_EXTRAP) && !(smin < 0.1);  // DEGENERACY_THRESHOLD
            const double alpha = extrap ? w.alpha : 0.0, ca = extrap ? 1.0 + alpha : 1.0;
            // pred_j = sum_i curr_i ca conj(U(j, i)) - alpha prev_j
            for (int j = 0; j < m; ++j)
                for (int i = tid; i < kd; i += TR_NT) warm[(size_t)j * KD + i] = make_double2(0.0, 0.0);
            __syncthreads();
            for (int e = tid; e < m * m; e += TR_NT) {
                const int i = e % m, j = e / m;
                const double2 u = Up[j + i * m];
                warm[(size_t)j * KD + perm[i]] = make_double2(u.x * ca, -u.y * ca);
            }
            if (extrap)
                for (int j = tid; j < m; j += TR_NT) warm[(size_t)j * KD + iP + j] = make_double2(-alpha, 0.0);
            nw = m;
        }
        // the learned start's vectors at a job's first k-point
        if (nE == 0)
            nw += w.nWc;  // the unit vectors e_{iW + j}, set with the start vectors below

        // ---- initialize_block: the warm vectors, then the start vectors
        // e_{iR + r} while the block is short, each B-orthonormalised against
        // the entries before it and rejected when its norm is not above 1e-12 ----
        // orthonormalize_fields of the predicted vectors followed by
        // initialize_block over them is one Gram-Schmidt pass over the
        // predicted vectors and then the start vectors: the first pass's
        // output is already B-orthonormal, so in exact arithmetic the second
        // keeps it and projects the start vectors against the same span.
        // M c for the candidates: computed for the predicted vectors, a
        // column of M for a unit vector
        const int nu0 = (nE > 0 && (w.flags & TR_PREDICT)) ? nw : 0;
        for (int e = tid; e < kd * (nw + nR - nu0); e += TR_NT) {
            const int i = e % kd, t = nu0 + e / kd;
            const int idx = t < nw ? (nE > 0 ? perm[t] : iW + t) : iR + (t - nw);
            warm[(size_t)t * KD + i] = make_double2(i == idx ? 1.0 : 0.0, 0.0);
            MC[(size_t)t * KD + i] = Mp[i + idx * KD];
        }
        __syncthreads();
        const int nent = orth(warm, nw + nR, nu0, w.m0, 1e-12, true, nw, ent, ment);
        const bool ran_out = nent < w.m0;
        if (w.flags & TR_GAMMA) {
            // add_vector(y0), then project each entry against it and renormalise
            __syncthreads();
            unit(tv, iY);
            __syncthreads();
            matvec(tv, tmv);
            const double yn = sqrt(fmax(cdot(tv, tmv).x, 0.0));
            if (yn > 1e-12 && yn >= 1e-10)
                for (int i = tid; i < kd; i += TR_NT) {
                    tv[i] = make_double2(tv[i].x / yn, tv[i].y / yn);
                    tmv[i] = make_double2(tmv[i].x / yn, tmv[i].y / yn);
                }
            // every entry against the same y0: the coefficients and the norms
            // are independent across entries, one per lane
            __syncthreads();
            dots(tmv, ent, nent);
            __syncthreads();
            for (int i = tid; i < kd; i += TR_NT) {
                const double2 y = tv[i], my = tmv[i];
#pragma unroll 4
                for (int k = 0; k < nent; ++k) {
                    const double2 c = cb[k];
                    double2 &e = ent[(size_t)k * KD + i], &me = ment[(size_t)k * KD + i];
                    e.x -= c.x * y.x - c.y * y.y;
                    e.y -= c.x * y.y + c.y * y.x;
                    me.x -= c.x * my.x - c.y * my.y;
                    me.y -= c.x * my.y + c.y * my.x;
                }
            }
            __syncthreads();
            for (int k = tid; k < nent; k += TR_NT) {
                const double2 *e = ent + (size_t)k * KD, *me = ment + (size_t)k * KD;
                double re = 0.0;
#pragma unroll 8
                for (int q = 0; q < kd; ++q) re += e[q].x * me[q].x + e[q].y * me[q].y;
                cb[k].x = sqrt(fmax(re, 0.0));
            }
            __syncthreads();
            for (int i = tid; i < kd; i += TR_NT)
                for (int k = 0; k < nent; ++k) {
                    const double nrm = cb[k].x;
                    if (nrm > 1e-12) {
                        double2 &e = ent[(size_t)k * KD + i];
                        e = make_double2(e.x / nrm, e.y / nrm);
                    }
                }
        }
        // the block's coefficients over K
        for (int e = 0; e < nent; ++e)
            for (int q = tid; q < kd; q += TR_NT) {
                const double2 x = ent[(size_t)e * KD + q];
                w.cx[2 * (q + (size_t)e * kd)] = x.x;
                w.cx[2 * (q + (size_t)e * kd) + 1] = x.y;
            }
        for (int e = nent; e < w.m0; ++e)  // a short block is an error the host reports; keep it finite
            for (int q = tid; q < 2 * kd; q += TR_NT) w.cx[2 * (size_t)e * kd + q] = 0.0;
        if (tid == 0) w.out[0] = ran_out ? -1 : nent;
    } else if (tid == 0) {
        w.out[0] = 0;
    }
    // history copy H_new_i = E_perm[i], and the permutation for the host's record
    for (int e = tid; e < nE * nE; e += TR_NT) {
        const int j = e % nE, i = e / nE;
        w.hist[2 * e] = j == perm[i] ? 1.0 : 0.0;
        w.hist[2 * e + 1] = 0.0;
    }
    for (int i = tid; i < nE; i += TR_NT) w.out[2 + i] = perm[i];
    if (tid == 0) w.out[1] = full_t;
}

}  // namespace

size_t transition_bytes(int KD, int NBW, int M0, int MM, int NC) {
    return TransLayout(KD, NBW, M0, MM, NC).total;
}

void launch_transition(const TransWork *dwork, int nwork, int KD, int NBW, int M0, int MM, int NC, bool scratch,
                       hipStream_t st) {
    if (nwork <= 0) return;
    if (MM > 32) throw std::runtime_error("transition: polar order above 32");
    // in LDS: up to 48 coefficient rows, 16 bands (the one-workgroup subspace path)
    static const size_t lds_max = transition_bytes(48, 16, 18, 16, 16 + 18 + 2);
    const size_t sh = scratch ? 0 : transition_bytes(KD, NBW, M0, MM, NC);
    if (sh > lds_max) throw std::runtime_error("transition: workspace exceeds the LDS allowance");
    static const bool attr_set = [] {
        BH_HIP_CHECK(hipFuncSetAttribute((const void *)transition_kernel<false>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_max));
        return true;
    }();
    (void)attr_set;
    if (scratch)
        hipLaunchKernelGGL(transition_kernel<true>, dim3(nwork), dim3(TR_NT), 0, st, dwork, KD, NBW, M0, MM, NC);
    else
        hipLaunchKernelGGL(transition_kernel<false>, dim3(nwork), dim3(TR_NT), sh, st, dwork, KD, NBW, M0, MM, NC);
    BH_HIP_CHECK(hipGetLastError());
}

}  // namespace bh
