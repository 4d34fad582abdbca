// Batched band-structure engine for one MI355X.
//
// A "slot" holds one band structure (bandstructure::run_core,
// drivers/bandstructure.rs:381-885) while it walks its k-path.  All slots
// advance together in steps; one step runs one LOBPCG iteration
// (eigensolver.rs:1684-2192) for every iterating slot with batched kernels,
// plus the k-point transitions (band tracking, subspace prediction, block
// initialisation) of slots that just converged, which run on the device in
// the same step as the new k-point's first iteration.  The host keeps the
// reference's per-instance control (convergence by relative eigenvalue
// change, soft locking with its degeneracy guard, the solve() exit rule) and
// the decisions that depend only on the k-path; every O(N) operation and the
// transition algebra run on the GPU.  One host synchronisation per step.
//
// Per-slot device layout (columns of N complex values, storage precision T),
// two copies b = 0, 1 of V and AV; the slot's parity selects the current one
// and the Ritz update writes the other:
//   V  [C + 1] : Z = [X (m) | W (w) | P (a)];  column C holds the Gamma mode y0
//   AV [C + 1] : A Z (A X and A W from the Ritz update, A P applied fresh)
// B Z is never stored: TE has B = I, and TM's B = eps is applied where a
// kernel reads B Z (residuals, Rayleigh quotients, Grams).
//   H  [2 nb]  : subspace history (previous / current converged eigenvectors)
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <exception>
#include <functional>
#include <mutex>
#include <thread>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <numeric>

#include "aux_kernels.h"
#include "blas.h"
#include "engine.h"
#include "fft_ops.h"
#include "transition.h"

namespace bh {

namespace {

// Pinned staging arena mirrored on the device: descriptors for one step are
// written on the host and uploaded with a single copy.
struct Staging {
    char *h = nullptr, *d = nullptr;
    size_t cap = 0, off = 0;
    void init(size_t bytes) {
        cap = bytes;
        BH_HIP_CHECK(hipHostMalloc((void **)&h, cap, hipHostMallocDefault));
        BH_HIP_CHECK(hipMalloc((void **)&d, cap));
    }
    void release() {
        if (h) (void)hipHostFree(h);
        if (d) (void)hipFree(d);
        h = d = nullptr;
    }
    template <typename X>
    X *alloc(size_t n, X **dev) {
        off = (off + 63) & ~size_t(63);
        if (off + n * sizeof(X) > cap) throw std::runtime_error("staging arena exhausted");
        X *hp = reinterpret_cast<X *>(h + off);
        *dev = reinterpret_cast<X *>(d + off);
        off += n * sizeof(X);
        return hp;
    }
    void upload(hipStream_t st) {
        if (off) BH_HIP_CHECK(hipMemcpyAsync(d, h, off, hipMemcpyHostToDevice, st));
    }
};

// Fixed pool of host threads for per-slot work that is independent across
// slots (k-point transitions, dielectric tables of new jobs).  The calling
// thread takes part; the first exception thrown by any item is rethrown.
//
// Batches are small (tens of items of ~10 us), so the pool is built for
// latency.  One parallel_for at a time.  A generation number and the next item index
// share one 64-bit word (gen << 32 | index), so a worker can claim an item
// (compare-and-swap) only for the generation it read; a worker that wakes late
// for a finished generation fails the claim and never touches a newer
// generation's items.  The caller works through items too and returns once
// every item has completed (not once every worker has woken up).  Workers
// spin briefly between batches, then sleep on a condition variable.
class WorkerPool {
  public:
    explicit WorkerPool(int nthreads) {
        for (int t = 0; t < nthreads; ++t) th_.emplace_back([this] { loop(); });
    }
    ~WorkerPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_.store(true);
        }
        cv_.notify_all();
        for (auto &t : th_) t.join();
    }
    void parallel_for(int n, const std::function<void(int)> &fn) {
        if (n <= 0) return;
        if (n == 1 || th_.empty()) {
            for (int i = 0; i < n; ++i) fn(i);
            return;
        }
        const uint64_t g = (word_.load(std::memory_order_relaxed) >> 32) + 1;
        Batch &b = batch_[g & 1];
        b.fn.store(&fn, std::memory_order_relaxed);
        b.n.store(n, std::memory_order_relaxed);
        b.done.store(0, std::memory_order_relaxed);
        err_ = nullptr;
        word_.store(g << 32, std::memory_order_release);  // publish generation g, index 0
        if (sleepers_.load(std::memory_order_acquire) > 0) {
            std::lock_guard<std::mutex> lk(mu_);
            cv_.notify_all();
        }
        work(g);
        while (b.done.load(std::memory_order_acquire) < n) std::this_thread::yield();
        if (err_) std::rethrow_exception(err_);
    }

  private:
    // fn and n are atomic only because a worker that read an old generation
    // may still compare against n while the slot is refilled; its claim then
    // fails, so it never calls the new fn
    struct Batch {
        std::atomic<const std::function<void(int)> *> fn{nullptr};
        std::atomic<int> n{0};
        std::atomic<int> done{0};
    };
    // claims and runs items of generation g until none is left
    void work(uint64_t g) {
        Batch &b = batch_[g & 1];
        uint64_t w = word_.load(std::memory_order_acquire);
        while ((w >> 32) == g) {
            const int i = (int)(w & 0xffffffffu);
            if (i >= b.n.load(std::memory_order_relaxed)) return;
            if (!word_.compare_exchange_weak(w, w + 1, std::memory_order_acq_rel, std::memory_order_acquire))
                continue;  // w reloaded
            try {
                (*b.fn.load(std::memory_order_relaxed))(i);
            } catch (...) {
                std::lock_guard<std::mutex> lk(mu_);
                if (!err_) err_ = std::current_exception();
            }
            b.done.fetch_add(1, std::memory_order_acq_rel);
            w = word_.load(std::memory_order_acquire);
        }
    }
    void loop() {
        uint64_t seen = 0;
        while (true) {
            uint64_t g = word_.load(std::memory_order_acquire) >> 32;
            for (int spin = 0; g == seen && spin < 20000 && !stop_.load(std::memory_order_relaxed); ++spin) {
                std::this_thread::yield();
                g = word_.load(std::memory_order_acquire) >> 32;
            }
            if (g == seen) {
                std::unique_lock<std::mutex> lk(mu_);
                sleepers_.fetch_add(1, std::memory_order_acq_rel);
                cv_.wait(lk, [&] { return stop_.load() || (word_.load(std::memory_order_acquire) >> 32) != seen; });
                sleepers_.fetch_sub(1, std::memory_order_acq_rel);
                if (stop_.load()) return;
                continue;
            }
            seen = g;
            work(g);
        }
    }
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_;
    std::atomic<uint64_t> word_{0};
    Batch batch_[2];
    std::atomic<int> sleepers_{0};
    std::atomic<bool> stop_{false};
    std::exception_ptr err_;
};

// A group of slots with its own stream, staging arena, Gram partials and
// readback buffers, so that groups can be planned on the host and run on the
// GPU concurrently.
struct Group {
    hipStream_t st{};
    hipEvent_t ev{};
    Staging stg;
    double *part = nullptr;  // Gram partial sums
    double *h_rel = nullptr;  // readback of the group's residual norms
    double *h_prec = nullptr; // readback of the first-iteration preconditioned residual norms
    double *rrws = nullptr;   // wide Rayleigh-Ritz workspace, one block per slot of the group
    double *h_pw = nullptr;   // readback of the group's plane-wave projections
    int *sweeps = nullptr;
    int s0 = 0, s1 = 0;
    bool inflight = false;
    bool has_iter = false;
    bool own_st = false;
    // event pairs around timed launches of the steps in flight
    std::vector<std::pair<hipEvent_t, hipEvent_t>> tev;
    size_t tev_used = 0;
    // adds the elapsed time of the completed pairs to *ms and recycles them
    void harvest(double *ms) {
        for (size_t i = 0; i < tev_used; ++i) {
            float e = 0.f;
            BH_HIP_CHECK(hipEventElapsedTime(&e, tev[i].first, tev[i].second));
            *ms += e;
        }
        tev_used = 0;
    }
};

// zero-filled device allocation, cleared on the given stream
template <typename T>
T *dalloc(size_t n, hipStream_t st) {
    T *p = nullptr;
    if (n == 0) n = 1;
    BH_HIP_CHECK(hipMalloc((void **)&p, n * sizeof(T)));
    BH_HIP_CHECK(hipMemsetAsync(p, 0, n * sizeof(T), st));
    return p;
}

// FREE: no job; ITER: solving at its k-point; DRAIN: the job's last solve has
// finished and its record waits for the step's results
enum SlotState { FREE = 0, ITER = 2, DRAIN = 3 };

// ints per slot in the Rayleigh-Ritz counter array (RRWork::cnt)
constexpr int CNT_STRIDE = 8;

struct Slot {
    int state = FREE;
    int job = -1;
    int k = 0;
    int nk = 0;
    // job constants
    int nb = 0, m0 = 0, pol = 0, precond = 0, exit_rule = 0, max_iter = 0;
    double tol = 0;
    bool prediction = true, extrapolation = true, tracking = true, reuse_gamma = true;
    DielectricTables diel;
    double y0val = 0;  // Gamma mode value 1 / ||1||_B
    // per-solve state
    int it = 0, m = 0, nW = 0;
    int par = 0;  // current copy of V / AV
    int last_p = 0;  // subspace width of the iteration in flight
    std::vector<char> soft;
    std::vector<char> soft_iter;  // soft-locked set when the current iteration's residuals are formed
    std::vector<double> lam, lam_prev;
    bool lam_prev_valid = false;
    bool gamma_defl = false;
    OpParams params{};
    // finished solve
    std::vector<double> res_eigs;
    int res_iters = 0;
    bool res_conv = false;
    std::vector<int> E_cols;  // column id in V (C == y0)
    // the k-point transition launched in this step: K = [E | P | R | y0 | Wc]
    // (E the finished eigenvectors, P the current history, R start vectors,
    // Wc learned-start vectors), the device builds the tracking permutation,
    // the history copy and the next block from K's Gram
    bool t_pending = false;    // planned, launched by the next execute()
    bool t_launched = false;   // launched; its status comes back with the step
    int kd = 0, nE = 0, nP = 0, nR = 0, t_nwc = 0, t_flags = 0, t_m0 = 0;
    int t_hist_new = -1, t_hist_old = -1;
    double t_alpha = 0;
    std::vector<double> t_om;  // finished frequencies (ascending) for the tracking
    bool fin_pending = false;  // finishing work of the solve (exit norms, plane-wave projections) this step
    // the finished k-point's record, completed when the step's results are back
    // (the tracking permutation orders its frequencies)
    struct Record {
        int kidx = 0, iters = 0, rel_iter = 0, rel_m = 0, closing_k = 0;
        bool conv = false, rel_fin = false, pw = false, track = false, save_first_gamma = false, closing = false;
        bool gamma = false;  // the finished k-point deflated y0
        std::vector<double> om, prec, rel;
        std::vector<int> rel_act;
        std::vector<char> rel_soft;  // soft-locked bands at the exit iteration
    };
    bool rec_pending = false;
    Record rec;
    // tracking / history
    bool have_prev = false;
    std::vector<double> prev_om;
    int hist_cur = -1;        // history buffer (0/1) holding the current eigenvectors
    double cur_k[2] = {0, 0}, prev_k[2] = {0, 0};
    bool hist_prev_valid = false;
    // first Gamma reuse
    bool have_first_gamma = false;
    std::vector<double> first_gamma;
    int first_gamma_iters = 0;
    std::vector<double> first_gamma_rel;
    int first_gamma_rel_iter = 0;
    std::vector<double> first_gamma_prec;
    // B-norms of the preconditioned residuals M^-1 r_i of the first iteration
    std::vector<double> prec_first;
    bool prec_pending = false;
    // relative residual B-norms: the latest log iteration's values (block
    // order) and the norms of an iteration in flight (read back next step)
    std::vector<double> rel_last;
    int rel_last_iter = 0;
    bool rel_pending = false;     // an iteration with norms is in flight
    bool rel_final = false;       // norms of the exit iteration are computed in the transition step
    int rel_pending_iter = 0;     // its 1-based iteration
    std::vector<int> rel_act;     // its active (non-soft-locked) bands
    std::vector<char> rel_soft;   // soft-locked bands at the exit iteration (rel_final)
    int rel_m = 0;                // block size of that iteration
    // learned start: warm vectors synthesised from the job's plane-wave
    // coefficients into the second history buffer, used by the first k-point
    int nWc = 0;
    // plane-wave projections of a finished solve's eigenvectors (readback pending)
    bool pw_pending = false;
    std::vector<double> first_gamma_pw;
};

bool is_gamma_frac(const std::array<double, 2> &k) { return k[0] * k[0] + k[1] * k[1] < 1e-8; }

// eigensolver.rs:251-259 should_log_iteration (iteration index 0-based): the
// iterations whose residual B-norms the reference computes
bool should_log_iteration(int iter) {
    const int n = iter + 1;
    return (n >= 1 && n <= 5) || n == 10 || n == 20 || n == 50 || n == 100 || (n > 100 && n % 50 == 0);
}

}  // namespace

template <typename T>
class Engine : public EngineBase {
  public:
    Engine(int nx_, int ny_, int nbmax_, const EngineOptions &opt) : nx(nx_), ny(ny_), nbmax(nbmax_), eopt(opt) {
        if (!grid_supported(nx, ny)) throw std::runtime_error("unsupported grid size");
        N = (long)nx * ny;
        m0max = nbmax + 2;
        C = std::max(2 * m0max, 3 * nbmax);
        // up to 16 bands (48 columns) the one-workgroup subspace kernels; up
        // to 32 bands (96 columns) the tiled Grams and the wide Rayleigh-Ritz
        if (C > 96) throw std::runtime_error("n_bands too large: at most 32 bands (96 subspace columns)");
        wide = C > GRAM_MAXC;
        CV = C + 1;
        S = std::max(1, opt.slots);
        NRMAX = m0max + 2;
        // leading dimension of the small matrices: the subspace width and the
        // transition basis [E | H | R | y0] (2 nb + m0 - nb + 3 columns)
        LD = wide ? ((std::max(C, 2 * nbmax + NRMAX + 1) + 15) / 16) * 16 : GRAM_MAXC;
        NCHMAX = (int)std::min<long>(wide ? 16 : 64, std::max<long>(1, N / 32));
        chunk_budget = wide ? 256 : 1024;
        if (opt.setup_stream) st = opt.setup_stream;
        else { BH_HIP_CHECK(hipStreamCreate(&st)); own_st = true; }
        V = dalloc<cplx<T>>((size_t)S * 2 * CV * N, st);
        AV = dalloc<cplx<T>>((size_t)S * 2 * CV * N, st);
        H = dalloc<cplx<T>>((size_t)S * 2 * nbmax * N, st);
        RND = dalloc<cplx<T>>((size_t)NRMAX * N, st);
        EPS = dalloc<T>((size_t)S * N, st);
        TEN = dalloc<T>((size_t)S * 4 * N, st);
        d_lam = dalloc<double>((size_t)S * LD, st);
        d_Y = dalloc<double>((size_t)S * LD * LD * 2, st);
        d_coef = dalloc<double>((size_t)S * LD * LD * 2, st);
        d_cnt = dalloc<int>(CNT_STRIDE * (size_t)S, st);
        d_rel = dalloc<double>((size_t)S * LD, st);
        d_prec = dalloc<double>((size_t)S * LD, st);
        if (opt.pw_half < 0 || opt.pw_half > 8) throw std::runtime_error("plane-wave half width must be 0..8");
        if (opt.pw_half > 0) d_pw = dalloc<double>((size_t)S * nbmax * pw_ng() * 2, st);
        // k-point transitions: block and history coefficients, status and
        // permutation, the exit iteration's residual norms
        d_tcx = dalloc<double>((size_t)S * LD * LD * 2, st);
        d_thist = dalloc<double>((size_t)S * nbmax * nbmax * 2, st);
        TOUT = 2 + LD;
        d_tout = dalloc<int>((size_t)S * TOUT, st);
        d_relf = dalloc<double>((size_t)S * LD, st);
        if (wide) {  // the transition kernel's workspace does not fit in LDS
            tbytes = (transition_bytes(LD, nbmax, m0max, nbmax, nbmax + m0max + 2) + 255) & ~size_t(255);
            d_tscr = dalloc<char>((size_t)S * tbytes, st);
        }
        BH_HIP_CHECK(hipHostMalloc((void **)&h_lam, sizeof(double) * S * LD, hipHostMallocDefault));
        BH_HIP_CHECK(hipHostMalloc((void **)&h_cnt, sizeof(int) * CNT_STRIDE * S, hipHostMallocDefault));
        BH_HIP_CHECK(hipHostMalloc((void **)&h_tout, sizeof(int) * TOUT * S, hipHostMallocDefault));
        BH_HIP_CHECK(hipHostMalloc((void **)&h_relf, sizeof(double) * LD * S, hipHostMallocDefault));
        int ng = opt.profile ? 1 : std::max(1, std::min(opt.groups, S));
        slots_per_group = (S + ng - 1) / ng;
        ng = (S + slots_per_group - 1) / slots_per_group;
        groups.resize(ng);
        for (int g = 0; g < ng; ++g) {
            Group &G = groups[g];
            G.s0 = g * slots_per_group;
            G.s1 = std::min(S, G.s0 + slots_per_group);
            const int gs = G.s1 - G.s0;
            if (g < 4 && opt.group_streams[g]) G.st = opt.group_streams[g];
            else { BH_HIP_CHECK(hipStreamCreate(&G.st)); G.own_st = true; }
            BH_HIP_CHECK(hipEventCreateWithFlags(&G.ev, hipEventDisableTiming));
            // per slot and step: dielectric tables of a new job plus descriptors / coefficients
            G.stg.init(((size_t)1 << 20) + (size_t)gs * (5 * (size_t)N * sizeof(T) + ((size_t)96 << 10)));
            // partial sums: two products (dual Gram) of up to 2048 chunk blocks
            G.part = dalloc<double>((size_t)2 * (chunk_budget + 2 * gs) * LD * LD * 2, st);
            if (wide) G.rrws = dalloc<double>((size_t)gs * rr_wide_bytes(C) / sizeof(double), st);
            BH_HIP_CHECK(hipHostMalloc((void **)&G.h_rel, sizeof(double) * gs * LD, hipHostMallocDefault));
            BH_HIP_CHECK(hipHostMalloc((void **)&G.h_prec, sizeof(double) * gs * LD, hipHostMallocDefault));
            if (d_pw)
                BH_HIP_CHECK(hipHostMalloc((void **)&G.h_pw, sizeof(double) * gs * nbmax * pw_ng() * 2,
                                           hipHostMallocDefault));
            G.sweeps = dalloc<int>(6, st);
        }
        // twiddle tables W_n^t = exp(-2 pi i t / n), computed in FP64 on the host
        std::vector<cplx<T>> tx(nx), ty(ny);
        for (int t = 0; t < nx; ++t) {
            double a = -2.0 * M_PI * t / nx;
            tx[t] = {(T)std::cos(a), (T)std::sin(a)};
        }
        for (int t = 0; t < ny; ++t) {
            double a = -2.0 * M_PI * t / ny;
            ty[t] = {(T)std::cos(a), (T)std::sin(a)};
        }
        twx = dalloc<cplx<T>>(nx, st);
        twy = dalloc<cplx<T>>(ny, st);
        BH_HIP_CHECK(hipMemcpyAsync(twx, tx.data(), sizeof(cplx<T>) * nx, hipMemcpyHostToDevice, st));
        BH_HIP_CHECK(hipMemcpyAsync(twy, ty.data(), sizeof(cplx<T>) * ny, hipMemcpyHostToDevice, st));
        {
            const std::vector<int> rx = digit_reversal(nx), ry = digit_reversal(ny);
            revx = dalloc<int>(nx, st);
            revy = dalloc<int>(ny, st);
            BH_HIP_CHECK(hipMemcpyAsync(revx, rx.data(), sizeof(int) * nx, hipMemcpyHostToDevice, st));
            BH_HIP_CHECK(hipMemcpyAsync(revy, ry.data(), sizeof(int) * ny, hipMemcpyHostToDevice, st));
            BH_HIP_CHECK(hipStreamSynchronize(st));  // the host tables go out of scope
        }
        big = grid_multipass(nx, ny);
        if (big) {  // an operator item applies to at most m0 columns of a slot
            TMPC = m0max;
            TMP = dalloc<cplx<T>>((size_t)S * TMPC * N, st);
        }
        launch_random<T>(RND, NRMAX, N, 1.0, st);
        BH_HIP_CHECK(hipStreamSynchronize(st));
        slots.resize(S);
        int hw = (int)std::thread::hardware_concurrency();
        int nt = opt.host_threads > 0 ? opt.host_threads : std::min(std::max(hw, 1), 16);
        pool.reset(new WorkerPool(std::max(nt - 1, 0)));
    }

    ~Engine() override {
        (void)hipStreamSynchronize(st);
        for (auto &G : groups) {
            (void)hipStreamSynchronize(G.st);
            (void)hipFree(G.part);
            if (G.rrws) (void)hipFree(G.rrws);
            (void)hipFree(G.sweeps);
            (void)hipHostFree(G.h_rel);
            (void)hipHostFree(G.h_prec);
            if (G.h_pw) (void)hipHostFree(G.h_pw);
            G.stg.release();
            for (auto &e : G.tev) {
                (void)hipEventDestroy(e.first);
                (void)hipEventDestroy(e.second);
            }
            (void)hipEventDestroy(G.ev);
            if (G.own_st) (void)hipStreamDestroy(G.st);
        }
        for (void *p : std::initializer_list<void *>{V, AV, H, RND, EPS, TEN, d_lam, d_Y, d_coef, d_cnt, d_rel, d_prec, d_pw, twx, twy, revx, revy, TMP,
                                                     d_tcx, d_thist, d_tout, d_relf, d_tscr})
            if (p) (void)hipFree(p);
        (void)hipHostFree(h_lam);
        (void)hipHostFree(h_cnt);
        (void)hipHostFree(h_tout);
        (void)hipHostFree(h_relf);
        if (own_st) (void)hipStreamDestroy(st);
    }

    void run(const std::vector<JobSpec> &jobs, std::vector<JobResult> &out, RunStats &stats) override;

  private:
    int nx, ny;
    long N;
    int nbmax, m0max, C, CV, S, LD, NRMAX, NCHMAX;
    bool wide = false;       // more than 48 subspace columns
    int chunk_budget = 1024;  // Gram row chunks per launch (partial blocks)
    EngineOptions eopt;
    hipStream_t st{};
    bool own_st = false;
    std::unique_ptr<WorkerPool> pool;
    cplx<T> *V{}, *AV{}, *H{}, *RND{};
    T *EPS{}, *TEN{};
    double *d_lam{}, *d_Y{}, *d_coef{};
    int *d_cnt{};
    double *d_rel{};  // relative residual B-norms, LD per slot
    double *d_prec{}; // first-iteration preconditioned residual B-norms, LD per slot
    double *d_pw{};   // plane-wave projections of finished solves, nbmax x (2h+1)^2 per slot
    double *h_lam{};
    int *h_cnt{};
    // k-point transitions (see Slot::t_pending)
    double *d_tcx{};    // LD x LD complex per slot: the new block over K
    double *d_thist{};  // nbmax x nbmax complex per slot: the history copy
    int *d_tout{}, *h_tout{};  // TOUT ints per slot: status, full rank, permutation
    int TOUT = 0;
    double *d_relf{}, *h_relf{};  // LD per slot: relative residual norms of a solve's exit iteration
    char *d_tscr{};     // wide subspaces: tbytes of transition workspace per slot
    size_t tbytes = 0;
    cplx<T> *twx{}, *twy{};
    int *revx{}, *revy{};     // digit-reversal tables (multi-pass operators)
    bool big = false;         // grid on the multi-pass operator path
    int TMPC = 0;             // second-field columns per slot for multi-pass TE / TP
    cplx<T> *TMP{};
    std::vector<Group> groups;
    int slots_per_group = 1;
    Group &group_of(int s) { return groups[s / slots_per_group]; }
    std::vector<Slot> slots;
    const std::vector<JobSpec> *jobs_{};
    std::vector<JobResult> *out_{};
    RunStats *stats_{};

    // column addressing
    // copy b of slot s; vcol/avcol address the current copy
    cplx<T> *vbuf(int s, int b, int j) { return V + (((size_t)s * 2 + b) * CV + j) * N; }
    cplx<T> *avbuf(int s, int b, int j) { return AV + (((size_t)s * 2 + b) * CV + j) * N; }
    cplx<T> *vcol(int s, int j) { return vbuf(s, slots[s].par, j); }
    cplx<T> *avcol(int s, int j) { return avbuf(s, slots[s].par, j); }
    // B weight of slot s: eps for TM, none (identity) for TE
    const T *bweight(int s) { return slots[s].pol == 0 ? eps(s) : nullptr; }
    cplx<T> *hcol(int s, int b, int j) { return H + ((size_t)s * 2 * nbmax + (size_t)b * nbmax + j) * N; }
    cplx<T> *rcol(int j) { return RND + (size_t)j * N; }
    T *eps(int s) { return EPS + (size_t)s * N; }
    cplx<T> *tmpcol(int s) { return TMP ? TMP + (size_t)s * TMPC * N : nullptr; }
    T *ten(int s) { return TEN + (size_t)s * 4 * N; }
    double *coef(int s) { return d_coef + (size_t)s * LD * LD * 2; }
    // algorithmic work of one Theta column: field in + out (+ the 2x2 inverse-eps
    // tensor for TE); FFT FLOPs 5 N log2 N per transform (2 for TM, 6 for TE)
    double theta_bytes_per_col(int pol) const {
        return (double)N * (2.0 * sizeof(cplx<T>) + (pol == 1 ? 4.0 * sizeof(T) : 0.0));
    }
    double theta_flops_per_col(int pol, double n_log) const {
        return pol == 1 ? 30.0 * n_log + 40.0 * N : 10.0 * n_log + 2.0 * N;
    }
    // FLOPs of one Rayleigh-Ritz problem of order n (complex multiply-add = 8):
    // Cholesky n^3/6, triangular inverse n^3/6, G2 T and T^H X n^3/2 each,
    // cyclic Jacobi 4.5 n^3 per sweep (n^2/2 rotations of two rows and two
    // columns of A and two columns of V), back-transform T Y n^2 k / 2
    // the Rayleigh-Ritz solve as run: Cholesky and the triangular inverse
    // (n^3/6 complex multiply-adds each), the projection (n^3), the Jacobi
    // sweeps counted (4.5 n^3 each) and the back-transform (n^2 k / 2)
    static double rr_flops(int n, int k, int sweeps) {
        const double n3 = (double)n * n * n;
        return 8.0 * (n3 / 3.0 + n3 + 4.5 * n3 * sweeps + 0.5 * (double)n * n * k);
    }
    // the same problem counted independently of the method: Cholesky, inverse
    // and projection (4 n^3 / 3 complex multiply-adds), a Hermitian
    // eigendecomposition with vectors at 30 n^3 real FLOPs, the back-transform
    static double rr_flops_algorithmic(int n, int k) {
        const double n3 = (double)n * n * n;
        return 8.0 * (4.0 * n3 / 3.0) + 30.0 * n3 + 4.0 * (double)n * n * k;
    }

    // --- per-step plans -----------------------------------------------------
    struct GramItem { int slot; std::vector<const cplx<T> *> a, b; bool weighted; int na, nb; const int *dev_na = nullptr, *dev_nb = nullptr; };
    struct StepPlan {
        std::vector<int> fill_y0;            // new jobs: set the Gamma mode column
        std::vector<int> table_upload;       // new jobs: upload eps / tensor tables
        std::vector<int> fin;                // finished solves: exit-iteration norms, plane-wave projections
        std::vector<int> trans;              // k-point transitions (Gram, device transition, new block)
        std::vector<int> iter;               // slots iterating this step
    };

    void start_job(int s, int job);
    void begin_transition(int s);
    void begin_block(int s, bool first_k);
    void complete_step_results(int s);
    void absorb(int s);
    void decide(int s);
    void finish_solve(int s, bool conv, int iters);
    void record_k(int s, int kidx, const std::vector<double> &om, int iters, bool conv, const std::vector<double> &prec,
                  const std::vector<double> &rel, int rel_iter, const double *pw = nullptr);
    // plane waves per projected vector, (2h + 1)^2
    int pw_ng() const { return (2 * eopt.pw_half + 1) * (2 * eopt.pw_half + 1); }
    double *pwbuf(int s) { return d_pw + (size_t)s * nbmax * pw_ng() * 2; }
    void take_norms(int s);
    void set_kpoint(int s);
    void execute(StepPlan &plan, Group &G);
};

template <typename T>
void Engine<T>::record_k(int s, int kidx, const std::vector<double> &om, int iters, bool conv,
                         const std::vector<double> &prec,
                         const std::vector<double> &rel, int rel_iter, const double *pw) {
    Slot &sl = slots[s];
    JobResult &r = (*out_)[sl.job];
    if (pw && !r.pw.empty()) {
        const size_t blk = (size_t)sl.nb * pw_ng() * 2;
        std::memcpy(r.pw.data() + (size_t)kidx * blk, pw, sizeof(double) * blk);
    }
    for (int b = 0; b < sl.nb; ++b) {
        r.omegas[(size_t)kidx * sl.nb + b] = b < (int)om.size() ? om[b] : NAN;
        r.rel_residuals[(size_t)kidx * sl.nb + b] = b < (int)rel.size() ? rel[b] : NAN;
        r.prec_first[(size_t)kidx * sl.nb + b] = b < (int)prec.size() ? prec[b] : NAN;
    }
    r.iters[kidx] = iters;
    r.converged[kidx] = conv ? 1 : 0;
    r.rel_iters[kidx] = rel_iter;
}

// The norms of the iteration in flight, from this group's readback.
template <typename T>
void Engine<T>::take_norms(int s) {
    Slot &sl = slots[s];
    if (!sl.rel_pending) return;
    const double *h = group_of(s).h_rel + (size_t)(s - group_of(s).s0) * LD;
    sl.rel_last.assign(std::min(sl.rel_m, sl.nb), 0.0);  // soft-locked bands: zero residual
    for (int b : sl.rel_act)
        if (b < (int)sl.rel_last.size()) sl.rel_last[b] = h[b];
    sl.rel_last_iter = sl.rel_pending_iter;
    sl.rel_pending = false;
    sl.rel_final = false;
}

template <typename T>
void Engine<T>::set_kpoint(int s) {
    Slot &sl = slots[s];
    const JobSpec &J = (*jobs_)[sl.job];
    const auto &kf = J.kpath[sl.k];
    OpParams &P = sl.params;
    P.b1[0] = sl.diel.b1[0]; P.b1[1] = sl.diel.b1[1];
    P.b2[0] = sl.diel.b2[0]; P.b2[1] = sl.diel.b2[1];
    // lattice.rs:274-280 fractional_to_cartesian
    P.bloch[0] = kf[0] * sl.diel.b1[0] + kf[1] * sl.diel.b2[0];
    P.bloch[1] = kf[0] * sl.diel.b1[1] + kf[1] * sl.diel.b2[1];
    P.pol = sl.pol;
    P.gamma = (std::fabs(P.bloch[0]) < 1e-12 && std::fabs(P.bloch[1]) < 1e-12) ? 1 : 0;
    P.precond = sl.precond;
    P.shift = 0.5 * spectral_smin(nx, ny, P.b1, P.b2, P.bloch);  // SHIFT_SMIN_FRACTION
    if (sl.pol == 0) {
        P.eps_eff = sl.diel.eps_mean;
        P.mass_floor = (std::isfinite(P.eps_eff) && P.eps_eff > 0) ? P.eps_eff * 1e-2 : 0.0;
    } else {
        P.eps_eff = sl.diel.eps_harm;
        P.mass_floor = 0.0;
    }
    sl.gamma_defl = P.gamma != 0;
}

template <typename T>
void Engine<T>::start_job(int s, int job) {
    Slot &sl = slots[s];
    const JobSpec &J = (*jobs_)[job];
    sl = Slot();
    sl.job = job;
    sl.k = 0;
    sl.nk = (int)J.kpath.size();
    sl.nb = J.n_bands;
    sl.m0 = std::max(J.block_size == 0 ? J.n_bands + 2 : J.block_size, std::max(J.n_bands, 1));
    if (sl.nb > nbmax || sl.m0 > m0max) throw std::runtime_error("job n_bands exceeds engine capacity");
    sl.pol = J.pol;
    sl.precond = J.precond < 0 ? (J.pol == 0 ? 1 : 2) : J.precond;  // diagnostics.rs:158-166
    sl.exit_rule = J.exit_rule;
    sl.max_iter = J.max_iter;
    sl.tol = J.tol;
    sl.prediction = J.prediction != 0;
    sl.extrapolation = J.extrapolation != 0;
    sl.tracking = J.tracking != 0;
    sl.reuse_gamma = J.reuse_gamma != 0;
    sl.diel = build_dielectric(J.geom, nx, ny, J.smoothing != 0);
    if (J.warm_count > 0) {
        if (J.pw_half < 0 || J.pw_half > 8) throw std::runtime_error("warm plane-wave half width must be 0..8");
        sl.nWc = std::min(J.warm_count, nbmax);
    }
    // create_gamma_mode (initialization.rs:166-224): u0 = 1, y0 = u0 / ||u0||_B
    double nsq = sl.pol == 0 ? sl.diel.eps_sum : (double)N;
    sl.y0val = 1.0 / (nsq > 0 ? std::sqrt(nsq) : 1.0);
    set_kpoint(s);
    sl.E_cols.clear();
}

template <typename T>
void Engine<T>::finish_solve(int s, bool conv, int iters) {
    Slot &sl = slots[s];
    // collect_all_eigenvalues / all_eigenvectors (eigensolver.rs:2362-2437)
    std::vector<std::pair<double, int>> pr;
    if (sl.gamma_defl) pr.push_back({0.0, C});
    for (int i = 0; i < (int)sl.lam.size() && i < sl.m; ++i) pr.push_back({sl.lam[i], i});
    std::stable_sort(pr.begin(), pr.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
    sl.res_eigs.clear();
    sl.E_cols.clear();
    for (int i = 0; i < (int)pr.size() && i < sl.nb; ++i) {
        sl.res_eigs.push_back(pr[i].first);
        sl.E_cols.push_back(pr[i].second);
    }
    sl.res_iters = iters;
    sl.res_conv = conv;
    stats_->iterations += iters;
    stats_->solves += 1;
    begin_transition(s);
}

// Results of the previous step's Rayleigh-Ritz for an iterating slot.
template <typename T>
void Engine<T>::absorb(int s) {
    Slot &sl = slots[s];
    if (sl.it == 0) return;
    take_norms(s);
    if (sl.prec_pending) {
        const double *h = group_of(s).h_prec + (size_t)(s - group_of(s).s0) * LD;
        sl.prec_first.assign(h, h + std::min(sl.m, sl.nb));
        sl.prec_pending = false;
    }
    const int *cnt = h_cnt + CNT_STRIDE * (size_t)s;
    int cx = cnt[0], cw = cnt[1];
    if (eopt.timed_kernel == TIMED_RR) {
        stats_->timed_flops += rr_flops_algorithmic(sl.last_p, cx + cw);
        stats_->timed_flops_executed += rr_flops(sl.last_p, cx + cw, cnt[3]);
    }
    stats_->rr_problems += 1;
    stats_->svqb_fallbacks += cnt[4];
    if (cnt[5] < sl.last_p) stats_->rank_drops += 1;
    sl.lam_prev = sl.lam;
    sl.lam_prev_valid = sl.it >= 2;
    if (cx > 0) {
        sl.lam.assign(h_lam + (size_t)s * LD, h_lam + (size_t)s * LD + cx);
        sl.m = cx;
    }
    sl.nW = cw;
}

// Convergence and soft-locking decisions (eigensolver.rs:1750-1864).
template <typename T>
void Engine<T>::decide(int s) {
    Slot &sl = slots[s];
    const int it = sl.it;
    // an exit at a log iteration: the reference computes that iteration's
    // residual norms before its exit test, from the soft-locked set before any
    // new locks; they are computed in the transition step
    auto exit_norms = [&](const std::vector<char> &soft0) {
        if (!should_log_iteration(it)) return;
        sl.rel_final = sl.rel_pending = true;
        sl.rel_pending_iter = it + 1;
        sl.rel_m = sl.m;
        sl.rel_soft = soft0;
        sl.rel_act.clear();
        for (int i = 0; i < sl.m; ++i)
            if (!(i < (int)soft0.size() && soft0[i])) sl.rel_act.push_back(i);
    };
    if (it >= sl.max_iter) {
        int nsoft = 0;
        for (char c : sl.soft) nsoft += c ? 1 : 0;
        finish_solve(s, (sl.gamma_defl ? 1 : 0) + nsoft >= sl.nb, sl.max_iter);
        return;
    }
    const int nbc = std::min(sl.nb, sl.m0);
    std::vector<double> ch;
    int n_conv = 0;
    if (it >= 2) {
        size_t kk = std::min(sl.lam.size(), sl.lam_prev.size());
        for (size_t i = 0; i < kk; ++i)
            ch.push_back(std::fabs(sl.lam[i] - sl.lam_prev[i]) / std::max(std::fabs(sl.lam_prev[i]), 1e-10));
        size_t ncheck = std::min((size_t)sl.m, ch.size());
        for (size_t i = 0; i < ncheck && (int)i < nbc; ++i) n_conv += ch[i] < sl.tol ? 1 : 0;
    }
    int n_soft = 0;
    for (char c : sl.soft) n_soft += c ? 1 : 0;
    const int n_hard = sl.gamma_defl ? 1 : 0;
    const int total = sl.exit_rule == 0 ? n_hard + n_soft + n_conv : n_soft;
    if (total >= sl.nb) {
        exit_norms(sl.soft);
        finish_solve(s, true, it + 1);
        return;
    }
    const std::vector<char> soft_before = sl.soft;
    // the reference forms this iteration's residuals (and deflates them)
    // before it soft-locks new bands; only P drops the newly locked ones
    sl.soft_iter = soft_before;
    if (it >= 2) {
        // soft_lock_bands with the 1e-4 degeneracy guard (eigensolver.rs:889-969)
        std::vector<int> want;
        size_t ncheck = std::min((size_t)sl.m, ch.size());
        for (size_t i = 0; i < ncheck; ++i)
            if (ch[i] < sl.tol) want.push_back((int)i);
        if (!want.empty()) {
            std::vector<char> lockset(sl.soft.size() + sl.lam.size() + 1, 0);
            for (int b : want) lockset[b] = 1;
            std::vector<int> safe;
            for (int b : want) {
                if (b >= (int)sl.lam.size() || b >= (int)sl.soft.size() || sl.soft[b]) continue;
                double li = sl.lam[b];
                bool ok = true;
                for (int j = 0; j < (int)sl.lam.size(); ++j) {
                    if (j == b || j >= (int)sl.soft.size() || sl.soft[j]) continue;
                    double mx = std::max(std::max(std::fabs(li), std::fabs(sl.lam[j])), 1e-10);
                    if (std::fabs(li - sl.lam[j]) / mx < 1e-4 && !lockset[j]) { ok = false; break; }
                }
                if (ok) safe.push_back(b);
            }
            int newly = 0;
            for (int b : safe)
                if (!sl.soft[b]) { sl.soft[b] = 1; ++newly; }
            n_soft += newly;
            if (newly > 0 && n_soft >= sl.nb) {
                exit_norms(soft_before);
                finish_solve(s, true, it + 1);
                return;
            }
        }
    }
    // continue iterating: the plan for this slot is built in execute()
}

// A solve has finished (finish_solve): plan the move to the next k-point.
// The reference's end-of-solve work (bandstructure.rs:755-806 tracking and
// recording, subspace_prediction.rs:248-380 history and prediction,
// initialization.rs:248-367 the next block) runs on the device in this step
// (transition_kernel over the Gram of K = [E | P | R | y0]); the host keeps
// the decisions that depend only on the k-path and the solve's own results,
// and records the k-point once the tracking permutation is back.
template <typename T>
void Engine<T>::begin_transition(int s) {
    Slot &sl = slots[s];
    const JobSpec &J = (*jobs_)[sl.job];
    const int nE = (int)sl.E_cols.size();
    const int nP = sl.hist_cur >= 0 ? sl.nb : 0;
    sl.nWc = 0;  // the learned start serves the first k-point only
    // the record of the finished k-point
    Slot::Record &R = sl.rec;
    R = Slot::Record();
    R.kidx = sl.k;
    R.om.resize(sl.res_eigs.size());
    for (size_t i = 0; i < R.om.size(); ++i) R.om[i] = sl.res_eigs[i] > 0 ? std::sqrt(sl.res_eigs[i]) : 0.0;
    R.iters = sl.res_iters;
    R.conv = sl.res_conv;
    R.prec = sl.prec_first;
    R.rel_fin = sl.rel_final;  // the exit iteration's norms are computed in this step
    R.gamma = sl.gamma_defl;
    if (R.rel_fin) {
        R.rel_act = sl.rel_act;
        R.rel_soft = sl.rel_soft;
        R.rel_m = sl.rel_m;
        R.rel_iter = sl.rel_pending_iter;
    } else {
        R.rel = sl.rel_last;
        R.rel_iter = sl.rel_last_iter;
    }
    R.pw = eopt.pw_half > 0 && nE > 0;
    R.track = sl.tracking && sl.have_prev && nP > 0 && nE > 0;
    R.save_first_gamma = sl.reuse_gamma && sl.k == 0 && is_gamma_frac(J.kpath[0]);
    int knext = sl.k + 1;
    // closing Gamma reuse (bandstructure.rs:580-601)
    const bool reuse = sl.reuse_gamma && sl.nk > 1 && is_gamma_frac(J.kpath[0]) && is_gamma_frac(J.kpath[sl.nk - 1]);
    if (knext == sl.nk - 1 && reuse && (sl.have_first_gamma || R.save_first_gamma)) {
        R.closing = true;
        R.closing_k = knext;
        knext = sl.nk;
    }
    const bool next_solve = knext < sl.nk;
    sl.rec_pending = true;
    sl.fin_pending = R.rel_fin || R.pw;
    sl.rel_pending = sl.rel_final = false;
    // K = [E | P | R | y0]
    sl.nE = nE;
    sl.nP = nP;
    sl.nR = std::min(nE > 0 ? (sl.m0 - std::min(sl.nb, sl.m0) + 2) : sl.m0 + 2, NRMAX);
    sl.kd = nE + nP + sl.nR + 1;
    if (sl.kd > LD) throw std::runtime_error("transition basis too wide");
    sl.t_nwc = 0;
    sl.t_flags = R.track ? TR_TRACK : 0;
    sl.t_om = R.om;
    sl.t_alpha = 0.0;
    sl.t_hist_old = sl.hist_cur;
    sl.t_hist_new = -1;
    if (!next_solve) {
        // the job ends: the device is needed only for the tracking permutation
        sl.state = DRAIN;
        sl.t_m0 = 0;
        sl.t_pending = R.track;
        if (!sl.t_pending && !sl.fin_pending) complete_step_results(s);
        return;
    }
    // history update (subspace_prediction.rs:248-262)
    const bool had_cur = sl.hist_cur >= 0;
    if (nE > 0) {
        sl.t_hist_new = had_cur ? 1 - sl.hist_cur : 0;
        if (had_cur) { sl.prev_k[0] = sl.cur_k[0]; sl.prev_k[1] = sl.cur_k[1]; }
        sl.hist_prev_valid = had_cur;
        const auto &kc = J.kpath[sl.k];
        sl.cur_k[0] = kc[0];
        sl.cur_k[1] = kc[1];
        if (sl.prediction && had_cur) {
            sl.t_flags |= TR_PREDICT;
            // extrapolation unless the path turns a corner (CORNER_DOT_THRESHOLD);
            // the kernel also drops it when the overlap is degenerate
            const auto &kn = J.kpath[knext];
            const double dkp[2] = {sl.cur_k[0] - sl.prev_k[0], sl.cur_k[1] - sl.prev_k[1]};
            const double dkn[2] = {kn[0] - sl.cur_k[0], kn[1] - sl.cur_k[1]};
            const bool corner = dkp[0] * dkn[0] + dkp[1] * dkn[1] < 0.0;
            if (sl.extrapolation && !corner) {
                sl.t_flags |= TR_EXTRAP;
                const double np_ = std::sqrt(dkp[0] * dkp[0] + dkp[1] * dkp[1]);
                const double nn_ = std::sqrt(dkn[0] * dkn[0] + dkn[1] * dkn[1]);
                sl.t_alpha = np_ > 1e-12 ? nn_ / np_ : 1.0;
            }
        }
        sl.hist_cur = sl.t_hist_new;
    }
    sl.k = knext;
    set_kpoint(s);
    begin_block(s, false);
}

// The start block of a solve (initialize_block, initialization.rs:248-367)
// built on the device this step: a new job's first k-point (K = [R | y0 | Wc])
// or the next k-point of a transition planned above.  The slot iterates in
// the same step.
template <typename T>
void Engine<T>::begin_block(int s, bool first_k) {
    Slot &sl = slots[s];
    if (first_k) {
        sl.nE = sl.nP = 0;
        sl.nR = std::min(sl.m0 + 2, NRMAX);
        sl.t_nwc = sl.nWc;
        sl.kd = sl.nR + 1 + sl.t_nwc;
        if (sl.kd > LD) throw std::runtime_error("transition basis too wide");
        sl.t_flags = 0;
        sl.t_om.clear();
        sl.t_alpha = 0.0;
        sl.t_hist_old = sl.t_hist_new = -1;
    }
    if (sl.gamma_defl) sl.t_flags |= TR_GAMMA;
    sl.t_m0 = sl.m0;
    sl.t_pending = true;
    sl.it = 0;
    sl.m = sl.m0;
    sl.rel_last.clear();
    sl.rel_last_iter = 0;
    sl.rel_pending = sl.rel_final = false;
    sl.prec_first.clear();
    sl.prec_pending = false;
    sl.nW = 0;
    sl.soft.assign(sl.m0, 0);
    sl.soft_iter.assign(sl.m0, 0);
    sl.lam.clear();
    sl.lam_prev.clear();
    sl.lam_prev_valid = false;
    sl.state = ITER;
}

// Results of the slot's transition and finished solve from the step that
// launched them: the block's status, the tracking permutation, and the
// record of the finished k-point (tracked frequencies, exit norms, plane-wave
// projections).
template <typename T>
void Engine<T>::complete_step_results(int s) {
    Slot &sl = slots[s];
    const int *to = h_tout + (size_t)TOUT * s;
    std::vector<int> perm;
    if (sl.t_launched) {
        sl.t_launched = false;
        if (to[0] == -1) throw std::runtime_error("ran out of start vectors during block initialisation");
        if (to[0] == -2) throw std::runtime_error("k-point transition outside the kernel's capacity");
        if (to[0] != sl.t_m0) throw std::runtime_error("k-point transition built a short block");
        if (sl.rec_pending && sl.rec.track) perm.assign(to + 2, to + 2 + sl.rec.om.size());
    }
    if (!sl.rec_pending) return;
    sl.rec_pending = false;
    Slot::Record &R = sl.rec;
    std::vector<double> om = R.om;
    for (size_t i = 0; i < perm.size(); ++i) {
        if (perm[i] < 0 || perm[i] >= (int)R.om.size()) throw std::runtime_error("invalid tracking permutation");
        om[i] = R.om[perm[i]];
    }
    sl.prev_om = om;
    sl.have_prev = true;
    std::vector<double> rel = R.rel;
    if (R.rel_fin) {  // soft-locked bands: zero residual
        rel.assign(std::min(R.rel_m, sl.nb), 0.0);
        const double *h = h_relf + (size_t)LD * s;
        for (int b : R.rel_act)
            if (b < (int)rel.size()) rel[b] = h[b];
    }
    // projections of the eigenvectors, in E order (ascending eigenvalue);
    // tracking permutes the frequencies only
    const double *pw = R.pw ? group_of(s).h_pw + (size_t)(s - group_of(s).s0) * nbmax * pw_ng() * 2 : nullptr;
    record_k(s, R.kidx, om, R.iters, R.conv, R.prec, rel, R.rel_iter, pw);
    if (R.save_first_gamma) {
        sl.first_gamma = om;
        sl.first_gamma_iters = R.iters;
        sl.first_gamma_rel = rel;
        sl.first_gamma_rel_iter = R.rel_iter;
        sl.first_gamma_prec = R.prec;
        if (pw) sl.first_gamma_pw.assign(pw, pw + (size_t)sl.nb * pw_ng() * 2);
        sl.have_first_gamma = true;
    }
    if (R.closing)
        record_k(s, R.closing_k, sl.first_gamma, sl.first_gamma_iters, true, sl.first_gamma_prec, sl.first_gamma_rel,
                 sl.first_gamma_rel_iter, sl.first_gamma_pw.empty() ? nullptr : sl.first_gamma_pw.data());
    if (sl.state == DRAIN) {
        sl.state = FREE;
        sl.job = -1;
    }
}

template <typename T>
void Engine<T>::run(const std::vector<JobSpec> &jobs, std::vector<JobResult> &out, RunStats &stats) {
    jobs_ = &jobs;
    out_ = &out;
    stats_ = &stats;
    out.assign(jobs.size(), JobResult());
    for (size_t j = 0; j < jobs.size(); ++j) {
        const JobSpec &J = jobs[j];
        if (J.nx != nx || J.ny != ny) throw std::runtime_error("job grid does not match engine grid");
        out[j].omegas.assign(J.kpath.size() * J.n_bands, NAN);
        out[j].iters.assign(J.kpath.size(), 0);
        out[j].converged.assign(J.kpath.size(), 0);
        out[j].rel_residuals.assign(J.kpath.size() * J.n_bands, NAN);
        out[j].rel_iters.assign(J.kpath.size(), 0);
        out[j].prec_first.assign(J.kpath.size() * J.n_bands, NAN);
        if (eopt.pw_half > 0) out[j].pw.assign(J.kpath.size() * J.n_bands * pw_ng() * 2, 0.0);
    }
    for (auto &sl : slots) sl = Slot();
    size_t next_job = 0;
    auto t0 = std::chrono::steady_clock::now();
    for (auto &G : groups) G.inflight = false;
    // Groups alternate: while the GPU runs one group's step, the host plans
    // the next; the groups' kernels overlap on the device.
    while (true) {
        bool launched = false;
        for (auto &G : groups) {
            if (G.inflight) {
                auto w0 = std::chrono::steady_clock::now();
                BH_HIP_CHECK(hipEventSynchronize(G.ev));
                stats.host_wait_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - w0).count();
                G.inflight = false;
            }
            auto p0 = std::chrono::steady_clock::now();
            StepPlan plan;
            // results of the group's previous step: transition status, tracking
            // permutations, records of finished k-points
            for (int s = G.s0; s < G.s1; ++s)
                if (slots[s].t_launched || slots[s].rec_pending) complete_step_results(s);
            for (int s = G.s0; s < G.s1; ++s)
                if (slots[s].state == ITER) {
                    absorb(s);
                    decide(s);
                }
            std::vector<std::pair<int, int>> starts;  // (slot, job)
            for (int s = G.s0; s < G.s1; ++s)
                if (slots[s].state == FREE && next_job < jobs.size()) starts.push_back({s, (int)next_job++});
            pool->parallel_for((int)starts.size(), [&](int i) { start_job(starts[i].first, starts[i].second); });
            for (auto &sj : starts) {
                const int s = sj.first;
                plan.table_upload.push_back(s);
                plan.fill_y0.push_back(s);
                begin_block(s, true);
            }
            for (int s = G.s0; s < G.s1; ++s) {
                Slot &sl = slots[s];
                if (sl.fin_pending) plan.fin.push_back(s);
                if (sl.t_pending) plan.trans.push_back(s);
                if (sl.state == ITER) plan.iter.push_back(s);
            }
            const bool any = !plan.iter.empty() || !plan.trans.empty() || !plan.fin.empty();
            if (!any) {
                stats.host_plan_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - p0).count();
                continue;
            }
            execute(plan, G);
            stats.host_plan_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - p0).count();
            stats.steps++;
            launched = true;
        }
        if (!launched) {
            bool busy = false;
            for (auto &G : groups) busy |= G.inflight;
            if (!busy) break;
        }
    }
    for (auto &G : groups) BH_HIP_CHECK(hipStreamSynchronize(G.st));
    for (auto &G : groups) G.harvest(&stats.timed_ms);
    stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

template <typename T>
void Engine<T>::execute(StepPlan &plan, Group &G) {
    Staging &stg = G.stg;
    const hipStream_t st = G.st;
    G.harvest(&stats_->timed_ms);  // this group's previous step has completed
    double *const d_part = G.part;
    int *const d_sweeps = G.sweeps;
    stg.off = 0;
    const bool need_b_eps = true;
    (void)need_b_eps;
    // OpParams for every slot (small), referenced by the operator work items
    OpParams *dparams;
    OpParams *hp = stg.alloc<OpParams>(S, &dparams);
    for (int s = 0; s < S; ++s) hp[s] = slots[s].params;

    // ---- new jobs: tables and the Gamma mode ----
    struct TableCopy { T *dst; T *src; size_t n; };
    std::vector<TableCopy> tcopies;
    for (int s : plan.table_upload) {
        Slot &sl = slots[s];
        T *de, *dt;
        T *he = stg.alloc<T>(N, &de);
        T *ht = stg.alloc<T>(4 * N, &dt);
        for (long i = 0; i < N; ++i) he[i] = (T)sl.diel.eps[i];
        for (long i = 0; i < 4 * N; ++i) ht[i] = (T)sl.diel.tensor[i];
        tcopies.push_back({eps(s), de, (size_t)N});
        tcopies.push_back({ten(s), dt, (size_t)4 * N});
    }

    // Finished solves and k-point transitions are planned before the
    // iteration section, which moves each iterating slot to its other V / AV
    // copy: they address the copy that holds the finished solve's X and AX.

    // ---- finished solves: the exit iteration's residual norms, plane-wave projections ----
    std::vector<ResidualWork<T>> res_fin;
    std::vector<DeflateWork<T>> def_fin;
    std::vector<int> def_fin_slot;
    std::vector<PwWork<T>> pw_proj;
    int max_r_fin = 0, max_d_fin = 0;
    for (int s : plan.fin) {
        Slot &sl = slots[s];
        sl.fin_pending = false;
        const Slot::Record &R = sl.rec;
        if (R.rel_fin) {
            // R = A X - B X lambda of the bands not soft-locked, into the free
            // columns [m, m + a), deflated against y0 and the soft-locked
            // bands, then their relative B-norms (into d_relf)
            const int m = R.rel_m, a = (int)R.rel_act.size();
            int *act_d;
            int *act_h = stg.alloc<int>(std::max(a, 1), &act_d);
            for (int i = 0; i < a; ++i) act_h[i] = R.rel_act[i];
            ResidualWork<T> rw{};
            rw.ax = avcol(s, 0);
            rw.bx = vcol(s, 0);
            rw.bw = bweight(s);
            rw.out = vcol(s, m);
            rw.act = act_d;
            rw.lam = d_lam + (size_t)s * LD;
            rw.stride = N;
            rw.count = a;
            max_r_fin = std::max(max_r_fin, a);
            std::vector<int> dcols;
            if (R.gamma) dcols.push_back(C);
            for (int j = 0; j < (int)R.rel_soft.size() && j < m; ++j)
                if (R.rel_soft[j]) dcols.push_back(j);
            const int nd = (int)dcols.size();
            if (nd > 0 && a > 0) {
                const cplx<T> **dd_d;
                const cplx<T> **dd = stg.alloc<const cplx<T> *>(nd, (const cplx<T> ***)&dd_d);
                for (int i = 0; i < nd; ++i) dd[i] = vcol(s, dcols[i]);
                DeflateWork<T> dw{};
                dw.d_ptrs = dd_d;
                dw.weight = bweight(s);
                dw.R = vcol(s, m);
                dw.stride = N;
                dw.nd = nd;
                dw.a = a;
                dw.ld = LD;
                dw.act = act_d;
                dw.lam = d_lam + (size_t)s * LD;
                dw.rel = d_relf + (size_t)s * LD;
                def_fin.push_back(dw);
                def_fin_slot.push_back(s);
                max_d_fin = std::max(max_d_fin, nd);
            } else {
                rw.rel = d_relf + (size_t)s * LD;
            }
            res_fin.push_back(rw);
        }
        if (R.pw) {  // projections of the finished solve's eigenvectors
            for (int i = 0; i < (int)sl.E_cols.size() && i < sl.nb; ++i) {
                PwWork<T> pw{};
                pw.x = vcol(s, sl.E_cols[i]);
                pw.c = pwbuf(s) + (size_t)i * pw_ng() * 2;
                pw.stride = N;
                pw.count = 1;
                pw_proj.push_back(pw);
            }
        }
    }

    // ---- k-point transitions: Gram of K, the device transition, history copy,
    // the new block X = K cx, its A X and Rayleigh quotients ----
    std::vector<GramWork<T>> g_t1;
    std::vector<TransWork> tw;
    std::vector<LincombWork<T>> lc_x;  // history copies and new blocks, one launch
    std::vector<OpWork<T>> op_x[2];
    std::vector<RayleighWork<T>> rq;
    int maxq_x = 0, maxkd = 0;
    for (int s : plan.trans) {
        Slot &sl = slots[s];
        sl.t_pending = false;
        sl.t_launched = true;
        const int kd = sl.kd, nE = sl.nE, nP = sl.nP, nR = sl.nR, m0 = sl.t_m0;
        maxkd = std::max(maxkd, kd);
        // K = [E | P | R | y0 | Wc], B-weighted Gram (lower triangle)
        const cplx<T> **kp_d;
        const cplx<T> **kp = stg.alloc<const cplx<T> *>(kd, (const cplx<T> ***)&kp_d);
        int q = 0;
        for (int i = 0; i < nE; ++i) kp[q++] = vcol(s, sl.E_cols[i]);
        for (int i = 0; i < nP; ++i) kp[q++] = hcol(s, sl.t_hist_old, i);
        for (int i = 0; i < nR; ++i) kp[q++] = rcol(i);
        kp[q++] = vcol(s, C);
        for (int i = 0; i < sl.t_nwc; ++i) kp[q++] = hcol(s, 1, i);
        GramWork<T> g{};
        g.a_ptrs = kp_d;
        g.b_ptrs = kp_d;
        g.weight = bweight(s);
        g.na = kd;
        g.nb = kd;
        g.ld = LD;
        g.lower_only = 1;
        TransWork t{};  // t.gram and t.nchunks: the Gram's partial sums, assigned below
        g_t1.push_back(g);
        double *om_d = nullptr, *pv_d = nullptr;
        if (!sl.t_om.empty()) {
            double *h = stg.alloc<double>(sl.t_om.size(), &om_d);
            std::memcpy(h, sl.t_om.data(), sizeof(double) * sl.t_om.size());
        }
        if (!sl.prev_om.empty()) {
            double *h = stg.alloc<double>(sl.prev_om.size(), &pv_d);
            std::memcpy(h, sl.prev_om.data(), sizeof(double) * sl.prev_om.size());
        }
        t.om = om_d;
        t.prev_om = pv_d;
        t.nprev = (int)sl.prev_om.size();
        t.cx = d_tcx + (size_t)s * LD * LD * 2;
        t.hist = d_thist + (size_t)s * nbmax * nbmax * 2;
        t.out = d_tout + (size_t)s * TOUT;
        t.scratch = d_tscr ? d_tscr + (size_t)s * tbytes : nullptr;
        t.alpha = sl.t_alpha;
        t.ld = LD;
        t.kd = kd;
        t.nE = nE;
        t.nP = nP;
        t.nR = nR;
        t.nWc = sl.t_nwc;
        t.nb = sl.nb;
        t.m0 = m0;
        t.flags = sl.t_flags;
        tw.push_back(t);
        if (m0 == 0) continue;  // the job ends: the tracking permutation only
        const int hn = sl.t_hist_new;
        if (nE > 0 && hn >= 0) {
            // history copy: H[hn][i] = E[perm[i]] through permutation coefficients
            const cplx<T> **pp_d;
            const cplx<T> **pp = stg.alloc<const cplx<T> *>(nE, (const cplx<T> ***)&pp_d);
            for (int i = 0; i < nE; ++i) pp[i] = vcol(s, sl.E_cols[i]);
            LincombWork<T> w{};
            w.a_ptrs = pp_d;
            w.na = nE;
            w.nq = nE;
            w.C = t.hist;
            w.ldc = nE;
            w.out = hcol(s, hn, 0);
            w.out_stride = N;
            w.mode = 0;
            lc_x.push_back(w);
        }
        // X = K cx into the slot's other V copy (K's columns stay untouched
        // while the history copy reads E in the same launch)
        const int nxt = 1 - sl.par;
        LincombWork<T> w{};
        w.a_ptrs = kp_d;
        w.na = kd;
        w.nq = m0;
        w.C = t.cx;
        w.ldc = kd;
        w.out = vbuf(s, nxt, 0);
        w.out_stride = N;
        lc_x.push_back(w);
        maxq_x = std::max(maxq_x, std::max(m0, nE));
        sl.par = nxt;
        OpWork<T> o{};
        o.in = vcol(s, 0);
        o.out = avcol(s, 0);
        o.eps = eps(s);
        o.tensor = ten(s);
        o.params = dparams + s;
        o.tmp = tmpcol(s);
        o.tmp_stride = N;
        o.in_stride = N;
        o.out_stride = N;
        o.count = m0;
        op_x[sl.pol].push_back(o);
        RayleighWork<T> rw{};
        rw.x = vcol(s, 0);
        rw.ax = avcol(s, 0);
        rw.bx = vcol(s, 0);
        rw.bw = bweight(s);
        rw.lam = d_lam + (size_t)s * LD;
        rw.stride = N;
        rw.count = m0;
        rq.push_back(rw);
    }

    // ---- iteration for ITER slots ----
    std::vector<ResidualWork<T>> res;
    bool any_norms = false;  // residual B-norms on log iterations, read back after the step
    std::vector<DeflateWork<T>> defR, defP;
    std::vector<int> defR_slot, defP_slot;  // slot of each deflation item
    std::vector<OpWork<T>> op_pre[2];  // [0] FD, [1] TP
    std::vector<NormWork<T>> prec_norms;  // first iterations: B-norms of M^-1 r
    std::vector<OpWork<T>> op_p[2];    // Theta on the new directions [TM, TE]
    long theta_cols[2] = {0, 0};
    std::vector<Gram2Work<T>> g_dual;
    std::vector<int> rr_slot;  // slot of each Rayleigh-Ritz item
    std::vector<RRWork> e_rr2;
    std::vector<LincombWork<T>> lc_ritz;
    int max_ritz = 0;
    std::vector<unsigned char> maskbuf;
    int max_a = 0, max_p = 0, max_d = 0, max_r = 0;
    struct DefMask { int slot; int nd; int a; std::vector<unsigned char> m; };
    std::vector<DefMask> masksP;
    for (int s : plan.iter) {
        Slot &sl = slots[s];
        const int m = sl.m;
        std::vector<int> act;
        for (int i = 0; i < m; ++i)
            if (!(i < (int)sl.soft.size() && sl.soft[i])) act.push_back(i);
        const int a = (int)act.size();
        // residual columns: the active bands, then those soft-locked in this
        // iteration's decision (their residuals are formed, deflated and
        // normed as the reference's are, but not taken into P)
        std::vector<int> rcols = act;
        for (int i = 0; i < m; ++i)
            if (i < (int)sl.soft.size() && sl.soft[i] && !(i < (int)sl.soft_iter.size() && sl.soft_iter[i]))
                rcols.push_back(i);
        const int ar = (int)rcols.size();
        const int w = sl.nW > 0 ? std::min(sl.nW, a) : 0;
        const int p = m + w + a;
        if (p > C || m + w + ar > C) throw std::runtime_error("subspace wider than the engine allows");
        max_a = std::max(max_a, a);
        max_r = std::max(max_r, ar);
        max_p = std::max(max_p, p);
        // X at [0, m), W at [m, m + nW) as the last Ritz update wrote them; the
        // first w of W are used and residuals overwrite the rest
        // residuals into P slots [m + w, m + w + ar)
        int *act_d;
        int *act_h = stg.alloc<int>(std::max(ar, 1), &act_d);
        for (int i = 0; i < ar; ++i) act_h[i] = rcols[i];
        ResidualWork<T> rw{};
        rw.ax = avcol(s, 0);
        rw.bx = vcol(s, 0);
        rw.bw = bweight(s);
        rw.out = vcol(s, m + w);
        rw.act = act_d;
        rw.lam = d_lam + (size_t)s * LD;
        rw.stride = N;
        rw.count = ar;
        // the reference's residual B-norms on log iterations (eigensolver.rs:1729-1737),
        // taken by the last kernel that writes the residuals
        const bool log_it = should_log_iteration(sl.it);
        // residual deflation against D_old = [y0 if Gamma] + the bands soft-locked
        // before this iteration (eigensolver.rs:1715-1722)
        std::vector<int> rdcols;
        if (sl.gamma_defl) rdcols.push_back(C);
        for (int j = 0; j < (int)sl.soft_iter.size() && j < m; ++j)
            if (sl.soft_iter[j]) rdcols.push_back(j);
        const int nrd = (int)rdcols.size();
        max_d = std::max(max_d, nrd);
        if (nrd > 0 && ar > 0) {
            const cplx<T> **dd_d;
            const cplx<T> **dd = stg.alloc<const cplx<T> *>(nrd, (const cplx<T> ***)&dd_d);
            for (int i = 0; i < nrd; ++i) dd[i] = vcol(s, rdcols[i]);
            // r_c -= sum_i d_i <B d_i, r_c>, <B d, r> = sum conj(d) w r with w the B weight
            DeflateWork<T> dw{};
            dw.d_ptrs = dd_d;
            dw.weight = bweight(s);
            dw.R = vcol(s, m + w);
            dw.stride = N;
            dw.nd = nrd;
            dw.a = ar;
            dw.ld = LD;
            if (log_it) {
                dw.act = act_d;
                dw.lam = d_lam + (size_t)s * LD;
                dw.rel = d_rel + (size_t)s * LD;
            }
            defR.push_back(dw);
            defR_slot.push_back(s);
        } else if (log_it) {
            rw.rel = d_rel + (size_t)s * LD;
        }
        res.push_back(rw);
        // direction deflation against D_new = [y0 if Gamma] + every soft-locked band
        std::vector<int> dcols;
        if (sl.gamma_defl) dcols.push_back(C);
        for (int j = 0; j < (int)sl.soft.size() && j < m; ++j)
            if (sl.soft[j]) dcols.push_back(j);
        const int nd = (int)dcols.size();
        max_d = std::max(max_d, nd);
        if (nd > 0 && a > 0) {
            const cplx<T> **dd_d;
            const cplx<T> **dd = stg.alloc<const cplx<T> *>(nd, (const cplx<T> ***)&dd_d);
            for (int i = 0; i < nd; ++i) dd[i] = vcol(s, dcols[i]);
            DeflateWork<T> dw{};
            dw.d_ptrs = dd_d;
            dw.weight = bweight(s);
            dw.R = vcol(s, m + w);
            dw.stride = N;
            dw.nd = nd;
            dw.a = a;
            dw.ld = LD;
            // soft part skipped for compacted P index c with soft[c]
            defP.push_back(dw);
            defP_slot.push_back(s);
            DefMask dm{s, nd, a, std::vector<unsigned char>((size_t)LD * LD, 1)};
            for (int c = 0; c < a; ++c)
                if (c < (int)sl.soft.size() && sl.soft[c])
                    for (int i = 0; i < nd; ++i)
                        if (dcols[i] != C) dm.m[i + (size_t)c * LD] = 0;
            masksP.push_back(std::move(dm));
        }
        if (log_it) {
            any_norms = true;
            sl.rel_pending = true;
            sl.rel_pending_iter = sl.it + 1;
            sl.rel_act = rcols;
            sl.rel_m = m;
        }
        // the first iteration: B-norms of the preconditioned residuals (how far
        // one preconditioned step moves each start vector; M = I without a
        // preconditioner), taken before the direction deflation
        if (sl.it == 0 && a > 0) {
            NormWork<T> nw{};
            nw.r = vcol(s, m + w);
            nw.bw = bweight(s);
            nw.act = act_d;
            nw.lam = nullptr;
            nw.rel = d_prec + (size_t)s * LD;
            nw.stride = N;
            nw.count = a;
            prec_norms.push_back(nw);
            sl.prec_pending = true;
        }
        // preconditioner (P in place)
        if (sl.precond == 1 || sl.precond == 2) {
            OpWork<T> o{};
            o.in = vcol(s, m + w);
            o.out = vcol(s, m + w);
            o.eps = eps(s);
            o.tensor = ten(s);
            o.params = dparams + s;
            o.tmp = tmpcol(s);
            o.tmp_stride = N;
            o.in_stride = N;
            o.out_stride = N;
            o.count = a;
            op_pre[sl.precond - 1].push_back(o);
        }
        // A P for the new directions; A X and A W come from the last Ritz update
        if (a > 0) {
            OpWork<T> o{};
            o.in = vcol(s, m + w);
            o.out = avcol(s, m + w);
            o.eps = eps(s);
            o.tensor = ten(s);
            o.params = dparams + s;
            o.tmp = tmpcol(s);
            o.tmp_stride = N;
            o.in_stride = N;
            o.out_stride = N;
            o.count = a;
            op_p[sl.pol].push_back(o);
            theta_cols[sl.pol] += a;
        }
        // both Rayleigh-Ritz Grams Z^H B Z and Z^H A Z in one pass
        Gram2Work<T> g2{};
        g2.Z = vcol(s, 0);
        g2.bweight = bweight(s);
        g2.AZ = avcol(s, 0);
        g2.stride = N;
        g2.n = p;
        g2.ld = LD;
        g_dual.push_back(g2);
        RRWork rw2{};
        rw2.lam = d_lam + (size_t)s * LD;
        rw2.C = d_Y + (size_t)s * LD * LD * 2;
        rw2.cnt = d_cnt + CNT_STRIDE * s;
        rw2.n = p;
        rw2.ld = LD;
        rw2.ldc = LD;
        rw2.nsplit = sl.nb;
        if (eopt.profile) rw2.sweeps_out = d_sweeps;
        e_rr2.push_back(rw2);
        rr_slot.push_back(s);
        sl.last_p = p;
        // Ritz update into the other copy: [X' | W'] = Z C, A[X' | W'] = (A Z) C,
        // B[X' | W'] = eps [X' | W'] (TM)
        const int nxt = 1 - sl.par;
        LincombWork<T> lx{};
        lx.A = vcol(s, 0);
        lx.a_stride = N;
        lx.na = p;
        lx.dev_nq = d_cnt + CNT_STRIDE * s + 2;
        lx.C = d_Y + (size_t)s * LD * LD * 2;
        lx.ldc = LD;
        lx.out = vbuf(s, nxt, 0);
        lx.out_stride = N;
        lc_ritz.push_back(lx);
        LincombWork<T> la = lx;
        la.A = avcol(s, 0);
        la.out = avbuf(s, nxt, 0);
        la.out2 = nullptr;
        la.w2 = nullptr;
        lc_ritz.push_back(la);
        max_ritz = std::max(max_ritz, std::min(p, 2 * sl.nb));
        sl.par = nxt;
        sl.it += 1;
    }

    std::vector<PwWork<T>> pw_syn;
    // learned start: the new jobs' warm vectors from their plane-wave coefficients
    int max_wc = 0, wc_half = 0;
    for (int s : plan.table_upload) {
        Slot &sl = slots[s];
        if (sl.nWc <= 0) continue;
        const JobSpec &J = (*jobs_)[sl.job];
        const size_t ngw = (size_t)(2 * J.pw_half + 1) * (2 * J.pw_half + 1);
        if (!pw_syn.empty() && J.pw_half != wc_half)
            throw std::runtime_error("jobs in one run must share the warm plane-wave half width");
        wc_half = J.pw_half;
        double *cd;
        double *ch = stg.alloc<double>(2 * ngw * sl.nWc, &cd);
        std::memcpy(ch, J.warm_pw.data(), sizeof(double) * 2 * ngw * sl.nWc);
        PwWork<T> pw{};
        pw.cin = cd;
        pw.out = hcol(s, 1, 0);
        pw.stride = N;
        pw.count = sl.nWc;
        pw_syn.push_back(pw);
        max_wc = std::max(max_wc, sl.nWc);
    }
    // ---- descriptor arrays into staging ----
    auto put = [&](auto &vec) {
        using W = typename std::remove_reference<decltype(vec)>::type::value_type;
        W *d = nullptr;
        if (vec.empty()) return d;
        W *h = stg.alloc<W>(vec.size(), &d);
        std::memcpy((void *)h, vec.data(), sizeof(W) * vec.size());
        return d;
    };
    // Gram partial buffers: item i of a launch uses partial block i
    auto nch_for = [&](size_t nwork) {
        long c = (long)((chunk_budget + nwork - 1) / std::max<size_t>(nwork, 1));
        c = std::min<long>(std::max<long>(c, 1), NCHMAX);
        return (int)c;
    };
    auto assign_partials = [&](std::vector<GramWork<T>> &v, int nch) {
        for (size_t i = 0; i < v.size(); ++i) v[i].partial = d_part + i * (size_t)nch * LD * LD * 2;
    };
    // the transition Gram's partials are summed into chunk 0 before the transition kernel loads M
    const int nch_dual = nch_for(g_dual.size()), nch_t1 = nch_for(g_t1.size());
    assign_partials(g_t1, nch_t1);
    for (size_t i = 0; i < g_t1.size(); ++i) {
        tw[i].gram = g_t1[i].partial;
        tw[i].nchunks = 1;
    }
    for (size_t i = 0; i < g_dual.size(); ++i) {  // two products per item
        g_dual[i].partial1 = d_part + 2 * i * (size_t)nch_dual * LD * LD * 2;
        g_dual[i].partial2 = g_dual[i].partial1 + (size_t)nch_dual * LD * LD * 2;
        e_rr2[i].partial1 = g_dual[i].partial1;
        e_rr2[i].partial2 = g_dual[i].partial2;
        // narrow: the dual Gram's partials are summed into chunk 0 before the Rayleigh-Ritz launch
        e_rr2[i].nchunks = wide ? nch_dual : 1;
        if (wide) e_rr2[i].ws = G.rrws + (size_t)(rr_slot[i] - G.s0) * (rr_wide_bytes(C) / sizeof(double));
    }
    // wide subspaces: the two Rayleigh-Ritz Grams as tiled Gram products
    // (G1 = Z^H (B Z) with the B weight on the right operand, G2 = Z^H (A Z))
    std::vector<GramWork<T>> g_wide;
    if (wide) {
        for (const Gram2Work<T> &d : g_dual) {
            GramWork<T> g{};
            g.A = d.Z;
            g.B = d.Z;
            g.a_stride = g.b_stride = d.stride;
            g.weight = d.bweight;
            g.na = g.nb = d.n;
            g.ld = d.ld;
            g.lower_only = 1;
            g.partial = d.partial1;
            g_wide.push_back(g);
            g.B = d.AZ;
            g.weight = nullptr;
            g.partial = d.partial2;
            g_wide.push_back(g);
        }
    }
    // P-deflation masks
    unsigned char *mask_d = nullptr;
    if (!masksP.empty()) {
        unsigned char *mh = stg.alloc<unsigned char>(masksP.size() * LD * LD, &mask_d);
        for (size_t i = 0; i < masksP.size(); ++i) std::memcpy(mh + i * LD * LD, masksP[i].m.data(), LD * LD);
    }
    for (size_t i = 0; i < defP.size(); ++i) defP[i].mask = mask_d + i * LD * LD;
    // Columns longer than the one-launch deflation holds in registers, or more
    // deflation vectors than it takes: the same deflation in three launches
    // (coefficients as a Gram over row chunks, their masked reduction into the
    // slot's coefficient block, r -= D C) plus the residual norms it owes.
    const bool defl_rows = N > DEFL_MAXROWS || max_d > DEFL_MAXD;
    struct DeflRows {
        std::vector<GramWork<T>> g;
        std::vector<ReduceWork> r;
        std::vector<LincombWork<T>> l;
        std::vector<NormWork<T>> n;
        int nch = 1;
        GramWork<T> *dg{};
        ReduceWork *dr{};
        LincombWork<T> *dl{};
        NormWork<T> *dn{};
    } drR, drP, drF;
    auto to_rows = [&](const std::vector<DeflateWork<T>> &items, const std::vector<int> &islot, DeflRows &o) {
        o.nch = nch_for(items.size());
        for (size_t i = 0; i < items.size(); ++i) {
            const DeflateWork<T> &d = items[i];
            double *part = d_part + i * (size_t)o.nch * LD * LD * 2;
            GramWork<T> g{};
            g.a_ptrs = d.d_ptrs;  // <B d_i, r_c> = sum conj(d_i) w r_c
            g.B = d.R;
            g.b_stride = d.stride;
            g.weight = d.weight;
            g.na = d.nd;
            g.nb = d.a;
            g.ld = LD;
            g.partial = part;
            o.g.push_back(g);
            o.r.push_back(ReduceWork{part, coef(islot[i]), d.mask, o.nch, 0});
            LincombWork<T> l{};
            l.a_ptrs = d.d_ptrs;
            l.na = d.nd;
            l.nq = d.a;
            l.C = coef(islot[i]);
            l.ldc = LD;
            l.base = d.R;
            l.base_stride = d.stride;
            l.out = d.R;
            l.out_stride = d.stride;
            l.mode = 1;
            o.l.push_back(l);
            if (d.rel) o.n.push_back(NormWork<T>{d.R, d.weight, d.act, d.lam, d.rel, d.stride, d.a, 0});
        }
    };
    // the finished solves' exit deflation, run before the transitions
    const bool defl_rows_fin = N > DEFL_MAXROWS || max_d_fin > DEFL_MAXD;
    if (defl_rows) {
        to_rows(defR, defR_slot, drR);
        to_rows(defP, defP_slot, drP);
    }
    if (defl_rows_fin) to_rows(def_fin, def_fin_slot, drF);
    for (DeflRows *o : {&drR, &drP, &drF}) {
        o->dg = put(o->g);
        o->dr = put(o->r);
        o->dl = put(o->l);
        o->dn = put(o->n);
    }
    auto deflate = [&](const DeflateWork<T> *dwork, const std::vector<DeflateWork<T>> &items, const DeflRows &o,
                       int maxa, int maxd, bool rows) {
        if (!rows) {
            launch_deflate<T>(dwork, (int)items.size(), N, maxa, maxd, st);
            return;
        }
        launch_gram<T>(o.dg, (int)o.g.size(), o.nch, N, std::max(maxd, 1), maxa, st, eopt.gram_mfma_min);
        launch_reduce(o.dr, (int)o.r.size(), LD, st);
        launch_lincomb<T>(o.dl, (int)o.l.size(), N, maxa, st, std::max(maxd, 1));
        launch_bnorm<T>(o.dn, (int)o.n.size(), maxa, N, st);
    };
    auto *d_res_fin = put(res_fin);
    auto *d_def_fin = put(def_fin);
    auto *d_tw = put(tw);
    auto *d_lc_x = put(lc_x);
    OpWork<T> *d_op_x[2] = {put(op_x[0]), put(op_x[1])};
    auto *d_rq = put(rq);
    auto *d_res = put(res);
    auto *d_prec_norms = put(prec_norms);
    auto *d_defR = put(defR);
    OpWork<T> *d_op_pre[2] = {put(op_pre[0]), put(op_pre[1])};
    auto *d_defP = put(defP);
    OpWork<T> *d_op_p[2] = {put(op_p[0]), put(op_p[1])};
    auto *d_g_dual = put(g_dual);
    auto *d_g_wide = put(g_wide);
    auto *d_e_rr2 = put(e_rr2);
    auto *d_lc_ritz = put(lc_ritz);
    auto *d_g_t1 = put(g_t1);
    auto *d_pw_proj = put(pw_proj);
    auto *d_pw_syn = put(pw_syn);
    stg.upload(st);

    // ---- launches (single stream, in dependency order) ----
    for (auto &tc : tcopies)
        BH_HIP_CHECK(hipMemcpyAsync(tc.dst, tc.src, sizeof(T) * tc.n, hipMemcpyDeviceToDevice, st));
    for (int s : plan.fill_y0) {
        Slot &sl = slots[s];
        for (int b = 0; b < 2; ++b)  // the Gamma mode sits in column C of both copies
            launch_fill<T>(vbuf(s, b, C), nullptr, eps(s), sl.y0val, 0.0, N, st);
    }
    launch_pw_synth<T>(d_pw_syn, (int)pw_syn.size(), max_wc, nx, ny, wc_half, st);
    // launches of the timed kernel class are bracketed by events on this stream
    // algorithmic HBM bytes of every launch: each column read or written once, tables and weights once
    const double colb = (double)N * sizeof(cplx<T>), wb = (double)N * sizeof(T);
    auto lc_bytes = [&](const std::vector<LincombWork<T>> &v) {
        double b = 0;
        for (auto &l : v) b += colb * (l.na + l.nq * (l.out2 ? 2.0 : 1.0));
        return b;
    };
    for (auto &tc : tcopies) stats_->alg_bytes += 2.0 * sizeof(T) * tc.n;
    stats_->alg_bytes += 2.0 * colb * plan.fill_y0.size();
    stats_->alg_bytes += lc_bytes(lc_x);
    for (auto &r : rq) stats_->alg_bytes += r.count * 2.0 * colb + (r.bw ? wb : 0.0);
    for (auto &r : res) stats_->alg_bytes += r.count * 3.0 * colb + (r.bw ? wb : 0.0);
    for (auto &r : res_fin) stats_->alg_bytes += r.count * 3.0 * colb + (r.bw ? wb : 0.0);
    for (auto &d : def_fin) stats_->alg_bytes += (double)N * ((2.0 * d.a + d.nd) * sizeof(cplx<T>) + (d.weight ? sizeof(T) : 0));
    for (int k = 0; k < 2; ++k)
        for (auto &o : op_pre[k])
            stats_->alg_bytes += o.count * (2.0 + (o.bout ? 1.0 : 0.0)) * colb + (k == 1 || o.bout ? wb : 0.0);
    for (auto &n : prec_norms) stats_->alg_bytes += n.count * colb + (n.bw ? wb : 0.0);
    for (auto &g : g_t1) stats_->alg_bytes += (g.na + g.nb) * colb + (g.weight ? wb : 0.0);
    stats_->alg_bytes += pw_proj.size() * colb;
    auto timed = [&](int cls, double bytes, double flops, auto &&launch) {
        stats_->alg_bytes += bytes;
        if (eopt.timed_kernel != cls) { launch(); return; }
        if (G.tev_used == G.tev.size()) {
            hipEvent_t a, b;
            BH_HIP_CHECK(hipEventCreate(&a));
            BH_HIP_CHECK(hipEventCreate(&b));
            G.tev.push_back({a, b});
        }
        auto &ev = G.tev[G.tev_used++];
        BH_HIP_CHECK(hipEventRecord(ev.first, st));
        launch();
        BH_HIP_CHECK(hipEventRecord(ev.second, st));
        stats_->timed_launches += 1;
        stats_->timed_bytes += bytes;
        stats_->timed_flops += flops;
        stats_->timed_flops_executed += flops;
    };
    const double n_log = (double)N * std::log2((double)N);
    // finished solves: exit-iteration norms and plane-wave projections (read X)
    launch_residual<T>(d_res_fin, (int)res_fin.size(), std::max(max_r_fin, 1), N, st);
    if (!def_fin.empty()) deflate(d_def_fin, def_fin, drF, max_r_fin, max_d_fin, defl_rows_fin);
    launch_pw_project<T>(d_pw_proj, (int)pw_proj.size(), 1, nx, ny, eopt.pw_half, st);
    // transitions: Gram of K, tracking / prediction / block on the device,
    // history copy (reads E), X = K cx (overwrites E's copy)
    if (!g_t1.empty()) {
        launch_gram<T>(d_g_t1, (int)g_t1.size(), nch_t1, N, maxkd, maxkd, st, eopt.gram_mfma_min);
        launch_gram_lower_reduce<T>(d_g_t1, (int)g_t1.size(), nch_t1, maxkd, st);
        // the workspace is laid out for this launch's largest item
        double bytes = 0;  // the Gram in, block and history coefficients out
        int KDl = 1, NBWl = 1, M0l = 1, MMl = 1, NCl = 1;
        for (const TransWork &t : tw) {
            bytes += 16.0 * ((double)t.kd * t.kd + (double)t.kd * t.m0 + (double)t.nE * t.nE);
            const int nw = std::max(std::min(t.nE, t.nb), t.nWc);
            KDl = std::max(KDl, t.kd);
            NBWl = std::max(NBWl, nw);
            M0l = std::max(M0l, t.m0);
            MMl = std::max(MMl, std::min(t.nE, t.nP));
            NCl = std::max(NCl, std::max(nw + t.nR, t.m0));
        }
        timed(TIMED_TRANS, bytes, 0.0, [&] {
            launch_transition(d_tw, (int)tw.size(), KDl, NBWl, M0l, MMl, NCl, d_tscr != nullptr, st);
        });
    }
    launch_lincomb<T>(d_lc_x, (int)lc_x.size(), N, std::max(maxq_x, 1), st, LD);
    for (int pol = 0; pol < 2; ++pol) {
        if (op_x[pol].empty()) continue;
        long cols = 0;
        for (auto &o : op_x[pol]) cols += o.count;
        timed(TIMED_THETA, (double)cols * theta_bytes_per_col(pol), (double)cols * theta_flops_per_col(pol, n_log),
              [&] {
                  launch_op<T>(pol == 0 ? MODE_THETA_TM : MODE_THETA_TE, nx, ny, d_op_x[pol], (int)op_x[pol].size(),
                               m0max, twx, twy, st, false, revx, revy);
              });
    }
    launch_rayleigh<T>(d_rq, (int)rq.size(), m0max, N, st);
    // iteration
    launch_residual<T>(d_res, (int)res.size(), std::max(max_r, 1), N, st);
    // algorithmic: each column read and written once, each vector and weight read once
    auto defl_bytes = [&](const std::vector<DeflateWork<T>> &v) {
        double b = 0;
        for (auto &d : v) b += (double)N * ((2.0 * d.a + d.nd) * sizeof(cplx<T>) + (d.weight ? sizeof(T) : 0));
        return b;
    };
    auto defl_flops = [&](const std::vector<DeflateWork<T>> &v) {
        double f = 0;
        for (auto &d : v) f += 2.0 * 8.0 * (double)N * d.a * d.nd;
        return f;
    };
    if (!defR.empty())
        timed(TIMED_DEFL, defl_bytes(defR), defl_flops(defR),
              [&] { deflate(d_defR, defR, drR, max_r, max_d, defl_rows); });
    const bool prec_fp32 = eopt.precision == 2;
    launch_op<T>(MODE_PREC_FD, nx, ny, d_op_pre[0], (int)op_pre[0].size(), std::max(max_a, 1), twx, twy, st,
                 prec_fp32, revx, revy);
    launch_op<T>(MODE_PREC_TP, nx, ny, d_op_pre[1], (int)op_pre[1].size(), std::max(max_a, 1), twx, twy, st,
                 prec_fp32, revx, revy);
    launch_bnorm<T>(d_prec_norms, (int)prec_norms.size(), std::max(max_a, 1), N, st);
    if (!defP.empty())
        timed(TIMED_DEFL, defl_bytes(defP), defl_flops(defP),
              [&] { deflate(d_defP, defP, drP, max_a, max_d, defl_rows); });
    for (int pol = 0; pol < 2; ++pol) {
        if (op_p[pol].empty()) continue;
        timed(TIMED_THETA, (double)theta_cols[pol] * theta_bytes_per_col(pol),
              (double)theta_cols[pol] * theta_flops_per_col(pol, n_log), [&] {
                  launch_op<T>(pol == 0 ? MODE_THETA_TM : MODE_THETA_TE, nx, ny, d_op_p[pol], (int)op_p[pol].size(),
                               std::max(max_a, 1), twx, twy, st, false, revx, revy);
              });
    }
    {
        double bytes = 0, flops = 0;
        for (auto &g2 : g_dual) {
            bytes += 2.0 * g2.n * (double)N * sizeof(cplx<T>) + (g2.bweight ? (double)N * sizeof(T) : 0.0);
            flops += 2.0 * 8.0 * (double)N * g2.n * (g2.n + 1) / 2.0;
        }
        timed(TIMED_GRAM2, bytes, flops,
              [&] {
                  if (wide)
                      launch_gram<T>(d_g_wide, (int)g_wide.size(), nch_dual, N, max_p, max_p, st, eopt.gram_mfma_min);
                  else {
                      launch_gram2<T>(d_g_dual, (int)g_dual.size(), nch_dual, N, std::max(max_p, 1), st);
                      launch_gram2_reduce<T>(d_g_dual, (int)g_dual.size(), nch_dual, max_p, st);
                  }
              });
    }
    {
        double bytes = 0;  // the two n x n Grams in, the n x 2nb coefficients out
        for (auto &g2 : g_dual)
            bytes += 2.0 * g2.n * g2.n * 16.0 + (double)g2.n * std::min(g2.n, 2 * nbmax) * 16.0;
        // FLOPs are added when the sweep counts come back (absorb)
        timed(TIMED_RR, bytes, 0.0, [&] { launch_rr(d_e_rr2, (int)e_rr2.size(), std::max(max_p, 1), st); });
    }
    {
        // algorithmic: each input column read once, each output written once
        // (the output count min(p, 2nb) is the device-side count's bound)
        double bytes = 0, flops = 0;
        for (auto &l : lc_ritz) {
            const double nq = std::min(l.na, 2 * nbmax);
            bytes += (double)N * sizeof(cplx<T>) * (l.na + nq * (l.out2 ? 2.0 : 1.0));
            flops += 8.0 * N * l.na * nq;
        }
        timed(TIMED_RITZ, bytes, flops,
              [&] { launch_lincomb<T>(d_lc_ritz, (int)lc_ritz.size(), N, std::max(max_ritz, 1), st, max_p); });
    }
    if (!pw_proj.empty()) {
        const int gs = G.s1 - G.s0;
        BH_HIP_CHECK(hipMemcpyAsync(G.h_pw, pwbuf(G.s0), sizeof(double) * gs * nbmax * pw_ng() * 2,
                                    hipMemcpyDeviceToHost, st));
    }
    if (!tw.empty()) {
        const int gs = G.s1 - G.s0;
        BH_HIP_CHECK(hipMemcpyAsync(h_tout + (size_t)TOUT * G.s0, d_tout + (size_t)TOUT * G.s0,
                                    sizeof(int) * TOUT * gs, hipMemcpyDeviceToHost, st));
    }
    if (!res_fin.empty()) {
        const int gs = G.s1 - G.s0;
        BH_HIP_CHECK(hipMemcpyAsync(h_relf + (size_t)LD * G.s0, d_relf + (size_t)LD * G.s0, sizeof(double) * gs * LD,
                                    hipMemcpyDeviceToHost, st));
    }
    if (any_norms) {
        const int gs = G.s1 - G.s0;
        BH_HIP_CHECK(hipMemcpyAsync(G.h_rel, d_rel + (size_t)G.s0 * LD, sizeof(double) * gs * LD,
                                    hipMemcpyDeviceToHost, st));
    }
    if (!prec_norms.empty()) {
        const int gs = G.s1 - G.s0;
        BH_HIP_CHECK(hipMemcpyAsync(G.h_prec, d_prec + (size_t)G.s0 * LD, sizeof(double) * gs * LD,
                                    hipMemcpyDeviceToHost, st));
    }
    if (!plan.iter.empty()) {
        const int gs = G.s1 - G.s0;
        BH_HIP_CHECK(hipMemcpyAsync(h_lam + (size_t)G.s0 * LD, d_lam + (size_t)G.s0 * LD, sizeof(double) * gs * LD,
                                    hipMemcpyDeviceToHost, st));
        BH_HIP_CHECK(hipMemcpyAsync(h_cnt + CNT_STRIDE * (size_t)G.s0, d_cnt + CNT_STRIDE * (size_t)G.s0, sizeof(int) * CNT_STRIDE * gs,
                                    hipMemcpyDeviceToHost, st));
    }
    BH_HIP_CHECK(hipEventRecord(G.ev, st));
    G.inflight = true;
    G.has_iter = !plan.iter.empty();
    if (!eopt.profile) return;
    BH_HIP_CHECK(hipStreamSynchronize(st));
    if (!plan.iter.empty()) {
        int sw[6];
        BH_HIP_CHECK(hipMemcpyAsync(sw, d_sweeps, sizeof(sw), hipMemcpyDeviceToHost, st));
        BH_HIP_CHECK(hipMemsetAsync(d_sweeps, 0, sizeof(sw), st));
        BH_HIP_CHECK(hipStreamSynchronize(st));
        stats_->eigh_sweeps[0] += sw[0];
        stats_->eigh_sweeps[1] += sw[3];
        stats_->eigh_max_sweeps[0] = std::max(stats_->eigh_max_sweeps[0], (long)sw[2]);
        stats_->eigh_max_sweeps[1] = std::max(stats_->eigh_max_sweeps[1], (long)sw[5]);
        stats_->eigh_launch_max_sum[0] += sw[2];
        stats_->eigh_launch_max_sum[1] += sw[5];
        stats_->eigh_calls[0] += (long)plan.iter.size();
        stats_->eigh_calls[1] += (long)plan.iter.size();
    }
}

std::unique_ptr<EngineBase> make_engine(int nx, int ny, int nb_max, const EngineOptions &opt) {
    if (opt.precision == 1) return std::unique_ptr<EngineBase>(new Engine<float>(nx, ny, nb_max, opt));
    return std::unique_ptr<EngineBase>(new Engine<double>(nx, ny, nb_max, opt));
}

template class Engine<double>;
template class Engine<float>;

}  // namespace bh
