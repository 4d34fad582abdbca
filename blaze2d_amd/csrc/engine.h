// Batched band-structure engine: many LOBPCG instances ("slots") advance in
// lockstep steps on one GPU; the host keeps the reference's per-instance
// control logic, the GPU does every O(N) operation.
#pragma once

#include <array>
#include <memory>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "blas.h"
#include "hostmath.h"

namespace bh {

enum TimedKernel { TIMED_NONE = 0, TIMED_THETA = 1, TIMED_GRAM2 = 2, TIMED_RR = 3, TIMED_RITZ = 4, TIMED_DEFL = 5,
                   TIMED_TRANS = 6 };

struct JobSpec {
    GeometrySpec geom;
    int nx = 0, ny = 0;
    int pol = 0;            // 0 TM, 1 TE
    int n_bands = 8;
    int max_iter = 200;
    double tol = 1e-6;
    int block_size = 0;     // 0 -> n_bands + 2
    int smoothing = 1;
    int precond = -1;       // -1 auto (TM FD, TE TP), 0 none, 1 FD, 2 TP
    int prediction = 1;
    int extrapolation = 1;
    int tracking = 1;
    int reuse_gamma = 1;
    int exit_rule = 0;      // 0 solve() accounting, 1 soft-locked only
    std::vector<std::array<double, 2>> kpath;
    // learned start: plane-wave coefficients of warm vectors for the first
    // k-point, warm_count x (2 pw_half + 1)^2 complex (interleaved)
    std::vector<double> warm_pw;
    int warm_count = 0;
    int pw_half = 0;
};

struct JobResult {
    std::vector<double> omegas;  // nk * n_bands
    std::vector<int> iters, converged;
    // relative residual B-norms of the first n_bands block entries at each
    // solve's last log iteration (nk * n_bands) and that iteration, 1-based (nk)
    std::vector<double> rel_residuals;
    std::vector<int> rel_iters;
    // per k-point and band: B-norm of the preconditioned residual M^-1 r_i of
    // the solve's first iteration (M = I without a preconditioner)
    std::vector<double> prec_first;
    // with EngineOptions::pw_half > 0: per k-point the n_bands converged
    // eigenvectors (ascending eigenvalue) projected on the (2 h + 1)^2 plane
    // waves G = n1 b1 + n2 b2, |n1|, |n2| <= h (nk * n_bands * (2h+1)^2 complex)
    std::vector<double> pw;
};

struct RunStats {
    long steps = 0;
    long iterations = 0;          // total LOBPCG iterations over all solves
    long solves = 0;              // k-point solves (excluding reused closing Gamma)
    double seconds = 0;
    // launches of the kernel class EngineOptions::timed_kernel, timed with
    // HIP events on the stream they run on, and their algorithmic work
    double timed_ms = 0;
    long timed_launches = 0;
    double timed_bytes = 0;       // algorithmic HBM bytes summed over those launches
    double timed_flops = 0;       // algorithmic FLOPs summed over those launches
    double timed_flops_executed = 0;  // the FLOPs the kernels ran (Rayleigh-Ritz: Jacobi sweeps as run)
    long eigh_calls[2] = {0, 0};     // [SVQB, Rayleigh-Ritz] problems solved (profile mode)
    long eigh_sweeps[2] = {0, 0};    // Jacobi sweeps summed over those problems
    long eigh_max_sweeps[2] = {0, 0};      // largest sweep count seen in one problem
    long eigh_launch_max_sum[2] = {0, 0};  // per launch maximum, summed over launches
    double host_wait_seconds = 0;          // blocked on group completion events
    double host_plan_seconds = 0;          // planning and enqueueing
    long rr_problems = 0;                  // Rayleigh-Ritz problems solved
    long svqb_fallbacks = 0;               // of them, solved on the SVQB path (Cholesky QR not provably equivalent)
    long rank_drops = 0;                   // of them, with a direction dropped (rank < subspace width)
    double alg_bytes = 0;                  // algorithmic HBM bytes of every launch (each column read or written once)
};

struct EngineOptions {
    int slots = 64;         // concurrent band structures
    int precision = 0;      // 0 fp64, 1 fp32 storage, 2 fp64 with FP32 preconditioners
    int profile = 0;        // time operator kernels with events (forces one group)
    int groups = 4;         // slot groups pipelined on separate streams
    int host_threads = 0;   // host threads for per-slot planning (0: hardware threads, at most 16)
    int timed_kernel = 0;   // TIMED_* class whose launches are timed with events (0: none)
    int pw_half = 0;        // > 0: project converged eigenvectors on the low plane waves (JobResult::pw)
    int gram_mfma_min = GRAM_MFMA_MIN_DEFAULT;  // FP64 Gram products at least this wide use the matrix cores
    // Streams owned by the caller.  HIP gives each new stream its own hardware
    // queue until GPU_MAX_HW_QUEUES is reached and then shares queues, which
    // serialises the groups; a caller that creates the group streams early
    // (before other streams and before any null-stream use) keeps them apart.
    hipStream_t setup_stream = nullptr;
    hipStream_t group_streams[4] = {nullptr, nullptr, nullptr, nullptr};
};

class EngineBase {
  public:
    virtual ~EngineBase() {}
    virtual void run(const std::vector<JobSpec> &jobs, std::vector<JobResult> &out, RunStats &stats) = 0;
};

std::unique_ptr<EngineBase> make_engine(int nx, int ny, int nb_max, const EngineOptions &opt);

}  // namespace bh
