// Host-side setup and small dense algebra for the band-solve engine.
//
// Per-geometry setup (dielectric sampling / analytic smoothing), per-k-point
// preconditioner statistics, and the O(nb^3) decisions between k-points
// (band tracking, subspace prediction, block initialisation) that act on
// small Gram matrices produced on the GPU.
#pragma once

#include <array>
#include <complex>
#include <vector>

namespace bh {

using zd = std::complex<double>;

struct AtomSpec {
    double pos[2];
    double radius;  // in units of |a1|
    double eps_inside;
};

struct GeometrySpec {
    double a1[2], a2[2];
    double eps_bg;
    std::vector<AtomSpec> atoms;
};

struct DielectricTables {
    int nx = 0, ny = 0;
    std::vector<double> eps, inv_eps;
    std::vector<double> tensor;  // [N][4]: xx, xy, yx, yy
    double b1[2], b2[2];
    double eps_mean = 1.0;       // arithmetic mean of eps (TM effective epsilon)
    double eps_harm = 1.0;       // harmonic mean (TE effective epsilon)
    double eps_sum = 0.0;
};

// dielectric.rs:146-210,473-574 (analytic smoothing) or raw node sampling.
DielectricTables build_dielectric(const GeometrySpec &g, int nx, int ny, bool smoothing);

// lattice.rs:170-180
void reciprocal(const GeometrySpec &g, double b1[2], double b2[2]);

// fourier_diagonal.rs:146-189: s_min over clamped |k+G|^2 > 1e-6
double spectral_smin(int nx, int ny, const double b1[2], const double b2[2], const double bloch[2]);

// maxwell.rs:1609-1690: the smallest |k+G|^2 after clamping each k+G to
// length >= sqrt(1e-9) (ThetaOperator::k_plus_g_sq_min)
double kpg_sq_min(int nx, int ny, const double b1[2], const double b2[2], const double bloch[2]);

// Hermitian eigendecomposition from the lower triangle of a (Householder
// tridiagonalisation, implicit-shift QL), ascending, col-major vectors.
void eigh(int n, const std::vector<zd> &a, std::vector<double> &w, std::vector<zd> &v);

// Polar factor U of an m x m col-major matrix (subspace_prediction.rs:491-531).
// Returns the smallest singular value; full_rank (optional) is set when every
// singular value passed the 1e-13 relative cut, i.e. U = O (O^H O)^-1/2 is the
// unique polar factor.
double polar(int m, const std::vector<zd> &o, std::vector<zd> &u, bool *full_rank = nullptr);

// band_tracking.rs hungarian_assignment
std::vector<int> hungarian(const std::vector<std::vector<double>> &w);

// band_tracking.rs track_bands_with_frequencies given the overlap O = prev^H B curr.
// polar_u / polar_smin / polar_full (optional) receive the polar() results of
// the overlap.
std::vector<int> track_bands(int m, const std::vector<zd> &overlap, const std::vector<double> &prev_om,
                             const std::vector<double> &curr_om, std::vector<zd> *polar_u = nullptr,
                             double *polar_smin = nullptr, bool *polar_full = nullptr);

// Coefficient-space vector algebra: vectors are columns of coefficients over a
// basis K whose (weighted) Gram matrix M = K^H B K is known (col-major kd x kd).
struct CoefSpace {
    int kd;
    const std::vector<zd> *M;                  // Hermitian kd x kd, column-major
    zd inner(const zd *a, const zd *b) const;  // a^H M b
    void apply(const zd *v, zd *out) const;    // out = M v
};

// a^H b over n entries
zd cdot(const zd *a, const zd *b, int n);

}  // namespace bh
