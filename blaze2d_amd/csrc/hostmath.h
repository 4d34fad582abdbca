// Host-side setup for the band-solve engine.
//
// Per-geometry setup (dielectric sampling / analytic smoothing) and per-k-point
// preconditioner statistics.
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

}  // namespace bh
