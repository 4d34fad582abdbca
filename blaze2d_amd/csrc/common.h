// Shared definitions for the gfx950 band-solve kernels and the host engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace bh {

// Complex value in storage precision T (interleaved re, im; 8 or 16 bytes).
template <typename T>
struct cplx {
    T x, y;
};

// The global-memory view of a device pointer taken from a work descriptor.
// Through it, loads and stores become global_* instructions; through the
// generic pointer they are flat_* ones, which also count against the LDS wait
// counter, so every wait for an LDS read would wait for them too.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T *gmem(T *p) {
    return (__attribute__((address_space(1))) T *)p;
}
// element i of a complex array through the global view (both parts in one access)
template <typename T>
__device__ __forceinline__ cplx<T> gload(const cplx<T> *p, long i) {
    const __attribute__((address_space(1))) cplx<T> *g = gmem(p) + i;
    return cplx<T>{g->x, g->y};
}
template <typename T>
__device__ __forceinline__ void gstore(cplx<T> *p, long i, cplx<T> v) {
    __attribute__((address_space(1))) cplx<T> *g = gmem(p) + i;
    g->x = v.x;
    g->y = v.y;
}

template <typename T>
__host__ __device__ inline cplx<T> cmul(cplx<T> a, cplx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__host__ __device__ inline cplx<T> cmulc(cplx<T> a, cplx<T> b) {  // a * conj(b)
    return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename T>
__host__ __device__ inline cplx<T> cadd(cplx<T> a, cplx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__host__ __device__ inline cplx<T> csub(cplx<T> a, cplx<T> b) { return {a.x - b.x, a.y - b.y}; }

// Complex multiply-accumulate as four FMAs (written out: the plain expression
// acc += a*b + c*d compiles to a multiply, an FMA and an add).
// (re, im) += a * b
__device__ inline void cmac(double &re, double &im, double ar, double ai, double br, double bi) {
    re = fma(ar, br, re);
    re = fma(-ai, bi, re);
    im = fma(ar, bi, im);
    im = fma(ai, br, im);
}
// (re, im) += conj(a) * b
__device__ inline void cmacc(double &re, double &im, double ar, double ai, double br, double bi) {
    re = fma(ar, br, re);
    re = fma(ai, bi, re);
    im = fma(ar, bi, im);
    im = fma(-ai, br, im);
}

struct hip_error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define BH_HIP_CHECK(expr)                                                                  \
    do {                                                                                    \
        hipError_t _e = (expr);                                                             \
        if (_e != hipSuccess)                                                               \
            throw ::bh::hip_error(std::string(#expr) + ": " + hipGetErrorString(_e) + " (" + \
                                  __FILE__ + ":" + std::to_string(__LINE__) + ")");         \
    } while (0)

// Per-slot operator parameters for one k-point (maxwell.rs:126-176,
// fourier_diagonal.rs:146-189, transverse_projection.rs:83-126).
struct OpParams {
    double b1[2], b2[2];   // reciprocal lattice vectors
    double bloch[2];       // Bloch wavevector (Cartesian)
    double shift;          // adaptive preconditioner shift 0.5 * s_min
    double eps_eff;        // FD: arithmetic (TM) / harmonic (TE) mean
    double mass_floor;     // FD: 0.01 * eps_eff for TM, 0 for TE
    int pol;               // 0 TM, 1 TE
    int gamma;             // |bloch| < 1e-12 component-wise
    int precond;           // 0 none, 1 fourier-diagonal, 2 transverse-projection
    int pad;
};

constexpr double KG_FLOOR = 1e-9;  // operators.rs:62 K_PLUS_G_NEAR_ZERO_FLOOR

}  // namespace bh
