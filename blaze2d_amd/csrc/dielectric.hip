// Dielectric tables on the GPU: one thread per grid node, many geometries per
// launch (Dielectric2D::from_geometry, dielectric.rs:146-210,473-574, with the
// analytic smoothing of analytic_geometry.rs:183-541).
//
// Each node is the centre of an axis-aligned Cartesian pixel of size
// (|a1|/nx, |a2|/ny).  For every periodic image (3 x 3) of every circular
// atom the pixel's fill fraction is exact when the pixel is fully inside or
// outside the circle and sampled on 16 x 16 sub-points otherwise.  The atom
// with the largest fill (first one on ties) gives eps_inside and the
// interface normal (pixel centre minus circle centre); the total fill is
// clamped to [0, 1].  Then
//   eps      = f eps_in + (1 - f) eps_bg,   <1/eps> = f / eps_in + (1 - f) / eps_bg,
//   tensor   = [t + d ny^2, -d nx ny, -d nx ny, t + d nx^2],  t = 1/eps, d = <1/eps> - t,
// isotropic (t = <1/eps>) away from interfaces (f <= 1e-6 or >= 1 - 1e-6).
// Without smoothing a node takes the eps of the first atom whose image
// contains it (geometry.rs:59-119).  The same per-node arithmetic runs on the
// host in build_dielectric; the tables agree to rounding.
#include "dielectric.h"

#include <vector>

namespace bh {

namespace {

struct DevGeom {
    double a1[2], a2[2];
    double eps_bg;
    int n_atoms, pad;
    double atoms[DIEL_MAX_ATOMS][4];  // fractional x, y, radius (units of |a1|), eps_inside
};

__device__ inline void frac_to_cart(const DevGeom &g, double f0, double f1, double &x, double &y) {
    x = g.a1[0] * f0 + g.a2[0] * f1;
    y = g.a1[1] * f0 + g.a2[1] * f1;
}

__device__ inline void unit_normal(double cx, double cy, double px, double py, double &nx, double &ny) {
    const double dx = px - cx, dy = py - cy, d = sqrt(dx * dx + dy * dy);
    if (d < 1e-12) {
        nx = 1.0;
        ny = 0.0;
    } else {
        nx = dx / d;
        ny = dy / d;
    }
}

// fill fraction of the pixel (centre rc, size rs) inside the circle (cc, r);
// has_normal and normal for partially covered pixels
__device__ double circle_pixel(double ccx, double ccy, double radius, double rcx, double rcy, double rsx, double rsy,
                               bool &has_normal, double &nxo, double &nyo) {
    has_normal = false;
    const double hw = rsx / 2.0, hh = rsy / 2.0, area_px = rsx * rsy;
    const double x0 = rcx - hw, x1 = rcx + hw, y0 = rcy - hh, y1 = rcy + hh;
    const double r2 = radius * radius;
    const double cx[4] = {x0, x1, x0, x1}, cy[4] = {y0, y0, y1, y1};
    int inside = 0;
    double far2 = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const double d2 = (cx[i] - ccx) * (cx[i] - ccx) + (cy[i] - ccy) * (cy[i] - ccy);
        inside += d2 <= r2;
        far2 = fmax(far2, d2);
    }
    if (inside == 4) return 1.0;
    const double px = fmin(fmax(ccx, x0), x1) - ccx, py = fmin(fmax(ccy, y0), y1) - ccy;
    if (px * px + py * py > r2) return 0.0;
    if (far2 <= r2) return 1.0;
    constexpr int NS = 16;
    const double sw = rsx / NS, sh = rsy / NS, sa = sw * sh;
    double area = 0.0;
    for (int iy = 0; iy < NS; ++iy)
        for (int ix = 0; ix < NS; ++ix) {
            const double sx = x0 + (ix + 0.5) * sw - ccx, sy = y0 + (iy + 0.5) * sh - ccy;
            if (sx * sx + sy * sy <= r2) area += sa;
        }
    const double f = fmin(fmax(area / area_px, 0.0), 1.0);
    if (f < 1e-10) return 0.0;
    if (f > 1.0 - 1e-10) return 1.0;
    has_normal = true;
    unit_normal(ccx, ccy, rcx, rcy, nxo, nyo);
    return f;
}

__device__ double sample_eps(const DevGeom &g, double f0, double f1, double clen) {
    auto wrap_signed = [](double d) {
        while (d >= 0.5) d -= 1.0;
        while (d < -0.5) d += 1.0;
        return d;
    };
    f0 -= floor(f0);
    f1 -= floor(f1);
    for (int a = 0; a < g.n_atoms; ++a) {
        const double r = g.atoms[a][2] * clen;
        const double dx = f0 - g.atoms[a][0], dy = f1 - g.atoms[a][1];
        for (int n1 = -1; n1 <= 1; ++n1)
            for (int n2 = -1; n2 <= 1; ++n2) {
                double x, y;
                frac_to_cart(g, wrap_signed(dx) + n1, wrap_signed(dy) + n2, x, y);
                if (x * x + y * y <= r * r) return g.atoms[a][3];
            }
    }
    return g.eps_bg;
}

__global__ void __launch_bounds__(256) dielectric_kernel(const DevGeom *__restrict__ geoms, int nx, int ny,
                                                         int smoothing, double *__restrict__ eps,
                                                         double *__restrict__ inv_eps, double *__restrict__ tensor) {
    const DevGeom &g = geoms[blockIdx.y];
    const int n = nx * ny;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int iy = i / nx, ix = i - iy * nx;
    const size_t o = (size_t)blockIdx.y * n + i;
    const double clen = sqrt(g.a1[0] * g.a1[0] + g.a1[1] * g.a1[1]);
    double e, ie, t[4];
    if (!smoothing) {
        e = sample_eps(g, (double)ix / nx, (double)iy / ny, clen);
        ie = 1.0 / e;
        t[0] = ie; t[1] = 0.0; t[2] = 0.0; t[3] = ie;
    } else {
        double pcx, pcy;
        frac_to_cart(g, (double)ix / nx, (double)iy / ny, pcx, pcy);
        const double psx = clen / nx, psy = sqrt(g.a2[0] * g.a2[0] + g.a2[1] * g.a2[1]) / ny;
        bool have = false;
        double best = 0.0, total = 0.0, bnx = 0.0, bny = 0.0, be = 0.0;
        for (int a = 0; a < g.n_atoms; ++a)
            for (int di = -1; di <= 1; ++di)
                for (int dj = -1; dj <= 1; ++dj) {
                    double ccx, ccy;
                    frac_to_cart(g, g.atoms[a][0] + di, g.atoms[a][1] + dj, ccx, ccy);
                    bool hn;
                    double nrx = 0.0, nry = 0.0;
                    const double fill = circle_pixel(ccx, ccy, g.atoms[a][2] * clen, pcx, pcy, psx, psy, hn, nrx, nry);
                    if (fill > 1e-10) {
                        total += fill;
                        if (!have || fill > best) {
                            if (!hn) unit_normal(ccx, ccy, pcx, pcy, nrx, nry);
                            best = fill;
                            bnx = nrx;
                            bny = nry;
                            be = g.atoms[a][3];
                            have = true;
                        }
                    }
                }
        total = fmin(fmax(total, 0.0), 1.0);
        if (have) {
            const double fl = fmin(best, total);
            e = fl * be + (1.0 - fl) * g.eps_bg;
            ie = fl / be + (1.0 - fl) / g.eps_bg;
            if (fl < 1e-10) {
                e = g.eps_bg;
                ie = 1.0 / g.eps_bg;
            } else if (fl > 1.0 - 1e-10) {
                e = be;
                ie = 1.0 / be;
            }
            const bool interface = fl > 1e-6 && fl < 1.0 - 1e-6;
            if (interface && fl >= 1e-10 && fl <= 1.0 - 1e-10) {
                const double it = 1.0 / e, dl = ie - it;
                t[0] = it + dl * bny * bny;
                t[1] = -dl * bnx * bny;
                t[2] = -dl * bnx * bny;
                t[3] = it + dl * bnx * bnx;
            } else {
                t[0] = ie; t[1] = 0.0; t[2] = 0.0; t[3] = ie;
            }
        } else {
            e = g.eps_bg;
            ie = 1.0 / g.eps_bg;
            t[0] = ie; t[1] = 0.0; t[2] = 0.0; t[3] = ie;
        }
    }
    if (eps) eps[o] = e;
    if (inv_eps) inv_eps[o] = ie;
    if (tensor) {
#pragma unroll
        for (int q = 0; q < 4; ++q) tensor[4 * o + q] = t[q];
    }
}

}  // namespace

void launch_dielectric(const DielGeom *host_geoms, int ngeom, int nx, int ny, bool smoothing, double *d_eps,
                       double *d_inv_eps, double *d_tensor, void *d_geoms, hipStream_t st, hipEvent_t ev0,
                       hipEvent_t ev1) {
    if (ngeom <= 0) return;
    std::vector<DevGeom> g(ngeom);
    for (int k = 0; k < ngeom; ++k) {
        const DielGeom &h = host_geoms[k];
        if (h.n_atoms < 0 || h.n_atoms > DIEL_MAX_ATOMS) throw std::runtime_error("dielectric: too many atoms");
        DevGeom &d = g[k];
        d = DevGeom{};
        for (int c = 0; c < 2; ++c) {
            d.a1[c] = h.a1[c];
            d.a2[c] = h.a2[c];
        }
        d.eps_bg = h.eps_bg;
        d.n_atoms = h.n_atoms;
        for (int a = 0; a < h.n_atoms; ++a)
            for (int c = 0; c < 4; ++c) d.atoms[a][c] = h.atoms[a][c];
    }
    BH_HIP_CHECK(hipMemcpyAsync(d_geoms, g.data(), sizeof(DevGeom) * ngeom, hipMemcpyHostToDevice, st));
    const int n = nx * ny;
    if (ev0) BH_HIP_CHECK(hipEventRecord(ev0, st));
    hipLaunchKernelGGL(dielectric_kernel, dim3((n + 255) / 256, ngeom), dim3(256), 0, st,
                       (const DevGeom *)d_geoms, nx, ny, smoothing ? 1 : 0, d_eps, d_inv_eps, d_tensor);
    BH_HIP_CHECK(hipGetLastError());
    if (ev1) BH_HIP_CHECK(hipEventRecord(ev1, st));
    // the host staging vector must outlive the copy
    BH_HIP_CHECK(hipStreamSynchronize(st));
}

size_t dielectric_geom_bytes(int ngeom) { return sizeof(DevGeom) * (size_t)ngeom; }

}  // namespace bh
