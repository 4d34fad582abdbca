// Host-side setup and small dense algebra for the band-solve engine.
#include "hostmath.h"

#include <algorithm>
#include <cmath>
#include <functional>
#include <limits>

namespace bh {

namespace {
constexpr double PI = 3.14159265358979323846;

void frac_to_cart(const GeometrySpec &g, const double f[2], double c[2]) {
    c[0] = g.a1[0] * f[0] + g.a2[0] * f[1];
    c[1] = g.a1[1] * f[0] + g.a2[1] * f[1];
}
double char_len(const GeometrySpec &g) { return std::sqrt(g.a1[0] * g.a1[0] + g.a1[1] * g.a1[1]); }

// geometry.rs:59-119 node sampling with 3x3 periodic images
double sample_eps(const GeometrySpec &g, const double fin[2]) {
    auto wrap_unit = [](double v) { return v - std::floor(v); };
    auto wrap_signed = [](double d) {
        while (d >= 0.5) d -= 1.0;
        while (d < -0.5) d += 1.0;
        return d;
    };
    double f[2] = {wrap_unit(fin[0]), wrap_unit(fin[1])};
    for (const AtomSpec &a : g.atoms) {
        double r = a.radius * char_len(g);
        double dx = f[0] - a.pos[0], dy = f[1] - a.pos[1];
        for (int n1 = -1; n1 <= 1; ++n1)
            for (int n2 = -1; n2 <= 1; ++n2) {
                double df[2] = {wrap_signed(dx) + n1, wrap_signed(dy) + n2}, dc[2];
                frac_to_cart(g, df, dc);
                if (dc[0] * dc[0] + dc[1] * dc[1] <= r * r) return a.eps_inside;
            }
    }
    return g.eps_bg;
}

void unit_normal(const double c[2], const double p[2], double n[2]) {
    double dx = p[0] - c[0], dy = p[1] - c[1], d = std::sqrt(dx * dx + dy * dy);
    if (d < 1e-12) { n[0] = 1; n[1] = 0; }
    else { n[0] = dx / d; n[1] = dy / d; }
}

// analytic_geometry.rs circle_rectangle_intersection: returns fill fraction,
// sets has_normal / normal for partial pixels.
double circle_pixel(const double cc[2], double radius, const double rc[2], const double rs[2], bool &has_normal,
                    double normal[2]) {
    has_normal = false;
    const double hw = rs[0] / 2.0, hh = rs[1] / 2.0, area_px = rs[0] * rs[1];
    const double x0 = rc[0] - hw, x1 = rc[0] + hw, y0 = rc[1] - hh, y1 = rc[1] + hh;
    const double r2 = radius * radius;
    const double cx[4] = {x0, x1, x0, x1}, cy[4] = {y0, y0, y1, y1};
    int inside = 0;
    double far2 = 0;
    for (int i = 0; i < 4; ++i) {
        double d2 = (cx[i] - cc[0]) * (cx[i] - cc[0]) + (cy[i] - cc[1]) * (cy[i] - cc[1]);
        inside += d2 <= r2;
        far2 = std::max(far2, d2);
    }
    if (inside == 4) return 1.0;
    double px = std::min(std::max(cc[0], x0), x1) - cc[0], py = std::min(std::max(cc[1], y0), y1) - cc[1];
    if (px * px + py * py > r2) return 0.0;
    if (far2 <= r2) return 1.0;
    const int ns = 16;
    const double sw = rs[0] / ns, sh = rs[1] / ns, sa = sw * sh;
    double area = 0;
    for (int iy = 0; iy < ns; ++iy)
        for (int ix = 0; ix < ns; ++ix) {
            double sx = x0 + (ix + 0.5) * sw - cc[0], sy = y0 + (iy + 0.5) * sh - cc[1];
            if (sx * sx + sy * sy <= r2) area += sa;
        }
    double f = std::min(std::max(area / area_px, 0.0), 1.0);
    if (f < 1e-10) return 0.0;
    if (f > 1.0 - 1e-10) return 1.0;
    has_normal = true;
    unit_normal(cc, rc, normal);
    return f;
}
}  // namespace

void reciprocal(const GeometrySpec &g, double b1[2], double b2[2]) {
    double det = g.a1[0] * g.a2[1] - g.a1[1] * g.a2[0];
    double inv = 2.0 * PI / det;
    b1[0] = g.a2[1] * inv; b1[1] = -g.a2[0] * inv;
    b2[0] = -g.a1[1] * inv; b2[1] = g.a1[0] * inv;
}

DielectricTables build_dielectric(const GeometrySpec &g, int nx, int ny, bool smoothing) {
    DielectricTables d;
    d.nx = nx; d.ny = ny;
    reciprocal(g, d.b1, d.b2);
    const size_t n = (size_t)nx * ny;
    d.eps.assign(n, g.eps_bg);
    d.inv_eps.assign(n, 1.0 / g.eps_bg);
    d.tensor.assign(4 * n, 0.0);
    if (!smoothing) {
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
                double f[2] = {(double)ix / nx, (double)iy / ny};
                double e = sample_eps(g, f);
                size_t i = (size_t)iy * nx + ix;
                d.eps[i] = e; d.inv_eps[i] = 1.0 / e;
                d.tensor[4 * i] = d.tensor[4 * i + 3] = 1.0 / e;
            }
    } else {
        struct Circ { double c[2], r, eps; };
        std::vector<Circ> circles;
        for (const AtomSpec &a : g.atoms)
            for (int di = -1; di <= 1; ++di)
                for (int dj = -1; dj <= 1; ++dj) {
                    double f[2] = {a.pos[0] + di, a.pos[1] + dj};
                    Circ c;
                    frac_to_cart(g, f, c.c);
                    c.r = a.radius * char_len(g);
                    c.eps = a.eps_inside;
                    circles.push_back(c);
                }
        const double ps[2] = {std::sqrt(g.a1[0] * g.a1[0] + g.a1[1] * g.a1[1]) / nx,
                              std::sqrt(g.a2[0] * g.a2[0] + g.a2[1] * g.a2[1]) / ny};
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
                size_t i = (size_t)iy * nx + ix;
                double f[2] = {(double)ix / nx, (double)iy / ny}, pc[2];
                frac_to_cart(g, f, pc);
                bool have = false;
                double best = 0, total = 0, bn[2] = {0, 0}, be = 0;
                for (const Circ &c : circles) {
                    bool hn;
                    double nrm[2];
                    double fill = circle_pixel(c.c, c.r, pc, ps, hn, nrm);
                    if (fill > 1e-10) {
                        total += fill;
                        if (!have || fill > best) {
                            if (!hn) unit_normal(c.c, pc, nrm);
                            best = fill; bn[0] = nrm[0]; bn[1] = nrm[1]; be = c.eps;
                            have = true;
                        }
                    }
                }
                total = std::min(std::max(total, 0.0), 1.0);
                double e, ie, t[4];
                if (have) {
                    // analytic_geometry.rs compute_smoothed_dielectric
                    double fl = std::min(best, total);
                    e = fl * be + (1.0 - fl) * g.eps_bg;
                    ie = fl / be + (1.0 - fl) / g.eps_bg;
                    if (fl < 1e-10) { e = g.eps_bg; ie = 1.0 / g.eps_bg; }
                    else if (fl > 1.0 - 1e-10) { e = be; ie = 1.0 / be; }
                    bool interface = fl > 1e-6 && fl < 1.0 - 1e-6;
                    if (interface && fl >= 1e-10 && fl <= 1.0 - 1e-10) {
                        double it = 1.0 / e, dl = ie - it;
                        t[0] = it + dl * bn[1] * bn[1];
                        t[1] = -dl * bn[0] * bn[1];
                        t[2] = -dl * bn[0] * bn[1];
                        t[3] = it + dl * bn[0] * bn[0];
                    } else {
                        t[0] = ie; t[1] = 0; t[2] = 0; t[3] = ie;
                    }
                } else {
                    e = g.eps_bg; ie = 1.0 / g.eps_bg;
                    t[0] = ie; t[1] = 0; t[2] = 0; t[3] = ie;
                }
                d.eps[i] = e; d.inv_eps[i] = ie;
                for (int q = 0; q < 4; ++q) d.tensor[4 * i + q] = t[q];
            }
    }
    double s = 0, si = 0;
    for (size_t i = 0; i < n; ++i) { s += d.eps[i]; si += d.inv_eps[i]; }
    d.eps_sum = s;
    d.eps_mean = s / n;
    d.eps_harm = si > 0 ? 1.0 / (si / n) : 1.0;
    return d;
}

double spectral_smin(int nx, int ny, const double b1[2], const double b2[2], const double bloch[2]) {
    double smin = std::numeric_limits<double>::infinity();
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix < nx; ++ix) {
            double n1 = ix <= nx / 2 ? ix : ix - nx, n2 = iy <= ny / 2 ? iy : iy - ny;
            double rx = n1 * b1[0] + n2 * b2[0] + bloch[0], ry = n1 * b1[1] + n2 * b2[1] + bloch[1];
            double raw = rx * rx + ry * ry, ksq;
            if (raw >= 1e-9) ksq = raw;
            else ksq = 1e-9;  // clamped magnitude sqrt(1e-9)
            if (ksq > 1e-6 && std::isfinite(ksq)) smin = std::min(smin, ksq);
        }
    return std::isfinite(smin) ? smin : 1.0;
}

double kpg_sq_min(int nx, int ny, const double b1[2], const double b2[2], const double bloch[2]) {
    const double floor_sq = 1e-9;
    double m = std::numeric_limits<double>::infinity();
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix < nx; ++ix) {
            const double n1 = ix <= nx / 2 ? ix : ix - nx, n2 = iy <= ny / 2 ? iy : iy - ny;
            double kx = n1 * b1[0] + n2 * b2[0] + bloch[0], ky = n1 * b1[1] + n2 * b2[1] + bloch[1];
            if (!std::isfinite(kx) || !std::isfinite(ky)) {
                kx = std::sqrt(floor_sq);
                ky = 0.0;
            } else {
                const double sq = kx * kx + ky * ky;
                if (sq == 0.0) {
                    kx = std::sqrt(floor_sq);
                    ky = 0.0;
                } else if (sq < floor_sq) {
                    const double sc = std::sqrt(floor_sq / sq);
                    kx *= sc;
                    ky *= sc;
                }
            }
            m = std::min(m, kx * kx + ky * ky);
        }
    return std::isfinite(m) ? m : 0.0;
}

}  // namespace bh
