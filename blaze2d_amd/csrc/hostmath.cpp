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

void eigh(int n, const std::vector<zd> &a, std::vector<double> &w, std::vector<zd> &V) {
    // A from the lower triangle, col-major
    std::vector<zd> A((size_t)n * n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            if (i > j) A[i + (size_t)j * n] = a[i + (size_t)j * n];
            else if (i < j) A[i + (size_t)j * n] = std::conj(a[j + (size_t)i * n]);
            else A[i + (size_t)i * n] = a[i + (size_t)i * n].real();
        }
    auto at = [&](int i, int j) -> zd & { return A[i + (size_t)j * n]; };
    // 1. Householder reduction to tridiagonal form, A = Q T Q^H.  Step k maps
    // x = A[k+1:, k] to alpha e_1 with the Hermitian reflector H = I - 2 u u^H,
    // u = (x - alpha e_1) / |x - alpha e_1|, alpha = -|x| x_0 / |x_0|; the
    // trailing block B becomes H B H = B - u q^H - q u^H with p = 2 B u,
    // q = p - (u^H p) u, and Q collects Q H_k.
    std::vector<zd> Q((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) Q[i + (size_t)i * n] = 1.0;
    std::vector<zd> u(n), p(n);
    for (int k = 0; k + 2 < n; ++k) {
        const int r0 = k + 1;
        double tail = 0.0;
        for (int i = r0 + 1; i < n; ++i) tail += std::norm(at(i, k));
        if (tail == 0.0) continue;  // already reduced; the off-diagonal stays complex
        const zd x0 = at(r0, k);
        const double ax0 = std::abs(x0), xn = std::sqrt(ax0 * ax0 + tail);
        const zd ph = ax0 > 0.0 ? x0 / ax0 : zd(1.0);
        const zd alpha = -xn * ph;
        // |x - alpha e_1|^2 = 2 |x| (|x| + |x_0|)
        const double un = 1.0 / std::sqrt(2.0 * xn * (xn + ax0));
        u[r0] = (x0 - alpha) * un;
        for (int i = r0 + 1; i < n; ++i) u[i] = at(i, k) * un;
        for (int i = r0; i < n; ++i) {
            zd s = 0.0;
            for (int j = r0; j < n; ++j) s += at(i, j) * u[j];
            p[i] = 2.0 * s;
        }
        double c = 0.0;  // u^H p, real
        for (int i = r0; i < n; ++i) c += (std::conj(u[i]) * p[i]).real();
        for (int i = r0; i < n; ++i) p[i] -= c * u[i];
        for (int j = r0; j < n; ++j) {
            const zd uj = std::conj(u[j]), pj = std::conj(p[j]);
            for (int i = r0; i < n; ++i) at(i, j) -= u[i] * pj + p[i] * uj;
        }
        at(r0, k) = alpha;
        at(k, r0) = std::conj(alpha);
        for (int i = r0 + 1; i < n; ++i) at(i, k) = at(k, i) = 0.0;
        for (int i = 0; i < n; ++i) {  // Q[:, r0:] -= 2 (Q[:, r0:] u) u^H
            zd s = 0.0;
            for (int j = r0; j < n; ++j) s += Q[i + (size_t)j * n] * u[j];
            s *= 2.0;
            for (int j = r0; j < n; ++j) Q[i + (size_t)j * n] -= s * std::conj(u[j]);
        }
    }
    // 2. T = D T' D^H with T' real symmetric: ph_0 = 1, ph_{k+1} = ph_k e_k / |e_k|
    std::vector<double> d(n), e(n, 0.0);
    std::vector<zd> phase(n, 1.0);
    for (int k = 0; k < n; ++k) d[k] = at(k, k).real();
    for (int k = 0; k + 1 < n; ++k) {
        const zd ek = at(k + 1, k);
        const double ae = std::abs(ek);
        e[k] = ae;
        phase[k + 1] = ae > 0.0 ? phase[k] * (ek / ae) : phase[k];
    }
    // 3. implicit-shift QL on T' (Wilkinson shift from the leading 2 x 2),
    // rotations accumulated in Z
    std::vector<double> Z((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) Z[i + (size_t)i * n] = 1.0;
    const double eps = std::numeric_limits<double>::epsilon();
    for (int l = 0; l < n; ++l) {
        for (int iter = 0; iter < 60; ++iter) {
            int m = l;
            for (; m + 1 < n; ++m)
                if (std::fabs(e[m]) <= eps * (std::fabs(d[m]) + std::fabs(d[m + 1]))) break;
            if (m == l) break;
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? r : -r));
            double s = 1.0, c = 1.0, shift = 0.0;
            bool underflow = false;
            for (int i = m - 1; i >= l; --i) {
                const double f = s * e[i], b = c * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {  // split: restart on the smaller problem
                    d[i + 1] -= shift;
                    e[m] = 0.0;
                    underflow = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - shift;
                r = (d[i] - g) * s + 2.0 * c * b;
                shift = s * r;
                d[i + 1] = g + shift;
                g = c * r - b;
                double *zi = Z.data() + (size_t)i * n, *zi1 = zi + n;
                for (int k = 0; k < n; ++k) {
                    const double zk1 = zi1[k], zk = zi[k];
                    zi1[k] = s * zk + c * zk1;
                    zi[k] = c * zk - s * zk1;
                }
            }
            if (underflow) continue;
            d[l] -= shift;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    // 4. ascending order; V = Q diag(phase) Z
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return d[x] < d[y]; });
    w.resize(n);
    V.assign((size_t)n * n, 0.0);
    for (int j = 0; j < n; ++j) {
        w[j] = d[idx[j]];
        const double *zj = Z.data() + (size_t)idx[j] * n;
        for (int k = 0; k < n; ++k) {
            const zd f = phase[k] * zj[k];
            if (f == 0.0) continue;
            const zd *qk = Q.data() + (size_t)k * n;
            for (int i = 0; i < n; ++i) V[i + (size_t)j * n] += qk[i] * f;
        }
    }
}

double polar(int m, const std::vector<zd> &O, std::vector<zd> &U, bool *full_rank) {
    std::vector<zd> OhO((size_t)m * m);
    for (int j = 0; j < m; ++j)
        for (int i = 0; i < m; ++i) {
            zd s = 0;
            for (int k = 0; k < m; ++k) s += std::conj(O[k + (size_t)i * m]) * O[k + (size_t)j * m];
            OhO[i + (size_t)j * m] = s;
        }
    std::vector<double> ev;
    std::vector<zd> V;
    eigh(m, OhO, ev, V);
    std::vector<double> sig(m);
    double smin = std::numeric_limits<double>::infinity(), smax = 0;
    for (int i = 0; i < m; ++i) {
        sig[i] = std::sqrt(std::max(ev[i], 0.0));
        smin = std::min(smin, sig[i]);
        smax = std::max(smax, sig[i]);
    }
    if (full_rank) *full_rank = true;
    std::vector<zd> W((size_t)m * m, 0.0);
    for (int j = m - 1; j >= 0; --j) {
        std::vector<zd> col(m, 0.0);
        if (sig[j] > 1e-13 * std::max(smax, 1e-300)) {
            for (int i = 0; i < m; ++i) {
                zd s = 0;
                for (int k = 0; k < m; ++k) s += O[i + (size_t)k * m] * V[k + (size_t)j * m];
                col[i] = s / sig[j];
            }
        } else {
            if (full_rank) *full_rank = false;
            for (int e = 0; e < m; ++e) {
                std::fill(col.begin(), col.end(), zd(0.0));
                col[e] = 1.0;
                for (int jj = j + 1; jj < m; ++jj) {
                    zd c = 0;
                    for (int i = 0; i < m; ++i) c += std::conj(W[i + (size_t)jj * m]) * col[i];
                    for (int i = 0; i < m; ++i) col[i] -= c * W[i + (size_t)jj * m];
                }
                double nn = 0;
                for (auto &v : col) nn += std::norm(v);
                if (nn > 0.25) {
                    nn = std::sqrt(nn);
                    for (auto &v : col) v /= nn;
                    break;
                }
            }
        }
        for (int i = 0; i < m; ++i) W[i + (size_t)j * m] = col[i];
    }
    U.assign((size_t)m * m, 0.0);
    for (int j = 0; j < m; ++j)
        for (int i = 0; i < m; ++i) {
            zd s = 0;
            for (int k = 0; k < m; ++k) s += W[i + (size_t)k * m] * std::conj(V[j + (size_t)k * m]);
            U[i + (size_t)j * m] = s;
        }
    return smin;
}

std::vector<int> hungarian(const std::vector<std::vector<double>> &w) {
    const int n = (int)w.size();
    if (n == 0) return {};
    if (n == 1) return {0};
    double mx = -std::numeric_limits<double>::infinity();
    for (auto &r : w)
        for (double v : r) mx = std::max(mx, v);
    const double INF = std::numeric_limits<double>::infinity();
    std::vector<double> u(n + 1, 0.0), v(n + 1, 0.0);
    std::vector<int> p(n + 1, 0), way(n + 1, 0);
    for (int i = 1; i <= n; ++i) {
        p[0] = i;
        int j0 = 0;
        std::vector<double> minv(n + 1, INF);
        std::vector<char> used(n + 1, 0);
        do {
            used[j0] = 1;
            int i0 = p[j0], j1 = 0;
            double delta = INF;
            for (int j = 1; j <= n; ++j)
                if (!used[j]) {
                    double cur = (mx - w[i0 - 1][j - 1]) - u[i0] - v[j];
                    if (cur < minv[j]) { minv[j] = cur; way[j] = j0; }
                    if (minv[j] < delta) { delta = minv[j]; j1 = j; }
                }
            for (int j = 0; j <= n; ++j) {
                if (used[j]) { u[p[j]] += delta; v[j] -= delta; }
                else minv[j] -= delta;
            }
            j0 = j1;
        } while (p[j0] != 0);
        do {
            int j1 = way[j0];
            p[j0] = p[j1];
            j0 = j1;
        } while (j0);
    }
    std::vector<int> perm(n);
    for (int j = 1; j <= n; ++j) perm[p[j] - 1] = j - 1;
    return perm;
}

std::vector<int> track_bands(int m, const std::vector<zd> &O, const std::vector<double> &pw,
                             const std::vector<double> &cw, std::vector<zd> *polar_u, double *polar_smin,
                             bool *polar_full) {
    std::vector<zd> U;
    bool full = false;
    const double smin = polar(m, O, U, &full);
    if (polar_u) *polar_u = U;
    if (polar_smin) *polar_smin = smin;
    if (polar_full) *polar_full = full;
    auto mag = [&](int i, int j) { return std::abs(U[i + (size_t)j * m]); };
    // detect_degenerate_bands_from_rotation
    std::vector<std::pair<int, int>> rb;
    if (m >= 2) {
        std::vector<char> amb(m);
        for (int i = 0; i < m; ++i) {
            std::vector<double> row(m);
            for (int j = 0; j < m; ++j) row[j] = mag(i, j);
            std::sort(row.begin(), row.end(), std::greater<double>());
            double ratio = row[0] > 1e-10 ? row[1] / row[0] : 0.0;
            amb[i] = row[0] < 0.7 || ratio > 0.5;
        }
        int bs = -1;
        for (int i = 0; i < m; ++i) {
            if (amb[i]) { if (bs < 0) bs = i; }
            else if (bs >= 0) { if (i - bs > 1) rb.push_back({bs, i - bs}); bs = -1; }
        }
        if (bs >= 0 && m - bs > 1) rb.push_back({bs, m - bs});
    }
    // split_blocks_by_frequency
    auto near = [](double a, double b) {
        double gap = std::fabs(a - b);
        if (gap <= 5e-3) return true;
        return gap / std::max(0.5 * (std::fabs(a) + std::fabs(b)), 1e-6) <= 0.015;
    };
    std::vector<std::pair<int, int>> db;
    for (auto &b : rb) {
        int st = b.first, sz = b.second;
        if (sz < 2 || st >= (int)pw.size()) continue;
        int en = std::min((int)pw.size(), st + sz);
        if (en - st < 2) continue;
        int cs = st, cl = 1;
        for (int i = st + 1; i < en; ++i) {
            if (near(pw[i - 1], pw[i])) ++cl;
            else { if (cl > 1) db.push_back({cs, cl}); cs = i; cl = 1; }
        }
        if (cl > 1) db.push_back({cs, cl});
    }
    std::vector<std::vector<double>> W(m, std::vector<double>(m));
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) W[i][j] = mag(i, j);
    std::vector<int> perm = hungarian(W);
    auto refine = [&](int st, int sz) {
        int en = st + sz;
        if (en > (int)perm.size() || en > (int)pw.size()) return;
        std::vector<int> ci(sz);
        for (int k = 0; k < sz; ++k) ci[k] = perm[st + k];
        for (int c : ci)
            if (c >= (int)cw.size()) return;
        std::vector<std::vector<double>> fw(sz, std::vector<double>(sz));
        for (int a = 0; a < sz; ++a)
            for (int b = 0; b < sz; ++b) fw[a][b] = -std::fabs(pw[st + a] - cw[ci[b]]);
        std::vector<int> lp = hungarian(fw);
        for (int li = 0; li < (int)lp.size(); ++li)
            if (li < sz && lp[li] < sz) perm[st + li] = ci[lp[li]];
    };
    for (auto &b : db) refine(b.first, b.second);
    // detect_frequency_mismatch_blocks
    std::vector<std::pair<int, int>> mb;
    int bs = -1;
    for (int i = 0; i < (int)perm.size(); ++i) {
        int c = perm[i];
        if (i >= (int)pw.size() || c >= (int)cw.size()) {
            if (bs >= 0) { if (i - bs > 1) mb.push_back({bs, i - bs}); bs = -1; }
            continue;
        }
        double mis = std::fabs(pw[i] - cw[c]);
        double sc = std::max(std::max(std::fabs(pw[i]), std::fabs(cw[c])), 1e-3);
        bool bad = mis > 5e-3 && mis / sc > 0.02;
        if (bad) { if (bs < 0) bs = i; }
        else if (bs >= 0) { if (i - bs > 1) mb.push_back({bs, i - bs}); bs = -1; }
    }
    if (bs >= 0 && (int)perm.size() - bs > 1) mb.push_back({bs, (int)perm.size() - bs});
    for (auto &b : mb) refine(b.first, b.second);
    return perm;
}

zd cdot(const zd *a, const zd *b, int n) {
    double re = 0.0, im = 0.0;
    for (int i = 0; i < n; ++i) {
        const double ar = a[i].real(), ai = a[i].imag(), br = b[i].real(), bi = b[i].imag();
        re += ar * br + ai * bi;
        im += ar * bi - ai * br;
    }
    return zd(re, im);
}

void CoefSpace::apply(const zd *v, zd *out) const {
    const std::vector<zd> &G = *M;
    std::vector<double> re(kd, 0.0), im(kd, 0.0);
    for (int j = 0; j < kd; ++j) {
        const double vr = v[j].real(), vi = v[j].imag();
        if (vr == 0.0 && vi == 0.0) continue;
        const zd *col = G.data() + (size_t)j * kd;
        for (int i = 0; i < kd; ++i) {
            const double gr = col[i].real(), gi = col[i].imag();
            re[i] += gr * vr - gi * vi;
            im[i] += gr * vi + gi * vr;
        }
    }
    for (int i = 0; i < kd; ++i) out[i] = zd(re[i], im[i]);
}

zd CoefSpace::inner(const zd *a, const zd *b) const {
    std::vector<zd> mb(kd);
    apply(b, mb.data());
    return cdot(a, mb.data(), kd);
}

}  // namespace bh
