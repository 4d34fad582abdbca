// C ABI of the gfx950 band-solve backend (include/blaze2d_hip.h).
#include "../../include/blaze2d_hip.h"

#include <chrono>
#include <cmath>
#include <cstring>
#include <exception>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "aux_kernels.h"
#include "blas.h"
#include "cnn.h"
#include "diagnostics.h"
#include "dielectric.h"
#include "engine.h"
#include "fft_ops.h"
#include "hostmath.h"

using namespace bh;

struct bh_context {
    int device = 0;
    hipStream_t st{};
    // band-solve slot groups run on st and these; created right after st so
    // that with HIP's default of 4 hardware queues per process each of four
    // groups has a queue of its own (HIP shares queues once GPU_MAX_HW_QUEUES
    // streams exist)
    hipStream_t gst[3]{};
    // twiddle tables per (n, precision)
    std::vector<std::pair<int, void *>> tw64, tw32;
    // digit-reversal tables per line length (multi-pass operators)
    std::vector<std::pair<int, int *>> rev;
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    // device address of the file-scope allocation-free scratch array on this
    // context's device (set once at context creation)
    char *noalloc_scratch = nullptr;
};

// Scratch of bh_gram_noalloc, the Gram entry point that allocates nothing:
// the work descriptor, the per-chunk partial sums (at most 64 chunks of a
// 48 x 48 complex block), their reduction and the packed p x q result.  One
// copy per device, held until the process ends; every call on a device shares
// it, so calls that overlap on one device race.
constexpr int NOALLOC_CHUNKS = 64;
constexpr size_t NOALLOC_BLOCK = sizeof(double) * 2 * GRAM_MAXC * GRAM_MAXC;
constexpr size_t NOALLOC_SCRATCH_BYTES = 256 + (NOALLOC_CHUNKS + 2) * NOALLOC_BLOCK;  // 2.4 MB
static __device__ __align__(256) char scratch[NOALLOC_SCRATCH_BYTES];

// out[i q + j] = red[i + j ld] (row-major p x q complex from the col-major reduction)
__global__ void gram_pack_kernel(const double *__restrict__ red, int ld, int p, int q, double *__restrict__ out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= p * q) return;
    const int i = e / q, j = e - i * q;
    out[2 * e] = red[2 * (i + (size_t)j * ld)];
    out[2 * e + 1] = red[2 * (i + (size_t)j * ld) + 1];
}

struct bh_buffer {
    bh_context *ctx;
    int nx, ny, precision, count;
    void *data;
    size_t n() const { return (size_t)nx * ny; }
    size_t esize() const { return precision == 1 ? sizeof(cplx<float>) : sizeof(cplx<double>); }
    char *col(int i) const { return (char *)data + (size_t)i * n() * esize(); }
};

struct bh_cnn {
    bh_context *ctx;
    Cnn *net;
};

struct bh_theta {
    bh_context *ctx;
    int nx, ny, pol, precond, precision;
    OpParams params{};
    OpParams *d_params = nullptr;
    void *d_eps = nullptr, *d_ten = nullptr;
};

namespace {
thread_local std::string g_err;

int fail(const std::string &m) {
    g_err = m;
    return -1;
}

#define BH_API_TRY try {
#define BH_API_CATCH \
    }                \
    catch (const std::exception &e) { return fail(e.what()); }

template <typename T>
const cplx<T> *twiddles(bh_context *ctx, int n) {
    auto &tab = std::is_same<T, double>::value ? ctx->tw64 : ctx->tw32;
    for (auto &p : tab)
        if (p.first == n) return (const cplx<T> *)p.second;
    std::vector<cplx<T>> h(n);
    for (int t = 0; t < n; ++t) {
        double a = -2.0 * M_PI * t / n;
        h[t] = {(T)std::cos(a), (T)std::sin(a)};
    }
    void *d = nullptr;
    BH_HIP_CHECK(hipMalloc(&d, sizeof(cplx<T>) * n));
    BH_HIP_CHECK(hipMemcpy(d, h.data(), sizeof(cplx<T>) * n, hipMemcpyHostToDevice));
    tab.push_back({n, d});
    return (const cplx<T> *)d;
}

void *ctx_scratch(bh_context *ctx, size_t bytes) {
    if (bytes > ctx->scratch_bytes) {
        if (ctx->scratch) BH_HIP_CHECK(hipFree(ctx->scratch));
        BH_HIP_CHECK(hipMalloc(&ctx->scratch, bytes));
        ctx->scratch_bytes = bytes;
    }
    return ctx->scratch;
}

const int *revtab(bh_context *ctx, int n) {
    for (auto &p : ctx->rev)
        if (p.first == n) return p.second;
    const std::vector<int> h = digit_reversal(n);
    int *d = nullptr;
    BH_HIP_CHECK(hipMalloc((void **)&d, sizeof(int) * n));
    BH_HIP_CHECK(hipMemcpy(d, h.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    ctx->rev.push_back({n, d});
    return d;
}

template <typename T>
void run_op(bh_context *ctx, int mode, int nx, int ny, const void *in, void *out, int count, const OpParams *dparams,
            const void *eps, const void *ten, void *bout) {
    OpWork<T> w{};
    w.in = (const cplx<T> *)in;
    w.out = (cplx<T> *)out;
    w.bout = (cplx<T> *)bout;
    w.eps = (const T *)eps;
    w.tensor = (const T *)ten;
    w.params = dparams;
    w.in_stride = (long)nx * ny;
    w.out_stride = (long)nx * ny;
    w.count = count;
    // multi-pass TE / TP: a second field per column behind the descriptor
    const bool two = grid_multipass(nx, ny) && (mode == MODE_THETA_TE || mode == MODE_PREC_TP);
    char *base = (char *)ctx_scratch(ctx, 256 + (two ? (size_t)count * nx * ny * sizeof(cplx<T>) : 0));
    OpWork<T> *dw = (OpWork<T> *)base;
    if (two) {
        w.tmp = (cplx<T> *)(base + 256);
        w.tmp_stride = (long)nx * ny;
    }
    BH_HIP_CHECK(hipMemcpyAsync(dw, &w, sizeof(w), hipMemcpyHostToDevice, ctx->st));
    launch_op<T>(mode, nx, ny, dw, 1, count, twiddles<T>(ctx, nx), twiddles<T>(ctx, ny), ctx->st, false,
                 revtab(ctx, nx), revtab(ctx, ny));
    BH_HIP_CHECK(hipStreamSynchronize(ctx->st));
}

template <typename T>
void run_gram(bh_context *ctx, const bh_buffer *x, int xi0, int p, const bh_buffer *y, int yi0, int q,
              double *out) {
    const long n = (long)x->n();
    const int nch = (int)std::min<long>(64, std::max<long>(1, n / 64));
    const int ld = GRAM_MAXC;
    char *base = (char *)ctx_scratch(ctx, sizeof(GramWork<T>) + 256 + sizeof(double) * 2 * ld * ld * (nch + 1));
    GramWork<T> w{};
    w.A = (const cplx<T> *)x->col(xi0);
    w.B = (const cplx<T> *)y->col(yi0);
    w.a_stride = n;
    w.b_stride = n;
    w.na = p;
    w.nb = q;
    w.ld = ld;
    double *part = (double *)(base + 256);
    double *red = part + 2 * (size_t)ld * ld * nch;
    w.partial = part;
    BH_HIP_CHECK(hipMemcpyAsync(base, &w, sizeof(w), hipMemcpyHostToDevice, ctx->st));
    launch_gram<T>((const GramWork<T> *)base, 1, nch, n, p, q, ctx->st);
    launch_gram_reduce(part, red, 1, nch, ld, 0, 0, ctx->st);
    std::vector<double> h(2 * (size_t)ld * ld);
    BH_HIP_CHECK(hipMemcpyAsync(h.data(), red, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx->st));
    BH_HIP_CHECK(hipStreamSynchronize(ctx->st));
    for (int i = 0; i < p; ++i)
        for (int j = 0; j < q; ++j) {
            out[2 * ((size_t)i * q + j)] = h[2 * (i + (size_t)j * ld)];
            out[2 * ((size_t)i * q + j) + 1] = h[2 * (i + (size_t)j * ld) + 1];
        }
}

template <typename T>
void run_lincomb(bh_context *ctx, const bh_buffer *qb, int qi0, int r, const double *coeffs, int m, bh_buffer *out,
                 int oi0) {
    const long n = (long)qb->n();
    char *base = (char *)ctx_scratch(ctx, 256 + sizeof(double) * 2 * (size_t)r * m);
    LincombWork<T> w{};
    w.A = (const cplx<T> *)qb->col(qi0);
    w.a_stride = n;
    w.na = r;
    w.nq = m;
    w.C = (const double *)(base + 256);
    w.ldc = r;
    w.out = (cplx<T> *)out->col(oi0);
    w.out_stride = n;
    BH_HIP_CHECK(hipMemcpyAsync(base, &w, sizeof(w), hipMemcpyHostToDevice, ctx->st));
    BH_HIP_CHECK(hipMemcpyAsync(base + 256, coeffs, sizeof(double) * 2 * (size_t)r * m, hipMemcpyHostToDevice, ctx->st));
    launch_lincomb<T>((const LincombWork<T> *)base, 1, n, m, ctx->st);
    BH_HIP_CHECK(hipStreamSynchronize(ctx->st));
}

bool check_idx(const bh_buffer *b, int first, int count) {
    return b && first >= 0 && count >= 0 && first + count <= b->count;
}

GeometrySpec to_geom(const bh_geometry &g) {
    GeometrySpec s;
    s.a1[0] = g.a1[0]; s.a1[1] = g.a1[1];
    s.a2[0] = g.a2[0]; s.a2[1] = g.a2[1];
    s.eps_bg = g.eps_bg;
    for (int i = 0; i < g.n_atoms && i < BH_MAX_ATOMS; ++i)
        s.atoms.push_back({{g.atoms[i][0], g.atoms[i][1]}, g.atoms[i][2], g.atoms[i][3]});
    return s;
}
}  // namespace

extern "C" {

const char *bh_version(void) { return "blaze2d_hip 0.1 (gfx950)"; }
const char *bh_last_error(void) { return g_err.c_str(); }

int bh_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int bh_context_create(int device, bh_context **out) {
    BH_API_TRY
    if (!out) return fail("null output");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= device) return fail("no HIP device available");
    BH_HIP_CHECK(hipSetDevice(device));
    auto *c = new bh_context();
    c->device = device;
    BH_HIP_CHECK(hipStreamCreate(&c->st));
    for (auto &g : c->gst) BH_HIP_CHECK(hipStreamCreate(&g));
    // the allocation-free entry points use the device address of `scratch` on
    // this device (the host-side name is only a stand-in)
    void *sp = nullptr;
    const hipError_t se = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
    if (se != hipSuccess) {
        for (auto &g : c->gst) (void)hipStreamDestroy(g);
        (void)hipStreamDestroy(c->st);
        delete c;
        return fail(std::string("bh_context_create: scratch symbol address: ") + hipGetErrorString(se));
    }
    c->noalloc_scratch = (char *)sp;
    *out = c;
    return 0;
    BH_API_CATCH
}

void bh_context_destroy(bh_context *ctx) {
    if (!ctx) return;
    (void)hipStreamSynchronize(ctx->st);
    for (auto &p : ctx->tw64) (void)hipFree(p.second);
    for (auto &p : ctx->tw32) (void)hipFree(p.second);
    for (auto &p : ctx->rev) (void)hipFree(p.second);
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    for (auto &g : ctx->gst) {
        (void)hipStreamSynchronize(g);
        (void)hipStreamDestroy(g);
    }
    (void)hipStreamDestroy(ctx->st);
    delete ctx;
}

int bh_synchronize(bh_context *ctx) {
    BH_API_TRY
    BH_HIP_CHECK(hipStreamSynchronize(ctx->st));
    return 0;
    BH_API_CATCH
}

int bh_buffer_alloc(bh_context *ctx, int nx, int ny, int precision, int count, bh_buffer **out) {
    BH_API_TRY
    if (!ctx || !out || nx <= 0 || ny <= 0 || count < 0 || (precision != 0 && precision != 1))
        return fail("bh_buffer_alloc: bad arguments");
    auto *b = new bh_buffer{ctx, nx, ny, precision, count, nullptr};
    size_t bytes = std::max<size_t>(1, b->n() * count * b->esize());
    BH_HIP_CHECK(hipMalloc(&b->data, bytes));
    BH_HIP_CHECK(hipMemset(b->data, 0, bytes));
    *out = b;
    return 0;
    BH_API_CATCH
}

void bh_buffer_free(bh_buffer *buf) {
    if (!buf) return;
    (void)hipFree(buf->data);
    delete buf;
}

int bh_buffer_upload(bh_buffer *b, int first, int count, const double *host) {
    BH_API_TRY
    if (!check_idx(b, first, count) || !host) return fail("bh_buffer_upload: bad range");
    size_t n = b->n() * count;
    if (b->precision == 0) {
        BH_HIP_CHECK(hipMemcpy(b->col(first), host, n * 16, hipMemcpyHostToDevice));
    } else {
        std::vector<float> f(2 * n);
        for (size_t i = 0; i < 2 * n; ++i) f[i] = (float)host[i];
        BH_HIP_CHECK(hipMemcpy(b->col(first), f.data(), n * 8, hipMemcpyHostToDevice));
    }
    return 0;
    BH_API_CATCH
}

int bh_buffer_download(const bh_buffer *b, int first, int count, double *host) {
    BH_API_TRY
    if (!check_idx(b, first, count) || !host) return fail("bh_buffer_download: bad range");
    BH_HIP_CHECK(hipStreamSynchronize(b->ctx->st));
    size_t n = b->n() * count;
    if (b->precision == 0) {
        BH_HIP_CHECK(hipMemcpy(host, b->col(first), n * 16, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> f(2 * n);
        BH_HIP_CHECK(hipMemcpy(f.data(), b->col(first), n * 8, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < 2 * n; ++i) host[i] = f[i];
    }
    return 0;
    BH_API_CATCH
}

int bh_fft2d(bh_context *ctx, bh_buffer *b, int first, int count, int direction) {
    BH_API_TRY
    if (!check_idx(b, first, count)) return fail("bh_fft2d: bad range");
    if (!grid_supported(b->nx, b->ny)) return fail("bh_fft2d: grid size not instantiated");
    if (count == 0) return 0;
    int mode = direction > 0 ? MODE_FFT_INV : MODE_FFT_FWD;
    // the FFT kernel reads all input before writing, but output may alias input only per column
    if (b->precision == 0) run_op<double>(ctx, mode, b->nx, b->ny, b->col(first), b->col(first), count, nullptr, nullptr, nullptr, nullptr);
    else run_op<float>(ctx, mode, b->nx, b->ny, b->col(first), b->col(first), count, nullptr, nullptr, nullptr, nullptr);
    return 0;
    BH_API_CATCH
}

int bh_scale(bh_context *ctx, bh_buffer *x, int ix, double re, double im) {
    BH_API_TRY
    if (!check_idx(x, ix, 1)) return fail("bh_scale: bad index");
    if (x->precision == 0) launch_scale<double>((cplx<double> *)x->col(ix), (long)x->n(), re, im, ctx->st);
    else launch_scale<float>((cplx<float> *)x->col(ix), (long)x->n(), re, im, ctx->st);
    BH_HIP_CHECK(hipStreamSynchronize(ctx->st));
    return 0;
    BH_API_CATCH
}

int bh_axpy(bh_context *ctx, double re, double im, const bh_buffer *x, int ix, bh_buffer *y, int iy) {
    BH_API_TRY
    if (!check_idx(x, ix, 1) || !check_idx(y, iy, 1) || x->n() != y->n() || x->precision != y->precision)
        return fail("bh_axpy: bad arguments");
    if (x->precision == 0)
        launch_axpy<double>((const cplx<double> *)x->col(ix), (cplx<double> *)y->col(iy), (long)x->n(), re, im, ctx->st);
    else
        launch_axpy<float>((const cplx<float> *)x->col(ix), (cplx<float> *)y->col(iy), (long)x->n(), re, im, ctx->st);
    BH_HIP_CHECK(hipStreamSynchronize(ctx->st));
    return 0;
    BH_API_CATCH
}

int bh_dot(bh_context *ctx, const bh_buffer *x, int ix, const bh_buffer *y, int iy, double out[2]) {
    BH_API_TRY
    if (!check_idx(x, ix, 1) || !check_idx(y, iy, 1) || x->n() != y->n() || x->precision != y->precision || !out)
        return fail("bh_dot: bad arguments");
    double *d = (double *)ctx_scratch(ctx, 16);
    if (x->precision == 0)
        launch_dot<double>((const cplx<double> *)x->col(ix), (const cplx<double> *)y->col(iy), (long)x->n(), d, ctx->st);
    else
        launch_dot<float>((const cplx<float> *)x->col(ix), (const cplx<float> *)y->col(iy), (long)x->n(), d, ctx->st);
    BH_HIP_CHECK(hipMemcpyAsync(out, d, 16, hipMemcpyDeviceToHost, ctx->st));
    BH_HIP_CHECK(hipStreamSynchronize(ctx->st));
    return 0;
    BH_API_CATCH
}

int bh_gram(bh_context *ctx, const bh_buffer *x, int xi0, int p, const bh_buffer *y, int yi0, int q, double *out) {
    BH_API_TRY
    if (!check_idx(x, xi0, p) || !check_idx(y, yi0, q) || x->n() != y->n() || x->precision != y->precision || !out)
        return fail("bh_gram: bad arguments");
    if (p > GRAM_MAXC || q > GRAM_MAXC) return fail("bh_gram: at most 48 columns per operand");
    if (p == 0 || q == 0) return 0;
    if (x->precision == 0) run_gram<double>(ctx, x, xi0, p, y, yi0, q, out);
    else run_gram<float>(ctx, x, xi0, p, y, yi0, q, out);
    return 0;
    BH_API_CATCH
}

int bh_gram_noalloc(bh_context *ctx, const bh_buffer *x, int xi0, int p, const bh_buffer *y, int yi0, int q,
                    double *out) {
    BH_API_TRY
    if (!ctx || !check_idx(x, xi0, p) || !check_idx(y, yi0, q) || x->n() != y->n() || x->precision != y->precision ||
        !out)
        return fail("bh_gram_noalloc: bad arguments");
    if (p > GRAM_MAXC || q > GRAM_MAXC) return fail("bh_gram_noalloc: at most 48 columns per operand");
    if (!ctx->noalloc_scratch) return fail("bh_gram_noalloc: context has no scratch address");
    if (p == 0 || q == 0) return 0;
    const long n = (long)x->n();
    const int nch = (int)std::min<long>(NOALLOC_CHUNKS, std::max<long>(1, n / 64));  // as bh_gram
    const int ld = GRAM_MAXC;
    char *base = ctx->noalloc_scratch;
    double *part = (double *)(base + 256);
    double *red = (double *)((char *)part + (size_t)nch * NOALLOC_BLOCK);
    double *packed = (double *)((char *)red + NOALLOC_BLOCK);
    if ((size_t)((char *)packed - base) + sizeof(double) * 2 * p * q > NOALLOC_SCRATCH_BYTES)
        return fail("bh_gram_noalloc: problem exceeds the 2.4 MB scratch");
    // every partial entry read below (i < p, j < q of each chunk) is written by
    // the Gram kernel, so the shared scratch needs no clearing between calls
    auto go = [&](auto tag) {
        using T = decltype(tag);
        GramWork<T> w{};
        w.A = (const cplx<T> *)x->col(xi0);
        w.B = (const cplx<T> *)y->col(yi0);
        w.a_stride = n;
        w.b_stride = n;
        w.na = p;
        w.nb = q;
        w.ld = ld;
        w.partial = part;
        BH_HIP_CHECK(hipMemcpyAsync(base, &w, sizeof(w), hipMemcpyHostToDevice, ctx->st));
        launch_gram<T>((const GramWork<T> *)base, 1, nch, n, p, q, ctx->st);
    };
    if (x->precision == 0) go(double());
    else go(float());
    launch_gram_reduce(part, red, 1, nch, ld, 0, 0, ctx->st);
    hipLaunchKernelGGL(gram_pack_kernel, dim3((p * q + 255) / 256), dim3(256), 0, ctx->st, red, ld, p, q, packed);
    BH_HIP_CHECK(hipGetLastError());
    BH_HIP_CHECK(hipMemcpyAsync(out, packed, sizeof(double) * 2 * p * q, hipMemcpyDeviceToHost, ctx->st));
    BH_HIP_CHECK(hipStreamSynchronize(ctx->st));
    return 0;
    BH_API_CATCH
}

int bh_lincomb(bh_context *ctx, const bh_buffer *qb, int qi0, int r, const double *coeffs, int m, bh_buffer *out,
               int oi0) {
    BH_API_TRY
    if (!check_idx(qb, qi0, r) || !check_idx(out, oi0, m) || qb->n() != out->n() || qb->precision != out->precision ||
        !coeffs)
        return fail("bh_lincomb: bad arguments");
    if (r > GRAM_MAXC) return fail("bh_lincomb: at most 48 input columns");
    if (qb == out && !(oi0 >= qi0 + r || oi0 + m <= qi0)) return fail("bh_lincomb: output overlaps input");
    if (m == 0) return 0;
    if (qb->precision == 0) run_lincomb<double>(ctx, qb, qi0, r, coeffs, m, out, oi0);
    else run_lincomb<float>(ctx, qb, qi0, r, coeffs, m, out, oi0);
    return 0;
    BH_API_CATCH
}

int bh_hermitian_eigh(bh_context *ctx, int n, int count, const double *a, double *w, double *v) {
    BH_API_TRY
    if (!ctx || n <= 0 || n > EIGH_MAXN || count < 0 || !a || !w || !v) return fail("bh_hermitian_eigh: bad arguments");
    if (count == 0) return 0;
    const size_t nn = (size_t)n * n;
    // device layout: [work items][partials (row-major G(i,j))][eigenvalues][eigenvectors]
    size_t bytes = 256 + count * sizeof(EighWork) + count * (2 * nn + n + 2 * nn) * sizeof(double);
    char *base = (char *)ctx_scratch(ctx, bytes);
    EighWork *dw = (EighWork *)(base + 256);
    double *dpart = (double *)(dw + count);
    double *dwv = dpart + 2 * nn * count;
    double *dv = dwv + (size_t)n * count;
    std::vector<double> hp(2 * nn * count);
    for (int c = 0; c < count; ++c)
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {  // row-major G(i, j) = A(i, j)
                hp[2 * (c * nn + (size_t)i * n + j)] = a[2 * (c * nn + i + (size_t)j * n)];
                hp[2 * (c * nn + (size_t)i * n + j) + 1] = a[2 * (c * nn + i + (size_t)j * n) + 1];
            }
    std::vector<EighWork> hw(count);
    for (int c = 0; c < count; ++c) {
        EighWork e{};
        e.partial = dpart + 2 * nn * c;
        e.w = dwv + (size_t)n * c;
        e.V = dv + 2 * nn * c;
        e.n = n;
        e.nchunks = 1;
        e.ld = n;
        e.mode = 0;
        e.ldv = n;
        e.nsplit = n;
        hw[c] = e;
    }
    BH_HIP_CHECK(hipMemcpyAsync(dw, hw.data(), sizeof(EighWork) * count, hipMemcpyHostToDevice, ctx->st));
    BH_HIP_CHECK(hipMemcpyAsync(dpart, hp.data(), sizeof(double) * hp.size(), hipMemcpyHostToDevice, ctx->st));
    launch_eigh(dw, count, n, ctx->st);
    BH_HIP_CHECK(hipMemcpyAsync(w, dwv, sizeof(double) * n * count, hipMemcpyDeviceToHost, ctx->st));
    BH_HIP_CHECK(hipMemcpyAsync(v, dv, sizeof(double) * 2 * nn * count, hipMemcpyDeviceToHost, ctx->st));
    BH_HIP_CHECK(hipStreamSynchronize(ctx->st));
    return 0;
    BH_API_CATCH
}

int bh_rayleigh_ritz(bh_context *ctx, int n, int count, int nsplit, const double *g1, const double *g2, double *lam,
                     double *coef, int *info) {
    BH_API_TRY
    if (!ctx || n <= 0 || n > 96 || count < 0 || nsplit <= 0 || !g1 || !g2 || !lam || !coef || !info)
        return fail("bh_rayleigh_ritz: bad arguments (n must be 1..96)");
    if (count == 0) return 0;
    BH_HIP_CHECK(hipSetDevice(ctx->device));
    const size_t nn = (size_t)n * n;
    // device layout: [work items][G1 partials][G2 partials][lam][C][counts], each
    // partial row-major G(i, j) with one chunk
    const size_t wbytes = ((count * sizeof(RRWork) + 255) / 256) * 256;
    const size_t wsb = n > GRAM_MAXC ? rr_wide_bytes(n) : 0;  // wide kernel workspace per problem
    const size_t bytes = wbytes + sizeof(double) * count * (2 * nn * 3 + n) + sizeof(int) * 8 * count + 16 + wsb * count;
    char *base = (char *)ctx_scratch(ctx, bytes);
    RRWork *dw = (RRWork *)base;
    double *dp1 = (double *)(base + wbytes);
    double *dp2 = dp1 + 2 * nn * count;
    double *dc = dp2 + 2 * nn * count;
    double *dl = dc + 2 * nn * count;
    int *dcnt = (int *)(dl + (size_t)n * count);
    double *dws = (double *)(((uintptr_t)(dcnt + 8 * count) + 15) & ~(uintptr_t)15);
    std::vector<double> h1(2 * nn * count), h2(2 * nn * count);
    for (int c = 0; c < count; ++c)
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                const size_t src = 2 * (c * nn + i + (size_t)j * n), dst = 2 * (c * nn + (size_t)i * n + j);
                h1[dst] = g1[src];
                h1[dst + 1] = g1[src + 1];
                h2[dst] = g2[src];
                h2[dst + 1] = g2[src + 1];
            }
    std::vector<RRWork> hw(count);
    for (int c = 0; c < count; ++c) {
        RRWork r{};
        r.partial1 = dp1 + 2 * nn * c;
        r.partial2 = dp2 + 2 * nn * c;
        r.lam = dl + (size_t)n * c;
        r.C = dc + 2 * nn * c;
        r.cnt = dcnt + 8 * c;
        r.n = n;
        r.nchunks = 1;
        r.ld = n;
        r.ldc = n;
        r.nsplit = nsplit;
        if (wsb) r.ws = dws + (wsb / sizeof(double)) * c;
        hw[c] = r;
    }
    BH_HIP_CHECK(hipMemsetAsync(dl, 0, sizeof(double) * n * count, ctx->st));
    BH_HIP_CHECK(hipMemcpyAsync(dw, hw.data(), sizeof(RRWork) * count, hipMemcpyHostToDevice, ctx->st));
    BH_HIP_CHECK(hipMemcpyAsync(dp1, h1.data(), sizeof(double) * h1.size(), hipMemcpyHostToDevice, ctx->st));
    BH_HIP_CHECK(hipMemcpyAsync(dp2, h2.data(), sizeof(double) * h2.size(), hipMemcpyHostToDevice, ctx->st));
    launch_rr(dw, count, n, ctx->st);
    std::vector<int> hc(8 * (size_t)count);
    BH_HIP_CHECK(hipMemcpyAsync(lam, dl, sizeof(double) * n * count, hipMemcpyDeviceToHost, ctx->st));
    BH_HIP_CHECK(hipMemcpyAsync(coef, dc, sizeof(double) * 2 * nn * count, hipMemcpyDeviceToHost, ctx->st));
    BH_HIP_CHECK(hipMemcpyAsync(hc.data(), dcnt, sizeof(int) * hc.size(), hipMemcpyDeviceToHost, ctx->st));
    BH_HIP_CHECK(hipStreamSynchronize(ctx->st));
    for (int c = 0; c < count; ++c) {
        const int *k = hc.data() + 8 * c;
        info[4 * c] = k[5];
        info[4 * c + 1] = k[4];
        info[4 * c + 2] = k[0];
        info[4 * c + 3] = k[1];
        // columns past the Ritz coefficients held the B-orthonormalising
        // transform while the kernel ran
        for (size_t e = 2 * ((size_t)(k[0] + k[1]) * n); e < 2 * nn; ++e) coef[2 * c * nn + e] = 0.0;
    }
    return 0;
    BH_API_CATCH
}

int bh_theta_create(bh_context *ctx, const bh_theta_desc *d, bh_theta **out) {
    BH_API_TRY
    if (!ctx || !d || !out || !d->eps) return fail("bh_theta_create: bad arguments");
    if (d->pol != 0 && d->pol != 1) return fail("bh_theta_create: pol must be 0 (TM) or 1 (TE)");
    if (d->precond < -1 || d->precond > 2) return fail("bh_theta_create: precond must be -1..2");
    if (d->precision != 0 && d->precision != 1) return fail("bh_theta_create: precision must be 0 or 1");
    if (!grid_supported(d->nx, d->ny)) return fail("bh_theta_create: grid size not instantiated");
    auto *op = new bh_theta();
    op->ctx = ctx;
    op->nx = d->nx;
    op->ny = d->ny;
    op->pol = d->pol;
    op->precision = d->precision;
    op->precond = d->precond < 0 ? (d->pol == 0 ? 1 : 2) : d->precond;
    const size_t n = (size_t)d->nx * d->ny;
    OpParams &P = op->params;
    for (int i = 0; i < 2; ++i) { P.b1[i] = d->b1[i]; P.b2[i] = d->b2[i]; P.bloch[i] = d->bloch[i]; }
    P.pol = d->pol;
    P.gamma = (std::fabs(P.bloch[0]) < 1e-12 && std::fabs(P.bloch[1]) < 1e-12) ? 1 : 0;
    P.precond = op->precond;
    P.shift = 0.5 * spectral_smin(d->nx, d->ny, P.b1, P.b2, P.bloch);
    double s = 0, si = 0;
    for (size_t i = 0; i < n; ++i) { s += d->eps[i]; si += 1.0 / d->eps[i]; }
    if (d->pol == 0) {
        P.eps_eff = s / n;
        P.mass_floor = (std::isfinite(P.eps_eff) && P.eps_eff > 0) ? P.eps_eff * 1e-2 : 0.0;
    } else {
        P.eps_eff = si > 0 ? 1.0 / (si / n) : 1.0;
        P.mass_floor = 0.0;
    }
    std::vector<double> ten(4 * n);
    for (size_t i = 0; i < n; ++i) {
        if (d->tensor) for (int t = 0; t < 4; ++t) ten[4 * i + t] = d->tensor[4 * i + t];
        else { ten[4 * i] = ten[4 * i + 3] = 1.0 / d->eps[i]; ten[4 * i + 1] = ten[4 * i + 2] = 0.0; }
    }
    BH_HIP_CHECK(hipMalloc((void **)&op->d_params, sizeof(OpParams)));
    BH_HIP_CHECK(hipMemcpy(op->d_params, &P, sizeof(OpParams), hipMemcpyHostToDevice));
    if (op->precision == 0) {
        BH_HIP_CHECK(hipMalloc(&op->d_eps, n * 8));
        BH_HIP_CHECK(hipMalloc(&op->d_ten, 4 * n * 8));
        BH_HIP_CHECK(hipMemcpy(op->d_eps, d->eps, n * 8, hipMemcpyHostToDevice));
        BH_HIP_CHECK(hipMemcpy(op->d_ten, ten.data(), 4 * n * 8, hipMemcpyHostToDevice));
    } else {
        std::vector<float> fe(n), ft(4 * n);
        for (size_t i = 0; i < n; ++i) fe[i] = (float)d->eps[i];
        for (size_t i = 0; i < 4 * n; ++i) ft[i] = (float)ten[i];
        BH_HIP_CHECK(hipMalloc(&op->d_eps, n * 4));
        BH_HIP_CHECK(hipMalloc(&op->d_ten, 4 * n * 4));
        BH_HIP_CHECK(hipMemcpy(op->d_eps, fe.data(), n * 4, hipMemcpyHostToDevice));
        BH_HIP_CHECK(hipMemcpy(op->d_ten, ft.data(), 4 * n * 4, hipMemcpyHostToDevice));
    }
    *out = op;
    return 0;
    BH_API_CATCH
}

void bh_theta_destroy(bh_theta *op) {
    if (!op) return;
    (void)hipFree(op->d_params);
    (void)hipFree(op->d_eps);
    (void)hipFree(op->d_ten);
    delete op;
}

static int theta_call(bh_theta *op, int mode, const bh_buffer *in, int ii0, bh_buffer *out, int oi0, int count) {
    if (!op || !check_idx(in, ii0, count) || !check_idx(out, oi0, count)) return fail("operator: bad range");
    if (in->nx != op->nx || in->ny != op->ny || out->nx != op->nx || out->ny != op->ny ||
        in->precision != op->precision || out->precision != op->precision)
        return fail("operator: buffer shape or precision mismatch");
    if (count == 0) return 0;
    if (op->precision == 0)
        run_op<double>(op->ctx, mode, op->nx, op->ny, in->col(ii0), out->col(oi0), count, op->d_params, op->d_eps,
                       op->d_ten, nullptr);
    else
        run_op<float>(op->ctx, mode, op->nx, op->ny, in->col(ii0), out->col(oi0), count, op->d_params, op->d_eps,
                      op->d_ten, nullptr);
    return 0;
}

int bh_theta_apply(bh_theta *op, const bh_buffer *in, int ii0, bh_buffer *out, int oi0, int count) {
    BH_API_TRY
    return theta_call(op, op->pol == 0 ? MODE_THETA_TM : MODE_THETA_TE, in, ii0, out, oi0, count);
    BH_API_CATCH
}

int bh_mass_apply(bh_theta *op, const bh_buffer *in, int ii0, bh_buffer *out, int oi0, int count) {
    BH_API_TRY
    if (!op || !check_idx(in, ii0, count) || !check_idx(out, oi0, count)) return fail("bh_mass_apply: bad range");
    if (in->nx != op->nx || in->ny != op->ny || out->nx != op->nx || out->ny != op->ny ||
        in->precision != op->precision || out->precision != op->precision)
        return fail("bh_mass_apply: buffer shape or precision mismatch");
    const long n = (long)in->n();
    // identity coefficients with a pointwise eps image (TM) via the lincomb kernel
    for (int c = 0; c < count; ++c) {
        double one[2] = {1.0, 0.0};
        if (op->pol == 1) {
            BH_HIP_CHECK(hipMemcpyAsync(out->col(oi0 + c), in->col(ii0 + c), n * in->esize(), hipMemcpyDeviceToDevice,
                                        op->ctx->st));
            continue;
        }
        char *base = (char *)ctx_scratch(op->ctx, 512);
        auto go = [&](auto tag) {
            using T = decltype(tag);
            LincombWork<T> w{};
            w.A = (const cplx<T> *)in->col(ii0 + c);
            w.a_stride = n;
            w.na = 1;
            w.nq = 1;
            w.C = (const double *)(base + 256);
            w.ldc = 1;
            // write the identity copy into the output and eps * copy into the same output
            w.out = (cplx<T> *)out->col(oi0 + c);
            w.out2 = (cplx<T> *)out->col(oi0 + c);
            w.w2 = (const T *)op->d_eps;
            w.out_stride = n;
            BH_HIP_CHECK(hipMemcpyAsync(base, &w, sizeof(w), hipMemcpyHostToDevice, op->ctx->st));
            BH_HIP_CHECK(hipMemcpyAsync(base + 256, one, 16, hipMemcpyHostToDevice, op->ctx->st));
            launch_lincomb<T>((const LincombWork<T> *)base, 1, n, 1, op->ctx->st);
            BH_HIP_CHECK(hipStreamSynchronize(op->ctx->st));
        };
        if (op->precision == 0) go(double());
        else go(float());
    }
    BH_HIP_CHECK(hipStreamSynchronize(op->ctx->st));
    return 0;
    BH_API_CATCH
}

int bh_precond_apply(bh_theta *op, bh_buffer *buf, int first, int count) {
    BH_API_TRY
    if (!op) return fail("bh_precond_apply: null operator");
    if (op->precond == 0) return 0;
    return theta_call(op, op->precond == 1 ? MODE_PREC_FD : MODE_PREC_TP, buf, first, buf, first, count);
    BH_API_CATCH
}

int bh_dielectric(const bh_geometry *geom, int nx, int ny, int smoothing, double *eps, double *inv_eps,
                  double *tensor) {
    BH_API_TRY
    if (!geom || nx <= 0 || ny <= 0) return fail("bh_dielectric: bad arguments");
    if (geom->n_atoms < 0 || geom->n_atoms > BH_MAX_ATOMS) return fail("bh_dielectric: n_atoms out of range");
    DielectricTables d = build_dielectric(to_geom(*geom), nx, ny, smoothing != 0);
    size_t n = (size_t)nx * ny;
    if (eps) std::memcpy(eps, d.eps.data(), n * 8);
    if (inv_eps) std::memcpy(inv_eps, d.inv_eps.data(), n * 8);
    if (tensor) std::memcpy(tensor, d.tensor.data(), 4 * n * 8);
    return 0;
    BH_API_CATCH
}

int bh_dielectric_batch(bh_context *ctx, const bh_geometry *geoms, int ngeom, int nx, int ny, int smoothing,
                        double *eps, double *inv_eps, double *tensor, double *gpu_ms) {
    BH_API_TRY
    if (!ctx || !geoms || ngeom <= 0 || nx <= 0 || ny <= 0) return fail("bh_dielectric_batch: bad arguments");
    if ((long)nx * ny > (1L << 26)) return fail("bh_dielectric_batch: grid too large");
    std::vector<DielGeom> hg(ngeom);
    for (int k = 0; k < ngeom; ++k) {
        const bh_geometry &g = geoms[k];
        if (g.n_atoms < 0 || g.n_atoms > BH_MAX_ATOMS) return fail("bh_dielectric_batch: n_atoms out of range");
        DielGeom &d = hg[k];
        d = DielGeom{};
        for (int c = 0; c < 2; ++c) {
            d.a1[c] = g.a1[c];
            d.a2[c] = g.a2[c];
        }
        d.eps_bg = g.eps_bg;
        d.n_atoms = g.n_atoms;
        for (int a = 0; a < g.n_atoms; ++a)
            for (int c = 0; c < 4; ++c) d.atoms[a][c] = g.atoms[a][c];
    }
    BH_HIP_CHECK(hipSetDevice(ctx->device));
    const size_t n = (size_t)nx * ny * ngeom;
    double *d_eps = nullptr, *d_inv = nullptr, *d_ten = nullptr;
    void *d_geo = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto release = [&] {
        if (d_eps) (void)hipFree(d_eps);
        if (d_inv) (void)hipFree(d_inv);
        if (d_ten) (void)hipFree(d_ten);
        if (d_geo) (void)hipFree(d_geo);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    };
    try {
        if (eps) BH_HIP_CHECK(hipMalloc(&d_eps, n * sizeof(double)));
        if (inv_eps) BH_HIP_CHECK(hipMalloc(&d_inv, n * sizeof(double)));
        if (tensor) BH_HIP_CHECK(hipMalloc(&d_ten, 4 * n * sizeof(double)));
        BH_HIP_CHECK(hipMalloc(&d_geo, dielectric_geom_bytes(ngeom)));
        BH_HIP_CHECK(hipEventCreate(&e0));
        BH_HIP_CHECK(hipEventCreate(&e1));
        launch_dielectric(hg.data(), ngeom, nx, ny, smoothing != 0, d_eps, d_inv, d_ten, d_geo, ctx->st, e0, e1);
        if (gpu_ms) {
            float ms = 0.0f;
            BH_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            *gpu_ms = ms;
        }
        if (eps) BH_HIP_CHECK(hipMemcpy(eps, d_eps, n * sizeof(double), hipMemcpyDeviceToHost));
        if (inv_eps) BH_HIP_CHECK(hipMemcpy(inv_eps, d_inv, n * sizeof(double), hipMemcpyDeviceToHost));
        if (tensor) BH_HIP_CHECK(hipMemcpy(tensor, d_ten, 4 * n * sizeof(double), hipMemcpyDeviceToHost));
    } catch (...) {
        release();
        throw;
    }
    release();
    return 0;
    BH_API_CATCH
}

DielGeom to_diel_geom(const bh_geometry &g) {
    DielGeom d{};
    for (int c = 0; c < 2; ++c) {
        d.a1[c] = g.a1[c];
        d.a2[c] = g.a2[c];
    }
    d.eps_bg = g.eps_bg;
    d.n_atoms = g.n_atoms;
    for (int a = 0; a < g.n_atoms; ++a)
        for (int c = 0; c < 4; ++c) d.atoms[a][c] = g.atoms[a][c];
    return d;
}

int bh_operator_diagnostics(bh_context *ctx, const bh_job *jobs, int njobs, int power_iterations,
                            bh_diagnostics *out, double *gpu_ms) {
    BH_API_TRY
    if (!ctx || !jobs || njobs <= 0 || !out) return fail("bh_operator_diagnostics: bad arguments");
    if (power_iterations < 1 || power_iterations > 10000)
        return fail("bh_operator_diagnostics: power_iterations out of range");
    const int nx = jobs[0].nx, ny = jobs[0].ny;
    for (int j = 0; j < njobs; ++j) {
        if (jobs[j].nx != nx || jobs[j].ny != ny) return fail("bh_operator_diagnostics: jobs must share the grid");
        if (jobs[j].geom.n_atoms < 0 || jobs[j].geom.n_atoms > BH_MAX_ATOMS)
            return fail("bh_operator_diagnostics: n_atoms out of range");
        if (jobs[j].pol != 0 && jobs[j].pol != 1) return fail("bh_operator_diagnostics: pol must be 0 or 1");
        if (jobs[j].nk < 0 || (jobs[j].nk > 0 && !jobs[j].kpath)) return fail("bh_operator_diagnostics: bad k-path");
    }
    if (!grid_supported(nx, ny)) return fail("bh_operator_diagnostics: grid size not instantiated");
    BH_HIP_CHECK(hipSetDevice(ctx->device));
    const long n = (long)nx * ny;
    // jobs with a non-Gamma k-point, grouped by (polarisation, preconditioner,
    // smoothing) so that each launch has one operator mode
    struct Act {
        int job, k, pol, prec, smooth;
    };
    std::vector<Act> act;
    for (int j = 0; j < njobs; ++j) {
        const bh_job &J = jobs[j];
        out[j] = bh_diagnostics{};
        out[j].k_index = -1;
        for (int k = 0; k < J.nk; ++k) {
            const double kx = J.kpath[2 * k], ky = J.kpath[2 * k + 1];
            if (kx * kx + ky * ky >= 1e-8) {
                out[j].k_index = k;
                break;
            }
        }
        if (out[j].k_index < 0) continue;
        const int prec = J.precond < 0 ? (J.pol == 0 ? 1 : 2) : J.precond;
        if (prec < 0 || prec > 2) return fail("bh_operator_diagnostics: unknown preconditioner");
        out[j].precond = prec;
        act.push_back({j, out[j].k_index, J.pol, prec, J.smoothing ? 1 : 0});
    }
    std::stable_sort(act.begin(), act.end(), [](const Act &a, const Act &b) {
        if (a.pol != b.pol) return a.pol < b.pol;
        if (a.prec != b.prec) return a.prec < b.prec;
        return a.smooth < b.smooth;
    });
    // jobs per chunk: four fields, eps and tensor per job, about 1 GiB at most
    const long per_job = n * (6 * 16 + 5 * 8);
    const int chunk = (int)std::max<long>(1, std::min<long>(4096, (1L << 30) / per_job));
    const cplx<double> *twx = twiddles<double>(ctx, nx), *twy = twiddles<double>(ctx, ny);
    const hipStream_t st = ctx->st;
    double *d_f = nullptr, *d_eps = nullptr, *d_ten = nullptr, *d_res = nullptr;
    void *d_geo = nullptr, *d_meta = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto release = [&] {
        for (void *p : {(void *)d_f, (void *)d_eps, (void *)d_ten, (void *)d_res, d_geo, d_meta})
            if (p) (void)hipFree(p);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    };
    try {
        const int cap = (int)std::min<size_t>(act.size(), (size_t)chunk);
        if (cap > 0) {
            // four fields per job, then two second-field columns for multi-pass TE / TP
            BH_HIP_CHECK(hipMalloc((void **)&d_f, (size_t)cap * 6 * n * 16));
            BH_HIP_CHECK(hipMalloc((void **)&d_eps, (size_t)cap * n * 8));
            BH_HIP_CHECK(hipMalloc((void **)&d_ten, (size_t)cap * n * 32));
            BH_HIP_CHECK(hipMalloc((void **)&d_res, (size_t)cap * 3 * 8));
            BH_HIP_CHECK(hipMalloc(&d_geo, dielectric_geom_bytes(cap)));
            // params, three OpWork lists (self-adjoint; power step from column 0 or 1),
            // the preconditioner lists and four DiagItem lists
            BH_HIP_CHECK(hipMalloc(&d_meta, (size_t)cap * (sizeof(OpParams) + 5 * sizeof(OpWork<double>) +
                                                           4 * sizeof(DiagItem))));
        }
        BH_HIP_CHECK(hipEventCreate(&e0));
        BH_HIP_CHECK(hipEventCreate(&e1));
        double total_ms = 0.0;
        for (size_t c0 = 0; c0 < act.size();) {
            // a chunk: same polarisation and preconditioner, at most cap jobs
            size_t c1 = c0 + 1;
            while (c1 < act.size() && c1 - c0 < (size_t)cap && act[c1].pol == act[c0].pol &&
                   act[c1].prec == act[c0].prec)
                ++c1;
            const int m = (int)(c1 - c0), pol = act[c0].pol, prec = act[c0].prec;
            // host parameters
            std::vector<OpParams> hp(m);
            std::vector<double> lam_min(m);
            for (int i = 0; i < m; ++i) {
                const bh_job &J = jobs[act[c0 + i].job];
                const GeometrySpec g = to_geom(J.geom);
                double b1[2], b2[2];
                reciprocal(g, b1, b2);
                const double kx = J.kpath[2 * act[c0 + i].k], ky = J.kpath[2 * act[c0 + i].k + 1];
                OpParams &P = hp[i];
                P = OpParams{};
                for (int c = 0; c < 2; ++c) {
                    P.b1[c] = b1[c];
                    P.b2[c] = b2[c];
                    P.bloch[c] = kx * b1[c] + ky * b2[c];
                }
                P.pol = pol;
                P.gamma = (std::fabs(P.bloch[0]) < 1e-12 && std::fabs(P.bloch[1]) < 1e-12) ? 1 : 0;
                P.precond = prec;
                P.shift = 0.5 * spectral_smin(nx, ny, P.b1, P.b2, P.bloch);
                lam_min[i] = kpg_sq_min(nx, ny, P.b1, P.b2, P.bloch);
            }
            char *meta = (char *)d_meta;
            OpParams *d_par = (OpParams *)meta;
            OpWork<double> *w_sa = (OpWork<double> *)(d_par + cap);
            OpWork<double> *w_p0 = w_sa + cap, *w_p1 = w_p0 + cap;  // Theta from column 0 / 1
            OpWork<double> *w_m0 = w_p1 + cap, *w_m1 = w_m0 + cap;  // preconditioner on column 1 / 0
            DiagItem *i_sa = (DiagItem *)(w_m1 + cap);
            DiagItem *i_n0 = i_sa + cap, *i_s0 = i_n0 + cap, *i_s1 = i_s0 + cap;
            std::vector<OpWork<double>> hw(5 * (size_t)m);
            std::vector<DiagItem> hi(4 * (size_t)m);
            auto col = [&](int i, int c) { return (cplx<double> *)d_f + ((size_t)i * 6 + c) * n; };
            for (int i = 0; i < m; ++i) {
                OpWork<double> w{};
                w.eps = d_eps + (size_t)i * n;
                w.tensor = d_ten + (size_t)i * n * 4;
                w.params = d_par + i;
                w.tmp = col(i, 4);
                w.tmp_stride = n;
                w.in_stride = n;
                w.out_stride = n;
                w.count = 1;
                OpWork<double> sa = w;  // x, y in columns 0, 2 -> Ax, Ay in 1, 3
                sa.in = col(i, 0);
                sa.out = col(i, 1);
                sa.in_stride = sa.out_stride = 2 * n;
                sa.count = 2;
                OpWork<double> p0 = w, p1 = w, m0 = w, m1 = w;
                p0.in = col(i, 0);
                p0.out = col(i, 1);
                p1.in = col(i, 1);
                p1.out = col(i, 0);
                m0.in = m0.out = col(i, 1);
                m1.in = m1.out = col(i, 0);
                hw[i] = sa;
                hw[m + i] = p0;
                hw[2 * m + i] = p1;
                hw[3 * m + i] = m0;
                hw[4 * m + i] = m1;
                double *r = d_res + 3 * i;  // lambda, pm lambda, self-adjoint error
                hi[i] = DiagItem{col(i, 0), col(i, 1), col(i, 2), col(i, 3), r + 2};
                hi[m + i] = DiagItem{col(i, 0), nullptr, nullptr, nullptr, nullptr};
                hi[2 * m + i] = DiagItem{col(i, 0), col(i, 1), nullptr, nullptr, r};
                hi[3 * m + i] = DiagItem{col(i, 1), col(i, 0), nullptr, nullptr, r};
            }
            BH_HIP_CHECK(hipMemcpyAsync(d_par, hp.data(), sizeof(OpParams) * m, hipMemcpyHostToDevice, st));
            BH_HIP_CHECK(hipMemcpyAsync(w_sa, hw.data(), sizeof(OpWork<double>) * m, hipMemcpyHostToDevice, st));
            BH_HIP_CHECK(hipMemcpyAsync(w_p0, hw.data() + m, sizeof(OpWork<double>) * m, hipMemcpyHostToDevice, st));
            BH_HIP_CHECK(hipMemcpyAsync(w_p1, hw.data() + 2 * m, sizeof(OpWork<double>) * m, hipMemcpyHostToDevice, st));
            BH_HIP_CHECK(hipMemcpyAsync(w_m0, hw.data() + 3 * m, sizeof(OpWork<double>) * m, hipMemcpyHostToDevice, st));
            BH_HIP_CHECK(hipMemcpyAsync(w_m1, hw.data() + 4 * m, sizeof(OpWork<double>) * m, hipMemcpyHostToDevice, st));
            for (int l = 0; l < 4; ++l)  // the four lists sit cap entries apart
                BH_HIP_CHECK(hipMemcpyAsync(i_sa + (size_t)l * cap, hi.data() + (size_t)l * m, sizeof(DiagItem) * m,
                                            hipMemcpyHostToDevice, st));
            std::vector<DielGeom> hg(m);
            for (int i = 0; i < m; ++i) hg[i] = to_diel_geom(jobs[act[c0 + i].job].geom);
            BH_HIP_CHECK(hipEventRecord(e0, st));
            // dielectric tables, one launch per run of equal smoothing
            for (int a = 0; a < m;) {
                int b = a + 1;
                while (b < m && act[c0 + b].smooth == act[c0 + a].smooth) ++b;
                launch_dielectric(hg.data() + a, b - a, nx, ny, act[c0 + a].smooth != 0, d_eps + (size_t)a * n,
                                  nullptr, d_ten + (size_t)a * n * 4, d_geo, st);
                a = b;
            }
            launch_diag_params(d_par, d_eps, m, n, st);
            const int mode_a = pol == 0 ? MODE_THETA_TM : MODE_THETA_TE;
            const int mode_m = prec == 1 ? MODE_PREC_FD : MODE_PREC_TP;
            // self-adjointness: Ax, Ay
            launch_diag_init(col(0, 0), col(0, 2), 6 * n, m, n, st);
            const int *rvx = revtab(ctx, nx), *rvy = revtab(ctx, ny);
            launch_op<double>(mode_a, nx, ny, w_sa, m, 2, twx, twy, st, false, rvx, rvy);
            launch_diag_reduce(i_sa, m, 1, n, st);
            // power iterations on Theta, then on M^-1 Theta, each from the normalised x
            std::vector<DiagItem> h2(2 * (size_t)m);  // lives until the stream is synchronised
            for (int phase = 0; phase < (prec ? 2 : 1); ++phase) {
                if (phase == 1) launch_diag_init(col(0, 0), col(0, 2), 6 * n, m, n, st);
                launch_diag_reduce(i_n0, m, 2, n, st);
                for (int t = 0; t < power_iterations; ++t) {
                    const bool odd = t & 1;
                    launch_op<double>(mode_a, nx, ny, odd ? w_p1 : w_p0, m, 1, twx, twy, st, false, rvx, rvy);
                    if (phase == 1)
                        launch_op<double>(mode_m, nx, ny, odd ? w_m1 : w_m0, m, 1, twx, twy, st, false, rvx, rvy);
                    // lambda into slot 0 (Theta) or 1 (M^-1 Theta) of each job's results
                    launch_diag_reduce(odd ? i_s1 : i_s0, m, 0, n, st);
                }
                if (phase == 0 && prec) {
                    // keep Theta's lambda: the preconditioned pass writes slot 1
                    for (int i = 0; i < m; ++i) {
                        h2[i] = hi[2 * m + i];
                        h2[m + i] = hi[3 * m + i];
                        h2[i].out = h2[m + i].out = d_res + 3 * i + 1;
                    }
                    BH_HIP_CHECK(hipMemcpyAsync(i_s0, h2.data(), sizeof(DiagItem) * m, hipMemcpyHostToDevice, st));
                    BH_HIP_CHECK(
                        hipMemcpyAsync(i_s1, h2.data() + m, sizeof(DiagItem) * m, hipMemcpyHostToDevice, st));
                }
            }
            BH_HIP_CHECK(hipEventRecord(e1, st));
            std::vector<double> res(3 * (size_t)m);
            BH_HIP_CHECK(hipMemcpyAsync(res.data(), d_res, sizeof(double) * 3 * m, hipMemcpyDeviceToHost, st));
            BH_HIP_CHECK(hipStreamSynchronize(st));
            float ms = 0.0f;
            BH_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            total_ms += ms;
            for (int i = 0; i < m; ++i) {
                bh_diagnostics &o = out[act[c0 + i].job];
                o.lambda_max = res[3 * i];
                o.lambda_min = lam_min[i];
                o.kappa = lam_min[i] > 1e-15 ? o.lambda_max / lam_min[i] : INFINITY;
                o.pm_lambda_max = prec ? res[3 * i + 1] : 0.0;
                o.pm_kappa = o.pm_lambda_max;
                o.self_adjoint_err = res[3 * i + 2];
            }
            c0 = c1;
        }
        if (gpu_ms) *gpu_ms = total_ms;
    } catch (...) {
        release();
        throw;
    }
    release();
    return 0;
    BH_API_CATCH
}

int bh_bandstructure_run(bh_context *ctx, const bh_job *jobs, int njobs, const bh_run_options *opts, double *omegas,
                         int *iters, int *converged, bh_run_stats *stats) {
    return bh_bandstructure_run_ex(ctx, jobs, njobs, opts, omegas, iters, converged, nullptr, stats);
}

int bh_bandstructure_run_ex(bh_context *ctx, const bh_job *jobs, int njobs, const bh_run_options *opts,
                            double *omegas, int *iters, int *converged, const bh_run_outputs *extra,
                            bh_run_stats *stats) {
    BH_API_TRY
    double *rel_residuals = extra ? extra->rel_residuals : nullptr;
    int *rel_residual_iters = extra ? extra->rel_residual_iters : nullptr;
    const int pw_half = extra && extra->pw_coefs ? extra->pw_half : 0;
    if (pw_half < 0 || pw_half > 8) return fail("bh_bandstructure_run_ex: pw_half must be 0..8");
    if (!ctx || (!jobs && njobs > 0) || njobs < 0) return fail("bh_bandstructure_run: bad arguments");
    if (njobs == 0) return 0;
    std::vector<JobSpec> js(njobs);
    for (int j = 0; j < njobs; ++j) {
        const bh_job &b = jobs[j];
        if (b.nk <= 0 || !b.kpath) return fail("bh_bandstructure_run: empty k-path");
        if (b.n_bands <= 0) return fail("bh_bandstructure_run: n_bands must be positive");
        if (b.pol != 0 && b.pol != 1) return fail("bh_bandstructure_run: pol must be 0 (TM) or 1 (TE)");
        if (b.precond < -1 || b.precond > 2) return fail("bh_bandstructure_run: precond must be -1..2");
        if (b.exit_rule != 0 && b.exit_rule != 1) return fail("bh_bandstructure_run: exit_rule must be 0 or 1");
        if (b.geom.n_atoms < 0 || b.geom.n_atoms > BH_MAX_ATOMS)
            return fail("bh_bandstructure_run: n_atoms out of range");
        if (b.block_size < 0) return fail("bh_bandstructure_run: block_size must be >= 0");
        if (!grid_supported(b.nx, b.ny)) return fail("bh_bandstructure_run: grid size not instantiated");
        JobSpec &s = js[j];
        s.geom = to_geom(b.geom);
        s.nx = b.nx;
        s.ny = b.ny;
        s.pol = b.pol;
        s.n_bands = b.n_bands;
        s.max_iter = b.max_iter > 0 ? b.max_iter : 200;
        s.tol = b.tol > 0 ? b.tol : 1e-6;
        s.block_size = b.block_size;
        s.smoothing = b.smoothing;
        s.precond = b.precond;
        s.prediction = b.subspace_prediction;
        s.extrapolation = b.extrapolation;
        s.tracking = b.band_tracking;
        s.reuse_gamma = b.reuse_gamma;
        s.exit_rule = b.exit_rule;
        for (int k = 0; k < b.nk; ++k) s.kpath.push_back({b.kpath[2 * k], b.kpath[2 * k + 1]});
        if (b.warm_pw && b.warm_count > 0) {
            if (b.pw_half < 0 || b.pw_half > 8) return fail("bh_bandstructure_run: warm pw_half must be 0..8");
            const size_t ngw = (size_t)(2 * b.pw_half + 1) * (2 * b.pw_half + 1);
            s.warm_pw.assign(b.warm_pw, b.warm_pw + 2 * ngw * b.warm_count);
            s.warm_count = b.warm_count;
            s.pw_half = b.pw_half;
        }
    }
    EngineOptions eo;
    const int slots_total = opts && opts->slots > 0 ? opts->slots : 64;
    eo.precision = opts ? opts->precision : 0;
    if (eo.precision < 0 || eo.precision > 2) return fail("bh_bandstructure_run: precision must be 0, 1 or 2");
    eo.profile = opts ? opts->profile : 0;
    eo.groups = opts && opts->groups > 0 ? opts->groups : 4;
    eo.timed_kernel = opts ? opts->timed_kernel : 0;
    eo.pw_half = pw_half;
    if (eo.timed_kernel < 0 || eo.timed_kernel > 6) return fail("bh_bandstructure_run: timed_kernel must be 0..6");
    eo.gram_mfma_min = opts && opts->gram_mfma_min > 0 ? opts->gram_mfma_min : GRAM_MFMA_MIN_DEFAULT;
    const int engines = opts && opts->engines > 0 ? opts->engines : 1;
    if (engines > 16) return fail("bh_bandstructure_run: engines must be 1..16");
    // partitions: the jobs of each grid, then the largest split in two (job
    // order kept) until there are `engines` of them; each runs as one engine
    // with its own slots, streams and host thread
    struct Part {
        int nx, ny;
        std::vector<int> idx;
        double work = 0;
        int slots = 1, nbmax = 0;
    };
    std::vector<Part> parts;
    {
        std::map<std::pair<int, int>, std::vector<int>> by_grid;
        for (int j = 0; j < njobs; ++j) by_grid[{js[j].nx, js[j].ny}].push_back(j);
        for (auto &kv : by_grid) parts.push_back(Part{kv.first.first, kv.first.second, kv.second});
    }
    while ((int)parts.size() < engines) {
        int big = -1;
        for (int p = 0; p < (int)parts.size(); ++p)
            if (parts[p].idx.size() >= 2 && (big < 0 || parts[p].idx.size() > parts[big].idx.size())) big = p;
        if (big < 0) break;
        Part a{parts[big].nx, parts[big].ny, {}}, b{parts[big].nx, parts[big].ny, {}};
        for (size_t i = 0; i < parts[big].idx.size(); ++i) (i % 2 ? b : a).idx.push_back(parts[big].idx[i]);
        parts[big] = a;
        parts.push_back(b);
    }
    double work_total = 0;
    for (auto &p : parts) {
        for (int j : p.idx) {
            const JobSpec &s = js[j];
            p.work += (double)s.kpath.size() * s.nx * s.ny * s.n_bands;
            p.nbmax = std::max(p.nbmax, std::max(s.n_bands, s.block_size > 0 ? s.block_size - 2 : 0));
        }
        work_total += p.work;
    }
    for (auto &p : parts) {
        const double share = parts.size() == 1 ? slots_total : slots_total * p.work / std::max(work_total, 1.0);
        p.slots = std::max(1, std::min((int)p.idx.size(), (int)std::lround(share)));
    }
    BH_HIP_CHECK(hipSetDevice(ctx->device));
    std::vector<std::vector<JobResult>> pres(parts.size());
    std::vector<RunStats> pstats(parts.size());
    std::vector<std::exception_ptr> perr(parts.size());
    auto run_part = [&](int p) {
        try {
            BH_HIP_CHECK(hipSetDevice(ctx->device));
            EngineOptions o = eo;
            o.slots = parts[p].slots;
            if (p == 0) {  // the first engine on the context's streams
                o.setup_stream = ctx->st;
                o.group_streams[0] = ctx->st;
                o.group_streams[1] = ctx->gst[0];
                o.group_streams[2] = ctx->gst[1];
                o.group_streams[3] = ctx->gst[2];
            }
            std::vector<JobSpec> pj;
            pj.reserve(parts[p].idx.size());
            for (int j : parts[p].idx) pj.push_back(js[j]);
            auto eng = make_engine(parts[p].nx, parts[p].ny, parts[p].nbmax, o);
            eng->run(pj, pres[p], pstats[p]);
        } catch (...) {
            perr[p] = std::current_exception();
        }
    };
    const auto t0 = std::chrono::steady_clock::now();
    {
        std::vector<std::thread> th;
        for (int p = 1; p < (int)parts.size(); ++p) th.emplace_back(run_part, p);
        run_part(0);
        for (auto &t : th) t.join();
    }
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (auto &e : perr)
        if (e) std::rethrow_exception(e);
    std::vector<const JobResult *> res(njobs);
    for (size_t p = 0; p < parts.size(); ++p)
        for (size_t i = 0; i < parts[p].idx.size(); ++i) res[parts[p].idx[i]] = &pres[p][i];
    size_t off = 0, koff = 0;
    for (int j = 0; j < njobs; ++j) {
        const JobResult &r = *res[j];
        if (omegas) std::memcpy(omegas + off, r.omegas.data(), sizeof(double) * r.omegas.size());
        if (rel_residuals)
            std::memcpy(rel_residuals + off, r.rel_residuals.data(), sizeof(double) * r.rel_residuals.size());
        if (extra && extra->prec_residuals)
            std::memcpy(extra->prec_residuals + off, r.prec_first.data(), sizeof(double) * r.prec_first.size());
        if (pw_half > 0) {
            const size_t ng2 = (size_t)2 * (2 * pw_half + 1) * (2 * pw_half + 1);
            std::memcpy(extra->pw_coefs + off * ng2, r.pw.data(), sizeof(double) * r.pw.size());
        }
        for (size_t k = 0; k < r.iters.size(); ++k) {
            if (iters) iters[koff + k] = r.iters[k];
            if (converged) converged[koff + k] = r.converged[k];
            if (rel_residual_iters) rel_residual_iters[koff + k] = r.rel_iters[k];
        }
        off += r.omegas.size();
        koff += r.iters.size();
    }
    if (stats) {
        *stats = bh_run_stats{};
        for (const RunStats &rs : pstats) {
            stats->steps += rs.steps;
            stats->iterations += rs.iterations;
            stats->solves += rs.solves;
            for (int p = 0; p < 2; ++p) {
                stats->eigh_calls[p] += rs.eigh_calls[p];
                stats->eigh_sweeps[p] += rs.eigh_sweeps[p];
                stats->eigh_max_sweeps[p] = std::max(stats->eigh_max_sweeps[p], (long)rs.eigh_max_sweeps[p]);
                stats->eigh_launch_max_sum[p] += rs.eigh_launch_max_sum[p];
            }
            stats->timed_ms += rs.timed_ms;
            stats->timed_launches += rs.timed_launches;
            stats->timed_bytes += rs.timed_bytes;
            stats->timed_flops += rs.timed_flops;
            stats->timed_flops_executed += rs.timed_flops_executed;
            stats->host_wait_seconds += rs.host_wait_seconds;
            stats->host_plan_seconds += rs.host_plan_seconds;
            stats->rr_problems += rs.rr_problems;
            stats->svqb_fallbacks += rs.svqb_fallbacks;
            stats->rank_drops += rs.rank_drops;
            stats->alg_bytes += rs.alg_bytes;
        }
        stats->seconds = parts.size() == 1 ? pstats[0].seconds : wall;
    }
    return 0;
    BH_API_CATCH
}

// ---- learned start ---------------------------------------------------------

int bh_cnn_create(bh_context *ctx, int height, int width, int n_vectors, int pw_half, int max_batch,
                  unsigned long long seed, bh_cnn **out) {
    BH_API_TRY
    if (!ctx || !out) return fail("bh_cnn_create: null argument");
    BH_HIP_CHECK(hipSetDevice(ctx->device));
    auto *c = new bh_cnn{ctx, nullptr};
    try {
        c->net = new Cnn(height, width, n_vectors, pw_half, max_batch, seed, ctx->st);
    } catch (...) {
        delete c;
        throw;
    }
    *out = c;
    return 0;
    BH_API_CATCH
}

void bh_cnn_destroy(bh_cnn *net) {
    if (!net) return;
    delete net->net;
    delete net;
}

long long bh_cnn_param_count(const bh_cnn *net) { return net ? net->net->param_count() : -1; }

int bh_cnn_get_params(bh_cnn *net, float *host) {
    BH_API_TRY
    if (!net || !host) return fail("bh_cnn_get_params: null argument");
    net->net->get_params(host);
    return 0;
    BH_API_CATCH
}

int bh_cnn_set_params(bh_cnn *net, const float *host) {
    BH_API_TRY
    if (!net || !host) return fail("bh_cnn_set_params: null argument");
    net->net->set_params(host);
    return 0;
    BH_API_CATCH
}

int bh_cnn_get_grads(bh_cnn *net, float *host) {
    BH_API_TRY
    if (!net || !host) return fail("bh_cnn_get_grads: null argument");
    net->net->get_grads(host);
    return 0;
    BH_API_CATCH
}

static int cnn_check(const bh_cnn *net, int batch, const float *grid, const float *embed, const char *who) {
    if (!net || !grid || !embed) return fail(std::string(who) + ": null argument");
    if (batch < 1 || batch > net->net->max_batch_size())
        return fail(std::string(who) + ": batch must be 1.." + std::to_string(net->net->max_batch_size()));
    return 0;
}

int bh_cnn_predict(bh_cnn *net, int batch, const float *grid, const float *embed, double *coefs) {
    BH_API_TRY
    if (cnn_check(net, batch, grid, embed, "bh_cnn_predict")) return -1;
    if (!coefs) return fail("bh_cnn_predict: null output");
    net->net->upload_inputs(batch, grid, embed);
    net->net->forward(batch);
    net->net->get_output(batch, coefs);
    return 0;
    BH_API_CATCH
}

int bh_cnn_predict_geometries(bh_cnn *net, int batch, const bh_geometry *geoms, int smoothing, const float *embed,
                              double *coefs) {
    BH_API_TRY
    if (!net || !geoms || !embed || !coefs) return fail("bh_cnn_predict_geometries: null argument");
    if (batch < 1 || batch > net->net->max_batch_size())
        return fail("bh_cnn_predict_geometries: batch must be 1.." + std::to_string(net->net->max_batch_size()));
    std::vector<DielGeom> hg(batch);
    for (int k = 0; k < batch; ++k) {
        const bh_geometry &g = geoms[k];
        if (g.n_atoms < 0 || g.n_atoms > BH_MAX_ATOMS) return fail("bh_cnn_predict_geometries: n_atoms out of range");
        DielGeom &d = hg[k];
        d = DielGeom{};
        for (int c = 0; c < 2; ++c) {
            d.a1[c] = g.a1[c];
            d.a2[c] = g.a2[c];
        }
        d.eps_bg = g.eps_bg;
        d.n_atoms = g.n_atoms;
        for (int a = 0; a < g.n_atoms; ++a)
            for (int c = 0; c < 4; ++c) d.atoms[a][c] = g.atoms[a][c];
    }
    net->net->inputs_from_geometries(batch, hg.data(), smoothing != 0, embed);
    net->net->forward(batch);
    net->net->get_output(batch, coefs);
    return 0;
    BH_API_CATCH
}

int bh_cnn_train_step(bh_cnn *net, int batch, const float *grid, const float *embed, const double *targets,
                      float learning_rate, double *loss) {
    BH_API_TRY
    if (cnn_check(net, batch, grid, embed, "bh_cnn_train_step")) return -1;
    if (!targets) return fail("bh_cnn_train_step: null targets");
    if (!(learning_rate >= 0.0f)) return fail("bh_cnn_train_step: learning rate must be >= 0");
    net->net->upload_inputs(batch, grid, embed);
    const double l = net->net->train_step(batch, targets, learning_rate);
    if (loss) *loss = l;
    return 0;
    BH_API_CATCH
}

int bh_cnn_set_graphs(bh_cnn *net, int enable) {
    BH_API_TRY
    if (!net || !net->net) return fail("bh_cnn_set_graphs: null network");
    if (enable != 0 && enable != 1) return fail("bh_cnn_set_graphs: enable must be 0 or 1");
    net->net->set_graphs(enable == 1);
    return 0;
    BH_API_CATCH
}

}  // extern "C"
