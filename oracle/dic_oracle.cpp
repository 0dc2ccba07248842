// dic_oracle.cpp — FP64 CPU restatement of geopyv-dev's subset-correlation
// hot path.  TEST INFRASTRUCTURE ONLY: this file is the correctness checker
// and the timed CPU baseline.  Nothing in the shipped library links it, and
// the product path never calls it.
//
// Every function follows the reference item cited beside it (paths relative
// to the geopyv-dev repository).  Loop orders and association follow the
// reference so that sums round the way the Rust FP64 code rounds; the build
// uses -ffp-contract=off so no FMA contraction changes that.
//
// Third-party arithmetic restated here (the crates are not vendored):
//   * rustfft 6.4.1 (unnormalised forward e^{-2πi jk/N}, inverse e^{+2πi jk/N})
//     -> a recursive mixed-radix DFT of our own; agreement is to round-off,
//     not bit-exact.
//   * nalgebra 0.33.3 try_inverse: closed-form cofactor inverse for 2x2/3x3,
//     LU with partial pivoting (reciprocal-multiply Gauss steps, unit-lower /
//     upper triangular solves) for larger matrices.
//   * ndarray 0.16.1 ArrayBase::sum over contiguous data: eight-way unrolled
//     fold (used for f_m and g_m, subset.rs:885,1162,1201,1274,1316).
//
// Parity status of each function is recorded in the header of oracle.py.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <complex>
#include <algorithm>
#include <chrono>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle_types.h"

typedef std::complex<double> cplx;

extern "C" {

// ---------------------------------------------------------------------------
// Status codes (mirror of include/geodic.h; error.rs:5-40 plus the two panics)
// ---------------------------------------------------------------------------
enum {
    ORC_OK = 0,
    ORC_INVALID_INPUT = 3,
    ORC_SINGULAR_HESSIAN = 10,
    ORC_SINGULAR_WARP = 11,
};

// ---------------------------------------------------------------------------
// ndarray sum: eight-way unrolled fold over a contiguous slice
// ---------------------------------------------------------------------------
static double unrolled_sum(const double* xs, size_t n) {
    double acc = 0.0;
    double p0 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0, p4 = 0.0, p5 = 0.0, p6 = 0.0, p7 = 0.0;
    size_t i = 0;
    while (n - i >= 8) {
        p0 = p0 + xs[i + 0]; p1 = p1 + xs[i + 1]; p2 = p2 + xs[i + 2]; p3 = p3 + xs[i + 3];
        p4 = p4 + xs[i + 4]; p5 = p5 + xs[i + 5]; p6 = p6 + xs[i + 6]; p7 = p7 + xs[i + 7];
        i += 8;
    }
    acc = acc + (p0 + p4);
    acc = acc + (p1 + p5);
    acc = acc + (p2 + p6);
    acc = acc + (p3 + p7);
    for (; i < n; ++i) acc = acc + xs[i];
    return acc;
}

double orc_unrolled_sum(const double* xs, int64_t n) { return unrolled_sum(xs, (size_t)n); }

// ---------------------------------------------------------------------------
// FFT (rustfft stand-in): recursive mixed-radix decimation in time
// ---------------------------------------------------------------------------
static void factorize(int n, std::vector<int>& f) {
    f.clear();
    for (int p = 2; n > 1; ++p) {
        while (n % p == 0) { f.push_back(p); n /= p; }
    }
}

struct FftPlan {
    int n;
    int sign;
    std::vector<int> radices;
    std::vector<cplx> tw;  // tw[k] = exp(sign*2πi k/n)
    FftPlan(int n_, int sign_) : n(n_), sign(sign_) {
        factorize(n, radices);
        tw.resize(n);
        for (int k = 0; k < n; ++k) {
            double a = sign * 2.0 * M_PI * (double)k / (double)n;
            tw[k] = cplx(std::cos(a), std::sin(a));
        }
    }
    // out[k] = Σ_j in[j*stride] ω^{jk}, length len = n / (twiddle stride tws)
    void rec(const cplx* in, cplx* out, int len, int stride, int tws, size_t level, cplx* scratch) const {
        if (len == 1) { out[0] = in[0]; return; }
        int p = radices[level];
        int m = len / p;
        // p sub-transforms of length m over the decimated inputs
        for (int r = 0; r < p; ++r) rec(in + (size_t)r * stride, out + (size_t)r * m, m, stride * p, tws * p, level + 1, scratch);
        // combine: X[k + q m] = Σ_r ω_len^{r(k + q m)} Y_r[k]
        for (int k = 0; k < m; ++k) {
            for (int r = 0; r < p; ++r) scratch[r] = out[(size_t)r * m + k];
            for (int q = 0; q < p; ++q) {
                cplx s(0.0, 0.0);
                int kk = k + q * m;
                for (int r = 0; r < p; ++r) {
                    long e = ((long)r * kk) % len;
                    s += scratch[r] * tw[(size_t)e * tws];
                }
                out[(size_t)kk] = s;
            }
        }
    }
    void exec(const cplx* in, cplx* out) const {
        std::vector<cplx> scratch(64);
        int pmax = 1;
        for (int p : radices) pmax = std::max(pmax, p);
        scratch.resize(pmax);
        rec(in, out, n, 1, 1, 0, scratch.data());
    }
};

// Unnormalised DFT of length n (sign -1 forward, +1 inverse), in place on
// split re/im arrays.  Exposed so tests can compare it with a direct DFT.
void orc_fft(double* re, double* im, int n, int sign) {
    FftPlan plan(n, sign);
    std::vector<cplx> a(n), b(n);
    for (int i = 0; i < n; ++i) a[i] = cplx(re[i], im[i]);
    plan.exec(a.data(), b.data());
    for (int i = 0; i < n; ++i) { re[i] = b[i].real(); im[i] = b[i].imag(); }
}

// num_complex Div: ((ac+bd) + (bc-ad)i) / (c²+d²)
static inline cplx cdiv(cplx a, cplx b) {
    double ns = b.real() * b.real() + b.imag() * b.imag();
    double re = a.real() * b.real() + a.imag() * b.imag();
    double im = a.imag() * b.real() - a.real() * b.imag();
    return cplx(re / ns, im / ns);
}

// build_kernel (image.rs:422-430): zero-phase layout of the 5-tap quintic kernel
static void build_kernel(int n, std::vector<cplx>& k) {
    const double kk[6] = {1.0 / 120.0, 13.0 / 60.0, 11.0 / 20.0, 13.0 / 60.0, 1.0 / 120.0, 0.0};
    k.assign(n, cplx(0.0, 0.0));
    k[0] = cplx(kk[2], 0.0);
    k[1] = cplx(kk[3], 0.0);
    k[2] = cplx(kk[4], 0.0);
    k[n - 2] = cplx(kk[0], 0.0);
    k[n - 1] = cplx(kk[1], 0.0);
}

// load_grayscale + gaussian_blur_5x5 + reflect101 (image.rs:450-527): the
// file loader's BT.709 integer grey of a decoded RGB8 frame (channels 3) or
// a grey frame (channels 1: to_rgb8 replicates it, and the formula returns
// the value), then the horizontal and vertical 5-tap Gaussian passes
// (sigma 1.1, taps exp(-x²/(2σ²)) normalised by their in-order sum), each
// sum in tap order, rounded half away from zero and clamped to 0..255.
void orc_load_convert(const uint8_t* src, int h, int w, int channels, double* out) {
    const double sigma = 1.1;
    double raw[5], k[5], sum = 0.0;
    for (int i = 0; i < 5; ++i) {
        const double x = (double)i - 2.0;
        raw[i] = std::exp(-x * x / (2.0 * sigma * sigma));
    }
    for (int i = 0; i < 5; ++i) sum += raw[i];
    for (int i = 0; i < 5; ++i) k[i] = raw[i] / sum;
    auto reflect101 = [](long idx, long len) { return idx < 0 ? -idx : (idx >= len ? 2 * (len - 1) - idx : idx); };
    std::vector<double> gs((size_t)h * w), hp((size_t)h * w);
    for (long y = 0; y < h; ++y)
        for (long x = 0; x < w; ++x) {
            const uint8_t* p = src + (y * w + x) * channels;
            const uint32_t r = p[0], g = channels == 3 ? p[1] : p[0], b = channels == 3 ? p[2] : p[0];
            gs[y * w + x] = (double)((2126 * r + 7152 * g + 722 * b + 5000) / 10000);
        }
    for (long i = 0; i < h; ++i)
        for (long j = 0; j < w; ++j) {
            double v = 0.0;
            for (long ki = 0; ki < 5; ++ki) v += k[ki] * gs[i * w + reflect101(j + ki - 2, w)];
            hp[i * w + j] = v;
        }
    for (long i = 0; i < h; ++i)
        for (long j = 0; j < w; ++j) {
            double v = 0.0;
            for (long ki = 0; ki < 5; ++ki) v += k[ki] * hp[reflect101(i + ki - 2, h) * w + j];
            out[i * w + j] = std::min(std::max(std::round(v), 0.0), 255.0);
        }
}

// pad_replicate (image.rs:436-443)
void orc_pad_replicate(const double* img, int rows, int cols, int border, double* out) {
    int pr = rows + 2 * border, pc = cols + 2 * border;
    for (int i = 0; i < pr; ++i) {
        int ri = std::min(std::max(i - border, 0), rows - 1);
        for (int j = 0; j < pc; ++j) {
            int ci = std::min(std::max(j - border, 0), cols - 1);
            out[(size_t)i * pc + j] = img[(size_t)ri * cols + ci];
        }
    }
}

// get_c (image.rs:252-327).  Row pass, transpose, column pass, transpose
// back, real part.  The imaginary round-off is carried through both passes
// as the reference does.  Threads split lines (image.rs:278,306); the two
// transposes are also split over threads here (serial in the reference).
void orc_get_c(const double* img, int rows, int cols, int border, double* out, int threads) {
    const int pr = rows + 2 * border, pc = cols + 2 * border;
    std::vector<double> padded((size_t)pr * pc);
    orc_pad_replicate(img, rows, cols, border, padded.data());
    std::vector<cplx> c((size_t)pr * pc);
    for (size_t i = 0; i < c.size(); ++i) c[i] = cplx(padded[i], 0.0);
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#else
    (void)threads;
#endif
    // row pass
    {
        FftPlan fwd(pc, -1), inv(pc, +1);
        std::vector<cplx> kx, kxf(pc);
        build_kernel(pc, kx);
        fwd.exec(kx.data(), kxf.data());
        const double scale = 1.0 / (double)pc;
#pragma omp parallel num_threads(threads)
        {
            std::vector<cplx> tmp(pc), tmp2(pc);
#pragma omp for schedule(static)
            for (int i = 0; i < pr; ++i) {
                cplx* row = &c[(size_t)i * pc];
                fwd.exec(row, tmp.data());
                for (int j = 0; j < pc; ++j) tmp[j] = cdiv(tmp[j], kxf[j]);
                inv.exec(tmp.data(), tmp2.data());
                for (int j = 0; j < pc; ++j) row[j] = cplx(tmp2[j].real() * scale, tmp2[j].imag() * scale);
            }
        }
    }
    std::vector<cplx> ct((size_t)pr * pc);
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int j = 0; j < pc; ++j)
        for (int i = 0; i < pr; ++i) ct[(size_t)j * pr + i] = c[(size_t)i * pc + j];
    // column pass
    {
        FftPlan fwd(pr, -1), inv(pr, +1);
        std::vector<cplx> ky, kyf(pr);
        build_kernel(pr, ky);
        fwd.exec(ky.data(), kyf.data());
        const double scale = 1.0 / (double)pr;
#pragma omp parallel num_threads(threads)
        {
            std::vector<cplx> tmp(pr), tmp2(pr);
#pragma omp for schedule(static)
            for (int j = 0; j < pc; ++j) {
                cplx* col = &ct[(size_t)j * pr];
                fwd.exec(col, tmp.data());
                for (int i = 0; i < pr; ++i) tmp[i] = cdiv(tmp[i], kyf[i]);
                inv.exec(tmp.data(), tmp2.data());
                for (int i = 0; i < pr; ++i) col[i] = cplx(tmp2[i].real() * scale, tmp2[i].imag() * scale);
            }
        }
    }
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int i = 0; i < pr; ++i)
        for (int j = 0; j < pc; ++j) out[(size_t)i * pc + j] = ct[(size_t)j * pr + i].real();
}

// ---------------------------------------------------------------------------
// B-spline evaluation (image.rs:115-230)
// ---------------------------------------------------------------------------
static const double QM[6][6] = {
    { 1.0 / 120.0,  13.0 / 60.0,  11.0 / 20.0,  13.0 / 60.0,  1.0 / 120.0,  0.0},
    {-1.0 / 24.0,  -5.0 / 12.0,   0.0,          5.0 / 12.0,   1.0 / 24.0,   0.0},
    { 1.0 / 12.0,   1.0 / 6.0,   -1.0 / 2.0,    1.0 / 6.0,    1.0 / 12.0,   0.0},
    {-1.0 / 12.0,   1.0 / 6.0,    0.0,         -1.0 / 6.0,    1.0 / 12.0,   0.0},
    { 1.0 / 24.0,  -1.0 / 6.0,    1.0 / 4.0,   -1.0 / 6.0,    1.0 / 24.0,   0.0},
    {-1.0 / 120.0,  1.0 / 24.0,  -1.0 / 12.0,   1.0 / 12.0,  -1.0 / 24.0,   1.0 / 120.0},
};

// qt_poly (image.rs:194-208)
static inline void qt_poly(double t, double out[6]) {
    double t2 = t * t, t3 = t2 * t, t4 = t3 * t;
    double p[6] = {1.0, t, t2, t3, t4, t4 * t};
    for (int k = 0; k < 6; ++k) {
        double s = 0.0;
        for (int r = 0; r < 6; ++r) s += QM[r][k] * p[r];
        out[k] = s;
    }
}

// qt_dpoly (image.rs:217-230)
static inline void qt_dpoly(double t, double out[6]) {
    double t2 = t * t, t3 = t2 * t;
    double p[6] = {0.0, 1.0, 2.0 * t, 3.0 * t2, 4.0 * t3, 5.0 * t2 * t2};
    for (int k = 0; k < 6; ++k) {
        double s = 0.0;
        for (int r = 0; r < 6; ++r) s += QM[r][k] * p[r];
        out[k] = s;
    }
}

struct Img {
    const double* coef;  // (h+2b) x (w+2b) row-major
    int h, w, border;
};

// Image::eval (image.rs:128-148)
static inline double img_eval(const Img& im, double x, double y) {
    x = std::min(std::max(x, 0.0), (double)im.w - 1e-9);
    y = std::min(std::max(y, 0.0), (double)im.h - 1e-9);
    size_t xf = (size_t)std::floor(x), yf = (size_t)std::floor(y);
    double a[6], b[6];
    qt_poly(y - (double)yf, a);
    qt_poly(x - (double)xf, b);
    size_t r0 = yf + im.border - 2, c0 = xf + im.border - 2;
    size_t cw = (size_t)im.w + 2 * im.border;
    double val = 0.0;
    for (int r = 0; r < 6; ++r) {
        const double* row = im.coef + (r0 + r) * cw + c0;
        double s = 0.0;
        for (int k = 0; k < 6; ++k) s += row[k] * b[k];
        val += a[r] * s;
    }
    return val;
}

// Image::eval_grad via rows (image.rs:115-124,155-178)
static inline void img_eval_grad(const Img& im, double x, double y, double* v, double* gx, double* gy) {
    x = std::min(std::max(x, 0.0), (double)im.w - 1e-9);
    y = std::min(std::max(y, 0.0), (double)im.h - 1e-9);
    size_t xf = (size_t)std::floor(x), yf = (size_t)std::floor(y);
    double a[6], da[6], b[6], db[6];
    qt_poly(y - (double)yf, a);
    qt_dpoly(y - (double)yf, da);
    qt_poly(x - (double)xf, b);
    qt_dpoly(x - (double)xf, db);
    size_t r0 = yf + im.border - 2, c0 = xf + im.border - 2;
    size_t cw = (size_t)im.w + 2 * im.border;
    double cb[6], cq[6];
    for (int r = 0; r < 6; ++r) {
        const double* row = im.coef + (r0 + r) * cw + c0;
        double s = 0.0, t = 0.0;
        for (int k = 0; k < 6; ++k) { s += row[k] * b[k]; t += row[k] * db[k]; }
        cb[r] = s; cq[r] = t;
    }
    double val = 0.0, ggx = 0.0, ggy = 0.0;
    for (int r = 0; r < 6; ++r) {
        val += a[r] * cb[r];
        ggx += a[r] * cq[r];
        ggy += da[r] * cb[r];
    }
    *v = val; *gx = ggx; *gy = ggy;
}

void orc_eval(const double* coef, int h, int w, int border, const double* xy, int64_t n, double* out) {
    Img im{coef, h, w, border};
    for (int64_t i = 0; i < n; ++i) out[i] = img_eval(im, xy[2 * i], xy[2 * i + 1]);
}

void orc_eval_grad(const double* coef, int h, int w, int border, const double* xy, int64_t n, double* out3) {
    Img im{coef, h, w, border};
    for (int64_t i = 0; i < n; ++i) img_eval_grad(im, xy[2 * i], xy[2 * i + 1], &out3[3 * i], &out3[3 * i + 1], &out3[3 * i + 2]);
}

// get_qcqt for one pixel (image.rs:339-366): Q · C_block · Qᵀ, block [dr*6+dc]
void orc_qcqt_block(const double* coef, int h, int w, int border, int i, int j, double* out36) {
    (void)h;
    size_t cw = (size_t)w + 2 * border;
    double cb[6][6], tmp[6][6];
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) cb[r][c] = coef[(size_t)(i + border - 2 + r) * cw + (j + border - 2 + c)];
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) {
            double s = 0.0;
            for (int k = 0; k < 6; ++k) s += QM[r][k] * cb[k][c];
            tmp[r][c] = s;
        }
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) {
            double s = 0.0;
            for (int k = 0; k < 6; ++k) s += tmp[r][k] * QM[c][k];
            out36[r * 6 + c] = s;
        }
}

// ---------------------------------------------------------------------------
// Local masks (masks.rs:58-263)
// ---------------------------------------------------------------------------
// shape: 0 circle, 1 square, 2 semicircle.  coords (n,2) [x_off, y_off];
// subset_mask (2s+1)^2 i32.  Returns n_px, or -1 for size 0.
int orc_mask_local(int shape, int size, double* coords, int32_t* subset_mask) {
    if (size <= 0) return -1;
    int side = 2 * size + 1;
    int n = 0;
    for (int row = 0; row < side; ++row) {
        for (int col = 0; col < side; ++col) {
            long x = (long)col - size, y = (long)row - size;
            bool keep;
            if (shape == 1) {
                keep = true;
            } else {
                double dist = std::sqrt((double)(x * x + y * y));
                keep = dist <= (double)size;
                if (shape == 2) keep = keep && y >= 0;
            }
            if (subset_mask) subset_mask[row * side + col] = keep ? 1 : 0;
            if (keep) {
                if (coords) { coords[2 * n] = (double)x; coords[2 * n + 1] = (double)y; }
                ++n;
            }
        }
    }
    return n;
}

// mask_update (masks.rs:199-263), including the edge quirk at 225-232 where
// an exclusive end at or beyond the edge also drops the last row/column.
int orc_mask_update(int shape, int size, double cx, double cy, const uint8_t* mask, int mh, int mw, double* coords_out) {
    int side = 2 * size + 1;
    std::vector<int32_t> sm((size_t)side * side);
    if (orc_mask_local(shape, size, nullptr, sm.data()) < 0) return -1;
    long center_x = (long)cx;  // `as isize` truncates toward zero
    long center_y = (long)cy;
    long ly = center_y - size, uy = center_y + size + 1;
    long lx = center_x - size, ux = center_x + size + 1;
    long lyp = 0, uyp = 0, lxp = 0, uxp = 0;
    if (lx < 0) { lxp = -lx; lx = 0; }
    if (ly < 0) { lyp = -ly; ly = 0; }
    if (ux >= mw) { uxp = ux + 1 - mw; ux = mw - 1; }
    if (uy >= mh) { uyp = uy + 1 - mh; uy = mh - 1; }
    std::vector<uint8_t> local((size_t)side * side, 0);
    long trow_end = side - uyp, tcol_end = side - uxp;
    // an empty or reversed window, or slices of different shapes, make the
    // reference's ndarray slice/assign panic
    if (ly > uy || lx > ux || trow_end < lyp || tcol_end < lxp) return -2;
    if (uy - ly != trow_end - lyp || ux - lx != tcol_end - lxp) return -2;
    for (long r = 0; r < uy - ly; ++r)
        for (long c = 0; c < ux - lx; ++c)
            local[(size_t)(lyp + r) * side + (lxp + c)] = mask[(size_t)(ly + r) * mw + (lx + c)];
    int n = 0;
    for (int row = 0; row < side; ++row)
        for (int col = 0; col < side; ++col)
            if (local[(size_t)row * side + col] != 0 && sm[(size_t)row * side + col] != 0) {
                if (coords_out) { coords_out[2 * n] = (double)(col - size); coords_out[2 * n + 1] = (double)(row - size); }
                ++n;
            }
    return n;
}

// ---------------------------------------------------------------------------
// Small dense linear algebra (nalgebra try_inverse semantics)
// ---------------------------------------------------------------------------
// Row-major m x m.  Returns 1 on success, 0 when singular.
static int inv_small(const double* M, int m, double* out) {
    auto A = [&](int r, int c) { return M[r * m + c]; };
    if (m == 1) {
        if (A(0, 0) == 0.0) return 0;
        out[0] = 1.0 / A(0, 0);
        return 1;
    }
    if (m == 2) {
        double det = A(0, 0) * A(1, 1) - A(1, 0) * A(0, 1);
        if (det == 0.0) return 0;
        out[0] = A(1, 1) / det;
        out[1] = -A(0, 1) / det;
        out[2] = -A(1, 0) / det;
        out[3] = A(0, 0) / det;
        return 1;
    }
    if (m == 3) {
        double m11 = A(1, 1) * A(2, 2) - A(2, 1) * A(1, 2);
        double m12 = A(1, 0) * A(2, 2) - A(2, 0) * A(1, 2);
        double m13 = A(1, 0) * A(2, 1) - A(2, 0) * A(1, 1);
        double det = A(0, 0) * m11 - A(0, 1) * m12 + A(0, 2) * m13;
        if (det == 0.0) return 0;
        out[0] = m11 / det;
        out[1] = -(A(0, 1) * A(2, 2) - A(2, 1) * A(0, 2)) / det;
        out[2] = (A(0, 1) * A(1, 2) - A(1, 1) * A(0, 2)) / det;
        out[3] = -m12 / det;
        out[4] = (A(0, 0) * A(2, 2) - A(2, 0) * A(0, 2)) / det;
        out[5] = -(A(0, 0) * A(1, 2) - A(1, 0) * A(0, 2)) / det;
        out[6] = m13 / det;
        out[7] = -(A(0, 0) * A(2, 1) - A(2, 0) * A(0, 1)) / det;
        out[8] = (A(0, 0) * A(1, 1) - A(1, 0) * A(0, 1)) / det;
        return 1;
    }
    // LU with partial pivoting (first max |.| in the column), then solve
    // against the permuted identity.
    std::vector<double> L(M, M + m * m);
    auto Lr = [&](int r, int c) -> double& { return L[r * m + c]; };
    std::vector<std::pair<int, int>> perm;
    for (int i = 0; i < m; ++i) {
        int piv = i;
        double best = std::fabs(Lr(i, i));
        for (int r = i + 1; r < m; ++r)
            if (std::fabs(Lr(r, i)) > best) { best = std::fabs(Lr(r, i)); piv = r; }
        double diag = Lr(piv, i);
        if (diag == 0.0) continue;
        if (piv != i) {
            perm.push_back({i, piv});
            for (int c = 0; c < m; ++c) std::swap(Lr(i, c), Lr(piv, c));
        }
        double inv_diag = 1.0 / diag;
        for (int r = i + 1; r < m; ++r) Lr(r, i) *= inv_diag;
        for (int k = i + 1; k < m; ++k) {
            double pk = Lr(i, k);
            for (int r = i + 1; r < m; ++r) Lr(r, k) = -pk * Lr(r, i) + Lr(r, k);
        }
    }
    for (int i = 0; i < m; ++i)
        if (Lr(i, i) == 0.0) return 0;
    std::vector<double> B((size_t)m * m, 0.0);
    for (int i = 0; i < m; ++i) B[i * m + i] = 1.0;
    for (auto& pr : perm)
        for (int c = 0; c < m; ++c) std::swap(B[pr.first * m + c], B[pr.second * m + c]);
    for (int c = 0; c < m; ++c) {
        for (int i = 0; i < m; ++i) {
            double coeff = B[i * m + c];
            for (int r = i + 1; r < m; ++r) B[r * m + c] = -coeff * Lr(r, i) + B[r * m + c];
        }
        for (int i = m - 1; i >= 0; --i) {
            double coeff = B[i * m + c] / Lr(i, i);
            B[i * m + c] = coeff;
            for (int r = 0; r < i; ++r) B[r * m + c] = -coeff * Lr(r, i) + B[r * m + c];
        }
    }
    std::memcpy(out, B.data(), sizeof(double) * m * m);
    return 1;
}

int orc_inverse(const double* M, int m, double* out) { return inv_small(M, m, out); }

// C = A B, row-major square, sum over k ascending (nalgebra gemm column axpy)
static void matmul(const double* A, const double* B, int m, double* C) {
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) {
            double s = A[i * m + 0] * B[0 * m + j];
            for (int k = 1; k < m; ++k) s = s + A[i * m + k] * B[k * m + j];
            C[i * m + j] = s;
        }
}

// ---------------------------------------------------------------------------
// Warp helpers (subset.rs:236-343,487-649)
// ---------------------------------------------------------------------------
static inline int p_len(int order) { return order == 0 ? 2 : (order == 1 ? 6 : 12); }

// apply_warp_into (subset.rs:262-300)
void orc_apply_warp(const double* coord, const double* p, int plen, const double* fc, int64_t n, double* gc) {
    double xc = coord[0], yc = coord[1];
    if (plen == 2) {
        for (int64_t i = 0; i < n; ++i) { gc[2 * i] = fc[2 * i] + p[0]; gc[2 * i + 1] = fc[2 * i + 1] + p[1]; }
    } else if (plen <= 7) {
        double u = p[0], v = p[1], ux = p[2], vx = p[3], uy = p[4], vy = p[5];
        for (int64_t i = 0; i < n; ++i) {
            double x = fc[2 * i], y = fc[2 * i + 1];
            double dx = x - xc, dy = y - yc;
            gc[2 * i] = x + u + ux * dx + uy * dy;
            gc[2 * i + 1] = y + v + vx * dx + vy * dy;
        }
    } else {
        double u = p[0], v = p[1], ux = p[2], vx = p[3], uy = p[4], vy = p[5];
        double uxx = p[6], vxx = p[7], uxy = p[8], vxy = p[9], uyy = p[10], vyy = p[11];
        for (int64_t i = 0; i < n; ++i) {
            double x = fc[2 * i], y = fc[2 * i + 1];
            double dx = x - xc, dy = y - yc;
            gc[2 * i] = x + u + ux * dx + uy * dy + 0.5 * uxx * dx * dx + uxy * dx * dy + 0.5 * uyy * dy * dy;
            gc[2 * i + 1] = y + v + vx * dx + vy * dy + 0.5 * vxx * dx * dx + vxy * dx * dy + 0.5 * vyy * dy * dy;
        }
    }
}

// steepest_descent (subset.rs:311-343); sdi is (n, m) row-major
void orc_steepest_descent(const double* coord, const double* coords, const double* grad, int64_t n, int order, double* sdi) {
    int m = p_len(order);
    for (int64_t i = 0; i < n; ++i) {
        double dx = coords[2 * i] - coord[0], dy = coords[2 * i + 1] - coord[1];
        double gx = grad[2 * i], gy = grad[2 * i + 1];
        double* s = sdi + i * m;
        s[0] = gx; s[1] = gy;
        if (m < 6) continue;
        s[2] = gx * dx; s[3] = gy * dx; s[4] = gx * dy; s[5] = gy * dy;
        if (m == 12) {
            s[6] = gx * 0.5 * dx * dx; s[7] = gy * 0.5 * dx * dx;
            s[8] = gx * dx * dy; s[9] = gy * dx * dy;
            s[10] = gx * 0.5 * dy * dy; s[11] = gy * 0.5 * dy * dy;
        }
    }
}

// compute_hessian (subset.rs:348-364)
void orc_hessian(const double* sdi, int64_t n, int m, double* H) {
    for (int i = 0; i < m; ++i)
        for (int j = i; j < m; ++j) {
            double dot = 0.0;
            for (int64_t k = 0; k < n; ++k) dot += sdi[k * m + i] * sdi[k * m + j];
            H[i * m + j] = dot;
            H[j * m + i] = dot;
        }
}

// convergence_norm (subset.rs:622-649)
double orc_convergence_norm(const double* d, int n, double size) {
    auto sq = [](double v) { return v * v; };
    if (n == 2) return std::sqrt(sq(d[0]) + sq(d[1]));
    if (n <= 7)
        return std::sqrt(sq(d[0]) + sq(d[1]) + sq(d[2] * size) + sq(d[3] * size) + sq(d[4] * size) + sq(d[5] * size));
    return std::sqrt(sq(d[0]) + sq(d[1]) + sq(d[2] * size) + sq(d[3] * size) + sq(d[4] * size) + sq(d[5] * size) +
                     sq(0.5 * d[6] * size * size) + sq(0.5 * d[7] * size * size) + sq(0.5 * d[8] * size * size) +
                     sq(0.5 * d[9] * size * size) + sq(0.5 * d[10] * size * size) + sq(0.5 * d[11] * size * size));
}

// s_terms of compose_icgn (subset.rs:547-571)
static void s_terms(const double* q, double s[18]) {
    double u = q[0], v = q[1], ux = q[2], vx = q[3], uy = q[4], vy = q[5];
    double uxx = q[6], vxx = q[7], uxy = q[8], vxy = q[9], uyy = q[10], vyy = q[11];
    s[0] = 2.0 * ux + ux * ux + u * uxx;
    s[1] = 2.0 * u * uxy + 2.0 * (1.0 + ux) * uy;
    s[2] = uy * uy + u * uyy;
    s[3] = 2.0 * u * (1.0 + ux);
    s[4] = 2.0 * u * uy;
    s[5] = u * u;
    s[6] = 0.5 * (v * uxx + 2.0 * (1.0 + ux) * vx + u * vxx);
    s[7] = uy * vx + ux * vy + v * uxy + u * vxy + vy + ux;
    s[8] = 0.5 * (v * uyy + 2.0 * uy * (1.0 + vy) + u * vyy);
    s[9] = v + v * ux + u * vx;
    s[10] = u + v * uy + u * vy;
    s[11] = u * v;
    s[12] = vx * vx + v * vxx;
    s[13] = 2.0 * v * vxy + 2.0 * vx * (1.0 + vy);
    s[14] = 2.0 * vy + vy * vy + v * vyy;
    s[15] = 2.0 * v * vx;
    s[16] = 2.0 * v * (1.0 + vy);
    s[17] = v * v;
}

static void lifted6(const double* q, double W[36]) {
    double s[18];
    s_terms(q, s);
    const double r[36] = {
        1.0 + s[0], s[1], s[2], s[3], s[4], s[5],
        s[6], 1.0 + s[7], s[8], s[9], s[10], s[11],
        s[12], s[13], 1.0 + s[14], s[15], s[16], s[17],
        0.5 * q[6], q[8], 0.5 * q[10], 1.0 + q[2], q[4], q[0],
        0.5 * q[7], q[9], 0.5 * q[11], q[3], 1.0 + q[5], q[1],
        0.0, 0.0, 0.0, 0.0, 0.0, 1.0,
    };
    std::memcpy(W, r, sizeof(r));
}

// compose_icgn (subset.rs:487-611).  Returns 0, or ORC_SINGULAR_WARP.
int orc_compose_icgn(const double* p, const double* dp, int plen, double* out) {
    if (plen == 2) {
        out[0] = p[0] - dp[0];
        out[1] = p[1] - dp[1];
        return ORC_OK;
    }
    if (plen == 6) {
        const double wo[9] = {1.0 + p[2], p[4], p[0], p[3], 1.0 + p[5], p[1], 0.0, 0.0, 1.0};
        const double wd[9] = {1.0 + dp[2], dp[4], dp[0], dp[3], 1.0 + dp[5], dp[1], 0.0, 0.0, 1.0};
        double wi[9], wn[9];
        if (!inv_small(wd, 3, wi)) return ORC_SINGULAR_WARP;
        matmul(wo, wi, 3, wn);
        out[0] = wn[0 * 3 + 2];
        out[1] = wn[1 * 3 + 2];
        out[2] = wn[0 * 3 + 0] - 1.0;
        out[3] = wn[1 * 3 + 0];
        out[4] = wn[0 * 3 + 1];
        out[5] = wn[1 * 3 + 1] - 1.0;
        return ORC_OK;
    }
    double wo[36], wd[36], wi[36], wn[36];
    lifted6(p, wo);
    lifted6(dp, wd);
    if (!inv_small(wd, 6, wi)) return ORC_SINGULAR_WARP;
    matmul(wo, wi, 6, wn);
    out[0] = wn[3 * 6 + 5];
    out[1] = wn[4 * 6 + 5];
    out[2] = wn[3 * 6 + 3] - 1.0;
    out[3] = wn[4 * 6 + 3];
    out[4] = wn[3 * 6 + 4];
    out[5] = wn[4 * 6 + 4] - 1.0;
    out[6] = wn[3 * 6 + 0] * 2.0;
    out[7] = wn[4 * 6 + 0] * 2.0;
    out[8] = wn[3 * 6 + 1];
    out[9] = wn[4 * 6 + 1];
    out[10] = wn[3 * 6 + 2] * 2.0;
    out[11] = wn[4 * 6 + 2] * 2.0;
    return ORC_OK;
}

// ---------------------------------------------------------------------------
// Warp-adequacy diagnostic (subset.rs:698-820)
// ---------------------------------------------------------------------------
// zn_slice (subset.rs:698-704)
void orc_zn_slice(const double* x, int64_t n, double* out) {
    double nn = (double)n;
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += x[i];
    double mean = s / nn;
    double v = 0.0;
    for (int64_t i = 0; i < n; ++i) { double d = x[i] - mean; v += d * d; }
    double var = v / nn;
    double sd = std::sqrt(std::max(var, 1e-18));
    for (int64_t i = 0; i < n; ++i) out[i] = (x[i] - mean) / sd;
}

// eta_component (subset.rs:713-727); modes (n, K) row-major
double orc_eta_component(const double* modes, int K, const double* grad_zn, const double* e_centered, int64_t n) {
    std::vector<double> raw(n), sc(n);
    double acc = 0.0;
    for (int k = 0; k < K; ++k) {
        for (int64_t i = 0; i < n; ++i) raw[i] = grad_zn[i] * modes[i * K + k];
        orc_zn_slice(raw.data(), n, sc.data());
        double s = 0.0;
        for (int64_t i = 0; i < n; ++i) s += sc[i] * e_centered[i];
        double proj = s / (double)n;
        acc += proj * proj;
    }
    return std::sqrt(acc);
}

// omitted_mode_diagnostic (subset.rs:748-820).  Returns 0 for order 0..2,
// -1 otherwise (None).
int orc_omitted_mode_diagnostic(const double* coord, int order, const double* f_coords, const double* grad_f,
                                const double* f, double f_m, double delta_f, const double* g, double g_m,
                                double delta_g, int64_t n, double* eta_u, double* eta_v) {
    if (order < 0 || order > 2) return -1;
    std::vector<double> e(n), ec(n), gx(n), gy(n), zx(n), zy(n);
    for (int64_t i = 0; i < n; ++i) e[i] = (f[i] - f_m) / delta_f - (g[i] - g_m) / delta_g;
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += e[i];
    double e_bar = s / (double)n;
    for (int64_t i = 0; i < n; ++i) ec[i] = e[i] - e_bar;
    for (int64_t i = 0; i < n; ++i) { gx[i] = grad_f[2 * i]; gy[i] = grad_f[2 * i + 1]; }
    orc_zn_slice(gx.data(), n, zx.data());
    orc_zn_slice(gy.data(), n, zy.data());
    int K = order + 2;
    std::vector<double> modes((size_t)n * K);
    for (int64_t i = 0; i < n; ++i) {
        double xi = f_coords[2 * i] - coord[0], ze = f_coords[2 * i + 1] - coord[1];
        double* mo = &modes[(size_t)i * K];
        if (order == 0) { mo[0] = xi; mo[1] = ze; }
        else if (order == 1) { mo[0] = xi * xi; mo[1] = xi * ze; mo[2] = ze * ze; }
        else { mo[0] = xi * xi * xi; mo[1] = xi * xi * ze; mo[2] = xi * ze * ze; mo[3] = ze * ze * ze; }
    }
    *eta_u = orc_eta_component(modes.data(), K, zx.data(), ec.data(), n);
    *eta_v = orc_eta_component(modes.data(), K, zy.data(), ec.data(), n);
    return 0;
}

// ---------------------------------------------------------------------------
// Subset::new (subset.rs:836-919)
// ---------------------------------------------------------------------------
struct SubsetRef {
    double coord[2];
    int order;
    int64_t n;
    std::vector<double> f_coords, f, grad_f;
    double f_m, delta_f, sssig, sigma;
};

static int subset_new(const Img& fim, const double* coord, const double* offsets, int64_t n, int order, SubsetRef& S) {
    if (n == 0) return ORC_INVALID_INPUT;
    S.coord[0] = coord[0]; S.coord[1] = coord[1];
    S.order = order;
    S.n = n;
    S.f_coords.resize(2 * n);
    S.f.resize(n);
    S.grad_f.resize(2 * n);
    for (int64_t i = 0; i < n; ++i) {
        S.f_coords[2 * i] = offsets[2 * i] + coord[0];
        S.f_coords[2 * i + 1] = offsets[2 * i + 1] + coord[1];
    }
    for (int64_t i = 0; i < n; ++i)
        img_eval_grad(fim, S.f_coords[2 * i], S.f_coords[2 * i + 1], &S.f[i], &S.grad_f[2 * i], &S.grad_f[2 * i + 1]);
    S.f_m = unrolled_sum(S.f.data(), n) / (double)n;
    double d = 0.0;
    for (int64_t i = 0; i < n; ++i) { double t = S.f[i] - S.f_m; d += t * t; }
    S.delta_f = std::sqrt(d);
    if (S.delta_f == 0.0) return ORC_INVALID_INPUT;
    double ss = 0.0;
    for (int64_t i = 0; i < n; ++i) ss += 0.5 * (S.grad_f[2 * i] * S.grad_f[2 * i] + S.grad_f[2 * i + 1] * S.grad_f[2 * i + 1]);
    S.sssig = ss;
    double v = 0.0;
    for (int64_t i = 0; i < n; ++i) { double t = S.f[i] - S.f_m; v += t * t; }
    S.sigma = std::sqrt(v / (double)n);
    return ORC_OK;
}

// OrcResult (oracle_types.h) is the result record shared with the Python
// binding (field order fixed).

// solve_icgn_result (subset.rs:1117-1219) and solve_fagn_result
// (subset.rs:1228-1334), deformable texture.  history (max_it x 4) optional.
static int subset_solve(const Img& gim, const SubsetRef& S, const double* p0, int p0_len, int method, double tolerance,
                        double max_norm, int max_iterations, OrcResult* R, double* history) {
    const int64_t n = S.n;
    const int m = p_len(S.order);
    const double size = std::sqrt((double)n);
    double p[12] = {0};
    for (int k = 0; k < m && k < p0_len; ++k) p[k] = p0 ? p0[k] : 0.0;
    std::vector<double> gc(2 * n), g(n, 0.0);
    std::vector<double> sdi((size_t)n * m), H((size_t)m * m), Hi((size_t)m * m);
    int it_done = 0;
    bool converged = false;
    double last_zncc = 0.0, last_znssd = 4.0, last_norm = NAN;
    if (method == 0) {
        orc_steepest_descent(S.coord, S.f_coords.data(), S.grad_f.data(), n, S.order, sdi.data());
        orc_hessian(sdi.data(), n, m, H.data());
        if (!inv_small(H.data(), m, Hi.data())) return ORC_SINGULAR_HESSIAN;
        std::vector<double> gz(m);
        for (int iteration = 1; iteration <= max_iterations; ++iteration) {
            orc_apply_warp(S.coord, p, m, S.f_coords.data(), n, gc.data());
            for (int64_t i = 0; i < n; ++i) g[i] = img_eval(gim, gc[2 * i], gc[2 * i + 1]);
            double g_m = unrolled_sum(g.data(), n) / (double)n;
            double dg = 0.0;
            for (int64_t i = 0; i < n; ++i) { double t = g[i] - g_m; dg += t * t; }
            double delta_g = std::sqrt(dg);
            // delta_p_and_znssd_icgn (subset.rs:418-446)
            double ratio = S.delta_f / delta_g;
            std::fill(gz.begin(), gz.end(), 0.0);
            double zacc = 0.0;
            for (int64_t i = 0; i < n; ++i) {
                double a = S.f[i] - S.f_m;
                double b = g[i] - g_m;
                double z = a / S.delta_f - b / delta_g;
                zacc += z * z;
                double bracket = a - ratio * b;
                for (int j = 0; j < m; ++j) gz[j] += sdi[(size_t)i * m + j] * bracket;
            }
            double dp[12];
            for (int r = 0; r < m; ++r) {
                double s = gz[0] * Hi[r * m + 0];
                for (int k = 1; k < m; ++k) s = s + gz[k] * Hi[r * m + k];
                dp[r] = -s;
            }
            double pn[12];
            int st = orc_compose_icgn(p, dp, m, pn);
            if (st) return st;
            std::memcpy(p, pn, sizeof(double) * m);
            double norm = orc_convergence_norm(dp, m, size);
            double c_zncc = 1.0 - zacc / 2.0;
            if (history) { history[4 * (iteration - 1)] = iteration; history[4 * (iteration - 1) + 1] = norm; history[4 * (iteration - 1) + 2] = c_zncc; history[4 * (iteration - 1) + 3] = zacc; }
            it_done = iteration; last_zncc = c_zncc; last_znssd = zacc; last_norm = norm;
            if (norm < max_norm) { converged = true; break; }
        }
    } else {
        std::vector<double> gg(2 * n), gz(m);
        for (int iteration = 1; iteration <= max_iterations; ++iteration) {
            double g_center[2] = {S.coord[0] + p[0], S.coord[1] + p[1]};
            orc_apply_warp(S.coord, p, m, S.f_coords.data(), n, gc.data());
            for (int64_t i = 0; i < n; ++i) g[i] = img_eval(gim, gc[2 * i], gc[2 * i + 1]);
            double g_m = unrolled_sum(g.data(), n) / (double)n;
            double dg = 0.0;
            for (int64_t i = 0; i < n; ++i) { double t = g[i] - g_m; dg += t * t; }
            double delta_g = std::sqrt(dg);
            for (int64_t i = 0; i < n; ++i) {
                double v;
                img_eval_grad(gim, gc[2 * i], gc[2 * i + 1], &v, &gg[2 * i], &gg[2 * i + 1]);
            }
            orc_steepest_descent(g_center, gc.data(), gg.data(), n, S.order, sdi.data());
            orc_hessian(sdi.data(), n, m, H.data());
            // delta_p_fagn (subset.rs:452-473)
            for (int j = 0; j < m; ++j) {
                double s = 0.0;
                for (int64_t i = 0; i < n; ++i)
                    s += sdi[(size_t)i * m + j] * ((S.f[i] - S.f_m) * (delta_g / S.delta_f) - (g[i] - g_m));
                gz[j] = s;
            }
            if (!inv_small(H.data(), m, Hi.data())) return ORC_SINGULAR_HESSIAN;
            double dp[12];
            for (int r = 0; r < m; ++r) {
                double s = gz[0] * Hi[r * m + 0];
                for (int k = 1; k < m; ++k) s = s + gz[k] * Hi[r * m + k];
                dp[r] = s;
            }
            for (int k = 0; k < m; ++k) p[k] += dp[k];
            double norm = orc_convergence_norm(dp, m, size);
            double zacc = 0.0;
            for (int64_t i = 0; i < n; ++i) {
                double v = (S.f[i] - S.f_m) / S.delta_f - (g[i] - g_m) / delta_g;
                zacc += v * v;
            }
            double c_zncc = 1.0 - zacc / 2.0;
            if (history) { history[4 * (iteration - 1)] = iteration; history[4 * (iteration - 1) + 1] = norm; history[4 * (iteration - 1) + 2] = c_zncc; history[4 * (iteration - 1) + 3] = zacc; }
            it_done = iteration; last_zncc = c_zncc; last_znssd = zacc; last_norm = norm;
            if (norm < max_norm) { converged = true; break; }
        }
    }
    std::memset(R->p, 0, sizeof(R->p));
    std::memcpy(R->p, p, sizeof(double) * m);
    R->c_zncc = last_zncc;
    R->c_znssd = last_znssd;
    R->norm = last_norm;
    R->iterations = it_done;
    R->converged = converged ? 1 : 0;
    R->quality_ok = (converged && last_zncc >= tolerance) ? 1 : 0;
    double g_m_final = unrolled_sum(g.data(), n) / (double)n;
    double dgf = 0.0;
    for (int64_t i = 0; i < n; ++i) { double t = g[i] - g_m_final; dgf += t * t; }
    double delta_g_final = std::sqrt(dgf);
    orc_omitted_mode_diagnostic(S.coord, S.order, S.f_coords.data(), S.grad_f.data(), S.f.data(), S.f_m, S.delta_f,
                                g.data(), g_m_final, delta_g_final, n, &R->eta_u, &R->eta_v);
    return ORC_OK;
}

// Subset::new followed by one solve.  f_out/grad_out/f_coords_out optional.
int orc_subset_solve(const double* fcoef, const double* gcoef, int h, int w, int border, const double* coord,
                     const double* offsets, int64_t n, int order, const double* p0, int p0_len, int method,
                     double tolerance, double max_norm, int max_iterations, OrcResult* R, double* history,
                     double* f_out, double* grad_out, double* f_coords_out) {
    Img fim{fcoef, h, w, border}, gim{gcoef, h, w, border};
    SubsetRef S;
    std::memset(R, 0, sizeof(*R));
    R->order = order;
    R->n_px = (int32_t)n;
    int st = subset_new(fim, coord, offsets, n, order, S);
    if (S.f.size() == (size_t)n && n > 0) {
        R->f_m = S.f_m; R->delta_f = S.delta_f; R->sssig = S.sssig; R->sigma = S.sigma;
        if (f_out) std::memcpy(f_out, S.f.data(), sizeof(double) * n);
        if (grad_out) std::memcpy(grad_out, S.grad_f.data(), sizeof(double) * 2 * n);
        if (f_coords_out) std::memcpy(f_coords_out, S.f_coords.data(), sizeof(double) * 2 * n);
    }
    if (st) { R->status = st; return st; }
    st = subset_solve(gim, S, p0, p0_len, method, tolerance, max_norm, max_iterations, R, history);
    R->status = st;
    return st;
}

// Subset::residual_map at warp p (subset.rs:1350-1392, deformable texture,
// residual_textured 1367-1392): Subset::new, the warped target intensities,
// g_m = Σg / n (ndarray's sum), Δg = √Σ(g - g_m)², then
// e_i = (f_i - f_m)/Δf - (g_i - g_m)/Δg.  p holds p_len entries of `order`.
int orc_residual_map(const double* fcoef, const double* gcoef, int h, int w, int border, const double* coord,
                     const double* offsets, int64_t n, int order, const double* p, double* out) {
    Img fim{fcoef, h, w, border}, gim{gcoef, h, w, border};
    SubsetRef S;
    const int st = subset_new(fim, coord, offsets, n, order, S);
    if (st) return st;
    const int m = p_len(order);
    std::vector<double> gc(2 * n), g(n);
    orc_apply_warp(S.coord, p, m, S.f_coords.data(), n, gc.data());
    for (int64_t i = 0; i < n; ++i) g[i] = img_eval(gim, gc[2 * i], gc[2 * i + 1]);
    const double g_m = unrolled_sum(g.data(), n) / (double)n;
    double dg = 0.0;
    for (int64_t i = 0; i < n; ++i) { const double t = g[i] - g_m; dg += t * t; }
    const double delta_g = std::sqrt(dg);
    for (int64_t i = 0; i < n; ++i) out[i] = (S.f[i] - S.f_m) / S.delta_f - (g[i] - g_m) / delta_g;
    return ORC_OK;
}

// Batch of subsets sharing one offset table per template.  The CPU baseline:
// one subset per task, as mesh.rs:1224-1233 parallelises a colour round.
// tpl_start has n_tpl+1 entries into tpl_xy (pairs of doubles).
int orc_batch_solve(const double* fcoef, const double* gcoef, int h, int w, int border, int64_t n_sub,
                    const double* coords, const int32_t* tpl_index, const double* tpl_xy, const int64_t* tpl_start,
                    const int32_t* orders, const double* p0 /* n x 12 */, int method, double tolerance,
                    double max_norm, int max_iterations, OrcResult* out, int threads) {
    Img fim{fcoef, h, w, border}, gim{gcoef, h, w, border};
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#else
    (void)threads;
#endif
#pragma omp parallel for num_threads(threads) schedule(dynamic, 4)
    for (int64_t s = 0; s < n_sub; ++s) {
        SubsetRef S;
        OrcResult* R = &out[s];
        std::memset(R, 0, sizeof(*R));
        int t = tpl_index[s];
        int64_t n = tpl_start[t + 1] - tpl_start[t];
        R->order = orders[s];
        R->n_px = (int32_t)n;
        int st = subset_new(fim, &coords[2 * s], &tpl_xy[2 * tpl_start[t]], n, orders[s], S);
        if (!st) {
            R->f_m = S.f_m; R->delta_f = S.delta_f; R->sssig = S.sssig; R->sigma = S.sigma;
            st = subset_solve(gim, S, &p0[12 * s], 12, method, tolerance, max_norm, max_iterations, R, nullptr);
        }
        R->status = st;
    }
    return 0;
}

int orc_result_size(void) { return (int)sizeof(OrcResult); }

double orc_wall_seconds(void) {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

}  // extern "C"
