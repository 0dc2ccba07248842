// B-spline prefilter on the GPU: get_c (image.rs:252-327).
//
// C solves the circulant system k*C = P separably, P being the image padded
// by `border` replicated pixels (image.rs:436-443) and k the zero-phase
// 5-tap quintic kernel [1/120, 13/60, 11/20, 13/60, 1/120] (image.rs:
// 422-430).  Each line is deconvolved spectrally: unnormalised forward DFT,
// division by K(w) = 11/20 + (13/30) cos w + (1/60) cos 2w in FP64, inverse
// DFT, times 1/n.  Rows first, then columns, as the reference does.
//
// MI355X layout:
//  * one workgroup per PAIR of lines: two real lines ride in one complex
//    transform (z = a + i b); the operator is real, so Re/Im of the result
//    are the two deconvolved lines;
//  * the line lives in LDS (16 B per complex sample); each stage of a
//    self-sorting Stockham mixed-radix transform computes every output from
//    its R inputs by Horner evaluation in the stage twiddle, so any length
//    whose prime factors are small works (552 = 2^3*3*23, 2088 = 2^3*3^2*29,
//    4136 = 2^3*11*47, 3040 = 2^5*5*19, 1064 = 2^3*7*19);
//  * twiddles come from one FP64 table per length, built on the host in
//    long double and cached in HBM;
//  * the row pass writes the FP64 intermediate row-major; the column pass
//    reads column pairs of it.  Column-pair workgroups are remapped so that
//    each XCD works on a contiguous band of columns and the 128-byte lines
//    it fetches are shared by the four neighbouring pairs in its own L2;
//  * the column pass writes C in FP64 and the FP32 copy the batched engine
//    reads, in the same kernel.
//
// The default path (DIC_PREFILTER_FIR) applies the same spectral division
// in its spatial form.  The inverse of the circulant k is itself circulant:
// its taps h_n[k] = (1/n) Σ_f cos(2πfk/n) / K(2πf/n) are the inverse DFT of
// 1/K, computed once per line length on the host in long double.  They fall
// by a factor ~0.43 per tap (the larger pole of the quintic B-spline,
// -0.4306), to ~1e-19 at |k| = 52, so each line is a 105-tap circular FIR
// whose dropped taps weigh less than 1e-15 of a pixel; a line shorter than
// 105 keeps one tap per residue and is the exact inverse.  That is ~40 %
// fewer FP64 operations than two transforms of lengths with prime factors
// 29 or 47 and needs no barriers between stages:
//  * rows: a workgroup stages 4 padded rows x 623 samples (replicate
//    padding and the circular wrap resolved while loading) in LDS, stored
//    phase-major so that each thread's 8 consecutive outputs read their
//    window without bank conflicts; the window rolls through registers;
//  * columns: each thread produces 16 consecutive outputs of one column from
//    a 127-row window read straight from the row-pass output (64 lanes = 64
//    adjacent columns, so every load is a 512-byte line), with workgroups
//    banded by column per XCD so the overlapping windows meet in one L2.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "internal.h"

namespace geodic {

namespace {

constexpr int kMaxStages = 24;
constexpr int kMaxOut = 8;  // outputs per thread per stage

struct FftStages {
    int n;
    int nst;
    int radix[kMaxStages];
    int ns[kMaxStages];  // product of radices before this stage
};

FftStages make_stages(int n) {
    FftStages s{};
    s.n = n;
    int m = n, ns = 1;
    for (int p = 2; m > 1; ++p) {
        while (m % p == 0) {
            s.radix[s.nst] = p;
            s.ns[s.nst] = ns;
            ns *= p;
            m /= p;
            ++s.nst;
        }
    }
    return s;
}

__device__ __forceinline__ double2 cmul_add(double2 a, double2 w, double2 c) {
    // a*w + c
    return make_double2(fma(a.x, w.x, fma(-a.y, w.y, c.x)), fma(a.x, w.y, fma(a.y, w.x, c.y)));
}

// In-LDS Stockham transform of length n, sign -1 (forward) or +1 (inverse).
// tw[e] = exp(-2πi e/n).
__device__ void fft_lds(double2* buf, const FftStages& st, const double2* __restrict__ tw, bool inverse) {
    const int n = st.n;
    const int nt = blockDim.x;
    const int tid = threadIdx.x;
    for (int s = 0; s < st.nst; ++s) {
        const int R = st.radix[s];
        const int Ns = st.ns[s];
        const int NsR = Ns * R;
        const int NR = n / R;
        const int tstride = n / NsR;
        double2 res[kMaxOut];
#pragma unroll
        for (int c = 0; c < kMaxOut; ++c) {
            const int o = tid + c * nt;
            if (o < n) {
                const int grp = o / NsR;
                const int rem = o - grp * NsR;
                const int jl = rem % Ns;
                const int j = grp * Ns + jl;
                double2 w = tw[(long)rem * tstride];
                if (inverse) w.y = -w.y;
                double2 acc = buf[j + (R - 1) * NR];
                for (int r = R - 2; r >= 0; --r) acc = cmul_add(acc, w, buf[j + r * NR]);
                res[c] = acc;
            }
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < kMaxOut; ++c) {
            const int o = tid + c * nt;
            if (o < n) buf[o] = res[c];
        }
        __syncthreads();
    }
}

// Spectral division by K(w_f), w_f = 2πf/n (FP64).
__device__ void spectral_divide(double2* buf, int n, const double2* __restrict__ tw) {
    for (int f = threadIdx.x; f < n; f += blockDim.x) {
        const double c1 = tw[f].x;
        const double c2 = tw[(int)(((long)2 * f) % n)].x;
        const double K = 11.0 / 20.0 + (13.0 / 30.0) * c1 + (1.0 / 60.0) * c2;
        double2 v = buf[f];
        v.x = v.x / K;
        v.y = v.y / K;
        buf[f] = v;
    }
    __syncthreads();
}

template <typename PixT>
__global__ void __launch_bounds__(1024) prefilter_rows_kernel(const PixT* __restrict__ pix, int64_t h, int64_t w,
                                                              int border, int64_t hp, int64_t wp, FftStages st,
                                                              const double2* __restrict__ tw,
                                                              double* __restrict__ out) {
    extern __shared__ double2 lds[];
    const int64_t ra = 2 * (int64_t)blockIdx.x;
    const int64_t rb = ra + 1;
    const bool hasb = rb < hp;
    const int64_t sa = min(max(ra - border, (int64_t)0), h - 1);
    const int64_t sb = min(max(rb - border, (int64_t)0), h - 1);
    for (int64_t j = threadIdx.x; j < wp; j += blockDim.x) {
        const int64_t sj = min(max(j - border, (int64_t)0), w - 1);
        const double va = (double)pix[sa * w + sj];
        const double vb = hasb ? (double)pix[sb * w + sj] : 0.0;
        lds[j] = make_double2(va, vb);
    }
    __syncthreads();
    fft_lds(lds, st, tw, false);
    spectral_divide(lds, (int)wp, tw);
    fft_lds(lds, st, tw, true);
    const double scale = 1.0 / (double)wp;
    for (int64_t j = threadIdx.x; j < wp; j += blockDim.x) {
        const double2 v = lds[j];
        out[ra * wp + j] = v.x * scale;
        if (hasb) out[rb * wp + j] = v.y * scale;
    }
}

__global__ void __launch_bounds__(1024) prefilter_cols_kernel(const double* __restrict__ tmp, int64_t hp, int64_t wp,
                                                              int npairs, int per_xcd, FftStages st,
                                                              const double2* __restrict__ tw,
                                                              double* __restrict__ c64, float* __restrict__ c32,
                                                              int64_t wp32) {
    extern __shared__ double2 lds[];
    // XCD-aware: dispatch puts workgroup b on XCD b % 8; give each XCD a
    // contiguous band of column pairs.
    const int b = blockIdx.x;
    const int pair = (b & 7) * per_xcd + (b >> 3);
    if (pair >= npairs) return;
    const int64_t ca = 2 * (int64_t)pair;
    const int64_t cb = ca + 1;
    const bool hasb = cb < wp;
    for (int64_t i = threadIdx.x; i < hp; i += blockDim.x) {
        const double va = tmp[i * wp + ca];
        const double vb = hasb ? tmp[i * wp + cb] : 0.0;
        lds[i] = make_double2(va, vb);
    }
    __syncthreads();
    fft_lds(lds, st, tw, false);
    spectral_divide(lds, (int)hp, tw);
    fft_lds(lds, st, tw, true);
    const double scale = 1.0 / (double)hp;
    for (int64_t i = threadIdx.x; i < hp; i += blockDim.x) {
        const double2 v = lds[i];
        const double a = v.x * scale;
        c64[i * wp + ca] = a;
        c32[i * wp32 + ca] = (float)a;
        if (hasb) {
            const double bb = v.y * scale;
            c64[i * wp + cb] = bb;
            c32[i * wp32 + cb] = (float)bb;
        }
    }
}

// ---- circular FIR form ----
constexpr int kFirHalf = 52;                         // taps k = -52..52
constexpr int kFirTaps = 112;                        // tap slots (105 used, zero-padded to whole chunks)
constexpr int kChunk = 8;                            // window samples per chunk (loads in flight together)
constexpr int kRowR = 8;                             // row pass: outputs per thread
constexpr int kRowRows = 4;                          // rows per workgroup, one wavefront each
constexpr int kRowTile = 64 * kRowR;                 // outputs per row per workgroup
constexpr int kRowChunks = (kRowR + kFirTaps - 1 + kChunk - 1) / kChunk;  // 15
constexpr int kRowWin = kRowTile - kRowR + kRowChunks * kChunk;          // 624 samples per row
constexpr int kRowS = kRowWin / kRowR;               // samples per phase (78)
constexpr int kColR = 16;                            // column pass: outputs per thread
constexpr int kColRows = 4 * kColR;                  // rows per workgroup (4 wavefronts)
constexpr int kColChunks = (kColR + kFirTaps - 1 + kChunk - 1) / kChunk; // 16
constexpr int kTapPad = 16;                          // zero taps before slot 0 (covers t = m - r >= -15)
constexpr int kTapLds = kTapPad + kColChunks * kChunk;                   // 144

// Taps in LDS at slot t + kTapPad, zero outside [0, kFirTaps): a chunk's
// products then need no bounds tests.
__device__ __forceinline__ void stage_taps(double* stap, const double* __restrict__ taps) {
    for (int i = threadIdx.x; i < kTapLds; i += blockDim.x) {
        const int t = i - kTapPad;
        stap[i] = (t >= 0 && t < kFirTaps) ? taps[t] : 0.0;
    }
}

// Row pass: out[i][j] = Σ_t taps[t] P[i][(j - k1 + t) mod wp], P the
// replicate-padded image (image.rs:436-443).
template <typename PixT>
__global__ void __launch_bounds__(256) fir_rows_kernel(const PixT* __restrict__ pix, int64_t h, int64_t w, int border,
                                                       int64_t hp, int64_t wp, int k1,
                                                       const double* __restrict__ taps, double* __restrict__ out) {
    __shared__ double win[kRowRows][kRowR * kRowS];
    __shared__ double stap[kTapLds];
    const int lane = threadIdx.x & 63, rr = threadIdx.x >> 6;
    const int64_t row = (int64_t)blockIdx.y * kRowRows + rr;
    const int64_t c0 = (int64_t)blockIdx.x * kRowTile;
    stage_taps(stap, taps);
    double* wr = win[rr];
    if (row < hp) {
        const int64_t si = min(max(row - border, (int64_t)0), h - 1);
        const PixT* src = pix + si * w;
        // padded column c0 - k1 + q lies in [-k1, wp + kRowWin): one wrap
        // either way when the row is at least as long as the window
        const bool one_wrap = wp >= kRowWin;
        for (int q = lane; q < kRowWin; q += 64) {
            int64_t j = c0 - k1 + q;
            if (one_wrap) {
                j = j < 0 ? j + wp : (j >= wp ? j - wp : j);
            } else {
                j %= wp;
                if (j < 0) j += wp;
            }
            const int64_t sj = min(max(j - border, (int64_t)0), w - 1);
            wr[(q % kRowR) * kRowS + q / kRowR] = (double)src[sj];
        }
    }
    __syncthreads();
    if (row >= hp) return;
    // window sample lane*R + m sits at phase m % R, slot lane + m / R; with
    // R = chunk, chunk c's sample u is at phase u, slot lane + c
    double acc[kRowR];
#pragma unroll
    for (int r = 0; r < kRowR; ++r) acc[r] = 0.0;
    const double* wl = wr + lane;
#pragma unroll 1
    for (int c = 0; c < kRowChunks; ++c) {
        double v[kChunk];
#pragma unroll
        for (int u = 0; u < kChunk; ++u) v[u] = wl[u * kRowS + c];
        const double* tp = stap + kTapPad + c * kChunk;
#pragma unroll
        for (int u = 0; u < kChunk; ++u)
#pragma unroll
            for (int r = 0; r < kRowR; ++r) acc[r] = fma(tp[u - r], v[u], acc[r]);
    }
    double* o = out + row * wp + c0 + (int64_t)lane * kRowR;
    const int64_t left = wp - c0 - (int64_t)lane * kRowR;
#pragma unroll
    for (int r = 0; r < kRowR; ++r)
        if (r < left) o[r] = acc[r];
}

// Column pass over the row-pass output, writing C in FP64 and FP32.  Tile
// t = (column band, row block), row blocks fastest, tiles dealt to XCDs in
// contiguous runs.  Window row m of a thread is (i0 - k1 + m) mod hp, at
// most one wrap when the column is at least as long as the window (else a
// remainder), computed per row so a chunk's loads are independent.
__global__ void __launch_bounds__(256) fir_cols_kernel(const double* __restrict__ tmp, int64_t hp, int64_t wp, int k1,
                                                       const double* __restrict__ taps, int row_tiles, int n_tiles,
                                                       int per_xcd, double* __restrict__ c64, float* __restrict__ c32,
                                                       int64_t wp32) {
    __shared__ double stap[kTapLds];
    const int b = blockIdx.x;
    const int tile = (b & 7) * per_xcd + (b >> 3);
    if (tile >= n_tiles) return;
    stage_taps(stap, taps);
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t c = (int64_t)(tile / row_tiles) * 64 + lane;
    const int64_t i0 = (int64_t)(tile % row_tiles) * kColRows + (int64_t)wave * kColR;
    if (i0 >= hp) return;
    const int64_t cl = min(c, wp - 1);  // lanes past the last column load a valid one and store nothing
    int64_t q0 = (i0 - k1) % hp;
    if (q0 < 0) q0 += hp;
    const bool one_wrap = hp >= kColChunks * kChunk;
    const double* col = tmp + cl;
    double acc[kColR];
#pragma unroll
    for (int r = 0; r < kColR; ++r) acc[r] = 0.0;
#pragma unroll 1
    for (int ch = 0; ch < kColChunks; ++ch) {
        double v[kChunk];
#pragma unroll
        for (int u = 0; u < kChunk; ++u) {
            int64_t q = q0 + ch * kChunk + u;
            q = one_wrap ? (q >= hp ? q - hp : q) : q % hp;
            v[u] = col[q * wp];
        }
        const double* tp = stap + kTapPad + ch * kChunk;
#pragma unroll
        for (int u = 0; u < kChunk; ++u)
#pragma unroll
            for (int r = 0; r < kColR; ++r) acc[r] = fma(tp[u - r], v[u], acc[r]);
    }
    if (c >= wp) return;
#pragma unroll
    for (int r = 0; r < kColR; ++r) {
        const int64_t i = i0 + r;
        if (i < hp) {
            c64[i * wp + c] = acc[r];
            c32[i * wp32 + c] = (float)acc[r];
        }
    }
}

// Taps of the circular FIR for line length n (device, kFirTaps doubles, slot
// t holding offset t - k1), cached per length like the twiddles.
int get_fir_taps(Engine* e, int n, const double** out, int* k1_out) {
    for (auto& pr : e->fir_taps)
        if (pr.first == n) {
            *out = pr.second;
            *k1_out = n >= 2 * kFirHalf + 1 ? kFirHalf : n / 2;
            return DIC_OK;
        }
    const int k1 = n >= 2 * kFirHalf + 1 ? kFirHalf : n / 2;
    const int nt = n >= 2 * kFirHalf + 1 ? 2 * kFirHalf + 1 : n;  // one tap per residue when short
    const int kmax = std::max(k1, nt - 1 - k1);
    // h[k] = (1/n) Σ_f cos(kθ_f) / K(θ_f), cos(kθ) by the Chebyshev recurrence
    std::vector<long double> hk(kmax + 1, 0.0L);
    const long double two_pi = 6.283185307179586476925286766559005768L;
    for (int f = 0; f < n; ++f) {
        const long double c1 = cosl(two_pi * (long double)f / (long double)n);
        const long double K = 11.0L / 20.0L + (13.0L / 30.0L) * c1 + (1.0L / 60.0L) * (2.0L * c1 * c1 - 1.0L);
        const long double rk = 1.0L / K;
        long double cm = 1.0L, cc = c1;
        hk[0] += rk;
        for (int k = 1; k <= kmax; ++k) {
            hk[k] += cc * rk;
            const long double nx = 2.0L * c1 * cc - cm;
            cm = cc;
            cc = nx;
        }
    }
    std::vector<double> h(kFirTaps, 0.0);
    for (int t = 0; t < nt; ++t) h[t] = (double)(hk[std::abs(t - k1)] / (long double)n);
    double* d = nullptr;
    DIC_HIP_TRY(hipMalloc(&d, sizeof(double) * kFirTaps));
    DIC_HIP_TRY(hipMemcpyAsync(d, h.data(), sizeof(double) * kFirTaps, hipMemcpyHostToDevice, e->stream));
    DIC_HIP_TRY(hipStreamSynchronize(e->stream));
    e->fir_taps.push_back({n, d});
    *out = d;
    *k1_out = k1;
    return DIC_OK;
}

int threads_for(int n) {
    int t = 64;
    while (t * kMaxOut < n && t < 1024) t *= 2;
    return t;
}

int get_twiddles(Engine* e, int n, const double2** out) {
    for (auto& pr : e->twiddles)
        if (pr.first == n) {
            *out = pr.second;
            return DIC_OK;
        }
    std::vector<double2> h(n);
    const long double two_pi = 6.283185307179586476925286766559005768L;
    for (int k = 0; k < n; ++k) {
        long double a = -two_pi * (long double)k / (long double)n;
        h[k] = make_double2((double)cosl(a), (double)sinl(a));
    }
    double2* d = nullptr;
    DIC_HIP_TRY(hipMalloc(&d, sizeof(double2) * n));
    DIC_HIP_TRY(hipMemcpyAsync(d, h.data(), sizeof(double2) * n, hipMemcpyHostToDevice, e->stream));
    DIC_HIP_TRY(hipStreamSynchronize(e->stream));
    e->twiddles.push_back({n, d});
    *out = d;
    return DIC_OK;
}

}  // namespace

namespace {

// The file loader's 5 Gaussian taps, sigma 1.1 (image.rs:483-491): exp of
// -x²/(2σ²) at x = -2..2, normalised by their sum taken in order.  Host
// libm, as the reference's f64::exp.
struct BlurTaps {
    double k[5];
};

BlurTaps loader_taps() {
    const double sigma = 1.1;
    BlurTaps t;
    double raw[5], sum = 0.0;
    for (int i = 0; i < 5; ++i) {
        const double x = (double)i - 2.0;
        raw[i] = std::exp(-x * x / (2.0 * sigma * sigma));
    }
    for (int i = 0; i < 5; ++i) sum += raw[i];
    for (int i = 0; i < 5; ++i) t.k[i] = raw[i] / sum;
    return t;
}

// The file loader's conversion of a decoded frame (image.rs:450-527): each
// pixel's BT.709 integer grey (2126 R + 7152 G + 722 B + 5000) / 10000
// (a grey frame is its own grey: the formula returns the value of an equal
// R, G, B), the horizontal then the vertical 5-tap pass with reflect-101
// borders in FP64, each sum in tap order without contraction, rounded half
// away from zero and clamped to 0..255.  One thread per output pixel,
// recomputing the five horizontal sums of its column window (the 25 grey
// samples come from cache); h, w >= 3 so one reflection stays in range.
template <int C>
__global__ void __launch_bounds__(256) loader_convert_kernel(const uint8_t* __restrict__ src, int64_t h, int64_t w,
                                                             BlurTaps t, uint8_t* __restrict__ out) {
#pragma clang fp contract(off)
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = blockIdx.y;
    if (j >= w) return;
    auto reflect = [](int64_t x, int64_t n) { return x < 0 ? -x : (x >= n ? 2 * (n - 1) - x : x); };
    int64_t cols[5];
    for (int kj = 0; kj < 5; ++kj) cols[kj] = reflect(j + kj - 2, w);
    double acc = 0.0;
    for (int ki = 0; ki < 5; ++ki) {
        const uint8_t* row = src + reflect(i + ki - 2, h) * w * C;
        double hs = 0.0;
        for (int kj = 0; kj < 5; ++kj) {
            const uint8_t* p = row + cols[kj] * C;
            const uint32_t g = C == 1 ? (uint32_t)p[0]
                                      : (2126u * p[0] + 7152u * p[1] + 722u * p[2] + 5000u) / 10000u;
            hs = hs + t.k[kj] * (double)g;
        }
        acc = acc + t.k[ki] * hs;
    }
    out[i * w + j] = (uint8_t)fmin(fmax(round(acc), 0.0), 255.0);
}

}  // namespace

namespace {

// New coefficients: the knot table of an image that solves read as a
// reference follows them, any other image's is stale until remade.
int knot_refresh(Engine* e, Image* img) {
    img->knot_valid = false;
    return img->knot_wanted ? knot_table_run(e, img) : DIC_OK;
}

}  // namespace

int prefilter_run(Engine* e, Image* img, const void* pix, int pix_type, int mem, double* tmp_own) {
    const int64_t h = img->h, w = img->w, hp = img->hp, wp = img->wp;
    if (e->prefilter_method == DIC_PREFILTER_FFT && (wp > (int64_t)kMaxOut * 1024 || hp > (int64_t)kMaxOut * 1024)) {
        set_error("prefilter: padded line longer than 8192 samples");
        return DIC_ERR_INVALID_INPUT;
    }
    const size_t pix_bytes = (size_t)h * w * pixel_bytes(pix_type);
    const void* dpix = pix;
    if (mem == DIC_MEM_HOST) {
        if (e->pix_stage_bytes < pix_bytes) {
            if (e->pix_stage) DIC_HIP_TRY(hipFree(e->pix_stage));
            e->pix_stage = nullptr;
            e->pix_stage_bytes = 0;
            DIC_HIP_TRY(hipMalloc(&e->pix_stage, pix_bytes));
            e->pix_stage_bytes = pix_bytes;
        }
        DIC_HIP_TRY(hipMemcpyAsync(e->pix_stage, pix, pix_bytes, hipMemcpyHostToDevice, e->stream));
        dpix = e->pix_stage;
    }
    const size_t tmp_bytes = (size_t)hp * wp * sizeof(double);
    if (!tmp_own && e->pre_tmp_bytes < tmp_bytes) {
        if (e->pre_tmp) DIC_HIP_TRY(hipFree(e->pre_tmp));
        e->pre_tmp = nullptr;
        e->pre_tmp_bytes = 0;
        DIC_HIP_TRY(hipMalloc(&e->pre_tmp, tmp_bytes));
        e->pre_tmp_bytes = tmp_bytes;
    }
    const bool fft = e->prefilter_method == DIC_PREFILTER_FFT;
    const double2 *tw_r = nullptr, *tw_c = nullptr;
    const double *taps_r = nullptr, *taps_c = nullptr;
    int k1_r = 0, k1_c = 0;
    FftStages sr{}, sc{};
    int st = DIC_OK;
    if (fft) {
        st = get_twiddles(e, (int)wp, &tw_r);
        if (!st) st = get_twiddles(e, (int)hp, &tw_c);
        if (st) return st;
        sr = make_stages((int)wp);
        sc = make_stages((int)hp);
        if (sr.nst > kMaxStages || sc.nst > kMaxStages) {
            set_error("prefilter: too many FFT stages");
            return DIC_ERR_INVALID_INPUT;
        }
    } else {
        st = get_fir_taps(e, (int)wp, &taps_r, &k1_r);
        if (!st) st = get_fir_taps(e, (int)hp, &taps_c, &k1_c);
        if (st) return st;
    }
    const int tr = threads_for((int)wp), tc = threads_for((int)hp);
    const size_t lr = sizeof(double2) * wp, lc = sizeof(double2) * hp;
    double* tmp = tmp_own ? tmp_own : e->pre_tmp;
    if (loader_pixels(pix_type) && !img->grey_dev) {
        const int st0 = dev_alloc(e, (void**)&img->grey_dev, (size_t)h * w);
        if (st0) return st0;
    }
    if (!e->capturing) DIC_HIP_TRY(hipEventRecord(e->ev_pre0, e->stream));
    if (loader_pixels(pix_type)) {
        // decoded file pixels: the loader's conversion, then the prefilter of
        // its 8-bit result
        const dim3 grid((unsigned)((w + 255) / 256), (unsigned)h);
        if (pix_type == DIC_PIX_RGB8_LOAD)
            loader_convert_kernel<3><<<grid, 256, 0, e->stream>>>((const uint8_t*)dpix, h, w, loader_taps(), img->grey_dev);
        else
            loader_convert_kernel<1><<<grid, 256, 0, e->stream>>>((const uint8_t*)dpix, h, w, loader_taps(), img->grey_dev);
        DIC_HIP_TRY(hipGetLastError());
        dpix = img->grey_dev;
        pix_type = DIC_PIX_U8;
    }
    if (!fft) {
        const dim3 rgrid((unsigned)((wp + kRowTile - 1) / kRowTile), (unsigned)((hp + kRowRows - 1) / kRowRows));
        if (pix_type == DIC_PIX_U8)
            fir_rows_kernel<uint8_t><<<rgrid, 256, 0, e->stream>>>((const uint8_t*)dpix, h, w, img->border, hp, wp,
                                                                   k1_r, taps_r, tmp);
        else
            fir_rows_kernel<double><<<rgrid, 256, 0, e->stream>>>((const double*)dpix, h, w, img->border, hp, wp,
                                                                  k1_r, taps_r, tmp);
        DIC_HIP_TRY(hipGetLastError());
        const int row_tiles = (int)((hp + kColRows - 1) / kColRows);
        const int n_tiles = row_tiles * (int)((wp + 63) / 64);
        const int per_xcd = (n_tiles + 7) / 8;
        fir_cols_kernel<<<per_xcd * 8, 256, 0, e->stream>>>(tmp, hp, wp, k1_c, taps_c, row_tiles, n_tiles, per_xcd,
                                                            img->coef64, img->coef32, img->wp32);
        DIC_HIP_TRY(hipGetLastError());
        const int kt = knot_refresh(e, img);
        if (kt) return kt;
        if (!e->capturing) {
            DIC_HIP_TRY(hipEventRecord(e->ev_pre1, e->stream));
            e->pre_recorded = true;
        }
        return DIC_OK;
    }
    const int row_blocks = (int)((hp + 1) / 2);
    if (pix_type == DIC_PIX_U8) {
        DIC_HIP_TRY(hipFuncSetAttribute((const void*)prefilter_rows_kernel<uint8_t>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lr));
        prefilter_rows_kernel<uint8_t><<<row_blocks, tr, lr, e->stream>>>((const uint8_t*)dpix, h, w, img->border, hp,
                                                                          wp, sr, tw_r, tmp);
    } else {
        DIC_HIP_TRY(hipFuncSetAttribute((const void*)prefilter_rows_kernel<double>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lr));
        prefilter_rows_kernel<double><<<row_blocks, tr, lr, e->stream>>>((const double*)dpix, h, w, img->border, hp,
                                                                         wp, sr, tw_r, tmp);
    }
    DIC_HIP_TRY(hipGetLastError());
    const int npairs = (int)((wp + 1) / 2);
    const int per_xcd = (npairs + 7) / 8;
    DIC_HIP_TRY(hipFuncSetAttribute((const void*)prefilter_cols_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lc));
    prefilter_cols_kernel<<<per_xcd * 8, tc, lc, e->stream>>>(tmp, hp, wp, npairs, per_xcd, sc, tw_c,
                                                              img->coef64, img->coef32, img->wp32);
    DIC_HIP_TRY(hipGetLastError());
    const int kt = knot_refresh(e, img);
    if (kt) return kt;
    if (!e->capturing) {
        DIC_HIP_TRY(hipEventRecord(e->ev_pre1, e->stream));
        e->pre_recorded = true;
    }
    return DIC_OK;
}

}  // namespace geodic
