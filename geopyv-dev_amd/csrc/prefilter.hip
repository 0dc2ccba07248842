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
#include <cmath>
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

int prefilter_run(Engine* e, Image* img, const void* pix, int pix_type, int mem, double* tmp_own) {
    const int64_t h = img->h, w = img->w, hp = img->hp, wp = img->wp;
    if (wp > (int64_t)kMaxOut * 1024 || hp > (int64_t)kMaxOut * 1024) {
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
    const double2 *tw_r = nullptr, *tw_c = nullptr;
    int st = get_twiddles(e, (int)wp, &tw_r);
    if (st) return st;
    st = get_twiddles(e, (int)hp, &tw_c);
    if (st) return st;
    FftStages sr = make_stages((int)wp), sc = make_stages((int)hp);
    if (sr.nst > kMaxStages || sc.nst > kMaxStages) {
        set_error("prefilter: too many FFT stages");
        return DIC_ERR_INVALID_INPUT;
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
    if (!e->capturing) {
        DIC_HIP_TRY(hipEventRecord(e->ev_pre1, e->stream));
        e->pre_recorded = true;
    }
    return DIC_OK;
}

}  // namespace geodic
