// Batched engine: Subset::new + the ICGN solve (subset.rs:836-919,
// 1117-1219) for many subsets, restructured for CDNA4.
//
//  * one workgroup of WPS wavefronts per subset, dispatched largest first
//    so uneven subsets balance across the CUs; thread t owns pixels
//    t, t + 64·WPS, ... and keeps their reference state in VGPRs for the
//    whole solve: f - f_m and ∇f in FP32 (P pixels per thread, walked by a
//    rolled loop whose uniform index the compiler turns into register
//    indexing, so nothing spills and nothing streams from HBM);
//  * every iteration stages the coefficient window the warped subset
//    touches (bounded analytically from the warp and the template extents)
//    from the FP32 copy of C into LDS with row-coalesced loads, and every
//    pixel gathers its 6x6 block from LDS; a window that does not fit the
//    LDS budget falls back to gathering from HBM for that iteration;
//  * positions are an integer base (the floor of the subset centre) plus an
//    FP32 offset relative to it, so the FP32 basis never sees an absolute
//    coordinate; interpolation and residuals are FP32;
//  * each iteration is ONE pass over the pixels.  With g' = g - f_m (a
//    shift by the reference mean keeps the sums well conditioned) the pass
//    accumulates, in FP64,
//        s1 = Σ g',  s2 = Σ g'²,  s3 = Σ a·g',  B_j = Σ sdi_j·g'   (a = f - f_m)
//    and the reference's per-pixel forms (subset.rs:1162-1176, 418-446)
//    follow exactly, summed in another order:
//        g_m = f_m + s1/n,   Δg² = s2 - s1²/n,
//        ∇ZNSSD_j = Σ sdi_j (a - (Δf/Δg)(g - g_m)) = A_j - (Δf/Δg)(B_j - (s1/n) S_j),
//        ZNSSD = 2 - 2 (s3 - (s1/n) Σa) / (Δf Δg),
//    with A_j = Σ sdi_j a and S_j = Σ sdi_j fixed at setup;
//  * every sum accumulates in FP64: on CDNA4 v_fma_f64 issues at the rate of
//    an unpacked FP32 FMA, so FP64 accumulation costs no more than a
//    compensated FP32 two-sum and is at least as accurate.  Wavefront sums
//    are DPP trees, cross-wave sums go through LDS in fixed wave order;
//  * the Hessian is factorised once (FP64 Cholesky) and each update
//    Δp = -(LLᵀ)⁻¹∇ZNSSD and the composition run in FP64;
//  * the final ZNSSD and warp-adequacy diagnostic re-evaluate g once at the
//    warp of the last iteration (the reference's final g buffer).
#pragma once

#include "bspline.h"
#include "linalg.h"
#include "../../include/geodic.h"

// Resident wavefronts per SIMD the register budget must allow: 4 (128
// VGPRs) for orders 0 and 1, 2 for the 12-parameter warp.
#ifndef GEODIC_BATCH_ATTRS
#define GEODIC_BATCH_ATTRS(WPS, OCC) __launch_bounds__(64 * (WPS)) __attribute__((amdgpu_waves_per_eu(OCC, 8)))
#endif
#ifndef GEODIC_DYN_LDS
#define GEODIC_DYN_LDS(T, name) extern __shared__ T name[]
#endif

namespace geodic {
namespace batch {

struct BatchArgs {
    const float* fC;
    const float* gC;
    int64_t wp;
    int border;
    int w, h;
    double wmax, hmax;
    int64_t n_sub;
    const double* coords;
    const int32_t* tpl_index;
    const int64_t* tpl_start;
    const int32_t* tpl_xy;
    const int32_t* tpl_ext;  // 4 per template: dxmin, dxmax, dymin, dymax
    const int32_t* perm;
    const double* p0;
    double* p_out;
    dic_subset_result* res;
    double* hist;
    int hist_cap;
    int max_it;
    double tol;
    double max_norm;
    int win_cap;  // floats of LDS for the coefficient window
};

constexpr int kRedSlots = 40;  // doubles per wavefront in the cross-wave reduction buffer

template <int M>
struct Modes {
    static constexpr int K = (M == 2) ? 2 : (M == 6 ? 3 : 4);
};

// warped local offset (relative to the subset centre), FP64, for bounds
template <int M>
__device__ __forceinline__ void warp_box(const double* p, const int32_t e[4], double& x0, double& x1, double& y0,
                                         double& y1) {
    const double cx[2] = {(double)e[0], (double)e[1]}, cy[2] = {(double)e[2], (double)e[3]};
    x0 = y0 = 1e300;
    x1 = y1 = -1e300;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            double ox, oy;
            if (M == 2) {
                ox = cx[a] + p[0];
                oy = cy[b] + p[1];
            } else {
                ox = cx[a] + p[0] + p[2] * cx[a] + p[4] * cy[b];
                oy = cy[b] + p[1] + p[3] * cx[a] + p[5] * cy[b];
            }
            x0 = fmin(x0, ox);
            x1 = fmax(x1, ox);
            y0 = fmin(y0, oy);
            y1 = fmax(y1, oy);
        }
    if (M == 12) {
        const double mx = fmax(fabs(cx[0]), fabs(cx[1])), my = fmax(fabs(cy[0]), fabs(cy[1]));
        const double qx = 0.5 * fabs(p[6]) * mx * mx + fabs(p[8]) * mx * my + 0.5 * fabs(p[10]) * my * my;
        const double qy = 0.5 * fabs(p[7]) * mx * mx + fabs(p[9]) * mx * my + 0.5 * fabs(p[11]) * my * my;
        x0 -= qx;
        x1 += qx;
        y0 -= qy;
        y1 += qy;
    }
}

struct Window {
    int x0, y0, ww, wh;  // origin in coefficient space and size
    bool lds;
};

// Window of absolute positions [ax0, ax1] x [ay0, ay1] after the clamp.
__device__ __forceinline__ Window make_window(double ax0, double ax1, double ay0, double ay1, const BatchArgs& A) {
    // margin covers the FP32 rounding of per-pixel positions
    const double e = 1e-3;
    ax0 = clamp_ref(ax0 - e, 0.0, A.wmax);
    ax1 = clamp_ref(ax1 + e, 0.0, A.wmax);
    ay0 = clamp_ref(ay0 - e, 0.0, A.hmax);
    ay1 = clamp_ref(ay1 + e, 0.0, A.hmax);
    Window W;
    // a NaN bound (diverged warp) gives an empty window: every gather then
    // goes to HBM through the bounds-safe index path of eval32
    if (!(ax0 <= ax1) || !(ay0 <= ay1)) {
        W.x0 = W.y0 = W.ww = W.wh = 0;
        W.lds = false;
        return W;
    }
    const int ix0 = (int)floor(ax0), ix1 = (int)floor(ax1);
    const int iy0 = (int)floor(ay0), iy1 = (int)floor(ay1);
    W.x0 = uni(ix0 + A.border - 2);
    W.y0 = uni(iy0 + A.border - 2);
    W.ww = uni(ix1 - ix0 + 6);
    W.wh = uni(iy1 - iy0 + 6);
    W.lds = (W.ww * W.wh) <= A.win_cap;
    return W;
}

// Stage C[y0:y0+wh, x0:x0+ww] into LDS: thread t takes column t % 64 of
// row block t / 64 (so each row's loads coalesce) and walks rows with a
// pointer increment; eight rows in flight per group.
template <int NT>
__device__ __forceinline__ void stage_window(float* win, const float* __restrict__ C, int64_t wp, const Window& W,
                                             int tid) {
    constexpr int RB = NT / 64;  // rows handled concurrently by the workgroup
    const int c0 = tid & 63, rb = tid >> 6;
    const float* __restrict__ src0 = C + (int64_t)W.y0 * wp + W.x0;
    const int wpi = (int)wp;  // row offsets within a window fit in 32 bits
    for (int cb = 0; cb < W.ww; cb += 64) {
        const int c = cb + c0;
        if (c < W.ww) {
            int r = rb;
            for (; r + 7 * RB < W.wh; r += 8 * RB) {
                float v[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) v[k] = src0[(r + k * RB) * wpi + c];
#pragma unroll
                for (int k = 0; k < 8; ++k) win[(r + k * RB) * W.ww + c] = v[k];
            }
            for (; r < W.wh; r += RB) win[r * W.ww + c] = src0[r * wpi + c];
        }
    }
}

// Cell and fraction along one axis: base + local offset, clamped into the
// image as image.rs:130-132 clamps (x to [0, n - 1e-9]).
__device__ __forceinline__ void cell_frac(int base, float local, int n, int& cell, float& t) {
    const float fl = floorf(local);
    // a NaN or huge offset lands in cell 0 with a NaN fraction (the
    // reference's saturating cast); the index always stays in range
    const int ci = (fl > -1.0e9f && fl < 1.0e9f) ? (int)fl : 0;
    cell = base + ci;
    t = local - fl;
    if (cell < 0) {
        cell = 0;
        t = 0.f;
    } else if (cell > n - 1) {
        cell = n - 1;
        t = 1.f;
    }
}

// FP32 value (and optionally gradient) at cell (cx, cy) + (tx, ty).
template <bool GRAD>
__device__ __forceinline__ float eval32(const float* win, const float* __restrict__ C, int64_t wp, const Window& W,
                                        int border, int cellx, int celly, float tx, float ty, float& gx, float& gy) {
    float a[6], b[6];
    basis_f32(ty, a);
    basis_f32(tx, b);
    const int lx = cellx + border - 2 - W.x0, ly = celly + border - 2 - W.y0;
    const bool in = W.lds && lx >= 0 && ly >= 0 && lx <= W.ww - 6 && ly <= W.wh - 6;
    float val = 0.f, ggx = 0.f, ggy = 0.f;
    float da[6], db[6];
    if (GRAD) {
        dbasis_f32(ty, da);
        dbasis_f32(tx, db);
    }
    // Two explicit gather paths so each compiles to its own address space
    // (ds_read from the staged window, global_load otherwise).
    if (in) {
        const float* base = win + ly * W.ww + lx;
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            const float* row = base + r * W.ww;
            float s = 0.f, t = 0.f;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                s = fmaf(row[k], b[k], s);
                if (GRAD) t = fmaf(row[k], db[k], t);
            }
            val = fmaf(a[r], s, val);
            if (GRAD) {
                ggx = fmaf(a[r], t, ggx);
                ggy = fmaf(da[r], s, ggy);
            }
        }
    } else {
        const float* __restrict__ base = C + (int64_t)(celly + border - 2) * wp + (cellx + border - 2);
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            const float* row = base + r * wp;
            float s = 0.f, t = 0.f;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                s = fmaf(row[k], b[k], s);
                if (GRAD) t = fmaf(row[k], db[k], t);
            }
            val = fmaf(a[r], s, val);
            if (GRAD) {
                ggx = fmaf(a[r], t, ggx);
                ggy = fmaf(da[r], s, ggy);
            }
        }
    }
    gx = ggx;
    gy = ggy;
    return val;
}

// Sums V values over the workgroup: DPP tree per wavefront, then the
// wavefront partials in wave order through LDS.  Result uniform (SGPRs).
template <int WPS, int V>
__device__ __forceinline__ void block_sum(double* v, double* sRed, int lane, int wave) {
    static_assert(V <= kRedSlots, "reduction buffer");
#pragma unroll
    for (int j = 0; j < V; ++j) v[j] = wave_sum_u(v[j]);
    if constexpr (WPS > 1) {
        if (lane == 0) {
#pragma unroll
            for (int j = 0; j < V; ++j) sRed[wave * kRedSlots + j] = v[j];
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < V; ++j) {
            double t = sRed[j];
#pragma unroll
            for (int w = 1; w < WPS; ++w) t += sRed[w * kRedSlots + j];
            v[j] = uni(t);
        }
        __syncthreads();
    }
}

template <int M>
__device__ __forceinline__ void sdi64(double gx, double gy, double dx, double dy, double* s) {
    s[0] = gx;
    s[1] = gy;
    if (M >= 6) {
        s[2] = gx * dx;
        s[3] = gy * dx;
        s[4] = gx * dy;
        s[5] = gy * dy;
    }
    if (M == 12) {
        const double hx = 0.5 * dx * dx, hy = 0.5 * dy * dy, xy = dx * dy;
        s[6] = gx * hx;
        s[7] = gy * hx;
        s[8] = gx * xy;
        s[9] = gy * xy;
        s[10] = gx * hy;
        s[11] = gy * hy;
    }
}

// Hessian rows [A0, A1) from the register-resident gradients.
template <int M, int P, int WPS, int A0, int A1>
__device__ __forceinline__ void hessian_rows(const float* gxr, const float* gyr, const int32_t* __restrict__ offs,
                                             int n, int tid, int lane, int wave, double* sH, double* sRed) {
    constexpr int NT = 64 * WPS;
    constexpr int NE = (A1 - A0) * M - ((A1 - 1) * A1 - (A0 - 1) * A0) / 2;
    double acc[NE > 0 ? NE : 1];
#pragma unroll
    for (int q = 0; q < NE; ++q) acc[q] = 0.0;
#pragma unroll 1
    for (int k = 0; k < P; ++k) {
        const int i = tid + NT * k;
        if (i < n) {
            const int32_t o = offs[i];
            double sd[M];
            sdi64<M>((double)gxr[k], (double)gyr[k], (double)unpack_dx(o), (double)unpack_dy(o), sd);
            int q = 0;
#pragma unroll
            for (int a = A0; a < A1; ++a)
#pragma unroll
                for (int b = a; b < M; ++b) {
                    acc[q] = fma(sd[a], sd[b], acc[q]);
                    ++q;
                }
        }
    }
    block_sum<WPS, NE>(acc, sRed, lane, wave);
    if (tid == 0) {
        int q = 0;
#pragma unroll
        for (int a = A0; a < A1; ++a)
#pragma unroll
            for (int b = a; b < M; ++b) {
                sH[a * M + b] = acc[q];
                sH[b * M + a] = acc[q];
                ++q;
            }
    }
}

// warped offset of pixel (dx, dy) relative to the integer base, FP32
template <int M>
__device__ __forceinline__ void warp_local32(const float* q, float fx, float fy, float dx, float dy, float& ox,
                                             float& oy) {
    // q[0..1] = u, v; q[2..5] = u_x, v_x, u_y, v_y; q[6..11] second order
    float wx = q[0], wy = q[1];
    if (M >= 6) {
        wx = fmaf(q[2], dx, fmaf(q[4], dy, wx));
        wy = fmaf(q[3], dx, fmaf(q[5], dy, wy));
    }
    if (M == 12) {
        const float hx = 0.5f * dx * dx, hy = 0.5f * dy * dy, xy = dx * dy;
        wx = fmaf(q[6], hx, fmaf(q[8], xy, fmaf(q[10], hy, wx)));
        wy = fmaf(q[7], hx, fmaf(q[9], xy, fmaf(q[11], hy, wy)));
    }
    ox = fx + dx + wx;
    oy = fy + dy + wy;
}

template <int M, int P, int WPS>
__global__ void GEODIC_BATCH_ATTRS(WPS, (M == 12 ? 1 : 3)) icgn_batch_kernel(BatchArgs A) {
    constexpr int NT = 64 * WPS;
    constexpr int NL = M * (M + 1) / 2;
    constexpr int K = Modes<M>::K;
    GEODIC_DYN_LDS(float, win);
    __shared__ double sH[M * M];
    __shared__ double sL[NL];
    __shared__ double sScr[144];
    __shared__ double sP[12];
    __shared__ double sRed[WPS * kRedSlots];
    __shared__ int sCtl[2];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const double qnan = __longlong_as_double(0x7ff8000000000000LL);

    // One workgroup per subset, dispatched largest first (perm).  A lane-0
    // atomic work counter broadcast with a shuffle is avoided on purpose:
    // the compiler restructures that divergent loop so that its back-edge
    // skips the atomic.
    const int64_t idx = blockIdx.x;
    if (idx >= A.n_sub) return;
    const int64_t s = A.perm[idx];
    const int t = A.tpl_index[s];
    const int64_t st = A.tpl_start[t];
    const int n = (int)(A.tpl_start[t + 1] - st);
    const int32_t* __restrict__ offs = A.tpl_xy + st;
    int32_t ext[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) ext[q] = A.tpl_ext[4 * t + q];
    const double xc = A.coords[2 * s], yc = A.coords[2 * s + 1];
    // integer base and FP32 fraction of the centre
    const double bxd = floor(xc), byd = floor(yc);
    const bool centre_ok = bxd > -1.0e9 && bxd < 1.0e9 && byd > -1.0e9 && byd < 1.0e9;
    const int bx = centre_ok ? (int)bxd : 0, by = centre_ok ? (int)byd : 0;
    const float fxc = (float)(xc - bxd), fyc = (float)(yc - byd);
    const double dn = (double)n;
    dic_subset_result* R = A.res + s;
    if (tid == 0) {
        R->n_px = n;
        R->order = (M == 2) ? 0 : (M == 6 ? 1 : 2);
        R->iterations = 0;
        R->converged = 0;
        R->quality_ok = 0;
        R->c_zncc = 0.0;
        R->c_znssd = 4.0;
        R->norm = qnan;
        R->eta_u = R->eta_v = 0.0;
        R->f_m = R->delta_f = R->sssig = R->sigma = 0.0;
        for (int k = 0; k < 12; ++k) R->p[k] = 0.0;
        R->status = 0;
    }
    if (n == 0 || n > NT * P || !centre_ok) {
        if (tid == 0) R->status = DIC_ERR_INVALID_INPUT;
        return;
    }

    float am[P], gxr[P], gyr[P];

    // ---- Subset::new on the staged reference window ----
    double f_m, delta_f, sa;
    double Aj[M], Sj[M];
    {
        const Window W = make_window(xc + ext[0], xc + ext[1], yc + ext[2], yc + ext[3], A);
        if (W.lds) stage_window<NT>(win, A.fC, A.wp, W, tid);
        __syncthreads();
        double red[2] = {0.0, 0.0};  // Σf, SSSIG
#pragma unroll 1
        for (int k = 0; k < P; ++k) {
            const int i = tid + NT * k;
            float v = 0.f, gx = 0.f, gy = 0.f;
            if (i < n) {
                const int32_t o = offs[i];
                int cx, cy;
                float tx, ty;
                cell_frac(bx, fxc + (float)unpack_dx(o), A.w, cx, tx);
                cell_frac(by, fyc + (float)unpack_dy(o), A.h, cy, ty);
                v = eval32<true>(win, A.fC, A.wp, W, A.border, cx, cy, tx, ty, gx, gy);
                red[0] += (double)v;
                red[1] += 0.5 * ((double)gx * gx + (double)gy * gy);
            }
            am[k] = v;
            gxr[k] = gx;
            gyr[k] = gy;
        }
        block_sum<WPS, 2>(red, sRed, lane, wave);
        f_m = uni(red[0] / dn);
        const double ss = red[1];
        double v2[3 + 2 * M];
#pragma unroll
        for (int j = 0; j < 3 + 2 * M; ++j) v2[j] = 0.0;
#pragma unroll 1
        for (int k = 0; k < P; ++k) {
            const int i = tid + NT * k;
            if (i < n) {
                const int32_t o = offs[i];
                const double a = (double)am[k] - f_m;
                v2[0] = fma(a, a, v2[0]);
                v2[1] += a;
                double sd[M];
                sdi64<M>((double)gxr[k], (double)gyr[k], (double)unpack_dx(o), (double)unpack_dy(o), sd);
#pragma unroll
                for (int j = 0; j < M; ++j) {
                    v2[3 + j] = fma(sd[j], a, v2[3 + j]);
                    v2[3 + M + j] += sd[j];
                }
                am[k] = (float)a;
            }
        }
        block_sum<WPS, 3 + 2 * M>(v2, sRed, lane, wave);
        const double sq = v2[0];
        sa = v2[1];
#pragma unroll
        for (int j = 0; j < M; ++j) {
            Aj[j] = v2[3 + j];
            Sj[j] = v2[3 + M + j];
        }
        delta_f = uni(sqrt(sq));
        if (tid == 0) {
            R->f_m = f_m;
            R->delta_f = delta_f;
            R->sssig = ss;
            R->sigma = sqrt(sq / dn);
        }
    }
    if (delta_f == 0.0) {
        if (tid == 0) R->status = DIC_ERR_INVALID_INPUT;
        return;
    }
    if constexpr (M == 12) {
        hessian_rows<M, P, WPS, 0, 3>(gxr, gyr, offs, n, tid, lane, wave, sH, sRed);
        hessian_rows<M, P, WPS, 3, 6>(gxr, gyr, offs, n, tid, lane, wave, sH, sRed);
        hessian_rows<M, P, WPS, 6, 12>(gxr, gyr, offs, n, tid, lane, wave, sH, sRed);
    } else if constexpr (M == 6) {
        hessian_rows<M, P, WPS, 0, 2>(gxr, gyr, offs, n, tid, lane, wave, sH, sRed);
        hessian_rows<M, P, WPS, 2, 6>(gxr, gyr, offs, n, tid, lane, wave, sH, sRed);
    } else {
        hessian_rows<M, P, WPS, 0, M>(gxr, gyr, offs, n, tid, lane, wave, sH, sRed);
    }
    __syncthreads();
    if (tid == 0) sCtl[0] = cholesky<M>(sH, sL) ? 0 : DIC_ERR_SINGULAR_HESSIAN;
    __syncthreads();
    if (sCtl[0]) {
        if (tid == 0) R->status = DIC_ERR_SINGULAR_HESSIAN;
        return;
    }
    const double size = sqrt(dn);
    const float f_m32 = (float)f_m;
    double p[M], p_last[M];
#pragma unroll
    for (int k = 0; k < M; ++k) p_last[k] = p[k] = uni(A.p0[12 * s + k]);

    int it_done = 0;
    bool converged = false;
    int status = 0;
    double g_m = 0.0, delta_g = 0.0, last_norm = 0.0;
    for (int it = 1; it <= A.max_it; ++it) {
        double bx0, bx1, by0, by1;
        warp_box<M>(p, ext, bx0, bx1, by0, by1);
        const Window W = make_window(xc + bx0, xc + bx1, yc + by0, yc + by1, A);
        if (W.lds) stage_window<NT>(win, A.gC, A.wp, W, tid);
        __syncthreads();
        float q[12];
#pragma unroll
        for (int k = 0; k < M; ++k) q[k] = (float)p[k];
        double acc[3 + M];
#pragma unroll
        for (int j = 0; j < 3 + M; ++j) acc[j] = 0.0;
#pragma unroll 1
        for (int k = 0; k < P; ++k) {
            const int i = tid + NT * k;
            if (i < n) {
                const int32_t o = offs[i];
                const float dxf = (float)unpack_dx(o), dyf = (float)unpack_dy(o);
                float ox, oy;
                warp_local32<M>(q, fxc, fyc, dxf, dyf, ox, oy);
                int cx, cy;
                float tx, ty;
                cell_frac(bx, ox, A.w, cx, tx);
                cell_frac(by, oy, A.h, cy, ty);
                float d0, d1;
                const float g = eval32<false>(win, A.gC, A.wp, W, A.border, cx, cy, tx, ty, d0, d1);
                const double gp = (double)(g - f_m32);
                const double a = (double)am[k];
                acc[0] += gp;
                acc[1] = fma(gp, gp, acc[1]);
                acc[2] = fma(a, gp, acc[2]);
                double sd[M];
                sdi64<M>((double)gxr[k], (double)gyr[k], (double)dxf, (double)dyf, sd);
#pragma unroll
                for (int j = 0; j < M; ++j) acc[3 + j] = fma(sd[j], gp, acc[3 + j]);
            }
        }
        block_sum<WPS, 3 + M>(acc, sRed, lane, wave);
        // g' above is g - (float)f_m; fold the rounding of the shift back in
        const double shift = (double)f_m32 - f_m;
        const double s1 = acc[0] + dn * shift;
        const double s2 = acc[1] + shift * (2.0 * acc[0] + dn * shift);
        const double s3 = acc[2] + shift * sa;
        double Bj[M];
#pragma unroll
        for (int j = 0; j < M; ++j) Bj[j] = acc[3 + j] + shift * Sj[j];
        const double mg = s1 / dn;
        g_m = uni(f_m + mg);
        delta_g = uni(sqrt(fmax(s2 - s1 * mg, 0.0)));
        const double ratio = delta_f / delta_g;
        double gz[M];
#pragma unroll
        for (int j = 0; j < M; ++j) gz[j] = Aj[j] - ratio * (Bj[j] - mg * Sj[j]);
        double dp[M];
        cholesky_solve_neg<M>(sL, gz, dp);
        double pn[M];
        bool ok;
        if constexpr (M == 12) {
            __syncthreads();
            if (tid == 0) {
                const bool c = compose_icgn_dev<M>(p, dp, pn, sScr);
                sCtl[0] = c ? 0 : DIC_ERR_SINGULAR_WARP;
                for (int k = 0; k < M; ++k) sP[k] = pn[k];
            }
            __syncthreads();
            ok = sCtl[0] == 0;
#pragma unroll
            for (int k = 0; k < M; ++k) pn[k] = sP[k];
        } else {
            ok = compose_icgn_dev<M>(p, dp, pn, sScr);
        }
        // every thread reads the window before it is restaged
        __syncthreads();
        if (!ok) {
            status = DIC_ERR_SINGULAR_WARP;
            break;
        }
#pragma unroll
        for (int k = 0; k < M; ++k) {
            p_last[k] = p[k];
            p[k] = uni(pn[k]);
        }
        last_norm = uni(convergence_norm_dev<M>(dp, size));
        it_done = it;
        if (A.hist && it <= A.hist_cap && tid == 0) {
            const double znssd = 2.0 - 2.0 * (s3 - mg * sa) / (delta_f * delta_g);
            double* hh = A.hist + ((int64_t)s * A.hist_cap + (it - 1)) * 4;
            hh[0] = (double)it;
            hh[1] = last_norm;
            hh[2] = 1.0 - znssd / 2.0;
            hh[3] = znssd;
        }
        if (last_norm < A.max_norm) {
            converged = true;
            break;
        }
    }
    if (status) {
        if (tid == 0) {
            R->status = status;
            R->iterations = it_done;
        }
        return;
    }

    // ---- final ZNSSD (subset.rs:1190-1193) and warp-adequacy diagnostic
    // (subset.rs:748-820) on g at the last iteration's warp; with no
    // iteration g is all zero and the reference's NaN propagates alike ----
    // e_i = a_i/Δf - (g_i - g_m)/Δg replaces a_i in place: nothing after
    // this point needs a_i itself, and no extra per-pixel registers are held
    const double rf = 1.0 / delta_f;
    const double rg = it_done ? 1.0 / delta_g : qnan;
    const double gmf = it_done ? g_m : 0.0;
    double r1[4] = {0.0, 0.0, 0.0, 0.0};  // Σz², Σe, Σgx, Σgy
    {
        Window W;
        W.lds = false;
        W.x0 = W.y0 = W.ww = W.wh = 0;
        if (it_done) {
            double bx0, bx1, by0, by1;
            warp_box<M>(p_last, ext, bx0, bx1, by0, by1);
            W = make_window(xc + bx0, xc + bx1, yc + by0, yc + by1, A);
            if (W.lds) stage_window<NT>(win, A.gC, A.wp, W, tid);
        }
        __syncthreads();
        float q[12];
#pragma unroll
        for (int k = 0; k < M; ++k) q[k] = (float)p_last[k];
#pragma unroll 1
        for (int k = 0; k < P; ++k) {
            const int i = tid + NT * k;
            if (i < n) {
                float g = 0.f;
                if (it_done) {
                    const int32_t o = offs[i];
                    float ox, oy;
                    warp_local32<M>(q, fxc, fyc, (float)unpack_dx(o), (float)unpack_dy(o), ox, oy);
                    int cx, cy;
                    float tx, ty;
                    cell_frac(bx, ox, A.w, cx, tx);
                    cell_frac(by, oy, A.h, cy, ty);
                    float d0, d1;
                    g = eval32<false>(win, A.gC, A.wp, W, A.border, cx, cy, tx, ty, d0, d1);
                }
                const double z = (double)am[k] * rf - ((double)g - gmf) * rg;
                r1[0] = fma(z, z, r1[0]);
                r1[1] += z;
                r1[2] += (double)gxr[k];
                r1[3] += (double)gyr[k];
                am[k] = (float)z;
            }
        }
    }
    block_sum<WPS, 4>(r1, sRed, lane, wave);
    const double znssd = it_done ? r1[0] : 4.0;
    const double zncc = it_done ? 1.0 - znssd / 2.0 : 0.0;
    const double e_bar = r1[1] / dn;
    const double mx = r1[2] / dn, my = r1[3] / dn;
    double r2[2] = {0.0, 0.0};
#pragma unroll 1
    for (int k = 0; k < P; ++k) {
        const int i = tid + NT * k;
        if (i < n) {
            const double a = (double)gxr[k] - mx, b = (double)gyr[k] - my;
            r2[0] = fma(a, a, r2[0]);
            r2[1] = fma(b, b, r2[1]);
        }
    }
    block_sum<WPS, 2>(r2, sRed, lane, wave);
    const double sdx = sqrt(fmax(r2[0] / dn, 1e-18)), sdy = sqrt(fmax(r2[1] / dn, 1e-18));
    double rs[2 * K], rv[2 * K], rp[2 * K];
#pragma unroll
    for (int q = 0; q < 2 * K; ++q) rs[q] = rv[q] = rp[q] = 0.0;
    for (int pass = 0; pass < 3; ++pass) {
#pragma unroll 1
        for (int k = 0; k < P; ++k) {
            const int i = tid + NT * k;
            if (i < n) {
                const int32_t o = offs[i];
                const double xi = unpack_dx(o), ze = unpack_dy(o);
                double mo[4];
                if (M == 2) {
                    mo[0] = xi;
                    mo[1] = ze;
                } else if (M == 6) {
                    mo[0] = xi * xi;
                    mo[1] = xi * ze;
                    mo[2] = ze * ze;
                } else {
                    mo[0] = xi * xi * xi;
                    mo[1] = xi * xi * ze;
                    mo[2] = xi * ze * ze;
                    mo[3] = ze * ze * ze;
                }
                const double zx = ((double)gxr[k] - mx) / sdx, zy = ((double)gyr[k] - my) / sdy;
                const double ec = (double)am[k] - e_bar;
#pragma unroll
                for (int q = 0; q < K; ++q) {
                    const double ra = zx * mo[q], rb = zy * mo[q];
                    if (pass == 0) {
                        rs[q] += ra;
                        rs[K + q] += rb;
                    } else if (pass == 1) {
                        const double a = ra - rs[q], b = rb - rs[K + q];
                        rv[q] = fma(a, a, rv[q]);
                        rv[K + q] = fma(b, b, rv[K + q]);
                    } else {
                        rp[q] = fma((ra - rs[q]) / rv[q], ec, rp[q]);
                        rp[K + q] = fma((rb - rs[K + q]) / rv[K + q], ec, rp[K + q]);
                    }
                }
            }
        }
        if (pass == 0) {
            block_sum<WPS, 2 * K>(rs, sRed, lane, wave);
#pragma unroll
            for (int q = 0; q < 2 * K; ++q) rs[q] /= dn;
        } else if (pass == 1) {
            block_sum<WPS, 2 * K>(rv, sRed, lane, wave);
#pragma unroll
            for (int q = 0; q < 2 * K; ++q) rv[q] = sqrt(fmax(rv[q] / dn, 1e-18));
        }
    }
    block_sum<WPS, 2 * K>(rp, sRed, lane, wave);
    double eu = 0.0, ev = 0.0;
#pragma unroll
    for (int q = 0; q < K; ++q) {
        const double a = rp[q] / dn, b = rp[K + q] / dn;
        eu += a * a;
        ev += b * b;
    }
    if (tid == 0) {
        for (int k = 0; k < 12; ++k) {
            R->p[k] = (k < M) ? p[k] : 0.0;
            A.p_out[12 * s + k] = (k < M) ? p[k] : 0.0;
        }
        R->c_zncc = zncc;
        R->c_znssd = znssd;
        R->norm = it_done ? last_norm : qnan;
        R->iterations = it_done;
        R->converged = converged ? 1 : 0;
        R->quality_ok = (converged && zncc >= A.tol) ? 1 : 0;
        R->eta_u = sqrt(eu);
        R->eta_v = sqrt(ev);
    }
}

}  // namespace batch
}  // namespace geodic
