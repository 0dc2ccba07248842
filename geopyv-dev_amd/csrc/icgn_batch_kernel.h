// Batched engine: the same Subset::new + ICGN solve (subset.rs:836-919,
// 1117-1219) restructured for CDNA4.
//
//  * one wavefront-sized workgroup per subset, dispatched largest first,
//    so uneven subsets balance across the 256 CUs;
//  * a subset is one wavefront; its pixels stay in VGPRs for the whole
//    solve (lane l owns pixels l, l+64, ...): reference intensity minus its
//    mean, reference gradient, and the current target intensity;
//  * every iteration stages the coefficient window the warped subset
//    touches (bounded analytically from the warp and the template extents)
//    from the FP32 copy of C into LDS with row-coalesced loads, and every
//    pixel gathers its 6x6 block from LDS; a window that does not fit
//    falls back to gathering from HBM for that iteration;
//  * positions are formed in FP64 (centre + warped offset) and only the
//    in-cell fraction goes to FP32, so the FP32 basis never sees an
//    absolute coordinate; interpolation and residuals are FP32;
//  * every sum (means, ZNSSD terms, Hessian, ∇ZNSSD) accumulates in FP64:
//    on CDNA4 a v_fma_f64 issues at the same rate as an unpacked FP32 FMA,
//    so FP64 accumulation costs nothing over a compensated FP32 two-sum and
//    is at least as accurate; cross-lane trees are FP64 too;
//  * the Hessian is factorised once (FP64 Cholesky in LDS) and each update
//    Δp = -(LLᵀ)⁻¹∇ZNSSD and the composition run in FP64.
#pragma once

#include "bspline.h"
#include "linalg.h"
#include "../../include/geodic.h"

#ifndef GEODIC_BATCH_WAVES
#define GEODIC_BATCH_WAVES 3  // resident waves per SIMD the register budget must allow
#endif
#ifndef GEODIC_BATCH_ATTRS
#define GEODIC_BATCH_ATTRS __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(GEODIC_BATCH_WAVES, 8)))
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
    float4* slab;            // per-pixel f, grad f (x, y), latest g
    const int64_t* slab_off;
    const double* p0;
    double* p_out;
    dic_subset_result* res;
    double* hist;
    int hist_cap;
    int max_it;
    double tol;
    double max_norm;
    int win_cap;
};

template <int M>
struct Modes {
    static constexpr int K = (M == 2) ? 2 : (M == 6 ? 3 : 4);
};

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

// warped local offset (relative to the subset centre)
template <int M>
__device__ __forceinline__ void warp_local(const double* p, double dx, double dy, double& ox, double& oy) {
    if (M == 2) {
        ox = dx + p[0];
        oy = dy + p[1];
        return;
    }
    ox = dx + p[0] + p[2] * dx + p[4] * dy;
    oy = dy + p[1] + p[3] * dx + p[5] * dy;
    if (M == 12) {
        const double hx = 0.5 * dx * dx, hy = 0.5 * dy * dy, xy = dx * dy;
        ox += p[6] * hx + p[8] * xy + p[10] * hy;
        oy += p[7] * hx + p[9] * xy + p[11] * hy;
    }
}

// Conservative box of the warped template (local offsets) for extents e.
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

// Stage C[y0:y0+wh, x0:x0+ww] into LDS.  The window is walked as one flat
// range (consecutive lanes take consecutive columns, so each row's loads
// coalesce) in groups of eight elements per lane, so eight loads are in
// flight before the first LDS store waits on its data.
__device__ __forceinline__ void stage_window(float* win, const float* __restrict__ C, int64_t wp, const Window& W,
                                             int lane) {
    const int total = W.ww * W.wh;
    const float inv = 1.0f / (float)W.ww;
    constexpr int G = 8;
    for (int base = 0; base < total; base += 64 * G) {
        float v[G];
#pragma unroll
        for (int k = 0; k < G; ++k) {
            const int q = base + lane + 64 * k;
            // exact row index for q < 2^22: the float quotient is off by far
            // less than the 0.5/ww margin to the next integer
            const int r = (int)(((float)q + 0.5f) * inv);
            const int c = q - r * W.ww;
            v[k] = (q < total) ? C[(int64_t)(W.y0 + r) * wp + W.x0 + c] : 0.f;
        }
#pragma unroll
        for (int k = 0; k < G; ++k) {
            const int q = base + lane + 64 * k;
            if (q < total) win[q] = v[k];
        }
    }
}

// Window of absolute positions [ax0, ax1] x [ay0, ay1] after the clamp.
__device__ __forceinline__ Window make_window(double ax0, double ax1, double ay0, double ay1, const BatchArgs& A) {
    const double e = 1e-6;
    ax0 = clamp_ref(ax0 - e, 0.0, A.wmax);
    ax1 = clamp_ref(ax1 + e, 0.0, A.wmax);
    ay0 = clamp_ref(ay0 - e, 0.0, A.hmax);
    ay1 = clamp_ref(ay1 + e, 0.0, A.hmax);
    Window W;
    // a NaN bound (diverged warp) gives an empty window: every gather then
    // goes to HBM through the NaN-safe index path of eval32
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

// FP32 value (and optionally gradient) at absolute (x, y).
template <bool GRAD>
__device__ __forceinline__ float eval32(const float* win, const float* __restrict__ C, int64_t wp, const Window& W,
                                        const BatchArgs& A, double x, double y, float& gx, float& gy) {
    x = clamp_ref(x, 0.0, A.wmax);
    y = clamp_ref(y, 0.0, A.hmax);
    // after the clamp only NaN fails x >= 0; it indexes cell 0 with a NaN
    // fraction, as the reference's saturating `as usize` cast does
    const double fx = (x >= 0.0) ? floor(x) : 0.0, fy = (y >= 0.0) ? floor(y) : 0.0;
    const float tx = (float)(x - fx), ty = (float)(y - fy);
    const int cx = (int)fx + A.border - 2, cy = (int)fy + A.border - 2;
    float a[6], b[6];
    basis_f32(ty, a);
    basis_f32(tx, b);
    const int lx = cx - W.x0, ly = cy - W.y0;
    const bool in = W.lds && lx >= 0 && ly >= 0 && lx + 6 <= W.ww && ly + 6 <= W.wh;
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
        const float* __restrict__ base = C + (int64_t)cy * wp + cx;
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

// Reference pass over the slab: Hessian rows [A0, A1) (bounded register use).
template <int M, int A0, int A1>
__device__ __forceinline__ void hessian_rows(const float4* __restrict__ slab, const int32_t* __restrict__ offs, int n,
                                             int lane, double* sH) {
    constexpr int NE = (A1 - A0) * M - ((A1 - 1) * A1 - (A0 - 1) * A0) / 2;
    double acc[NE > 0 ? NE : 1];
#pragma unroll
    for (int q = 0; q < NE; ++q) acc[q] = 0.0;
    #pragma unroll 4
            for (int i = lane; i < n; i += 64) {
        const int32_t o = offs[i];
        const float4 v = slab[i];
        double sd[M];
        sdi64<M>((double)v.y, (double)v.z, (double)unpack_dx(o), (double)unpack_dy(o), sd);
        int q = 0;
#pragma unroll
        for (int a = A0; a < A1; ++a)
#pragma unroll
            for (int b = a; b < M; ++b) {
                acc[q] = fma(sd[a], sd[b], acc[q]);
                ++q;
            }
    }
    int q = 0;
#pragma unroll
    for (int a = A0; a < A1; ++a)
#pragma unroll
        for (int b = a; b < M; ++b) {
            const double v = wave_sum_u(acc[q]);
            ++q;
            if (lane == 0) {
                sH[a * M + b] = v;
                sH[b * M + a] = v;
            }
        }
}

// One wavefront per subset; the per-pixel state (reference intensity,
// reference gradient, latest target intensity) streams from an FP32 slab
// in HBM (16 B per pixel, read once per iteration, cache-resident while the
// wave works on it), so the pixel loop stays rolled and register use stays
// low enough for several resident waves per SIMD.
//
// Each iteration is ONE pass over the pixels.  With g' = g - f_m (shift by
// the reference mean, which keeps the sums well conditioned) the pass
// accumulates, in FP64,
//     s1 = Σ g',  s2 = Σ g'²,  s3 = Σ a·g',  B_j = Σ sdi_j·g'      (a = f - f_m)
// and then, exactly as the reference's per-pixel forms (subset.rs:1162-1176,
// 418-446) but summed in a different order,
//     g_m = f_m + s1/n,  Δg² = s2 - s1²/n,
//     ∇ZNSSD_j = Σ sdi_j (a - (Δf/Δg)(g - g_m)) = A_j - (Δf/Δg)(B_j - (s1/n) S_j),
//     ZNSSD = 2 - 2 (s3 - (s1/n) Σa) / (Δf Δg),
// where A_j = Σ sdi_j a and S_j = Σ sdi_j are loop invariants from setup.
template <int M>
__global__ void GEODIC_BATCH_ATTRS icgn_batch_kernel(BatchArgs A) {
    constexpr int NL = M * (M + 1) / 2;
    constexpr int K = Modes<M>::K;
    GEODIC_DYN_LDS(float, win);
    __shared__ double sH[M * M];
    __shared__ double sL[NL];
    __shared__ double sScr[144];
    __shared__ double sP[12];
    __shared__ int sCtl[2];
    const int lane = threadIdx.x;
    const double qnan = __longlong_as_double(0x7ff8000000000000LL);

    {
        // One workgroup per subset, dispatched largest first (perm).  A
        // lane-0 atomic work counter broadcast with a shuffle is avoided on
        // purpose: the compiler restructures that divergent loop so that its
        // back-edge skips the atomic.
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
        const double dn = (double)n;
        float4* __restrict__ slab = A.slab + A.slab_off[s];
        dic_subset_result* R = A.res + s;
        if (lane == 0) {
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
        if (n == 0) {
            if (lane == 0) R->status = DIC_ERR_INVALID_INPUT;
            return;
        }

        // ---- Subset::new on the staged reference window ----
        double f_m, delta_f, sa;
        double Aj[M], Sj[M];
        {
            const Window W = make_window(xc + ext[0], xc + ext[1], yc + ext[2], yc + ext[3], A);
            if (W.lds) stage_window(win, A.fC, A.wp, W, lane);
            __syncthreads();
            double sf = 0.0, ss = 0.0;
            #pragma unroll 4
            for (int i = lane; i < n; i += 64) {
                const int32_t o = offs[i];
                float gx, gy;
                const float v = eval32<true>(win, A.fC, A.wp, W, A, xc + unpack_dx(o), yc + unpack_dy(o), gx, gy);
                slab[i] = make_float4(v, gx, gy, 0.f);
                sf += (double)v;
                ss += 0.5 * ((double)gx * gx + (double)gy * gy);
            }
            __syncthreads();
            f_m = wave_sum_u(sf) / dn;
            ss = wave_sum_u(ss);
            double sq = 0.0, sa_l = 0.0;
#pragma unroll
            for (int j = 0; j < M; ++j) Aj[j] = Sj[j] = 0.0;
            #pragma unroll 4
            for (int i = lane; i < n; i += 64) {
                const int32_t o = offs[i];
                const float4 v = slab[i];
                const double a = (double)v.x - f_m;
                sq = fma(a, a, sq);
                sa_l += a;
                double sd[M];
                sdi64<M>((double)v.y, (double)v.z, (double)unpack_dx(o), (double)unpack_dy(o), sd);
#pragma unroll
                for (int j = 0; j < M; ++j) {
                    Aj[j] = fma(sd[j], a, Aj[j]);
                    Sj[j] += sd[j];
                }
            }
            sq = wave_sum_u(sq);
            sa = wave_sum_u(sa_l);
#pragma unroll
            for (int j = 0; j < M; ++j) {
                Aj[j] = wave_sum_u(Aj[j]);
                Sj[j] = wave_sum_u(Sj[j]);
            }
            delta_f = sqrt(sq);
            if (lane == 0) {
                R->f_m = f_m;
                R->delta_f = delta_f;
                R->sssig = ss;
                R->sigma = sqrt(sq / dn);
            }
        }
        if (delta_f == 0.0) {
            if (lane == 0) R->status = DIC_ERR_INVALID_INPUT;
            return;
        }
        if constexpr (M == 12) {
            hessian_rows<M, 0, 3>(slab, offs, n, lane, sH);
            hessian_rows<M, 3, 6>(slab, offs, n, lane, sH);
            hessian_rows<M, 6, 12>(slab, offs, n, lane, sH);
        } else {
            hessian_rows<M, 0, M>(slab, offs, n, lane, sH);
        }
        __syncthreads();
        if (lane == 0) sCtl[0] = cholesky<M>(sH, sL) ? 0 : DIC_ERR_SINGULAR_HESSIAN;
        __syncthreads();
        if (sCtl[0]) {
            if (lane == 0) R->status = DIC_ERR_SINGULAR_HESSIAN;
            return;
        }
        const double size = sqrt(dn);
        double p[M];
#pragma unroll
        for (int k = 0; k < M; ++k) p[k] = uni(A.p0[12 * s + k]);

        int it_done = 0;
        bool converged = false;
        int status = 0;
        double g_m = 0.0, delta_g = 0.0, last_norm = 0.0;
        for (int it = 1; it <= A.max_it; ++it) {
            double bx0, bx1, by0, by1;
            warp_box<M>(p, ext, bx0, bx1, by0, by1);
            const Window W = make_window(xc + bx0, xc + bx1, yc + by0, yc + by1, A);
            if (W.lds) stage_window(win, A.gC, A.wp, W, lane);
            __syncthreads();
            double s1 = 0.0, s2 = 0.0, s3 = 0.0;
            double Bj[M];
#pragma unroll
            for (int j = 0; j < M; ++j) Bj[j] = 0.0;
            #pragma unroll 4
            for (int i = lane; i < n; i += 64) {
                const int32_t o = offs[i];
                const double dx = (double)unpack_dx(o), dy = (double)unpack_dy(o);
                const float4 v = slab[i];
                double ox, oy;
                warp_local<M>(p, dx, dy, ox, oy);
                float d0, d1;
                const float g = eval32<false>(win, A.gC, A.wp, W, A, xc + ox, yc + oy, d0, d1);
                slab[i].w = g;
                const double gp = (double)g - f_m;
                const double a = (double)v.x - f_m;
                s1 += gp;
                s2 = fma(gp, gp, s2);
                s3 = fma(a, gp, s3);
                double sd[M];
                sdi64<M>((double)v.y, (double)v.z, dx, dy, sd);
#pragma unroll
                for (int j = 0; j < M; ++j) Bj[j] = fma(sd[j], gp, Bj[j]);
            }
            __syncthreads();
            s1 = wave_sum_u(s1);
            s2 = wave_sum_u(s2);
            s3 = wave_sum_u(s3);
#pragma unroll
            for (int j = 0; j < M; ++j) Bj[j] = wave_sum_u(Bj[j]);
            const double mg = s1 / dn;
            g_m = f_m + mg;
            delta_g = sqrt(fmax(s2 - s1 * mg, 0.0));
            const double ratio = delta_f / delta_g;
            double gz[M];
#pragma unroll
            for (int j = 0; j < M; ++j) gz[j] = Aj[j] - ratio * (Bj[j] - mg * Sj[j]);
            double dp[M];
            cholesky_solve_neg<M>(sL, gz, dp);
            double pn[M];
            bool ok;
            if constexpr (M == 12) {
                if (lane == 0) {
                    const bool c = compose_icgn_dev<M>(p, dp, pn, sScr);
                    sCtl[0] = c ? 0 : DIC_ERR_SINGULAR_WARP;
                    for (int k = 0; k < M; ++k) sP[k] = pn[k];
                }
                __syncthreads();
                ok = sCtl[0] == 0;
#pragma unroll
                for (int k = 0; k < M; ++k) pn[k] = sP[k];
                __syncthreads();
            } else {
                ok = compose_icgn_dev<M>(p, dp, pn, sScr);
            }
            if (!ok) {
                status = DIC_ERR_SINGULAR_WARP;
                break;
            }
#pragma unroll
            for (int k = 0; k < M; ++k) p[k] = uni(pn[k]);
            last_norm = uni(convergence_norm_dev<M>(dp, size));
            it_done = it;
            if (A.hist && it <= A.hist_cap && lane == 0) {
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
            if (lane == 0) {
                R->status = status;
                R->iterations = it_done;
            }
            return;
        }
        // ZNSSD of the last g (subset.rs:1190-1193) and the warp-adequacy
        // diagnostic (subset.rs:748-820); with no iteration g is all zero and
        // the reference's NaN propagates the same way here.
        const double rf = 1.0 / delta_f;
        const double rg = it_done ? 1.0 / delta_g : qnan;
        const double gmf = it_done ? g_m : 0.0;
        double zz = 0.0, se = 0.0, sx = 0.0, sy = 0.0;
        #pragma unroll 4
            for (int i = lane; i < n; i += 64) {
            const float4 v = slab[i];
            const double z = ((double)v.x - f_m) * rf - ((double)v.w - gmf) * rg;
            zz = fma(z, z, zz);
            se += z;
            sx += (double)v.y;
            sy += (double)v.z;
        }
        const double znssd = it_done ? wave_sum_u(zz) : 4.0;
        const double zncc = it_done ? 1.0 - znssd / 2.0 : 0.0;
        const double e_bar = wave_sum_u(se) / dn;
        const double mx = wave_sum_u(sx) / dn, my = wave_sum_u(sy) / dn;
        double vx = 0.0, vy = 0.0;
        #pragma unroll 4
            for (int i = lane; i < n; i += 64) {
            const float4 v = slab[i];
            const double a = (double)v.y - mx, b = (double)v.z - my;
            vx = fma(a, a, vx);
            vy = fma(b, b, vy);
        }
        const double sdx = sqrt(fmax(wave_sum_u(vx) / dn, 1e-18)), sdy = sqrt(fmax(wave_sum_u(vy) / dn, 1e-18));
        double rs[2 * K], rv[2 * K], rp[2 * K];
#pragma unroll
        for (int q = 0; q < 2 * K; ++q) rs[q] = rv[q] = rp[q] = 0.0;
        for (int pass = 0; pass < 3; ++pass) {
            #pragma unroll 4
            for (int i = lane; i < n; i += 64) {
                const int32_t o = offs[i];
                const float4 v = slab[i];
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
                const double zx = ((double)v.y - mx) / sdx, zy = ((double)v.z - my) / sdy;
                const double ec = (((double)v.x - f_m) * rf - ((double)v.w - gmf) * rg) - e_bar;
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
            if (pass == 0) {
#pragma unroll
                for (int q = 0; q < 2 * K; ++q) rs[q] = wave_sum_u(rs[q]) / dn;
            } else if (pass == 1) {
#pragma unroll
                for (int q = 0; q < 2 * K; ++q) rv[q] = sqrt(fmax(wave_sum_u(rv[q]) / dn, 1e-18));
            }
        }
        double eu = 0.0, ev = 0.0;
#pragma unroll
        for (int q = 0; q < K; ++q) {
            const double a = wave_sum_u(rp[q]) / dn, b = wave_sum_u(rp[K + q]) / dn;
            eu += a * a;
            ev += b * b;
        }
        if (lane == 0) {
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
}

}  // namespace batch
}  // namespace geodic
