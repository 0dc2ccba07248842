// Batched engine: Subset::new + the ICGN solve (subset.rs:836-919,
// 1117-1219) for many subsets, restructured for CDNA4.
//
//  * one workgroup of WPS wavefronts per subset, dispatched largest first
//    so uneven subsets balance across the CUs; thread t owns pixels
//    t, t + 64·WPS, ... and keeps their reference state in VGPRs for the
//    whole solve: a = f - f_m and ∇f in FP32 (P pixels per thread, walked
//    by a rolled loop whose uniform index the compiler turns into register
//    indexing); the template offsets sit in LDS as FP32 pairs;
//  * the coefficient window the warped subset touches (bounded analytically
//    from the warp and the template extents, plus a few pixels of slack) is
//    staged from the FP32 copy of C into LDS with 16-byte loads, and kept
//    until an iteration's warp leaves it; every pixel gathers its 6x6 block
//    from LDS.  When the whole warped subset lies inside the image and the
//    window, the pixel loop runs without clamps or bounds tests (the fast
//    path); otherwise every position is clamped as image.rs:130-132 does and
//    blocks outside the window are read from HBM;
//  * positions are an integer base (the floor of the subset centre) plus an
//    FP32 offset, so the FP32 basis never sees an absolute coordinate; at
//    setup every pixel shares the centre's fraction, so the basis is
//    computed once per subset;
//  * each iteration is ONE pass over the pixels.  With g' = g - f_m (a
//    shift by the reference mean keeps the sums well conditioned) the pass
//    accumulates
//        s1 = Σ g',  s2 = Σ g'²,  s3 = Σ a·g',  B_j = Σ sdi_j·g'
//    and the reference's per-pixel forms (subset.rs:1162-1176, 418-446)
//    follow exactly, summed in another order:
//        g_m = f_m + s1/n,   Δg² = s2 - s1²/n,
//        ∇ZNSSD_j = Σ sdi_j (a - (Δf/Δg)(g - g_m)) = A_j - (Δf/Δg)(B_j - (s1/n) S_j),
//        ZNSSD = 2 - 2 (s3 - (s1/n) Σa) / (Δf Δg),
//    with A_j = Σ sdi_j a and S_j = Σ sdi_j fixed at setup;
//  * per-pixel arithmetic and each thread's partial sums (at most P terms)
//    are FP32; the partials are widened to FP64 for the wavefront DPP trees
//    and the cross-wavefront sums (LDS, fixed wave order), so a sum's error
//    is that of a dozen FP32 additions;
//  * the Hessian is factorised and inverted once (FP64 Cholesky, one column
//    of H⁻¹ per lane); for orders 0 and 1 the constant parts H⁻¹A and H⁻¹S
//    of every update are formed then too, so each update
//    Δp = -H⁻¹∇ZNSSD = -(H⁻¹A - (Δf/Δg)(H⁻¹B - (s1/n) H⁻¹S)) is one
//    dot product and three operations per lane; it and the composition run
//    in FP64;
//  * the warp-adequacy diagnostic (subset.rs:748-820) needs only reference
//    data except for the final residual e: its gradient standardisation and
//    mode statistics are computed at setup, and the final pass that
//    re-evaluates g at the last iteration's warp accumulates Σe², Σe and the
//    projections Σ sc_q·e together (Σ sc_q·(e - ē) = Σ sc_q·e - ē Σ sc_q).
#pragma once

#include <cstddef>

#include "bspline.h"
#include "linalg.h"
#include "../../include/geodic.h"

#ifndef GEODIC_BATCH_ATTRS
#define GEODIC_BATCH_ATTRS(WPS, OCC) __launch_bounds__(64 * (WPS)) __attribute__((amdgpu_waves_per_eu(OCC, 8)))
#endif
// Resident wavefronts per SIMD the register budget is sized for (orders 0
// and 1; the 12-parameter warp takes what it needs).
#ifndef GEODIC_BATCH_OCC
#define GEODIC_BATCH_OCC 3
#endif
#ifndef GEODIC_DYN_LDS
#define GEODIC_DYN_LDS(T, name) extern __shared__ T name[]
#endif

namespace geodic {
namespace batch {

// What a workgroup needs to start its subset, in one record so that it
// arrives in one round trip: the centre, the template (its pixels
// tpl_xy[st, st + n) and extents dxmin, dxmax, dymin, dymax).
struct SubMeta {
    double xc, yc;
    int64_t st;
    int32_t n, t;
    int32_t ext[4];
};
static_assert(sizeof(SubMeta) == 48, "SubMeta layout");

struct BatchArgs {
    const float* fC;
    const float* gC;
    int64_t pitch;  // row pitch of the FP32 coefficient arrays, a multiple of 4
    int border;
    int w, h;
    double wmax, hmax;
    int64_t n_sub;
    const SubMeta* meta;     // per subset
    const int32_t* tpl_xy;   // offset pairs in the engine's pixel order, as OffXY (two offh_t)
    const int32_t* perm;
    const double* p0;
    double* p_out;
    dic_subset_result* res;
    double* hist;
    int hist_cap;
    int max_it;
    double tol;
    double max_norm;
    // max_norm² and (6 max_norm)², set by the host: kernel arguments stay
    // in scalar registers, where products formed in the kernel would take
    // vector registers for the whole solve
    double max_norm2, diag_norm2;
    int win_cap;   // floats of LDS for the coefficient window (a multiple of 4)
    int win_slack; // pixels of slack around a restaged window
    // list solves: the initial warps (12 each) and result records in
    // dispatch order (device-mapped host memory), else null
    const double* p0_list;
    dic_subset_result* res_list;
    // list solves: a flag per dispatch position, raised (after its record)
    // by the workgroup that solved it, else null
    int32_t* done_list;
    // f, ∂x f and ∂y f of the reference at every pixel (row pitch w), the
    // knot form below evaluated once per image, else null
    const float4* knot;
};

constexpr int kRedSlots = 42;  // doubles per wavefront in the cross-wave reduction buffer

template <int M>
struct Modes {
    static constexpr int K = (M == 2) ? 2 : (M == 6 ? 3 : 4);
};

// warped local offset (relative to the subset centre), FP64, for bounds
template <int M>
__device__ __forceinline__ void warp_box(const double* p, const int32_t e[4], double& x0, double& x1, double& y0,
                                         double& y1) {
    const double cx[2] = {(double)e[0], (double)e[1]}, cy[2] = {(double)e[2], (double)e[3]};
    if constexpr (M == 6) {
        // the affine image of the box is bounded axis by axis: the extreme
        // x over the corners is p0 + the extremes of (1 + p2) cx and of
        // p4 cy (a shorter dependent chain than the four corners)
        const double ax0 = (1.0 + p[2]) * cx[0], ax1 = (1.0 + p[2]) * cx[1];
        const double bx0 = p[4] * cy[0], bx1 = p[4] * cy[1];
        const double ay0 = p[3] * cx[0], ay1 = p[3] * cx[1];
        const double by0 = (1.0 + p[5]) * cy[0], by1 = (1.0 + p[5]) * cy[1];
        x0 = p[0] + (fmin(ax0, ax1) + fmin(bx0, bx1));
        x1 = p[0] + (fmax(ax0, ax1) + fmax(bx0, bx1));
        y0 = p[1] + (fmin(ay0, ay1) + fmin(by0, by1));
        y1 = p[1] + (fmax(ay0, ay1) + fmax(by0, by1));
        return;
    }
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

// Staged coefficient window: rows [y0, y0 + wh), columns [x0, x0 + ww) of
// C (x0 and ww multiples of 4), row stride ls floats in LDS.  Pixels whose
// cell lies in [cx_lo, cx_hi] x [cy_lo, cy_hi] find their whole 6x6 block
// inside.  lds == false: nothing staged, every gather reads HBM.
struct Window {
    int x0, y0, ww, wh, ls;
    int cx_lo, cx_hi, cy_lo, cy_hi;
    bool lds;
};

__device__ __forceinline__ Window no_window() {
    Window W;
    W.x0 = W.y0 = W.ww = W.wh = W.ls = 0;
    W.cx_lo = W.cy_lo = 1;
    W.cx_hi = W.cy_hi = 0;
    W.lds = false;
    return W;
}

// Cells touched by the absolute box [ax0, ax1] x [ay0, ay1] after the clamp
// (image.rs:130-131); false for a NaN box.
__device__ __forceinline__ bool box_cells(double ax0, double ax1, double ay0, double ay1, const BatchArgs& A,
                                          int c[4]) {
    if (ax0 >= 0.0 && ax1 <= A.wmax && ay0 >= 0.0 && ay1 <= A.hmax && ax0 <= ax1 && ay0 <= ay1) {
        // inside the image: the clamp is the identity
        c[0] = uni((int)floor(ax0));
        c[1] = uni((int)floor(ax1));
        c[2] = uni((int)floor(ay0));
        c[3] = uni((int)floor(ay1));
        return true;
    }
    ax0 = clamp_ref(ax0, 0.0, A.wmax);
    ax1 = clamp_ref(ax1, 0.0, A.wmax);
    ay0 = clamp_ref(ay0, 0.0, A.hmax);
    ay1 = clamp_ref(ay1, 0.0, A.hmax);
    if (!(ax0 <= ax1) || !(ay0 <= ay1)) return false;
    c[0] = uni((int)floor(ax0));
    c[1] = uni((int)floor(ax1));
    c[2] = uni((int)floor(ay0));
    c[3] = uni((int)floor(ay1));
    return true;
}

// Window holding the blocks of cells [c0, c1] x [c2, c3] (all in the image).
// Columns round out to 16-byte boundaries; the FP32 rows are pitched to
// 16 bytes, so x0 + ww never passes the pitch.  S > 0 fixes the LDS row
// stride at compile time (gathers then address rows by immediate offsets);
// a wider window does not fit such a layout.
template <int S>
__device__ __forceinline__ Window window_for(const int c[4], const BatchArgs& A) {
    Window W;
    W.x0 = (c[0] + A.border - 2) & ~3;
    const int x1 = c[1] + A.border + 3;
    W.ww = (x1 - W.x0 + 4) & ~3;
    W.y0 = c[2] + A.border - 2;
    W.wh = c[3] - c[2] + 6;
    W.cx_lo = W.x0 - A.border + 2;
    W.cx_hi = W.x0 + W.ww - A.border - 4;
    W.cy_lo = c[2];
    W.cy_hi = c[3];
    W.ls = S > 0 ? S : W.ww;
    W.lds = W.ww <= W.ls && W.ls * W.wh <= A.win_cap;
    return W;
}

__device__ __forceinline__ bool covers(const Window& W, const int c[4]) {
    return W.lds && c[0] >= W.cx_lo && c[1] <= W.cx_hi && c[2] >= W.cy_lo && c[3] <= W.cy_hi;
}

// Stage the window into LDS, less `shift`: thread t copies 16-byte chunks
// t, t + NT, ... of the row-major chunk grid, four loads in flight.
template <int NT>
__device__ __forceinline__ void stage_window(float* win, const float* __restrict__ C, int64_t pitch, const Window& W,
                                             int tid, float shift) {
    const int cpr = W.ww >> 2;  // chunks per row
    const int total = cpr * W.wh;
    const float* __restrict__ src0 = C + (int64_t)W.y0 * pitch + W.x0;
    const int pi = (int)pitch;  // offsets inside a window fit in 32 bits
    int r = tid / cpr;
    int c = tid - r * cpr;
    const int dr = NT / cpr, dc = NT - (NT / cpr) * cpr;
    int k = tid;
    for (; k + 3 * NT < total; k += 4 * NT) {
        int rr[4], cc[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            rr[j] = r;
            cc[j] = c;
            r += dr;
            c += dc;
            if (c >= cpr) {
                c -= cpr;
                ++r;
            }
        }
        float4 v[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) v[j] = *(const float4*)(src0 + rr[j] * pi + 4 * cc[j]);
#pragma unroll
        for (int j = 0; j < 4; ++j)
            *(float4*)(win + rr[j] * W.ls + 4 * cc[j]) =
                make_float4(v[j].x - shift, v[j].y - shift, v[j].z - shift, v[j].w - shift);
    }
    for (; k < total; k += NT) {
        const float4 v = *(const float4*)(src0 + r * pi + 4 * c);
        *(float4*)(win + r * W.ls + 4 * c) = make_float4(v.x - shift, v.y - shift, v.z - shift, v.w - shift);
        r += dr;
        c += dc;
        if (c >= cpr) {
            c -= cpr;
            ++r;
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

// 6x6 contraction Σ_r a_r Σ_k blk[r][k] b_k, row stride `stride`.
__device__ __forceinline__ float contract6(const float* blk, int stride, const float a[6], const float b[6]) {
    float val = 0.f;
#pragma unroll
    for (int r = 0; r < 6; ++r) {
        const float* row = blk + r * stride;
        float s = 0.f;
#pragma unroll
        for (int k = 0; k < 6; ++k) s = fmaf(row[k], b[k], s);
        val = fmaf(a[r], s, val);
    }
    return val;
}

// The 6x6 block into registers, every read issued before any use (the
// scheduling barrier keeps the basis arithmetic that follows from being
// interleaved with the reads, so the 18 LDS round trips overlap).
__device__ __forceinline__ void gather6(const float* blk, int stride, float c[36]) {
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int k = 0; k < 6; ++k) c[6 * r + k] = blk[r * stride + k];
    GEODIC_SCHED_BARRIER();
}

// The same with a compile-time row stride S: rows 0-3 are immediate offsets
// from blk, rows 4-5 from blk4 = blk + 4 S (DS offsets reach 255 dwords).
// blk4 is passed in, computed from the run-time stride, so the compiler
// keeps one second base instead of re-deriving an address per read.
template <int S>
__device__ __forceinline__ void gather6s(const float* blk, const float* blk4, float c[36]) {
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int k = 0; k < 6; ++k) c[6 * r + k] = blk[r * S + k];
#pragma unroll
    for (int r = 4; r < 6; ++r)
#pragma unroll
        for (int k = 0; k < 6; ++k) c[6 * r + k] = blk4[(r - 4) * S + k];
    GEODIC_SCHED_BARRIER();
}

// Gather a block at LDS index idx with row stride S (compile time) or ls.
template <int S>
__device__ __forceinline__ void gather_block(const float* win, int idx, int ls, float c[36]) {
    if constexpr (S > 0) {
        gather6s<S>(win + idx, win + idx + 4 * ls, c);
    } else {
        gather6(win + idx, ls, c);
    }
}

// The 5x5 block at the start of a 6x6 one (rows and columns 0-4), every
// read issued before any use, with a compile-time row stride S (rows 4 on
// from blk4 = blk + 4 S) or the run-time stride ls.
template <int S>
__device__ __forceinline__ void gather_block5(const float* win, int idx, int ls, float c[25]) {
    const float* blk = win + idx;
    if constexpr (S > 0) {
        const float* blk4 = blk + 4 * ls;
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int k = 0; k < 5; ++k) c[5 * r + k] = blk[r * S + k];
#pragma unroll
        for (int k = 0; k < 5; ++k) c[20 + k] = blk4[k];
    } else {
#pragma unroll
        for (int r = 0; r < 5; ++r)
#pragma unroll
            for (int k = 0; k < 5; ++k) c[5 * r + k] = blk[r * ls + k];
    }
    GEODIC_SCHED_BARRIER();
}

// Value, ∂x and ∂y at a knot (zero fraction on both axes), from the 5x5
// block of its cell: there the value basis is (1/120, 13/60, 11/20, 13/60,
// 1/120, 0), symmetric, and the derivative basis (-1/24, -5/12, 0, 5/12,
// 1/24, 0), antisymmetric (the Q matrix's first two rows, image.rs:183-190,
// at t = 0), so each row adds or subtracts its mirror pairs first.
__device__ __forceinline__ float knot_grad(const float c[25], float& gx, float& gy) {
    constexpr float b0 = (float)GQ00, b1 = (float)GQ01, b2 = (float)GQ02;
    constexpr float d0 = (float)GQ14, d1 = (float)GQ13;
    float s[5], t[5];
#pragma unroll
    for (int r = 0; r < 5; ++r) {
        const float* q = c + 5 * r;
        s[r] = fmaf(b0, q[0] + q[4], fmaf(b1, q[1] + q[3], b2 * q[2]));
        t[r] = fmaf(d0, q[4] - q[0], d1 * (q[3] - q[1]));
    }
    gx = fmaf(b0, t[0] + t[4], fmaf(b1, t[1] + t[3], b2 * t[2]));
    gy = fmaf(d0, s[4] - s[0], d1 * (s[3] - s[1]));
    return fmaf(b0, s[0] + s[4], fmaf(b1, s[1] + s[3], b2 * s[2]));
}

__device__ __forceinline__ float contract6r(const float c[36], const float a[6], const float b[6]) {
    float val = 0.f;
#pragma unroll
    for (int r = 0; r < 6; ++r) {
        float s = c[6 * r] * b[0];
#pragma unroll
        for (int k = 1; k < 6; ++k) s = fmaf(c[6 * r + k], b[k], s);
        val = (r == 0) ? a[0] * s : fmaf(a[r], s, val);
    }
    return val;
}

__device__ __forceinline__ float contract6r_grad(const float c[36], const float a[6], const float b[6],
                                                 const float da[6], const float db[6], float& gx, float& gy) {
    float val = 0.f, ggx = 0.f, ggy = 0.f;
#pragma unroll
    for (int r = 0; r < 6; ++r) {
        float s = c[6 * r] * b[0], t = c[6 * r] * db[0];
#pragma unroll
        for (int k = 1; k < 6; ++k) {
            s = fmaf(c[6 * r + k], b[k], s);
            t = fmaf(c[6 * r + k], db[k], t);
        }
        val = fmaf(a[r], s, val);
        ggx = fmaf(a[r], t, ggx);
        ggy = fmaf(da[r], s, ggy);
    }
    gx = ggx;
    gy = ggy;
    return val;
}

// Value, ∂x and ∂y from one 6x6 block: ∂x with the derivative basis in x,
// ∂y with it in y (image.rs:155-178).
__device__ __forceinline__ float contract6_grad(const float* blk, int stride, const float a[6], const float b[6],
                                                const float da[6], const float db[6], float& gx, float& gy) {
    float val = 0.f, ggx = 0.f, ggy = 0.f;
#pragma unroll
    for (int r = 0; r < 6; ++r) {
        const float* row = blk + r * stride;
        float s = 0.f, t = 0.f;
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            s = fmaf(row[k], b[k], s);
            t = fmaf(row[k], db[k], t);
        }
        val = fmaf(a[r], s, val);
        ggx = fmaf(a[r], t, ggx);
        ggy = fmaf(da[r], s, ggy);
    }
    gx = ggx;
    gy = ggy;
    return val;
}

// General gather at cell (cx, cy) + (tx, ty), less `shift`: LDS when the
// block is inside the window (staged already less `shift`), HBM otherwise
// (two explicit paths so each compiles to its own address space).  The
// basis sums to one, so subtracting the shift from the value equals
// subtracting it from every coefficient.
template <bool GRAD>
__device__ __forceinline__ float eval32(const float* win, const float* __restrict__ C, int64_t pitch, const Window& W,
                                        int border, int cellx, int celly, float tx, float ty, float shift, float& gx,
                                        float& gy) {
    float a[6], b[6], da[6], db[6];
    basis_f32(ty, a);
    basis_f32(tx, b);
    if (GRAD) {
        dbasis_f32(ty, da);
        dbasis_f32(tx, db);
    }
    const bool in = W.lds && cellx >= W.cx_lo && cellx <= W.cx_hi && celly >= W.cy_lo && celly <= W.cy_hi;
    if (in) {
        const float* blk = win + (celly - W.cy_lo) * W.ls + (cellx - W.cx_lo);
        if (GRAD) return contract6_grad(blk, W.ls, a, b, da, db, gx, gy);
        return contract6(blk, W.ls, a, b);
    }
    const float* __restrict__ blk = C + (int64_t)(celly + border - 2) * pitch + (cellx + border - 2);
    if (GRAD) return contract6_grad(blk, (int)pitch, a, b, da, db, gx, gy) - shift;
    return contract6(blk, (int)pitch, a, b) - shift;
}

// LDS writes of some lanes made visible to the other lanes of the same
// wavefront.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Rows of kRedSlots doubles in the reduction buffer: one per wavefront for
// its partials and, with several wavefronts, one per wavefront for totals.
template <int WPS>
constexpr int red_rows() {
    return WPS == 1 ? 1 : 2 * WPS;
}

// The totals of V values from the wavefronts' partial rows of sRed: lane j
// adds value j over the wavefronts in wave order into `tot`, and every lane
// then reads the totals back as lane-invariant loads.
template <int WPS, int V>
__device__ __forceinline__ void sum_rows(const double* sRed, double* tot, int lane, double* v) {
    if constexpr (WPS == 2 && V <= 8) {
        // two rows: every lane adds them itself (the same additions), one
        // LDS round trip fewer
#pragma unroll
        for (int j = 0; j < V; ++j) v[j] = sRed[j] + sRed[kRedSlots + j];
        return;
    }
    if (lane < V) {
        double t = sRed[lane];
#pragma unroll
        for (int q = 1; q < WPS; ++q) t += sRed[q * kRedSlots + lane];
        tot[lane] = t;
    }
    wave_lds_sync();
#pragma unroll
    for (int j = 0; j < V; ++j) v[j] = tot[j];
}

// Sums V FP64 values over the workgroup: a transposed DPP/permlane tree per
// wavefront whose results the lanes holding them write to the wavefront's
// row of sRed (no lane reads into scalar registers), then value j summed
// over the wavefronts in wave order by lane j into the wavefront's totals
// row, read back by every lane.  Every wave gets the same values; the
// additions do not depend on scheduling.
template <int WPS, int V>
__device__ __forceinline__ void block_sum(double* v, double* sRed, int lane, int wave) {
    static_assert(V <= kRedSlots, "reduction buffer");
    wave_sums_to<V>(v, sRed + wave * kRedSlots, lane);
    if constexpr (WPS == 1) {
        wave_lds_sync();
#pragma unroll
        for (int j = 0; j < V; ++j) v[j] = sRed[j];
    } else {
        __syncthreads();
        sum_rows<WPS, V>(sRed, sRed + (WPS + wave) * kRedSlots, lane, v);
        __syncthreads();  // every wavefront has read the partials before any writes again
    }
}

// The same sums delivered to wavefront 0 only (the one that then runs the
// serial update); returns true there.  Same addition order as block_sum.
// The partials rows are written again only after the caller's next
// workgroup barrier.
template <int WPS, int V>
__device__ __forceinline__ bool block_sum_lead(double* v, double* sRed, int lane, int wave) {
    static_assert(V <= kRedSlots, "reduction buffer");
    wave_sums_to<V>(v, sRed + wave * kRedSlots, lane);
    if constexpr (WPS == 1) {
        wave_lds_sync();
#pragma unroll
        for (int j = 0; j < V; ++j) v[j] = sRed[j];
    } else {
        __syncthreads();
        if (wave == 0) sum_rows<WPS, V>(sRed, sRed + WPS * kRedSlots, lane, v);
    }
    return wave == 0;
}

// FP32 per-thread partials widened to FP64 and summed over the workgroup,
// delivered to wavefront 0 only (returns true there), as block_sum_lead.
template <int WPS, int V>
__device__ __forceinline__ bool block_sum_lead32(const float* f, double* v, double* sRed, int lane, int wave) {
#pragma unroll
    for (int j = 0; j < V; ++j) v[j] = (double)f[j];
    return block_sum_lead<WPS, V>(v, sRed, lane, wave);
}

// FP32 per-thread partials widened to FP64 and summed over the workgroup.
template <int WPS, int V>
__device__ __forceinline__ void block_sum32(const float* f, double* v, double* sRed, int lane, int wave) {
#pragma unroll
    for (int j = 0; j < V; ++j) v[j] = (double)f[j];
    block_sum<WPS, V>(v, sRed, lane, wave);
}

// steepest-descent row of one pixel, FP32 (subset.rs:311-333)
template <int M>
__device__ __forceinline__ void sdi32(float gx, float gy, float dx, float dy, float* s) {
    s[0] = gx;
    s[1] = gy;
    if (M >= 6) {
        s[2] = gx * dx;
        s[3] = gy * dx;
        s[4] = gx * dy;
        s[5] = gy * dy;
    }
    if (M == 12) {
        const float hx = 0.5f * dx * dx, hy = 0.5f * dy * dy, xy = dx * dy;
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
__device__ __forceinline__ void hessian_rows(const float* gxr, const float* gyr, const OffXY* sXY, int n, int tid,
                                             int lane, int wave, double* sH, double* sRed) {
    constexpr int NT = 64 * WPS;
    constexpr int NE = (A1 - A0) * M - ((A1 - 1) * A1 - (A0 - 1) * A0) / 2;
    constexpr int PU = P <= 4 ? P : 1;  // a few pixels per thread: overlap them
    // the rolled loop runs the slots the template fills (a uniform count)
    const int KP = PU == P ? P : min(P, (n + NT - 1) / NT);
    float acc[NE > 0 ? NE : 1];
#pragma unroll
    for (int q = 0; q < NE; ++q) acc[q] = 0.f;
#pragma unroll PU
    for (int k = 0; k < KP; ++k) {
        const int i = tid + NT * k;
        if (i < n) {
            const float2 d = off_xy(sXY[i]);
            float sd[M];
            sdi32<M>(gxr[k], gyr[k], d.x, d.y, sd);
            int q = 0;
#pragma unroll
            for (int a = A0; a < A1; ++a)
#pragma unroll
                for (int b = a; b < M; ++b) {
                    acc[q] = fmaf(sd[a], sd[b], acc[q]);
                    ++q;
                }
        }
    }
    double v[NE > 0 ? NE : 1];
    block_sum32<WPS, NE>(acc, v, sRed, lane, wave);
    if (tid == 0) {
        int q = 0;
#pragma unroll
        for (int a = A0; a < A1; ++a)
#pragma unroll
            for (int b = a; b < M; ++b) {
                sH[a * M + b] = v[q];
                sH[b * M + a] = v[q];
                ++q;
            }
    }
}

// Warped position relative to the integer base, FP32.  q2p = 1 + u_x and
// q5p = 1 + v_y fold the identity into the gradient terms; h[] holds the
// second-order coefficients already halved where the warp halves them.
template <int M>
__device__ __forceinline__ void warp_pos(const float* q, float c0x, float c0y, float dx, float dy, float& ox,
                                         float& oy) {
    if (M == 2) {
        ox = c0x + dx;
        oy = c0y + dy;
    } else if (M == 6) {
        ox = fmaf(q[2], dx, fmaf(q[4], dy, c0x));
        oy = fmaf(q[3], dx, fmaf(q[5], dy, c0y));
    } else {
        const float xx = dx * dx, xy = dx * dy, yy = dy * dy;
        ox = fmaf(q[2], dx, fmaf(q[4], dy, fmaf(q[6], xx, fmaf(q[8], xy, fmaf(q[10], yy, c0x)))));
        oy = fmaf(q[3], dx, fmaf(q[5], dy, fmaf(q[7], xx, fmaf(q[9], xy, fmaf(q[11], yy, c0y)))));
    }
}

// Per-iteration FP32 accumulators: s1, s2, s3, B_0..B_{M-1}; without S3
// the ZNSSD's s3 is left out and B_j follow s2.
template <int M, bool S3 = true>
__device__ __forceinline__ void accumulate(float* acc, float gp, float a, float gx, float gy, float dx, float dy) {
    constexpr int B = S3 ? 3 : 2;  // B_j at acc[B + j]
    acc[0] += gp;
    acc[1] = fmaf(gp, gp, acc[1]);
    if constexpr (S3) acc[2] = fmaf(a, gp, acc[2]);
    const float tx = gx * gp, ty = gy * gp;
    acc[B] += tx;
    acc[B + 1] += ty;
    if (M >= 6) {
        acc[B + 2] = fmaf(tx, dx, acc[B + 2]);
        acc[B + 3] = fmaf(ty, dx, acc[B + 3]);
        acc[B + 4] = fmaf(tx, dy, acc[B + 4]);
        acc[B + 5] = fmaf(ty, dy, acc[B + 5]);
    }
    if (M == 12) {
        const float hx = 0.5f * dx * dx, hy = 0.5f * dy * dy, xy = dx * dy;
        acc[B + 6] = fmaf(tx, hx, acc[B + 6]);
        acc[B + 7] = fmaf(ty, hx, acc[B + 7]);
        acc[B + 8] = fmaf(tx, xy, acc[B + 8]);
        acc[B + 9] = fmaf(ty, xy, acc[B + 9]);
        acc[B + 10] = fmaf(tx, hy, acc[B + 10]);
        acc[B + 11] = fmaf(ty, hy, acc[B + 11]);
    }
}

// The omitted-mode monomials of one pixel (subset.rs:760-775).
template <int M>
__device__ __forceinline__ void modes32(float xi, float ze, float* mo) {
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
}

// x = -(L Lᵀ)⁻¹ b with the packed lower factor and its reciprocal diagonal.
template <int M>
__device__ __forceinline__ void chol_solve_neg_rd(const double* L, const double* rd, const double* b, double* x) {
    double y[M];
#pragma unroll
    for (int i = 0; i < M; ++i) {
        double s = b[i];
#pragma unroll
        for (int k = 0; k < i; ++k) s = fma(-L[i * (i + 1) / 2 + k], y[k], s);
        y[i] = s * rd[i];
    }
#pragma unroll
    for (int i = M - 1; i >= 0; --i) {
        double s = y[i];
#pragma unroll
        for (int k = i + 1; k < M; ++k) s = fma(-L[k * (k + 1) / 2 + i], x[k], s);
        x[i] = s * rd[i];
    }
#pragma unroll
    for (int i = 0; i < M; ++i) x[i] = -x[i];
}

// H⁻¹ of the symmetric positive definite Hessian given as its upper
// triangle, row by row (hu[(a, b)], a <= b), through the Cholesky factor
// L Lᵀ = H (one reciprocal per column, every step in registers) and the two
// substitutions of column c of the identity, run by lane c (lanes >= M
// solve column 0 and store nothing), which stores column c of Hi; false
// when a pivot is not positive (Hi is then meaningless).  Every lane
// factorises alike, so a column's value does not depend on the lane.
template <int M>
__device__ __forceinline__ bool chol_inverse_cols(const double* hu, double* Hi, int lane) {
    double L[M][M], rd[M];
    bool ok = true;
#pragma unroll
    for (int j = 0; j < M; ++j) {
        double d = hu[j * M - j * (j - 1) / 2];
#pragma unroll
        for (int k = 0; k < j; ++k) d = fma(-L[j][k], L[j][k], d);
        ok = ok && d > 0.0;
        const double l = sqrt(d);
        rd[j] = 1.0 / l;
#pragma unroll
        for (int i = j + 1; i < M; ++i) {
            double t = hu[j * M - j * (j - 1) / 2 + (i - j)];
#pragma unroll
            for (int k = 0; k < j; ++k) t = fma(-L[i][k], L[j][k], t);
            L[i][j] = t * rd[j];
        }
    }
    const int c = lane < M ? lane : 0;
    double y[M], x[M];
#pragma unroll
    for (int i = 0; i < M; ++i) {
        double t = (i == c) ? 1.0 : 0.0;
#pragma unroll
        for (int k = 0; k < i; ++k) t = fma(-L[i][k], y[k], t);
        y[i] = t * rd[i];
    }
#pragma unroll
    for (int i = M - 1; i >= 0; --i) {
        double t = y[i];
#pragma unroll
        for (int k = i + 1; k < M; ++k) t = fma(-L[k][i], x[k], t);
        x[i] = t * rd[i];
    }
    if (lane < M) {
#pragma unroll
        for (int i = 0; i < M; ++i) Hi[i * M + c] = x[i];
    }
    return ok;
}


// The factorisation of cholesky<M> (L packed row-major lower triangle, rd
// = reciprocal diagonal) by the lanes of one wavefront, column by column:
// the diagonal's square root, the column below it divided by it (lane i
// takes row i), then the trailing block less the column's outer product
// (one entry per lane).  Every entry meets the same subtractions in the
// same order as in the one-lane form.  A (row-major, LDS) is overwritten;
// every lane of the wavefront calls it; the result is the same in all.
template <int M>
__device__ __forceinline__ bool cholesky_lanes(double* A, double* L, double* rd, int lane) {
#pragma unroll
    for (int k = 0; k < M; ++k) {
        const double akk = A[k * M + k];
        if (!(akk > 0.0)) return false;
        const double d = sqrt(akk);
        if (lane == 0) {
            L[k * (k + 1) / 2 + k] = d;
            rd[k] = 1.0 / d;
        }
        if (lane > k && lane < M) {
            const double lik = A[lane * M + k] / d;
            A[lane * M + k] = lik;
            L[lane * (lane + 1) / 2 + k] = lik;
        }
        wave_lds_sync();
        const int T = M - 1 - k;  // trailing block T x T, lower triangle used (k unrolled: a constant)
#pragma unroll
        for (int e0 = 0; e0 < T * T; e0 += 64) {
            const int e = e0 + lane;
            if (e < T * T) {
                const int i = k + 1 + e / T, j = k + 1 + e % T;
                if (j <= i) A[i * M + j] -= A[i * M + k] * A[j * M + k];
            }
        }
        wave_lds_sync();
    }
    return true;
}

// Column c of (LLᵀ)⁻¹ from the factor (rd = reciprocal diagonal), written to
// Hinv (row-major): forward then backward substitution of e_c.  The factor
// is read at lane-invariant addresses, so lanes solving different columns
// share each LDS read.
template <int M>
__device__ __forceinline__ void chol_inverse_column(const double* L, const double* rd, int c, double* Hinv) {
    double x[M];
#pragma unroll
    for (int i = 0; i < M; ++i) {
        double s = (i == c) ? 1.0 : 0.0;
#pragma unroll
        for (int k = 0; k < i; ++k) s = fma(-L[i * (i + 1) / 2 + k], x[k], s);
        x[i] = s * rd[i];
    }
#pragma unroll
    for (int i = M - 1; i >= 0; --i) {
        double s = x[i];
#pragma unroll
        for (int k = i + 1; k < M; ++k) s = fma(-L[k * (k + 1) / 2 + i], x[k], s);
        x[i] = s * rd[i];
    }
#pragma unroll
    for (int i = 0; i < M; ++i) Hinv[i * M + c] = x[i];
}

// Order-1 composition W(p)·W(dp)⁻¹ (subset.rs:499-528) with one reciprocal
// of the determinant; false when W(dp) is singular.  No contraction, so
// that every instantiation (whole batches and lists alike) rounds alike.
__device__ __forceinline__ bool compose6_fast(const double* p, const double* dp, double* out) {
#pragma clang fp contract(off)
    const double a = 1.0 + dp[2], b = dp[4], c = dp[0];
    const double d = dp[3], e = 1.0 + dp[5], f = dp[1];
    const double det = a * e - b * d;
    if (det == 0.0) return false;
    const double r = rcp_nr(det);
    // inverse of [[a b c] [d e f] [0 0 1]]
    const double i00 = e * r, i01 = -b * r, i02 = (b * f - c * e) * r;
    const double i10 = -d * r, i11 = a * r, i12 = (c * d - a * f) * r;
    const double o00 = 1.0 + p[2], o01 = p[4], o02 = p[0];
    const double o10 = p[3], o11 = 1.0 + p[5], o12 = p[1];
    const double n00 = o00 * i00 + o01 * i10, n01 = o00 * i01 + o01 * i11, n02 = o00 * i02 + o01 * i12 + o02;
    const double n10 = o10 * i00 + o11 * i10, n11 = o10 * i01 + o11 * i11, n12 = o10 * i02 + o11 * i12 + o12;
    out[0] = n02;
    out[1] = n12;
    out[2] = n00 - 1.0;
    out[3] = n10;
    out[4] = n01;
    out[5] = n11 - 1.0;
    return true;
}

// What one iteration needs, decided by the lead wavefront from the FP64
// warp and published through LDS: the window to use (and whether to stage
// it first), whether the fast path applies, and the FP32 warp.
struct Plan {
    int x0, y0, ww, wh, cx_lo, cx_hi, cy_lo, cy_hi;
    int ls, lds, restage, fast;
    int diag;  // predicted last iteration: the pass also sums the diagnostic's projections
    float q[12];  // q[2] = 1 + u_x, q[5] = 1 + v_y, second-order terms pre-halved
    float c0x, c0y;  // centre fraction + (u, v): the warped centre relative to the base
};

template <int M, int S>
__device__ __forceinline__ Plan make_plan(const double* p, const int32_t ext[4], double xc, double yc, float fxc,
                                          float fyc, const Window& Wcur, const BatchArgs& A) {
    const double margin = 1e-3;  // covers the FP32 rounding of per-pixel positions
    double bx0, bx1, by0, by1;
    warp_box<M>(p, ext, bx0, bx1, by0, by1);
    const double ax0 = xc + bx0 - margin, ax1 = xc + bx1 + margin;
    const double ay0 = yc + by0 - margin, ay1 = yc + by1 + margin;
    int c[4];
    const bool box_ok = box_cells(ax0, ax1, ay0, ay1, A, c);
    Window W = Wcur;
    bool restage = false;
    if (!box_ok) {
        W = no_window();
    } else if (!covers(Wcur, c)) {
        // box plus slack, else the bare box, else no window
        int cs[4];
        const double sl = (double)A.win_slack;
        W = no_window();
        if (box_cells(ax0 - sl, ax1 + sl, ay0 - sl, ay1 + sl, A, cs)) W = window_for<S>(cs, A);
        if (!W.lds) W = window_for<S>(c, A);
        restage = W.lds;
    }
    Plan P;
    P.x0 = W.x0;
    P.y0 = W.y0;
    P.ww = W.ww;
    P.wh = W.wh;
    P.ls = W.ls;
    P.cx_lo = W.cx_lo;
    P.cx_hi = W.cx_hi;
    P.cy_lo = W.cy_lo;
    P.cy_hi = W.cy_hi;
    P.lds = W.lds ? 1 : 0;
    P.restage = restage ? 1 : 0;
    P.fast = (box_ok && covers(W, c) && ax0 >= 0.0 && ax1 <= A.wmax && ay0 >= 0.0 && ay1 <= A.hmax) ? 1 : 0;
    P.diag = 0;
#pragma unroll
    for (int k = 0; k < 12; ++k) P.q[k] = (k < M) ? (float)p[k] : 0.f;
    if (M >= 6) {
        P.q[2] = (float)(1.0 + p[2]);
        P.q[5] = (float)(1.0 + p[5]);
    }
    if (M == 12) {
        P.q[6] = (float)(0.5 * p[6]);
        P.q[7] = (float)(0.5 * p[7]);
        P.q[10] = (float)(0.5 * p[10]);
        P.q[11] = (float)(0.5 * p[11]);
    }
    P.c0x = fxc + (float)p[0];
    P.c0y = fyc + (float)p[1];
    return P;
}

__device__ __forceinline__ Window plan_window(const Plan& P) {
    Window W;
    W.x0 = P.x0;
    W.y0 = P.y0;
    W.ww = P.ww;
    W.wh = P.wh;
    W.ls = P.ls;
    W.cx_lo = P.cx_lo;
    W.cx_hi = P.cx_hi;
    W.cy_lo = P.cy_lo;
    W.cy_hi = P.cy_hi;
    W.lds = P.lds != 0;
    return W;
}

// Plan read back from LDS: the integers into scalar registers, the warp as
// plain (lane-invariant) loads.
__device__ __forceinline__ Plan load_plan(const Plan* sp) {
    Plan P;
    const int* si = (const int*)sp;
    int* di = (int*)&P;
    constexpr int NI = (int)(offsetof(Plan, q) / 4);
#pragma unroll
    for (int k = 0; k < NI; ++k) di[k] = uni(si[k]);
#pragma unroll
    for (int k = NI; k < (int)(sizeof(Plan) / 4); ++k) di[k] = si[k];
    return P;
}

// One pass over the pixels at plan `pl` (its window W staged less f_m):
// acc[0, 3 + M) gets s1, s2, s3, B_j (without S3: s1, s2, B_j in
// acc[0, 2 + M)), and with DIAG acc[3 + M, 3 + M + 2K) gets the
// diagnostic's projection sums Σ c·mode·g' per gradient axis (c = ∇f less
// its mean (mx, my), g' = g - f_m).  Every form performs the same FP32
// operations on the sums it shares with the others.
template <int M, int P, int WPS, int S, bool DIAG, bool S3 = true>
__device__ __forceinline__ void pixel_pass(float* acc, const Plan& pl, const Window& W, const float (&am)[P],
                                           const float (&gxr)[P], const float (&gyr)[P], const OffXY* sXY, int n,
                                           int tid, int bx, int by, float f_m32, float mx, float my,
                                           const float* win, const BatchArgs& A) {
    constexpr int NT = 64 * WPS;
    constexpr int PU = P <= 4 ? P : 1;
    constexpr int K = Modes<M>::K;
    static_assert(S3 || !DIAG, "a diagnostic pass sums s3");
    // the rolled loop runs the slots the template fills (a uniform count)
    const int KP = PU == P ? P : min(P, (n + NT - 1) / NT);
    constexpr int NA = (S3 ? 3 : 2) + M;
    constexpr int NV = DIAG ? NA + 2 * K : NA;
#pragma unroll
    for (int j = 0; j < NV; ++j) acc[j] = 0.f;
    const auto diag = [&](float gp, float2 d, int k) {
        if constexpr (DIAG) {
            const float cx = gxr[k] - mx, cy = gyr[k] - my;
            float mo[4];
            modes32<M>(d.x, d.y, mo);
#pragma unroll
            for (int q = 0; q < K; ++q) {
                const float t = gp * mo[q];
                acc[NA + q] = fmaf(cx, t, acc[NA + q]);
                acc[NA + K + q] = fmaf(cy, t, acc[NA + K + q]);
            }
        }
    };
    if (pl.fast) {
        // positions less 1/2, rounded to nearest: the cell and the
        // centred basis coordinate s in [-1/2, 1/2] in two operations
        const float c0xh = uni(pl.c0x - 0.5f), c0yh = uni(pl.c0y - 0.5f);
        const int LS = S > 0 ? S : W.ls;
        const int cb = (by - W.cy_lo) * LS + (bx - W.cx_lo);
        const float wwf = (float)LS;
        const auto pixel = [&](int k, int i) {
            const float2 d = off_xy(sXY[i]);
            float ox, oy;
            warp_pos<M>(pl.q, c0xh, c0yh, d.x, d.y, ox, oy);
            const float rx = rintf(ox), ry = rintf(oy);
            // offsets inside the window are small integers: exact in FP32
            float cc[36];
            gather_block<S>(win, cb + (int)fmaf(ry, wwf, rx), W.ls, cc);
            float a[6], b[6];
            basis_f32_s(oy - ry, a);
            basis_f32_s(ox - rx, b);
            const float gp = contract6r(cc, a, b);
            accumulate<M, S3>(acc, gp, am[k], gxr[k], gyr[k], d.x, d.y);
            diag(gp, d, k);
        };
        if constexpr (PU == P) {
#pragma unroll
            for (int k = 0; k < P; ++k) {
                const int i = tid + NT * k;
                if (i < n) pixel(k, i);
            }
        } else {
            // slots every lane fills need no per-lane test
            const int KF = n / NT;
#pragma unroll 1
            for (int k = 0; k < KF; ++k) pixel(k, tid + NT * k);
            if (KF < KP) {
                const int i = tid + NT * KF;
                if (i < n) pixel(KF, i);
            }
        }
    } else {
#pragma unroll PU
        for (int k = 0; k < KP; ++k) {
            const int i = tid + NT * k;
            if (i < n) {
                const float2 d = off_xy(sXY[i]);
                float ox, oy;
                warp_pos<M>(pl.q, pl.c0x, pl.c0y, d.x, d.y, ox, oy);
                int cx, cy;
                float tx, ty;
                cell_frac(bx, ox, A.w, cx, tx);
                cell_frac(by, oy, A.h, cy, ty);
                float d0, d1;
                const float gp = eval32<false>(win, A.gC, A.pitch, W, A.border, cx, cy, tx, ty, f_m32, d0, d1);
                accumulate<M, S3>(acc, gp, am[k], gxr[k], gyr[k], d.x, d.y);
                diag(gp, d, k);
            }
        }
    }
}

// ZNSSD from the pass sums (subset.rs:418-446): 2 - 2 (s3 - mg Σa)/(Δf Δg),
// contraction off so that every call site rounds alike.
__device__ __forceinline__ double znssd_sums(double s3, double mg, double sa, double delta_f, double dg) {
#pragma clang fp contract(off)
    return 2.0 - 2.0 * (s3 - mg * sa) / (delta_f * dg);
}

// The warp-adequacy scores (subset.rs:748-820) from the projection sums
// G_l = Σ c·mode·g' of the last iteration's pass (uniform), lane l < 2K
// taking projection l = (axis l / K, mode l % K).  With z = c / sd, the
// standardised column sc = (z·mode - rs) / spread and e the zero-normalised
// residual, Σ sc·(e - ē) = (Σ c·mode·e / sd - rs Σe) / spread (Σ sc = 0),
// and Σ c·mode·e = (Σ c·mode·a)/Δf - (G - (g_m - f_m')·Σ c·mode)/Δg, so
// projection l = dc[l] - rg·(dc[2K + l]·G_l - mgp·dc[4K + l]) with the
// reference-only terms dc from setup; mgp = g_m - f_m' (f_m' the FP32
// shift of the window), rg = 1/Δg.  Returns η_u², η_v² (sums of squares
// in mode order) in every lane.
template <int K>
__device__ __forceinline__ void diag_eta2(const double* G, const double* dc, double rg, double mgp, int lane,
                                          double& eu, double& ev) {
    double g = 0.0;
#pragma unroll
    for (int l = 0; l < 2 * K; ++l) g = (lane == l) ? G[l] : g;
    const int l = lane < 2 * K ? lane : 0;
    const double pr = dc[l] - rg * (dc[2 * K + l] * g - mgp * dc[4 * K + l]);
    const double p2 = pr * pr;
    eu = 0.0;
    ev = 0.0;
#pragma unroll
    for (int q = 0; q < K; ++q) {
        eu += readlane_f64(p2, q);
        ev += readlane_f64(p2, K + q);
    }
}

// LIST: a list solve, reading the initial warps from A.p0_list and writing
// the result records to A.res_list in dispatch order (a separate
// instantiation, so whole-batch solves compile exactly as without it)
template <int M, int P, int WPS, int S, bool LIST = false>
__global__ void GEODIC_BATCH_ATTRS(WPS, (M == 12 ? 1 : (P <= 4 ? 2 : GEODIC_BATCH_OCC))) icgn_batch_kernel(BatchArgs A) {
    constexpr int NT = 64 * WPS;
    // With a few pixels per thread (short lists spread over many wavefronts
    // for latency) the pixel loops are unrolled so that the pixels' LDS
    // reads and arithmetic overlap; otherwise they stay rolled.
    constexpr int PU = P <= 4 ? P : 1;
    constexpr int NL = M * (M + 1) / 2;
    constexpr int K = Modes<M>::K;
    constexpr int NA = 3 + M;                          // per-iteration sums
    constexpr int NH1 = (M <= 6) ? NL : 0;             // Hessian entries summed in the first setup pass
    constexpr bool kHinv = true;                       // the update is one dot product per lane with H⁻¹
    constexpr int V1 = 2 + 2 * M + NH1;                // first setup pass
    constexpr int V2 = 4 + 6 * K;                      // second setup pass: f and the diagnostic's reference sums
    constexpr int NV = NA + 2 * K;                     // a diagnostic pass: the iteration sums + projections
    // The last iteration's pass also sums the diagnostic's projections when
    // it is predicted: the previous update's norm below 6·max_norm (or the
    // iteration cap next) foretells the last iteration for ~99 % of the
    // subsets; a missed one runs that pass once more after the loop.  The
    // second-order warp keeps its register budget and always runs it after.
#ifndef GEODIC_PREDICT
#define GEODIC_PREDICT 1
#endif
    constexpr bool kPredict = GEODIC_PREDICT && M <= 6;
    // With the prediction, a pass not predicted to be the last leaves out
    // s3 (it serves only the last iteration's ZNSSD and the history): one
    // sum fewer per pixel and per tree.  History runs every pass as a
    // predicted one; a missed prediction's ZNSSD comes from the pass after
    // the loop, which evaluates the same warp with the same operations.
    constexpr bool kS3 = !kPredict;
    constexpr int NA0 = (kS3 ? 3 : 2) + M;  // sums of a pass not predicted to be the last
    GEODIC_DYN_LDS(float, lds);
    float* win = lds;
    OffXY* sXY = (OffXY*)(lds + A.win_cap);
    __shared__ double sH[M == 12 ? M * M : 1];  // order 2: the Hessian, then its factor's workspace
    __shared__ double sL[M == 12 ? NL : 1];
    __shared__ double sRd[M == 12 ? M : 1];
    __shared__ double sAS[2 * M];               // A_j, S_j
    __shared__ float sDg[2];                    // mean gradient mx, my
    __shared__ double sDc[6 * K];               // the diagnostic's reference terms (diag_eta2)
    __shared__ double sFin[1 + 2 * K];          // last iteration: ZNSSD, the projections' sums
    __shared__ double sP[12];                   // the current warp (its only copy)
    __shared__ double sK[5];                    // f_m, Δf, Σa, 1/n, √n: read where needed
    __shared__ double sDp[M];                   // the update (order 2), -H⁻¹∇ZNSSD (orders 0, 1)
    __shared__ double sPn[M == 12 ? 12 : 1];    // the composed warp's entries (order 2)
    __shared__ double sHv[kHinv ? M * M : 1];  // the Hessian's inverse
    __shared__ double sHAS[M <= 6 ? 2 * M : 1];  // H⁻¹A, H⁻¹S
    __shared__ double sG[4];  // g_m, Δg, squared norm of the last update, its mean s1/n
    __shared__ double sRed[red_rows<WPS>() * kRedSlots];
    __shared__ int sCtl[3];
    __shared__ Plan sPlan;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const double qnan = __longlong_as_double(0x7ff8000000000000LL);
    const double margin = 1e-3;  // covers the FP32 rounding of per-pixel positions

    // One workgroup per subset, dispatched largest first (perm).  A lane-0
    // atomic work counter broadcast with a shuffle is avoided on purpose:
    // the compiler restructures that divergent loop so that its back-edge
    // skips the atomic.
    const int64_t idx = blockIdx.x;
    if (idx >= A.n_sub) return;
    // the lead wavefront issues the read of its initial warp as early as
    // it can (a list's sits in host memory beside the dispatch order), so
    // that the first plan does not wait for another round trip
    double p0v = 0.0;
    if constexpr (LIST) {
        if (wave == 0 && lane < M) p0v = A.p0_list[12 * idx + lane];
    }
    const int64_t s = A.perm[idx];
    if constexpr (!LIST) {
        if (wave == 0 && lane < M) p0v = A.p0[12 * s + lane];
    }
    const SubMeta& md = A.meta[s];
    const int64_t st = md.st;
    const int n = md.n;
    // the rolled loops run the slots the template fills (a uniform count)
    const int KP = PU == P ? P : min(P, (n + NT - 1) / NT);
    const int32_t* __restrict__ offs = A.tpl_xy + st;
    int32_t ext[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) ext[q] = md.ext[q];
    const double xc = md.xc, yc = md.yc;
    // integer base and FP32 fraction of the centre
    const double bxd = floor(xc), byd = floor(yc);
    const bool centre_ok = bxd > -1.0e9 && bxd < 1.0e9 && byd > -1.0e9 && byd < 1.0e9;
    const int bx = centre_ok ? (int)bxd : 0, by = centre_ok ? (int)byd : 0;
    const float fxc = uni((float)(xc - bxd)), fyc = uni((float)(yc - byd));
    const double dn = (double)n;
    dic_subset_result* R = LIST ? A.res_list + idx : A.res + s;
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
    // list solves: the workgroup's record is complete; raised with a
    // system-scope release so that a host that sees the flag sees the record
    const auto list_done = [&]() {
        if constexpr (LIST) {
            if (tid == 0 && A.done_list) __hip_atomic_store(A.done_list + idx, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    };
    if (n == 0 || n > NT * P || !centre_ok) {
        if (tid == 0) R->status = DIC_ERR_INVALID_INPUT;
        list_done();
        return;
    }
    // the offsets arrive as the pairs the LDS keeps: a plain copy
    for (int i = tid; i < n; i += NT) reinterpret_cast<int32_t*>(sXY)[i] = offs[i];

    float am[P], gxr[P], gyr[P];

    // ---- Subset::new on the staged reference window: one pass evaluates f
    // and ∇f and sums everything that does not need f_m, using the
    // provisional shift c: Σ sdi_j (f - f_m) = Σ sdi_j (f - c) - (f_m - c) S_j ----
    const float cshift = 127.5f;
    double f_m, delta_f, sa;
    {
        const double ax0 = xc + ext[0] - margin, ax1 = xc + ext[1] + margin;
        const double ay0 = yc + ext[2] - margin, ay1 = yc + ext[3] + margin;
        int c[4];
        Window W = no_window();
        const bool inside = ax0 >= 0.0 && ax1 <= A.wmax && ay0 >= 0.0 && ay1 <= A.hmax;
        // every pixel sits at an integer offset from the centre: in the fast
        // path one basis serves the whole subset, and a centre on a whole
        // pixel puts every pixel on a knot, where the 5x5 form applies
        const bool knot = fxc == 0.f && fyc == 0.f;
        // knots inside the image: f and ∇f come from the reference's knot
        // table (the same operations, done once per image), read from the
        // offsets in HBM so that the loads overlap the copy and the barrier
        const bool table = knot && inside && A.knot != nullptr;
        if (table) {
            const OffXY* __restrict__ oxy = reinterpret_cast<const OffXY*>(offs);
#pragma unroll PU
            for (int k = 0; k < KP; ++k) {
                const int i = tid + NT * k;
                float4 t = make_float4(0.f, 0.f, 0.f, 0.f);
                if (i < n) {
                    const float2 d = off_xy(oxy[i]);
                    t = A.knot[(int64_t)(by + (int)d.y) * A.w + (bx + (int)d.x)];
                }
                am[k] = t.x;
                gxr[k] = t.y;
                gyr[k] = t.z;
            }
        } else {
            if (box_cells(ax0, ax1, ay0, ay1, A, c)) W = window_for<S>(c, A);
            if (W.lds) stage_window<NT>(win, A.fC, A.pitch, W, tid, 0.f);
        }
        __syncthreads();
        const bool fast = W.lds && inside;
        float u[V1];
#pragma unroll
        for (int j = 0; j < V1; ++j) u[j] = 0.f;
        float ba[6], bb[6], bda[6], bdb[6];
        if (!knot) {
            basis_f32(fyc, ba);
            basis_f32(fxc, bb);
            dbasis_f32(fyc, bda);
            dbasis_f32(fxc, bdb);
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                ba[q] = uni(ba[q]);
                bb[q] = uni(bb[q]);
                bda[q] = uni(bda[q]);
                bdb[q] = uni(bdb[q]);
            }
        }
        const int LS = S > 0 ? S : W.ls;
        const int cb = (by - W.cy_lo) * LS + (bx - W.cx_lo);
        const float wwf = (float)LS;
#pragma unroll PU
        for (int k = 0; k < KP; ++k) {
            const int i = tid + NT * k;
            float v = 0.f, gx = 0.f, gy = 0.f;
            if (i < n) {
                const float2 d = off_xy(sXY[i]);
                if (table) {
                    v = am[k];
                    gx = gxr[k];
                    gy = gyr[k];
                } else if (fast && knot) {
                    float cc[25];
                    gather_block5<S>(win, cb + (int)fmaf(d.y, wwf, d.x), W.ls, cc);
                    v = knot_grad(cc, gx, gy);
                } else if (fast) {
                    float cc[36];
                    gather_block<S>(win, cb + (int)fmaf(d.y, wwf, d.x), W.ls, cc);
                    v = contract6r_grad(cc, ba, bb, bda, bdb, gx, gy);
                } else {
                    int cx, cy;
                    float tx, ty;
                    cell_frac(bx, fxc + d.x, A.w, cx, tx);
                    cell_frac(by, fyc + d.y, A.h, cy, ty);
                    v = eval32<true>(win, A.fC, A.pitch, W, A.border, cx, cy, tx, ty, 0.f, gx, gy);
                }
                const float a = v - cshift;
                u[0] += v;
                u[1] = fmaf(0.5f, fmaf(gx, gx, gy * gy), u[1]);
                float sd[M];
                sdi32<M>(gx, gy, d.x, d.y, sd);
#pragma unroll
                for (int j = 0; j < M; ++j) {
                    u[2 + j] = fmaf(sd[j], a, u[2 + j]);
                    u[2 + M + j] += sd[j];
                }
                if constexpr (NH1 > 0) {
                    int q = 2 + 2 * M;
#pragma unroll
                    for (int aa = 0; aa < M; ++aa)
#pragma unroll
                        for (int b2 = aa; b2 < M; ++b2) {
                            u[q] = fmaf(sd[aa], sd[b2], u[q]);
                            ++q;
                        }
                }
                v = a;
            }
            am[k] = v;
            gxr[k] = gx;
            gyr[k] = gy;
        }
        double w1[V1];
        block_sum32<WPS, V1>(u, w1, sRed, lane, wave);
        f_m = uni(w1[0] / dn);
        const double dlt = f_m - (double)cshift;
        if (tid == 0) {
            R->f_m = f_m;
            R->sssig = w1[1];
#pragma unroll
            for (int j = 0; j < M; ++j) {
                sAS[j] = w1[2 + j] - dlt * w1[2 + M + j];
                sAS[M + j] = w1[2 + M + j];
            }
        }
        if constexpr (NH1 > 0) {
            // orders 0 and 1: the lead wavefront factorises and inverts the
            // Hessian (summed in the pass above) in registers, while the
            // other wavefronts go on to the second pass
            if (wave == 0) {
                const bool ok = chol_inverse_cols<M>(w1 + 2 + 2 * M, sHv, lane);
                if (lane == 0) sCtl[0] = ok ? 0 : DIC_ERR_SINGULAR_HESSIAN;
                // row j of H⁻¹ times A and S, lane j: the constant parts of
                // every update (A_j = Σ sdi_j (f - f_m), S_j = Σ sdi_j)
                wave_lds_sync();
                if (lane < M) {
                    double ha = 0.0, hs = 0.0;
#pragma unroll
                    for (int k = 0; k < M; ++k) {
                        const double h = sHv[lane * M + k];
                        ha = fma(h, w1[2 + k] - dlt * w1[2 + M + k], ha);
                        hs = fma(h, w1[2 + M + k], hs);
                    }
                    sHAS[lane] = ha;
                    sHAS[M + lane] = hs;
                }
            }
        }
        if constexpr (M == 12) {
            hessian_rows<M, P, WPS, 0, 4>(gxr, gyr, sXY, n, tid, lane, wave, sH, sRed);
            hessian_rows<M, P, WPS, 4, 12>(gxr, gyr, sXY, n, tid, lane, wave, sH, sRed);
        }
        // (the factorisation's status reaches every wavefront through the
        // barriers of the next reduction)
        if constexpr (M == 12) {
            // order 2: the lead wavefront factorises over its lanes, then
            // H⁻¹ = L⁻ᵀL⁻¹ once, lane c solving for column c, so each
            // iteration's update is one dot product per lane instead of two
            // dependent substitutions
            __syncthreads();
            if (wave == 0) {
                const bool ok = cholesky_lanes<M>(sH, sL, sRd, lane);
                if (lane == 0) sCtl[0] = ok ? 0 : DIC_ERR_SINGULAR_HESSIAN;
            }
            __syncthreads();
            if (wave == 0 && lane < M && sCtl[0] == 0) chol_inverse_column<M>(sL, sRd, lane, sHv);
        }
        // ... and the lead wavefront plans the first iteration
        if (wave == 0) {
            double p0[M];
#pragma unroll
            for (int k = 0; k < M; ++k) p0[k] = readlane_f64(p0v, k);
            Plan pl = make_plan<M, S>(p0, ext, xc, yc, fxc, fyc, no_window(), A);
            pl.diag = (kPredict && (A.max_it == 1 || A.hist)) ? 1 : 0;
            if (lane == 0) {
                sCtl[2] = 0;
                sPlan = pl;
                for (int k = 0; k < M; ++k) sP[k] = p0[k];
            }
        }

        // ---- second pass: a = f - f_m in place, Σa, Σa², and the
        // diagnostic's reference sums: per gradient axis (∇f less its mean:
        // c) Σc², Σ c·mode, Σ (c·mode)², Σ c·mode·a (subset.rs:698-727,
        // 786-818) ----
        const float d32 = (float)dlt;
        const float mx = uni((float)(w1[2 + M] / dn));  // Σ sdi_0 = Σ ∂x f
        const float my = uni((float)(w1[3 + M] / dn));  // Σ sdi_1 = Σ ∂y f
        float v2[V2];
#pragma unroll
        for (int j = 0; j < V2; ++j) v2[j] = 0.f;
#pragma unroll PU
        for (int k = 0; k < KP; ++k) {
            const int i = tid + NT * k;
            if (i < n) {
                const float a = am[k] - d32;
                am[k] = a;
                v2[0] += a;
                v2[1] = fmaf(a, a, v2[1]);
                const float cx = gxr[k] - mx, cy = gyr[k] - my;
                v2[2] = fmaf(cx, cx, v2[2]);
                v2[3] = fmaf(cy, cy, v2[3]);
                const float2 d = off_xy(sXY[i]);
                float mo[4];
                modes32<M>(d.x, d.y, mo);
#pragma unroll
                for (int q = 0; q < K; ++q) {
                    // explicit FMAs throughout: every instantiation rounds alike
                    const float xm = cx * mo[q], ym = cy * mo[q];
                    v2[4 + q] += xm;
                    v2[4 + K + q] += ym;
                    v2[4 + 2 * K + q] = fmaf(xm, xm, v2[4 + 2 * K + q]);
                    v2[4 + 3 * K + q] = fmaf(ym, ym, v2[4 + 3 * K + q]);
                    v2[4 + 4 * K + q] = fmaf(xm, a, v2[4 + 4 * K + q]);
                    v2[4 + 5 * K + q] = fmaf(ym, a, v2[4 + 5 * K + q]);
                }
            }
        }
        double w2[V2];
        block_sum32<WPS, V2>(v2, w2, sRed, lane, wave);
        sa = w2[0];
        const double sq = w2[1];
        delta_f = uni(sqrt(sq));
        if (wave == 0) {
            // the diagnostic's reference terms, projection l on lane l: with
            // sd the gradient's spread, rs and spread the mean and spread of
            // z·mode (z = c / sd), α = 1/(sd·spread·n), β = rs/(spread·n):
            // dc = (α Σ c·mode·a - β Σa)/Δf, α, α Σ c·mode
            if (lane < 2 * K) {
                const double rn = 1.0 / dn;
                const double isd = 1.0 / sqrt(fmax((lane < K ? w2[2] : w2[3]) * rn, 1e-18));
                double scm = 0.0, scm2 = 0.0, scma = 0.0;
#pragma unroll
                for (int j = 0; j < 2 * K; ++j) {
                    if (j == lane) {
                        scm = w2[4 + j];
                        scm2 = w2[4 + 2 * K + j];
                        scma = w2[4 + 4 * K + j];
                    }
                }
                const double rs = scm * isd * rn;
                const double var = scm2 * isd * isd * rn - rs * rs;
                const double isp = 1.0 / sqrt(fmax(var, 1e-18));
                const double al = isd * isp * rn, be = rs * isp * rn;
                sDc[lane] = (al * scma - be * sa) / delta_f;
                sDc[2 * K + lane] = al;
                sDc[4 * K + lane] = al * scm;
            }
        }
        if (tid == 0) {
            R->delta_f = delta_f;
            R->sigma = sqrt(sq / dn);
            sDg[0] = mx;
            sDg[1] = my;
        }
    }
    if (delta_f == 0.0) {
        if (tid == 0) R->status = DIC_ERR_INVALID_INPUT;
        list_done();
        return;
    }
    if (sCtl[0]) {
        if (tid == 0) R->status = DIC_ERR_SINGULAR_HESSIAN;
        list_done();
        return;
    }

    // the per-subset constants and the current warp (sP) stay in LDS
    // through the iterations, out of the registers the pixel loop needs
    if (tid == 0) {
        sK[0] = f_m;
        sK[1] = delta_f;
        sK[2] = sa;
        sK[3] = 1.0 / dn;
        sK[4] = sqrt(dn);
    }
    const float f_m32 = uni((float)f_m);

    int it_done = 0;
    bool converged = false;
    int status = 0;
    double g_m = 0.0, delta_g = 0.0, last_norm = 0.0;
    Plan pl = load_plan(&sPlan);
    for (int it = 1; it <= A.max_it; ++it) {
        const Window W = plan_window(pl);
        if (pl.restage) {
            // the target window holds C - f_m, so gathers give g' = g - f_m
            stage_window<NT>(win, A.gC, A.pitch, W, tid, f_m32);
            __syncthreads();
        }
        // the lead wavefront alone runs the update; the others wait at the
        // barrier below, leaving their SIMD slots to other workgroups
        const bool dg_it = kPredict && pl.diag;
        double sum[NV];
        bool lead;
        if (dg_it) {
            float acc[NV];
            if constexpr (kPredict)
                pixel_pass<M, P, WPS, S, true>(acc, pl, W, am, gxr, gyr, sXY, n, tid, bx, by, f_m32, sDg[0], sDg[1],
                                               win, A);
            lead = block_sum_lead32<WPS, NV>(acc, sum, sRed, lane, wave);
            // the projections' sums wait in LDS in case this pass is the last
            if (lead && lane == 0) {
#pragma unroll
                for (int l = 0; l < 2 * K; ++l) sFin[1 + l] = sum[NA + l];
            }
        } else {
            float acc[NA0];
            pixel_pass<M, P, WPS, S, false, kS3>(acc, pl, W, am, gxr, gyr, sXY, n, tid, bx, by, f_m32, 0.f, 0.f, win,
                                                  A);
            lead = block_sum_lead32<WPS, NA0>(acc, sum, sRed, lane, wave);
            if constexpr (!kS3) {
                // the layout of a pass with s3 (none here)
#pragma unroll
                for (int j = M - 1; j >= 0; --j) sum[3 + j] = sum[2 + j];
                sum[2] = 0.0;
            }
        }
        const bool has_s3 = kS3 || dg_it;
        if (lead) {
            const double f_m = sK[0], delta_f = sK[1], sa = sK[2], rn = sK[3], size = sK[4];
            const double max_norm2 = A.max_norm2;
            // g' above is g - (float)f_m; fold the rounding of the shift back in
            const double shift = (double)f_m32 - f_m;
            const double s1 = sum[0] + dn * shift;
            const double s2 = sum[1] + shift * (2.0 * sum[0] + dn * shift);
            const double s3 = fma(shift, sa, sum[2]);
            const double mg = s1 * rn;
            const double gm = f_m + mg;
            const double var = fmax(s2 - s1 * mg, 0.0);
            // 1/Δg: a zero spread keeps Δf/Δg = Δf/0 (inf, or NaN for Δf = 0)
            const double rdg = var > 0.0 ? rsqrt_nr(var) : __builtin_inf();
            const double dg = var > 0.0 ? var * rdg : 0.0;
            const double ratio = delta_f * rdg;
            double pn[M];
            bool ok;
            double nrm2;
            if constexpr (M == 12) {
                // lane j solves component j; the update then sits in LDS,
                // and the whole wavefront composes it with the warp
                double gz[M];
#pragma unroll
                for (int j = 0; j < M; ++j)
                    gz[j] = sAS[j] - ratio * (sum[3 + j] + shift * sAS[M + j] - mg * sAS[M + j]);
                if (lane < M) {
                    // three chains of four: a shorter dependent path
                    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
#pragma unroll
                    for (int k = 0; k < M / 3; ++k) {
                        a0 = fma(sHv[lane * M + k], gz[k], a0);
                        a1 = fma(sHv[lane * M + M / 3 + k], gz[M / 3 + k], a1);
                        a2 = fma(sHv[lane * M + 2 * M / 3 + k], gz[2 * M / 3 + k], a2);
                    }
                    sDp[lane] = -((a0 + a1) + a2);
                }
                __builtin_amdgcn_wave_barrier();
                ok = compose12_lanes(sP, sDp, lane, pn, sPn);
                nrm2 = convergence_norm2_dev<M>(sDp, size);
            } else {
                // lane j: component j of -H⁻¹∇ZNSSD, then broadcast
                double dp[M];
                double acc = 0.0;
                // with ∇ZNSSD = A - ratio (B' - mg S), B' = B + shift S:
                // H⁻¹∇ZNSSD = H⁻¹A - ratio (H⁻¹B + (shift - mg) H⁻¹S)
                if (lane < M) {
                    // two chains of half the terms: a shorter dependent path
                    double h0 = 0.0, h1 = 0.0;
#pragma unroll
                    for (int k = 0; k < M / 2; ++k) {
                        h0 = fma(sHv[lane * M + k], sum[3 + k], h0);
                        h1 = fma(sHv[lane * M + M / 2 + k], sum[3 + M / 2 + k], h1);
                    }
                    const double hb = h0 + h1;
                    const double hs = sHAS[M + lane];
                    acc = sHAS[lane] - ratio * fma(shift - mg, hs, hb);
                }
                // through LDS: lane-invariant loads instead of lane reads
                // into scalar registers
                if (lane < M) sDp[lane] = acc;
                wave_lds_sync();
#pragma unroll
                for (int k = 0; k < M; ++k) dp[k] = -sDp[k];
                double p[M];
#pragma unroll
                for (int k = 0; k < M; ++k) p[k] = sP[k];
                if constexpr (M == 6)
                    ok = compose6_fast(p, dp, pn);
                else
                    ok = compose_icgn_dev<M>(p, dp, pn, nullptr);  // order 0: p - dp
                // squared norm against the squared limit: no square root per iteration
                nrm2 = convergence_norm2_dev<M>(dp, size);
            }
            const bool conv = ok && nrm2 < max_norm2;
            const bool last = ok && (conv || it == A.max_it);
            const bool keep_h = ok && A.hist && it <= A.hist_cap;
            // ZNSSD at this iteration's warp (subset.rs:418-446 in sums); a
            // last pass without s3 leaves it to the pass after the loop
            double znssd = 0.0;
            if ((last || keep_h) && has_s3) znssd = znssd_sums(s3, mg, sa, delta_f, dg);
            if (last && lane == 0) {
                if (has_s3) sFin[0] = znssd;
                if (dg_it) sCtl[2] = 1;  // its projections' sums are in sFin
            }
            if (keep_h && lane == 0) {
                const double nrm = sqrt(nrm2);
                double* hh = A.hist + ((int64_t)s * A.hist_cap + (it - 1)) * 4;
                hh[0] = (double)it;
                hh[1] = nrm;
                hh[2] = 1.0 - znssd / 2.0;
                hh[3] = znssd;
            }
            // plan the next iteration while the other wavefronts wait
            Plan nx = pl;
            if (ok && !conv && it < A.max_it) {
                nx = make_plan<M, S>(pn, ext, xc, yc, fxc, fyc, W, A);
                nx.diag = (kPredict && (A.hist || nrm2 < A.diag_norm2 || it + 1 == A.max_it)) ? 1 : 0;
            }
            if (lane == 0) {
#pragma unroll
                for (int k = 0; k < M; ++k) sP[k] = pn[k];
                sCtl[0] = ok ? 0 : DIC_ERR_SINGULAR_WARP;
                sCtl[1] = conv ? 1 : 0;
                if (last) {
                    sG[0] = gm;
                    sG[1] = dg;
                    sG[2] = nrm2;
                    sG[3] = mg;
                }
                sPlan = nx;
            }
        }
        if constexpr (WPS > 1) {
            __syncthreads();
        } else {
            __builtin_amdgcn_wave_barrier();
        }
        // the flags and the next plan in one round of LDS reads (a last
        // iteration left its own plan there, so the pass after the loop
        // reads the same one)
        const int st0 = sCtl[0], st1 = sCtl[1];
        pl = load_plan(&sPlan);
        if (st0) {
            status = st0;
            break;
        }
        it_done = it;
        if (st1) {
            converged = true;
            break;
        }
    }
    if (status) {
        if (tid == 0) {
            R->status = status;
            R->iterations = it_done;
        }
        list_done();
        return;
    }
    if (it_done) {
        g_m = uni(sG[0]);
        delta_g = uni(sG[1]);
        last_norm = uni(sqrt(sG[2]));
    }

    // ---- the diagnostic (subset.rs:1196-1210) on g at the last iteration's
    // warp: its projections were summed in that pass when it was predicted,
    // else that pass runs once more here (its plan and staged window are
    // still current) with the same operations; with no iteration g is all
    // zero and the reference's NaN propagates alike ----
    double eu = 0.0, ev = 0.0;
    if (it_done) {
        double sum[NV];
        bool lead = wave == 0;
        if (!sCtl[2]) {
            const Window W = plan_window(pl);
            float acc[NV];
            pixel_pass<M, P, WPS, S, true>(acc, pl, W, am, gxr, gyr, sXY, n, tid, bx, by, f_m32, sDg[0], sDg[1], win,
                                           A);
            lead = block_sum_lead32<WPS, NV>(acc, sum, sRed, lane, wave);
            if constexpr (!kS3) {
                // the last iteration's pass left out s3: its ZNSSD from this
                // one, with that iteration's mean and spread
                if (lead && lane == 0) {
                    const double shift = (double)f_m32 - f_m;
                    sFin[0] = znssd_sums(fma(shift, sa, sum[2]), sG[3], sa, delta_f, sG[1]);
                }
            }
        } else if (lead) {
#pragma unroll
            for (int l = 0; l < 2 * K; ++l) sum[NA + l] = sFin[1 + l];
        }
        if (lead) diag_eta2<K>(sum + NA, sDc, 1.0 / delta_g, g_m - (double)f_m32, lane, eu, ev);
    }
    if (tid == 0) {
        const double znssd = it_done ? sFin[0] : 4.0;
        const double zncc = it_done ? 1.0 - znssd / 2.0 : 0.0;
        for (int k = 0; k < 12; ++k) {
            const double pk = k < M ? sP[k] : 0.0;
            R->p[k] = pk;
            A.p_out[12 * s + k] = pk;
        }
        R->c_zncc = zncc;
        R->c_znssd = znssd;
        R->norm = it_done ? last_norm : qnan;
        R->iterations = it_done;
        R->converged = converged ? 1 : 0;
        R->quality_ok = (converged && zncc >= A.tol) ? 1 : 0;
        // with no iteration e is NaN (g_m, Δg undefined), as in the reference
        R->eta_u = it_done ? sqrt(eu) : qnan;
        R->eta_v = it_done ? sqrt(ev) : qnan;
    }
    list_done();
}

}  // namespace batch
}  // namespace geodic
