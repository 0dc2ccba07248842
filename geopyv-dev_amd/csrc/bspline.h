// Bi-quintic B-spline basis and coefficient-block evaluation, device side.
//
// C is the (h+2b) x (w+2b) coefficient array; the 6x6 block of pixel
// (yf, xf) starts at C[yf+b-2, xf+b-2] (image.rs:48-51).  A query (x, y)
// is clamped to [0, w-1e-9] x [0, h-1e-9] before the floor (image.rs:
// 130-131), so out-of-range queries saturate to the boundary value.
#pragma once

#ifndef GEODIC_SIMT
#include <hip/hip_runtime.h>
#endif
#include <cstdint>

// Nothing is scheduled across this point (device compiler only).
#ifndef GEODIC_SCHED_BARRIER
#define GEODIC_SCHED_BARRIER() __builtin_amdgcn_sched_barrier(0)
#endif

namespace geodic {

// Q (image.rs:183-190): row r = power of t, column k = tap.
#define GQ00 (1.0 / 120.0)
#define GQ01 (13.0 / 60.0)
#define GQ02 (11.0 / 20.0)
#define GQ03 (13.0 / 60.0)
#define GQ04 (1.0 / 120.0)
#define GQ10 (-1.0 / 24.0)
#define GQ11 (-5.0 / 12.0)
#define GQ13 (5.0 / 12.0)
#define GQ14 (1.0 / 24.0)
#define GQ20 (1.0 / 12.0)
#define GQ21 (1.0 / 6.0)
#define GQ22 (-1.0 / 2.0)
#define GQ23 (1.0 / 6.0)
#define GQ24 (1.0 / 12.0)
#define GQ30 (-1.0 / 12.0)
#define GQ31 (1.0 / 6.0)
#define GQ33 (-1.0 / 6.0)
#define GQ34 (1.0 / 12.0)
#define GQ40 (1.0 / 24.0)
#define GQ41 (-1.0 / 6.0)
#define GQ42 (1.0 / 4.0)
#define GQ43 (-1.0 / 6.0)
#define GQ44 (1.0 / 24.0)
#define GQ50 (-1.0 / 120.0)
#define GQ51 (1.0 / 24.0)
#define GQ52 (-1.0 / 12.0)
#define GQ53 (1.0 / 12.0)
#define GQ54 (-1.0 / 24.0)
#define GQ55 (1.0 / 120.0)

__device__ __forceinline__ double clamp_ref(double v, double lo, double hi) {
    // f64::clamp: NaN passes through
    return v < lo ? lo : (v > hi ? hi : v);
}

// Qᵀ·[1, t, ..., t⁵] with the reference's power-sum association
// (image.rs:194-208): out[k] = Σ_r Q[r][k]·t^r, r ascending, from 0.0.
__device__ __forceinline__ void qt_poly_ref(double t, double o[6]) {
#pragma clang fp contract(off)
    const double t2 = t * t, t3 = t2 * t, t4 = t3 * t, t5 = t4 * t;
    o[0] = ((((0.0 + GQ00) + GQ10 * t) + GQ20 * t2) + GQ30 * t3 + GQ40 * t4) + GQ50 * t5;
    o[1] = ((((0.0 + GQ01) + GQ11 * t) + GQ21 * t2) + GQ31 * t3 + GQ41 * t4) + GQ51 * t5;
    o[2] = ((((0.0 + GQ02) + 0.0 * t) + GQ22 * t2) + 0.0 * t3 + GQ42 * t4) + GQ52 * t5;
    o[3] = ((((0.0 + GQ03) + GQ13 * t) + GQ23 * t2) + GQ33 * t3 + GQ43 * t4) + GQ53 * t5;
    o[4] = ((((0.0 + GQ04) + GQ14 * t) + GQ24 * t2) + GQ34 * t3 + GQ44 * t4) + GQ54 * t5;
    o[5] = GQ55 * t5;
}

// Qᵀ·[0, 1, 2t, 3t², 4t³, 5t⁴] (image.rs:217-230)
__device__ __forceinline__ void qt_dpoly_ref(double t, double o[6]) {
#pragma clang fp contract(off)
    const double t2 = t * t, t3 = t2 * t;
    const double p1 = 1.0, p2 = 2.0 * t, p3 = 3.0 * t2, p4 = 4.0 * t3, p5 = 5.0 * t2 * t2;
    o[0] = (((GQ10 * p1 + GQ20 * p2) + GQ30 * p3) + GQ40 * p4) + GQ50 * p5;
    o[1] = (((GQ11 * p1 + GQ21 * p2) + GQ31 * p3) + GQ41 * p4) + GQ51 * p5;
    o[2] = (((0.0 * p1 + GQ22 * p2) + 0.0 * p3) + GQ42 * p4) + GQ52 * p5;
    o[3] = (((GQ13 * p1 + GQ23 * p2) + GQ33 * p3) + GQ43 * p4) + GQ53 * p5;
    o[4] = (((GQ14 * p1 + GQ24 * p2) + GQ34 * p3) + GQ44 * p4) + GQ54 * p5;
    o[5] = GQ55 * p5;
}

// FP64 value from the global coefficient array (image.rs:128-148).
__device__ __forceinline__ double eval_f64(const double* __restrict__ C, int64_t wp, int border, double wmax,
                                           double hmax, double x, double y) {
#pragma clang fp contract(off)
    x = clamp_ref(x, 0.0, wmax);
    y = clamp_ref(y, 0.0, hmax);
    // after the clamp only NaN fails x >= 0: it indexes cell 0 with a NaN
    // fraction, as the reference's saturating `as usize` cast does
    const double fx = (x >= 0.0) ? floor(x) : 0.0, fy = (y >= 0.0) ? floor(y) : 0.0;
    double a[6], b[6];
    qt_poly_ref(y - fy, a);
    qt_poly_ref(x - fx, b);
    const double* base = C + ((int64_t)fy + border - 2) * wp + ((int64_t)fx + border - 2);
    double val = 0.0;
#pragma unroll
    for (int r = 0; r < 6; ++r) {
        const double* row = base + r * wp;
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < 6; ++k) s += row[k] * b[k];
        val += a[r] * s;
    }
    return val;
}

// FP64 value and gradient (image.rs:115-124,155-178).
__device__ __forceinline__ void eval_grad_f64(const double* __restrict__ C, int64_t wp, int border, double wmax,
                                              double hmax, double x, double y, double& v, double& gx, double& gy) {
#pragma clang fp contract(off)
    x = clamp_ref(x, 0.0, wmax);
    y = clamp_ref(y, 0.0, hmax);
    // after the clamp only NaN fails x >= 0: it indexes cell 0 with a NaN
    // fraction, as the reference's saturating `as usize` cast does
    const double fx = (x >= 0.0) ? floor(x) : 0.0, fy = (y >= 0.0) ? floor(y) : 0.0;
    double a[6], da[6], b[6], db[6];
    qt_poly_ref(y - fy, a);
    qt_dpoly_ref(y - fy, da);
    qt_poly_ref(x - fx, b);
    qt_dpoly_ref(x - fx, db);
    const double* base = C + ((int64_t)fy + border - 2) * wp + ((int64_t)fx + border - 2);
    double val = 0.0, ggx = 0.0, ggy = 0.0;
#pragma unroll
    for (int r = 0; r < 6; ++r) {
        const double* row = base + r * wp;
        double s = 0.0, t = 0.0;
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            s += row[k] * b[k];
            t += row[k] * db[k];
        }
        val += a[r] * s;
        ggx += a[r] * t;
        ggy += da[r] * s;
    }
    v = val;
    gx = ggx;
    gy = ggy;
}

// FP32 basis: the same polynomials as qt_poly, evaluated about the cell
// centre.  With s = t - 1/2 and z = s², each mirror pair b_k, b_{5-k}
// (b_{5-k}(t) = b_k(1-t)) splits into an even part E and an odd part O,
// b_k = E + O, b_{5-k} = E - O, each a quadratic in z: 23 operations for the
// six weights instead of 28 in power form.
__device__ __forceinline__ void basis_f32_s(float s, float b[6]) {
    const float z = s * s;
    const float e0 = fmaf(fmaf(z, 1.f / 48.f, 1.f / 96.f), z, 1.f / 3840.f);
    const float o0 = s * fmaf(fmaf(z, -1.f / 120.f, -1.f / 48.f), z, -1.f / 384.f);
    const float e1 = fmaf(fmaf(z, -1.f / 16.f, 7.f / 32.f), z, 79.f / 1280.f);
    const float o1 = s * fmaf(fmaf(z, 1.f / 24.f, -1.f / 16.f), z, -25.f / 128.f);
    const float e2 = fmaf(fmaf(z, 1.f / 24.f, -11.f / 48.f), z, 841.f / 1920.f);
    const float o2 = s * fmaf(fmaf(z, -1.f / 12.f, 7.f / 24.f), z, -77.f / 192.f);
    b[0] = e0 + o0;
    b[5] = e0 - o0;
    b[1] = e1 + o1;
    b[4] = e1 - o1;
    b[2] = e2 + o2;
    b[3] = e2 - o2;
}

__device__ __forceinline__ void basis_f32(float t, float b[6]) { basis_f32_s(t - 0.5f, b); }

__device__ __forceinline__ void dbasis_f32(float t, float d[6]) {
    d[0] = (float)GQ10 + t * (2.f * (float)GQ20 + t * (3.f * (float)GQ30 + t * (4.f * (float)GQ40 + t * 5.f * (float)GQ50)));
    d[1] = (float)GQ11 + t * (2.f * (float)GQ21 + t * (3.f * (float)GQ31 + t * (4.f * (float)GQ41 + t * 5.f * (float)GQ51)));
    d[2] = t * (2.f * (float)GQ22 + t * t * (4.f * (float)GQ42 + t * 5.f * (float)GQ52));
    d[3] = (float)GQ13 + t * (2.f * (float)GQ23 + t * (3.f * (float)GQ33 + t * (4.f * (float)GQ43 + t * 5.f * (float)GQ53)));
    d[4] = (float)GQ14 + t * (2.f * (float)GQ24 + t * (3.f * (float)GQ34 + t * (4.f * (float)GQ44 + t * 5.f * (float)GQ54)));
    const float t2 = t * t;
    d[5] = 5.f * (float)GQ55 * (t2 * t2);
}

// Value known to be equal on every lane, moved to scalar registers.
__device__ __forceinline__ double uni(double v) {
    const int lo = __builtin_amdgcn_readfirstlane(__double2loint(v));
    const int hi = __builtin_amdgcn_readfirstlane(__double2hiint(v));
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ float uni(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// Wavefront sum with DPP row operations (no LDS traffic): pairs and quads
// by quad permutes, rows of 16 by row shifts, then the row broadcasts carry
// rows 0-1 and 2-3 together; lane 63 ends with the total, which is read
// into scalar registers.  The addition tree is fixed, so the result does not
// depend on anything but the 64 inputs.  Needs every lane active.
template <int CTRL, int ROW_MASK, bool BOUND_CTRL>
__device__ __forceinline__ double dpp_move(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROW_MASK, 0xf, BOUND_CTRL);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROW_MASK, 0xf, BOUND_CTRL);
    return __hiloint2double(hi, lo);
}

// The single-value tree: lane 63 ends with the total.
__device__ __forceinline__ double wave_tree1(double v) {
    v += dpp_move<0xb1, 0xf, true>(v);   // quad_perm [1,0,3,2]
    v += dpp_move<0x4e, 0xf, true>(v);   // quad_perm [2,3,0,1]
    v += dpp_move<0x114, 0xf, true>(v);  // row_shr:4
    v += dpp_move<0x118, 0xf, true>(v);  // row_shr:8
    v += dpp_move<0x142, 0xa, false>(v); // row_bcast:15 into rows 1 and 3
    v += dpp_move<0x143, 0xc, false>(v); // row_bcast:31 into rows 2 and 3
    return v;
}

__device__ __forceinline__ double wave_sum_u(double v) {
    v = wave_tree1(v);
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
    return __hiloint2double(hi, lo);
}

// Several wavefront sums at once (transposed tree).  Groups of eight
// values go through: a permlane32 swap pairing two values (lanes 0-31 then
// carry the first, 32-63 the second), a permlane16 swap pairing two such
// registers (one value per 16-lane row), a bank-masked row rotation by 8
// pairing two of those (one value per 8-lane half-row), and three DPP steps
// inside each half-row.  Eight sums cost 35 VALU operations instead of 8 x 18
// for separate trees; leftover values use the single-value tree.  Value
// 8u + 4h + o of a group ends in lane 16·row(o) + 8h + 4, row = (0, 2, 1, 3)[o].
// The addition tree is fixed.  Needs every lane active.
template <int CTRL, int BANK_MASK>
__device__ __forceinline__ double dpp_merge(double old, double src) {
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(old), __double2loint(src), CTRL, 0xf, BANK_MASK, false);
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(old), __double2hiint(src), CTRL, 0xf, BANK_MASK, false);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double swap32_sum(double x, double y) {
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)__double2loint(x), (unsigned)__double2loint(y), false,
                                                     false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)__double2hiint(x), (unsigned)__double2hiint(y), false,
                                                     false);
    return __hiloint2double((int)hi[0], (int)lo[0]) + __hiloint2double((int)hi[1], (int)lo[1]);
}

__device__ __forceinline__ double swap16_sum(double x, double y) {
    const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)__double2loint(x), (unsigned)__double2loint(y), false,
                                                     false);
    const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)__double2hiint(x), (unsigned)__double2hiint(y), false,
                                                     false);
    return __hiloint2double((int)hi[0], (int)lo[0]) + __hiloint2double((int)hi[1], (int)lo[1]);
}

// Eight wavefront sums at once: the register whose lane 16·row(o) + 8h + 4
// holds the sum of value 4h + o, row = (0, 2, 1, 3)[o].
__device__ __forceinline__ double wave_tree8(const double* x) {
    const double r0 = swap32_sum(x[0], x[1]), r1 = swap32_sum(x[2], x[3]);
    const double r2 = swap32_sum(x[4], x[5]), r3 = swap32_sum(x[6], x[7]);
    const double t0 = swap16_sum(r0, r1), t1 = swap16_sum(r2, r3);
    // lanes 0-7 of each row: t0's value; lanes 8-15: t1's
    double w = dpp_merge<0x128, 0xc>(t0, t1) + dpp_merge<0x128, 0x3>(t1, t0);  // row_ror:8
    w += dpp_move<0x114, 0xf, true>(w);  // row_shr:4 (lanes 4-7, 12-15 pair correctly)
    w += dpp_move<0x4e, 0xf, true>(w);   // quad_perm [2,3,0,1]
    w += dpp_move<0xb1, 0xf, true>(w);   // quad_perm [1,0,3,2]
    return w;
}

// The V wavefront sums of in[], each written to dst[j] by the lane that
// holds it (no lane reads into scalar registers): groups of eight through
// wave_tree8, a remainder of three or more as a padded group, else
// single-value trees.  Needs every lane active.
template <int V>
__device__ __forceinline__ void wave_sums_to(const double* in, double* dst, int lane) {
    constexpr int G = V / 8, R = V % 8;
    const bool holder = (lane & 7) == 4;
    const int row = lane >> 4;
    const int slot = 4 * ((lane >> 3) & 1) + (row == 1 ? 2 : (row == 2 ? 1 : row));  // 4h + o
#pragma unroll
    for (int u = 0; u < G; ++u) {
        const double w = wave_tree8(in + 8 * u);
        if (holder) dst[8 * u + slot] = w;
    }
    if constexpr (R >= 3) {
        double x[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) x[j] = j < R ? in[8 * G + j] : 0.0;
        const double w = wave_tree8(x);
        if (holder && slot < R) dst[8 * G + slot] = w;
    } else {
#pragma unroll
        for (int j = 8 * G; j < V; ++j) {
            const double t = wave_tree1(in[j]);
            if (lane == 63) dst[j] = t;
        }
    }
}


// A template offset pair as stored in LDS: two half-precision numbers
// (offsets are integers far inside the exact range of FP16), so the FP32
// arithmetic that uses them can take them as mixed-precision operands.
#ifdef GEODIC_SIMT
typedef int16_t offh_t;
#else
typedef _Float16 offh_t;
#endif
struct OffXY {
    offh_t x, y;
};

__device__ __forceinline__ float2 off_xy(const OffXY o) { return make_float2((float)o.x, (float)o.y); }

__device__ __forceinline__ int unpack_dx(int32_t v) { return (int)(int16_t)(v & 0xffff); }
__device__ __forceinline__ int unpack_dy(int32_t v) { return (int)(int16_t)((uint32_t)v >> 16); }

template <int M>
struct PLen {};

}  // namespace geodic
