// Small dense FP64 linear algebra for the per-subset Gauss-Newton update,
// usable from device code (one lane or all lanes redundantly).
#pragma once

#include <cstdint>

#ifndef GEODIC_SIMT
#include <hip/hip_runtime.h>
#endif

namespace geodic {

// Inverse with nalgebra's try_inverse semantics (the rounding the reference
// gets from nalgebra 0.33.3): closed-form cofactors for m <= 3, otherwise LU
// with partial pivoting (first max |.|, reciprocal-multiply Gauss steps),
// then unit-lower and upper triangular solves against the permuted
// identity.  Row-major; A is overwritten.  Returns false when singular.
template <int M>
__device__ __forceinline__ bool inverse_ref(double* A, double* out) {
#pragma clang fp contract(off)
    if (M == 2) {
        const double det = A[0] * A[3] - A[2] * A[1];
        if (det == 0.0) return false;
        out[0] = A[3] / det;
        out[1] = -A[1] / det;
        out[2] = -A[2] / det;
        out[3] = A[0] / det;
        return true;
    }
    if (M == 3) {
        const double m11 = A[4] * A[8] - A[7] * A[5];
        const double m12 = A[3] * A[8] - A[6] * A[5];
        const double m13 = A[3] * A[7] - A[6] * A[4];
        const double det = A[0] * m11 - A[1] * m12 + A[2] * m13;
        if (det == 0.0) return false;
        out[0] = m11 / det;
        out[1] = -(A[1] * A[8] - A[7] * A[2]) / det;
        out[2] = (A[1] * A[5] - A[4] * A[2]) / det;
        out[3] = -m12 / det;
        out[4] = (A[0] * A[8] - A[6] * A[2]) / det;
        out[5] = -(A[0] * A[5] - A[3] * A[2]) / det;
        out[6] = m13 / det;
        out[7] = -(A[0] * A[7] - A[6] * A[1]) / det;
        out[8] = (A[0] * A[4] - A[3] * A[1]) / det;
        return true;
    }
    // row i was exchanged with row piv(i) >= i (4 bits each; piv(i) = i: no
    // exchange), replayed on the identity in the same order afterwards
    uint64_t pivots = 0;
#pragma unroll 1
    for (int i = 0; i < M; ++i) {
        int piv = i;
        double best = fabs(A[i * M + i]);
        for (int r = i + 1; r < M; ++r) {
            const double v = fabs(A[r * M + i]);
            if (v > best) {
                best = v;
                piv = r;
            }
        }
        const double diag = A[piv * M + i];
        if (diag == 0.0) {
            pivots |= (uint64_t)i << (4 * i);
            continue;
        }
        pivots |= (uint64_t)piv << (4 * i);
        if (piv != i) {
            for (int c = 0; c < M; ++c) {
                const double t = A[i * M + c];
                A[i * M + c] = A[piv * M + c];
                A[piv * M + c] = t;
            }
        }
        const double inv_diag = 1.0 / diag;
        for (int r = i + 1; r < M; ++r) A[r * M + i] *= inv_diag;
#pragma unroll 1
        for (int k = i + 1; k < M; ++k) {
            const double pk = A[i * M + k];
            for (int r = i + 1; r < M; ++r) A[r * M + k] = -pk * A[r * M + i] + A[r * M + k];
        }
    }
    for (int i = 0; i < M; ++i)
        if (A[i * M + i] == 0.0) return false;
    for (int i = 0; i < M * M; ++i) out[i] = 0.0;
    for (int i = 0; i < M; ++i) out[i * M + i] = 1.0;
#pragma unroll 1
    for (int i = 0; i < M; ++i) {
        const int piv = (int)((pivots >> (4 * i)) & 15);
        if (piv != i)
            for (int c = 0; c < M; ++c) {
                const double t = out[i * M + c];
                out[i * M + c] = out[piv * M + c];
                out[piv * M + c] = t;
            }
    }
#pragma unroll 1
    for (int c = 0; c < M; ++c) {
        for (int i = 0; i < M; ++i) {
            const double coeff = out[i * M + c];
            for (int r = i + 1; r < M; ++r) out[r * M + c] = -coeff * A[r * M + i] + out[r * M + c];
        }
        for (int i = M - 1; i >= 0; --i) {
            const double coeff = out[i * M + c] / A[i * M + i];
            out[i * M + c] = coeff;
            for (int r = 0; r < i; ++r) out[r * M + c] = -coeff * A[r * M + i] + out[r * M + c];
        }
    }
    return true;
}

// C = A·B (row-major, m x m), k ascending.
template <int M>
__device__ void matmul_ref(const double* A, const double* B, double* C) {
#pragma clang fp contract(off)
#pragma unroll
    for (int i = 0; i < M; ++i)
#pragma unroll
        for (int j = 0; j < M; ++j) {
            double s = A[i * M] * B[j];
#pragma unroll
            for (int k = 1; k < M; ++k) s = s + A[i * M + k] * B[k * M + j];
            C[i * M + j] = s;
        }
}

// s_terms of compose_icgn (subset.rs:547-571) and the lifted 6x6 warp.
__device__ inline void lifted6(const double* q, double* W) {
#pragma clang fp contract(off)
    const double u = q[0], v = q[1], ux = q[2], vx = q[3], uy = q[4], vy = q[5];
    const double uxx = q[6], vxx = q[7], uxy = q[8], vxy = q[9], uyy = q[10], vyy = q[11];
    const double s0 = 2.0 * ux + ux * ux + u * uxx;
    const double s1 = 2.0 * u * uxy + 2.0 * (1.0 + ux) * uy;
    const double s2 = uy * uy + u * uyy;
    const double s3 = 2.0 * u * (1.0 + ux);
    const double s4 = 2.0 * u * uy;
    const double s5 = u * u;
    const double s6 = 0.5 * (v * uxx + 2.0 * (1.0 + ux) * vx + u * vxx);
    const double s7 = uy * vx + ux * vy + v * uxy + u * vxy + vy + ux;
    const double s8 = 0.5 * (v * uyy + 2.0 * uy * (1.0 + vy) + u * vyy);
    const double s9 = v + v * ux + u * vx;
    const double s10 = u + v * uy + u * vy;
    const double s11 = u * v;
    const double s12 = vx * vx + v * vxx;
    const double s13 = 2.0 * v * vxy + 2.0 * vx * (1.0 + vy);
    const double s14 = 2.0 * vy + vy * vy + v * vyy;
    const double s15 = 2.0 * v * vx;
    const double s16 = 2.0 * v * (1.0 + vy);
    const double s17 = v * v;
    W[0] = 1.0 + s0; W[1] = s1; W[2] = s2; W[3] = s3; W[4] = s4; W[5] = s5;
    W[6] = s6; W[7] = 1.0 + s7; W[8] = s8; W[9] = s9; W[10] = s10; W[11] = s11;
    W[12] = s12; W[13] = s13; W[14] = 1.0 + s14; W[15] = s15; W[16] = s16; W[17] = s17;
    W[18] = 0.5 * uxx; W[19] = uxy; W[20] = 0.5 * uyy; W[21] = 1.0 + ux; W[22] = uy; W[23] = u;
    W[24] = 0.5 * vxx; W[25] = vxy; W[26] = 0.5 * vyy; W[27] = vx; W[28] = 1.0 + vy; W[29] = v;
    W[30] = 0.0; W[31] = 0.0; W[32] = 0.0; W[33] = 0.0; W[34] = 0.0; W[35] = 1.0;
}

// compose_icgn (subset.rs:487-611): p_new = W(p)·W(dp)^-1.  scratch holds
// 4*36 doubles for order 2.  Returns false when W(dp) is singular.
template <int M>
__device__ bool compose_icgn_dev(const double* p, const double* dp, double* out, double* scratch) {
#pragma clang fp contract(off)
    if (M == 2) {
        out[0] = p[0] - dp[0];
        out[1] = p[1] - dp[1];
        return true;
    }
    if (M == 6) {
        double wo[9] = {1.0 + p[2], p[4], p[0], p[3], 1.0 + p[5], p[1], 0.0, 0.0, 1.0};
        double wd[9] = {1.0 + dp[2], dp[4], dp[0], dp[3], 1.0 + dp[5], dp[1], 0.0, 0.0, 1.0};
        double wi[9], wn[9];
        if (!inverse_ref<3>(wd, wi)) return false;
        matmul_ref<3>(wo, wi, wn);
        out[0] = wn[2];
        out[1] = wn[5];
        out[2] = wn[0] - 1.0;
        out[3] = wn[3];
        out[4] = wn[1];
        out[5] = wn[4] - 1.0;
        return true;
    }
    double* wo = scratch;
    double* wd = scratch + 36;
    double* wi = scratch + 72;
    double* wn = scratch + 108;
    lifted6(p, wo);
    lifted6(dp, wd);
    if (!inverse_ref<6>(wd, wi)) return false;
    matmul_ref<6>(wo, wi, wn);
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
    return true;
}

// A double broadcast from lane l (two 32-bit readlanes).
__device__ __forceinline__ double readlane_f64(double v, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

// 1/x and 1/sqrt(x) from the hardware estimates (v_rcp_f64, v_rsq_f64)
// refined by two Newton steps each: within an ulp or two of the correctly
// rounded result, in a shorter dependent chain than the IEEE division and
// square root sequences.  rsqrt_nr wants x > 0 normal (the coupled step
// never forms y², so it does not overflow for small x).
__device__ __forceinline__ double rcp_nr(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}
__device__ __forceinline__ double rsqrt_nr(double x) {
    double y = __builtin_amdgcn_rsq(x);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const double e = fma(-(x * y), y, 1.0);
        y = fma(0.5 * y, e, y);
    }
    return y;
}


// Order-2 composition W(p)·W(dp)⁻¹ (subset.rs:547-611) spread over a
// wavefront.  The warp is read from rows 3 and 4 of the product only
// (subset.rs:596-609), and those rows are X·W(dp)⁻¹ with X rows 3 and 4 of
// W(p): so lanes 8r + c (r < 6, c < 8) hold the augmented matrix
// [W(dp)ᵀ | Xᵀ] and Gauss-Jordan elimination with threshold pivoting (one
// column per step: the diagonal, or the column's largest entry found by
// readlanes when the diagonal is small, rows exchanged and eliminated by
// shuffles) leaves the two rows in columns 6 and 7.  Each lane
// builds its own entry of the lifted matrices (the s-terms of
// subset.rs:547-571), so no lane holds a matrix.  p and dp are read
// lane-invariantly (LDS); false when W(dp) is singular.
__device__ __forceinline__ double lifted6_entry(const double* q, int e) {
#pragma clang fp contract(off)
    const double u = q[0], v = q[1], ux = q[2], vx = q[3], uy = q[4], vy = q[5];
    const double uxx = q[6], vxx = q[7], uxy = q[8], vxy = q[9], uyy = q[10], vyy = q[11];
    double w = 0.0;
    w = e == 0 ? 1.0 + (2.0 * ux + ux * ux + u * uxx) : w;
    w = e == 1 ? 2.0 * u * uxy + 2.0 * (1.0 + ux) * uy : w;
    w = e == 2 ? uy * uy + u * uyy : w;
    w = e == 3 ? 2.0 * u * (1.0 + ux) : w;
    w = e == 4 ? 2.0 * u * uy : w;
    w = e == 5 ? u * u : w;
    w = e == 6 ? 0.5 * (v * uxx + 2.0 * (1.0 + ux) * vx + u * vxx) : w;
    w = e == 7 ? 1.0 + (uy * vx + ux * vy + v * uxy + u * vxy + vy + ux) : w;
    w = e == 8 ? 0.5 * (v * uyy + 2.0 * uy * (1.0 + vy) + u * vyy) : w;
    w = e == 9 ? v + v * ux + u * vx : w;
    w = e == 10 ? u + v * uy + u * vy : w;
    w = e == 11 ? u * v : w;
    w = e == 12 ? vx * vx + v * vxx : w;
    w = e == 13 ? 2.0 * v * vxy + 2.0 * vx * (1.0 + vy) : w;
    w = e == 14 ? 1.0 + (2.0 * vy + vy * vy + v * vyy) : w;
    w = e == 15 ? 2.0 * v * vx : w;
    w = e == 16 ? 2.0 * v * (1.0 + vy) : w;
    w = e == 17 ? v * v : w;
    w = e == 18 ? 0.5 * uxx : w;
    w = e == 19 ? uxy : w;
    w = e == 20 ? 0.5 * uyy : w;
    w = e == 21 ? 1.0 + ux : w;
    w = e == 22 ? uy : w;
    w = e == 23 ? u : w;
    w = e == 24 ? 0.5 * vxx : w;
    w = e == 25 ? vxy : w;
    w = e == 26 ? 0.5 * vyy : w;
    w = e == 27 ? vx : w;
    w = e == 28 ? 1.0 + vy : w;
    w = e == 29 ? v : w;
    w = e == 35 ? 1.0 : w;
    return w;
}

// Entry (r, c) of [W(dp)ᵀ | Xᵀ]: the quadratic rows 0-2 of W(dp) (the
// s-terms) by lifted6_entry over their 18 positions only; rows 3-5 of W(dp)
// and the two rows of W(p) in X are the warp's own terms, picked directly.
__device__ __forceinline__ double lifted_aug_entry(const double* p, const double* dp, int r, int c) {
#pragma clang fp contract(off)
    if (c < 3) return lifted6_entry(dp, 6 * c + r);
    if (c == 5) return r == 5 ? 1.0 : 0.0;
    const double* q = c < 5 ? dp : p;
    const bool vrow = c == 4 || c == 7;  // row 4 (the v terms), else row 3 (the u terms)
    const double a0 = vrow ? q[7] : q[6], a1 = vrow ? q[9] : q[8], a2 = vrow ? q[11] : q[10];
    const double a3 = vrow ? q[3] : q[2], a4 = vrow ? q[5] : q[4], a5 = vrow ? q[1] : q[0];
    double w = 0.5 * a0;
    w = r == 1 ? a1 : w;
    w = r == 2 ? 0.5 * a2 : w;
    w = r == 3 ? (vrow ? a3 : 1.0 + a3) : w;
    w = r == 4 ? (vrow ? 1.0 + a4 : a4) : w;
    w = r == 5 ? a5 : w;
    return w;
}

__device__ __forceinline__ bool compose12_lanes(const double* p, const double* dp, int lane, double* out,
                                                 double* sOut) {
    const int r = lane >> 3, c = lane & 7;
    const bool live = lane < 48;
    // G[r][c] = W(dp)[c][r] for c < 6; X[0][r], X[1][r] (= W(p)[18 + r], W(p)[24 + r]) for c = 6, 7
    double v = 0.0;
    if (live) v = lifted_aug_entry(p, dp, r, c);
    // row 5 of W(dp) is [0 0 0 0 0 1], so column 5 of G is e5 and stays so
    // while row 5 is never exchanged: the last step then has pivot 1 and
    // nothing to eliminate
    bool moved5 = false;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        if (k == 5 && !moved5) break;
        // threshold pivoting: W(dp) is the identity plus terms of the order
        // of the update, so the diagonal serves unless it falls below 1/4
        // (multipliers then stay below 4 times the column's entries); only
        // then is the column searched for its largest entry
        int piv = k;
        double best = fabs(readlane_f64(v, 8 * k + k));
        if (!(best >= 0.25)) {
#pragma unroll
            for (int i = k + 1; i < 6; ++i) {
                const double a = fabs(readlane_f64(v, 8 * i + k));
                if (a > best) {
                    best = a;
                    piv = i;
                }
            }
            if (!(best > 0.0)) return false;
        }
        moved5 = moved5 || piv == 5;
        if (piv != k) {
            const int src = ((r == k) ? piv : (r == piv ? k : r)) * 8 + c;
            const double t = __shfl(v, live ? src : lane, 64);
            v = live ? t : v;
        }
        // the pivot row travels unnormalised, so the reciprocal of the
        // pivot runs while the two exchanges are in flight
        const double dk = readlane_f64(v, 8 * k + k);
        const double f = __shfl(v, live ? 8 * r + k : lane, 64);  // G[r][k]
        const double g = __shfl(v, live ? 8 * k + c : lane, 64);  // G[k][c]
        const double rp = rcp_nr(dk);
        if (live && r == k) v *= rp;
        if (live && r != k) v = fma(-(f * rp), g, v);
    }
    // the entries through LDS (written by the lanes that hold them, read
    // back lane-invariantly) instead of lane reads into scalar registers
    if (live && c >= 6) sOut[2 * r + (c - 6)] = v;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    out[0] = sOut[10];
    out[1] = sOut[11];
    out[2] = sOut[6] - 1.0;
    out[3] = sOut[7];
    out[4] = sOut[8];
    out[5] = sOut[9] - 1.0;
    out[6] = sOut[0] * 2.0;
    out[7] = sOut[1] * 2.0;
    out[8] = sOut[2];
    out[9] = sOut[3];
    out[10] = sOut[4] * 2.0;
    out[11] = sOut[5] * 2.0;
    return true;
}

// inverse_ref over the lanes of a workgroup, with the same operations on
// every entry in the same order as the one-lane version (so the same
// rounding): every lane repeats the pivot search from LDS, the row swaps
// and the Gauss steps of the trailing block are spread over lanes, and lane
// c < M carries column c of the inverse through both triangular solves.  A
// and out are in LDS (A is overwritten); every thread of the workgroup calls
// it, and the result is the same in every lane.
template <int M>
__device__ bool inverse_ref_lanes(double* A, double* out, int lane) {
#pragma clang fp contract(off)
    if constexpr (M <= 3) {
        double t[M * M];
        const bool ok = inverse_ref<M>(A, t);
        __syncthreads();
        if (ok && lane == 0)
            for (int i = 0; i < M * M; ++i) out[i] = t[i];
        __syncthreads();
        return ok;
    } else {
        int pos = lane;  // row of the 1 in column `lane` of the permuted identity
        for (int i = 0; i < M; ++i) {
            int piv = i;
            double best = fabs(A[i * M + i]);
            for (int r = i + 1; r < M; ++r) {
                const double v = fabs(A[r * M + i]);
                if (v > best) {
                    best = v;
                    piv = r;
                }
            }
            const double diag = A[piv * M + i];
            __syncthreads();
            if (diag == 0.0) continue;
            if (piv != i) {
                if (lane < M) {
                    const double a = A[i * M + lane];
                    A[i * M + lane] = A[piv * M + lane];
                    A[piv * M + lane] = a;
                }
                pos = pos == i ? piv : (pos == piv ? i : pos);
                __syncthreads();
            }
            const double inv_diag = 1.0 / diag;
            const int rest = M - 1 - i;
            if (lane < rest) A[(i + 1 + lane) * M + i] *= inv_diag;
            __syncthreads();
            for (int e = lane; e < rest * rest; e += 64) {
                const int r = i + 1 + e / rest, k = i + 1 + e % rest;
                A[r * M + k] = -A[i * M + k] * A[r * M + i] + A[r * M + k];
            }
            __syncthreads();
        }
        bool ok = true;
        for (int i = 0; i < M; ++i) ok = ok && A[i * M + i] != 0.0;
        if (!ok) return false;
        if (lane < M) {
            double x[M];
#pragma unroll
            for (int r = 0; r < M; ++r) x[r] = r == pos ? 1.0 : 0.0;
#pragma unroll
            for (int i = 0; i < M; ++i) {
                const double coeff = x[i];
#pragma unroll
                for (int r = i + 1; r < M; ++r) x[r] = -coeff * A[r * M + i] + x[r];
            }
#pragma unroll
            for (int i = M - 1; i >= 0; --i) {
                const double coeff = x[i] / A[i * M + i];
                x[i] = coeff;
#pragma unroll
                for (int r = 0; r < i; ++r) x[r] = -coeff * A[r * M + i] + x[r];
            }
#pragma unroll
            for (int r = 0; r < M; ++r) out[r * M + lane] = x[r];
        }
        __syncthreads();
        return true;
    }
}

// compose_icgn_dev<12> over the lanes of a workgroup with the same rounding:
// the lifted W(dp) one entry per lane, inverted by inverse_ref_lanes<6>, and
// rows 3 and 4 of W(p)·W(dp)⁻¹ (the only rows the warp is read from,
// subset.rs:596-609) one entry per lane.  p and dp are read
// lane-invariantly; scratch is 96 doubles of LDS.  Every thread of the
// workgroup calls it; false when W(dp) is singular.
__device__ inline bool compose12_ref_lanes(const double* p, const double* dp, double* out, double* scratch,
                                           int lane) {
#pragma clang fp contract(off)
    double* wd = scratch;
    double* wi = scratch + 36;
    double* wo = scratch + 72;  // rows 3 and 4 of W(p)
    double* wn = scratch + 84;  // rows 3 and 4 of the product
    if (lane < 36) wd[lane] = lifted6_entry(dp, lane);
    else if (lane < 48) wo[lane - 36] = lifted6_entry(p, 18 + (lane - 36));
    __syncthreads();
    if (!inverse_ref_lanes<6>(wd, wi, lane)) return false;
    if (lane < 12) {
        const int row = lane / 6, j = lane % 6;
        double s = wo[row * 6] * wi[j];
#pragma unroll
        for (int k = 1; k < 6; ++k) s = s + wo[row * 6 + k] * wi[k * 6 + j];
        wn[lane] = s;
    }
    __syncthreads();
    out[0] = wn[5];
    out[1] = wn[6 + 5];
    out[2] = wn[3] - 1.0;
    out[3] = wn[6 + 3];
    out[4] = wn[4];
    out[5] = wn[6 + 4] - 1.0;
    out[6] = wn[0] * 2.0;
    out[7] = wn[6] * 2.0;
    out[8] = wn[1];
    out[9] = wn[6 + 1];
    out[10] = wn[2] * 2.0;
    out[11] = wn[6 + 2] * 2.0;
    __syncthreads();
    return true;
}

// Square of convergence_norm (subset.rs:622-649), for comparisons.
template <int M>
__device__ inline double convergence_norm2_dev(const double* d, double size) {
    double s = d[0] * d[0] + d[1] * d[1];
    if (M >= 6) {
#pragma unroll
        for (int k = 2; k < 6; ++k) {
            const double t = d[k] * size;
            s = fma(t, t, s);
        }
    }
    if (M == 12) {
        const double h = 0.5 * size * size;
#pragma unroll
        for (int k = 6; k < 12; ++k) {
            const double t = d[k] * h;
            s = fma(t, t, s);
        }
    }
    return s;
}

// convergence_norm (subset.rs:622-649)
template <int M>
__device__ inline double convergence_norm_dev(const double* d, double size) {
#pragma clang fp contract(off)
    if (M == 2) return sqrt(d[0] * d[0] + d[1] * d[1]);
    double s = d[0] * d[0] + d[1] * d[1];
#pragma unroll
    for (int k = 2; k < 6; ++k) {
        const double t = d[k] * size;
        s = s + t * t;
    }
    if (M == 12) {
#pragma unroll
        for (int k = 6; k < 12; ++k) {
            const double t = 0.5 * d[k] * size * size;
            s = s + t * t;
        }
    }
    return sqrt(s);
}

// Cholesky factor of an SPD m x m matrix (row-major, lower triangle packed
// row by row: L[i*(i+1)/2 + j]).  Returns false if not positive definite.
template <int M>
__device__ bool cholesky(const double* H, double* L) {
    for (int i = 0; i < M; ++i) {
        for (int j = 0; j <= i; ++j) {
            double s = H[i * M + j];
            for (int k = 0; k < j; ++k) s -= L[i * (i + 1) / 2 + k] * L[j * (j + 1) / 2 + k];
            if (i == j) {
                if (!(s > 0.0)) return false;
                L[i * (i + 1) / 2 + i] = sqrt(s);
            } else {
                L[i * (i + 1) / 2 + j] = s / L[j * (j + 1) / 2 + j];
            }
        }
    }
    return true;
}

}  // namespace geodic
