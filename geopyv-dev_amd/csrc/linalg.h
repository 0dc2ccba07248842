// Small dense FP64 linear algebra for the per-subset Gauss-Newton update,
// usable from device code (one lane or all lanes redundantly).
#pragma once

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
__device__ bool inverse_ref(double* A, double* out) {
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
    int perm_a[M], perm_b[M];
    int np = 0;
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
        if (diag == 0.0) continue;
        if (piv != i) {
            perm_a[np] = i;
            perm_b[np] = piv;
            ++np;
            for (int c = 0; c < M; ++c) {
                const double t = A[i * M + c];
                A[i * M + c] = A[piv * M + c];
                A[piv * M + c] = t;
            }
        }
        const double inv_diag = 1.0 / diag;
        for (int r = i + 1; r < M; ++r) A[r * M + i] *= inv_diag;
        for (int k = i + 1; k < M; ++k) {
            const double pk = A[i * M + k];
            for (int r = i + 1; r < M; ++r) A[r * M + k] = -pk * A[r * M + i] + A[r * M + k];
        }
    }
    for (int i = 0; i < M; ++i)
        if (A[i * M + i] == 0.0) return false;
    for (int i = 0; i < M * M; ++i) out[i] = 0.0;
    for (int i = 0; i < M; ++i) out[i * M + i] = 1.0;
    for (int q = 0; q < np; ++q)
        for (int c = 0; c < M; ++c) {
            const double t = out[perm_a[q] * M + c];
            out[perm_a[q] * M + c] = out[perm_b[q] * M + c];
            out[perm_b[q] * M + c] = t;
        }
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

// x = -(L Lᵀ)^-1 b
template <int M>
__device__ void cholesky_solve_neg(const double* L, const double* b, double* x) {
    double y[M];
#pragma unroll
    for (int i = 0; i < M; ++i) {
        double s = b[i];
#pragma unroll
        for (int k = 0; k < i; ++k) s -= L[i * (i + 1) / 2 + k] * y[k];
        y[i] = s / L[i * (i + 1) / 2 + i];
    }
#pragma unroll
    for (int i = M - 1; i >= 0; --i) {
        double s = y[i];
#pragma unroll
        for (int k = i + 1; k < M; ++k) s -= L[k * (k + 1) / 2 + i] * x[k];
        x[i] = s / L[i * (i + 1) / 2 + i];
    }
#pragma unroll
    for (int i = 0; i < M; ++i) x[i] = -x[i];
}

}  // namespace geodic
