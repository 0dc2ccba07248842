// Plain FP64 drop-in: Subset::new (subset.rs:836-919) followed by
// solve_icgn_result (subset.rs:1117-1219) or solve_fagn_result
// (subset.rs:1228-1334) for every subset of a batch.
//
// One wavefront (one 64-thread workgroup) per subset.  Arithmetic is FP64
// throughout and keeps the reference's pass structure: per-pixel f, grad f
// and g live in an HBM scratch slab (the reference's heap vectors), each
// sum is its own pass, the Hessian is inverted with nalgebra's LU
// semantics, and the update Δp = -H⁻¹·∇ZNSSD and the composition are done
// by one lane in the reference's order.  Sums are wavefront trees (fixed
// order) instead of sequential folds, so results agree with the FP64
// oracle to round-off, not bit for bit.
#include "bspline.h"
#include "internal.h"
#include "linalg.h"

namespace geodic {

namespace {

struct PlainArgs {
    const double* fC;
    const double* gC;
    int64_t wp;
    int border;
    double wmax, hmax;
    int64_t n_sub;
    const int32_t* list;  // block b solves subset list[b] (every subset in order when null)
    const double* coords;
    const int32_t* orders;
    const int32_t* tpl_index;
    const int64_t* tpl_start;
    const int32_t* tpl_xy;
    const int64_t* scr_off;  // pixels before each subset's arrays
    double* scratch;
    int arrays;              // per-pixel doubles: 4 (ICGN) or 6 (FAGN)
    const double* p0;
    double* p_out;
    dic_subset_result* res;
    double* hist;
    int hist_cap;
    int method;
    int max_it;
    double tol;
    double max_norm;
    const double* p0_list;        // with a list: warps in list order (else null)
    dic_subset_result* res_list;  // with a list: results in list order (else null)
};

template <int M>
struct Modes {
    static constexpr int K = (M == 2) ? 2 : (M == 6 ? 3 : 4);
};

template <int M>
__device__ __forceinline__ void sdi_row(double gx, double gy, double dx, double dy, double* s) {
#pragma clang fp contract(off)
    // steepest_descent (subset.rs:311-343)
    s[0] = gx;
    s[1] = gy;
    if (M >= 6) {
        s[2] = gx * dx;
        s[3] = gy * dx;
        s[4] = gx * dy;
        s[5] = gy * dy;
    }
    if (M == 12) {
        s[6] = gx * 0.5 * dx * dx;
        s[7] = gy * 0.5 * dx * dx;
        s[8] = gx * dx * dy;
        s[9] = gy * dx * dy;
        s[10] = gx * 0.5 * dy * dy;
        s[11] = gy * 0.5 * dy * dy;
    }
}

template <int M>
__device__ __forceinline__ void warp_point(const double* p, double x, double y, double xc, double yc, double& ox,
                                           double& oy) {
#pragma clang fp contract(off)
    // apply_warp_into (subset.rs:262-300)
    if (M == 2) {
        ox = x + p[0];
        oy = y + p[1];
        return;
    }
    const double dx = x - xc, dy = y - yc;
    if (M == 6) {
        ox = x + p[0] + p[2] * dx + p[4] * dy;
        oy = y + p[1] + p[3] * dx + p[5] * dy;
    } else {
        ox = x + p[0] + p[2] * dx + p[4] * dy + 0.5 * p[6] * dx * dx + p[8] * dx * dy + 0.5 * p[10] * dy * dy;
        oy = y + p[1] + p[3] * dx + p[5] * dy + 0.5 * p[7] * dx * dx + p[9] * dx * dy + 0.5 * p[11] * dy * dy;
    }
}

template <int M>
__device__ __forceinline__ void mode_row(double xi, double ze, double* mo) {
#pragma clang fp contract(off)
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

// The order-2 update over the lanes of the wavefront (subset.rs:1146-1160):
// the Hessian inverse (FAGN), dp one component per lane and the ICGN
// composition, with the one-lane rounding; lane 0 then keeps the books.
__device__ __forceinline__ void plain_update12(const PlainArgs& A, int64_t s, int it, int lane, const double* gz,
                                            double zacc, double size, const double* p, double* sH, double* sHi,
                                            double* sScr, double* sDp, double* sP, double* sRes, int* sCtl) {
#pragma clang fp contract(off)
    constexpr int M = 12;
    int status = 0;
    double pn[12];
    if (A.method == DIC_METHOD_FAGN && !inverse_ref_lanes<M>(sH, sHi, lane)) status = DIC_ERR_SINGULAR_HESSIAN;
    if (!status) {
        if (lane < M) {
            double acc2 = gz[0] * sHi[lane * M];
            for (int k = 1; k < M; ++k) acc2 = acc2 + gz[k] * sHi[lane * M + k];
            sDp[lane] = (A.method == DIC_METHOD_ICGN) ? -acc2 : acc2;
        }
        __syncthreads();
        if (A.method == DIC_METHOD_ICGN && !compose12_ref_lanes(p, sDp, pn, sScr, lane)) status = DIC_ERR_SINGULAR_WARP;
    }
    if (lane == 0) {
        if (!status) {
            double dp[12];
            for (int k = 0; k < M; ++k) dp[k] = sDp[k];
            if (A.method == DIC_METHOD_FAGN)
                for (int k = 0; k < M; ++k) pn[k] = p[k] + dp[k];
            for (int k = 0; k < M; ++k) sP[k] = pn[k];
            const double norm = convergence_norm_dev<M>(dp, size);
            const double c_zncc = 1.0 - zacc / 2.0;
            if (A.hist && it <= A.hist_cap) {
                double* h = A.hist + ((int64_t)s * A.hist_cap + (it - 1)) * 4;
                h[0] = (double)it;
                h[1] = norm;
                h[2] = c_zncc;
                h[3] = zacc;
            }
            sRes[0] = c_zncc;
            sRes[1] = zacc;
            sRes[2] = norm;
            sCtl[3] = norm < A.max_norm ? 1 : 0;
        }
        sCtl[1] = status;
    }
}

// LIST: a list solve with the warps in A.p0_list and the result records
// to A.res_list in list order (a separate instantiation, so whole-batch
// solves compile exactly as without it)
template <int M, bool LIST = false>
__global__ void __launch_bounds__(64) icgn_plain_kernel(PlainArgs A) {
#pragma clang fp contract(off)
    constexpr int NP = M * (M + 1) / 2;
    constexpr int K = Modes<M>::K;
    __shared__ double sH[M * M];
    __shared__ double sHi[M * M];
    __shared__ double sScr[144];
    __shared__ double sP[12];
    __shared__ double sDp[12];
    __shared__ int sCtl[4];  // [0] stop flag, [1] status, [2] iterations, [3] converged
    __shared__ double sRes[4];  // last zncc, znssd, norm

    if ((int64_t)blockIdx.x >= A.n_sub) return;
    const int64_t s = A.list ? (int64_t)A.list[blockIdx.x] : (int64_t)blockIdx.x;
    const int lane = threadIdx.x;
    const int t = A.tpl_index[s];
    const int64_t st = A.tpl_start[t];
    const int64_t n = A.tpl_start[t + 1] - st;
    const int32_t* offs = A.tpl_xy + st;
    const double xc = A.coords[2 * s], yc = A.coords[2 * s + 1];
    const double dn = (double)n;
    dic_subset_result* R = LIST ? A.res_list + blockIdx.x : A.res + s;

    if (lane == 0) {
        for (int k = 0; k < 12; ++k) R->p[k] = 0.0;
        R->n_px = (int32_t)n;
        R->order = (M == 2) ? 0 : (M == 6 ? 1 : 2);
        R->iterations = 0;
        R->converged = 0;
        R->quality_ok = 0;
        R->status = 0;
        R->c_zncc = 0.0;
        R->c_znssd = 4.0;
        R->norm = __longlong_as_double(0x7ff8000000000000LL);
        R->eta_u = R->eta_v = 0.0;
        R->f_m = R->delta_f = R->sssig = R->sigma = 0.0;
    }
    if (n == 0) {
        if (lane == 0) R->status = DIC_ERR_INVALID_INPUT;
        return;
    }
    double* f = A.scratch + (int64_t)A.arrays * A.scr_off[s];
    double* gxa = f + n;
    double* gya = gxa + n;
    double* g = gya + n;
    double* ggx = g + n;  // FAGN only
    double* ggy = ggx + n;

    // ---- Subset::new: reference intensities and gradients ----
    double part = 0.0;
    for (int64_t i = lane; i < n; i += 64) {
        const int32_t o = offs[i];
        const double x = (double)unpack_dx(o) + xc, y = (double)unpack_dy(o) + yc;
        double v, gx, gy;
        eval_grad_f64(A.fC, A.wp, A.border, A.wmax, A.hmax, x, y, v, gx, gy);
        f[i] = v;
        gxa[i] = gx;
        gya[i] = gy;
        g[i] = 0.0;
        part += v;
    }
    const double f_m = wave_sum_u(part) / dn;
    part = 0.0;
    for (int64_t i = lane; i < n; i += 64) {
        const double d = f[i] - f_m;
        part += d * d;
    }
    const double sq = wave_sum_u(part);
    const double delta_f = sqrt(sq);
    part = 0.0;
    for (int64_t i = lane; i < n; i += 64) part += 0.5 * (gxa[i] * gxa[i] + gya[i] * gya[i]);
    const double sssig = wave_sum_u(part);
    const double sigma = sqrt(sq / dn);
    if (lane == 0) {
        R->f_m = f_m;
        R->delta_f = delta_f;
        R->sssig = sssig;
        R->sigma = sigma;
    }
    if (delta_f == 0.0) {
        if (lane == 0) R->status = DIC_ERR_INVALID_INPUT;
        return;
    }
    const double size = sqrt(dn);

    double p[12];
    const double* p0row = LIST ? A.p0_list + 12 * (int64_t)blockIdx.x : A.p0 + 12 * s;
#pragma unroll
    for (int k = 0; k < 12; ++k) p[k] = (k < M) ? p0row[k] : 0.0;

    if (A.method == DIC_METHOD_ICGN) {
        // SDI + Hessian, loop-invariant (subset.rs:1141-1143)
        double acc[NP];
#pragma unroll
        for (int q = 0; q < NP; ++q) acc[q] = 0.0;
        for (int64_t i = lane; i < n; i += 64) {
            const int32_t o = offs[i];
            const double x = (double)unpack_dx(o) + xc, y = (double)unpack_dy(o) + yc;
            double sd[M];
            sdi_row<M>(gxa[i], gya[i], x - xc, y - yc, sd);
            int q = 0;
#pragma unroll
            for (int a = 0; a < M; ++a)
#pragma unroll
                for (int b = a; b < M; ++b) acc[q++] += sd[a] * sd[b];
        }
        {
            int q = 0;
#pragma unroll
            for (int a = 0; a < M; ++a)
#pragma unroll
                for (int b = a; b < M; ++b) {
                    const double v = wave_sum_u(acc[q++]);
                    if (lane == 0) {
                        sH[a * M + b] = v;
                        sH[b * M + a] = v;
                    }
                }
        }
        __syncthreads();
        if constexpr (M == 12) {
            if (!inverse_ref_lanes<M>(sH, sHi, lane)) {
                if (lane == 0) R->status = DIC_ERR_SINGULAR_HESSIAN;
                return;
            }
        } else {
            if (lane == 0) sCtl[1] = inverse_ref<M>(sH, sHi) ? 0 : DIC_ERR_SINGULAR_HESSIAN;
            __syncthreads();
            if (sCtl[1] != 0) {
                if (lane == 0) R->status = DIC_ERR_SINGULAR_HESSIAN;
                return;
            }
        }
    }

    int it_done = 0;
    bool converged = false;
    for (int it = 1; it <= A.max_it; ++it) {
        double g_cx = xc + p[0], g_cy = yc + p[1];
        // warp + target intensities (and gradients for FAGN)
        part = 0.0;
        for (int64_t i = lane; i < n; i += 64) {
            const int32_t o = offs[i];
            const double x = (double)unpack_dx(o) + xc, y = (double)unpack_dy(o) + yc;
            double wx, wy;
            warp_point<M>(p, x, y, xc, yc, wx, wy);
            double v;
            if (A.method == DIC_METHOD_ICGN) {
                v = eval_f64(A.gC, A.wp, A.border, A.wmax, A.hmax, wx, wy);
            } else {
                double a, b;
                eval_grad_f64(A.gC, A.wp, A.border, A.wmax, A.hmax, wx, wy, v, a, b);
                ggx[i] = a;
                ggy[i] = b;
            }
            g[i] = v;
            part += v;
        }
        const double g_m = wave_sum_u(part) / dn;
        part = 0.0;
        for (int64_t i = lane; i < n; i += 64) {
            const double d = g[i] - g_m;
            part += d * d;
        }
        const double delta_g = sqrt(wave_sum_u(part));

        double gz[M];
#pragma unroll
        for (int j = 0; j < M; ++j) gz[j] = 0.0;
        double zacc = 0.0;
        if (A.method == DIC_METHOD_ICGN) {
            // delta_p_and_znssd_icgn (subset.rs:418-446)
            const double ratio = delta_f / delta_g;
            for (int64_t i = lane; i < n; i += 64) {
                const int32_t o = offs[i];
                const double x = (double)unpack_dx(o) + xc, y = (double)unpack_dy(o) + yc;
                const double a = f[i] - f_m;
                const double b = g[i] - g_m;
                const double z = a / delta_f - b / delta_g;
                zacc += z * z;
                const double bracket = a - ratio * b;
                double sd[M];
                sdi_row<M>(gxa[i], gya[i], x - xc, y - yc, sd);
#pragma unroll
                for (int j = 0; j < M; ++j) gz[j] += sd[j] * bracket;
            }
        } else {
            // Hessian at the target (subset.rs:1285-1286) + delta_p_fagn (452-473)
            double acc[NP];
#pragma unroll
            for (int q = 0; q < NP; ++q) acc[q] = 0.0;
            for (int64_t i = lane; i < n; i += 64) {
                const int32_t o = offs[i];
                const double x = (double)unpack_dx(o) + xc, y = (double)unpack_dy(o) + yc;
                double wx, wy;
                warp_point<M>(p, x, y, xc, yc, wx, wy);
                double sd[M];
                sdi_row<M>(ggx[i], ggy[i], wx - g_cx, wy - g_cy, sd);
                int q = 0;
#pragma unroll
                for (int a = 0; a < M; ++a)
#pragma unroll
                    for (int b = a; b < M; ++b) acc[q++] += sd[a] * sd[b];
                const double r = (f[i] - f_m) * (delta_g / delta_f) - (g[i] - g_m);
#pragma unroll
                for (int j = 0; j < M; ++j) gz[j] += sd[j] * r;
                const double z = (f[i] - f_m) / delta_f - (g[i] - g_m) / delta_g;
                zacc += z * z;
            }
            int q = 0;
#pragma unroll
            for (int a = 0; a < M; ++a)
#pragma unroll
                for (int b = a; b < M; ++b) {
                    const double v = wave_sum_u(acc[q++]);
                    if (lane == 0) {
                        sH[a * M + b] = v;
                        sH[b * M + a] = v;
                    }
                }
        }
#pragma unroll
        for (int j = 0; j < M; ++j) gz[j] = wave_sum_u(gz[j]);
        zacc = wave_sum_u(zacc);
        __syncthreads();
        // the update (subset.rs:1146-1160): order 2 over the lanes of the
        // wavefront; lower orders on lane 0 alone, their few operations
        // costing less than the barriers a lane-parallel form needs
        if constexpr (M == 12) {
            plain_update12(A, s, it, lane, gz, zacc, size, p, sH, sHi, sScr, sDp, sP, sRes, sCtl);
        } else {
            if (lane == 0) {
                int status = 0;
                double dp[12];
                if (A.method == DIC_METHOD_FAGN && !inverse_ref<M>(sH, sHi)) status = DIC_ERR_SINGULAR_HESSIAN;
                if (!status) {
                    for (int r = 0; r < M; ++r) {
                        double acc2 = gz[0] * sHi[r * M];
                        for (int k = 1; k < M; ++k) acc2 = acc2 + gz[k] * sHi[r * M + k];
                        dp[r] = (A.method == DIC_METHOD_ICGN) ? -acc2 : acc2;
                    }
                    double pn[12];
                    if (A.method == DIC_METHOD_ICGN) {
                        if (!compose_icgn_dev<M>(p, dp, pn, sScr)) status = DIC_ERR_SINGULAR_WARP;
                    } else {
                        for (int k = 0; k < M; ++k) pn[k] = p[k] + dp[k];
                    }
                    if (!status) {
                        for (int k = 0; k < M; ++k) sP[k] = pn[k];
                        const double norm = convergence_norm_dev<M>(dp, size);
                        const double c_zncc = 1.0 - zacc / 2.0;
                        if (A.hist && it <= A.hist_cap) {
                            double* h = A.hist + ((int64_t)s * A.hist_cap + (it - 1)) * 4;
                            h[0] = (double)it;
                            h[1] = norm;
                            h[2] = c_zncc;
                            h[3] = zacc;
                        }
                        sRes[0] = c_zncc;
                        sRes[1] = zacc;
                        sRes[2] = norm;
                        sCtl[3] = norm < A.max_norm ? 1 : 0;
                    }
                }
                sCtl[1] = status;
            }
        }
        __syncthreads();
        if (sCtl[1] != 0) {
            if (lane == 0) {
                R->status = sCtl[1];
                R->iterations = it_done;
            }
            return;
        }
#pragma unroll
        for (int k = 0; k < M; ++k) p[k] = sP[k];
        it_done = it;
        if (sCtl[3]) {
            converged = true;
            break;
        }
        __syncthreads();
    }

    // ---- results (subset.rs:1190-1218) ----
    const double c_zncc = it_done ? sRes[0] : 0.0;
    const double c_znssd = it_done ? sRes[1] : 4.0;
    // warp-adequacy diagnostic on the final g (subset.rs:1201-1210)
    part = 0.0;
    for (int64_t i = lane; i < n; i += 64) part += g[i];
    const double g_mf = wave_sum_u(part) / dn;
    part = 0.0;
    for (int64_t i = lane; i < n; i += 64) {
        const double d = g[i] - g_mf;
        part += d * d;
    }
    const double delta_gf = sqrt(wave_sum_u(part));
    double se = 0.0, sx = 0.0, sy = 0.0;
    for (int64_t i = lane; i < n; i += 64) {
        se += (f[i] - f_m) / delta_f - (g[i] - g_mf) / delta_gf;
        sx += gxa[i];
        sy += gya[i];
    }
    const double e_bar = wave_sum_u(se) / dn;
    const double mx = wave_sum_u(sx) / dn, my = wave_sum_u(sy) / dn;
    double vx = 0.0, vy = 0.0;
    for (int64_t i = lane; i < n; i += 64) {
        const double a = gxa[i] - mx, b = gya[i] - my;
        vx += a * a;
        vy += b * b;
    }
    const double sdx = sqrt(fmax(wave_sum_u(vx) / dn, 1e-18)), sdy = sqrt(fmax(wave_sum_u(vy) / dn, 1e-18));
    double rsum[2 * K], rvar[2 * K], rproj[2 * K];
#pragma unroll
    for (int k = 0; k < 2 * K; ++k) rsum[k] = rvar[k] = rproj[k] = 0.0;
    for (int64_t i = lane; i < n; i += 64) {
        const int32_t o = offs[i];
        const double x = (double)unpack_dx(o) + xc, y = (double)unpack_dy(o) + yc;
        double mo[K];
        mode_row<M>(x - xc, y - yc, mo);
        const double zx = (gxa[i] - mx) / sdx, zy = (gya[i] - my) / sdy;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            rsum[k] += zx * mo[k];
            rsum[K + k] += zy * mo[k];
        }
    }
#pragma unroll
    for (int k = 0; k < 2 * K; ++k) rsum[k] = wave_sum_u(rsum[k]) / dn;
    for (int64_t i = lane; i < n; i += 64) {
        const int32_t o = offs[i];
        const double x = (double)unpack_dx(o) + xc, y = (double)unpack_dy(o) + yc;
        double mo[K];
        mode_row<M>(x - xc, y - yc, mo);
        const double zx = (gxa[i] - mx) / sdx, zy = (gya[i] - my) / sdy;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const double a = zx * mo[k] - rsum[k], b = zy * mo[k] - rsum[K + k];
            rvar[k] += a * a;
            rvar[K + k] += b * b;
        }
    }
#pragma unroll
    for (int k = 0; k < 2 * K; ++k) rvar[k] = sqrt(fmax(wave_sum_u(rvar[k]) / dn, 1e-18));
    for (int64_t i = lane; i < n; i += 64) {
        const int32_t o = offs[i];
        const double x = (double)unpack_dx(o) + xc, y = (double)unpack_dy(o) + yc;
        double mo[K];
        mode_row<M>(x - xc, y - yc, mo);
        const double zx = (gxa[i] - mx) / sdx, zy = (gya[i] - my) / sdy;
        const double ec = ((f[i] - f_m) / delta_f - (g[i] - g_mf) / delta_gf) - e_bar;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            rproj[k] += ((zx * mo[k] - rsum[k]) / rvar[k]) * ec;
            rproj[K + k] += ((zy * mo[k] - rsum[K + k]) / rvar[K + k]) * ec;
        }
    }
    double eu = 0.0, ev = 0.0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double a = wave_sum_u(rproj[k]) / dn, b = wave_sum_u(rproj[K + k]) / dn;
        eu += a * a;
        ev += b * b;
    }
    if (lane == 0) {
        for (int k = 0; k < 12; ++k) R->p[k] = (k < M) ? p[k] : 0.0;
        for (int k = 0; k < 12; ++k) A.p_out[12 * s + k] = (k < M) ? p[k] : 0.0;
        R->c_zncc = c_zncc;
        R->c_znssd = c_znssd;
        R->norm = it_done ? sRes[2] : __longlong_as_double(0x7ff8000000000000LL);
        R->iterations = it_done;
        R->converged = converged ? 1 : 0;
        R->quality_ok = (converged && c_zncc >= A.tol) ? 1 : 0;
        R->eta_u = sqrt(eu);
        R->eta_v = sqrt(ev);
        R->status = 0;
    }
}

}  // namespace

std::vector<OrderRun> order_runs(const Batch* b, const int32_t* h, int64_t count) {
    std::vector<OrderRun> runs;
    for (int64_t k = 0; k < count;) {
        const int o = b->h_orders[h[k]];
        int64_t j = k;
        while (j < count && b->h_orders[h[j]] == o) ++j;
        runs.push_back({k, j - k, o});
        k = j;
    }
    return runs;
}

int launch_icgn_plain(Batch* b, const Image* f, const Image* g, const dic_solve_params* prm, const int32_t* list,
                      const int32_t* h_list, int64_t count, const double* p0_list, dic_subset_result* res_list,
                      double* ext_scratch, size_t ext_doubles) {
    Engine* e = b->eng;
    const int arrays = plain_arrays(prm->method);
    const size_t need = (size_t)b->total_px * arrays;
    double* scratch = ext_scratch;
    if (scratch) {
        if (need > ext_doubles) {
            set_error("invalid input: the batch's per-pixel arrays exceed the scratch");
            return DIC_ERR_INVALID_INPUT;
        }
    } else {
        if (b->scratch_doubles < need) {
            if (b->d_scratch) dev_free(e, b->d_scratch, b->scratch_doubles * sizeof(double));
            b->d_scratch = nullptr;
            b->scratch_doubles = 0;
            int st = dev_alloc(e, (void**)&b->d_scratch, need * sizeof(double));
            if (st) return st;
            b->scratch_doubles = need;
        }
        scratch = b->d_scratch;
    }
    PlainArgs A;
    A.fC = f->coef64;
    A.gC = g->coef64;
    A.wp = f->wp;
    A.border = f->border;
    A.wmax = (double)f->w - 1e-9;
    A.hmax = (double)f->h - 1e-9;
    A.n_sub = list ? count : b->n;
    A.list = list;
    A.coords = b->d_coords;
    A.orders = b->d_orders;
    A.tpl_index = b->d_tpl_index;
    A.tpl_start = b->d_tpl_start;
    A.tpl_xy = b->d_tpl_xy;
    A.scr_off = b->d_scr_off;
    A.scratch = scratch;
    A.arrays = arrays;
    A.p0 = b->d_p0;
    A.p0_list = list ? p0_list : nullptr;
    A.res_list = list ? res_list : nullptr;
    A.p_out = b->d_p;
    A.res = b->d_res;
    A.hist = prm->want_history ? b->d_hist : nullptr;
    A.hist_cap = prm->want_history ? b->hist_cap : 0;
    A.method = prm->method;
    A.max_it = prm->max_iterations;
    A.tol = prm->tolerance;
    A.max_norm = prm->max_norm;
    if (A.n_sub == 0) return DIC_OK;
    // one launch per warp order: the whole batch in one launch when it has
    // one order, else its subsets grouped by order; a list by its runs
    auto launch = [&](int order, const PlainArgs& Ao) {
        const dim3 grid((unsigned)Ao.n_sub), block(64);
        const bool lst = Ao.res_list != nullptr;
        if (order == 0)
            lst ? icgn_plain_kernel<2, true><<<grid, block, 0, e->stream>>>(Ao)
                : icgn_plain_kernel<2, false><<<grid, block, 0, e->stream>>>(Ao);
        else if (order == 1)
            lst ? icgn_plain_kernel<6, true><<<grid, block, 0, e->stream>>>(Ao)
                : icgn_plain_kernel<6, false><<<grid, block, 0, e->stream>>>(Ao);
        else
            lst ? icgn_plain_kernel<12, true><<<grid, block, 0, e->stream>>>(Ao)
                : icgn_plain_kernel<12, false><<<grid, block, 0, e->stream>>>(Ao);
        return hipGetLastError();
    };
    if (!b->mixed_orders) {
        DIC_HIP_TRY(launch(b->max_order, A));
    } else if (!list) {
        for (int o = 0; o < 3; ++o) {
            PlainArgs Ao = A;
            Ao.list = b->d_by_order + b->by_order_off[o];
            Ao.n_sub = b->by_order_off[o + 1] - b->by_order_off[o];
            if (Ao.n_sub) DIC_HIP_TRY(launch(o, Ao));
        }
    } else {
        if (!h_list) {
            set_error("launch_icgn_plain: a list over mixed warp orders needs its host copy");
            return DIC_ERR_INVALID_INPUT;
        }
        for (const OrderRun& r : order_runs(b, h_list, count)) {
            PlainArgs Ao = A;
            Ao.list = list + r.start;
            Ao.n_sub = r.count;
            if (Ao.p0_list) Ao.p0_list = p0_list + 12 * r.start;
            if (Ao.res_list) Ao.res_list = res_list + r.start;
            DIC_HIP_TRY(launch(r.order, Ao));
        }
    }
    return DIC_OK;
}

// ---------------------------------------------------------------------------
// Subset::residual_map (subset.rs:1350-1392, deformable texture): the
// per-pixel zero-normalised residual e_i = (f_i - f_m)/Δf - (g_i - g_m)/Δg
// at warp p, FP64 with the reference's association, one 64-lane wavefront
// per listed subset.  The reference intensities pass through the output
// row; g is evaluated once per use (sums, spread, output) rather than kept.
// ---------------------------------------------------------------------------
namespace {

struct ResidualArgs {
    const double* fC;
    const double* gC;
    int64_t wp;
    int border;
    double wmax, hmax;
    int64_t count;
    const int32_t* list;
    const double* coords;
    const int32_t* tpl_index;
    const int64_t* tpl_start;
    const int32_t* tpl_xy;
    const double* p;        // warps: 12 per listed subset, or per subset (p_by_subset)
    int p_by_subset;
    const int64_t* out_off;  // start of each listed subset's row in out
    double* out;
};

template <int M>
__global__ void __launch_bounds__(64) residual_kernel(ResidualArgs A) {
#pragma clang fp contract(off)
    const int64_t k = blockIdx.x;
    if (k >= A.count) return;
    const int64_t s = A.list[k];
    const int lane = threadIdx.x;
    const int t = A.tpl_index[s];
    const int64_t st = A.tpl_start[t];
    const int64_t n = A.tpl_start[t + 1] - st;
    const int32_t* offs = A.tpl_xy + st;
    const double xc = A.coords[2 * s], yc = A.coords[2 * s + 1];
    const double dn = (double)n;
    double* e = A.out + A.out_off[k];
    const double* pr = A.p + 12 * (A.p_by_subset ? s : k);
    double p[12];
#pragma unroll
    for (int q = 0; q < 12; ++q) p[q] = (q < M) ? pr[q] : 0.0;
    auto pos = [&](int64_t i, double& x, double& y) {
        const int32_t o = offs[i];
        x = (double)unpack_dx(o) + xc;
        y = (double)unpack_dy(o) + yc;
    };
    auto g_at = [&](int64_t i) {
        double x, y, wx, wy;
        pos(i, x, y);
        warp_point<M>(p, x, y, xc, yc, wx, wy);
        return eval_f64(A.gC, A.wp, A.border, A.wmax, A.hmax, wx, wy);
    };
    // f, f_m and Δf (subset.rs:877-886)
    double part = 0.0;
    for (int64_t i = lane; i < n; i += 64) {
        double x, y;
        pos(i, x, y);
        const double v = eval_f64(A.fC, A.wp, A.border, A.wmax, A.hmax, x, y);
        e[i] = v;
        part += v;
    }
    const double f_m = wave_sum_u(part) / dn;
    part = 0.0;
    for (int64_t i = lane; i < n; i += 64) {
        const double d = e[i] - f_m;
        part += d * d;
    }
    const double delta_f = sqrt(wave_sum_u(part));
    // g at the warped positions, g_m and Δg (subset.rs:1385-1388)
    part = 0.0;
    for (int64_t i = lane; i < n; i += 64) part += g_at(i);
    const double g_m = wave_sum_u(part) / dn;
    part = 0.0;
    for (int64_t i = lane; i < n; i += 64) {
        const double d = g_at(i) - g_m;
        part += d * d;
    }
    const double delta_g = sqrt(wave_sum_u(part));
    for (int64_t i = lane; i < n; i += 64) e[i] = (e[i] - f_m) / delta_f - (g_at(i) - g_m) / delta_g;
}

}  // namespace

int launch_residual(Batch* b, const Image* f, const Image* g, int order, const int32_t* d_list, int64_t count,
                    const double* d_p, bool p_by_subset, const int64_t* d_off, double* d_out) {
    Engine* e = b->eng;
    ResidualArgs A;
    A.fC = f->coef64;
    A.gC = g->coef64;
    A.wp = f->wp;
    A.border = f->border;
    A.wmax = (double)f->w - 1e-9;
    A.hmax = (double)f->h - 1e-9;
    A.count = count;
    A.list = d_list;
    A.coords = b->d_coords;
    A.tpl_index = b->d_tpl_index;
    A.tpl_start = b->d_tpl_start;
    A.tpl_xy = b->d_tpl_xy;
    A.p = d_p;
    A.p_by_subset = p_by_subset ? 1 : 0;
    A.out_off = d_off;
    A.out = d_out;
    if (count == 0) return DIC_OK;
    const dim3 grid((unsigned)count), block(64);
    if (order == 0)
        residual_kernel<2><<<grid, block, 0, e->stream>>>(A);
    else if (order == 1)
        residual_kernel<6><<<grid, block, 0, e->stream>>>(A);
    else
        residual_kernel<12><<<grid, block, 0, e->stream>>>(A);
    DIC_HIP_TRY(hipGetLastError());
    return DIC_OK;
}

// ---------------------------------------------------------------------------
// Fixed scratch of dic_batch_solve_noalloc: a file-scope device array that
// every GPU the library runs on holds for the life of the process.  In host
// code `scratch` names a host-side stand-in, not a device address, so host
// code never passes it to a kernel or a copy: the device address of each
// GPU's own copy is looked up once per device (noalloc_scratch_init, called
// by dic_engine_create with the device current) and used from then on.
// ---------------------------------------------------------------------------
namespace {
__device__ __align__(256) double scratch[kNoallocScratchBytes / sizeof(double)];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

int noalloc_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) {
        set_error("device index beyond the library's device table");
        return DIC_ERR_DEVICE;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return DIC_OK;
    void* ptr = nullptr;
    DIC_HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) {
        set_error("hipGetSymbolAddress(scratch) returned a null address");
        return DIC_ERR_DEVICE;
    }
    g_scratch_ptr[device] = static_cast<double*>(ptr);
    return DIC_OK;
}

double* noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[device];
}

}  // namespace geodic
