// Runs the batched-engine kernel (icgn_batch_kernel.h) through the CPU SIMT
// emulation on a synthetic speckle pair and compares every subset with the
// FP64 oracle.  Built with AddressSanitizer/UBSan by tests/simt/Makefile.
//
//   simt_driver ORDER RADIUS NSIDE MAXIT [WIN_SLACK] [WPS]
// exit 0 when every subset matches (status, convergence, |du|,|dv| < 1e-4 px,
// ZNCC within 1e-6, η within 1 % + 1e-6), 1 otherwise.
#include "simt_hip.h"

#include <cstdio>
#include <cstdlib>
#include <type_traits>

#include "../../geopyv-dev_amd/csrc/icgn_batch_kernel.h"

extern "C" {
void orc_get_c(const double* img, int rows, int cols, int border, double* out, int threads);
int orc_mask_local(int shape, int size, double* coords, int32_t* subset_mask);
struct OrcResult {
    double p[12];
    double c_zncc, c_znssd, norm, eta_u, eta_v, f_m, delta_f, sssig, sigma;
    int32_t iterations, converged, quality_ok, status, n_px, order;
};
int orc_batch_solve(const double* fcoef, const double* gcoef, int h, int w, int border, int64_t n_sub,
                    const double* coords, const int32_t* tpl_index, const double* tpl_xy, const int64_t* tpl_start,
                    const int32_t* orders, const double* p0, int method, double tolerance, double max_norm,
                    int max_iterations, OrcResult* out, int threads);
}

static double speckle(double x, double y, const std::vector<double>& cx, const std::vector<double>& cy) {
    double s = 0.0;
    for (size_t k = 0; k < cx.size(); ++k) {
        const double dx = x - cx[k], dy = y - cy[k];
        s += std::exp(-(dx * dx + dy * dy) / 8.0);
    }
    return 200.0 * s;
}

int main(int argc, char** argv) {
    const int order = argc > 1 ? atoi(argv[1]) : 1;
    const int radius = argc > 2 ? atoi(argv[2]) : 10;
    const int nside = argc > 3 ? atoi(argv[3]) : 3;
    const int max_it = argc > 4 ? atoi(argv[4]) : 50;
    const int slack = argc > 5 ? atoi(argv[5]) : 3;
    const int wps = argc > 6 ? atoi(argv[6]) : 1;
    // list mode: warps read from and results written to list buffers in
    // dispatch order; at eight wavefronts the four-pixel list kernel
    const int list = argc > 7 ? atoi(argv[7]) : 0;
    const int H = 96, W = 96, B = 20;
    const int HP = H + 2 * B, WP = W + 2 * B;
    uint64_t st = 12345;
    auto rnd = [&]() {
        st = st * 6364136223846793005ULL + 1442695040888963407ULL;
        return (double)(st >> 11) / 9007199254740992.0;
    };
    std::vector<double> cx(300), cy(300);
    for (int k = 0; k < 300; ++k) {
        cx[k] = rnd() * W;
        cy[k] = rnd() * H;
    }
    std::vector<double> f((size_t)H * W), g((size_t)H * W);
    const double u = 0.37, v = -0.52, ux = 0.004, vy = -0.003;
    for (int i = 0; i < H; ++i)
        for (int j = 0; j < W; ++j) {
            f[(size_t)i * W + j] = speckle(j, i, cx, cy);
            // target(x) = f(X) where x = X + u + ux (X - 48), y = Y + v + vy (Y - 48)
            const double X = (j - u + 48.0 * ux) / (1.0 + ux);
            const double Y = (i - v + 48.0 * vy) / (1.0 + vy);
            g[(size_t)i * W + j] = speckle(X, Y, cx, cy);
        }
    std::vector<double> Cf((size_t)HP * WP), Cg((size_t)HP * WP);
    orc_get_c(f.data(), H, W, B, Cf.data(), 1);
    orc_get_c(g.data(), H, W, B, Cg.data(), 1);
    // FP32 copies with rows pitched to 4 floats, as the engine stores them
    const int P32 = (WP + 3) & ~3;
    std::vector<float> Cf32((size_t)HP * P32, 0.f), Cg32((size_t)HP * P32, 0.f);
    for (int i = 0; i < HP; ++i)
        for (int j = 0; j < WP; ++j) {
            Cf32[(size_t)i * P32 + j] = (float)Cf[(size_t)i * WP + j];
            Cg32[(size_t)i * P32 + j] = (float)Cg[(size_t)i * WP + j];
        }

    // template and subsets
    const int side = 2 * radius + 1;
    std::vector<double> tc((size_t)side * side * 2);
    const int npx = orc_mask_local(0, radius, tc.data(), nullptr);
    std::vector<int32_t> packed(npx);
    int32_t ext[4] = {0, 0, 0, 0};
    for (int i = 0; i < npx; ++i) {
        const int dx = (int)tc[2 * i], dy = (int)tc[2 * i + 1];
        packed[i] = (int32_t)(((uint32_t)(uint16_t)dx) | ((uint32_t)(uint16_t)dy << 16));
        if (i == 0) {
            ext[0] = ext[1] = dx;
            ext[2] = ext[3] = dy;
        }
        ext[0] = std::min(ext[0], dx);
        ext[1] = std::max(ext[1], dx);
        ext[2] = std::min(ext[2], dy);
        ext[3] = std::max(ext[3], dy);
    }
    const int n = nside * nside;
    std::vector<double> coords(2 * n), p0(12 * n, 0.0), pout(12 * n, 0.0);
    for (int a = 0; a < nside; ++a)
        for (int b = 0; b < nside; ++b) {
            const int s = a * nside + b;
            coords[2 * s] = 30.0 + 18.0 * b + (s % 2 ? 0.25 : 0.0);
            coords[2 * s + 1] = 30.0 + 18.0 * a;
        }
    std::vector<int32_t> tpl_index(n, 0), perm(n), orders(n, order);
    for (int s = 0; s < n; ++s) perm[s] = s;
    std::vector<int64_t> tpl_start = {0, npx};
    std::vector<dic_subset_result> res(n);
    const int M = order == 0 ? 2 : (order == 1 ? 6 : 12);
    std::vector<double> hist((size_t)n * max_it * 4 + 4, 0.0);

    geodic::batch::BatchArgs A{};
    A.fC = Cf32.data();
    A.gC = Cg32.data();
    A.pitch = P32;
    A.border = B;
    A.w = W;
    A.h = H;
    A.wmax = W - 1e-9;
    A.hmax = H - 1e-9;
    A.n_sub = n;
    std::vector<geodic::batch::SubMeta> meta(n);
    for (int s = 0; s < n; ++s) {
        meta[s].xc = coords[2 * s];
        meta[s].yc = coords[2 * s + 1];
        meta[s].st = 0;
        meta[s].n = npx;
        meta[s].t = 0;
        for (int q = 0; q < 4; ++q) meta[s].ext[q] = ext[q];
    }
    A.meta = meta.data();
    A.tpl_xy = packed.data();
    A.perm = perm.data();
    A.p0 = p0.data();
    A.p_out = pout.data();
    A.res = res.data();
    std::vector<double> p0l(p0);  // dispatch order = subset order here
    std::vector<dic_subset_result> resl(n);
    std::vector<int32_t> done(n, 0);  // list mode: each workgroup's flag, raised after its record
    if (list) {
        A.p0_list = p0l.data();
        A.res_list = resl.data();
        A.done_list = done.data();
    }
    // SIMT_KNOT_TABLE=1: the reference's knot table (f, ∂x f, ∂y f at every
    // pixel by knot_grad on its 5x5 block, as the engine's table kernel
    // makes it), so subsets centred on a whole pixel read it in setup
    std::vector<float4> knot;
    const char* kt = std::getenv("SIMT_KNOT_TABLE");
    if (kt && kt[0] == '1') {
        knot.resize((size_t)H * W);
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                float c[25];
                for (int r = 0; r < 5; ++r)
                    for (int k = 0; k < 5; ++k) c[5 * r + k] = Cf32[(size_t)(y + B - 2 + r) * P32 + (x + B - 2 + k)];
                float gx, gy;
                const float val = geodic::batch::knot_grad(c, gx, gy);
                knot[(size_t)y * W + x] = make_float4(val, gx, gy, 0.f);
            }
        A.knot = knot.data();
    }
    A.hist = hist.data();
    A.hist_cap = max_it;
    A.max_it = max_it;
    A.tol = 0.75;
    A.max_norm = 1e-5;
    A.max_norm2 = A.max_norm * A.max_norm;
    A.diag_norm2 = 36.0 * A.max_norm2;
    const int wside = (ext[1] - ext[0] + 1) + 2 * slack + 8;
    const int wrow = (wside + 3 + 3) & ~3;
    const bool s72 = wrow <= 72;
    A.win_cap = (s72 ? 72 : wrow) * wside;
    A.win_slack = slack;
    const size_t lds = sizeof(float) * A.win_cap + 4 * (size_t)npx;
    // P = 12 pixels per thread: one wavefront holds 768 pixels
    auto run = [&](auto mtag) {
        constexpr int MM = decltype(mtag)::value;
        auto go = [&](auto stag) {
            constexpr int SS = decltype(stag)::value;
            if (wps == 1)
                simt::launch(n, 64, lds, [&] { geodic::batch::icgn_batch_kernel<MM, 12, 1, SS>(A); });
            else if (wps == 2)
                simt::launch(n, 128, lds, [&] { geodic::batch::icgn_batch_kernel<MM, 12, 2, SS>(A); });
            else if (wps == 4)
                simt::launch(n, 256, lds, [&] { geodic::batch::icgn_batch_kernel<MM, 12, 4, SS>(A); });
            else if (wps == 8 && list && npx <= 4 * 512)
                simt::launch(n, 512, lds, [&] { geodic::batch::icgn_batch_kernel<MM, 4, 8, SS, true>(A); });
            else
                std::abort();
        };
        if (s72)
            go(std::integral_constant<int, 72>{});
        else
            go(std::integral_constant<int, 0>{});
    };
    if (M == 2)
        run(std::integral_constant<int, 2>{});
    else if (M == 6)
        run(std::integral_constant<int, 6>{});
    else
        run(std::integral_constant<int, 12>{});

    std::vector<double> xy(2 * npx);
    for (int i = 0; i < 2 * npx; ++i) xy[i] = tc[i];
    std::vector<OrcResult> orc(n);
    orc_batch_solve(Cf.data(), Cg.data(), H, W, B, n, coords.data(), tpl_index.data(), xy.data(), tpl_start.data(),
                    orders.data(), p0.data(), 0, 0.75, 1e-5, max_it, orc.data(), 1);
    if (list) res = resl;
    int bad = 0;
    if (list)
        for (int s = 0; s < n; ++s)
            if (done[s] != 1) {
                printf("s=%d: done flag %d\n", s, done[s]);
                ++bad;
            }
    double dumax = 0.0, dsmax = 0.0;
    for (int s = 0; s < n; ++s) {
        const double du = std::max(std::fabs(res[s].p[0] - orc[s].p[0]), std::fabs(res[s].p[1] - orc[s].p[1]));
        double ds = 0.0;
        for (int k = 2; k < M; ++k) ds = std::max(ds, std::fabs(res[s].p[k] - orc[s].p[k]));
        dumax = std::max(dumax, du);
        dsmax = std::max(dsmax, ds);
        const double deta = std::max(std::fabs(res[s].eta_u - orc[s].eta_u), std::fabs(res[s].eta_v - orc[s].eta_v));
        const double eta_tol = 1e-6 + 1e-2 * std::max(std::fabs(orc[s].eta_u), std::fabs(orc[s].eta_v));
        const bool ok = res[s].status == orc[s].status && du < 1e-4 && res[s].converged == orc[s].converged &&
                        std::fabs(res[s].c_zncc - orc[s].c_zncc) < 1e-6 && deta < eta_tol;
        if (!ok) ++bad;
        printf("s=%d st=%d/%d it=%d/%d conv=%d/%d u=%.8f/%.8f v=%.8f/%.8f zncc=%.9f/%.9f eta=%.3e/%.3e\n", s,
               res[s].status, orc[s].status, res[s].iterations, orc[s].iterations, res[s].converged, orc[s].converged,
               res[s].p[0], orc[s].p[0], res[s].p[1], orc[s].p[1], res[s].c_zncc, orc[s].c_zncc, res[s].eta_u,
               orc[s].eta_u);
    }
    printf("order %d radius %d subsets %d: max|du| %.3e max|dstrain| %.3e mismatched %d\n", order, radius, n,
           dumax, dsmax, bad);
    if (argc > 8) {
        // the raw result records, for bit-for-bit comparisons between builds
        FILE* fo = fopen(argv[8], "wb");
        if (!fo || fwrite(res.data(), sizeof(dic_subset_result), res.size(), fo) != res.size()) return 2;
        fclose(fo);
    }
    return bad ? 1 : 0;
}
