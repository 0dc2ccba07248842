// Batched engine launch (the kernel is in icgn_batch_kernel.h).
#include "internal.h"
#include "icgn_batch_kernel.h"

namespace geodic {

namespace {

using batch::BatchArgs;
using batch::icgn_batch_kernel;
using batch::kRedSlots;

// P < 12 (a few pixels per thread) serves lists only.
template <int M, int P, int WPS, int S>
int launch_mpw(const BatchArgs& A, int blocks, size_t lds, hipStream_t stream) {
    if constexpr (P < 12) {
        const void* k = (const void*)icgn_batch_kernel<M, P, WPS, S, true>;
        DIC_HIP_TRY(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        icgn_batch_kernel<M, P, WPS, S, true><<<blocks, 64 * WPS, lds, stream>>>(A);
    } else {
        const void* k = A.res_list ? (const void*)icgn_batch_kernel<M, P, WPS, S, true>
                                   : (const void*)icgn_batch_kernel<M, P, WPS, S, false>;
        DIC_HIP_TRY(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        if (A.res_list)
            icgn_batch_kernel<M, P, WPS, S, true><<<blocks, 64 * WPS, lds, stream>>>(A);
        else
            icgn_batch_kernel<M, P, WPS, S, false><<<blocks, 64 * WPS, lds, stream>>>(A);
    }
    DIC_HIP_TRY(hipGetLastError());
    return DIC_OK;
}

// LDS row stride: 72 floats when every window of the class fits it (rows
// then sit at immediate offsets, and 72 = 8 mod 32 spreads an 8x4 block of
// pixels, one half-wavefront, over all 32 banks), else the window's own
// width.
template <int M, int WPS>
int launch_ms(const BatchArgs& A, int blocks, size_t lds, bool s72, hipStream_t stream) {
    return s72 ? launch_mpw<M, 12, WPS, 72>(A, blocks, lds, stream)
               : launch_mpw<M, 12, WPS, 0>(A, blocks, lds, stream);
}

template <int M>
int launch_m(const BatchArgs& A, int wps, int blocks, size_t lds, bool s72, hipStream_t stream, bool few) {
    // a list spread over eight wavefronts per subset with at most four
    // pixels per thread: the unrolled instantiation
    if (few && wps == 8)
        return s72 ? launch_mpw<M, 4, 8, 72>(A, blocks, lds, stream) : launch_mpw<M, 4, 8, 0>(A, blocks, lds, stream);
    switch (wps) {
        case 1: return launch_ms<M, 1>(A, blocks, lds, s72, stream);
        case 2: return launch_ms<M, 2>(A, blocks, lds, s72, stream);
        case 4: return launch_ms<M, 4>(A, blocks, lds, s72, stream);
        case 8: return launch_ms<M, 8>(A, blocks, lds, s72, stream);
        default: break;
    }
    set_error("invalid input: waves per subset must be 1, 2, 4 or 8");
    return DIC_ERR_INVALID_INPUT;
}

// Fewest wavefronts that hold n pixels at 12 per thread (the footprint that
// stays in registers without spilling).
int min_wps(int64_t n) { return n <= 64 * 12 ? 1 : (n <= 128 * 12 ? 2 : (n <= 256 * 12 ? 4 : 8)); }

}  // namespace

// Size classes of a batch in dispatch order (the order sorts by warp order,
// then size, so each class is a contiguous run): subsets of one warp order
// that need the same number of wavefronts launch together; an explicit
// schedule makes one class per warp order.
std::vector<SizeClass> size_classes(const Batch* b, const int32_t* h_perm, int64_t n) {
    if (!h_perm) {
        h_perm = b->h_perm.data();
        n = b->n;
    }
    std::vector<SizeClass> out;
    auto npx = [&](int64_t s) {
        const int t = b->h_tpl_index[s];
        return b->h_tpl_start[t + 1] - b->h_tpl_start[t];
    };
    auto side = [&](int64_t s) {
        const int32_t* e = &b->h_ext[4 * (size_t)b->h_tpl_index[s]];
        return std::max(e[1] - e[0] + 1, e[3] - e[2] + 1);
    };
    for (int64_t i = 0; i < n;) {
        const int w = b->waves_per_subset > 0 ? b->waves_per_subset : min_wps(npx(h_perm[i]));
        const int o = b->h_orders[h_perm[i]];
        SizeClass c;
        c.start = i;
        c.order = o;
        c.base_wps = w;
        c.wps = w;
        c.slack = b->win_slack;
        if (b->waves_per_subset == 0 && b->tuned_wps[o][w] > 0) {
            c.wps = b->tuned_wps[o][w];
            c.slack = b->tuned_slack[o][w];
        }
        c.max_n = 0;
        c.max_side = 0;
        while (i < n && b->h_orders[h_perm[i]] == o &&
               (b->waves_per_subset > 0 || min_wps(npx(h_perm[i])) == w)) {
            c.max_n = std::max<int64_t>(c.max_n, npx(h_perm[i]));
            c.max_side = std::max(c.max_side, side(h_perm[i]));
            ++i;
        }
        c.count = i - c.start;
        out.push_back(c);
    }
    return out;
}

namespace {

// f, ∂x f and ∂y f at every pixel (knot_grad on the 5x5 block of the pixel's
// cell, the batched engine's own setup arithmetic, so a solve reads the
// same values it would compute); 64 x 4 pixels per workgroup
__global__ void __launch_bounds__(256) knot_table_kernel(const float* __restrict__ C, int64_t pitch, int border, int w,
                                                         int h, float4* __restrict__ out) {
    const int x = (int)blockIdx.x * 64 + (int)(threadIdx.x & 63);
    const int y = (int)blockIdx.y * 4 + (int)(threadIdx.x >> 6);
    if (x >= w || y >= h) return;
    const float* __restrict__ blk = C + (int64_t)(y + border - 2) * pitch + (x + border - 2);
    float c[25];
#pragma unroll
    for (int r = 0; r < 5; ++r)
#pragma unroll
        for (int k = 0; k < 5; ++k) c[5 * r + k] = blk[r * pitch + k];
    float gx, gy;
    const float v = batch::knot_grad(c, gx, gy);
    out[(int64_t)y * w + x] = make_float4(v, gx, gy, 0.f);
}

BatchArgs common_args(Batch* b, const Image* f, const Image* g, const dic_solve_params* prm) {
    BatchArgs A;
    A.fC = f->coef32;
    A.gC = g->coef32;
    A.pitch = f->wp32;
    A.border = f->border;
    A.w = (int)f->w;
    A.h = (int)f->h;
    A.wmax = (double)f->w - 1e-9;
    A.hmax = (double)f->h - 1e-9;
    A.meta = reinterpret_cast<const batch::SubMeta*>(b->d_meta);
    A.tpl_xy = b->d_tpl_xy_eng;
    A.p0 = b->d_p0;
    A.p_out = b->d_p;
    A.res = b->d_res;
    A.p0_list = nullptr;
    A.res_list = nullptr;
    A.done_list = nullptr;
    A.knot = f->knot_valid ? f->knot : nullptr;
    A.hist = prm->want_history ? b->d_hist : nullptr;
    A.hist_cap = prm->want_history ? b->hist_cap : 0;
    A.max_it = prm->max_iterations;
    A.tol = prm->tolerance;
    A.max_norm = prm->max_norm;
    A.max_norm2 = prm->max_norm * prm->max_norm;
    A.diag_norm2 = 36.0 * A.max_norm2;  // (6 max_norm)²: the last-iteration prediction's threshold
    return A;
}

int launch_class(Batch* b, const BatchArgs& A, const SizeClass& c, hipStream_t st, const int32_t* d_perm = nullptr) {
    if (c.max_n > 64 * 12 * c.wps) {
        set_error("invalid input: subset larger than the wavefronts per subset can hold");
        return DIC_ERR_INVALID_INPUT;
    }
    BatchArgs Ac = A;
    Ac.perm = (d_perm ? d_perm : b->d_order_perm) + c.start;
    if (A.p0_list) Ac.p0_list = A.p0_list + 12 * c.start;
    if (A.res_list) Ac.res_list = A.res_list + c.start;
    if (A.done_list) Ac.done_list = A.done_list + c.start;
    Ac.n_sub = c.count;
    Ac.win_slack = c.slack;
    // LDS: the coefficient window (template box, slack, taps, cell rounding
    // and a little strain), then the template offsets as FP16 pairs
    const int side = c.max_side + 2 * c.slack + 8;
    const int row = (side + 3 + 3) & ~3;  // plus the 16-byte rounding of the first column
    const bool s72 = row <= 72;
    Ac.win_cap = (s72 ? 72 : row) * side;
    const size_t lds = sizeof(float) * (size_t)Ac.win_cap + 4 * (size_t)c.max_n;
    const bool few = A.res_list && c.max_n <= 64 * 4 * c.wps;
    if (c.order == 0) return launch_m<2>(Ac, c.wps, (int)c.count, lds, s72, st, few);
    if (c.order == 1) return launch_m<6>(Ac, c.wps, (int)c.count, lds, s72, st, few);
    return launch_m<12>(Ac, c.wps, (int)c.count, lds, s72, st, few);
}

}  // namespace

int knot_table_run(Engine* e, Image* im) {
    const dim3 grid((unsigned)((im->w + 63) / 64), (unsigned)((im->h + 3) / 4));
    knot_table_kernel<<<grid, 256, 0, e->stream>>>(im->coef32, im->wp32, im->border, (int)im->w, (int)im->h,
                                                   im->knot);
    DIC_HIP_TRY(hipGetLastError());
    im->knot_valid = true;
    return DIC_OK;
}

int knot_table_ensure(Engine* e, Image* im) {
    if (im->knot_valid) return DIC_OK;
    if (!im->knot) {
        if (e->capturing) return DIC_OK;
        const int st = dev_alloc(e, (void**)&im->knot, (size_t)im->h * im->w * sizeof(float4));
        if (st) return st;
    }
    im->knot_wanted = true;
    return knot_table_run(e, im);
}

int launch_icgn_batch(Batch* b, const Image* f, const Image* g, const dic_solve_params* prm, const int32_t* d_perm,
                      const int32_t* h_perm, int64_t count, bool latency, const double* p0_list,
                      dic_subset_result* res_list, int32_t* done_list) {
    Engine* e = b->eng;
    {
        const int st = knot_table_ensure(e, const_cast<Image*>(f));
        if (st) return st;
    }
    BatchArgs A = common_args(b, f, g, prm);
    A.p0_list = p0_list;
    A.res_list = res_list;
    A.done_list = done_list;
    std::vector<SizeClass> cls = size_classes(b, h_perm, count);
    if (latency && b->waves_per_subset == 0)
        for (SizeClass& c : cls)
            while (c.wps < 8 && c.count * 2 * c.wps <= 8 * (int64_t)e->num_cus) c.wps *= 2;
    // the first class runs on the engine's stream, the others on the
    // auxiliary streams, forked from and joined back into it
    const int nstreams = std::min<int>((int)cls.size(), 1 + Engine::kAux);
    if (nstreams > 1) {
        DIC_HIP_TRY(hipEventRecord(e->ev_fork, e->stream));
        for (int k = 0; k + 1 < nstreams; ++k) DIC_HIP_TRY(hipStreamWaitEvent(e->aux[k], e->ev_fork, 0));
    }
    int rc = DIC_OK;
    for (size_t k = 0; k < cls.size() && !rc; ++k) {
        hipStream_t st = (k % nstreams == 0) ? e->stream : e->aux[k % nstreams - 1];
        rc = launch_class(b, A, cls[k], st, d_perm);
    }
    // every forked stream is joined back, also after a failed launch, so
    // that a wait on the engine's stream covers the classes already running
    // (and a capture that started on it can end)
    for (int k = 0; k + 1 < nstreams; ++k) {
        if (hipEventRecord(e->ev_join[k], e->aux[k]) != hipSuccess ||
            hipStreamWaitEvent(e->stream, e->ev_join[k], 0) != hipSuccess) {
            if (!rc) set_error("launch_icgn_batch: joining a size-class stream failed");
            rc = rc ? rc : DIC_ERR_DEVICE;
        }
    }
    return rc;
}

// Every class alone on the engine's stream, for each candidate (fewest
// wavefronts or twice that, window slack 1 or 3 pixels), best of three
// timed runs; the fastest is kept for later solves.  The results the
// batch holds afterwards are those of the last candidate.
int autotune_icgn_batch(Batch* b, const Image* f, const Image* g, const dic_solve_params* prm) {
    Engine* e = b->eng;
    {
        const int st = knot_table_ensure(e, const_cast<Image*>(f));
        if (st) return st;
    }
    const BatchArgs A = common_args(b, f, g, prm);
    for (int o = 0; o < 3; ++o)
        for (int w = 0; w < 9; ++w) b->tuned_wps[o][w] = 0;
    const int saved = b->waves_per_subset;
    b->waves_per_subset = 0;
    const std::vector<SizeClass> cls = size_classes(b);
    b->waves_per_subset = saved;
    hipEvent_t t0, t1;
    DIC_HIP_TRY(hipEventCreate(&t0));
    DIC_HIP_TRY(hipEventCreate(&t1));
    int rc = DIC_OK;
    for (const SizeClass& c0 : cls) {
        // the untuned schedule (fewest wavefronts, the batch's slack) is
        // timed first and kept unless a candidate beats it by more than 3 %,
        // so that timing noise between near-equal candidates does not flip
        // the choice from run to run
        float best = 1e30f;
        int bw = c0.base_wps, bs = b->win_slack;
        std::vector<std::pair<int, int>> cand = {{c0.base_wps, b->win_slack}};
        for (int w = c0.base_wps; w <= std::min(8, 2 * c0.base_wps); w *= 2)
            for (int sl : {1, 3})
                if (!(w == c0.base_wps && sl == b->win_slack)) cand.push_back({w, sl});
        for (size_t ci = 0; ci < cand.size() && !rc; ++ci) {
            SizeClass c = c0;
            c.wps = cand[ci].first;
            c.slack = cand[ci].second;
            float tmin = 1e30f;
            for (int rep = 0; rep < 3 && !rc; ++rep) {
                if (hipEventRecord(t0, e->stream) != hipSuccess) rc = DIC_ERR_DEVICE;
                if (!rc) rc = launch_class(b, A, c, e->stream);
                if (!rc && hipEventRecord(t1, e->stream) != hipSuccess) rc = DIC_ERR_DEVICE;
                float ms = 0.f;
                if (!rc && (hipEventSynchronize(t1) != hipSuccess || hipEventElapsedTime(&ms, t0, t1) != hipSuccess))
                    rc = DIC_ERR_DEVICE;
                tmin = std::min(tmin, ms);
            }
            if (!rc && (ci == 0 || tmin < 0.97f * best)) {
                best = tmin;
                bw = c.wps;
                bs = c.slack;
            }
        }
        if (rc) break;
        b->tuned_wps[c0.order][c0.base_wps] = bw;
        b->tuned_slack[c0.order][c0.base_wps] = bs;
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (rc == DIC_ERR_DEVICE) set_error("autotune: device error while timing a candidate");
    return rc;
}

}  // namespace geodic
