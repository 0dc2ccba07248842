// Batched engine launch (the kernel is in icgn_batch_kernel.h).
#include "internal.h"
#include "icgn_batch_kernel.h"

namespace geodic {

namespace {

using batch::BatchArgs;
using batch::icgn_batch_kernel;

template <int M>
int launch_m(Batch* b, BatchArgs& A, int blocks, size_t lds) {
    Engine* e = b->eng;
    DIC_HIP_TRY(hipFuncSetAttribute((const void*)icgn_batch_kernel<M>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds));
    icgn_batch_kernel<M><<<blocks, 64, lds, e->stream>>>(A);
    DIC_HIP_TRY(hipGetLastError());
    return DIC_OK;
}

}  // namespace

int launch_icgn_batch(Batch* b, const Image* f, const Image* g, const dic_solve_params* prm) {
    Engine* e = b->eng;
    // FP32 slab: 4 floats per pixel, offsets are the template prefix sums
    const size_t need = (size_t)b->total_px * 4;
    if (b->slab_floats < need) {
        if (b->d_slab) dev_free(e, b->d_slab, b->slab_floats * sizeof(float));
        b->d_slab = nullptr;
        b->slab_floats = 0;
        int st = dev_alloc(e, (void**)&b->d_slab, need * sizeof(float));
        if (st) return st;
        b->slab_floats = need;
    }
    if (!b->slab_off_ready) {
        std::vector<int64_t> off(b->n);
        int64_t acc = 0;
        for (int64_t s = 0; s < b->n; ++s) {
            off[s] = acc;
            const int t = b->h_tpl_index[s];
            acc += b->h_tpl_start[t + 1] - b->h_tpl_start[t];
        }
        DIC_HIP_TRY(hipMemcpyAsync(b->d_slab_off, off.data(), sizeof(int64_t) * b->n, hipMemcpyHostToDevice,
                                   e->stream));
        DIC_HIP_TRY(hipStreamSynchronize(e->stream));
        b->slab_off_ready = true;
    }
    BatchArgs A;
    A.fC = f->coef32;
    A.gC = g->coef32;
    A.wp = f->wp;
    A.border = f->border;
    A.w = (int)f->w;
    A.h = (int)f->h;
    A.wmax = (double)f->w - 1e-9;
    A.hmax = (double)f->h - 1e-9;
    A.n_sub = b->n;
    A.coords = b->d_coords;
    A.tpl_index = b->d_tpl_index;
    A.tpl_start = b->d_tpl_start;
    A.tpl_xy = b->d_tpl_xy;
    A.tpl_ext = b->d_tpl_ext;
    A.perm = b->d_order_perm;
    A.slab = (float4*)b->d_slab;
    A.slab_off = b->d_slab_off;
    A.p0 = b->d_p0;
    A.p_out = b->d_p;
    A.res = b->d_res;
    A.hist = prm->want_history ? b->d_hist : nullptr;
    A.hist_cap = prm->want_history ? b->hist_cap : 0;
    A.max_it = prm->max_iterations;
    A.tol = prm->tolerance;
    A.max_norm = prm->max_norm;
    // LDS window: the template box plus a few pixels of deformation slack
    const int side = b->max_side + 6 + 2 * b->win_slack;
    A.win_cap = side * side;
    const size_t lds = sizeof(float) * (size_t)A.win_cap;
    const int blocks = (int)b->n;
    const int order = b->max_order;
    if (order == 0) return launch_m<2>(b, A, blocks, lds);
    if (order == 1) return launch_m<6>(b, A, blocks, lds);
    return launch_m<12>(b, A, blocks, lds);
}

}  // namespace geodic
