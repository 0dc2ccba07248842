// Batched engine launch (the kernel is in icgn_batch_kernel.h).
#include "internal.h"
#include "icgn_batch_kernel.h"

namespace geodic {

namespace {

using batch::BatchArgs;
using batch::icgn_batch_kernel;
using batch::kRedSlots;

template <int M, int P, int WPS, int S>
int launch_mpw(Batch* b, BatchArgs& A, int blocks, size_t lds) {
    Engine* e = b->eng;
    DIC_HIP_TRY(hipFuncSetAttribute((const void*)icgn_batch_kernel<M, P, WPS, S>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    icgn_batch_kernel<M, P, WPS, S><<<blocks, 64 * WPS, lds, e->stream>>>(A);
    DIC_HIP_TRY(hipGetLastError());
    return DIC_OK;
}

// LDS row stride: 72 floats when every window of the batch fits it (rows
// then sit at immediate offsets, and 72 = 8 mod 64 spreads an 8x8 tile of
// pixels over all 64 banks), else the window's own width.
template <int M, int WPS>
int launch_ms(Batch* b, BatchArgs& A, int blocks, size_t lds, bool s72) {
    return s72 ? launch_mpw<M, 12, WPS, 72>(b, A, blocks, lds) : launch_mpw<M, 12, WPS, 0>(b, A, blocks, lds);
}

// Enough wavefronts per subset to hold the largest subset at 12 pixels per
// thread (the footprint that stays in registers without spilling).
template <int M>
int launch_m(Batch* b, BatchArgs& A, int blocks, size_t lds, bool s72) {
    const int n = b->max_n;
    int wps = b->waves_per_subset;
    if (wps <= 0) wps = n <= 64 * 12 ? 1 : (n <= 128 * 12 ? 2 : (n <= 256 * 12 ? 4 : 8));
    if (n > 64 * 12 * wps) {
        set_error("invalid input: subset larger than the wavefronts per subset can hold");
        return DIC_ERR_INVALID_INPUT;
    }
    switch (wps) {
        case 1: return launch_ms<M, 1>(b, A, blocks, lds, s72);
        case 2: return launch_ms<M, 2>(b, A, blocks, lds, s72);
        case 4: return launch_ms<M, 4>(b, A, blocks, lds, s72);
        case 8: return launch_ms<M, 8>(b, A, blocks, lds, s72);
        default: break;
    }
    set_error("invalid input: waves per subset must be 1, 2, 4 or 8");
    return DIC_ERR_INVALID_INPUT;
}

}  // namespace

int launch_icgn_batch(Batch* b, const Image* f, const Image* g, const dic_solve_params* prm) {
    BatchArgs A;
    A.fC = f->coef32;
    A.gC = g->coef32;
    A.pitch = f->wp32;
    A.border = f->border;
    A.w = (int)f->w;
    A.h = (int)f->h;
    A.wmax = (double)f->w - 1e-9;
    A.hmax = (double)f->h - 1e-9;
    A.n_sub = b->n;
    A.coords = b->d_coords;
    A.tpl_index = b->d_tpl_index;
    A.tpl_start = b->d_tpl_start;
    A.tpl_xy = b->d_tpl_xy_eng;
    A.tpl_ext = b->d_tpl_ext;
    A.perm = b->d_order_perm;
    A.p0 = b->d_p0;
    A.p_out = b->d_p;
    A.res = b->d_res;
    A.hist = prm->want_history ? b->d_hist : nullptr;
    A.hist_cap = prm->want_history ? b->hist_cap : 0;
    A.max_it = prm->max_iterations;
    A.tol = prm->tolerance;
    A.max_norm = prm->max_norm;
    // LDS: the coefficient window (template box, block taps, 16-byte
    // rounding and the restaging slack, with room for moderate strain)
    // followed by the template offsets as FP32 pairs
    // LDS: the coefficient window (template box, slack, taps, cell rounding
    // and a little strain), then the template offsets as FP16 pairs
    const int side = b->max_side + 2 * b->win_slack + 8;
    const int row = (side + 3 + 3) & ~3;  // plus the 16-byte rounding of the first column
    const bool s72 = row <= 72;
    A.win_cap = (s72 ? 72 : row) * side;
    A.win_slack = b->win_slack;
    const size_t lds = sizeof(float) * (size_t)A.win_cap + 4 * (size_t)b->max_n;
    const int blocks = (int)b->n;
    const int order = b->max_order;
    if (order == 0) return launch_m<2>(b, A, blocks, lds, s72);
    if (order == 1) return launch_m<6>(b, A, blocks, lds, s72);
    return launch_m<12>(b, A, blocks, lds, s72);
}

}  // namespace geodic
