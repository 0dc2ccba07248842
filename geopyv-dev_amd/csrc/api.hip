// C ABI of the engine (include/geodic.h): engine, images, batches.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <string>

#include "internal.h"

namespace geodic {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int dev_alloc(Engine* e, void** p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) bytes = 16;
    hipError_t st = hipMalloc(p, bytes);
    if (st != hipSuccess) {
        *p = nullptr;
        set_error(std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(st));
        return st == hipErrorOutOfMemory ? DIC_ERR_OUT_OF_MEMORY : DIC_ERR_DEVICE;
    }
    e->bytes_in_use += (int64_t)bytes;
    return DIC_OK;
}

void dev_free(Engine* e, void* p, size_t bytes) {
    if (!p) return;
    if (bytes == 0) bytes = 16;
    (void)hipFree(p);
    e->bytes_in_use -= (int64_t)bytes;
}


// Argument checks shared by the solve entry points.
int validate_solve(const Batch* b, const Image* f, const Image* g, const dic_solve_params* prm) {
    const Engine* e = b->eng;
    CHECK_ARG(f->eng->device == e->device && g->eng->device == e->device,
              "invalid input: images live on another device");
    CHECK_ARG(f->h == g->h && f->w == g->w && f->border == g->border,
              "invalid input: reference and target images differ in shape or border");
    CHECK_ARG(prm->method == DIC_METHOD_ICGN || prm->method == DIC_METHOD_FAGN, "invalid input: method");
    CHECK_ARG(prm->precision == DIC_PREC_PLAIN_F64 || prm->precision == DIC_PREC_COMP_F32, "invalid input: precision");
    CHECK_ARG(prm->max_iterations >= 0 && prm->max_iterations <= 100000, "invalid input: max_iterations");
    CHECK_ARG(!(prm->precision == DIC_PREC_COMP_F32 && prm->method == DIC_METHOD_FAGN),
              "invalid input: the batched engine runs ICGN; use DIC_PREC_PLAIN_F64 for FAGN");
    CHECK_ARG(b->max_n <= 32767 * 2, "invalid input: template too large");
    return DIC_OK;
}

}  // namespace geodic

using namespace geodic;

struct dic_frame_graph {
    Engine* e = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double* tmp = nullptr;  // the graph's own prefilter row-pass buffer
    size_t tmp_bytes = 0;
};

extern "C" {

const char* dic_last_error(void) { return g_last_error.c_str(); }

int dic_engine_create(int device, dic_engine** out) {
    CHECK_ARG(out, "dic_engine_create: out is NULL");
    *out = nullptr;
    int count = 0;
    DIC_HIP_TRY(hipGetDeviceCount(&count));
    CHECK_ARG(device >= 0 && device < count, "dic_engine_create: no such HIP device");
    DIC_HIP_TRY(hipSetDevice(device));
    dic_engine* d = new dic_engine();
    Engine* e = &d->e;
    e->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        e->num_cus = prop.multiProcessorCount;
    // default-priority streams only (mixed stream priorities are avoided):
    // the engine's stream plus three that run size classes of a batch
    // alongside it
    bool ok = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreate(&e->ev_pre0) == hipSuccess && hipEventCreate(&e->ev_pre1) == hipSuccess &&
              hipEventCreate(&e->ev_sol0) == hipSuccess && hipEventCreate(&e->ev_sol1) == hipSuccess &&
              hipEventCreateWithFlags(&e->ev_fork, hipEventDisableTiming) == hipSuccess;
    for (int k = 0; ok && k < Engine::kAux; ++k)
        ok = hipStreamCreateWithFlags(&e->aux[k], hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&e->ev_join[k], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        set_error("dic_engine_create: stream/event creation failed");
        dic_engine_destroy(d);
        return DIC_ERR_DEVICE;
    }
    // device address of the fixed scratch of dic_batch_solve_noalloc on
    // this device (once per device, with the device current)
    const int sst = noalloc_scratch_init(device);
    if (sst) {
        dic_engine_destroy(d);
        return sst;
    }
    *out = d;
    return DIC_OK;
}

void dic_engine_destroy(dic_engine* d) {
    if (!d) return;
    Engine* e = &d->e;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    for (auto& pr : e->twiddles) (void)hipFree(pr.second);
    for (auto& pr : e->fir_taps) (void)hipFree(pr.second);
    if (e->pre_tmp) (void)hipFree(e->pre_tmp);
    if (e->pix_stage) (void)hipFree(e->pix_stage);
    for (hipEvent_t ev : {e->ev_pre0, e->ev_pre1, e->ev_sol0, e->ev_sol1})
        if (ev) (void)hipEventDestroy(ev);
    if (e->ev_fork) (void)hipEventDestroy(e->ev_fork);
    for (int k = 0; k < Engine::kAux; ++k) {
        if (e->aux[k]) {
            (void)hipStreamSynchronize(e->aux[k]);
            (void)hipStreamDestroy(e->aux[k]);
        }
        if (e->ev_join[k]) (void)hipEventDestroy(e->ev_join[k]);
    }
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete d;
}

int dic_engine_sync(dic_engine* d) {
    CHECK_ARG(d, "engine is NULL");
    DIC_HIP_TRY(hipSetDevice(d->e.device));
    DIC_HIP_TRY(hipStreamSynchronize(d->e.stream));
    return DIC_OK;
}

int dic_engine_set_prefilter(dic_engine* d, int method) {
    CHECK_ARG(d, "engine is NULL");
    CHECK_ARG(method == DIC_PREFILTER_FIR || method == DIC_PREFILTER_FFT, "invalid input: prefilter method");
    std::lock_guard<std::mutex> lk(d->e.mu);
    d->e.prefilter_method = method;
    return DIC_OK;
}

int dic_engine_timing(dic_engine* d, double* prefilter_ms, double* solve_ms) {
    CHECK_ARG(d, "engine is NULL");
    Engine* e = &d->e;
    DIC_HIP_TRY(hipSetDevice(e->device));
    DIC_HIP_TRY(hipStreamSynchronize(e->stream));
    float ms = 0.f;
    if (prefilter_ms) {
        *prefilter_ms = 0.0;
        if (e->pre_recorded) {
            DIC_HIP_TRY(hipEventElapsedTime(&ms, e->ev_pre0, e->ev_pre1));
            *prefilter_ms = ms;
        }
    }
    if (solve_ms) {
        *solve_ms = 0.0;
        if (e->sol_recorded) {
            DIC_HIP_TRY(hipEventElapsedTime(&ms, e->ev_sol0, e->ev_sol1));
            *solve_ms = ms;
        }
    }
    return DIC_OK;
}

int64_t dic_engine_bytes_in_use(dic_engine* d) { return d ? d->e.bytes_in_use.load() : 0; }

// ---------------------------------------------------------------------------
// images
// ---------------------------------------------------------------------------
int dic_image_create(dic_engine* d, const void* pixels, int pix_type, int mem, int64_t h, int64_t w, int border,
                     dic_image** out) {
    CHECK_ARG(d && out && pixels, "dic_image_create: NULL argument");
    *out = nullptr;
    CHECK_ARG(h >= 1 && w >= 1, "invalid input: image must be at least 1x1");
    CHECK_ARG(border >= 3, "invalid input: border must be >= 3");
    CHECK_ARG(pixel_bytes(pix_type) > 0, "invalid input: pixel type");
    // reflect-101 reflects once (image.rs:519-528): the 5-tap blur needs 3 pixels
    CHECK_ARG(!loader_pixels(pix_type) || (h >= 3 && w >= 3), "invalid input: a loaded image must be at least 3x3");
    CHECK_ARG(mem == DIC_MEM_HOST || mem == DIC_MEM_DEVICE, "invalid input: memory kind");
    Engine* e = &d->e;
    DIC_HIP_TRY(hipSetDevice(e->device));
    std::lock_guard<std::mutex> lk(e->mu);
    dic_image* di = new dic_image();
    Image* im = &di->im;
    im->eng = e;
    im->h = h;
    im->w = w;
    im->border = border;
    im->hp = h + 2 * border;
    im->wp = w + 2 * border;
    im->wp32 = (im->wp + 3) & ~(int64_t)3;
    const size_t np = (size_t)im->hp * im->wp;
    const size_t np32 = (size_t)im->hp * im->wp32;
    int st = dev_alloc(e, (void**)&im->coef64, np * sizeof(double));
    if (!st) st = dev_alloc(e, (void**)&im->coef32, np32 * sizeof(float));
    // the pitch padding of the FP32 copy is never written by the prefilter;
    // window loads may read it, so give it a defined value
    if (!st && im->wp32 != im->wp && hipMemsetAsync(im->coef32, 0, np32 * sizeof(float), e->stream) != hipSuccess) {
        set_error("dic_image_create: clearing the FP32 coefficients failed");
        st = DIC_ERR_DEVICE;
    }
    if (!st) st = prefilter_run(e, im, pixels, pix_type, mem);
    if (!st && mem == DIC_MEM_HOST && hipStreamSynchronize(e->stream) != hipSuccess) {
        set_error("dic_image_create: synchronisation failed");
        st = DIC_ERR_DEVICE;
    }
    if (st) {
        dev_free(e, im->coef64, np * sizeof(double));
        dev_free(e, im->coef32, np32 * sizeof(float));
        dev_free(e, im->grey_dev, (size_t)h * w);
        delete di;
        return st;
    }
    *out = di;
    return DIC_OK;
}

int dic_image_update(dic_image* di, const void* pixels, int pix_type, int mem) {
    CHECK_ARG(di && pixels, "dic_image_update: NULL argument");
    CHECK_ARG(pixel_bytes(pix_type) > 0, "invalid input: pixel type");
    CHECK_ARG(mem == DIC_MEM_HOST || mem == DIC_MEM_DEVICE, "invalid input: memory kind");
    Image* im = &di->im;
    CHECK_ARG(!loader_pixels(pix_type) || (im->h >= 3 && im->w >= 3), "invalid input: a loaded image must be at least 3x3");
    Engine* e = im->eng;
    DIC_HIP_TRY(hipSetDevice(e->device));
    std::lock_guard<std::mutex> lk(e->mu);
    int st = prefilter_run(e, im, pixels, pix_type, mem);
    if (st) return st;
    if (mem == DIC_MEM_HOST) DIC_HIP_TRY(hipStreamSynchronize(e->stream));
    return DIC_OK;
}

int dic_image_grey(const dic_image* di, double* out) {
    CHECK_ARG(di && out, "dic_image_grey: NULL argument");
    const Image* im = &di->im;
    CHECK_ARG(im->grey_dev, "invalid input: the image was not loaded from file pixels");
    Engine* e = im->eng;
    DIC_HIP_TRY(hipSetDevice(e->device));
    std::lock_guard<std::mutex> lk(e->mu);
    std::vector<uint8_t> g((size_t)im->h * im->w);
    DIC_HIP_TRY(hipMemcpyAsync(g.data(), im->grey_dev, g.size(), hipMemcpyDeviceToHost, e->stream));
    DIC_HIP_TRY(hipStreamSynchronize(e->stream));
    for (size_t i = 0; i < g.size(); ++i) out[i] = (double)g[i];
    return DIC_OK;
}

int dic_image_coef(const dic_image* di, double* out) {
    CHECK_ARG(di && out, "dic_image_coef: NULL argument");
    const Image* im = &di->im;
    Engine* e = im->eng;
    DIC_HIP_TRY(hipSetDevice(e->device));
    std::lock_guard<std::mutex> lk(e->mu);
    DIC_HIP_TRY(hipMemcpyAsync(out, im->coef64, sizeof(double) * im->hp * im->wp, hipMemcpyDeviceToHost, e->stream));
    DIC_HIP_TRY(hipStreamSynchronize(e->stream));
    return DIC_OK;
}

int dic_image_info(const dic_image* di, int64_t* h, int64_t* w, int* border) {
    CHECK_ARG(di, "dic_image_info: NULL image");
    if (h) *h = di->im.h;
    if (w) *w = di->im.w;
    if (border) *border = di->im.border;
    return DIC_OK;
}

void dic_image_destroy(dic_image* di) {
    if (!di) return;
    Image* im = &di->im;
    Engine* e = im->eng;
    (void)hipSetDevice(e->device);
    std::lock_guard<std::mutex> lk(e->mu);
    (void)hipStreamSynchronize(e->stream);
    const size_t np = (size_t)im->hp * im->wp;
    dev_free(e, im->coef64, np * sizeof(double));
    dev_free(e, im->coef32, (size_t)im->hp * im->wp32 * sizeof(float));
    dev_free(e, im->pix_dev, im->pix_dev_bytes);
    dev_free(e, im->grey_dev, (size_t)im->h * im->w);
    dev_free(e, im->knot, (size_t)im->h * im->w * sizeof(float4));
    delete di;
}

// ---------------------------------------------------------------------------
// batches
// ---------------------------------------------------------------------------
// Bytes of the batched engine's per-subset start record (batch::SubMeta:
// centre x, y as doubles, template start int64, size and index int32,
// extents 4 x int32).
constexpr size_t kMetaBytes = 48;

// Staging layout of a list solve of up to `cap` subsets: the list, its
// dispatch order (int32 each, 8-byte aligned), the initial warps, then (at
// list_io_bytes) the results.
static size_t list_ints_bytes(int64_t cap) { return ((size_t)cap * 4 + 7) & ~(size_t)7; }
static size_t list_io_bytes(int64_t cap) { return 2 * list_ints_bytes(cap) + sizeof(double) * 12 * (size_t)cap; }

static void batch_free(Batch* b) {
    Engine* e = b->eng;
    const int64_t n = b->n;
    dev_free(e, b->d_coords, sizeof(double) * 2 * n);
    dev_free(e, b->d_orders, sizeof(int32_t) * n);
    dev_free(e, b->d_tpl_index, sizeof(int32_t) * n);
    dev_free(e, b->d_tpl_start, sizeof(int64_t) * (b->n_tpl + 1));
    dev_free(e, b->d_tpl_xy, sizeof(int32_t) * (b->h_tpl_start.empty() ? 0 : b->h_tpl_start.back()));
    dev_free(e, b->d_tpl_xy_eng, sizeof(int32_t) * (b->h_tpl_start.empty() ? 0 : b->h_tpl_start.back()));
    dev_free(e, b->d_tpl_ext, sizeof(int32_t) * 4 * b->n_tpl);
    dev_free(e, b->d_meta, kMetaBytes * n);
    dev_free(e, b->d_scr_off, sizeof(int64_t) * n);
    dev_free(e, b->d_p0, sizeof(double) * 12 * n);
    dev_free(e, b->d_p, sizeof(double) * 12 * n);
    dev_free(e, b->d_res, sizeof(dic_subset_result) * n);
    dev_free(e, b->d_hist, sizeof(double) * 4 * n * b->hist_cap);
    dev_free(e, b->d_scratch, sizeof(double) * b->scratch_doubles);
    dev_free(e, b->d_order_perm, sizeof(int32_t) * n);
    dev_free(e, b->d_by_order, sizeof(int32_t) * (b->mixed_orders ? n : 0));
    if (b->h_list_io) (void)hipHostFree(b->h_list_io);
}

namespace {

// Pixel order of a template for the batched engine.  The window rows in
// LDS are 72 floats apart and LDS serves a wavefront's reads in two halves
// of 32 lanes over 32 banks, so pixel (dx, dy) starts in bank
// (dx + 8 dy) mod 32 and each run of 32 consecutive pixels (one half-wave)
// should cover 32 distinct banks.  Whole 8x4 blocks of the template do
// (and keep a compact footprint, so a near-uniform warp shifts them
// together); the pixels of partial blocks are dealt into runs of 32 taking
// one pixel per bank class, fullest classes first.  Sums over the subset
// do not depend on the order beyond rounding.
// IEEE half-precision bits of an integer, rounded to nearest even (exact
// up to 2048 in magnitude; template offsets are far smaller).
uint16_t half_bits(int v) {
    if (v == 0) return 0;
    const uint16_t sign = v < 0 ? 0x8000 : 0;
    const uint32_t a = (uint32_t)std::abs(v);
    int e = 31 - __builtin_clz(a);  // a in [2^e, 2^(e+1))
    uint32_t m;
    if (e <= 10) {
        m = (a << (10 - e)) & 0x3ffu;
    } else {
        const uint32_t sh = (uint32_t)e - 10, r = a & ((1u << sh) - 1), h = 1u << (sh - 1);
        m = (a >> sh) & 0x3ffu;
        if (r > h || (r == h && (m & 1u))) {
            if (++m == 0x400u) {
                m = 0;
                ++e;
            }
        }
    }
    return (uint16_t)(sign | (uint16_t)((e + 15) << 10) | (uint16_t)m);
}

void engine_pixel_order(const int32_t* in, int64_t n, int32_t* out) {
    auto ux = [](int32_t v) { return (int)(int16_t)(v & 0xffff); };
    auto uy = [](int32_t v) { return (int)(int16_t)((uint32_t)v >> 16); };
    if (n == 0) return;
    int x0 = ux(in[0]), y0 = uy(in[0]);
    for (int64_t i = 1; i < n; ++i) {
        x0 = std::min(x0, ux(in[i]));
        y0 = std::min(y0, uy(in[i]));
    }
    std::map<std::pair<int, int>, std::vector<int32_t>> blocks;
    for (int64_t i = 0; i < n; ++i) blocks[{(uy(in[i]) - y0) >> 2, (ux(in[i]) - x0) >> 3}].push_back(in[i]);
    int64_t k = 0;
    std::vector<std::vector<int32_t>> bank(32);
    for (auto& kv : blocks) {
        if (kv.second.size() == 32) {
            for (int32_t v : kv.second) out[k++] = v;
        } else {
            for (int32_t v : kv.second) bank[(unsigned)(ux(v) + 8 * uy(v)) & 31].push_back(v);
        }
    }
    for (auto& b : bank) std::reverse(b.begin(), b.end());
    while (k < n) {
        std::vector<int> ord(32);
        std::iota(ord.begin(), ord.end(), 0);
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int c) { return bank[a].size() > bank[c].size(); });
        for (int b : ord)
            if (!bank[b].empty()) {
                out[k++] = bank[b].back();
                bank[b].pop_back();
            }
    }
}

}  // namespace

int dic_batch_create(dic_engine* d, int64_t n_subsets, const double* coords, const int32_t* orders,
                     const int32_t* tpl_index, int32_t n_templates, const int64_t* tpl_start, const int16_t* tpl_xy,
                     dic_batch** out) {
    CHECK_ARG(d && out, "dic_batch_create: NULL argument");
    *out = nullptr;
    CHECK_ARG(n_subsets >= 1, "invalid input: batch needs at least one subset");
    CHECK_ARG(n_subsets <= (int64_t)1 << 30, "invalid input: batch too large");
    CHECK_ARG(coords && orders && tpl_index && tpl_start && tpl_xy, "dic_batch_create: NULL array");
    CHECK_ARG(n_templates >= 1, "invalid input: no templates");
    CHECK_ARG(tpl_start[0] == 0, "invalid input: tpl_start[0] must be 0");
    for (int t = 0; t < n_templates; ++t)
        CHECK_ARG(tpl_start[t + 1] >= tpl_start[t], "invalid input: tpl_start must be non-decreasing");
    int32_t max_order = 0;
    bool mixed = false;
    for (int64_t s = 0; s < n_subsets; ++s) {
        CHECK_ARG(orders[s] >= 0 && orders[s] <= 2, "invalid input: subset_order must be 0, 1, or 2");
        CHECK_ARG(tpl_index[s] >= 0 && tpl_index[s] < n_templates, "invalid input: template index out of range");
        max_order = std::max(max_order, orders[s]);
        mixed = mixed || orders[s] != orders[0];
    }
    Engine* e = &d->e;
    DIC_HIP_TRY(hipSetDevice(e->device));
    std::lock_guard<std::mutex> lk(e->mu);
    dic_batch* db = new dic_batch();
    Batch* b = &db->b;
    b->eng = e;
    b->n = n_subsets;
    b->n_tpl = n_templates;
    b->max_order = max_order;
    b->mixed_orders = mixed;
    b->h_orders.assign(orders, orders + n_subsets);
    b->h_tpl_start.assign(tpl_start, tpl_start + n_templates + 1);
    b->h_tpl_index.assign(tpl_index, tpl_index + n_subsets);
    const int64_t n_off = tpl_start[n_templates];
    std::vector<int32_t> packed(n_off);
    std::vector<int32_t> ext(4 * (size_t)n_templates, 0);
    for (int t = 0; t < n_templates; ++t) {
        int x0 = 0, x1 = 0, y0 = 0, y1 = 0;
        for (int64_t i = tpl_start[t]; i < tpl_start[t + 1]; ++i) {
            const int dx = tpl_xy[2 * i], dy = tpl_xy[2 * i + 1];
            if (i == tpl_start[t]) {
                x0 = x1 = dx;
                y0 = y1 = dy;
            } else {
                x0 = std::min(x0, dx);
                x1 = std::max(x1, dx);
                y0 = std::min(y0, dy);
                y1 = std::max(y1, dy);
            }
            packed[i] = (int32_t)(((uint32_t)(uint16_t)dx) | ((uint32_t)(uint16_t)dy << 16));
        }
        ext[4 * t] = x0;
        ext[4 * t + 1] = x1;
        ext[4 * t + 2] = y0;
        ext[4 * t + 3] = y1;
        const int64_t nt = tpl_start[t + 1] - tpl_start[t];
        b->max_n = std::max<int32_t>(b->max_n, (int32_t)nt);
        b->max_side = std::max(b->max_side, std::max(x1 - x0 + 1, y1 - y0 + 1));
    }
    // the plain kernel's per-pixel arrays: subset s's start at (pixels of
    // the subsets before it) x (arrays per pixel)
    b->h_scr_off.resize(n_subsets);
    for (int64_t s = 0; s < n_subsets; ++s) {
        const int t = tpl_index[s];
        b->h_scr_off[s] = b->total_px;
        b->total_px += tpl_start[t + 1] - tpl_start[t];
    }
    b->h_coords.assign(coords, coords + 2 * n_subsets);
    b->h_ext = ext;
    // provisional dispatch order (highest warp order, then largest subsets
    // first); refined for the image shape at the first solve
    std::vector<int32_t> perm(n_subsets);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](int32_t a, int32_t c) {
        if (orders[a] != orders[c]) return orders[a] > orders[c];
        const int64_t na = tpl_start[tpl_index[a] + 1] - tpl_start[tpl_index[a]];
        const int64_t nc = tpl_start[tpl_index[c] + 1] - tpl_start[tpl_index[c]];
        return na > nc;
    });
    // mixed orders: the subsets grouped by order for the plain kernel's
    // one launch per order
    std::vector<int32_t> by_order;
    if (mixed) {
        by_order.reserve(n_subsets);
        for (int o = 0; o < 3; ++o) {
            b->by_order_off[o] = (int64_t)by_order.size();
            for (int64_t s = 0; s < n_subsets; ++s)
                if (orders[s] == o) by_order.push_back((int32_t)s);
        }
        b->by_order_off[3] = n_subsets;
    }
    int st = DIC_OK;
    auto A = [&](void** p, size_t bytes) {
        if (!st) st = dev_alloc(e, p, bytes);
    };
    A((void**)&b->d_coords, sizeof(double) * 2 * n_subsets);
    A((void**)&b->d_orders, sizeof(int32_t) * n_subsets);
    A((void**)&b->d_tpl_index, sizeof(int32_t) * n_subsets);
    A((void**)&b->d_tpl_start, sizeof(int64_t) * (n_templates + 1));
    A((void**)&b->d_tpl_xy, sizeof(int32_t) * n_off);
    A((void**)&b->d_tpl_xy_eng, sizeof(int32_t) * n_off);
    A((void**)&b->d_tpl_ext, sizeof(int32_t) * 4 * n_templates);
    A((void**)&b->d_scr_off, sizeof(int64_t) * n_subsets);
    A((void**)&b->d_p0, sizeof(double) * 12 * n_subsets);
    A((void**)&b->d_p, sizeof(double) * 12 * n_subsets);
    A((void**)&b->d_res, sizeof(dic_subset_result) * n_subsets);
    A((void**)&b->d_order_perm, sizeof(int32_t) * n_subsets);
    if (mixed) A((void**)&b->d_by_order, sizeof(int32_t) * n_subsets);
    A(&b->d_meta, kMetaBytes * n_subsets);
    auto H2D = [&](void* dst, const void* src, size_t bytes) {
        if (!st && bytes && hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, e->stream) != hipSuccess) {
            set_error("dic_batch_create: upload failed");
            st = DIC_ERR_DEVICE;
        }
    };
    H2D(b->d_coords, coords, sizeof(double) * 2 * n_subsets);
    H2D(b->d_orders, orders, sizeof(int32_t) * n_subsets);
    H2D(b->d_tpl_index, tpl_index, sizeof(int32_t) * n_subsets);
    H2D(b->d_tpl_start, tpl_start, sizeof(int64_t) * (n_templates + 1));
    H2D(b->d_tpl_xy, packed.data(), sizeof(int32_t) * n_off);
    // the batched engine's copy: bank-friendly order, each offset pair as
    // the two half-precision numbers its kernels keep in LDS
    std::vector<int32_t> packed_eng(n_off);
    for (int t = 0; t < n_templates; ++t) engine_pixel_order(packed.data() + tpl_start[t], tpl_start[t + 1] - tpl_start[t],
                                                            packed_eng.data() + tpl_start[t]);
    for (int64_t k = 0; k < n_off; ++k) {
        const int32_t o = packed_eng[k];
        packed_eng[k] = (int32_t)((uint32_t)half_bits((int16_t)(o & 0xffff)) |
                                  ((uint32_t)half_bits((int16_t)((uint32_t)o >> 16)) << 16));
    }
    H2D(b->d_tpl_xy_eng, packed_eng.data(), sizeof(int32_t) * n_off);
    H2D(b->d_tpl_ext, ext.data(), sizeof(int32_t) * 4 * n_templates);
    H2D(b->d_order_perm, perm.data(), sizeof(int32_t) * n_subsets);
    if (mixed) H2D(b->d_by_order, by_order.data(), sizeof(int32_t) * n_subsets);
    H2D(b->d_scr_off, b->h_scr_off.data(), sizeof(int64_t) * n_subsets);
    // the batched engine's start record per subset (centre, template
    // start, size and extents)
    std::vector<char> meta(kMetaBytes * n_subsets);
    for (int64_t s = 0; s < n_subsets; ++s) {
        char* m = meta.data() + kMetaBytes * s;
        const int t = tpl_index[s];
        const int64_t st0 = tpl_start[t];
        const int32_t nt = (int32_t)(tpl_start[t + 1] - tpl_start[t]);
        std::memcpy(m, &coords[2 * s], 16);
        std::memcpy(m + 16, &st0, 8);
        std::memcpy(m + 24, &nt, 4);
        std::memcpy(m + 28, &t, 4);
        std::memcpy(m + 32, &ext[4 * (size_t)t], 16);
    }
    H2D(b->d_meta, meta.data(), meta.size());
    if (!st && hipMemsetAsync(b->d_p0, 0, sizeof(double) * 12 * n_subsets, e->stream) != hipSuccess) st = DIC_ERR_DEVICE;
    if (!st && hipMemsetAsync(b->d_p, 0, sizeof(double) * 12 * n_subsets, e->stream) != hipSuccess) st = DIC_ERR_DEVICE;
    if (!st && hipMemsetAsync(b->d_res, 0, sizeof(dic_subset_result) * n_subsets, e->stream) != hipSuccess)
        st = DIC_ERR_DEVICE;
    // host vectors must stay alive until the copies land
    if (!st && hipStreamSynchronize(e->stream) != hipSuccess) st = DIC_ERR_DEVICE;
    if (st) {
        batch_free(b);
        delete db;
        return st;
    }
    *out = db;
    return DIC_OK;
}

int dic_batch_set_p0(dic_batch* db, const double* p0, int mem) {
    CHECK_ARG(db && p0, "dic_batch_set_p0: NULL argument");
    Batch* b = &db->b;
    Engine* e = b->eng;
    DIC_HIP_TRY(hipSetDevice(e->device));
    std::lock_guard<std::mutex> lk(e->mu);
    DIC_HIP_TRY(hipMemcpyAsync(b->d_p0, p0, sizeof(double) * 12 * b->n,
                               mem == DIC_MEM_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, e->stream));
    if (mem == DIC_MEM_HOST) DIC_HIP_TRY(hipStreamSynchronize(e->stream));
    return DIC_OK;
}

int dic_batch_reset_warps(dic_batch* db, const double* p, int mem) {
    CHECK_ARG(db && p, "dic_batch_reset_warps: NULL argument");
    Batch* b = &db->b;
    Engine* e = b->eng;
    DIC_HIP_TRY(hipSetDevice(e->device));
    std::lock_guard<std::mutex> lk(e->mu);
    const hipMemcpyKind k = mem == DIC_MEM_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
    DIC_HIP_TRY(hipMemcpyAsync(b->d_p0, p, sizeof(double) * 12 * b->n, k, e->stream));
    DIC_HIP_TRY(hipMemcpyAsync(b->d_p, p, sizeof(double) * 12 * b->n, k, e->stream));
    if (mem == DIC_MEM_HOST) DIC_HIP_TRY(hipStreamSynchronize(e->stream));
    return DIC_OK;
}

int dic_batch_set_schedule(dic_batch* db, int win_slack, int waves_per_subset) {
    CHECK_ARG(db, "dic_batch_set_schedule: NULL batch");
    CHECK_ARG(win_slack >= 0 && win_slack <= 64, "invalid input: win_slack must be in 0..64");
    CHECK_ARG(waves_per_subset == 0 || waves_per_subset == 1 || waves_per_subset == 2 || waves_per_subset == 4 ||
                  waves_per_subset == 8,
              "invalid input: waves_per_subset must be 0 (automatic), 1, 2, 4 or 8");
    std::lock_guard<std::mutex> lk(db->b.eng->mu);
    db->b.win_slack = win_slack;
    db->b.waves_per_subset = waves_per_subset;
    return DIC_OK;
}

namespace {

// Dispatch order for an image of w x h.  Highest warp order first (each
// order's classes are separate launches), then largest subsets first.
// Within an order and a size, the subsets (in their given, usually raster, order) are cut into
// eight contiguous bands and dealt out so that dispatch slot b takes band
// b mod 8: the hardware sends workgroup b to XCD b mod 8, so each XCD works
// through one band and neighbouring subsets share its L2 (an eighth of the
// HBM traffic of raster order).  Inside a band the subsets go in order of
// the distance of their template box from the image edge, nearest first:
// clamped pixels make the edge subsets the slow ones (late convergence or
// the iteration cap), and started late they would decide the launch's tail.
int refresh_order(Batch* b, int64_t w, int64_t h) {
    if (b->perm_w == w && b->perm_h == h) return DIC_OK;
    const int64_t n = b->n;
    std::vector<int64_t> npx(n);
    std::vector<double> dist(n);
    for (int64_t s = 0; s < n; ++s) {
        const int t = b->h_tpl_index[s];
        npx[s] = b->h_tpl_start[t + 1] - b->h_tpl_start[t];
        const double xc = b->h_coords[2 * s], yc = b->h_coords[2 * s + 1];
        const int32_t* e = &b->h_ext[4 * (size_t)t];
        const double d = std::min(std::min(xc + e[0], (double)(w - 1) - (xc + e[1])),
                                  std::min(yc + e[2], (double)(h - 1) - (yc + e[3])));
        dist[s] = std::isfinite(d) ? d : -1e300;  // non-finite centres first
    }
    std::vector<int32_t> by_size(n);
    std::iota(by_size.begin(), by_size.end(), 0);
    const std::vector<int32_t>& ord = b->h_orders;
    std::stable_sort(by_size.begin(), by_size.end(), [&](int32_t a, int32_t c) {
        if (ord[a] != ord[c]) return ord[a] > ord[c];
        return npx[a] > npx[c];
    });
    b->h_perm.clear();
    b->h_perm.reserve(n);
    for (int64_t g0 = 0; g0 < n;) {
        int64_t g1 = g0;
        while (g1 < n && npx[by_size[g1]] == npx[by_size[g0]] && ord[by_size[g1]] == ord[by_size[g0]]) ++g1;
        const int64_t len = g1 - g0;
        std::vector<std::vector<int32_t>> band(8);
        for (int x = 0; x < 8; ++x) {
            const int64_t a = g0 + len * x / 8, c = g0 + len * (x + 1) / 8;
            for (int64_t i = a; i < c; ++i) band[x].push_back(by_size[i]);
            std::stable_sort(band[x].begin(), band[x].end(), [&](int32_t u, int32_t q) { return dist[u] < dist[q]; });
        }
        for (size_t j = 0; j < band[7].size() + 1; ++j)
            for (int x = 0; x < 8; ++x)
                if (j < band[x].size()) b->h_perm.push_back(band[x][j]);
        g0 = g1;
    }
    if ((int64_t)b->h_perm.size() != n) {
        set_error("dispatch order: internal error");
        return DIC_ERR_DEVICE;
    }
    Engine* e = b->eng;
    DIC_HIP_TRY(hipMemcpyAsync(b->d_order_perm, b->h_perm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice,
                               e->stream));
    DIC_HIP_TRY(hipStreamSynchronize(e->stream));
    b->perm_w = w;
    b->perm_h = h;
    return DIC_OK;
}

}  // namespace

namespace {

// The solve's device work on the engine's stream (timed by the engine's
// events unless a graph is being recorded).
int solve_enqueue(Batch* b, const Image* f, const Image* g, const dic_solve_params* prm) {
    Engine* e = b->eng;
    if (!e->capturing) DIC_HIP_TRY(hipEventRecord(e->ev_sol0, e->stream));
    int st = prm->precision == DIC_PREC_PLAIN_F64 ? launch_icgn_plain(b, f, g, prm) : launch_icgn_batch(b, f, g, prm);
    if (st) return st;
    if (!e->capturing) {
        DIC_HIP_TRY(hipEventRecord(e->ev_sol1, e->stream));
        e->sol_recorded = true;
    }
    return DIC_OK;
}

}  // namespace

int dic_batch_solve(dic_batch* db, const dic_image* f, const dic_image* g, const dic_solve_params* prm) {
    CHECK_ARG(db && f && g && prm, "dic_batch_solve: NULL argument");
    Batch* b = &db->b;
    Engine* e = b->eng;
    const int vst = validate_solve(b, &f->im, &g->im, prm);
    if (vst) return vst;
    DIC_HIP_TRY(hipSetDevice(e->device));
    std::lock_guard<std::mutex> lk(e->mu);
    if (prm->want_history) {
        const int cap = std::max(prm->max_iterations, 1);
        if (b->hist_cap != cap) {
            dev_free(e, b->d_hist, sizeof(double) * 4 * b->n * b->hist_cap);
            b->d_hist = nullptr;
            b->hist_cap = 0;
            int st = dev_alloc(e, (void**)&b->d_hist, sizeof(double) * 4 * b->n * cap);
            if (st) return st;
            b->hist_cap = cap;
        }
        DIC_HIP_TRY(hipMemsetAsync(b->d_hist, 0, sizeof(double) * 4 * b->n * b->hist_cap, e->stream));
    }
    int st = refresh_order(b, f->im.w, f->im.h);
    if (st) return st;
    st = solve_enqueue(b, &f->im, &g->im, prm);
    if (st) return st;
    b->last_precision = prm->precision;
    b->last_method = prm->method;
    b->last_max_it = prm->max_iterations;
    b->last_history = prm->want_history != 0;
    b->arrays_valid = prm->precision == DIC_PREC_PLAIN_F64;
    return DIC_OK;
}

int dic_batch_solve_noalloc(dic_batch* db, const dic_image* f, const dic_image* g, const dic_solve_params* prm) {
    CHECK_ARG(db && f && g && prm, "dic_batch_solve_noalloc: NULL argument");
    Batch* b = &db->b;
    Engine* e = b->eng;
    const int vst = validate_solve(b, &f->im, &g->im, prm);
    if (vst) return vst;
    CHECK_ARG(prm->precision == DIC_PREC_PLAIN_F64, "invalid input: dic_batch_solve_noalloc runs DIC_PREC_PLAIN_F64");
    CHECK_ARG(!prm->want_history, "invalid input: dic_batch_solve_noalloc records no history");
    const size_t need = (size_t)b->total_px * plain_arrays(prm->method) * sizeof(double);
    if (need > kNoallocScratchBytes) {
        set_error("invalid input: dic_batch_solve_noalloc: the batch's per-pixel arrays need " + std::to_string(need) +
                  " bytes, beyond the fixed scratch of " + std::to_string(kNoallocScratchBytes) + " bytes");
        return DIC_ERR_INVALID_INPUT;
    }
    DIC_HIP_TRY(hipSetDevice(e->device));
    std::lock_guard<std::mutex> lk(e->mu);
    double* ptr = noalloc_scratch(e->device);
    if (!ptr) {
        set_error("dic_batch_solve_noalloc: the fixed scratch was not set up on this device");
        return DIC_ERR_DEVICE;
    }
    if (!e->capturing) DIC_HIP_TRY(hipEventRecord(e->ev_sol0, e->stream));
    int st = launch_icgn_plain(b, &f->im, &g->im, prm, nullptr, nullptr, 0, nullptr, nullptr, ptr,
                               kNoallocScratchBytes / sizeof(double));
    if (st) return st;
    if (!e->capturing) {
        DIC_HIP_TRY(hipEventRecord(e->ev_sol1, e->stream));
        e->sol_recorded = true;
    }
    b->last_precision = prm->precision;
    b->last_method = prm->method;
    b->last_max_it = prm->max_iterations;
    b->last_history = false;
    b->arrays_valid = false;  // the per-pixel arrays stay in the shared scratch
    return DIC_OK;
}

int dic_batch_autotune(dic_batch* db, const dic_image* f, const dic_image* g, const dic_solve_params* prm) {
    CHECK_ARG(db && f && g && prm, "dic_batch_autotune: NULL argument");
    Batch* b = &db->b;
    Engine* e = b->eng;
    CHECK_ARG(f->im.eng->device == e->device && g->im.eng->device == e->device,
              "invalid input: images live on another device");
    CHECK_ARG(f->im.h == g->im.h && f->im.w == g->im.w && f->im.border == g->im.border,
              "invalid input: reference and target images differ in shape or border");
    CHECK_ARG(prm->max_iterations >= 0 && prm->max_iterations <= 100000, "invalid input: max_iterations");
    DIC_HIP_TRY(hipSetDevice(e->device));
    std::lock_guard<std::mutex> lk(e->mu);
    int st = refresh_order(b, f->im.w, f->im.h);
    if (st) return st;
    dic_solve_params p = *prm;
    p.want_history = 0;
    p.precision = DIC_PREC_COMP_F32;
    p.method = DIC_METHOD_ICGN;
    st = autotune_icgn_batch(b, &f->im, &g->im, &p);
    if (st) return st;
    DIC_HIP_TRY(hipStreamSynchronize(e->stream));
    return DIC_OK;
}

int dic_batch_get_schedule(dic_batch* db, int32_t* wps, int32_t* slack) {
    CHECK_ARG(db && wps && slack, "dic_batch_get_schedule: NULL argument");
    Batch* b = &db->b;
    std::lock_guard<std::mutex> lk(b->eng->mu);
    const int o = b->max_order;
    for (int k = 0; k < 4; ++k) {
        const int base = 1 << k;
        if (b->waves_per_subset > 0) {
            wps[k] = b->waves_per_subset;
            slack[k] = b->win_slack;
        } else if (b->tuned_wps[o][base] > 0) {
            wps[k] = b->tuned_wps[o][base];
            slack[k] = b->tuned_slack[o][base];
        } else {
            wps[k] = base;
            slack[k] = b->win_slack;
        }
    }
    return DIC_OK;
}

int dic_batch_results(dic_batch* db, dic_subset_result* out) {
    CHECK_ARG(db && out, "dic_batch_results: NULL argument");
    Batch* b = &db->b;
    Engine* e = b->eng;
    DIC_HIP_TRY(hipSetDevice(e->device));
    std::lock_guard<std::mutex> lk(e->mu);
    DIC_HIP_TRY(hipMemcpyAsync(out, b->d_res, sizeof(dic_subset_result) * b->n, hipMemcpyDeviceToHost, e->stream));
    DIC_HIP_TRY(hipStreamSynchronize(e->stream));
    return DIC_OK;
}

int dic_batch_history(dic_batch* db, double* out) {
    CHECK_ARG(db && out, "dic_batch_history: NULL argument");
    Batch* b = &db->b;
    Engine* e = b->eng;
    DIC_HIP_TRY(hipSetDevice(e->device));
    std::lock_guard<std::mutex> lk(e->mu);
    if (!b->last_history || !b->d_hist) {
        set_error("not solved: no history was recorded (want_history = 0)");
        return DIC_ERR_NOT_SOLVED;
    }
    DIC_HIP_TRY(hipMemcpyAsync(out, b->d_hist, sizeof(double) * 4 * b->n * b->hist_cap, hipMemcpyDeviceToHost,
                               e->stream));
    DIC_HIP_TRY(hipStreamSynchronize(e->stream));
    return DIC_OK;
}

int dic_batch_subset_arrays(dic_batch* db, int64_t s, double* f, double* grad_f) {
    CHECK_ARG(db, "dic_batch_subset_arrays: NULL batch");
    Batch* b = &db->b;
    CHECK_ARG(s >= 0 && s < b->n, "invalid input: subset index out of range");
    Engine* e = b->eng;
    DIC_HIP_TRY(hipSetDevice(e->device));
    std::lock_guard<std::mutex> lk(e->mu);
    if (!b->arrays_valid || !b->d_scratch) {
        set_error("not solved: per-pixel arrays need a whole-batch DIC_PREC_PLAIN_F64 solve (dic_batch_solve) "
                  "as the last solve of the batch");
        return DIC_ERR_NOT_SOLVED;
    }
    const int t = b->h_tpl_index[s];
    const int64_t n = b->h_tpl_start[t + 1] - b->h_tpl_start[t];
    std::vector<double> buf(3 * n);
    const int64_t at = (int64_t)plain_arrays(b->last_method) * b->h_scr_off[s];
    DIC_HIP_TRY(hipMemcpyAsync(buf.data(), b->d_scratch + at, sizeof(double) * 3 * n,
                               hipMemcpyDeviceToHost, e->stream));
    DIC_HIP_TRY(hipStreamSynchronize(e->stream));
    if (f) std::memcpy(f, buf.data(), sizeof(double) * n);
    if (grad_f)
        for (int64_t i = 0; i < n; ++i) {
            grad_f[2 * i] = buf[n + i];
            grad_f[2 * i + 1] = buf[2 * n + i];
        }
    return DIC_OK;
}

int dic_batch_residual_map(dic_batch* db, const dic_image* f, const dic_image* g, const int32_t* idx, int64_t count,
                           const double* p, double* out) {
    CHECK_ARG(db && f && g, "dic_batch_residual_map: NULL argument");
    CHECK_ARG(count >= 0, "invalid input: negative list length");
    CHECK_ARG(count == 0 || (idx && out), "dic_batch_residual_map: NULL array");
    Batch* b = &db->b;
    Engine* e = b->eng;
    CHECK_ARG(f->im.eng->device == e->device && g->im.eng->device == e->device,
              "invalid input: images live on another device");
    CHECK_ARG(f->im.h == g->im.h && f->im.w == g->im.w && f->im.border == g->im.border,
              "invalid input: reference and target images differ in shape or border");
    std::vector<int64_t> off(count + 1, 0);
    for (int64_t k = 0; k < count; ++k) {
        CHECK_ARG(idx[k] >= 0 && idx[k] < b->n, "invalid input: subset index out of range");
        const int t = b->h_tpl_index[idx[k]];
        off[k + 1] = off[k] + (b->h_tpl_start[t + 1] - b->h_tpl_start[t]);
    }
    DIC_HIP_TRY(hipSetDevice(e->device));
    std::lock_guard<std::mutex> lk(e->mu);
    if (!p && b->last_precision < 0) {
        set_error("not solved: residual_map: the batch has no stored solution");
        return DIC_ERR_NOT_SOLVED;
    }
    if (count == 0) return DIC_OK;
    // launched grouped by warp order (one launch per order); each listed
    // subset's row stays where the list puts it
    std::vector<int32_t> pos(count);
    std::iota(pos.begin(), pos.end(), 0);
    if (b->mixed_orders)
        std::stable_sort(pos.begin(), pos.end(),
                         [&](int32_t a, int32_t c) { return b->h_orders[idx[a]] > b->h_orders[idx[c]]; });
    std::vector<int32_t> idx_s(count);
    std::vector<int64_t> off_s(count);
    std::vector<double> p_s(p ? 12 * (size_t)count : 0);
    for (int64_t q = 0; q < count; ++q) {
        idx_s[q] = idx[pos[q]];
        off_s[q] = off[pos[q]];
        if (p) std::memcpy(&p_s[12 * q], p + 12 * (int64_t)pos[q], sizeof(double) * 12);
    }
    // one device block: list, offsets, warps (when given), residuals
    const size_t list_b = ((size_t)count * 4 + 7) & ~(size_t)7;
    const size_t off_b = sizeof(int64_t) * (size_t)count;
    const size_t p_b = p ? sizeof(double) * 12 * (size_t)count : 0;
    const size_t out_b = sizeof(double) * (size_t)off[count];
    char* d = nullptr;
    int st = dev_alloc(e, (void**)&d, list_b + off_b + p_b + out_b);
    if (st) return st;
    auto fail = [&](const char* what) {
        set_error(std::string("dic_batch_residual_map: ") + what);
        (void)hipStreamSynchronize(e->stream);
        dev_free(e, d, list_b + off_b + p_b + out_b);
        return DIC_ERR_DEVICE;
    };
    if (hipMemcpyAsync(d, idx_s.data(), sizeof(int32_t) * count, hipMemcpyHostToDevice, e->stream) != hipSuccess ||
        hipMemcpyAsync(d + list_b, off_s.data(), off_b, hipMemcpyHostToDevice, e->stream) != hipSuccess ||
        (p && hipMemcpyAsync(d + list_b + off_b, p_s.data(), p_b, hipMemcpyHostToDevice, e->stream) != hipSuccess))
        return fail("upload failed");
    const int32_t* d_list = reinterpret_cast<const int32_t*>(d);
    const int64_t* d_off = reinterpret_cast<const int64_t*>(d + list_b);
    const double* d_pl = reinterpret_cast<const double*>(d + list_b + off_b);
    double* d_out = reinterpret_cast<double*>(d + list_b + off_b + p_b);
    for (const OrderRun& r : order_runs(b, idx_s.data(), count)) {
        st = launch_residual(b, &f->im, &g->im, r.order, d_list + r.start, r.count, p ? d_pl + 12 * r.start : b->d_p,
                             p == nullptr, d_off + r.start, d_out);
        if (st) break;
    }
    if (st) {
        (void)hipStreamSynchronize(e->stream);
        dev_free(e, d, list_b + off_b + p_b + out_b);
        return st;
    }
    if (hipMemcpyAsync(out, d + list_b + off_b + p_b, out_b, hipMemcpyDeviceToHost, e->stream) != hipSuccess ||
        hipStreamSynchronize(e->stream) != hipSuccess)
        return fail("download failed");
    dev_free(e, d, list_b + off_b + p_b + out_b);
    return DIC_OK;
}

int dic_batch_device_p(dic_batch* db, const double** out) {
    CHECK_ARG(db && out, "dic_batch_device_p: NULL argument");
    *out = db->b.d_p;
    return DIC_OK;
}

void dic_batch_destroy(dic_batch* db) {
    if (!db) return;
    Batch* b = &db->b;
    (void)hipSetDevice(b->eng->device);
    {
        std::lock_guard<std::mutex> lk(b->eng->mu);
        (void)hipStreamSynchronize(b->eng->stream);
        batch_free(b);
    }
    delete db;
}

int dic_image_stage(dic_image* di, const void* pixels, int pix_type, int mem) {
    CHECK_ARG(di && pixels, "dic_image_stage: NULL argument");
    CHECK_ARG(pixel_bytes(pix_type) > 0, "invalid input: pixel type");
    CHECK_ARG(mem == DIC_MEM_HOST || mem == DIC_MEM_DEVICE, "invalid input: memory kind");
    Image* im = &di->im;
    CHECK_ARG(!loader_pixels(pix_type) || (im->h >= 3 && im->w >= 3), "invalid input: a loaded image must be at least 3x3");
    Engine* e = im->eng;
    DIC_HIP_TRY(hipSetDevice(e->device));
    std::lock_guard<std::mutex> lk(e->mu);
    const size_t bytes = (size_t)im->h * im->w * pixel_bytes(pix_type);
    // a frame graph may hold this buffer: its type (and so its size) is
    // fixed by the first staging
    CHECK_ARG(!im->pix_dev || im->pix_dev_type == pix_type,
              "invalid input: staged pixel type differs from the first staging");
    if (!im->pix_dev) {
        int st = dev_alloc(e, &im->pix_dev, bytes);
        if (st) return st;
        im->pix_dev_bytes = bytes;
        im->pix_dev_type = pix_type;
    }
    DIC_HIP_TRY(hipMemcpyAsync(im->pix_dev, pixels, bytes,
                               mem == DIC_MEM_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, e->stream));
    if (mem == DIC_MEM_HOST) DIC_HIP_TRY(hipStreamSynchronize(e->stream));
    return DIC_OK;
}

int dic_frame_graph_create(dic_batch* db, const dic_image* f, dic_image* g, int warm_start,
                           const dic_solve_params* prm, dic_frame_graph** out) {
    CHECK_ARG(db && f && g && prm && out, "dic_frame_graph_create: NULL argument");
    *out = nullptr;
    Batch* b = &db->b;
    Engine* e = b->eng;
    CHECK_ARG(f->im.eng == e && g->im.eng == e, "invalid input: images belong to another engine");
    CHECK_ARG(f->im.h == g->im.h && f->im.w == g->im.w && f->im.border == g->im.border,
              "invalid input: reference and target images differ in shape or border");
    CHECK_ARG(g->im.pix_dev != nullptr, "invalid input: stage the target's pixels (dic_image_stage) first");
    CHECK_ARG(prm->precision == DIC_PREC_COMP_F32 && prm->method == DIC_METHOD_ICGN,
              "invalid input: frame graphs run the batched engine (DIC_PREC_COMP_F32, ICGN)");
    CHECK_ARG(!prm->want_history, "invalid input: frame graphs record no history");
    CHECK_ARG(prm->max_iterations >= 0 && prm->max_iterations <= 100000, "invalid input: max_iterations");
    DIC_HIP_TRY(hipSetDevice(e->device));
    std::lock_guard<std::mutex> lk(e->mu);
    Image* gi = &g->im;
    // everything that allocates or synchronises happens before the capture
    int st = refresh_order(b, f->im.w, f->im.h);
    if (st) return st;
    dic_frame_graph* gr = new dic_frame_graph();
    gr->e = e;
    gr->tmp_bytes = (size_t)gi->hp * gi->wp * sizeof(double);
    st = dev_alloc(e, (void**)&gr->tmp, gr->tmp_bytes);
    if (!st) st = prefilter_run(e, gi, gi->pix_dev, gi->pix_dev_type, DIC_MEM_DEVICE, gr->tmp);  // twiddles, attributes
    if (!st) st = solve_enqueue(b, &f->im, gi, prm);
    if (!st && hipStreamSynchronize(e->stream) != hipSuccess) st = DIC_ERR_DEVICE;
    if (!st) {
        if (hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            set_error("dic_frame_graph_create: stream capture could not start");
            st = DIC_ERR_DEVICE;
        } else {
            e->capturing = true;
            st = prefilter_run(e, gi, gi->pix_dev, gi->pix_dev_type, DIC_MEM_DEVICE, gr->tmp);
            if (!st && warm_start &&
                hipMemcpyAsync(b->d_p0, b->d_p, sizeof(double) * 12 * b->n, hipMemcpyDeviceToDevice, e->stream) !=
                    hipSuccess)
                st = DIC_ERR_DEVICE;
            if (!st) st = solve_enqueue(b, &f->im, gi, prm);
            e->capturing = false;
            hipGraph_t graph = nullptr;
            const hipError_t ce = hipStreamEndCapture(e->stream, &graph);
            if (!st && (ce != hipSuccess || !graph)) {
                set_error("dic_frame_graph_create: stream capture failed");
                st = DIC_ERR_DEVICE;
            }
            gr->graph = graph;
            if (!st && hipGraphInstantiate(&gr->exec, gr->graph, nullptr, nullptr, 0) != hipSuccess) {
                set_error("dic_frame_graph_create: graph instantiation failed");
                st = DIC_ERR_DEVICE;
            }
        }
    }
    if (st) {
        if (gr->exec) (void)hipGraphExecDestroy(gr->exec);
        if (gr->graph) (void)hipGraphDestroy(gr->graph);
        dev_free(e, gr->tmp, gr->tmp_bytes);
        delete gr;
        return st;
    }
    *out = gr;
    return DIC_OK;
}

int dic_frame_graph_launch(dic_frame_graph* gr) {
    CHECK_ARG(gr, "dic_frame_graph_launch: NULL graph");
    Engine* e = gr->e;
    DIC_HIP_TRY(hipSetDevice(e->device));
    std::lock_guard<std::mutex> lk(e->mu);
    DIC_HIP_TRY(hipGraphLaunch(gr->exec, e->stream));
    return DIC_OK;
}

void dic_frame_graph_destroy(dic_frame_graph* gr) {
    if (!gr) return;
    Engine* e = gr->e;
    (void)hipSetDevice(e->device);
    (void)hipStreamSynchronize(e->stream);
    if (gr->exec) (void)hipGraphExecDestroy(gr->exec);
    if (gr->graph) (void)hipGraphDestroy(gr->graph);
    dev_free(e, gr->tmp, gr->tmp_bytes);
    delete gr;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// list solves: some subsets of a batch, each from its own initial warp
// ---------------------------------------------------------------------------
static_assert(sizeof(dic_subset_result) % 8 == 0, "result record must be whole 8-byte words");

namespace geodic {

int solve_list(Batch* b, const Image* f, const Image* g, const dic_solve_params* prm, const int32_t* idx,
               int64_t count, const double* p0, dic_subset_result* out) {
    Engine* e = b->eng;
    if (count == 0) return DIC_OK;
    // every index in range and listed once (two workgroups must not share a subset)
    if (b->list_mark.size() != (size_t)b->n) b->list_mark.assign(b->n, 0);
    if (++b->list_stamp == 0) {
        std::fill(b->list_mark.begin(), b->list_mark.end(), 0u);
        b->list_stamp = 1;
    }
    for (int64_t k = 0; k < count; ++k) {
        CHECK_ARG(idx[k] >= 0 && idx[k] < b->n, "invalid input: list index out of range");
        CHECK_ARG(b->list_mark[idx[k]] != b->list_stamp, "invalid input: a subset is listed twice");
        b->list_mark[idx[k]] = b->list_stamp;
    }
    // staging in pinned, device-mapped host memory: the kernels read the
    // list and its warps there and write the results there, so a list
    // costs its launches and one synchronisation, no copies
    if (count > b->list_cap) {
        int64_t cap = 256;
        while (cap < count) cap *= 2;
        cap = std::min<int64_t>(cap, b->n);
        // a list's launch may still be retiring after its flags went up
        if (b->h_list_io) {
            DIC_HIP_TRY(hipStreamSynchronize(e->stream));
            (void)hipHostFree(b->h_list_io);
        }
        b->h_list_io = nullptr;
        b->d_list_io = nullptr;
        b->list_cap = 0;
        DIC_HIP_TRY(hipHostMalloc((void**)&b->h_list_io,
                                  list_io_bytes(cap) + sizeof(dic_subset_result) * cap + list_ints_bytes(cap)));
        DIC_HIP_TRY(hipHostGetDevicePointer((void**)&b->d_list_io, b->h_list_io, 0));
        b->list_cap = cap;
    }
    const size_t ib = list_ints_bytes(b->list_cap);
    const size_t res_off = list_io_bytes(b->list_cap);
    int32_t* h_idx = reinterpret_cast<int32_t*>(b->h_list_io);
    int32_t* h_ord = reinterpret_cast<int32_t*>(b->h_list_io + ib);
    double* h_p0 = reinterpret_cast<double*>(b->h_list_io + 2 * ib);
    const dic_subset_result* h_res = reinterpret_cast<const dic_subset_result*>(b->h_list_io + res_off);
    const int32_t* d_idx = reinterpret_cast<const int32_t*>(b->d_list_io);
    const int32_t* d_ord = reinterpret_cast<const int32_t*>(b->d_list_io + ib);
    const double* d_src = reinterpret_cast<const double*>(b->d_list_io + 2 * ib);
    dic_subset_result* d_res = reinterpret_cast<dic_subset_result*>(b->d_list_io + res_off);
    // the batched engine's done flags, one per dispatch position
    const size_t done_off = res_off + sizeof(dic_subset_result) * (size_t)b->list_cap;
    int32_t* h_done = reinterpret_cast<int32_t*>(b->h_list_io + done_off);
    int32_t* d_done = reinterpret_cast<int32_t*>(b->d_list_io + done_off);
    // the kernels read each listed subset's warp from the staging and write
    // its result record there, in their dispatch order: the list's own order
    // within each warp order for the plain kernel (one launch per order),
    // by warp order and then template size (largest first) for the batched
    // engine, whose size classes need that
    int st = DIC_OK;
    std::vector<int32_t>& pos = b->list_pos;
    pos.resize(count);
    for (int64_t k = 0; k < count; ++k) pos[k] = (int32_t)k;
    const std::vector<int32_t>& ord = b->h_orders;
    if (prm->precision == DIC_PREC_PLAIN_F64) {
        if (b->mixed_orders)
            std::stable_sort(pos.begin(), pos.end(), [&](int32_t a, int32_t c) { return ord[idx[a]] > ord[idx[c]]; });
        for (int64_t q = 0; q < count; ++q) {
            h_idx[q] = idx[pos[q]];
            std::memcpy(h_p0 + 12 * q, p0 + 12 * (int64_t)pos[q], sizeof(double) * 12);
        }
        // the listed subsets' per-pixel arrays overwrite those of the
        // stored whole-batch results
        b->arrays_valid = false;
        st = launch_icgn_plain(b, f, g, prm, d_idx, h_idx, count, d_src, d_res);
        if (st) {
            (void)hipStreamSynchronize(e->stream);  // nothing may still write the staging
            return st;
        }
        DIC_HIP_TRY(hipStreamSynchronize(e->stream));
        for (int64_t q = 0; q < count; ++q) out[pos[q]] = h_res[q];
    } else {
        auto npx = [&](int32_t k) {
            const int t = b->h_tpl_index[idx[k]];
            return b->h_tpl_start[t + 1] - b->h_tpl_start[t];
        };
        std::stable_sort(pos.begin(), pos.end(), [&](int32_t a, int32_t c) {
            if (ord[idx[a]] != ord[idx[c]]) return ord[idx[a]] > ord[idx[c]];
            return npx(a) > npx(c);
        });
        for (int64_t q = 0; q < count; ++q) {
            h_ord[q] = idx[pos[q]];
            std::memcpy(h_p0 + 12 * q, p0 + 12 * (int64_t)pos[q], sizeof(double) * 12);
        }
        for (int64_t q = 0; q < count; ++q) __atomic_store_n(&h_done[q], 0, __ATOMIC_RELAXED);
        st = launch_icgn_batch(b, f, g, prm, d_ord, h_ord, count, true, d_src, d_res, d_done);
        if (st) {
            (void)hipStreamSynchronize(e->stream);  // nothing may still write the staging
            return st;
        }
        // every workgroup raises its flag after its record, and reads the
        // staging only before that: once all flags are up the records are
        // complete and the staging is free, though the launch may not have
        // retired yet (what the engine enqueues next waits for it on the
        // stream).  Polling them sees the list end sooner than a stream
        // synchronisation does; past a time limit, or on an error, the
        // stream's own synchronisation decides.
        {
            const auto t0 = std::chrono::steady_clock::now();
            int64_t q = 0;
            for (uint64_t spin = 1;; ++spin) {
                while (q < count && __atomic_load_n(&h_done[q], __ATOMIC_ACQUIRE) != 0) ++q;
                if (q == count) break;
                if ((spin & 4095) == 0 &&
                    std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(200))
                    break;
            }
            if (q < count) DIC_HIP_TRY(hipStreamSynchronize(e->stream));
        }
        for (int64_t q = 0; q < count; ++q) out[pos[q]] = h_res[q];
    }
    // the batch's stored result records, history and their solve settings
    // are the last whole-batch solve's: a list leaves them as they were
    return DIC_OK;
}

}  // namespace geodic

extern "C" {

int dic_batch_solve_list(dic_batch* db, const dic_image* f, const dic_image* g, const dic_solve_params* prm,
                         const int32_t* idx, int64_t count, const double* p0, dic_subset_result* out) {
    CHECK_ARG(db && f && g && prm, "dic_batch_solve_list: NULL argument");
    CHECK_ARG(count >= 0, "invalid input: negative list length");
    CHECK_ARG(count == 0 || (idx && p0 && out), "dic_batch_solve_list: NULL array");
    Batch* b = &db->b;
    Engine* e = b->eng;
    const int vst = validate_solve(b, &f->im, &g->im, prm);
    if (vst) return vst;
    CHECK_ARG(!prm->want_history, "invalid input: list solves record no history");
    DIC_HIP_TRY(hipSetDevice(e->device));
    std::lock_guard<std::mutex> lk(e->mu);
    return solve_list(b, &f->im, &g->im, prm, idx, count, p0, out);
}

}  // extern "C"
