// Internal declarations shared by the engine's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/geodic.h"

namespace geodic {

// Thread-local message for dic_last_error().
void set_error(const std::string& msg);

// Argument check of an entry point: records msg and returns
// DIC_ERR_INVALID_INPUT when cond fails.
#define CHECK_ARG(cond, msg)                  \
    do {                                      \
        if (!(cond)) {                        \
            ::geodic::set_error(std::string(msg)); \
            return DIC_ERR_INVALID_INPUT;     \
        }                                     \
    } while (0)
#define DIC_HIP_TRY(expr)                                                                   \
    do {                                                                                    \
        hipError_t _e = (expr);                                                             \
        if (_e != hipSuccess) {                                                             \
            ::geodic::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));         \
            return (_e == hipErrorOutOfMemory) ? DIC_ERR_OUT_OF_MEMORY : DIC_ERR_DEVICE;    \
        }                                                                                   \
    } while (0)

struct Engine {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_pre0 = nullptr, ev_pre1 = nullptr, ev_sol0 = nullptr, ev_sol1 = nullptr;
    static constexpr int kAux = 3;
    hipStream_t aux[kAux] = {nullptr, nullptr, nullptr};  // size classes of a batch run here too
    hipEvent_t ev_fork = nullptr, ev_join[kAux] = {nullptr, nullptr, nullptr};
    bool pre_recorded = false, sol_recorded = false;
    std::atomic<int64_t> bytes_in_use{0};  // updated by allocations and frees of any thread
    int num_cus = 256;
    bool capturing = false;  // recording a graph: no timing events, no host synchronisation
    std::mutex mu;
    // cached twiddle tables per FFT length (device, complex double)
    std::vector<std::pair<int, double2*>> twiddles;
    // cached circular-FIR taps per line length (device)
    std::vector<std::pair<int, double*>> fir_taps;
    int prefilter_method = DIC_PREFILTER_FIR;
    // prefilter intermediate (row-pass output), grown on demand
    double* pre_tmp = nullptr;
    size_t pre_tmp_bytes = 0;
    // pixel staging for host uploads
    void* pix_stage = nullptr;
    size_t pix_stage_bytes = 0;
};

struct Image {
    Engine* eng = nullptr;
    int64_t h = 0, w = 0;
    int border = 0;
    int64_t hp = 0, wp = 0;  // padded shape
    int64_t wp32 = 0;        // row pitch of the FP32 copy: wp rounded up to 4 floats (16-byte rows)
    double* coef64 = nullptr;
    float* coef32 = nullptr;
    void* pix_dev = nullptr;  // staged frame pixels (dic_image_stage), h x w of pix_type
    size_t pix_dev_bytes = 0;
    int pix_dev_type = -1;
    uint8_t* grey_dev = nullptr;  // the file loader's converted frame (h x w), loader pixel types only
    // f, ∂x f, ∂y f at every pixel (h x w float4), made by the first
    // batched solve that reads the image as a reference and remade by every
    // prefilter of it from then on
    float4* knot = nullptr;
    bool knot_wanted = false;
    bool knot_valid = false;
};

// Bytes per pixel of a pixel type; 0 for an unknown type.
inline size_t pixel_bytes(int pix_type) {
    switch (pix_type) {
        case DIC_PIX_U8: return 1;
        case DIC_PIX_F64: return 8;
        case DIC_PIX_RGB8_LOAD: return 3;
        case DIC_PIX_GREY8_LOAD: return 1;
        default: return 0;
    }
}
inline bool loader_pixels(int pix_type) { return pix_type == DIC_PIX_RGB8_LOAD || pix_type == DIC_PIX_GREY8_LOAD; }

struct Batch {
    Engine* eng = nullptr;
    int64_t n = 0;
    int32_t n_tpl = 0;
    int64_t total_px = 0;     // Σ n_px over subsets (scratch sizing)
    int32_t max_n = 0;        // largest template
    int32_t max_order = 0;    // highest warp order in the batch
    bool mixed_orders = false;  // the subsets do not all share one warp order
    // device arrays
    double* d_coords = nullptr;     // n x 2
    int32_t* d_orders = nullptr;    // n
    int32_t* d_tpl_index = nullptr; // n
    int64_t* d_tpl_start = nullptr; // n_tpl + 1
    int32_t* d_tpl_xy = nullptr;    // packed (dx & 0xffff) | (dy << 16), reference scan order
    int32_t* d_tpl_xy_eng = nullptr;  // the same pixels in LDS-bank-friendly order, as half-precision pairs (batched engine)
    int32_t* d_tpl_ext = nullptr;   // n_tpl x 4: dxmin, dxmax, dymin, dymax
    void* d_meta = nullptr;         // n x 48 B: per-subset start record of the batched engine
    int32_t max_side = 0;           // largest template box side
    int32_t win_slack = 3;          // deformation slack of the LDS window (pixels per side)
    int32_t waves_per_subset = 0;   // batched engine: wavefronts per subset (0 = smallest that fits)
    // autotuned wavefronts and slack by warp order and a class's fewest wavefronts (0 = none)
    int32_t tuned_wps[3][9] = {};
    int32_t tuned_slack[3][9] = {};
    int64_t* d_scr_off = nullptr;   // n: h_scr_off on the device
    double* d_p0 = nullptr;         // n x 12
    double* d_p = nullptr;          // n x 12 final warps
    dic_subset_result* d_res = nullptr;
    double* d_hist = nullptr;       // n x max_it x 4 (allocated on demand)
    int32_t hist_cap = 0;
    double* d_scratch = nullptr;    // plain path per-pixel arrays
    size_t scratch_doubles = 0;
    int32_t* d_order_perm = nullptr;  // dispatch order: warp order descending, template size descending, then
                                      // nearest the image edge first
    int32_t* d_by_order = nullptr;    // mixed orders: the subsets grouped by warp order, ascending index
    int64_t by_order_off[4] = {0, 0, 0, 0};  // group o is d_by_order[by_order_off[o], by_order_off[o + 1])
    std::vector<int32_t> h_orders;    // warp order per subset
    std::vector<double> h_coords;       // host copies for the dispatch order
    std::vector<int32_t> h_ext;
    std::vector<int32_t> h_perm;
    int64_t perm_w = -1, perm_h = -1;   // image shape the order was made for
    // host copies
    std::vector<int64_t> h_tpl_start;
    std::vector<int32_t> h_tpl_index;
    std::vector<int64_t> h_scr_off;   // plain path: Σ n_px of the subsets before each (pixels; a
                                      // subset's arrays start at that times the arrays per pixel)
    bool arrays_valid = false;        // d_scratch holds every subset's f and grad f of the stored results
    // list solves (dic_batch_solve_list): pinned, device-mapped staging of
    // the list, its dispatch order, the initial warps and the results,
    // grown on demand to list_cap subsets
    char* h_list_io = nullptr;
    char* d_list_io = nullptr;                // the same memory as the device sees it
    int64_t list_cap = 0;
    std::vector<uint32_t> list_mark;          // duplicate check
    std::vector<int32_t> list_pos;            // list positions in dispatch order
    uint32_t list_stamp = 0;
    int last_precision = -1;
    int last_method = -1;
    int last_max_it = 0;
    bool last_history = false;
};

// prefilter.hip
// tmp_own: a caller-owned row-pass buffer (hp x wp doubles) instead of the
// engine's, for work captured into graphs
int prefilter_run(Engine* e, Image* img, const void* pix, int pix_type, int mem, double* tmp_own = nullptr);
// icgn_plain.hip
// list: device array of `count` subset indices to solve (all subsets when
// null); p0_list / res_list (with a list): the listed subsets' initial warps
// (12 each) and result records in list order, device-visible (the batch's
// own p0 and results are then neither read nor written for them)
// ext_scratch: a caller-provided per-pixel scratch of ext_doubles doubles
// instead of the batch's own (grown on demand)
// h_list: the same list as the host reads it; a list that mixes warp
// orders must be grouped by order (each run is one launch)
int launch_icgn_plain(Batch* b, const Image* f, const Image* g, const dic_solve_params* prm,
                      const int32_t* list = nullptr, const int32_t* h_list = nullptr, int64_t count = 0,
                      const double* p0_list = nullptr, dic_subset_result* res_list = nullptr,
                      double* ext_scratch = nullptr, size_t ext_doubles = 0);
// Subset::residual_map for the `count` listed subsets (device list, all of
// warp order `order`), at the warps d_p (12 per listed subset, or per
// subset of the batch when p_by_subset), written to d_out from d_off[k] for
// listed subset k.
int launch_residual(Batch* b, const Image* f, const Image* g, int order, const int32_t* d_list, int64_t count,
                    const double* d_p, bool p_by_subset, const int64_t* d_off, double* d_out);
// Runs of equal warp order in the host list h[0, count): (start, length, order).
struct OrderRun {
    int64_t start, count;
    int order;
};
std::vector<OrderRun> order_runs(const Batch* b, const int32_t* h, int64_t count);
// per-pixel doubles of the plain kernel's scratch: f, grad f, g (+ grad g for FAGN)
inline int plain_arrays(int method) { return method == DIC_METHOD_FAGN ? 6 : 4; }
// The library's fixed device scratch for dic_batch_solve_noalloc: its
// device address on the current device, looked up once per device (at
// engine creation); kNoallocScratchBytes is its size.
constexpr size_t kNoallocScratchBytes = (size_t)64 << 20;
int noalloc_scratch_init(int device);
double* noalloc_scratch(int device);
// icgn_batch.hip
struct SizeClass {
    int64_t start, count;  // run of the dispatch order
    int order;             // warp order of the class's subsets
    int base_wps;          // fewest wavefronts that hold the class's subsets
    int wps;               // wavefronts per subset used
    int slack;             // LDS window slack used
    int64_t max_n;
    int max_side;
};
// classes of the n subsets h_perm lists (sorted by size); the batch's own
// dispatch order when h_perm is null
std::vector<SizeClass> size_classes(const Batch* b, const int32_t* h_perm = nullptr, int64_t n = 0);
// d_perm/h_perm/count: a dispatch order over some of the batch's subsets
// (device and host copies); the whole batch in its own order when null
// latency: a small list; each class takes more wavefronts per subset
// while the whole list still fits on the chip at two wavefronts per SIMD.
// p0_list / res_list: as for launch_icgn_plain, in the dispatch order
int launch_icgn_batch(Batch* b, const Image* f, const Image* g, const dic_solve_params* prm,
                      const int32_t* d_perm = nullptr, const int32_t* h_perm = nullptr, int64_t count = 0,
                      bool latency = false, const double* p0_list = nullptr, dic_subset_result* res_list = nullptr,
                      int32_t* done_list = nullptr);
// the image's knot table: knot_table_ensure makes it when it is missing or
// stale (allocating the first time, except under graph capture, where a
// missing table leaves the solves on the window path); knot_table_run
// recomputes it from the FP32 coefficients on the engine's stream
int knot_table_ensure(Engine* e, Image* im);
int knot_table_run(Engine* e, Image* im);
// times the candidate schedules of every size class and keeps the fastest
int autotune_icgn_batch(Batch* b, const Image* f, const Image* g, const dic_solve_params* prm);

// api.hip: solves the listed subsets (each at most once) from p0 (count x
// 12, host) and copies their results, in list order, to out (host).
// Synchronous.  The caller holds the engine's lock.
int solve_list(Batch* b, const Image* f, const Image* g, const dic_solve_params* prm, const int32_t* idx,
               int64_t count, const double* p0, dic_subset_result* out);

// api.hip: argument checks shared by the solve entry points
int validate_solve(const Batch* b, const Image* f, const Image* g, const dic_solve_params* prm);

// device memory accounting
int dev_alloc(Engine* e, void** p, size_t bytes);
void dev_free(Engine* e, void* p, size_t bytes);

}  // namespace geodic

// The opaque handles of the C ABI.
struct dic_engine {
    geodic::Engine e;
};
struct dic_image {
    geodic::Image im;
};
struct dic_batch {
    geodic::Batch b;
};
