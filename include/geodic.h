/* geodic.h — C ABI of the MI355X DIC subset-correlation engine.
 *
 * Plain pointers and sizes only; no C++ or framework types cross this
 * boundary and no call unwinds.  Every entry point returns a dic_status
 * (0 = success); dic_last_error() gives the message of the last failure on
 * the calling thread.
 *
 * The engine replaces the FP64 CPU hot path of geopyv-dev (Rust crate
 * `geopyv-dev` 0.2.4).  Each declaration names the reference interface it
 * stands in for (path:line in that repository).  The PyO3 module
 * `_geopyv_dev` is what binds those interfaces today; the Python mirror in
 * geopyv-dev_amd/geopyv_dev binds these symbols with ctypes instead.
 */
#ifndef GEODIC_H
#define GEODIC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes: src/error.rs:5-40 one-to-one, then the two panics of
 *      subset.rs (1143 Hessian, 520/595 W_delta) and device failures ---- */
typedef enum {
    DIC_OK = 0,
    DIC_ERR_IMAGE_LOAD = 1,          /* Error::ImageLoad */
    DIC_ERR_FILE_NOT_FOUND = 2,      /* Error::FileNotFound */
    DIC_ERR_INVALID_INPUT = 3,       /* Error::InvalidInput */
    DIC_ERR_MESH_GENERATION = 4,     /* Error::MeshGeneration */
    DIC_ERR_IO = 5,                  /* Error::Io */
    DIC_ERR_INVALID_MAGIC = 6,       /* Error::InvalidMagic */
    DIC_ERR_UNSUPPORTED_VERSION = 7, /* Error::UnsupportedVersion */
    DIC_ERR_NOT_SOLVED = 8,          /* Error::NotSolved */
    DIC_ERR_CANCELLED = 9,           /* Error::Cancelled */
    DIC_ERR_SINGULAR_HESSIAN = 10,   /* panic "ICGN Hessian singular", subset.rs:1143 */
    DIC_ERR_SINGULAR_WARP = 11,      /* panic "W_delta singular", subset.rs:520,595 */
    DIC_ERR_DEVICE = 20,             /* HIP runtime failure */
    DIC_ERR_OUT_OF_MEMORY = 21,      /* device allocation or HBM budget exceeded */
    DIC_ERR_BUSY = 22,               /* serving queue full */
    DIC_PENDING = 23                 /* dic_poll: the job has not finished (not an error) */
} dic_status;

/* Pixel element types for image upload.  U8 and F64 are grey images taken
 * as they are (Image::from_array, image.rs:83-86).  RGB8_LOAD (h, w, 3
 * interleaved) and GREY8_LOAD (h, w) are decoded file pixels: the file
 * loader's conversion (load_grayscale + gaussian_blur_5x5, image.rs:450-527:
 * BT.709 integer grey (2126 R + 7152 G + 722 B + 5000) / 10000, a 5x5
 * Gaussian with sigma 1.1 and reflect-101 borders in FP64, rounded and
 * clamped to 0-255) runs on the device before the prefilter, as
 * Image::from_file (image.rs:69-77) does; these need h, w >= 3. */
enum { DIC_PIX_U8 = 0, DIC_PIX_F64 = 1, DIC_PIX_RGB8_LOAD = 2, DIC_PIX_GREY8_LOAD = 3 };
/* Where a buffer lives. */
enum { DIC_MEM_HOST = 0, DIC_MEM_DEVICE = 1 };
/* Local mask shapes: masks.rs:21-31 MaskShape. */
enum { DIC_SHAPE_CIRCLE = 0, DIC_SHAPE_SQUARE = 1, DIC_SHAPE_SEMICIRCLE = 2 };
/* Per-node kernel: mesh.rs:285 SolveMethod. */
enum { DIC_METHOD_ICGN = 0, DIC_METHOD_FAGN = 1 };
/* Engine precision.  PLAIN_F64 is the drop-in (FP64 throughout, the
 * reference's pass structure); COMP_F32 is the batched engine (FP32
 * interpolation from an LDS-staged FP32 coefficient window, FP64
 * accumulation, FP64 Cholesky update). */
enum { DIC_PREC_PLAIN_F64 = 0, DIC_PREC_COMP_F32 = 1 };

typedef struct dic_engine dic_engine;
typedef struct dic_image dic_image;
typedef struct dic_batch dic_batch;

/* Solve settings: the arguments of Subset::solve_icgn / solve_fagn
 * (subset.rs:992-1022), plus the engine choice. */
typedef struct {
    int32_t method;          /* DIC_METHOD_* */
    int32_t precision;       /* DIC_PREC_* */
    int32_t max_iterations;  /* `max_iterations` */
    int32_t want_history;    /* record (iteration, norm, zncc, znssd) per iteration */
    double tolerance;        /* `tolerance`: quality_ok needs c_zncc >= tolerance */
    double max_norm;         /* `max_norm`: converged when ||dp|| < max_norm */
} dic_solve_params;

/* One subset's outcome: SolveResult (subset.rs:146-191) plus the reference
 * quantities Subset::new stores (subset.rs:885-899). */
typedef struct {
    double p[12];            /* warp, p_len = 2/6/12 entries used */
    double c_zncc, c_znssd;
    double norm;             /* last convergence norm (NaN if no iteration) */
    double eta_u, eta_v;     /* eta_omitted (subset.rs:748-820) */
    double f_m, delta_f, sssig, sigma;
    int32_t iterations, converged, quality_ok, status;
    int32_t n_px, order;
} dic_subset_result;

/* ---- engine ---- */
/* Creates the engine on HIP device `device` with its own stream. */
int dic_engine_create(int device, dic_engine** out);
void dic_engine_destroy(dic_engine* e);
int dic_engine_sync(dic_engine* e);
/* How images are prefiltered (get_c, image.rs:252-327): both apply the
 * spectral division by K(w) of every padded row, then every column.
 * DIC_PREFILTER_FIR (the default) as a circular convolution with the inverse
 * DFT of 1/K (105 taps; exact for lines shorter than that), DIC_PREFILTER_FFT
 * by forward transform, division and inverse transform of each line, as the
 * reference computes it.  Results agree to ~1e-13 of a grey level.  No
 * reference counterpart: the choice of evaluation is this library's. */
#define DIC_PREFILTER_FIR 0
#define DIC_PREFILTER_FFT 1
int dic_engine_set_prefilter(dic_engine* e, int method);
/* Kernel time of the last prefilter and the last batch solve, measured with
 * HIP events on the engine's stream (milliseconds). */
int dic_engine_timing(dic_engine* e, double* prefilter_ms, double* solve_ms);
/* Thread-local message of the last failure ("" if none). */
const char* dic_last_error(void);
/* Device bytes currently held by the engine's images and batches. */
int64_t dic_engine_bytes_in_use(dic_engine* e);

/* ---- images: Image::from_array (image.rs:83-86) -> get_c (image.rs:252-327) ----
 * `pixels` is (h, w) row-major, x = column, y = row, values 0-255.  The
 * coefficient array C is (h+2b, w+2b) FP64 on the device, with an FP32 copy
 * for the batched engine.  border must be >= 3 (image.rs:52). */
int dic_image_create(dic_engine* e, const void* pixels, int pix_type, int mem, int64_t h, int64_t w,
                     int border, dic_image** out);
/* New frame of the same shape into an existing image (Sequence::solve's
 * per-frame Image::from_file, sequence.rs:914): reuses its device buffers. */
int dic_image_update(dic_image* img, const void* pixels, int pix_type, int mem);
/* Copies a frame's pixels (host or device memory, the image's shape) into
 * the image's own device pixel buffer, the source a frame graph prefilters;
 * the pixel type is fixed by the first call.  Host copies complete before
 * return.  Per-frame ingest of sequence.rs:693-694,914 (the decode stays on
 * the host). */
int dic_image_stage(dic_image* img, const void* pixels, int pix_type, int mem);
/* The grey pixels an image loaded from file pixels (DIC_PIX_*_LOAD) was
 * prefiltered from, h*w doubles (Image.image_gs, python/src/py_image.rs);
 * DIC_ERR_INVALID_INPUT for an image given as grey pixels. */
int dic_image_grey(const dic_image* img, double* out);
/* Copies C back to the host: (h+2b)*(w+2b) doubles (Image.coef, image.rs:51). */
int dic_image_coef(const dic_image* img, double* out);
int dic_image_info(const dic_image* img, int64_t* h, int64_t* w, int* border);
void dic_image_destroy(dic_image* img);

/* ---- masks: LocalMask::new (masks.rs:58-187) and mask_update
 *      (masks.rs:199-263), host-side and bit-exact ----
 * coords_out holds (2*size+1)^2 pairs [x_off, y_off]; subset_mask_out
 * (2*size+1)^2 int32 (either may be NULL). */
int dic_mask_local(int shape, int size, double* coords_out, int32_t* subset_mask_out, int64_t* n_out);
int dic_mask_update(int shape, int size, double cx, double cy, const uint8_t* mask, int64_t mh, int64_t mw,
                    double* coords_out, int64_t* n_out);

/* ---- batched subset solve: Subset::new (subset.rs:836-919) followed by
 *      solve_icgn_result (subset.rs:1117-1219) or solve_fagn_result
 *      (subset.rs:1228-1334), for many subsets in one launch ----
 * Pixel lists are shared templates: template t holds the integer offsets
 * tpl_xy[2*tpl_start[t] .. 2*tpl_start[t+1]) as (dx, dy) pairs, in the
 * reference's row-major scan order; subset s uses template tpl_index[s]
 * about centre coords[2s], coords[2s+1] with warp order orders[s].  Warp
 * orders may differ between subsets: a solve then launches each order's
 * size classes side by side (the batched engine) or one launch per order
 * (the FP64 kernel), and list solves and residual maps group their lists
 * by order the same way; results are per subset either way. */
int dic_batch_create(dic_engine* e, int64_t n_subsets, const double* coords, const int32_t* orders,
                     const int32_t* tpl_index, int32_t n_templates, const int64_t* tpl_start,
                     const int16_t* tpl_xy, dic_batch** out);
/* Initial warps, n_subsets x 12 doubles (missing entries zero; p0 is
 * truncated or zero-padded to p_len as subset.rs:1135-1136 does). */
int dic_batch_set_p0(dic_batch* b, const double* p0, int mem);
/* Queues the solve on the engine stream; returns without waiting.  The
 * images may belong to another engine of the same device provided that
 * engine has been synchronised since it last wrote them (a shared cache of
 * reference images serving several engines). */
int dic_batch_solve(dic_batch* b, const dic_image* f_img, const dic_image* g_img, const dic_solve_params* prm);
/* No counterpart in the reference.  For callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): dic_batch_solve for DIC_PREC_PLAIN_F64,
 * want_history 0, whose per-pixel scratch (f, grad f and g per pixel, plus
 * grad g for FAGN: 32 or 48 bytes per template pixel summed over the batch)
 * lives in a fixed 64 MiB device array of the library instead of memory the
 * batch allocates; a larger batch is rejected with DIC_ERR_INVALID_INPUT.
 * It allocates nothing (the batch and images are allocated when created).
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (several engines or threads) race; each call writes every scratch
 * element before reading it.  Results as dic_batch_solve; per-pixel arrays
 * (dic_batch_subset_arrays) are not kept. */
int dic_batch_solve_noalloc(dic_batch* b, const dic_image* f_img, const dic_image* g_img, const dic_solve_params* prm);
/* Waits, then copies n_subsets results to the host. */
int dic_batch_results(dic_batch* b, dic_subset_result* out);
/* Waits, then copies n_subsets*max_iterations*4 history doubles
 * (iteration, norm, c_zncc, c_znssd); rows past `iterations` are zero. */
int dic_batch_history(dic_batch* b, double* out);
/* Reference intensities f (n_px) and gradients grad_f (n_px x 2) of one
 * subset (subset.rs:877-884); available after a whole-batch
 * DIC_PREC_PLAIN_F64 dic_batch_solve, until the batch's next solve of any
 * other kind (then DIC_ERR_NOT_SOLVED). */
int dic_batch_subset_arrays(dic_batch* b, int64_t subset, double* f, double* grad_f);
/* Schedule of the batched engine (DIC_PREC_COMP_F32): LDS window slack in
 * pixels per side beyond the template box, and wavefronts per subset for the
 * whole batch (1, 2, 4 or 8, overriding the autotuned choice; 0 = per size
 * class: the autotuned choice, else the fewest wavefronts that hold the
 * class's subsets at 12 pixels per thread).  No reference counterpart (the
 * reference has no device schedule). */
int dic_batch_set_schedule(dic_batch* b, int win_slack, int waves_per_subset);
/* Autotuner of the batched engine: for each size class of the batch (subsets
 * needing 1, 2, 4 or 8 wavefronts) times the candidate schedules (that many
 * wavefronts or twice as many; window slack 1 or 3 px) on these images and
 * keeps the fastest for later DIC_PREC_COMP_F32 solves.  Results held by the
 * batch afterwards are unspecified until the next solve.  No reference
 * counterpart (the reference has no device schedule). */
int dic_batch_autotune(dic_batch* b, const dic_image* f_img, const dic_image* g_img, const dic_solve_params* prm);
/* The schedule the batched engine runs each size class of the batch with:
 * wps[k] wavefronts per subset and slack[k] pixels of window slack for the
 * class of subsets needing 1, 2, 4 and 8 wavefronts (k = 0..3): the
 * autotuned choice if any, else the fewest wavefronts and the batch's
 * slack; with wavefronts set by dic_batch_set_schedule, that count for
 * every class.  A batch of mixed warp orders reports the classes of its
 * highest order (each order is tuned on its own).  No reference
 * counterpart (the reference has no device schedule). */
int dic_batch_get_schedule(dic_batch* b, int32_t* wps, int32_t* slack);
/* Solves only the listed subsets of a batch (each index at most once), each
 * from its row of p0 (count x 12, host; missing entries zero), and writes
 * their results to out (count records, host) in list order.  The batch's
 * stored warps of the listed subsets become the solved ones; its stored
 * result records (dic_batch_results) are left as they were, for listed and
 * other subsets alike.  Waits for the device.  No history
 * (prm->want_history must be 0).  This is one round of node solves of the
 * reference's mesh: the per-node Subset::new + solve_icgn_result calls of
 * a layer-RG colour (mesh.rs:1224-1233) or of one serial RG step
 * (mesh.rs:1133-1140). */
int dic_batch_solve_list(dic_batch* b, const dic_image* f_img, const dic_image* g_img, const dic_solve_params* prm,
                         const int32_t* idx, int64_t count, const double* p0, dic_subset_result* out);
/* ---- frame graphs: the per-frame steps of an incremental sequence
 *      (sequence.rs:624-968, one target frame) captured once as a HIP graph
 *      and replayed per frame: prefilter of g_img from its staged pixels,
 *      then (warm_start != 0) p0 := the batch's final warps, then the batched
 *      ICGN solve with prm (DIC_PREC_COMP_F32, no history).  Stage each new
 *      frame with dic_image_stage before the launch; updating f_img in place
 *      (dic_image_update, a reference update) keeps the graph valid.  The
 *      graph holds the batch's and images' device buffers: destroy it first.
 *      No reference counterpart (the reference loops on the host). ---- */
typedef struct dic_frame_graph dic_frame_graph;
/* Sets both the initial and the final warps (n_subsets x 12), so that the
 * next warm-started frame starts from p (e.g. zeros after a reference
 * update). */
int dic_batch_reset_warps(dic_batch* b, const double* p, int mem);
/* Creating a graph processes the staged frame once, eagerly (with the p0 the
 * batch holds), then records the graph without running it. */
int dic_frame_graph_create(dic_batch* b, const dic_image* f_img, dic_image* g_img, int warm_start,
                           const dic_solve_params* prm, dic_frame_graph** out);
/* Enqueues one frame on the engine's stream (dic_engine_sync waits for it). */
int dic_frame_graph_launch(dic_frame_graph* g);
void dic_frame_graph_destroy(dic_frame_graph* g);
/* Subset::residual_map (subset.rs:1350-1392): for each listed subset the
 * per-pixel zero-normalised residual e_i = (f_i - f_m)/Δf - (g_i - g_m)/Δg
 * at warp p (count x 12 doubles, host), or at the batch's stored final warps
 * when p is NULL (DIC_ERR_NOT_SOLVED if the batch was never solved),
 * evaluated in FP64 on the device.  out (host) receives the subsets' rows
 * one after another in list order, n_px doubles each in the template's
 * pixel order.  Waits for the device. */
int dic_batch_residual_map(dic_batch* b, const dic_image* f_img, const dic_image* g_img, const int32_t* idx,
                           int64_t count, const double* p, double* out);
/* Final warps as a device pointer (n_subsets x 12 doubles) for chaining
 * batches without a host round trip. */
int dic_batch_device_p(dic_batch* b, const double** out);
void dic_batch_destroy(dic_batch* b);

/* ---- mesh solve: Mesh::solve (mesh.rs:609-857) for uniform masking and
 *      deformable texture, on a mesh given as nodes and elements ----
 * nodes (n_nodes x 2, [x, y]); elements (n_elements x 3 for mesh_order 1,
 * x 6 for mesh_order 2: corners then midsides m01, m12, m20), int64 node
 * indices.  Node i's pixel list is template tpl_index[i] (as for
 * dic_batch_create): the local mask clipped by the region of interest, as
 * subset_at builds it (mesh.rs:575-593). */
enum { DIC_PRECOND_RG = 0, DIC_PRECOND_LAYER_RG = 1 };  /* Preconditioning, mesh.rs:299-303 */
typedef struct {
    dic_solve_params solve;   /* SolveConfig (mesh.rs:222-248): method, max_iterations, tolerance,
                                 max_norm, plus the engine precision; want_history must be 0 */
    int32_t subset_order;     /* SolveConfig.subset_order: 1 or 2 */
    int32_t preconditioning;  /* SolveConfig.preconditioning: DIC_PRECOND_* */
    int32_t override_active;  /* SolveConfig.override_active: keep a mesh with nodes short of quality_ok */
    int32_t batch_factor;     /* LayerRgConfig.batch_factor (mesh.rs:339-356) */
    int32_t workers;          /* LayerRgConfig.max_workers; 0 = the device's compute units */
    int32_t reserved;
    double root_rel_eps;      /* LayerRgConfig.root_rel_eps */
    double seed_coord[2];     /* SeedConfig.coord (mesh.rs:269-278) */
    double seed_warp[12];     /* SeedConfig.warp, zero padded to 12 */
    double seed_tolerance;    /* SeedConfig.tolerance */
} dic_mesh_params;
/* Caller-owned outputs (MeshSolution, mesh.rs:54-113). */
typedef struct {
    double* p;               /* n_nodes x 12 warps (p_len = 6 * subset_order used) */
    double* c_zncc;          /* n_nodes, clamped at 0 (mesh.rs:2396) */
    int32_t* iterations;     /* n_nodes */
    double* norms;           /* n_nodes: last convergence norm, 0 without iterations */
    int32_t* quality_ok;     /* n_nodes */
    double* warps;           /* n_elements x 12: element_strains (mesh.rs:1552-1665) */
    int64_t seed_node;       /* find_seed_node (mesh.rs:1682-1697) */
    int64_t n_node_solves;   /* subset solves run, retries and corrections included */
    int64_t n_lists;         /* device list solves launched */
    int64_t n_unsolvable;    /* nodes without quality_ok after corrections */
    double list_seconds;     /* wall time spent in the list solves (the rest is host traversal) */
} dic_mesh_output;
/* Progress and cancellation observer of a mesh solve: the SolveProgress
 * trait (src/progress.rs:16-38) that Mesh::solve takes (mesh.rs:610-616).
 * Every callback is made from the thread that called dic_mesh_solve, never
 * from the device or a worker; any pointer may be NULL.
 *   mesh_start(user, n)  a mesh solve over n nodes is starting;
 *   node_done(user)      one more node has been solved and stored (the
 *                        seed, then every node of the propagation, in the
 *                        order they are stored: mesh.rs:674,1143,1246);
 *   cancelled(user)      polled after every node_done and before every list
 *                        solve of the corrections; nonzero stops the solve
 *                        with DIC_ERR_CANCELLED and nothing written to the
 *                        outputs (mesh.rs:608-609).
 * The sequence loop's pair_start (sequence.rs:788) belongs to the caller's
 * frame loop above this ABI. */
typedef struct {
    void* user;
    void (*mesh_start)(void* user, int64_t n_nodes);
    void (*node_done)(void* user);
    int32_t (*cancelled)(void* user);
} dic_progress;
/* Seed, RG or layer-RG propagation, corrections, the quality gate, element
 * strains and the compatibility check.  Returns DIC_ERR_INVALID_INPUT for
 * an unsolvable mesh (unless override_active) or a folded element
 * (det F <= 0), DIC_ERR_MESH_GENERATION for a degenerate element, a subset
 * status for a failed node solve (as the reference's `?` does), and
 * DIC_ERR_CANCELLED when `progress` asks for it; the outputs are filled
 * whenever the propagation itself completed.  progress may be NULL (no
 * observer). */
int dic_mesh_solve(dic_engine* e, const dic_image* f_img, const dic_image* g_img, int64_t n_nodes,
                   const double* nodes, int64_t n_elements, int32_t mesh_order, const int64_t* elements,
                   const int32_t* tpl_index, int32_t n_templates, const int64_t* tpl_start, const int16_t* tpl_xy,
                   const dic_mesh_params* prm, const dic_progress* progress, dic_mesh_output* out);
/* element_strains (mesh.rs:1552-1665) alone, host-side: warps
 * (n_elements x 12) from node displacements (n_nodes x 2). */
int dic_mesh_element_strains(int64_t n_nodes, const double* nodes, int64_t n_elements, int32_t mesh_order,
                             const int64_t* elements, const double* displacements, double* warps);

/* ---- serving: continuous batching of correlation jobs on one GPU.  No
 *      reference counterpart (geopyv-dev has no server); it is the submit /
 *      poll path a serving caller binds instead of the Python loop ----
 * dic_submit queues a job (a reference/target pair, its subsets and solve
 * settings) and returns a ticket at once.  n_slots worker threads, each
 * with its own engine and stream, admit the queued jobs first come first
 * served as a slot frees and the job's device bytes fit within hbm_budget
 * beside the jobs in flight and the cached references; a job too large for
 * the budget alone runs when nothing else does.  Admitted jobs overlap on
 * the device with those already running.  Reference images stay
 * prefiltered on the device in a cache keyed by ref_key, evicted least
 * recently used first and only when no job holds them.  A job's records
 * do not depend on what else runs (its kernels read its own buffers and the
 * read-only reference coefficients).  All calls are thread-safe. */
typedef struct dic_server dic_server;
typedef struct {
    int32_t device;
    int32_t n_slots;     /* jobs on the device at once, >= 1 */
    int32_t max_queued;  /* waiting jobs beyond which dic_submit returns DIC_ERR_BUSY (0: no limit) */
    int32_t reserved;    /* 0 */
    int64_t hbm_budget;  /* device bytes for jobs in flight plus cached references */
} dic_server_params;
typedef struct {
    uint64_t ref_key;        /* jobs with equal keys carry equal reference pixels */
    const void* ref_pixels;  /* h x w host pixels of pix_type (DIC_PIX_U8 or DIC_PIX_F64) */
    const void* tar_pixels;  /* the target, same shape and type */
    int32_t pix_type;
    int32_t border;          /* >= 3 (image.rs:52) */
    int64_t h, w;
    int64_t n_subsets;       /* subsets as for dic_batch_create */
    const double* coords;
    const int32_t* orders;
    const int32_t* tpl_index;
    int32_t n_templates;
    int32_t reserved;        /* 0 */
    const int64_t* tpl_start;
    const int16_t* tpl_xy;
    const double* p0;        /* n_subsets x 12 initial warps, or NULL for zero */
    dic_solve_params params; /* want_history must be 0 */
} dic_job;
int dic_server_create(const dic_server_params* prm, dic_server** out);
/* Copies the job's arrays (the caller may free them on return), queues it
 * and sets *ticket.  DIC_ERR_BUSY when max_queued jobs are waiting. */
int dic_submit(dic_server* s, const dic_job* job, uint64_t* ticket);
/* Waits up to wait_ms (0: not at all, negative: until it finishes) for the
 * job.  Finished: its n_subsets records go to out in subset order,
 * *latency_ms (may be NULL) gets the time from dic_submit to the records
 * on the host, the ticket is released and the job's status is returned
 * (DIC_OK, or what its image, batch or solve call returned, with the
 * message in dic_last_error).  Not finished: DIC_PENDING, out untouched.
 * Unknown or released ticket: DIC_ERR_INVALID_INPUT. */
int dic_poll(dic_server* s, uint64_t ticket, int32_t wait_ms, dic_subset_result* out, double* latency_ms);
/* Jobs submitted, finished and not yet polled, and device bytes held by
 * the cached references (any may be NULL). */
int dic_server_stats(dic_server* s, int64_t* queued, int64_t* running, int64_t* finished, int64_t* cache_bytes);
/* Finishes the queued and running jobs, stops the workers and frees the
 * server (records not polled are dropped). */
void dic_server_destroy(dic_server* s);

#ifdef __cplusplus
}
#endif
#endif /* GEODIC_H */
