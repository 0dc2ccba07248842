// Serving loop behind the C ABI (dic_server_*, dic_submit, dic_poll):
// continuous batching of correlation jobs on one GPU.  No reference
// counterpart; the survey's boundary asks for an async submit plus poll
// path for a serving caller.
//
//  * n_slots worker threads, each owning an engine (its own stream); a job
//    runs on the slot that admitted it: target upload and prefilter, batch
//    creation, the solve, the records back to the host;
//  * admission is first come first served: the oldest queued job starts
//    when a slot is free and its device bytes fit within the budget beside
//    the jobs in flight and the cached references (unused references are
//    evicted, least recently used first, to make room); a job too large for
//    the budget alone runs when nothing else holds device memory;
//  * references are prefiltered once on the cache engine and shared
//    read-only by the slots (each reference is synchronised before any
//    slot reads it, as dic_batch_solve requires of images of another
//    engine);
//  * a job's records depend only on its own inputs: its kernels read its
//    own buffers and the read-only reference coefficients.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "internal.h"

namespace geodic {
namespace {

using Clock = std::chrono::steady_clock;

struct QueuedJob {
    uint64_t ticket = 0;
    dic_job job{};  // its pointers name the copies below
    std::vector<unsigned char> ref, tar;
    std::vector<double> coords, p0;
    std::vector<int32_t> orders, tpl_index;
    std::vector<int64_t> tpl_start;
    std::vector<int16_t> tpl_xy;
    int64_t bytes = 0;      // device bytes of the job besides its reference
    int64_t ref_bytes = 0;  // device bytes of its reference image
    Clock::time_point submitted;
};

struct Finished {
    int status = DIC_OK;
    std::string error;
    double latency_ms = 0.0;
    std::vector<dic_subset_result> records;
};

struct CachedRef {
    uint64_t key = 0;
    dic_image* img = nullptr;
    int64_t bytes = 0;
    int users = 0;
};

int64_t image_bytes(int64_t h, int64_t w, int border) {
    const int64_t hp = h + 2 * (int64_t)border, wp = w + 2 * (int64_t)border;
    return hp * (wp * 8 + ((wp + 3) & ~(int64_t)3) * 4);  // FP64 coefficients + the FP32 copy
}

}  // namespace

struct ServerImpl {
    dic_server_params prm{};
    std::vector<dic_engine*> engines;
    dic_engine* cache_engine = nullptr;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::unique_ptr<QueuedJob>> queue;
    std::unordered_map<uint64_t, int> state;  // ticket -> 0 queued, 1 running, 2 finished
    std::unordered_map<uint64_t, Finished> finished;
    std::list<CachedRef> cache;               // most recently used first
    int64_t in_flight = 0, cache_bytes = 0;
    int64_t running = 0;
    std::vector<int> free_slots;
    bool closing = false;
    uint64_t next_ticket = 1;
    std::vector<std::thread> workers;

    // Under mu: can the oldest queued job start now?  Evicts unused
    // references (their images go to `drop`, destroyed outside the lock).
    bool admit_front(std::vector<dic_image*>& drop) {
        if (free_slots.empty() || queue.empty()) return false;
        const QueuedJob& j = *queue.front();
        bool cached = false;
        for (const CachedRef& c : cache) cached = cached || c.key == j.job.ref_key;
        const int64_t need = j.bytes + (cached ? 0 : j.ref_bytes);
        int64_t free_b = prm.hbm_budget - in_flight - cache_bytes;
        for (auto it = cache.end(); need > free_b && it != cache.begin();) {
            --it;
            if (it->users == 0 && it->key != j.job.ref_key) {
                free_b += it->bytes;
                cache_bytes -= it->bytes;
                drop.push_back(it->img);
                it = cache.erase(it);
            }
        }
        if (need <= free_b) return true;
        bool idle = in_flight == 0;
        for (const CachedRef& c : cache) idle = idle && c.users == 0;
        return idle;
    }

    // The job's reference image, prefiltered and synchronised, its cache
    // entry held until release_ref.
    int acquire_ref(const QueuedJob& j, dic_image** out) {
        const dic_job& jb = j.job;
        {
            std::lock_guard<std::mutex> lk(mu);
            for (auto it = cache.begin(); it != cache.end(); ++it)
                if (it->key == jb.ref_key) {
                    ++it->users;
                    cache.splice(cache.begin(), cache, it);
                    *out = cache.front().img;
                    return DIC_OK;
                }
        }
        dic_image* img = nullptr;
        int st = dic_image_create(cache_engine, jb.ref_pixels, jb.pix_type, DIC_MEM_HOST, jb.h, jb.w, jb.border, &img);
        if (!st) st = dic_engine_sync(cache_engine);
        if (st) {
            if (img) dic_image_destroy(img);
            return st;
        }
        dic_image* dup = nullptr;
        {
            std::lock_guard<std::mutex> lk(mu);
            for (auto it = cache.begin(); it != cache.end(); ++it)
                if (it->key == jb.ref_key) {  // another slot made it meanwhile
                    ++it->users;
                    cache.splice(cache.begin(), cache, it);
                    dup = img;
                    img = cache.front().img;
                    break;
                }
            if (!dup) {
                cache.push_front(CachedRef{jb.ref_key, img, j.ref_bytes, 1});
                cache_bytes += j.ref_bytes;
            }
        }
        if (dup) dic_image_destroy(dup);
        *out = img;
        return DIC_OK;
    }

    void release_ref(uint64_t key) {
        std::lock_guard<std::mutex> lk(mu);
        for (CachedRef& c : cache)
            if (c.key == key) --c.users;
        cv.notify_all();
    }

    void run(const QueuedJob& j, dic_engine* eng, Finished& fin) {
        const dic_job& jb = j.job;
        dic_image* f = nullptr;
        dic_image* g = nullptr;
        dic_batch* b = nullptr;
        int st = acquire_ref(j, &f);
        const bool held = st == DIC_OK;
        if (!st) st = dic_image_create(eng, jb.tar_pixels, jb.pix_type, DIC_MEM_HOST, jb.h, jb.w, jb.border, &g);
        if (!st)
            st = dic_batch_create(eng, jb.n_subsets, jb.coords, jb.orders, jb.tpl_index, jb.n_templates,
                                  jb.tpl_start, jb.tpl_xy, &b);
        if (!st && jb.p0) st = dic_batch_set_p0(b, jb.p0, DIC_MEM_HOST);
        if (!st) st = dic_batch_solve(b, f, g, &jb.params);
        if (!st) st = dic_engine_sync(eng);
        if (!st) {
            fin.records.resize((size_t)jb.n_subsets);
            st = dic_batch_results(b, fin.records.data());
        }
        if (st) {
            fin.error = dic_last_error();
            fin.records.clear();
        }
        if (b) dic_batch_destroy(b);
        if (g) dic_image_destroy(g);
        if (held) release_ref(jb.ref_key);
        fin.status = st;
        fin.latency_ms = std::chrono::duration<double, std::milli>(Clock::now() - j.submitted).count();
    }

    void worker() {
        for (;;) {
            std::unique_ptr<QueuedJob> job;
            int slot = -1;
            bool exiting = false;
            std::vector<dic_image*> drop;  // evicted references, destroyed outside the lock
            {
                std::unique_lock<std::mutex> lk(mu);
                bool ok = false;
                for (;;) {
                    ok = admit_front(drop);
                    if (ok || (closing && queue.empty()) || !drop.empty()) break;
                    cv.wait(lk);
                }
                if (ok) {
                    job = std::move(queue.front());
                    queue.pop_front();
                    slot = free_slots.back();
                    free_slots.pop_back();
                    in_flight += job->bytes;
                    ++running;
                    state[job->ticket] = 1;
                } else {
                    exiting = closing && queue.empty();
                }
            }
            for (dic_image* im : drop) dic_image_destroy(im);
            if (exiting) return;
            if (!job) continue;
            Finished fin;
            try {
                run(*job, engines[(size_t)slot], fin);
            } catch (...) {  // no exception leaves the library
                fin.status = DIC_ERR_OUT_OF_MEMORY;
                fin.error = "serving: host allocation failed while running a job";
                fin.records.clear();
            }
            std::lock_guard<std::mutex> lk(mu);
            in_flight -= job->bytes;
            --running;
            free_slots.push_back(slot);
            state[job->ticket] = 2;
            finished[job->ticket] = std::move(fin);
            cv.notify_all();
        }
    }
};

}  // namespace geodic

struct dic_server {
    geodic::ServerImpl s;
};

using geodic::set_error;

extern "C" {

int dic_server_create(const dic_server_params* prm, dic_server** out) {
    CHECK_ARG(prm && out, "dic_server_create: NULL argument");
    *out = nullptr;
    CHECK_ARG(prm->n_slots >= 1 && prm->n_slots <= 64, "invalid input: n_slots must be 1..64");
    CHECK_ARG(prm->max_queued >= 0, "invalid input: max_queued must be >= 0");
    CHECK_ARG(prm->hbm_budget > 0, "invalid input: hbm_budget must be positive");
    std::unique_ptr<dic_server> ds;
    try {
        ds.reset(new dic_server());
    } catch (...) {
        set_error("dic_server_create: host allocation failed");
        return DIC_ERR_OUT_OF_MEMORY;
    }
    geodic::ServerImpl& s = ds->s;
    s.prm = *prm;
    auto cleanup = [&]() {
        for (dic_engine* e : s.engines) dic_engine_destroy(e);
        if (s.cache_engine) dic_engine_destroy(s.cache_engine);
    };
    for (int k = 0; k < prm->n_slots; ++k) {
        dic_engine* e = nullptr;
        const int st = dic_engine_create(prm->device, &e);
        if (st) {
            cleanup();
            return st;
        }
        s.engines.push_back(e);
        s.free_slots.push_back(k);
    }
    const int st = dic_engine_create(prm->device, &s.cache_engine);
    if (st) {
        cleanup();
        return st;
    }
    try {
        for (int k = 0; k < prm->n_slots; ++k) s.workers.emplace_back([&s] { s.worker(); });
    } catch (...) {
        {
            std::lock_guard<std::mutex> lk(s.mu);
            s.closing = true;
        }
        s.cv.notify_all();
        for (std::thread& t : s.workers) t.join();
        cleanup();
        set_error("dic_server_create: could not start the worker threads");
        return DIC_ERR_OUT_OF_MEMORY;
    }
    *out = ds.release();
    return DIC_OK;
}

int dic_submit(dic_server* ds, const dic_job* job, uint64_t* ticket) {
    CHECK_ARG(ds && job && ticket, "dic_submit: NULL argument");
    const dic_job& j = *job;
    CHECK_ARG(j.pix_type == DIC_PIX_U8 || j.pix_type == DIC_PIX_F64, "invalid input: job pixels must be U8 or F64");
    CHECK_ARG(j.ref_pixels && j.tar_pixels, "dic_submit: NULL pixels");
    CHECK_ARG(j.h >= 1 && j.w >= 1 && j.h <= (1 << 20) && j.w <= (1 << 20), "invalid input: image shape");
    CHECK_ARG(j.border >= 3, "invalid input: border must be >= 3");
    CHECK_ARG(j.n_subsets >= 1 && j.n_subsets <= ((int64_t)1 << 30), "invalid input: subset count");
    CHECK_ARG(j.coords && j.orders && j.tpl_index && j.tpl_start && j.tpl_xy, "dic_submit: NULL array");
    CHECK_ARG(j.n_templates >= 1, "invalid input: no templates");
    CHECK_ARG(j.tpl_start[0] == 0, "invalid input: tpl_start[0] must be 0");
    for (int32_t t = 0; t < j.n_templates; ++t)
        CHECK_ARG(j.tpl_start[t + 1] >= j.tpl_start[t], "invalid input: tpl_start must be non-decreasing");
    CHECK_ARG(!j.params.want_history, "invalid input: served jobs record no history");
    for (int64_t s = 0; s < j.n_subsets; ++s)
        CHECK_ARG(j.tpl_index[s] >= 0 && j.tpl_index[s] < j.n_templates,
                  "invalid input: template index out of range");
    geodic::ServerImpl& s = ds->s;
    std::unique_ptr<geodic::QueuedJob> q;
    try {
        q.reset(new geodic::QueuedJob());
        const size_t px = (size_t)j.h * (size_t)j.w * (j.pix_type == DIC_PIX_U8 ? 1 : 8);
        const size_t n = (size_t)j.n_subsets;
        const size_t n_off = (size_t)j.tpl_start[j.n_templates];
        q->ref.assign((const unsigned char*)j.ref_pixels, (const unsigned char*)j.ref_pixels + px);
        q->tar.assign((const unsigned char*)j.tar_pixels, (const unsigned char*)j.tar_pixels + px);
        q->coords.assign(j.coords, j.coords + 2 * n);
        q->orders.assign(j.orders, j.orders + n);
        q->tpl_index.assign(j.tpl_index, j.tpl_index + n);
        q->tpl_start.assign(j.tpl_start, j.tpl_start + j.n_templates + 1);
        q->tpl_xy.assign(j.tpl_xy, j.tpl_xy + 2 * n_off);
        if (j.p0) q->p0.assign(j.p0, j.p0 + 12 * n);
        q->job = j;
        q->job.ref_pixels = q->ref.data();
        q->job.tar_pixels = q->tar.data();
        q->job.coords = q->coords.data();
        q->job.orders = q->orders.data();
        q->job.tpl_index = q->tpl_index.data();
        q->job.tpl_start = q->tpl_start.data();
        q->job.tpl_xy = q->tpl_xy.data();
        q->job.p0 = j.p0 ? q->p0.data() : nullptr;
        // what the job allocates besides its reference: the target's
        // coefficients, its slot's row-pass buffer and staged pixels, the
        // batch (per subset: centre, warps, result, dispatch order and start
        // record; the templates twice) and a margin
        const int64_t hp = j.h + 2 * (int64_t)j.border, wp = j.w + 2 * (int64_t)j.border;
        q->ref_bytes = geodic::image_bytes(j.h, j.w, j.border) + j.h * j.w * 16;  // and its knot table
        q->bytes = q->ref_bytes + hp * wp * 8 + j.h * j.w * 8 + (int64_t)n * 512 + (int64_t)n_off * 8 + (1 << 20);
    } catch (...) {
        set_error("dic_submit: host allocation failed while copying the job");
        return DIC_ERR_OUT_OF_MEMORY;
    }
    q->submitted = geodic::Clock::now();
    {
        std::lock_guard<std::mutex> lk(s.mu);
        if (s.closing) {
            set_error("dic_submit: the server is shutting down");
            return DIC_ERR_INVALID_INPUT;
        }
        if (s.prm.max_queued > 0 && (int64_t)s.queue.size() >= s.prm.max_queued) {
            set_error("busy: the serving queue is full");
            return DIC_ERR_BUSY;
        }
        q->ticket = s.next_ticket++;
        *ticket = q->ticket;
        s.state[q->ticket] = 0;
        s.queue.push_back(std::move(q));
    }
    s.cv.notify_all();
    return DIC_OK;
}

int dic_poll(dic_server* ds, uint64_t ticket, int32_t wait_ms, dic_subset_result* out, double* latency_ms) {
    CHECK_ARG(ds && out, "dic_poll: NULL argument");
    geodic::ServerImpl& s = ds->s;
    std::unique_lock<std::mutex> lk(s.mu);
    auto known = [&] { return s.state.count(ticket) != 0; };
    auto done = [&] { return s.finished.count(ticket) != 0; };
    CHECK_ARG(known(), "invalid input: dic_poll: unknown or released ticket");
    if (!done()) {
        if (wait_ms < 0)
            s.cv.wait(lk, done);
        else if (wait_ms > 0)
            s.cv.wait_for(lk, std::chrono::milliseconds(wait_ms), done);
        if (!done()) return DIC_PENDING;
    }
    geodic::Finished fin = std::move(s.finished[ticket]);
    s.finished.erase(ticket);
    s.state.erase(ticket);
    lk.unlock();
    if (latency_ms) *latency_ms = fin.latency_ms;
    if (fin.status) {
        set_error(fin.error);
        return fin.status;
    }
    std::memcpy(out, fin.records.data(), sizeof(dic_subset_result) * fin.records.size());
    return DIC_OK;
}

int dic_server_stats(dic_server* ds, int64_t* queued, int64_t* running, int64_t* finished, int64_t* cache_bytes) {
    CHECK_ARG(ds, "dic_server_stats: NULL argument");
    geodic::ServerImpl& s = ds->s;
    std::lock_guard<std::mutex> lk(s.mu);
    if (queued) *queued = (int64_t)s.queue.size();
    if (running) *running = s.running;
    if (finished) *finished = (int64_t)s.finished.size();
    if (cache_bytes) *cache_bytes = s.cache_bytes;
    return DIC_OK;
}

void dic_server_destroy(dic_server* ds) {
    if (!ds) return;
    geodic::ServerImpl& s = ds->s;
    {
        std::lock_guard<std::mutex> lk(s.mu);
        s.closing = true;
    }
    s.cv.notify_all();
    for (std::thread& t : s.workers) t.join();
    for (geodic::CachedRef& c : s.cache) dic_image_destroy(c.img);
    s.cache.clear();
    for (dic_engine* e : s.engines) dic_engine_destroy(e);
    if (s.cache_engine) dic_engine_destroy(s.cache_engine);
    delete ds;
}

}  // extern "C"
