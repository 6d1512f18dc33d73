// lbm355_pool.cpp - serving pool: many cases resident on one GPU, admitted
// against a device-memory budget and stepped together.
//
// The reference runs one case per process and queues cases through Slurm
// (scripts/slurm_run_cases.py:104-198, scripts/run_all_cases.py).  Here one
// worker thread owns the device: requests wait in FIFO order until their
// footprint fits the budget, admitted cases of one collision/storage type are
// stepped by one fused launch per step over their concatenated brick lists
// (a batch), batches of different types are queued on their own streams in
// the same tick, and finished cases leave between ticks so that waiting ones
// can take their memory (continuous batching).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/lbm355.h"
#include "lbm355_internal.h"

// from lbm355_host.cpp: batch stepping split into enqueue and wait, and the
// library's thread-local error message
int batch_enqueue(lbm355_batch* b, int64_t n);
int batch_finish(lbm355_batch* b);
int lbm355_set_error(int code, const std::string& msg);
int64_t lbm355_case_step_bytes(const lbm355_case* c);

namespace {

int pool_fail(int code, const std::string& msg) { return lbm355_set_error(code, msg); }

struct Request {
    int64_t id = 0;
    lbm355_case_desc desc{};
    std::vector<int32_t> geometry;
    int64_t steps = 0, report_every = 0;
    int keep = 0;
    int64_t footprint = 0;
    lbm355_case* c = nullptr;
    int64_t active_cells = 0;
    lbm355_request_status st{};
    bool admitted = false;  // its footprint is counted in the pool's used bytes
    bool taken = false;
};

double now_s(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct lbm355_pool {
    std::mutex mu;
    std::condition_variable wake, changed;
    std::thread worker;
    bool stop = false;
    int device = 0;
    int64_t budget = 0, used = 0;
    int tick = 64;
    int graph = 1;
    std::chrono::steady_clock::time_point t0;
    std::map<int64_t, std::unique_ptr<Request>> reqs;
    std::deque<int64_t> queue;
    std::vector<Request*> running;  // worker-owned
    int64_t next_id = 1;
    int64_t n_done = 0, n_failed = 0, n_running = 0;
    int64_t cell_updates = 0, ticks = 0;
    double busy = 0.0;
    std::string worker_error;

    void run();
    void finish(Request* r, int state, int error);
};

void lbm355_pool::finish(Request* r, int state, int error)
{
    // called with the lock held; the worker's last library error explains a failure
    if (state == LBM355_REQ_FAILED) {
        const char* m = error == LBM355_ENOMEM && r->st.t_start == 0.0 ? "footprint exceeds the pool budget"
                                                                        : lbm355_last_error();
        std::strncpy(r->st.message, m, sizeof r->st.message - 1);
    }
    r->st.state = state;
    r->st.error = error;
    r->st.t_done = now_s(t0);
    if (state == LBM355_REQ_DONE) ++n_done;
    else ++n_failed;
    if (r->admitted && (!r->keep || state != LBM355_REQ_DONE)) {
        used -= r->footprint;
        r->admitted = false;
    }
    changed.notify_all();
}

void lbm355_pool::run()
{
    (void)hipSetDevice(device);
    std::vector<lbm355_batch*> batches;
    std::vector<lbm355_case*> solo;  // fd-force cases: their pre-kernels are per case
    bool rebuild = false;
    auto drop_batches = [&]() {
        for (lbm355_batch* b : batches) lbm355_batch_destroy(b);
        batches.clear();
        solo.clear();
    };
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        wake.wait(lk, [&] { return stop || !queue.empty() || !running.empty(); });
        if (stop) break;
        // FIFO admission against the budget; a request larger than the whole
        // budget fails instead of blocking the queue
        std::vector<Request*> admit;
        while (!queue.empty()) {
            Request* r = reqs[queue.front()].get();
            if (r->footprint > budget) {
                queue.pop_front();
                finish(r, LBM355_REQ_FAILED, LBM355_ENOMEM);
                continue;
            }
            if (used + r->footprint > budget) break;
            queue.pop_front();
            used += r->footprint;
            r->admitted = true;
            r->st.state = LBM355_REQ_RUNNING;
            r->st.t_start = now_s(t0);
            admit.push_back(r);
        }
        lk.unlock();

        for (Request* r : admit) {
            r->desc.geometry = r->geometry.data();
            const int rc = lbm355_case_create(&r->desc, &r->c);
            lbm355_case_info info;
            std::memset(&info, 0, sizeof info);
            if (rc == LBM355_OK) lbm355_case_info_get(r->c, &info);
            if (rc != LBM355_OK) {
                // zero-area cases terminate like sim_tcpc (exit code 2)
                if (r->c) lbm355_case_destroy(r->c);
                r->c = nullptr;
                lk.lock();
                finish(r, LBM355_REQ_FAILED, rc);
                lk.unlock();
                continue;
            }
            r->active_cells = info.active_cells;
            r->st.active_cells = info.active_cells;
            r->st.device_bytes = info.device_bytes;
            running.push_back(r);
            rebuild = true;
        }
        if (running.empty()) {
            lk.lock();
            n_running = 0;
            continue;
        }
        if (rebuild) {
            drop_batches();
            std::map<std::pair<int, int>, std::vector<lbm355_case*>> groups;
            for (Request* r : running) {
                if (r->desc.cy_enabled && r->desc.cy_model == LBM355_CY_FD_FORCE) solo.push_back(r->c);
                else groups[{r->desc.storage, r->desc.collision}].push_back(r->c);
            }
            int rc = LBM355_OK;
            for (auto& g : groups) {
                lbm355_batch* b = nullptr;
                rc = lbm355_batch_create(g.second.data(), int(g.second.size()), graph, &b);
                if (rc != LBM355_OK) break;
                batches.push_back(b);
            }
            if (rc != LBM355_OK) {
                lk.lock();
                worker_error = lbm355_last_error();
                for (Request* r : running) {
                    lbm355_case_destroy(r->c);
                    r->c = nullptr;
                    finish(r, LBM355_REQ_FAILED, rc);
                }
                running.clear();
                n_running = 0;
                drop_batches();
                continue;
            }
            rebuild = false;
        }
        // one tick: the same step count for every running case, cut at the
        // next completion or report point
        int64_t n = tick;
        for (Request* r : running) {
            n = std::min(n, r->steps - r->st.steps_done);
            if (r->report_every > 0) n = std::min(n, r->report_every - r->st.steps_done % r->report_every);
        }
        n = std::max<int64_t>(n, 1);
        const auto t_tick = std::chrono::steady_clock::now();
        int rc = LBM355_OK;
        // while all running cases fit the Infinity Cache together, the
        // batches of a tick run concurrently on their streams; beyond that
        // they run one after another, so that each batch's cache-sized groups
        // do not evict each other
        int64_t resident = 0;
        for (Request* r : running) resident += lbm355_case_step_bytes(r->c);
        const bool serial = resident > lbm355::kBatchGroupBytes;
        for (lbm355_batch* b : batches) {
            if (rc == LBM355_OK) rc = batch_enqueue(b, n);
            if (serial) {
                const int r2 = batch_finish(b);
                if (rc == LBM355_OK) rc = r2;
            }
        }
        for (lbm355_case* c : solo) {
            if (rc == LBM355_OK) rc = lbm355_case_step(c, n);  // queued on the case's stream
            if (serial) {
                const int r2 = lbm355_case_sync(c);
                if (rc == LBM355_OK) rc = r2;
            }
        }
        for (lbm355_batch* b : batches) {
            const int r2 = batch_finish(b);
            if (rc == LBM355_OK) rc = r2;
        }
        for (lbm355_case* c : solo) {
            const int r2 = lbm355_case_sync(c);
            if (rc == LBM355_OK) rc = r2;
        }
        const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_tick).count();
        std::vector<Request*> still;
        std::vector<std::pair<Request*, int>> ended;
        int64_t updates = 0;
        for (Request* r : running) {
            if (rc != LBM355_OK) { ended.push_back({r, rc}); continue; }
            r->st.steps_done += n;
            updates += r->active_cells * n;
            const bool report = r->report_every > 0 && r->st.steps_done % r->report_every == 0;
            const bool done = r->st.steps_done >= r->steps;
            if (report || done) {
                double fl[4] = {0, 0, 0, 0};
                if (lbm355_case_flux(r->c, 0, fl) == LBM355_OK) {
                    std::memcpy(r->st.flux, fl, sizeof fl);
                    if (r->st.t_first_report == 0.0) r->st.t_first_report = now_s(t0);
                }
            }
            if (done) {
                double mf[4] = {0, 0, 0, 0};
                if (r->desc.stats_every > 0 && lbm355_case_flux(r->c, 1, mf) == LBM355_OK)
                    std::memcpy(r->st.mean_flux, mf, sizeof mf);
                ended.push_back({r, LBM355_OK});
            } else {
                still.push_back(r);
            }
        }
        for (auto& e : ended)
            if (!(e.first->keep && e.second == LBM355_OK)) {
                lbm355_case_destroy(e.first->c);
                e.first->c = nullptr;
            }
        if (!ended.empty()) {
            rebuild = true;
            drop_batches();  // a batch must not outlive one of its cases
        }
        running.swap(still);
        lk.lock();
        ticks += 1;
        busy += dt;
        cell_updates += updates;
        n_running = int64_t(running.size());
        for (auto& e : ended) finish(e.first, e.second == LBM355_OK ? LBM355_REQ_DONE : LBM355_REQ_FAILED, e.second);
        changed.notify_all();
    }
    lk.unlock();
    drop_batches();
    for (Request* r : running) lbm355_case_destroy(r->c);
    running.clear();
}

extern "C" {

int lbm355_pool_create(int64_t budget_bytes, int steps_per_tick, int graph, lbm355_pool** out)
{
    if (!out) return pool_fail(LBM355_EINVAL, "null out");
    *out = nullptr;
    if (steps_per_tick <= 0) steps_per_tick = 64;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return pool_fail(LBM355_EDEVICE, "no HIP device");
    if (budget_bytes <= 0) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return pool_fail(LBM355_EDEVICE, "hipMemGetInfo");
        budget_bytes = int64_t(double(free_b) * 0.9);
    }
    std::unique_ptr<lbm355_pool> p(new (std::nothrow) lbm355_pool());
    if (!p) return pool_fail(LBM355_ENOMEM, "host allocation");
    p->device = dev;
    p->budget = budget_bytes;
    p->tick = steps_per_tick;
    p->graph = graph ? 1 : 0;
    p->t0 = std::chrono::steady_clock::now();
    lbm355_pool* raw = p.get();
    p->worker = std::thread([raw] { raw->run(); });
    *out = p.release();
    return LBM355_OK;
}

int lbm355_pool_submit(lbm355_pool* p, const lbm355_case_desc* desc, int64_t steps, int64_t report_every, int keep,
                       int64_t* id)
{
    if (!p || !desc || !id || !desc->geometry || steps <= 0) return pool_fail(LBM355_EINVAL, "bad submit arguments");
    if (desc->storage == LBM355_BF16_OFFSET)
        return pool_fail(LBM355_EINVAL, "bfloat16-offset storage is not served: its velocity error against FP64 "
                                        "(4.3e-3 on the 512x384x256 junction) exceeds the 1e-3 bar; use FP16-offset");
    std::unique_ptr<Request> r(new (std::nothrow) Request());
    if (!r) return pool_fail(LBM355_ENOMEM, "host allocation");
    int64_t fp = 0;
    const int rc = lbm355_case_footprint(desc, &fp);
    if (rc != LBM355_OK) return pool_fail(rc, lbm355_last_error());
    const int64_t n = int64_t(desc->nx) * desc->ny * desc->nz;
    r->desc = *desc;
    r->geometry.assign(desc->geometry, desc->geometry + n);
    r->desc.geometry = nullptr;
    r->steps = steps;
    r->report_every = report_every > 0 ? report_every : 0;
    r->keep = keep ? 1 : 0;
    r->footprint = fp;
    r->st.state = LBM355_REQ_QUEUED;
    r->st.steps = steps;
    std::lock_guard<std::mutex> g(p->mu);
    r->id = p->next_id++;
    r->st.t_submit = now_s(p->t0);
    *id = r->id;
    p->queue.push_back(r->id);
    p->reqs[r->id] = std::move(r);
    p->wake.notify_all();
    return LBM355_OK;
}

int lbm355_pool_poll(lbm355_pool* p, int64_t id, lbm355_request_status* st)
{
    if (!p) return pool_fail(LBM355_EINVAL, "null pool");
    std::lock_guard<std::mutex> g(p->mu);
    auto it = p->reqs.find(id);
    if (it == p->reqs.end()) return pool_fail(LBM355_ENOTFOUND, "unknown request id");
    if (st) *st = it->second->st;
    const int s = it->second->st.state;
    return s == LBM355_REQ_DONE || s == LBM355_REQ_FAILED ? LBM355_OK : LBM355_EBUSY;
}

int lbm355_pool_wait(lbm355_pool* p, int64_t id, double timeout_s, lbm355_request_status* st)
{
    if (!p) return pool_fail(LBM355_EINVAL, "null pool");
    std::unique_lock<std::mutex> lk(p->mu);
    auto it = p->reqs.find(id);
    if (it == p->reqs.end()) return pool_fail(LBM355_ENOTFOUND, "unknown request id");
    Request* r = it->second.get();
    auto ready = [&] { return r->st.state == LBM355_REQ_DONE || r->st.state == LBM355_REQ_FAILED; };
    if (timeout_s < 0) p->changed.wait(lk, ready);
    else p->changed.wait_for(lk, std::chrono::duration<double>(timeout_s), ready);
    if (st) *st = r->st;
    return ready() ? LBM355_OK : LBM355_EBUSY;
}

int lbm355_pool_drain(lbm355_pool* p, double timeout_s)
{
    if (!p) return pool_fail(LBM355_EINVAL, "null pool");
    std::unique_lock<std::mutex> lk(p->mu);
    auto idle = [&] { return p->queue.empty() && p->n_running == 0 && p->n_done + p->n_failed == p->next_id - 1; };
    if (timeout_s < 0) p->changed.wait(lk, idle);
    else p->changed.wait_for(lk, std::chrono::duration<double>(timeout_s), idle);
    return idle() ? LBM355_OK : LBM355_EBUSY;
}

int lbm355_pool_take(lbm355_pool* p, int64_t id, lbm355_case** out)
{
    if (!p || !out) return pool_fail(LBM355_EINVAL, "null argument");
    std::lock_guard<std::mutex> g(p->mu);
    auto it = p->reqs.find(id);
    if (it == p->reqs.end()) return pool_fail(LBM355_ENOTFOUND, "unknown request id");
    Request* r = it->second.get();
    if (r->st.state != LBM355_REQ_DONE) return pool_fail(LBM355_EBUSY, "request not done");
    if (!r->keep || r->taken || !r->c) return pool_fail(LBM355_EINVAL, "request was not submitted with keep=1");
    *out = r->c;
    r->c = nullptr;
    r->taken = true;
    if (r->admitted) p->used -= r->footprint;  // the caller owns the memory from now on
    r->admitted = false;
    p->wake.notify_all();
    return LBM355_OK;
}

int lbm355_pool_stats(lbm355_pool* p, lbm355_pool_info* st)
{
    if (!p || !st) return pool_fail(LBM355_EINVAL, "null argument");
    std::lock_guard<std::mutex> g(p->mu);
    std::memset(st, 0, sizeof *st);
    st->submitted = p->next_id - 1;
    st->queued = int64_t(p->queue.size());
    st->running = p->n_running;
    st->done = p->n_done;
    st->failed = p->n_failed;
    st->budget_bytes = p->budget;
    st->used_bytes = p->used;
    st->cell_updates = p->cell_updates;
    st->ticks = p->ticks;
    st->busy_seconds = p->busy;
    st->clock = now_s(p->t0);
    return LBM355_OK;
}

int lbm355_pool_destroy(lbm355_pool* p)
{
    if (!p) return LBM355_OK;
    {
        std::lock_guard<std::mutex> g(p->mu);
        p->stop = true;
        p->wake.notify_all();
    }
    if (p->worker.joinable()) p->worker.join();
    for (auto& kv : p->reqs)
        if (kv.second->c) lbm355_case_destroy(kv.second->c);  // kept and never taken
    delete p;
    return LBM355_OK;
}

}  // extern "C"
