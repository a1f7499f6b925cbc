// lt_server.cpp -- a map-serving loop on one GPU: resident lens models, requests admitted by HBM
// footprint as stream slots and memory free up (continuous batching: work of different requests
// overlaps on the device), results kept in device memory until the caller fetches them.
//
// The reference has no serving loop; its map drivers (utils/exec/main.cpp:397-545,
// utils/maps/main.cpp:428-587) call one map function after another, each allocating, computing and
// copying back synchronously. Here one dispatcher thread owns all launches:
//   submit()  -> queue (FIFO, with bounded bypass so that small requests can pass a large one
//                that does not fit yet)
//   dispatch  -> admit while a slot is free and the request's planes fit in the budget, allocate
//                the planes from a size-keyed cache, launch lt_map_device on the slot's stream,
//                record an event
//   poll      -> events that completed mark their tickets done and free their slots
//   fetch()   -> caller thread: wait, copy planes to host on its own stream, release memory.
// Every pixel is computed by the same code whatever else runs, so a request's result is the same
// alone and in a mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "lenstool_hip/lt_capi.h"
#include "lt_internal.h"

using namespace lt;

namespace {

using Clock = std::chrono::steady_clock;

enum class State { Queued, Running, Done, Failed };

struct Ticket {
    long long id = 0;
    lt_request req{};
    State state = State::Queued;
    int status = LT_OK;
    std::string err;
    int nplanes = 0;
    size_t plane_bytes = 0;
    double* planes[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    int slot = -1;
    hipEvent_t done_ev = nullptr;
    int bypassed = 0;
    double halo_px = 0;
    Clock::time_point t_submit, t_start, t_done, t_fetched;
};

struct ServedModel {
    int nhalos = 0;
    // one packed model per use; null when the lens has no such kernel (e.g. type 14 and gradients)
    lt_model use_model[4] = {nullptr, nullptr, nullptr, nullptr};  // gradient, hessian, kmap, potential
    int use_status[4] = {LT_OK, LT_OK, LT_OK, LT_OK};
    std::string use_err[4];
};

int use_index(int kind)
{
    switch (kind) {
    case LT_MAP_GRADIENT:
    case LT_MAP_DPL: return 0;
    case LT_MAP_MASS: return 2;
    case LT_MAP_POT: return 3;
    default: return 1;
    }
}

int use_kind(int idx)
{
    static const int k[4] = {LT_MAP_GRADIENT, LT_MAP_HESSIAN, LT_MAP_MASS, LT_MAP_POT};
    return k[idx];
}

int planes_of(int kind)
{
    switch (kind) {
    case LT_MAP_HESSIAN: return 4;
    case LT_MAP_FIELDS: return 5;
    case LT_MAP_GRADIENT:
    case LT_MAP_DPL: return 2;
    default: return 1;
    }
}

double ms_between(Clock::time_point a, Clock::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

double percentile(std::vector<double> v, double q)
{
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = q / 100.0 * (double)(v.size() - 1);
    const size_t i = (size_t)pos;
    const double f = pos - (double)i;
    return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
}

}  // namespace

struct lt_server_s {
    int device = 0;
    long long budget = 0;
    Clock::time_point t0;
    std::vector<hipStream_t> slots;
    std::vector<bool> slot_busy;
    std::vector<std::unique_ptr<ServedModel>> models;

    std::mutex mu;
    std::condition_variable cv_work;  // dispatcher wakes on submit / release / stop
    std::condition_variable cv_done;  // waiters wake on completion
    bool stop = false;
    long long next_id = 1;
    std::deque<long long> queue;
    std::unordered_map<long long, std::unique_ptr<Ticket>> tickets;
    std::vector<long long> running;

    // device memory: in use by tickets + cached free blocks, both counted against the budget
    long long bytes_used = 0, bytes_cached = 0, bytes_peak = 0;
    std::multimap<size_t, void*> cache;

    // statistics over completed requests since the last reset
    std::vector<double> lat_done, lat_fetch;
    long long n_submitted = 0, n_completed = 0, n_failed = 0;
    double halo_px_done = 0;
    Clock::time_point stats_t0;

    std::thread worker;

    // ---- memory (called with mu held, from the dispatcher) ----------------------------------------
    void* take(size_t bytes)
    {
        auto it = cache.find(bytes);
        if (it != cache.end()) {
            void* p = it->second;
            cache.erase(it);
            bytes_cached -= (long long)bytes;
            bytes_used += (long long)bytes;
            bytes_peak = std::max(bytes_peak, bytes_used + bytes_cached);
            return p;
        }
        // make room: drop cached blocks of other sizes
        while (!cache.empty() && bytes_used + bytes_cached + (long long)bytes > budget) {
            auto victim = cache.begin();
            (void)hipFree(victim->second);
            bytes_cached -= (long long)victim->first;
            cache.erase(victim);
        }
        void* p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        bytes_used += (long long)bytes;
        bytes_peak = std::max(bytes_peak, bytes_used + bytes_cached);
        return p;
    }

    void give_back(void* p, size_t bytes)
    {
        bytes_used -= (long long)bytes;
        bytes_cached += (long long)bytes;
        cache.emplace(bytes, p);
    }

    bool fits(const Ticket& t) const
    {
        // cached blocks can be freed to make room
        return bytes_used + (long long)(t.plane_bytes * t.nplanes) <= budget;
    }

    // ---- dispatcher -------------------------------------------------------------------------------
    void fail_ticket(Ticket& t, int st, const std::string& msg)
    {
        t.state = State::Failed;
        t.status = st;
        t.err = msg;
        t.t_done = Clock::now();
        ++n_failed;
    }

    void launch(Ticket& t, int slot)
    {
        for (int k = 0; k < t.nplanes; ++k) {
            t.planes[k] = (double*)take(t.plane_bytes);
            if (!t.planes[k]) {
                for (int j = 0; j < k; ++j) give_back(t.planes[j], t.plane_bytes);
                for (int j = 0; j < 5; ++j) t.planes[j] = nullptr;
                fail_ticket(t, LT_ERR_NOMEM, "device allocation of the map planes failed");
                return;
            }
        }
        ServedModel& m = *models[t.req.model];
        const int ui = use_index(t.req.kind);
        lt_grid_desc g = t.req.grid;
        g.pitch = g.nx;
        t.slot = slot;
        t.t_start = Clock::now();
        int st = lt_map_device(m.use_model[ui], t.req.kind, t.req.mode, t.planes[0], t.planes[1], t.planes[2],
                               t.planes[3], t.planes[4], t.req.r, t.req.mult, &g, t.req.precision, 0, slots[slot]);
        if (st == LT_OK && hipEventRecord(t.done_ev, slots[slot]) != hipSuccess) st = LT_ERR_HIP;
        if (st != LT_OK) {
            for (int k = 0; k < t.nplanes; ++k) give_back(t.planes[k], t.plane_bytes);
            for (int k = 0; k < 5; ++k) t.planes[k] = nullptr;
            fail_ticket(t, st, lt_last_error());
            return;
        }
        t.state = State::Running;
        slot_busy[slot] = true;
        running.push_back(t.id);
    }

    bool poll_running()
    {
        bool any = false;
        for (size_t i = 0; i < running.size();) {
            Ticket& t = *tickets[running[i]];
            hipError_t q = hipEventQuery(t.done_ev);
            if (q == hipErrorNotReady) {
                ++i;
                continue;
            }
            t.t_done = Clock::now();
            slot_busy[t.slot] = false;
            if (q == hipSuccess) {
                t.state = State::Done;
                ++n_completed;
                halo_px_done += t.halo_px;
                lat_done.push_back(ms_between(t.t_submit, t.t_done));
            } else {
                (void)hipGetLastError();
                fail_ticket(t, LT_ERR_HIP, hipGetErrorString(q));
            }
            running[i] = running.back();
            running.pop_back();
            any = true;
        }
        return any;
    }

    bool admit()
    {
        bool any = false;
        for (auto it = queue.begin(); it != queue.end();) {
            int slot = -1;
            for (size_t s = 0; s < slots.size(); ++s)
                if (!slot_busy[s]) {
                    slot = (int)s;
                    break;
                }
            if (slot < 0) break;
            Ticket& t = *tickets[*it];
            if (!fits(t)) {
                // bounded bypass: once the head has been passed over 8 times, nothing else goes first
                if (it == queue.begin() && ++t.bypassed > 8) break;
                ++it;
                continue;
            }
            launch(t, slot);
            it = queue.erase(it);
            any = true;
        }
        return any;
    }

    void run()
    {
        (void)hipSetDevice(device);
        std::unique_lock<std::mutex> lk(mu);
        while (!stop) {
            bool done = poll_running();
            bool admitted = admit();
            if (done || admitted) cv_done.notify_all();
            if (!running.empty())
                cv_work.wait_for(lk, std::chrono::microseconds(50));
            else if (queue.empty())
                cv_work.wait(lk, [&] { return stop || !queue.empty(); });
            else
                cv_work.wait_for(lk, std::chrono::microseconds(200));  // queued, waiting for memory
        }
    }
};

extern "C" {

int lt_server_create(lt_server* out, int nslots, long long hbm_budget_bytes)
{
    if (!out) return set_error(LT_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (nslots < 1 || nslots > 64) return set_error(LT_ERR_INVALID, "nslots must be 1..64");
    if (hbm_budget_bytes <= 0) return set_error(LT_ERR_INVALID, "HBM budget must be positive");
    int dev, st;
    if ((st = current_device(&dev))) return st;
    auto s = std::make_unique<lt_server_s>();
    s->device = dev;
    s->budget = hbm_budget_bytes;
    s->t0 = s->stats_t0 = Clock::now();
    for (int i = 0; i < nslots; ++i) {
        hipStream_t h;
        if (hipStreamCreateWithFlags(&h, hipStreamNonBlocking) != hipSuccess) {
            for (auto x : s->slots) (void)hipStreamDestroy(x);
            return set_error(LT_ERR_HIP, "stream creation failed");
        }
        s->slots.push_back(h);
    }
    s->slot_busy.assign(nslots, false);
    lt_server_s* raw = s.release();
    raw->worker = std::thread([raw] { raw->run(); });
    *out = raw;
    return LT_OK;
}

int lt_server_add_model(lt_server s, const lt_potential_soa* lens, int nhalos, int* model_id)
{
    if (!s || !lens || !model_id) return set_error(LT_ERR_INVALID, "NULL argument");
    (void)hipSetDevice(s->device);
    auto m = std::make_unique<ServedModel>();
    m->nhalos = nhalos;
    int ok = 0;
    for (int u = 0; u < 4; ++u) {
        if (u == 2 && !lens->dlsds) {
            m->use_status[u] = LT_ERR_INVALID;
            m->use_err[u] = "no dlsds: mass maps unavailable";
            continue;
        }
        int st = lt_model_create(&m->use_model[u], lens, nhalos, use_kind(u));
        m->use_status[u] = st;
        if (st) m->use_err[u] = lt_last_error();
        else ++ok;
    }
    if (!ok) return set_error(m->use_status[1], "model has no usable kernel: " + m->use_err[1]);
    std::lock_guard<std::mutex> lk(s->mu);
    s->models.push_back(std::move(m));
    *model_id = (int)s->models.size() - 1;
    return LT_OK;
}

int lt_server_submit(lt_server s, const lt_request* req, long long* ticket)
{
    if (!s || !req || !ticket) return set_error(LT_ERR_INVALID, "NULL argument");
    std::unique_lock<std::mutex> lk(s->mu);
    if (req->model < 0 || req->model >= (int)s->models.size()) return set_error(LT_ERR_NOT_FOUND, "unknown model");
    const ServedModel& m = *s->models[req->model];
    const int ui = use_index(req->kind);
    if (m.use_status[ui] != LT_OK) return set_error(m.use_status[ui], m.use_err[ui]);
    if (req->grid.nx <= 0 || req->grid.ny <= 0) return set_error(LT_ERR_INVALID, "bad grid");
    if (req->precision != LT_PREC_FP64 && req->precision != LT_PREC_FP32C)
        return set_error(LT_ERR_INVALID, "precision must be LT_PREC_FP64 or LT_PREC_FP32C");
    auto t = std::make_unique<Ticket>();
    t->id = s->next_id++;
    t->req = *req;
    t->nplanes = planes_of(req->kind);
    t->plane_bytes = (size_t)req->grid.nx * (size_t)req->grid.ny * sizeof(double);
    if ((long long)(t->plane_bytes * t->nplanes) > s->budget)
        return set_error(LT_ERR_CAPACITY, "request needs more device memory than the server's budget");
    t->halo_px = (double)m.nhalos * req->grid.nx * req->grid.ny;
    if (hipEventCreateWithFlags(&t->done_ev, hipEventDisableTiming) != hipSuccess)
        return set_error(LT_ERR_HIP, "event creation failed");
    t->t_submit = Clock::now();
    *ticket = t->id;
    s->queue.push_back(t->id);
    s->tickets.emplace(t->id, std::move(t));
    ++s->n_submitted;
    lk.unlock();
    s->cv_work.notify_one();
    return LT_OK;
}

int lt_server_wait(lt_server s, long long ticket, double timeout_ms)
{
    if (!s) return set_error(LT_ERR_INVALID, "server is NULL");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return set_error(LT_ERR_NOT_FOUND, "unknown or released ticket");
    Ticket* t = it->second.get();
    auto pred = [&] { return t->state == State::Done || t->state == State::Failed; };
    if (timeout_ms < 0) s->cv_done.wait(lk, pred);
    else if (!s->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), pred))
        return set_error(LT_ERR_BUSY, "not finished");
    if (t->state == State::Failed) return set_error(t->status, t->err);
    return LT_OK;
}

int lt_server_fetch(lt_server s, long long ticket, double* out0, double* out1, double* out2, double* out3,
                    double* out4, lt_request_times* times)
{
    int st = lt_server_wait(s, ticket, -1.0);
    std::unique_ptr<Ticket> t;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->tickets.find(ticket);
        if (it == s->tickets.end()) return st ? st : set_error(LT_ERR_NOT_FOUND, "unknown ticket");
        t = std::move(it->second);
        s->tickets.erase(it);
    }
    double* outs[5] = {out0, out1, out2, out3, out4};
    int cst = st;
    if (st == LT_OK) {
        (void)hipSetDevice(s->device);
        hipStream_t fs = nullptr;
        hipError_t e = hipStreamCreateWithFlags(&fs, hipStreamNonBlocking);
        for (int k = 0; k < t->nplanes && e == hipSuccess; ++k)
            if (outs[k]) e = hipMemcpyAsync(outs[k], t->planes[k], t->plane_bytes, hipMemcpyDeviceToHost, fs);
        if (e == hipSuccess) e = hipStreamSynchronize(fs);
        if (fs) (void)hipStreamDestroy(fs);
        if (e != hipSuccess) cst = hip_error(e, "result copy");
    }
    t->t_fetched = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        for (int k = 0; k < t->nplanes; ++k)
            if (t->planes[k]) s->give_back(t->planes[k], t->plane_bytes);
        if (st == LT_OK) s->lat_fetch.push_back(ms_between(t->t_submit, t->t_fetched));
    }
    if (times) {
        times->submit_ms = ms_between(s->t0, t->t_submit);
        times->start_ms = t->state == State::Done ? ms_between(s->t0, t->t_start) : -1.0;
        times->done_ms = ms_between(s->t0, t->t_done);
        times->fetched_ms = ms_between(s->t0, t->t_fetched);
    }
    (void)hipEventDestroy(t->done_ev);
    s->cv_work.notify_one();
    return cst;
}

int lt_server_get_stats(lt_server s, lt_server_stats* out)
{
    if (!s || !out) return set_error(LT_ERR_INVALID, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    out->submitted = s->n_submitted;
    out->completed = s->n_completed;
    out->failed = s->n_failed;
    out->queued = (long long)s->queue.size();
    out->running = (long long)s->running.size();
    out->p50_ms = percentile(s->lat_done, 50);
    out->p99_ms = percentile(s->lat_done, 99);
    out->p50_fetch_ms = percentile(s->lat_fetch, 50);
    out->p99_fetch_ms = percentile(s->lat_fetch, 99);
    out->halo_px = s->halo_px_done;
    out->elapsed_ms = ms_between(s->stats_t0, Clock::now());
    out->bytes_in_use = s->bytes_used;
    out->bytes_peak = s->bytes_peak;
    return LT_OK;
}

int lt_server_reset_stats(lt_server s)
{
    if (!s) return set_error(LT_ERR_INVALID, "server is NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    s->lat_done.clear();
    s->lat_fetch.clear();
    s->n_submitted = s->n_completed = s->n_failed = 0;
    s->halo_px_done = 0;
    s->bytes_peak = s->bytes_used + s->bytes_cached;
    s->stats_t0 = Clock::now();
    return LT_OK;
}

int lt_server_destroy(lt_server s)
{
    if (!s) return LT_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    (void)hipSetDevice(s->device);
    for (auto h : s->slots) (void)hipStreamSynchronize(h);
    for (auto& kv : s->tickets) {
        Ticket& t = *kv.second;
        for (int k = 0; k < t.nplanes; ++k)
            if (t.planes[k]) (void)hipFree(t.planes[k]);
        if (t.done_ev) (void)hipEventDestroy(t.done_ev);
    }
    for (auto& kv : s->cache) (void)hipFree(kv.second);
    for (auto& m : s->models)
        for (auto h : m->use_model)
            if (h) lt_model_destroy(h);
    for (auto h : s->slots) (void)hipStreamDestroy(h);
    delete s;
    return LT_OK;
}

}  // extern "C"
