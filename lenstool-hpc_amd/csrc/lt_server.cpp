// lt_server.cpp -- a map-serving loop on one GPU: resident lens models, requests admitted by HBM
// footprint as stream slots and memory free up (continuous batching: work of different requests
// overlaps on the device), results kept in device memory until the caller fetches them.
//
// The reference has no serving loop; its map drivers (utils/exec/main.cpp:397-545,
// utils/maps/main.cpp:428-587) call one map function after another, each allocating, computing and
// copying back synchronously. Here one dispatcher thread owns all launches:
//   submit()  -> queue (FIFO, with bounded bypass so that small requests can pass a large one
//                that does not fit yet)
//   dispatch  -> admit while a slot is free and the request's planes fit in the budget, carve
//                them out of one device arena reserved at creation (so that serving never calls
//                hipMalloc or hipFree, which can stall the device), launch lt_map_device on the
//                slot's stream, record an event
//   poll      -> events that completed mark their tickets done and free their slots
//   fetch()   -> caller thread: wait, copy planes to host on a pooled stream, release memory.
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
    size_t block_off = 0, block_len = 0;  // the planes' block in the arena (block_len 0: none)
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

constexpr size_t kAlign = 256;

size_t align_up(size_t n) { return (n + kAlign - 1) / kAlign * kAlign; }

// First-fit allocator over one device block: free ranges by offset, coalesced on release.
struct Arena {
    char* base = nullptr;
    size_t size = 0;
    std::map<size_t, size_t> free_;  // offset -> length

    void reset(char* b, size_t n)
    {
        base = b;
        size = n;
        free_.clear();
        if (n) free_.emplace(0, n);
    }
    size_t largest() const
    {
        size_t m = 0;
        for (const auto& f : free_) m = std::max(m, f.second);
        return m;
    }
    bool alloc(size_t need, size_t* off)
    {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < need) continue;
            *off = it->first;
            const size_t rest = it->second - need, at = it->first + need;
            free_.erase(it);
            if (rest) free_.emplace(at, rest);
            return true;
        }
        return false;
    }
    void release(size_t off, size_t len)
    {
        auto it = free_.emplace(off, len).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
};

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

    // device memory: one arena of `budget` bytes; bytes_used = blocks held by tickets
    Arena arena;
    long long bytes_used = 0, bytes_peak = 0;
    // streams for the result copies of fetch() (callers may fetch from several threads)
    std::mutex fetch_mu;
    std::vector<hipStream_t> fetch_streams;

    // statistics over completed requests since the last reset
    std::vector<double> lat_done, lat_fetch;
    long long n_submitted = 0, n_completed = 0, n_failed = 0;
    double halo_px_done = 0;
    Clock::time_point stats_t0;

    std::thread worker;

    // ---- memory (called with mu held) -------------------------------------------------------------
    static size_t block_bytes(const Ticket& t) { return (size_t)t.nplanes * align_up(t.plane_bytes); }

    bool take(Ticket& t)
    {
        const size_t need = block_bytes(t);
        size_t off;
        if (!arena.alloc(need, &off)) return false;
        t.block_off = off;
        t.block_len = need;
        for (int k = 0; k < t.nplanes; ++k)
            t.planes[k] = (double*)(arena.base + off + (size_t)k * align_up(t.plane_bytes));
        bytes_used += (long long)need;
        bytes_peak = std::max(bytes_peak, bytes_used);
        return true;
    }

    void give_back(Ticket& t)
    {
        if (!t.block_len) return;
        arena.release(t.block_off, t.block_len);
        bytes_used -= (long long)t.block_len;
        t.block_len = 0;
        for (int k = 0; k < 5; ++k) t.planes[k] = nullptr;
    }

    bool fits(const Ticket& t) const { return arena.largest() >= block_bytes(t); }

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
        if (!take(t)) {
            fail_ticket(t, LT_ERR_NOMEM, "no room for the map planes in the server's arena");
            return;
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
            give_back(t);
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
    void* arena = nullptr;
    if (hipMalloc(&arena, (size_t)hbm_budget_bytes) != hipSuccess) {
        (void)hipGetLastError();
        for (auto x : s->slots) (void)hipStreamDestroy(x);
        return set_error(LT_ERR_NOMEM, "could not reserve the HBM budget (" + std::to_string(hbm_budget_bytes) +
                                           " bytes) for the server's arena");
    }
    s->arena.reset((char*)arena, (size_t)hbm_budget_bytes);
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
    if ((long long)lt_server_s::block_bytes(*t) > s->budget)
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
        hipError_t e = hipSuccess;
        {
            std::lock_guard<std::mutex> fl(s->fetch_mu);
            if (!s->fetch_streams.empty()) {
                fs = s->fetch_streams.back();
                s->fetch_streams.pop_back();
            }
        }
        if (!fs) e = hipStreamCreateWithFlags(&fs, hipStreamNonBlocking);
        for (int k = 0; k < t->nplanes && e == hipSuccess; ++k)
            if (outs[k]) e = hipMemcpyAsync(outs[k], t->planes[k], t->plane_bytes, hipMemcpyDeviceToHost, fs);
        if (e == hipSuccess) e = hipStreamSynchronize(fs);
        if (fs) {
            std::lock_guard<std::mutex> fl(s->fetch_mu);
            s->fetch_streams.push_back(fs);
        }
        if (e != hipSuccess) cst = hip_error(e, "result copy");
    }
    t->t_fetched = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->give_back(*t);
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
    s->bytes_peak = s->bytes_used;
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
        if (t.done_ev) (void)hipEventDestroy(t.done_ev);
    }
    if (s->arena.base) (void)hipFree(s->arena.base);
    for (auto h : s->fetch_streams) (void)hipStreamDestroy(h);
    for (auto& m : s->models)
        for (auto h : m->use_model)
            if (h) lt_model_destroy(h);
    for (auto h : s->slots) (void)hipStreamDestroy(h);
    delete s;
    return LT_OK;
}

}  // extern "C"
