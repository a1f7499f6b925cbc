// lt_server.cpp -- a map-serving loop on one GPU: resident lens models, requests admitted by HBM
// footprint, cut into row tiles that are interleaved over the stream slots (continuous batching:
// work of different requests overlaps on the device, and a small request does not wait behind the
// whole of a large one), results kept in device memory until the caller fetches them.
//
// The reference has no serving loop; its map drivers (utils/exec/main.cpp:397-545,
// utils/maps/main.cpp:428-587) call one map function after another, each allocating, computing and
// copying back synchronously. Here one dispatcher thread owns all launches:
//   submit()  -> queue, in submission order
//   admit     -> a request whose planes fit in the arena reserved at creation gets its block
//                (lt_server_policy.h: a later request goes ahead of an earlier one that does not fit
//                only when that cannot starve the earlier one), so serving never calls hipMalloc or
//                hipFree, which can stall the device
//   tiles     -> each free slot takes the next row tile of the admitted requests in turn
//                (round robin), launched with lt_map_device on the slot's stream (rows through
//                jstart: every pixel is computed by the same code as in a whole-grid launch), and an
//                event marks the tile's end
//   poll      -> a request whose tiles have all completed is done
//   fetch()   -> caller thread: wait, copy planes to host on a pooled stream, release the block.
// A request's result is therefore the same alone and in any mix.
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
#include "lt_server_policy.h"

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
    bool has_block = false;
    double* planes[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    int rows_per_tile = 0, next_row = 0, tiles_total = 0, tiles_issued = 0, tiles_done = 0;
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

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t ev = nullptr;
    long long ticket = 0;  // 0: idle
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

// Rows per tile: about kTileWork Hessian-equivalent halo x pixels per launch (~2 ms of FP64 Hessian
// work on the MI355X), from the measured cost of each kernel family per halo x pixel relative to
// the FP64 Hessian; at least 4 rows (one work-group row).
constexpr double kTileWork = 1.0e9;

double relative_cost(int kind, int precision)
{
    const bool f32 = precision == LT_PREC_FP32C;
    switch (kind) {
    case LT_MAP_GRADIENT:
    case LT_MAP_DPL: return f32 ? 0.85 : 2.0;
    case LT_MAP_POT: return f32 ? 1.35 : 3.9;
    default: return f32 ? 0.6 : 1.0;
    }
}

int tile_rows(const lt_request& r, int nhalos)
{
    const double per_row = (double)r.grid.nx * std::max(nhalos, 1) * relative_cost(r.kind, r.precision);
    long long rows = (long long)(kTileWork / per_row);
    rows = std::max(4LL, rows / 4 * 4);
    return (int)std::min<long long>(rows, r.grid.ny);
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

}  // namespace

struct lt_server_s {
    int device = 0;
    long long budget = 0;
    Clock::time_point t0;
    std::vector<Slot> slots;
    std::vector<std::unique_ptr<ServedModel>> models;

    std::mutex mu;
    std::condition_variable cv_work;  // dispatcher wakes on submit / release / stop
    std::condition_variable cv_done;  // waiters wake on completion, admission and release
    bool stop = false;
    long long next_id = 1;
    std::deque<long long> queue;                // submitted, not admitted (submission order)
    std::vector<long long> active;              // admitted, rows left to issue (admission order)
    size_t rr = 0;                              // round-robin position in `active`
    std::unordered_map<long long, std::unique_ptr<Ticket>> tickets;
    long long passes = 0;                       // dispatcher passes (for the stuck-wait check)
    int fetching = 0;                           // fetches between their wait and their release

    // device memory: one arena of `budget` bytes (lt_server_policy.h)
    char* base = nullptr;
    Admission adm;
    long long bytes_peak = 0;
    // streams for the result copies of fetch() (callers may fetch from several threads)
    std::mutex fetch_mu;
    std::vector<hipStream_t> fetch_streams;

    // statistics over completed requests since the last reset
    std::vector<double> lat_done, lat_fetch;
    long long n_submitted = 0, n_completed = 0, n_failed = 0;
    double halo_px_done = 0;
    Clock::time_point stats_t0;

    std::thread worker;

    static size_t block_bytes(const Ticket& t) { return (size_t)t.nplanes * align_up(t.plane_bytes); }

    // ---- dispatcher (all with mu held) ---------------------------------------------------------------
    void fail_ticket(Ticket& t, int st, const std::string& msg)
    {
        if (t.state == State::Failed) return;
        t.state = State::Failed;
        t.status = st;
        t.err = msg;
        t.t_done = Clock::now();
        ++n_failed;
    }

    bool busy_with(long long id) const
    {
        for (const Slot& s : slots)
            if (s.ticket == id) return true;
        return false;
    }

    // a request whose tiles are all accounted for (or that failed with none in flight) is finished
    void settle(Ticket& t)
    {
        if (t.state == State::Running && t.tiles_done == t.tiles_total) {
            t.state = State::Done;
            t.t_done = Clock::now();
            ++n_completed;
            halo_px_done += t.halo_px;
            lat_done.push_back(ms_between(t.t_submit, t.t_done));
        }
    }

    bool poll_slots()
    {
        bool any = false;
        for (Slot& s : slots) {
            if (!s.ticket) continue;
            const hipError_t q = hipEventQuery(s.ev);
            if (q == hipErrorNotReady) continue;
            Ticket& t = *tickets[s.ticket];
            s.ticket = 0;
            any = true;
            ++t.tiles_done;
            if (q != hipSuccess) {
                (void)hipGetLastError();
                fail_ticket(t, LT_ERR_HIP, hipGetErrorString(q));
            }
            settle(t);
        }
        return any;
    }

    bool admit()
    {
        const std::vector<long long> ids =
            adm.pass(queue, [this](long long id) { return block_bytes(*tickets[id]); });
        for (long long id : ids) {
            Ticket& t = *tickets[id];
            size_t off = 0;
            adm.block_of(id, &off);
            t.has_block = true;
            for (int k = 0; k < t.nplanes; ++k) t.planes[k] = (double*)(base + off + (size_t)k * align_up(t.plane_bytes));
            t.state = State::Running;
            t.t_start = Clock::now();
            active.push_back(id);
        }
        bytes_peak = std::max(bytes_peak, (long long)adm.used());
        return !ids.empty();
    }

    // launch the next row tile of request t on slot s
    void launch_tile(Ticket& t, Slot& s)
    {
        ServedModel& m = *models[t.req.model];
        const int ui = use_index(t.req.kind);
        const int rows = std::min(t.rows_per_tile, t.req.grid.ny - t.next_row);
        lt_grid_desc g = t.req.grid;
        g.pitch = g.nx;
        g.ny = rows;
        g.jstart = t.req.grid.jstart + t.next_row;
        const size_t off = (size_t)t.next_row * (size_t)g.nx;
        double* p[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
        for (int k = 0; k < t.nplanes; ++k) p[k] = t.planes[k] + off;
        t.next_row += rows;
        ++t.tiles_issued;
        int st = lt_map_device(m.use_model[ui], t.req.kind, t.req.mode, p[0], p[1], p[2], p[3], p[4], t.req.r,
                               t.req.mult, &g, t.req.precision, 0, s.stream);
        if (st == LT_OK && hipEventRecord(s.ev, s.stream) != hipSuccess) st = LT_ERR_HIP;
        if (st != LT_OK) {
            // nothing queued for this tile: count it as ended, issue no more of this request
            ++t.tiles_done;
            t.tiles_total = t.tiles_issued;
            t.next_row = t.req.grid.ny;
            fail_ticket(t, st, lt_last_error());
            return;
        }
        s.ticket = t.id;
    }

    bool schedule()
    {
        bool any = false;
        for (Slot& s : slots) {
            if (s.ticket) continue;
            // next request (round robin) with rows left to issue
            Ticket* pick = nullptr;
            for (size_t k = 0; k < active.size() && !pick; ++k) {
                const size_t i = (rr + k) % active.size();
                Ticket& t = *tickets[active[i]];
                if (t.state == State::Running && t.next_row < t.req.grid.ny) {
                    pick = &t;
                    rr = i + 1;
                }
            }
            if (!pick) break;
            launch_tile(*pick, s);
            any = true;
        }
        // requests with nothing left to issue leave the round robin
        active.erase(std::remove_if(active.begin(), active.end(),
                                    [this](long long id) {
                                        const Ticket& t = *tickets[id];
                                        return t.state != State::Running || t.next_row >= t.req.grid.ny;
                                    }),
                     active.end());
        if (rr >= active.size()) rr = 0;
        // failed requests with no tile in flight are settled; their blocks wait for fetch
        return any;
    }

    bool idle_slots() const
    {
        for (const Slot& s : slots)
            if (s.ticket) return false;
        return true;
    }

    void run()
    {
        (void)hipSetDevice(device);
        std::unique_lock<std::mutex> lk(mu);
        while (!stop) {
            const bool done = poll_slots();
            const bool admitted = admit();
            const bool launched = schedule();
            ++passes;
            if (done || admitted || launched) cv_done.notify_all();
            if (!idle_slots())
                cv_work.wait_for(lk, std::chrono::microseconds(50));
            else if (queue.empty() && active.empty())
                cv_work.wait(lk, [&] { return stop || !queue.empty(); });
            else
                cv_work.wait_for(lk, std::chrono::microseconds(200));  // queued, waiting for memory
        }
    }

    // a queued request that nothing in progress can admit: only a fetch of another result could
    bool stuck(const Ticket& t) const
    {
        return t.state == State::Queued && idle_slots() && active.empty() && fetching == 0;
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
    auto drop_slots = [&]() {
        for (auto& x : s->slots) {
            if (x.ev) (void)hipEventDestroy(x.ev);
            if (x.stream) (void)hipStreamDestroy(x.stream);
        }
    };
    for (int i = 0; i < nslots; ++i) {
        Slot sl;
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            if (sl.stream) (void)hipStreamDestroy(sl.stream);
            drop_slots();
            return set_error(LT_ERR_HIP, "stream / event creation failed");
        }
        s->slots.push_back(sl);
    }
    void* arena = nullptr;
    if (hipMalloc(&arena, (size_t)hbm_budget_bytes) != hipSuccess) {
        (void)hipGetLastError();
        drop_slots();
        return set_error(LT_ERR_NOMEM, "could not reserve the HBM budget (" + std::to_string(hbm_budget_bytes) +
                                           " bytes) for the server's arena");
    }
    s->base = (char*)arena;
    s->adm.reset((size_t)hbm_budget_bytes);
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
    t->rows_per_tile = tile_rows(*req, m.nhalos);
    t->tiles_total = (req->grid.ny + t->rows_per_tile - 1) / t->rows_per_tile;
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
    // a failed request is finished once none of its tiles is still writing into its block
    auto finished = [&] {
        return t->state == State::Done || (t->state == State::Failed && t->tiles_done == t->tiles_issued);
    };
    const auto deadline = Clock::now() + std::chrono::duration<double, std::milli>(timeout_ms < 0 ? 0 : timeout_ms);
    long long stuck_since = -1;
    while (!finished()) {
        // a request that no work in progress can admit waits for some other result to be fetched; if
        // that stays so over two dispatcher passes with no fetch under way, waiting would never end
        if (s->stuck(*t)) {
            if (stuck_since < 0) stuck_since = s->passes;
            else if (s->passes >= stuck_since + 2)
                return set_error(LT_ERR_BUSY, "request cannot be admitted until results submitted before it are "
                                              "fetched (its planes do not fit in the arena beside them)");
        } else {
            stuck_since = -1;
        }
        if (timeout_ms >= 0 && Clock::now() >= deadline) return set_error(LT_ERR_BUSY, "not finished");
        s->cv_done.wait_for(lk, std::chrono::milliseconds(1));
        // another thread may have fetched (and so released) the ticket while this one slept
        it = s->tickets.find(ticket);
        if (it == s->tickets.end()) return set_error(LT_ERR_NOT_FOUND, "ticket released while waiting");
        t = it->second.get();
    }
    if (t->state == State::Failed) return set_error(t->status, t->err);
    return LT_OK;
}

int lt_server_fetch(lt_server s, long long ticket, double* out0, double* out1, double* out2, double* out3,
                    double* out4, lt_request_times* times)
{
    int st = lt_server_wait(s, ticket, -1.0);
    if (st == LT_ERR_BUSY || st == LT_ERR_NOT_FOUND || st == LT_ERR_INVALID) return st;
    std::unique_ptr<Ticket> t;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->tickets.find(ticket);
        if (it == s->tickets.end()) return st ? st : set_error(LT_ERR_NOT_FOUND, "unknown ticket");
        t = std::move(it->second);
        s->tickets.erase(it);
        ++s->fetching;
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
        if (t->has_block) s->adm.release(t->id);
        --s->fetching;
        if (st == LT_OK) s->lat_fetch.push_back(ms_between(t->t_submit, t->t_fetched));
    }
    if (times) {
        times->submit_ms = ms_between(s->t0, t->t_submit);
        times->start_ms = t->state == State::Done ? ms_between(s->t0, t->t_start) : -1.0;
        times->done_ms = ms_between(s->t0, t->t_done);
        times->fetched_ms = ms_between(s->t0, t->t_fetched);
    }
    s->cv_work.notify_one();
    s->cv_done.notify_all();
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
    long long busy = 0;
    for (const Slot& sl : s->slots) busy += sl.ticket != 0;
    out->running = busy;
    out->p50_ms = percentile(s->lat_done, 50);
    out->p99_ms = percentile(s->lat_done, 99);
    out->p50_fetch_ms = percentile(s->lat_fetch, 50);
    out->p99_fetch_ms = percentile(s->lat_fetch, 99);
    out->halo_px = s->halo_px_done;
    out->elapsed_ms = ms_between(s->stats_t0, Clock::now());
    out->bytes_in_use = (long long)s->adm.used();
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
    s->bytes_peak = (long long)s->adm.used();
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
    for (auto& sl : s->slots) (void)hipStreamSynchronize(sl.stream);
    if (s->base) (void)hipFree(s->base);
    for (auto h : s->fetch_streams) (void)hipStreamDestroy(h);
    for (auto& m : s->models)
        for (auto h : m->use_model)
            if (h) lt_model_destroy(h);
    for (auto& sl : s->slots) {
        (void)hipEventDestroy(sl.ev);
        (void)hipStreamDestroy(sl.stream);
    }
    delete s;
    return LT_OK;
}

}  // extern "C"
