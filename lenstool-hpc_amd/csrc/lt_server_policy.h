// lt_server_policy.h -- memory admission of the map-serving loop (lt_server.cpp), kept free of HIP so
// that the CPU tests can drive it (tests/cpp/server_policy_check.cpp).
//
// Result planes of admitted requests are carved out of one arena and stay there until the caller
// fetches them. A request that does not fit waits in the queue; a later request that fits may go
// ahead of it, but only when that cannot starve any request queued before it: for every request Q
// still queued ahead of the candidate, Q must fit in the arena holding only the blocks of requests
// submitted after Q (the candidate's block included, at the place it was given). A caller that
// fetches in submission order has freed everything submitted before Q by the time it waits for Q,
// and the requests after Q hold exactly those blocks, so Q is then admitted: such a caller never
// waits forever. (Without the rule: B (50 %) runs, A (60 %) waits, C (30 %) goes ahead; B is
// fetched, A still finds no 60 % block while C, fetched only after A, holds its 30 %.)
#pragma once

#include <cstddef>
#include <deque>
#include <functional>
#include <iterator>
#include <map>
#include <utility>
#include <vector>

namespace lt {

// First-fit allocator over [0, size): free ranges by offset, coalesced on release.
class Arena {
public:
    void reset(size_t n)
    {
        size_ = n;
        free_.clear();
        if (n) free_.emplace(0, n);
    }
    size_t size() const { return size_; }
    size_t largest() const
    {
        size_t m = 0;
        for (const auto& f : free_) m = m > f.second ? m : f.second;
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
    // mark [off, off + len) used; it must lie inside one free range
    bool occupy(size_t off, size_t len)
    {
        auto it = free_.upper_bound(off);
        if (it == free_.begin()) return false;
        --it;
        const size_t f0 = it->first, f1 = it->first + it->second;
        if (off < f0 || off + len > f1) return false;
        free_.erase(it);
        if (off > f0) free_.emplace(f0, off - f0);
        if (off + len < f1) free_.emplace(off + len, f1 - off - len);
        return true;
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

private:
    size_t size_ = 0;
    std::map<size_t, size_t> free_;  // offset -> length
};

// Blocks held by admitted requests (until fetched) and the admission rule above. Request ids grow
// with submission order.
class Admission {
public:
    void reset(size_t budget)
    {
        arena_.reset(budget);
        held_.clear();
        used_ = 0;
    }
    size_t used() const { return used_; }
    size_t largest_free() const { return arena_.largest(); }
    bool block_of(long long id, size_t* off) const
    {
        auto it = held_.find(id);
        if (it == held_.end()) return false;
        *off = it->second.first;
        return true;
    }

    // One pass over `queue` (ids in submission order, none admitted yet): admits, in order, every
    // request the rule allows and removes it from the queue. need(id) = bytes of its block.
    std::vector<long long> pass(std::deque<long long>& queue, const std::function<size_t(long long)>& need)
    {
        std::vector<long long> admitted;
        std::vector<std::pair<long long, size_t>> waiting;  // (id, need) queued ahead, not admitted
        for (auto it = queue.begin(); it != queue.end();) {
            const long long id = *it;
            const size_t n = need(id);
            size_t off = 0;
            if (!arena_.alloc(n, &off)) {
                waiting.emplace_back(id, n);
                ++it;
                continue;
            }
            held_[id] = {off, n};
            bool ok = true;
            for (const auto& q : waiting)
                if (!fits_after(q.first, q.second)) {
                    ok = false;
                    break;
                }
            if (!ok) {
                held_.erase(id);
                arena_.release(off, n);
                waiting.emplace_back(id, n);
                ++it;
                continue;
            }
            used_ += n;
            admitted.push_back(id);
            it = queue.erase(it);
        }
        return admitted;
    }

    // the request's result was fetched (or it failed): its block returns to the arena
    void release(long long id)
    {
        auto it = held_.find(id);
        if (it == held_.end()) return;
        arena_.release(it->second.first, it->second.second);
        used_ -= it->second.second;
        held_.erase(it);
    }

private:
    // would a request of `need` bytes fit if only the blocks of requests submitted after `id` were held?
    bool fits_after(long long id, size_t need) const
    {
        Arena a;
        a.reset(arena_.size());
        for (auto it = held_.upper_bound(id); it != held_.end(); ++it) a.occupy(it->second.first, it->second.second);
        return a.largest() >= need;
    }

    Arena arena_;
    std::map<long long, std::pair<size_t, size_t>> held_;  // id -> (offset, length)
    size_t used_ = 0;
};

}  // namespace lt
