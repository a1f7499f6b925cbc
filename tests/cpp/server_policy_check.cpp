// server_policy_check.cpp -- drives the serving loop's memory admission (lenstool-hpc_amd/csrc/
// lt_server_policy.h) on the CPU: a caller submits requests of random sizes and fetches them in
// submission order; memory returns to the arena only when a result is fetched. With the admission
// rule the caller never waits for a request that can no longer be admitted; the plain first-fit
// bypass (a later request takes any block that fits) does, e.g. on the sizes 50, 60, 30 % of the
// budget. Prints "rule <deadlocks> <scenarios> <max used> naive <deadlocks>" and exits 0 when the
// rule had no deadlock, never exceeded the budget, and the naive policy deadlocked on 50/60/30.
#include <cstdio>
#include <deque>
#include <random>
#include <vector>

#include "lt_server_policy.h"

using lt::Admission;
using lt::Arena;

// first fit with unlimited bypass (what the rule replaces)
struct Naive {
    Arena a;
    std::vector<std::pair<size_t, size_t>> blocks;  // by id: (off, len), len 0 = none
    void reset(size_t b, size_t n)
    {
        a.reset(b);
        blocks.assign(n, {0, 0});
    }
    std::vector<long long> pass(std::deque<long long>& q, const std::vector<size_t>& need)
    {
        std::vector<long long> out;
        for (auto it = q.begin(); it != q.end();) {
            size_t off;
            if (a.alloc(need[*it], &off)) {
                blocks[*it] = {off, need[*it]};
                out.push_back(*it);
                it = q.erase(it);
            } else {
                ++it;
            }
        }
        return out;
    }
    void release(long long id) { a.release(blocks[id].first, blocks[id].second); }
};

// returns true if the caller deadlocked. Submission runs `ahead` requests in front of the fetches.
template <class P, class Pass>
static bool scenario(P& pol, int n, int ahead, Pass pass)
{
    std::deque<long long> queue;
    std::vector<bool> admitted(n, false);
    int submitted = 0;
    for (int k = 0; k < n; ++k) {
        while (submitted < n && submitted <= k + ahead) queue.push_back(submitted++);
        for (long long id : pass(queue)) admitted[id] = true;
        if (!admitted[k]) return true;  // nothing else can free memory: the caller waits on k forever
        pol.release(k);
    }
    return false;
}

int main()
{
    const size_t B = 1000;
    // the sizes of the example: 50 %, 60 %, 30 %, all submitted before the first fetch
    const std::vector<size_t> ex = {500, 600, 300};
    Naive nv;
    nv.reset(B, ex.size());
    const bool naive_dead = scenario(nv, (int)ex.size(), 2, [&](std::deque<long long>& q) { return nv.pass(q, ex); });
    Admission ad;
    ad.reset(B);
    size_t mu = 0;
    bool rule_dead = scenario(ad, (int)ex.size(), 2, [&](std::deque<long long>& q) {
        auto r = ad.pass(q, [&](long long id) { return ex[id]; });
        if (ad.used() > mu) mu = ad.used();
        return r;
    });
    int dead = rule_dead ? 1 : 0, naive = 0, total = 1;
    std::mt19937_64 rng(7);
    for (int s = 0; s < 20000; ++s) {
        const int n = 2 + (int)(rng() % 12);
        std::vector<size_t> need(n);
        for (auto& v : need) v = 1 + rng() % B;
        const int ahead = (int)(rng() % (n + 1));
        ad.reset(B);
        dead += scenario(ad, n, ahead, [&](std::deque<long long>& q) {
            auto r = ad.pass(q, [&](long long id) { return need[id]; });
            if (ad.used() > mu) mu = ad.used();
            return r;
        });
        nv.reset(B, need.size());
        naive += scenario(nv, n, ahead, [&](std::deque<long long>& q) { return nv.pass(q, need); });
        ++total;
    }
    std::printf("rule %d %d %zu naive %d example_naive %d\n", dead, total, mu, naive, (int)naive_dead);
    return (dead == 0 && mu <= B && naive_dead) ? 0 : 1;
}
