// Randomised check of the serving arena (first-fit, coalescing): after every
// operation, live ranges are disjoint and inside the arena, the free map has
// no two touching ranges, and free + used bytes equal the arena size.  Exits
// non-zero with a message on the first violation.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "arena.h"

struct Live {
    uint64_t off, bytes;
};

static int fail(const char *what, int step) {
    std::printf("FAIL step %d: %s\n", step, what);
    return 1;
}

int main(int argc, char **argv) {
    const uint64_t size = 1ull << 24;
    const int steps = argc > 1 ? std::atoi(argv[1]) : 20000;
    mlmq::Arena a;
    a.reset(nullptr, size);
    std::mt19937_64 rng(42);
    std::vector<Live> live;
    for (int step = 0; step < steps; ++step) {
        const bool do_alloc = live.empty() || (rng() % 100) < 55;
        if (do_alloc) {
            const uint64_t want = 1 + rng() % (size / 32);
            uint64_t off = 0;
            if (a.alloc(want, &off)) {
                if (off % mlmq::kArenaAlign) return fail("unaligned offset", step);
                live.push_back({off, mlmq::round_up(want)});
            } else {
                // first fit must not refuse when some free range is large enough
                for (auto &kv : a.free_)
                    if (kv.second >= mlmq::round_up(want)) return fail("refused a fitting request", step);
            }
        } else {
            const size_t k = rng() % live.size();
            a.release(live[k].off, live[k].bytes);
            live[k] = live.back();
            live.pop_back();
        }
        uint64_t used = 0;
        for (size_t i = 0; i < live.size(); ++i) {
            used += live[i].bytes;
            if (live[i].off + live[i].bytes > size) return fail("range outside the arena", step);
            for (size_t j = i + 1; j < live.size(); ++j)
                if (live[i].off < live[j].off + live[j].bytes && live[j].off < live[i].off + live[i].bytes)
                    return fail("overlapping live ranges", step);
        }
        if (used != a.used) return fail("used bytes disagree", step);
        uint64_t free_bytes = 0, prev_end = ~0ull;
        for (auto &kv : a.free_) {
            if (kv.first == prev_end) return fail("touching free ranges not coalesced", step);
            prev_end = kv.first + kv.second;
            free_bytes += kv.second;
            for (const Live &l : live)
                if (l.off < kv.first + kv.second && kv.first < l.off + l.bytes)
                    return fail("free range overlaps a live range", step);
        }
        if (free_bytes + used != size) return fail("bytes lost", step);
        if (a.peak < a.used) return fail("peak below used", step);
    }
    for (const Live &l : live) a.release(l.off, l.bytes);
    if (a.free_.size() != 1 || a.free_.begin()->second != size || a.used != 0)
        return fail("arena not whole after releasing everything", steps);
    std::printf("ok %d steps\n", steps);
    return 0;
}
