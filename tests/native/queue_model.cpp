// CPU model of the L2 delta-bucket queue and the credit termination protocol
// (mlmq_tpds_withl3_amd/csrc/l2_delta_queue.h and run_worker in
// sssp_kernels.hip), for tests that run without a GPU.
//
// Each wave is a user-level context (ucontext) running the same rules as the
// device code, lane by lane: the bucket choice and clamp, the per-bucket
// reservation, the CAS publication with a bounded spin, the one-snapshot pop
// with persistent claims, the window advance, the credit return when a wave
// runs out of work, and the done-first termination check.  Every access to
// shared memory (a load, store, fetch_add, CAS, exchange or atomicMin of one
// lane) is one step, and a seeded scheduler picks which wave takes the next
// step, so every run is one reproducible interleaving of the waves.
// Accesses of one wave happen in program order (the model does not reorder
// a wave's accesses to different addresses).
//
// Ghost state (invisible to the waves) lets the model check what the device
// cannot observe: the number of records created and not yet fully expanded
// must be zero at the instant the termination check raises exit.
//
// Variants reproduce protocol choices that the device code makes differently
// from its earlier versions and from the reference:
//   * advance_rule 1: a wave holding a claim on the head bucket does not
//     advance the window (the rule before the fix: it can stall);
//   * tails_first 1: the termination check reads the reservations before the
//     credits (the reference's order: it can exit early);
//   * wide_claims 1: a claim may cover more slots than the wave's batch holds
//     (the rule before the fix: the ring can lap under it and strand records).
#include <ucontext.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <deque>
#include <queue>
#include <random>
#include <vector>

namespace {

constexpr uint64_t kEmpty = ~0ull;
constexpr uint32_t kErrNone = 0, kErrOverflow = 1;
constexpr uint32_t kUnreached = 0x7FFFFFFFu;
constexpr int kLanes = 64;
constexpr uint32_t kNarrowQueue = 8;

uint64_t make_rec(uint32_t v, uint32_t d) { return (uint64_t)d << 32 | v; }
uint32_t rec_v(uint64_t r) { return (uint32_t)r; }
// Light/heavy split: a heavy-phase record carries this flag in its distance
// word and is queued one bucket later than its distance.
constexpr uint32_t kHeavyPhase = 0x80000000u;
uint32_t rec_d(uint64_t r) { return (uint32_t)(r >> 32) & ~kHeavyPhase; }
bool rec_heavy(uint64_t r) { return (r >> 32) & kHeavyPhase; }

struct Shared {
    uint32_t nb = 1, window = 1, cap_log2 = 6, delta = 1, batch = 8, spin_limit = 64;
    std::vector<uint32_t> tail, head, taken;
    uint32_t base = 0, done = 0, exit = 0, err = kErrNone;
    std::vector<uint64_t> slots;
    std::vector<uint32_t> dist;
    uint64_t *slot(uint32_t b, uint32_t pos) {
        return &slots[((uint64_t)b << cap_log2) + (pos & ((1u << cap_log2) - 1))];
    }
};

struct Model;

struct Wave {
    ucontext_t ctx{};
    std::vector<char> stack;
    Model *m = nullptr;
    int id = 0;
    bool finished = false;
    std::mt19937_64 rng;
    // claims: bucket b -> (start, pending mask)
    std::vector<uint32_t> cl_start;
    std::vector<uint64_t> cl_mask;
    uint32_t base_hint = 0;
};

struct Model {
    Shared q;
    std::vector<Wave> waves;
    ucontext_t sched{};
    int cur = -1;
    uint64_t steps = 0, max_steps = 0;
    bool out_of_steps = false;
    // scripted interleaving: the first steps go to these waves in order (an
    // entry naming a finished wave is drawn at random instead)
    std::vector<int32_t> script;
    // ghost state
    int64_t outstanding = 0;        // records created and not yet fully expanded
    uint64_t exit_violations = 0;   // exits raised while work was outstanding
    uint64_t exits_raised = 0;
    // settings of the workload
    int advance_rule = 0;   // 1: a claim on the head bucket blocks the advance
    int tails_first = 0;    // 1: termination check reads the tails before done
    uint32_t slow_leader = 1;   // wave 0 takes a step 1/slow_leader as often as the others
    uint32_t pop_max = 32, min_read = 8, spill = 64;
    uint32_t light = 0;   // 1: light/heavy edge split
    uint32_t room = kLanes;   // publication mode: records a consumer takes per pop
    // 1: a claim takes `batch` slots even where that is more than the wave's
    // batch holds (the earlier rule, which can strand records; see l2_pop)
    uint32_t wide_claims = 0;
    // graph (solve mode)
    const int64_t *row = nullptr;
    const int32_t *col = nullptr, *w = nullptr;
    // publication mode
    uint32_t producers = 0, pub_batch = 0, dist0 = 0, span = 1, total = 0;
    std::vector<uint32_t> seen;
    uint32_t consumed = 0, bad = 0, push_fail = 0;

    // One step of the current wave: hand control to the scheduler.
    void yield() {
        Wave &wv = waves[cur];
        swapcontext(&wv.ctx, &sched);
    }
    bool stop() const { return out_of_steps; }
};

Model *g_model = nullptr;

// ---- shared-memory accesses, one step each ----
uint32_t ld(uint32_t *p) {
    g_model->yield();
    return *p;
}
void st(uint32_t *p, uint32_t v) {
    g_model->yield();
    *p = v;
}
uint32_t fadd(uint32_t *p, uint32_t v) {
    g_model->yield();
    const uint32_t o = *p;
    *p = o + v;
    return o;
}
bool cas32(uint32_t *p, uint32_t expect, uint32_t v) {
    g_model->yield();
    if (*p != expect) return false;
    *p = v;
    return true;
}
bool cas64(uint64_t *p, uint64_t expect, uint64_t v) {
    g_model->yield();
    if (*p != expect) return false;
    *p = v;
    return true;
}
uint64_t xchg64(uint64_t *p, uint64_t v) {
    g_model->yield();
    const uint64_t o = *p;
    *p = v;
    return o;
}
uint32_t amin(uint32_t *p, uint32_t v) {
    g_model->yield();
    const uint32_t o = *p;
    if (v < o) *p = v;
    return o;
}

// Lanes of one instruction complete in any order: a random permutation.
std::vector<int> lane_order(Wave &wv, const std::vector<int> &lanes) {
    std::vector<int> o = lanes;
    std::shuffle(o.begin(), o.end(), wv.rng);
    return o;
}

void raise_error(Shared &q, uint32_t code) {
    cas32(&q.err, kErrNone, code);
    st(&q.exit, 1u);
}

// l2_push: records of the lanes in `recs` (at most 64).  False on overflow.
bool l2_push(Wave &wv, const std::vector<uint64_t> &recs) {
    Shared &q = g_model->q;
    if (recs.empty()) return true;
    const uint32_t base = ld(&q.base);
    const int n = (int)recs.size();
    std::vector<uint32_t> b(n);
    for (int i = 0; i < n; ++i) {
        const uint32_t bd = rec_d(recs[i]) + (rec_heavy(recs[i]) ? q.delta : 0u);
        int32_t rel = (int32_t)(bd / q.delta) - (int32_t)base;
        rel = rel < 0 ? 0 : rel;
        rel = rel >= (int32_t)q.nb ? (int32_t)q.nb - 1 : rel;
        b[i] = (base + (uint32_t)rel) % q.nb;
    }
    // group by bucket: rank within the group, one reservation per group
    std::vector<uint32_t> rank(n), leader(n);
    std::vector<int> leaders;
    std::vector<uint32_t> cnt(n, 0);
    for (int i = 0; i < n; ++i) {
        int l = -1;
        for (int j = 0; j < i; ++j)
            if (b[j] == b[i]) {
                l = (int)leader[j];
                break;
            }
        if (l < 0) {
            l = i;
            leaders.push_back(i);
        }
        leader[i] = (uint32_t)l;
        rank[i] = cnt[l]++;
    }
    std::vector<uint32_t> pos0(n, 0);
    for (int l : lane_order(wv, leaders)) pos0[l] = fadd(&q.tail[b[l]], cnt[l]);
    bool ok = true;
    std::vector<int> lanes(n);
    for (int i = 0; i < n; ++i) lanes[i] = i;
    for (int i : lane_order(wv, lanes)) {
        uint64_t *slot = q.slot(b[i], pos0[leader[i]] + rank[i]);
        uint32_t spins = 0;
        while (!cas64(slot, kEmpty, recs[i])) {
            if (++spins > q.spin_limit || g_model->stop()) {
                ok = false;
                break;
            }
        }
        if (!ok) break;
    }
    if (!ok) {
        raise_error(q, kErrOverflow);
        return false;
    }
    return true;
}

// l2_pop: up to `room` records into `stage`.  `recs` is the most records a
// batch of this wave holds (the room of an idle wave): a new claim takes at
// most that many slots, so an idle wave tries every slot it still holds.
uint32_t l2_pop(Wave &wv, std::vector<uint64_t> &stage, uint32_t room, uint32_t recs) {
    Shared &q = g_model->q;
    const uint32_t nb = q.nb;
    const bool wide = nb > kNarrowQueue;
    std::vector<bool> rd(nb);
    std::vector<int> rdl;
    for (uint32_t b = 0; b < nb; ++b) {
        rd[b] = !wide || (b + nb - wv.base_hint % nb) % nb <= q.window;
        if (rd[b]) rdl.push_back((int)b);
    }
    // one round trip: every read bucket's counters, and the epoch
    std::vector<uint32_t> t(nb, 0), h(nb, 0), k(nb, 0);
    std::vector<int> ops;
    for (int b : rdl)
        for (int c = 0; c < 3; ++c) ops.push_back(b * 3 + c);
    ops.push_back(-1);
    uint32_t base = 0;
    for (int op : lane_order(wv, ops)) {
        if (op < 0) {
            base = ld(&q.base);
            continue;
        }
        const int b = op / 3, c = op % 3;
        if (c == 0) t[b] = ld(&q.tail[b]);
        else if (c == 1) h[b] = ld(&q.head[b]);
        else k[b] = ld(&q.taken[b]);
    }
    std::vector<bool> had_claim(nb);
    for (uint32_t b = 0; b < nb; ++b) had_claim[b] = wv.cl_mask[b] != 0;
    uint32_t got = 0;
    for (uint32_t i = 0; i < q.window && got == 0; ++i) {
        const uint32_t b = (base + i) % nb;
        const bool avail_b = rd[b] && (int32_t)(t[b] - h[b]) > 0;
        if (!had_claim[b] && !avail_b) continue;
        uint32_t s = wv.cl_start[b];
        uint64_t mk = wv.cl_mask[b];
        if (!had_claim[b]) {
            const uint32_t avail = t[b] - h[b];
            uint32_t kk = avail < q.batch ? avail : q.batch;
            if (!g_model->wide_claims) kk = std::min(kk, recs);
            s = fadd(&q.head[b], kk);
            mk = kk >= 64 ? ~0ull : ((1ull << kk) - 1);
        }
        uint64_t try_mask = mk;
        while ((uint32_t)__builtin_popcountll(try_mask) > room)
            try_mask &= ~(1ull << (63 - __builtin_clzll(try_mask)));
        std::vector<int> lanes;
        for (int l = 0; l < kLanes; ++l)
            if ((try_mask >> l) & 1) lanes.push_back(l);
        std::vector<uint64_t> v(kLanes, kEmpty);
        for (int l : lane_order(wv, lanes)) v[l] = xchg64(q.slot(b, s + l), kEmpty);
        uint64_t tk = 0;
        for (int l = 0; l < kLanes; ++l)
            if (v[l] != kEmpty) tk |= 1ull << l;
        mk &= ~tk;
        wv.cl_start[b] = s;
        wv.cl_mask[b] = mk;
        if (tk) {
            for (int l = 0; l < kLanes; ++l)
                if (v[l] != kEmpty) stage.push_back(v[l]);
            got = (uint32_t)__builtin_popcountll(tk);
            fadd(&q.taken[b], got);
        }
    }
    // window advance
    const uint32_t b0 = base % nb;
    if (rd[b0] && t[b0] - k[b0] == 0 && !(g_model->advance_rule == 1 && had_claim[b0])) {
        uint32_t total = 0;
        for (uint32_t b = 0; b < nb; ++b) total += rd[b] ? t[b] - k[b] : 0u;
        if (wide && total == 0) {
            std::vector<int> rest;
            for (uint32_t b = 0; b < nb; ++b)
                if (!rd[b]) rest.push_back((int)b);
            std::vector<uint32_t> t2(nb, 0), k2(nb, 0);
            std::vector<int> ops2;
            for (int b : rest) {
                ops2.push_back(b * 2);
                ops2.push_back(b * 2 + 1);
            }
            for (int op : lane_order(wv, ops2)) {
                if (op % 2 == 0) t2[op / 2] = ld(&q.tail[op / 2]);
                else k2[op / 2] = ld(&q.taken[op / 2]);
            }
            for (int b : rest) total += t2[b] - k2[b];
        }
        if (total > 0) cas32(&q.base, base, base + 1);
    }
    wv.base_hint = base;
    return got;
}

// Termination check of the leader wave.
bool check_done(Wave &wv) {
    Shared &q = g_model->q;
    std::vector<int> bs(q.nb);
    for (uint32_t b = 0; b < q.nb; ++b) bs[b] = (int)b;
    uint32_t done = 0, reserved = 0;
    if (!g_model->tails_first) done = ld(&q.done);
    for (int b : lane_order(wv, bs)) reserved += ld(&q.tail[b]);
    if (g_model->tails_first) done = ld(&q.done);
    if (reserved == done) {
        // ghost check: nothing may be queued or in expansion at this instant
        ++g_model->exits_raised;
        if (g_model->outstanding != 0) ++g_model->exit_violations;
        st(&q.exit, 1u);
        return true;
    }
    return false;
}

// ---- solve mode: the worker loop (run_worker) over a CSR graph ----
void l1_push_slf(std::deque<uint64_t> &l1, uint64_t r) {
    if (!l1.empty() && rec_d(r) < rec_d(l1.front())) l1.push_front(r);
    else l1.push_back(r);
}

void solve_wave(Wave &wv) {
    Model &m = *g_model;
    Shared &q = m.q;
    std::deque<uint64_t> l1;
    uint32_t credit = 0;
    uint64_t batches = 0;
    for (;;) {
        if (m.stop()) return;
        std::vector<uint64_t> in;
        // A batch holds at most 64 records, or 32 when a pop takes at most 32
        // (the device then runs two lanes per record): L2 top-ups stop there too.
        const uint32_t recs = m.pop_max <= 32 ? 32u : (uint32_t)kLanes;
        while (!l1.empty() && in.size() < std::min(m.pop_max, recs)) {
            in.push_back(l1.front());
            l1.pop_front();
        }
        if (in.size() < std::min(m.min_read, recs)) {
            std::vector<uint64_t> stage;
            const uint32_t got = l2_pop(wv, stage, recs - (uint32_t)in.size(), recs);
            credit += got;
            in.insert(in.end(), stage.begin(), stage.end());
        }
        if (in.empty()) {
            if (credit) {
                fadd(&q.done, credit);
                credit = 0;
            }
            if (wv.id == 0 && check_done(wv)) return;
            if (ld(&q.exit)) return;
            continue;
        }
        ++batches;
        std::vector<uint64_t> out, heavy;
        for (uint64_t r : in) {
            const uint32_t v = rec_v(r), d = rec_d(r);
            const bool hphase = m.light && rec_heavy(r);
            const uint32_t cur = ld(&q.dist[v]);
            if (d <= cur) {
                bool need_heavy = false;
                for (int64_t e = m.row[v]; e < m.row[v + 1]; ++e) {
                    if (m.light) {
                        const bool hv = (uint32_t)m.w[e] >= q.delta;
                        if (hv && !hphase) need_heavy = true;
                        if (hv != hphase) continue;
                    }
                    const uint32_t u = (uint32_t)m.col[e];
                    const uint32_t nd = d + (uint32_t)m.w[e];
                    const uint32_t old = amin(&q.dist[u], nd);
                    if (nd < old) {
                        out.push_back(make_rec(u, nd));
                        ++m.outstanding;   // ghost: a new record exists
                    }
                }
                if (need_heavy) {
                    heavy.push_back(make_rec(v, d | kHeavyPhase));
                    ++m.outstanding;
                }
            }
        }
        for (uint64_t r : out) l1_push_slf(l1, r);
        // heavy-phase records go to L2, one bucket later than their distance
        for (size_t i = 0; i < heavy.size(); i += kLanes) {
            std::vector<uint64_t> chunk(heavy.begin() + i,
                                        heavy.begin() + std::min(heavy.size(), i + kLanes));
            if (!l2_push(wv, chunk)) return;
        }
        // the batch's records are expanded: their children are held by L1
        m.outstanding -= (int64_t)in.size();
        if (l1.size() > m.spill) {
            std::vector<uint64_t> all(l1.begin(), l1.end());
            l1.clear();
            for (size_t i = 0; i < all.size(); i += kLanes) {
                std::vector<uint64_t> chunk(all.begin() + i,
                                            all.begin() + std::min(all.size(), i + kLanes));
                if (!l2_push(wv, chunk)) return;
            }
        }
        if ((batches & 31) == 0 && ld(&q.exit)) return;
    }
}

// ---- publication mode (the device queue check's scenario) ----
uint32_t dist_of(uint32_t id, uint32_t dist0, uint32_t span) {
    return dist0 + (uint32_t)(((uint64_t)id * 2654435761ull) % span);
}

void publication_wave(Wave &wv) {
    Model &m = *g_model;
    if ((uint32_t)wv.id < m.producers) {
        std::vector<uint64_t> recs;
        for (uint32_t l = 0; l < m.pub_batch; ++l) {
            const uint32_t id = (uint32_t)wv.id * m.pub_batch + l;
            recs.push_back(make_rec(id, dist_of(id, m.dist0, m.span)));
        }
        if (!l2_push(wv, recs)) ++m.push_fail;
        return;
    }
    for (;;) {
        if (m.stop()) return;
        if (ld(&m.consumed) >= m.total) return;
        std::vector<uint64_t> stage;
        const uint32_t got = l2_pop(wv, stage, m.room, m.room);
        for (uint64_t r : stage) {
            const uint32_t v = rec_v(r);
            if (v < m.total && rec_d(r) == dist_of(v, m.dist0, m.span)) ++m.seen[v];
            else ++m.bad;
        }
        if (got) fadd(&m.consumed, got);
    }
}

int g_mode = 0;   // 0 solve, 1 publication

void wave_entry(int id) {
    Wave &wv = g_model->waves[id];
    if (g_mode == 0) solve_wave(wv);
    else publication_wave(wv);
    wv.finished = true;
    // back to the scheduler for good
    swapcontext(&wv.ctx, &g_model->sched);
}

// Run every wave to completion under the seeded scheduler.  Returns false if
// the step budget ran out first.
bool schedule(Model &m, uint64_t seed) {
    std::mt19937_64 rng(seed);
    for (size_t i = 0; i < m.waves.size(); ++i) {
        Wave &wv = m.waves[i];
        wv.m = &m;
        wv.id = (int)i;
        wv.rng.seed(seed * 7919 + i + 1);
        wv.stack.assign(1 << 18, 0);
        wv.cl_start.assign(m.q.nb, 0);
        wv.cl_mask.assign(m.q.nb, 0);
        getcontext(&wv.ctx);
        wv.ctx.uc_stack.ss_sp = wv.stack.data();
        wv.ctx.uc_stack.ss_size = wv.stack.size();
        wv.ctx.uc_link = nullptr;
        makecontext(&wv.ctx, (void (*)())wave_entry, 1, (int)i);
    }
    g_model = &m;
    std::vector<int> live;
    for (;;) {
        live.clear();
        for (size_t i = 0; i < m.waves.size(); ++i)
            if (!m.waves[i].finished) live.push_back((int)i);
        if (live.empty()) break;
        if (++m.steps > m.max_steps) m.out_of_steps = true;   // waves now return at their next check
        int pick = live[rng() % live.size()];
        if (m.steps <= m.script.size()) {
            const int32_t sp = m.script[m.steps - 1];
            if (sp >= 0 && (size_t)sp < m.waves.size() && !m.waves[(size_t)sp].finished) pick = sp;
        }
        // a slow leader: its turn is passed on to another wave (drawn again)
        // with probability 1 - 1/slow_leader
        if (pick == 0 && live.size() > 1 && m.slow_leader > 1 && rng() % m.slow_leader != 0)
            pick = live[1 + rng() % (live.size() - 1)];
        m.cur = pick;
        swapcontext(&m.sched, &m.waves[m.cur].ctx);
    }
    g_model = nullptr;
    return !m.out_of_steps;
}

void init_queue(Shared &q, uint32_t nb, uint32_t window, uint32_t cap_log2, uint32_t delta,
                uint32_t batch, uint32_t spin_limit, uint32_t base) {
    q.nb = nb;
    q.window = window;
    q.cap_log2 = cap_log2;
    q.delta = delta;
    q.batch = batch;
    q.spin_limit = spin_limit;
    q.tail.assign(nb, 0);
    q.head.assign(nb, 0);
    q.taken.assign(nb, 0);
    q.base = base;
    q.slots.assign((size_t)nb << cap_log2, kEmpty);
}

}  // namespace

extern "C" {

// One solve of the graph (row int64[n+1], col, w int32) from `src` by
// `waves` (1..512) model waves.  cfg: {buckets, window, cap_log2, delta, claim,
// spin_limit, pop_max, min_read, spill, advance_rule, tails_first,
// slow_leader (wave 0 steps 1/slow_leader as often as the others; 0 or 1 = as
// often), light (1: light/heavy edge split), wide_claims (1: the earlier claim
// rule)}.
// out (uint64[12]): {finished (0 = step budget ran out), steps, err,
// exits_raised, exit_violations, sum tail, sum taken, done, slots not empty,
// outstanding at the end, exit flag, 0}.  dist_out: n distances.
int qm_solve(int32_t n, const int64_t *row, const int32_t *col, const int32_t *w, int32_t src,
             int32_t waves, const uint32_t *cfg, uint64_t seed, uint64_t max_steps,
             uint32_t *dist_out, uint64_t *out) {
    if (n <= 0 || src < 0 || src >= n || waves < 1 || waves > 512 || cfg[0] < 1 || cfg[0] > 64 ||
        cfg[1] < 1 || cfg[1] > cfg[0] || cfg[2] > 20 || cfg[3] == 0 || cfg[4] < 1 || cfg[4] > 64)
        return -1;
    Model m;
    g_mode = 0;
    init_queue(m.q, cfg[0], cfg[1], cfg[2], cfg[3], cfg[4], cfg[5], 0);
    m.pop_max = cfg[6];
    m.min_read = cfg[7];
    m.spill = cfg[8];
    m.advance_rule = (int)cfg[9];
    m.tails_first = (int)cfg[10];
    m.slow_leader = cfg[11] > 1 ? cfg[11] : 1;
    m.light = cfg[12];
    m.wide_claims = cfg[13];
    m.row = row;
    m.col = col;
    m.w = w;
    m.q.dist.assign((size_t)n, kUnreached);
    m.q.dist[src] = 0;
    // seed: (src, 0) in slot 0 of bucket 0, one reservation
    m.q.slots[0] = make_rec((uint32_t)src, 0);
    m.q.tail[0] = 1;
    m.outstanding = 1;
    m.max_steps = max_steps;
    m.waves.resize((size_t)waves);
    const bool fin = schedule(m, seed);
    uint64_t t = 0, k = 0, full = 0;
    for (uint32_t b = 0; b < m.q.nb; ++b) {
        t += m.q.tail[b];
        k += m.q.taken[b];
    }
    for (uint64_t s : m.q.slots) full += s != kEmpty;
    const uint64_t o[12] = {fin ? 1u : 0u, m.steps, m.q.err, m.exits_raised, m.exit_violations, t, k,
                            m.q.done, full, (uint64_t)m.outstanding, m.q.exit, 0};
    std::memcpy(out, o, sizeof o);
    std::memcpy(dist_out, m.q.dist.data(), sizeof(uint32_t) * (size_t)n);
    return 0;
}

// The device queue check's scenario: `producers` waves push `batch` records
// each while `consumers` waves pop until all are consumed.  cfg as qm_solve
// (pop/min_read/spill/tails_first unused), plus cfg[11] = initial epoch,
// cfg[12] = records a consumer takes per pop (1..64; 0 = 64), cfg[13] as
// qm_solve's wide_claims.
// out (uint64[8]): {finished, steps, consumed, bad, push_fail, records seen
// != 1, sum tail, sum taken}.
// `script` (may be null): the wave (producers first, then consumers) that
// takes each of the first `script_len` steps.
int qm_publication_scripted(uint32_t producers, uint32_t batch, uint32_t consumers, uint32_t dist0,
                            uint32_t span, const uint32_t *cfg, uint64_t seed, uint64_t max_steps,
                            const int32_t *script, uint32_t script_len, uint64_t *out) {
    if (producers < 1 || consumers < 1 || producers + consumers > 256 || batch < 1 || batch > 64 ||
        span == 0 || cfg[0] < 1 || cfg[0] > 64 || cfg[1] < 1 || cfg[1] > cfg[0] || cfg[2] > 20 ||
        cfg[3] == 0 || cfg[4] < 1 || cfg[4] > 64 || cfg[12] > 64)
        return -1;
    Model m;
    g_mode = 1;
    init_queue(m.q, cfg[0], cfg[1], cfg[2], cfg[3], cfg[4], cfg[5], cfg[11]);
    m.advance_rule = (int)cfg[9];
    m.room = cfg[12] ? cfg[12] : (uint32_t)kLanes;
    m.wide_claims = cfg[13];
    m.producers = producers;
    m.pub_batch = batch;
    m.dist0 = dist0;
    m.span = span;
    m.total = producers * batch;
    m.seen.assign(m.total, 0);
    m.max_steps = max_steps;
    if (script) m.script.assign(script, script + script_len);
    m.waves.resize(producers + consumers);
    const bool fin = schedule(m, seed);
    uint64_t not_once = 0, t = 0, k = 0;
    for (uint32_t c : m.seen) not_once += c != 1;
    for (uint32_t b = 0; b < m.q.nb; ++b) {
        t += m.q.tail[b];
        k += m.q.taken[b];
    }
    const uint64_t o[8] = {fin ? 1u : 0u, m.steps, m.consumed, m.bad, m.push_fail, not_once, t, k};
    std::memcpy(out, o, sizeof o);
    return 0;
}

int qm_publication(uint32_t producers, uint32_t batch, uint32_t consumers, uint32_t dist0,
                   uint32_t span, const uint32_t *cfg, uint64_t seed, uint64_t max_steps,
                   uint64_t *out) {
    return qm_publication_scripted(producers, batch, consumers, dist0, span, cfg, seed, max_steps,
                                   nullptr, 0, out);
}

}  // extern "C"
