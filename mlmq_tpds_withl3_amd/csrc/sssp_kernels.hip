// Persistent MLMQ SSSP kernels for gfx950 (CDNA4, wave64).
//
// One grid of role-free persistent workgroups (kWavesPerBlock waves each).
// Every wave is a worker running the loop of the reference's
// work_block_kernel (SS/SSSP/sssp_run.cu:507-584):
//   1. pop up to 64 records from its L1 SLF deque; below `min_read` records,
//      top up from the L2 delta-bucket queue (core/src/ml_queue.cu:223-319);
//   2. relax every out-edge of the batch (simple_process, :143-413) with a
//      load-balanced wave-wide edge loop and agent-scope atomicMin;
//   3. push improved (vertex, dist) records through L1, spilling to L2
//      (core/src/ml_queue.cu:321-355);
//   4. return the L2 credit when the wave runs out of work (the reference
//      returns it whenever L1 is empty, :565-572).
// Global wave 0 also detects termination while idle, replacing the separate
// manager kernel (:623-674).  Queue publication needs no manager at all (see
// l2_delta_queue.h), so no wave waits on a particular other workgroup: the
// solve completes even if only part of the grid is resident.
//
// S = evidence/l3_supplement_20260921_v2/single_source/ (paths above are in S).
#include <atomic>

#include "hop_block.h"
#include "l1_slf.h"
#include "solver_internal.h"

namespace mlmq {

struct SolveArgs {
    const uint32_t *row;    // n + 1 CSR offsets
    const uint32_t *col;    // m destinations
    const uint32_t *wgt;    // m weights
    const uint4 *blocks;    // n hop blocks (4 x uint4 each)
    const uint2 *hubs;      // hub table: (vertex, chunk size) per hub id
    uint32_t *dist;         // n distances
    L2Queue q;
    uint32_t n;
    // Small settings packed two words, to spare scalar registers in the
    // persistent loop: lims = min_read | pop_max << 8 | spill << 16 |
    // max_sleep << 24 (each <= 64); flags = fifo | split << 1 |
    // stale_check << 2 | prefetch << 4 | defer_heavy << 5 | light_split << 6 |
    // donate_min << 8 | csr_stream << 16 | packed << 17.
    uint32_t lims;
    uint32_t flags;
    __device__ __forceinline__ uint32_t min_read() const { return lims & 0xFFu; }
    __device__ __forceinline__ uint32_t pop_max() const { return (lims >> 8) & 0xFFu; }
    __device__ __forceinline__ uint32_t spill() const { return (lims >> 16) & 0xFFu; }
    __device__ __forceinline__ uint32_t max_sleep() const { return lims >> 24; }
    __device__ __forceinline__ uint32_t fifo() const { return flags & 1u; }
    __device__ __forceinline__ bool split() const { return (flags >> 1) & 1u; }
    __device__ __forceinline__ uint32_t stale_check() const { return (flags >> 2) & 3u; }
    __device__ __forceinline__ bool prefetch() const { return (flags >> 4) & 1u; }
    __device__ __forceinline__ bool defer_heavy() const { return (flags >> 5) & 1u; }
    __device__ __forceinline__ bool light_split() const { return (flags >> 6) & 1u; }
    __device__ __forceinline__ uint32_t donate_min() const { return (flags >> 8) & 0xFFu; }
    __device__ __forceinline__ bool csr_stream() const { return (flags >> 16) & 1u; }
    // vertex-id and weight mask of a packed graph's records and hop blocks
    __device__ __forceinline__ uint32_t id_mask() const {
        return (flags >> 17) & 1u ? (1u << kBackShift) - 1u : 0xFFFFFFFFu;
    }
    uint64_t timeout_ticks;
};

template <bool kProf>
struct PhaseClock {
    uint64_t t;
    __device__ __forceinline__ void start() {
        if (kProf) t = clock64();
    }
    __device__ __forceinline__ void lap(WaveStats &st, uint32_t slot) {
        if (kProf) {
            const uint64_t now = clock64();
            st.clk[slot] += now - t;
            t = now;
        }
    }
};

// A CSR column or weight of a heavy vertex.  With csr_stream the load is
// non-temporal: an edge list is read once per expansion, and keeping it out
// of the XCD's L2 leaves room for the distance lines its pre-checks and
// atomics hit.
__device__ __forceinline__ uint32_t ld_csr(const uint32_t *p, bool stream) {
    return stream ? __builtin_nontemporal_load(p) : *p;
}

__device__ __forceinline__ uint32_t word_of(const HopBlock &b, int w) {
    const uint4 &q = b.q[w >> 2];
    switch (w & 3) {
    case 0: return q.x;
    case 1: return q.y;
    case 2: return q.z;
    default: return q.w;
    }
}

// Heavy records (more out-edges than a hop block holds) read their columns
// and weights from the CSR in two tiers:
//   1. each full 64-edge piece of a record is relaxed by the whole wave, lane
//      i taking edge i (coalesced loads).  The owner of piece j is the
//      highest lane whose first piece index is <= j: one ballot and a
//      find-last-bit, no LDS;
//   2. the tails (deg mod 64 edges of each record) are spread over the lanes
//      by a DPP scan of their lengths, the owner of a slot found as a DPP
//      running maximum of start markers written to LDS.
// Each iteration handles kHeavyUnroll groups of 64 slots: all their
// column/weight loads, then all their pre-checks, then all their atomicMins,
// are in flight together.  Tier 1 also loads the next iteration's columns
// and weights while this one's pre-checks and atomics are in flight, so an
// iteration costs two round trips instead of three.
constexpr int kHeavyUnroll = 4;
static_assert(kHeavyUnroll * 64 < 1024, "relax_slots packs per-wave sums in 10-bit fields");

// Relax kHeavyUnroll slots per lane whose destinations c[] (valid vertex
// ids on every lane, active or not) and tentative distances nd[] are loaded;
// push the improved destinations.  `mid()` runs after the pre-check loads are
// issued and before they are waited for; any loads it issues stay in flight.
template <bool kSLF, int kBlk, uint32_t kCap, int S, typename Mid>
__device__ __forceinline__ bool relax_slots(const SolveArgs &a, const bool (&act)[kHeavyUnroll],
                                            const uint32_t (&c)[kHeavyUnroll],
                                            const uint32_t (&nd)[kHeavyUnroll], uint64_t *ring,
                                            HopBlock *blk, L1State &l1, WaveStats &st,
                                            Mid &&mid) {
    // Cached pre-check: a destination's distance read through this CU's and
    // XCD's caches may be stale, but distances only fall, so a cached value
    // <= nd proves the atomicMin could not lower it.  Most relaxations out of
    // high-degree vertices fail, and skipping their atomics (each a trip to
    // the memory-side atomic unit) is what bounds this loop; the check costs
    // one cache round trip.
    // (unpredicated, so the number of loads in flight is fixed and the wait
    // below leaves mid()'s loads alone)
    uint32_t pre[kHeavyUnroll];
#pragma unroll
    for (int j = 0; j < kHeavyUnroll; ++j) pre[j] = a.dist[c[j]];
    mid();
#pragma unroll
    for (int j = 0; j < kHeavyUnroll; ++j) asm volatile("" : "+v"(pre[j]));
    uint32_t pv[kHeavyUnroll];
    uint32_t nat = 0;
#pragma unroll
    for (int j = 0; j < kHeavyUnroll; ++j) {
        pv[j] = 0;
        const bool go = act[j] && nd[j] < pre[j];
        nat += go ? 1u : 0u;
        if (go) pv[j] = fetch_min(a.dist + c[j], nd[j]);
    }
#pragma unroll
    for (int j = 0; j < kHeavyUnroll; ++j) asm volatile("" : "+v"(pv[j]));
    uint64_t out[kHeavyUnroll];
    uint32_t bits = 0, nact = 0;
#pragma unroll
    for (int j = 0; j < kHeavyUnroll; ++j) {
        if (act[j] && nd[j] < pv[j]) bits |= 1u << j;
        nact += act[j] ? 1u : 0u;
        out[j] = make_rec(c[j], nd[j]);
    }
    // relaxations, skipped atomics and improvements: 10-bit fields (each at
    // most kHeavyUnroll * 64 per wave)
    const uint32_t sums =
        wave_sum(nact | (nact - nat) << 10 | (uint32_t)__builtin_popcount(bits) << 20);
    st.v[kStatRelax] += sums & 0x3FFu;
    st.v[kStatAtomic] += (sums >> 10) & 0x3FFu;   // skipped; see flush_stats
    st.v[kStatImprove] += sums >> 20;
    bool ok = true;
    if (sums >> 20) {
        if (kSLF) {
            ok = l1_push<kHeavyUnroll, kBlk ? 2 : 0, kCap, S>(ring, l1, out, bits, a.spill(), a.fifo(),
                                                     a.q, st, blk);
        } else {
#pragma unroll
            for (int j = 0; j < kHeavyUnroll; ++j)
                if (ok) ok = l2_push<S>(a.q, out[j], (bits >> j) & 1, st);
        }
    }
    return ok;
}

template <bool kSLF, int kBlk, uint32_t kCap = kL1Cap, int S = kAgentScope>
__device__ bool relax_heavy(const SolveArgs &a, bool heavy, uint32_t d, uint32_t hstart,
                            uint32_t hdeg, uint64_t *ring, HopBlock *blk, L1State &l1,
                            uint32_t *map, WaveStats &st) {
    const uint32_t lane = lane_id();
    const uint32_t deg = heavy ? hdeg : 0u;
    bool ok = true;

    // tier 1: full 64-edge pieces, one per 64 slots
    const uint32_t npc = deg / kWave;
    uint32_t ptotal = 0;
    const uint32_t pex = wave_excl_scan(npc, ptotal);   // first piece index per lane
    // Issue the column/weight loads of pieces [pb, pb + kHeavyUnroll).
    // Unpredicated: pieces past the end read edge 0 (there is one) and are
    // marked inactive.
    auto load_pieces = [&](uint32_t pb, bool (&act)[kHeavyUnroll], uint32_t (&c)[kHeavyUnroll],
                           uint32_t (&w)[kHeavyUnroll], uint32_t (&o_d)[kHeavyUnroll]) {
#pragma unroll
        for (int j = 0; j < kHeavyUnroll; ++j) {
            const uint32_t pj = pb + j;
            act[j] = pj < ptotal;
            const uint64_t m = ballot(npc > 0 && pex <= pj);
            const uint32_t ow = m ? 63u - (uint32_t)__builtin_clzll(m) : 0u;
            const uint32_t e =
                act[j] ? read_lane(hstart, ow) + (pj - read_lane(pex, ow)) * kWave + lane : 0u;
            o_d[j] = read_lane(d, ow);
            c[j] = ld_csr(a.col + e, a.csr_stream());
            w[j] = ld_csr(a.wgt + e, a.csr_stream());
        }
    };
    if (ptotal) {
        bool act[kHeavyUnroll];
        uint32_t c[kHeavyUnroll], w[kHeavyUnroll], o_d[kHeavyUnroll];
        load_pieces(0, act, c, w, o_d);
        for (uint32_t pb = 0;;) {
#pragma unroll
            for (int j = 0; j < kHeavyUnroll; ++j) asm volatile("" : "+v"(c[j]), "+v"(w[j]));
            uint32_t nd[kHeavyUnroll];
#pragma unroll
            for (int j = 0; j < kHeavyUnroll; ++j) nd[j] = w[j] + o_d[j];
            const uint32_t nb = pb + kHeavyUnroll;
            bool act2[kHeavyUnroll];
            uint32_t c2[kHeavyUnroll], w2[kHeavyUnroll], o_d2[kHeavyUnroll];
            ok = relax_slots<kSLF, kBlk, kCap, S>(a, act, c, nd, ring, blk, l1, st,
                                         [&]() { load_pieces(nb, act2, c2, w2, o_d2); });
            if (!ok || nb >= ptotal) break;
#pragma unroll
            for (int j = 0; j < kHeavyUnroll; ++j) {
                act[j] = act2[j];
                c[j] = c2[j];
                w[j] = w2[j];
                o_d[j] = o_d2[j];
            }
            pb = nb;
        }
    }

    // tier 2: the tails, spread over the lanes
    const uint32_t tdeg = deg % kWave;
    uint32_t total = 0;
    const uint32_t excl = wave_excl_scan(tdeg, total);
    // per lane: offset of its tail's first edge relative to its first slot
    const uint32_t off = hstart + (deg - tdeg) - excl;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < total && ok; base += kHeavyUnroll * kWave) {
        uint32_t o[kHeavyUnroll];
        bool act[kHeavyUnroll];
#pragma unroll
        for (int j = 0; j < kHeavyUnroll; ++j) {
            const uint32_t cb = base + j * kWave;
            act[j] = cb + lane < total;
            map[lane] = 0;
            wave_sync();
            if (tdeg > 0 && excl >= cb && excl - cb < (uint32_t)kWave) map[excl - cb] = lane + 1;
            wave_sync();
            const uint32_t o1 = max(carry, wave_incl_max(map[lane]));
            carry = read_lane(o1, kWave - 1);
            o[j] = act[j] ? o1 - 1 : 0u;
        }
        uint32_t c[kHeavyUnroll], nd[kHeavyUnroll], o_d[kHeavyUnroll];
#pragma unroll
        for (int j = 0; j < kHeavyUnroll; ++j) {
            const uint32_t o_off = (uint32_t)__shfl((int)off, (int)o[j], kWave);
            o_d[j] = (uint32_t)__shfl((int)d, (int)o[j], kWave);
            c[j] = 0;
            nd[j] = 0;
            if (act[j]) {
                const uint32_t e = o_off + base + j * kWave + lane;
                c[j] = ld_csr(a.col + e, a.csr_stream());
                nd[j] = ld_csr(a.wgt + e, a.csr_stream());
            }
        }
#pragma unroll
        for (int j = 0; j < kHeavyUnroll; ++j) asm volatile("" : "+v"(c[j]), "+v"(nd[j]));
#pragma unroll
        for (int j = 0; j < kHeavyUnroll; ++j) nd[j] += o_d[j];
        ok = relax_slots<kSLF, kBlk, kCap, S>(a, act, c, nd, ring, blk, l1, st, []() {});
    }
    return ok;
}

// Push the chunk records of the hubs this batch split: lane i owns
// `nextra` chunks (indices 1..nextra) of hub `hub` at distance `d`.  The
// chunks are spread over the lanes 64 at a time (owner of a slot = running
// maximum of start markers, as in relax_heavy) and pushed to L2.
template <int S = kAgentScope>
__device__ bool push_chunks(const SolveArgs &a, uint32_t hub, uint32_t d, uint32_t nextra,
                            uint32_t *map, WaveStats &st) {
    const uint32_t lane = lane_id();
    uint32_t total = 0;
    const uint32_t excl = wave_excl_scan(nextra, total);
    st.v[kStatChunks] += total;
    uint32_t carry = 0;
    for (uint32_t cb = 0; cb < total; cb += kWave) {
        map[lane] = 0;
        wave_sync();
        if (nextra > 0 && excl >= cb && excl - cb < (uint32_t)kWave) map[excl - cb] = lane + 1;
        wave_sync();
        const uint32_t o1 = max(carry, wave_incl_max(map[lane]));
        carry = read_lane(o1, kWave - 1);
        const bool act = cb + lane < total;
        const int o = act ? (int)o1 - 1 : 0;
        const uint32_t oh = (uint32_t)__shfl((int)hub, o, kWave);
        const uint32_t od = (uint32_t)__shfl((int)d, o, kWave);
        const uint32_t oe = (uint32_t)__shfl((int)excl, o, kWave);
        const uint64_t rec = make_rec(chunk_vertex_word(oh, 1 + cb + lane - oe), od);
        if (!l2_push<S>(a.q, rec, act, st)) return false;
    }
    return true;
}

// Relax all out-edges of the records in lanes [0, nin) (simple_process,
// SS/SSSP/sssp_run.cu:143-413), one record per lane, in two round trips:
//   1. the record's current distance (stale filter, reference :174: process
//      on <=) and its 64-byte hop block, loaded together;
//   2. one agent-scope atomicMin per inline out-edge, all issued back to
//      back (up to 7 per lane) before any result is used.
// A (dst, nd) record is pushed only if its atomicMin strictly lowered dst's
// distance (:211-243).  The batch's pushes go to L1 in one pass.
// kPre puts a cached pre-check of every destination's distance before its
// atomicMin (a third round trip, see relax_slots): the many-query kernels,
// which keep the whole device busy, are bound by the memory-side atomic
// traffic and gain from it; the single-query kernel, bound by the latency
// of its critical path, does not.
// kPair: lane pairs.  Record i sits in lanes 2i and 2i + 1 (at most 32
// records a batch); the even lane loads the first half of the hop block
// (degree, CSR start and inline edges 0-2), the odd lane the second half
// (edges 3-6), and each relaxes its own edges, so a batch runs 4 edge slots
// per lane instead of 7 and fills the wave when a pop takes 32 records or
// fewer.  The even lane speaks for the record: it issues the stale checks,
// owns the CSR edges of heavier vertices and the heavy-phase record; its
// results reach the odd lane by a DPP quad permutation, issued with every
// lane active (a DPP whose source lane is masked off reads 0).
template <bool kSLF, bool kProf, bool kHot, bool kExt, bool kPre, uint32_t kCap = kL1Cap,
          bool kPair = false, int S = kAgentScope>
__device__ bool process_batch(const SolveArgs &a, uint64_t rec, uint32_t nin, bool hot,
                              const uint32_t (&hw)[16], uint64_t *ring, HopBlock *blk, L1State &l1,
                              uint32_t *map, uint32_t (&pf)[kInlineEdges], uint32_t &sunk,
                              uint32_t &hungry, WaveStats &st, PhaseClock<kProf> &clk) {
    static_assert(!kPair || (kSLF && !kHot), "lane pairs need the record-only L1");
    constexpr int kSlots = kPair ? 4 : (int)kInlineEdges;
    const uint32_t lane = lane_id();
    const uint32_t half = kPair ? (lane & 1u) : 0u;
    const bool valid = (kPair ? (lane >> 1) : lane) < nin;
    const bool lead = valid && half == 0;
    uint32_t v = rec_vertex(rec);
    // Light/heavy split: a heavy-phase record relaxes only its vertex's edges
    // of weight >= delta (the reference has no such split; see SolveArgs).
    const bool hphase = (kExt && a.light_split()) && (rec_dist(rec) & kHeavyPhase) != 0;
    const uint32_t d = rec_dist(rec) & ~kHeavyPhase;
    // A chunk record names (hub id, chunk index); one extra round trip to the
    // hub table, taken only by batches that hold chunks.
    const bool chunk = valid && (v & kChunkFlag) != 0;
    // Back edge (hop_block.h): on a packed graph the bits above the vertex id
    // name the slot + 1 of the edge back to the vertex whose expansion
    // created this record; that edge is not relaxed.
    const uint32_t idm = a.id_mask();
    const uint32_t vword = v;   // with the back-edge bits, for the heavy-phase record
    const uint32_t back = chunk ? 0u : (v & ~idm) >> kBackShift;
    uint32_t ci = 0;
    if (ballot(chunk)) {
        if (chunk) {
            ci = v & ((1u << kChunkBits) - 1);
            v = a.hubs[(v & ~kChunkFlag) >> kChunkBits].x;
        }
    }
    if (!chunk) v &= idm;
    // Stale filter (reference :174, process on <=).  Distances only fall, so
    // a value read through this XCD's L2 (a.stale_check() == 1) is never below
    // the coherent one: the filter then lets some stale records through but
    // never drops a live one.  0 = no filter (every record is expanded).
    // 3 = late: the distance is read in the atomics' round trip instead (see
    // below), so this round trip waits only for the block.  Hot lanes
    // (block-carrying L1) already hold their block: no round trip; their
    // stale check is always late.
    // The coherent reads (modes 2 and 3) are atomicMin(dist[v], d), which
    // returns the current distance and never changes it: a record (v, d) is
    // created when an atomicMin lowers dist[v] to d, and distances only fall,
    // so dist[v] <= d from then on.  An agent-scope load refetches the whole
    // 128-byte line from memory on every call; the atomic is a 32-byte
    // request executed at the memory side (road6m: 0.87 GB fewer bytes read
    // per query, 6.96 -> 6.50 ms).
    uint32_t cur = kUnreached;
    HopBlock b = empty_block();
    if (valid && !hot) {
        // The block first: the compiler waits for every outstanding load at
        // the join of the stale-check modes' branches, and with the block
        // already issued that wait covers both loads in one round trip.
        if (kPair) {
            // this lane's half of the block: words 0-7 or 8-15
            const uint4 *p = a.blocks + 4ull * v + 2u * half;
            b.q[0] = p[0];
            b.q[1] = p[1];
            if (half == 0) {
                if (a.stale_check() == 2) cur = fetch_min(a.dist + v, d);
                else if (a.stale_check() == 1) cur = a.dist[v];
            }
        } else {
            b = load_block(a.blocks, v);
            if (a.stale_check() == 2) cur = fetch_min(a.dist + v, d);
            else if (a.stale_check() == 1) cur = a.dist[v];
        }
    }
    if (kPair) cur = dpp<0xA0>(cur);   // quad_perm [0,0,2,2]: the even lane's value
    if (kHot) {
        // word-wise select of register values (keeps both sources out of memory)
        uint32_t *bw = reinterpret_cast<uint32_t *>(&b);
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            uint32_t h = hw[i];
            asm volatile("" : "+v"(h));
            bw[i] = hot ? h : bw[i];
        }
    }
    const uint32_t bdeg = kPair ? dpp<0xA0>(b.q[0].x) : b.deg();
    // idle-wave count for work donation, in the same round trip
    hungry = a.donate_min() ? uniform(ld_relaxed<S>(a.q.ctl + kHungryOff)) : 0u;
    // bucket epoch, for deferring heavy records (same round trip)
    const uint32_t epoch = (kExt && a.defer_heavy()) ? uniform(ld_relaxed<S>(a.q.ctl + kBaseOff)) : 0u;
    bool keep = valid && (hot || d <= cur);
    st.v[kStatStale] += popc(ballot((kPair ? lead : valid) && !keep));
    // Deferral: a heavy record whose distance lies beyond the head bucket goes
    // back to L2 (into its own bucket) instead of being expanded now, so hub
    // edge lists are expanded in bucket order and rarely twice.
    if (kExt && a.defer_heavy()) {
        const bool defer = keep && !chunk && half == 0 && bdeg > kInlineEdges &&
                           d / a.q.delta > epoch;
        if (ballot(defer)) {
            st.v[kStatSpare1] += popc(ballot(defer));
            if (!l2_push<S>(a.q, make_rec(v, d), defer, st)) return false;
            keep = keep && !defer;
        }
    }
    uint32_t deg = keep ? bdeg : 0u;
    uint32_t hstart = b.start();
    uint32_t nextra = 0;   // chunks of a hub popped here that go to L2
    const uint32_t hub_word = b.q[0].z, csize = b.q[0].w;
    // chunk lanes always take the CSR path (their edges are not inline)
    const bool csr_path = chunk || bdeg > kInlineEdges;
    bool need_heavy = false;   // a light-phase record whose vertex has heavy edges
    // The light/heavy split first: the record's phase selects its part of the
    // adjacency (sorted by weight: the first b.q[1].x edges are light).  A
    // hub's part is then chunked like a whole list; with the split only the
    // heavy part is (chunk records carry kHeavyPhase).
    if ((kExt && a.light_split()) && csr_path) {
        const uint32_t nl = min(b.q[1].x, deg);
        if (hphase) {
            hstart += nl;
            deg -= nl;
        } else {
            need_heavy = deg > nl && half == 0;
            deg = nl;
        }
    }
    if (chunk) {
        const uint32_t lo = ci * csize;
        deg = lo < deg ? min(csize, deg - lo) : 0u;
        hstart += lo;
    } else if (a.split() && deg > kInlineEdges && hub_word != 0 &&
               (!(kExt && a.light_split()) || hphase)) {
        nextra = (deg - 1) / csize;
        deg = min(deg, csize);
    }
    if (half) nextra = 0;   // (the odd lane's header words are edges)
    bool heavy = csr_path && deg > 0 && half == 0;
    const uint32_t ideg = csr_path ? 0u : deg;
    clk.lap(st, kClkRow);
    if (ballot(nextra > 0)) {
        if (!push_chunks<S>(a, hub_word - 1, d | (hphase ? kHeavyPhase : 0u), nextra, map, st))
            return false;
    }

    // Slot k holds edge k; with lane pairs slot k holds the lane's words 2k
    // and 2k + 1: edge k - 1 on the even lane (slot 0 is the header), edge
    // k + 3 on the odd lane.
    // cw[k]: the destination word as stored (on a packed graph with the
    // child's back-edge bits, which its record keeps); c[k]: the vertex.
    uint32_t cw[kSlots], c[kSlots], nd[kSlots], pv[kSlots];
    uint32_t on = 0;   // bit k: slot k's edge is relaxed by this record
#pragma unroll
    for (int k = 0; k < kSlots; ++k) {
        cw[k] = word_of(b, kPair ? 2 * k : 2 + 2 * k);
        c[k] = cw[k] & idm;
        const uint32_t wk = word_of(b, kPair ? 2 * k + 1 : 3 + 2 * k);
        nd[k] = d + wk;
        const uint32_t eidx = kPair ? (uint32_t)k + (half ? 3u : 0xFFFFFFFFu) : (uint32_t)k;
        bool e = eidx < ideg;
        if ((kExt && a.light_split())) {
            const bool hv = wk >= a.q.delta;
            need_heavy = need_heavy || (e && hv && !hphase);
            e = e && hv == hphase;
        }
        if (e) on |= 1u << k;
    }
    // The back edge's slot bit (edge back - 1: slot back - 1, or with lane
    // pairs slot back on the even lane for edges 0-2 and slot back - 4 on
    // the odd lane for edges 3-6; back = 0 gives no bit).
    const uint32_t seen = on;   // examined: relaxed or skipped as the back edge
    on &= ~(kPair ? ((1u << back) >> (half ? 4u : 0u)) & (half ? 0xFu : 0xEu)
                  : (1u << back) >> 1);
    if (kPair && (kExt && a.light_split())) {
        // the odd lane's heavy inline edges also need the heavy-phase record
        const uint32_t nh = need_heavy ? 1u : 0u;
        const uint32_t other = dpp<0xF5>(nh);   // quad_perm [1,1,3,3]
        need_heavy = lead & ((nh | other) != 0u);
    }
    // Block-carrying L1: load every relaxed child's whole block in this
    // round trip; the improved children take it into L1 with them, and the
    // batch that pops them needs no global load.
    HopBlock cb[kHot ? kInlineEdges : 1];
    if (kHot) {
#pragma unroll
        for (int k = 0; k < (kHot ? (int)kInlineEdges : 0); ++k)
            cb[k] = load_block(a.blocks, ((on >> k) & 1) ? c[k] : 0u);
    }
    uint32_t go = on;   // bit k: edge k issues its atomicMin
    if (kPre) {
        uint32_t pre[kSlots];
#pragma unroll
        for (int k = 0; k < kSlots; ++k) pre[k] = ((on >> k) & 1) ? a.dist[c[k]] : 0u;
#pragma unroll
        for (int k = 0; k < kSlots; ++k) asm volatile("" : "+v"(pre[k]));
#pragma unroll
        for (int k = 0; k < kSlots; ++k)
            if (nd[k] >= pre[k]) go &= ~(1u << k);
    }
#pragma unroll
    for (int k = 0; k < kSlots; ++k) {
        pv[k] = 0;
        if ((go >> k) & 1) pv[k] = fetch_min(a.dist + c[k], nd[k]);
    }
    // Late stale check, in the same round trip (issued before the touches
    // below, so waiting for it leaves them in flight).  A record whose vertex
    // has since dropped below d has a fresher record queued; that one relaxes
    // every child to a strictly lower distance and pushes it, so the children
    // this stale expansion improved need not be pushed, and its CSR edges
    // need not be relaxed at all.  The read is the value-preserving
    // atomicMin described above.
    const bool late_chk = (kHot && hot && a.stale_check()) || (valid && a.stale_check() == 3);
    uint32_t cur2 = kUnreached;
    if (late_chk && half == 0) cur2 = fetch_min(a.dist + v, d);
    // Prefetch, in the same round trip as the atomics: touch every relaxed
    // child's hop block, so the batch that later pops the child finds the
    // line in this XCD's L2 (its dist line is already there from the
    // atomic).  Issued after the atomics and without a branch (lanes without
    // an edge, and all lanes when a.prefetch() is off, touch vertex 0's
    // block), so waiting for the atomics' results takes a fixed vmcnt that
    // leaves these loads in flight.  A register's previous word is folded
    // into `sunk` just before the register is reloaded, one batch after its
    // load (vmcnt counts in issue order, so this batch's block-load wait
    // already covered it).  The compiler still moves these loop-carried words
    // between registers at the loop's back edge, which waits for the loads
    // there; the touches pay for themselves regardless (without them the
    // batch's block loads miss the L2 more often).
    if (!kHot) {
#pragma unroll
        for (int k = 0; k < kSlots; ++k) {
            const uint32_t t = (a.prefetch() && ((on >> k) & 1)) ? c[k] : 0u;
            sunk ^= pf[k];
            pf[k] = reinterpret_cast<const uint32_t *>(a.blocks + 4ull * t)[0];
        }
    }
    // Keep every result opaque until all atomics are issued; otherwise the
    // compiler sinks each comparison into its predicated block and waits
    // for the atomics one at a time.
#pragma unroll
    for (int k = 0; k < kSlots; ++k) asm volatile("" : "+v"(pv[k]));
    asm volatile("" : "+v"(cur2));
    if (kPair) cur2 = dpp<0xA0>(cur2);
    uint64_t out[kSlots];
    uint32_t bits = 0;   // bit k: slot k's edge lowered its destination's distance
#pragma unroll
    for (int k = 0; k < kSlots; ++k) {
        if (((go >> k) & 1) && nd[k] < pv[k]) bits |= 1u << k;
        out[k] = make_rec(cw[k], nd[k]);
    }
    {
        const bool late_stale = late_chk && cur2 < d;
        st.v[kStatStale] += popc(ballot(late_stale && half == 0));
        if (late_stale) {
            bits = 0;
            heavy = false;
            need_heavy = false;
        }
    }
    // Edges examined (relaxed, or skipped as back edges), skipped back edges
    // and improvements of the batch in one wave reduction: 10-bit fields, at
    // most 64 x 7 per batch.  (Skipped edges count with the atomics the heavy
    // loop's pre-check skipped, see flush_stats.)
    const uint32_t sums = wave_sum((uint32_t)__builtin_popcount(seen) |
                                   ((uint32_t)__builtin_popcount(seen & ~on) << 10) |
                                   ((uint32_t)__builtin_popcount(bits) << 20));
    st.v[kStatRelax] += sums & 0x3FFu;
    st.v[kStatAtomic] += (sums >> 10) & 0x3FFu;
    st.v[kStatImprove] += sums >> 20;
    clk.lap(st, kClkAtomic);

    bool ok = true;
    if (sums >> 20) {
        if (kSLF) {
            ok = l1_push<kSlots, kHot ? 1 : 0, kCap, S>(ring, l1, out, bits, a.spill(), a.fifo(), a.q,
                                               st, blk, cb);
        } else {
#pragma unroll
            for (int k = 0; k < kSlots; ++k)
                if (ok) ok = l2_push<S>(a.q, out[k], (bits >> k) & 1, st);
        }
        clk.lap(st, kClkFlush);
    }
    if (ok && ballot(heavy)) {
        ok = relax_heavy<kSLF, kHot ? 1 : 0, kCap, S>(a, heavy, d, hstart, deg, ring, blk, l1, map, st);
        clk.lap(st, kClkEdges);
    }
    // The heavy edges wait for a heavy-phase record one bucket later: by then
    // the light phase of this bucket has settled the vertex's distance, so
    // the heavy edges are relaxed once instead of once per improvement.
    if (ok && ballot(need_heavy)) ok = l2_push<S>(a.q, make_rec(vword, d | kHeavyPhase), need_heavy, st);
    return ok;
}

// The worker loop of one wave for one query (reference work_block_kernel
// loop, SS/SSSP/sssp_run.cu:507-584).  Returns when the query's exit flag is
// seen, after an error, or at the deadline.  `leader` waves also run the
// termination check while idle.  With lane pairs (kPair) a batch holds at
// most 32 records, two lanes each: pops and L2 top-ups take at most 32.
template <bool kSLF, bool kProf, bool kHot = false, bool kExt = false, bool kPre = false,
          uint32_t kCap = kL1Cap, bool kPair = false, int S = kAgentScope>
__device__ void run_worker(const SolveArgs &a, bool leader, uint64_t *ring, uint32_t *map,
                           uint64_t *stage, WaveStats &st, PhaseClock<kProf> &clk,
                           uint64_t deadline, HopBlock *blk = nullptr) {
    const uint32_t lane = lane_id();
    const L2Queue &q = a.q;
    constexpr uint32_t kRecs = kPair ? kWave / 2 : kWave;   // records a batch can hold
    L1State l1{0u, 0u};
    // Records taken from L2 whose work is not finished.  Credits are returned
    // only when the wave runs out of work (the idle path below): a wave with
    // work cannot let the solve end anyway, and returning them whenever L1
    // drains mid-run put a non-returning atomic in flight that the next
    // batch's loads waited behind (vmcnt counts in issue order).  They are
    // returned with a release: termination needs every tail reservation a
    // record's expansion made (and the children written into the slots) to
    // be visible before its credit, which the quiescence check reads with an
    // acquire.  The release writes back this XCD's L2 and waits for the
    // wave's loads in flight, but the wave is idle here.
    uint32_t credit = 0;
    uint32_t cl_start = 0;      // lane b: start of the claim on bucket b
    uint64_t cl_mask = 0;       // lane b: pending slots of that claim
    uint32_t sleep = 0;
    uint32_t pf[kInlineEdges];  // prefetch results, consumed one batch late
    uint32_t sunk = 0;
    uint32_t base_hint = 0;     // L2 epoch seen by this wave's last pop
#pragma unroll
    for (int k = 0; k < (int)kInlineEdges; ++k) pf[k] = 0;
    bool idle_listed = false;   // counted in the control block's hungry-wave count

    for (;;) {
        uint64_t rec = kEmpty;
        uint32_t nin = 0;
        bool hot = false;
        uint32_t hw[16];
        if (kHot) {
            nin = l1_pop_hot(ring, blk, l1, a.pop_max(), rec, hot, hw);
            clk.lap(st, kClkL1Pop);
        } else if (kPair) {
            nin = l1_pop_pair<kCap>(ring, l1, min(a.pop_max(), kRecs), rec);
            clk.lap(st, kClkL1Pop);
        } else if (kSLF) {
            nin = l1_pop<kCap>(ring, l1, a.pop_max(), rec);
            clk.lap(st, kClkL1Pop);
        }
        // Top up from L2 below min_read.  A block-carrying L1 may hold back
        // cold records behind a hot run; it tops up only once it is empty, so
        // a hot batch stays free of global loads.
        if (nin < min(a.min_read(), kRecs) && (!kHot || l1.size == 0)) {
            const uint32_t got =
                l2_pop<kExt, S>(q, cl_start, cl_mask, stage, kRecs - nin, kRecs, st, base_hint);
            if (got) {
                credit += got;
                const uint32_t r = kPair ? lane >> 1 : lane;
                if (r >= nin && r < nin + got) rec = stage[r - nin];
                nin += got;
            }
            clk.lap(st, kClkL2Pop);
        }
        if (nin == 0) {
            if (credit) {
                if (lane == 0)
                    __hip_atomic_fetch_add(q.ctl + kDoneOff, credit, __ATOMIC_RELEASE, S);
                credit = 0;
            }
            if (a.donate_min() && !idle_listed) {
                if (lane == 0) fetch_add(q.ctl + kHungryOff, 1u);
                idle_listed = true;
            }
            st.v[kStatIdle] += 1;
            if (leader && l2_check_done<S>(q)) break;
            if (uniform(ld_relaxed<S>(q.ctl + kExitOff))) break;
            if (wall_clock64() > deadline) {
                raise_error<S>(q, kErrTimeout);
                break;
            }
            sleep = sleep ? (sleep * 2 < a.max_sleep() ? sleep * 2 : a.max_sleep()) : 1;
            for (uint32_t k = 0; k < sleep; ++k) __builtin_amdgcn_s_sleep(4);
            clk.lap(st, kClkIdle);
            continue;
        }
        sleep = 0;
        if (idle_listed) {
            if (lane == 0) fetch_add(q.ctl + kHungryOff, 0xFFFFFFFFu);
            idle_listed = false;
        }
        st.v[kStatBatches] += 1;
        uint32_t hungry = 0;
        if (!process_batch<kSLF, kProf, kHot, kExt, kPre, kCap, kPair, S>(a, rec, nin, hot, hw, ring, blk, l1, map, pf, sunk, hungry,
                                              st, clk))
            break;
        if (kSLF && hungry && l1.size > a.donate_min()) {
            // (uniform: hungry is a scalar load result)
            if (!l1_donate<kCap, S>(ring, l1, uniform(l1.size) / 2, q, st)) break;
            clk.lap(st, kClkFlush);
        }
        if ((st.v[kStatBatches] & 31) == 0) {
            if (uniform(ld_relaxed<S>(q.ctl + kExitOff))) break;
            if (wall_clock64() > deadline) {
                raise_error<S>(q, kErrTimeout);
                break;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < (int)kInlineEdges; ++k) sunk ^= pf[k];
    st.v[kStatSpare2] += popc(ballot(sunk == 0x5EEDF00Du));
}

__device__ __forceinline__ void flush_stats(uint32_t *ctl, WaveStats &st) {
    // kStatAtomic counts the atomics the heavy loop's pre-check skipped
    st.v[kStatAtomic] = st.v[kStatRelax] - st.v[kStatAtomic];
    if (lane_id() == 0) {
        uint64_t *stats = reinterpret_cast<uint64_t *>(ctl + kStatsOff);
#pragma unroll
        for (uint32_t i = 0; i < kNumStats; ++i) {
            const uint64_t x = (i >= kClkL1Pop && i <= kClkLaps) ? st.clk[i] : (uint64_t)st.v[i];
            if (x) add64(stats + i, x);
        }
    }
}

// kHot: block-carrying L1 (each wave's ring paired with 512 hop-block slots in
// LDS, 32 KiB per wave; about 150 KiB per 4-wave workgroup, so one
// workgroup per CU).
template <int kWaves, bool kSLF, bool kProf, bool kHot, bool kExt = false, bool kPair = false>
__global__ __launch_bounds__(kWaves * 64) void sssp_persistent_kernel(SolveArgs a) {
    __shared__ uint64_t s_ring[kWaves][kL1Alloc];
    __shared__ uint32_t s_map[kWaves][kWave];
    __shared__ uint64_t s_stage[kWaves][kWave];
    __shared__ uint4 s_blk[kHot ? kWaves : 1][kHot ? 4 * kL1Cap : 1];

    const uint32_t w = threadIdx.x / kWave;
    const uint32_t gwave = blockIdx.x * kWaves + w;

    WaveStats st;
    st.clear();
    PhaseClock<kProf> clk;
    const uint64_t t_begin = kProf ? clock64() : 0;
    clk.start();
    if (kProf) {
        for (int i = 0; i < 16; ++i) clk.lap(st, kClkLaps);
    }
    const uint64_t deadline = wall_clock64() + a.timeout_ticks;
    run_worker<kSLF, kProf, kHot, kExt, false, kL1Cap, kPair>(a, gwave == 0, s_ring[w], s_map[w], s_stage[w], st, clk, deadline,
                                  kHot ? reinterpret_cast<HopBlock *>(s_blk[w]) : nullptr);
    if (kProf) st.clk[kClkTotal] = clock64() - t_begin;
    flush_stats(a.q.ctl, st);
}

// ---- many queries in one persistent launch ----
//
// Every workgroup owns a private L2 partition (control block + bucket rings)
// and works through queries taken from a global counter, one at a time, with
// all of its waves.  Between queries the workgroup resets its control block
// and seeds the next source behind a workgroup barrier; its rings are empty
// after every completed query (each reserved slot was taken), and the
// distance rows were filled for all queries before the launch.  The query a
// record belongs to is the one its partition is serving, so records stay
// 8 bytes.
struct BatchArgs {
    SolveArgs base;            // graph, queue parameters, timeout
    uint32_t *ctl_all;         // grid x kCtlWords
    uint64_t *slots_all;       // grid x (nbuckets << cap_log2)
    uint32_t *dist_all;        // nq x n
    const uint32_t *sources;   // nq
    uint32_t nq;
    uint32_t *gctl;            // launch-wide: [0] next query, kErrOff error, kStatsOff stats
    uint32_t *qstatus;         // per query: device error code (0 = ok)
    uint64_t *qtime;           // per query: wall-clock ticks at start and end
};

// kStats = false leaves the statistics out: the counters then need no
// registers (the per-wave counters otherwise spill into VGPR lanes and bound
// the waves per SIMD at 4; without them it is 5).
template <int kWaves, bool kProf, bool kExt = false, uint32_t kCap = kL1Cap, bool kStats = true,
          bool kPair = false>
__global__ __launch_bounds__(kWaves * 64) void sssp_batch_kernel(BatchArgs b) {
    __shared__ uint64_t s_ring[kWaves][kCap + 64];
    __shared__ uint32_t s_map[kWaves][kWave];
    __shared__ uint64_t s_stage[kWaves][kWave];
    __shared__ uint32_t s_query;

    const uint32_t w = threadIdx.x / kWave;
    SolveArgs a = b.base;
    a.q.ctl = b.ctl_all + (uint64_t)blockIdx.x * kCtlWords;
    a.q.slots = b.slots_all + ((uint64_t)blockIdx.x * a.q.nbuckets << a.q.cap_log2);

    WaveStats st;
    st.clear();
    PhaseClock<kProf> clk;
    clk.start();
    const uint64_t deadline = wall_clock64() + a.timeout_ticks;

    for (;;) {
        if (threadIdx.x == 0) {
            uint32_t qi = fetch_add(b.gctl, 1u);
            if (ld_relaxed(b.gctl + kErrOff) != 0) qi = b.nq;
            s_query = qi;
        }
        __syncthreads();
        const uint32_t qi = s_query;
        if (qi >= b.nq) break;
        for (uint32_t i = threadIdx.x; i < kStatsOff; i += blockDim.x)
            a.q.ctl[i] = i == kTailOff ? 1u : 0u;
        if (threadIdx.x == 0) {
            a.q.slots[0] = make_rec(b.sources[qi], 0u);
            b.qtime[2 * qi] = wall_clock64();
        }
        __threadfence();
        __syncthreads();
        a.dist = b.dist_all + (uint64_t)qi * a.n;
        run_worker<true, kProf, false, kExt, true, kCap, kPair, kGroupScope>(a, w == 0, s_ring[w], s_map[w], s_stage[w], st,
                                                   clk, deadline);
        __syncthreads();
        if (threadIdx.x == 0) {
            b.qtime[2 * qi + 1] = wall_clock64();
            uint32_t e = ld_relaxed(a.q.ctl + kErrOff);
            if (e == kErrNone) {
                // the partition's accounting must balance, as after a single solve
                uint32_t t = 0, k = 0;
                for (uint32_t i = 0; i < a.q.nbuckets; ++i) {
                    t += ld_relaxed(a.q.ctl + kTailOff + i * kStride);
                    k += ld_relaxed(a.q.ctl + kTakenOff + i * kStride);
                }
                if (t != k || t != ld_relaxed(a.q.ctl + kDoneOff)) e = kErrInternal;
            }
            b.qstatus[qi] = e;
            if (e != kErrNone) cas32(b.gctl + kErrOff, kErrNone, e);
        }
        __syncthreads();
    }
    if (kStats) flush_stats(b.gctl, st);
}

// ---- serving: requests of any resident graphs, several workgroups each ----
//
// A serving launch carries a table of requests.  Each request names its own
// graph and settings (a full SolveArgs: CSR, hop blocks, queue parameters)
// and owns a slot handed out by the host's admission control: a distance row
// and a queue partition (control block + buckets).  Requests of different
// graphs and sizes share one launch.  serve_reset_kernel prepares the slots
// (row fill, control reset, empty-slot fill of fresh partitions, seed
// record); sssp_serve_kernel gives every request its `wg_count` workgroups,
// which share its partition; serve_verify_kernel compares finished rows with
// expected rows when the caller registered them.
struct ServeRequest {
    SolveArgs a;              // graph, partition (dist, q.ctl, q.slots) and settings
    uint32_t src;
    uint32_t status;          // out: device error code
    uint32_t fresh;           // partition content unknown: fill with empty slots
    uint32_t mismatches;      // out (verification)
    uint32_t wg_first;        // first workgroup of this request in the launch
    uint32_t wg_count;        // workgroups sharing its partition
    const uint32_t *expect;   // optional expected row
    uint64_t t_start;         // out: wall-clock ticks of the first workgroup's start
    uint64_t t_end;           // out: latest workgroup end
};

// grid (x: row chunks, y: requests)
__global__ __launch_bounds__(256) void serve_reset_kernel(ServeRequest *reqs) {
    const ServeRequest &r = reqs[blockIdx.y];
    const uint32_t n = r.a.n;
    const uint64_t part_slots = (uint64_t)r.a.q.nbuckets << r.a.q.cap_log2;
    const uint32_t src = r.src;
    uint32_t *dist = r.a.dist;
    uint32_t *ctl = r.a.q.ctl;
    uint64_t *slots = r.a.q.slots;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t t0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (uint64_t i = t0; i < n; i += stride) dist[i] = i == src ? 0u : kUnreached;
    if (r.fresh)
        for (uint64_t i = t0 + 1; i < part_slots; i += stride) slots[i] = kEmpty;
    if (blockIdx.x == 0) {
        for (uint32_t i = threadIdx.x; i < kStatsOff; i += blockDim.x) ctl[i] = i == kTailOff ? 1u : 0u;
        if (threadIdx.x == 0) slots[0] = make_rec(src, 0u);
    }
}

template <int kWaves, bool kExt = false>
__global__ __launch_bounds__(kWaves * 64) void sssp_serve_kernel(ServeRequest *reqs, uint32_t nreq) {
    __shared__ uint64_t s_ring[kWaves][kL1Alloc];
    __shared__ uint32_t s_map[kWaves][kWave];
    __shared__ uint64_t s_stage[kWaves][kWave];

    const uint32_t w = threadIdx.x / kWave;
    // this workgroup's request: the last one whose first workgroup is <= ours
    uint32_t lo = 0, hi = nreq;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (reqs[mid].wg_first <= blockIdx.x) lo = mid;
        else hi = mid;
    }
    ServeRequest *r = reqs + lo;
    const uint32_t part = blockIdx.x - r->wg_first;
    const SolveArgs a = r->a;
    if (part == 0 && threadIdx.x == 0) r->t_start = wall_clock64();

    WaveStats st;
    st.clear();
    PhaseClock<false> clk;
    const uint64_t deadline = wall_clock64() + a.timeout_ticks;
    // lane pairs per request, from its own pop size (uniform in the workgroup)
    if (lane_pairs_for(a.pop_max()))
        run_worker<true, false, false, kExt, true, kL1Cap, true>(a, part == 0 && w == 0, s_ring[w], s_map[w],
                                                                 s_stage[w], st, clk, deadline);
    else
        run_worker<true, false, false, kExt, true, kL1Cap, false>(a, part == 0 && w == 0, s_ring[w], s_map[w],
                                                                  s_stage[w], st, clk, deadline);
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        if (part == 0) {
            // Every record pushed was claimed and credited: exit was raised by
            // the quiescence check, so the control block no longer changes.
            uint32_t e = ld_relaxed(a.q.ctl + kErrOff);
            if (e == kErrNone) {
                uint32_t t = 0, k = 0;
                for (uint32_t i = 0; i < a.q.nbuckets; ++i) {
                    t += ld_relaxed(a.q.ctl + kTailOff + i * kStride);
                    k += ld_relaxed(a.q.ctl + kTakenOff + i * kStride);
                }
                if (t != k || t != ld_relaxed(a.q.ctl + kDoneOff)) e = kErrInternal;
            }
            r->status = e;
        }
        __hip_atomic_fetch_max(&r->t_end, (uint64_t)wall_clock64(), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    }
}

// grid (x: row chunks, y: requests)
__global__ __launch_bounds__(256) void serve_verify_kernel(ServeRequest *reqs) {
    ServeRequest *r = reqs + blockIdx.y;
    const uint32_t *expect = r->expect;
    if (!expect) return;
    const uint32_t *dist = r->a.dist;
    const uint32_t n = r->a.n;
    uint32_t bad = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        bad += dist[i] != expect[i] ? 1u : 0u;
    bad = wave_sum(bad);
    if (lane_id() == 0 && bad) atomicAdd(&r->mismatches, bad);
}

// ---- live serving: one persistent launch fed through a request ring ----
//
// The host keeps a table of request entries and a ring of tickets in pinned,
// coherent host memory.  To submit a request it writes the entry (graph,
// slot, settings, source, number of parts k), names the entry in the ring
// for tickets [first, first + k), then advances `published` by k.  Block 0
// of the launch is a dispatcher: it mirrors `published` and the host's stop
// flag into device memory, so that the waiting workgroups poll device memory
// instead of the bus.  Every other workgroup takes the next ticket from a
// device counter, waits until it is published, and serves part
// (ticket - first) of that ticket's request:
//   1. fill its share of the request's distance row (and of the queue
//      partition if fresh); part 0 also clears the control block and places
//      the seed;
//   2. wait until all k parts have filled their shares (tickets are handed
//      out in order, so every part of a published request is claimed as soon
//      as workgroups free up; no request waits on a later one);
//   3. run the worker loop on the request's partition (part 0's first wave
//      runs the quiescence check);
//   4. compare its share of the row with the expected row, if registered;
//   5. count itself finished; the last part writes status, mismatch count
//      and clock ticks into the host entry, then its `done` word.
// A request thus completes, and its result can be taken, while the launch
// keeps serving others, and a request submitted while the launch runs joins
// it without a new launch.
//
// Slot reuse inside one launch: the parts write the row with plain stores, so
// before a part counts its share done it releases at agent scope (which
// writes back its XCD's L2), and after the barrier every part acquires at
// agent scope (which drops its XCD's stale lines).  The pre-check of the
// relaxation loop reads distances through the cache and relies on them only
// ever falling; a line left from the previous request on the same slot would
// break that, and the acquire removes it.
//
// Every wait is bounded: ticket waits by the stop flag and the launch's
// lifetime, the barrier and the worker loop by the request's watchdog, the
// dispatcher by the lifetime.  A request whose parts do not all arrive before
// its watchdog fires completes with the timeout error.
struct StreamEntry {
    ServeRequest r;           // graph, slot and settings; status / mismatches / ticks written back
    uint32_t first;           // first ticket of the request
    uint32_t parts;           // workgroups serving it
    uint32_t done;            // 0 while pending; 1 when the result is written
    uint32_t seq;             // host's sequence number of the request (diagnostics)
};
constexpr uint32_t kStreamEntryWords = sizeof(StreamEntry) / 4;
static_assert(sizeof(StreamEntry) % 8 == 0, "entry words");

struct StreamHostCtl {        // host memory, written by the host
    uint32_t published;       // tickets published
    uint32_t stop;            // 1: serve what is published, then exit
    uint32_t stopped_at;      // written by the dispatcher: tickets below it were served
    uint32_t alive;           // written by the dispatcher: 1 while running
};

struct StreamSync {           // device memory, one per entry; zero between requests
    uint32_t arrive;          // parts whose share of the reset is done
    uint32_t finished;        // parts done
    uint32_t mismatches;
    uint32_t status;
    uint64_t t_start;
};

struct StreamArgs {
    StreamHostCtl *host;      // host memory
    StreamEntry *entries;     // host memory
    const uint32_t *ring;     // host memory: entry index of ticket t at t % ring_size
    uint32_t ring_size;       // power of two
    uint32_t *dev;            // device: [0] claimed, [1] published mirror, [2] stop mirror
    StreamSync *sync;         // device: per entry
    uint64_t lifetime_ticks;  // wall-clock bound of the whole launch
};

__device__ __forceinline__ uint32_t ld_sys(const uint32_t *p) {
    return __hip_atomic_load(const_cast<uint32_t *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ void st_sys(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ void st_sys_release(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

constexpr uint32_t kNoTicket = 0xFFFFFFFFu;

__device__ void stream_dispatch(const StreamArgs &s, uint64_t end) {
    if (threadIdx.x != 0) return;
    st_sys(&s.host->alive, 1u);
    uint32_t last = 0;
    for (;;) {
        // relaxed: the workgroups read the entries with loads that bypass the
        // caches, after this value; no cache needs invalidating here
        const uint32_t pub = ld_sys(&s.host->published);
        if (pub != last) {
            __hip_atomic_store(s.dev + 1, pub, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            last = pub;
        }
        if (ld_sys(&s.host->stop) || wall_clock64() > end) {
            // tickets below `last` are (or will be) served; later ones are not
            const uint32_t pub2 = ld_sys(&s.host->published);
            if (pub2 != last) {
                __hip_atomic_store(s.dev + 1, pub2, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                last = pub2;
            }
            st_sys(&s.host->stopped_at, last);
            __hip_atomic_store(s.dev + 2, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            break;
        }
        __builtin_amdgcn_s_sleep(32);
    }
    st_sys_release(&s.host->alive, 0u);
}

template <int kWaves>
__global__ __launch_bounds__(kWaves * 64) void sssp_stream_kernel(StreamArgs s) {
    __shared__ uint64_t s_ring[kWaves][kL1Alloc];
    __shared__ uint32_t s_map[kWaves][kWave];
    __shared__ uint64_t s_stage[kWaves][kWave];
    __shared__ uint64_t s_entry64[kStreamEntryWords / 2];   // 8-byte aligned copy of the entry
    __shared__ uint32_t s_ticket, s_index, s_go;
    uint32_t *s_entry = reinterpret_cast<uint32_t *>(s_entry64);

    const uint64_t end = wall_clock64() + s.lifetime_ticks;
    if (blockIdx.x == 0) {
        stream_dispatch(s, end);
        return;
    }
    const uint32_t w = threadIdx.x / kWave;
    WaveStats st;
    st.clear();
    PhaseClock<false> clk;
    for (;;) {
        // 1. a ticket, published
        if (threadIdx.x == 0) {
            uint32_t t = fetch_add(s.dev, 1u);
            uint32_t nap = 1;
            for (;;) {
                const uint32_t pub = ld_relaxed(s.dev + 1);
                if ((int32_t)(t - pub) < 0) break;
                if (ld_relaxed(s.dev + 2) || wall_clock64() > end) {
                    t = kNoTicket;
                    break;
                }
                for (uint32_t k = 0; k < nap; ++k) __builtin_amdgcn_s_sleep(8);
                nap = nap < 128 ? nap * 2 : 128;   // at most ~30 us between polls
            }
            s_ticket = t;
            if (t != kNoTicket) s_index = ld_sys(s.ring + (t & (s.ring_size - 1)));
        }
        __syncthreads();
        const uint32_t t = s_ticket;
        if (t == kNoTicket) break;
        const uint32_t ei = s_index;
        // 2. the request entry, copied from host memory (loads that bypass the caches)
        StreamEntry *E = s.entries + ei;
        for (uint32_t i = threadIdx.x; i < kStreamEntryWords; i += blockDim.x)
            s_entry[i] = ld_sys(reinterpret_cast<const uint32_t *>(E) + i);
        __syncthreads();
        const StreamEntry &se = *reinterpret_cast<const StreamEntry *>(s_entry);
        const SolveArgs a = se.r.a;
        const uint32_t k = se.parts, part = t - se.first;
        StreamSync *sy = s.sync + ei;
        const uint64_t deadline = wall_clock64() + a.timeout_ticks;
        if (part == 0 && threadIdx.x == 0)
            __hip_atomic_store(&sy->t_start, (uint64_t)wall_clock64(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // 3. this part's share of the reset
        {
            const uint64_t n = a.n, lo = n * part / k, hi = n * (part + 1) / k;
            for (uint64_t i = lo + threadIdx.x; i < hi; i += blockDim.x)
                a.dist[i] = i == se.r.src ? 0u : kUnreached;
            if (se.r.fresh) {
                const uint64_t ps = (uint64_t)a.q.nbuckets << a.q.cap_log2;
                const uint64_t slo = ps * part / k, shi = ps * (part + 1) / k;
                for (uint64_t i = (slo ? slo : 1) + threadIdx.x; i < shi; i += blockDim.x) a.q.slots[i] = kEmpty;
            }
            if (part == 0) {
                for (uint32_t i = threadIdx.x; i < kStatsOff; i += blockDim.x)
                    a.q.ctl[i] = i == kTailOff ? 1u : 0u;
                if (threadIdx.x == 0) a.q.slots[0] = make_rec(se.r.src, 0u);
            }
        }
        __threadfence();
        __syncthreads();
        if (threadIdx.x == 0) {
            __hip_atomic_fetch_add(&sy->arrive, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            uint32_t go = 1, nap = 1;
            while (ld_relaxed(&sy->arrive) < k) {
                if (wall_clock64() > deadline) {
                    go = 0;
                    break;
                }
                for (uint32_t z = 0; z < nap; ++z) __builtin_amdgcn_s_sleep(4);
                nap = nap < 32 ? nap * 2 : 32;
            }
            s_go = go;
        }
        __syncthreads();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        // 4. the worker loop
        if (s_go) {
            if (lane_pairs_for(a.pop_max()))
                run_worker<true, false, false, true, true, kL1Cap, true>(a, part == 0 && w == 0, s_ring[w], s_map[w],
                                                                         s_stage[w], st, clk, deadline);
            else
                run_worker<true, false, false, true, true, kL1Cap, false>(a, part == 0 && w == 0, s_ring[w], s_map[w],
                                                                          s_stage[w], st, clk, deadline);
        } else if (threadIdx.x == 0) {
            cas32(&sy->status, kErrNone, kErrTimeout);
            raise_error(a.q, kErrTimeout);
        }
        __threadfence();
        __syncthreads();
        // 5. verification of this part's share
        const uint32_t err = uniform(ld_relaxed(a.q.ctl + kErrOff));
        if (s_go && err == kErrNone && se.r.expect) {
            const uint64_t n = a.n, lo = n * part / k, hi = n * (part + 1) / k;
            uint32_t bad = 0;
            for (uint64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) bad += a.dist[i] != se.r.expect[i] ? 1u : 0u;
            bad = wave_sum(bad);
            if (lane_id() == 0 && bad) atomicAdd(&sy->mismatches, bad);
        }
        // 6. finish
        if (threadIdx.x == 0 && part == 0) {
            uint32_t e = err;
            if (e == kErrNone) {
                uint32_t tt = 0, kk = 0;
                for (uint32_t i = 0; i < a.q.nbuckets; ++i) {
                    tt += ld_relaxed(a.q.ctl + kTailOff + i * kStride);
                    kk += ld_relaxed(a.q.ctl + kTakenOff + i * kStride);
                }
                if (tt != kk || tt != ld_relaxed(a.q.ctl + kDoneOff)) e = kErrInternal;
            }
            cas32(&sy->status, kErrNone, e);
        }
        __threadfence();
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint32_t before =
                __hip_atomic_fetch_add(&sy->finished, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            if (before == k - 1) {
                // the last part: result into the host entry, then `done`
                const uint32_t status = ld_relaxed(&sy->status);
                const uint32_t mism = ld_relaxed(&sy->mismatches);
                const uint64_t t0 = __hip_atomic_load(&sy->t_start, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const uint64_t t1 = wall_clock64();
                st_sys(&E->r.status, status);
                st_sys(&E->r.mismatches, mism);
                st_sys(reinterpret_cast<uint32_t *>(&E->r.t_start), (uint32_t)t0);
                st_sys(reinterpret_cast<uint32_t *>(&E->r.t_start) + 1, (uint32_t)(t0 >> 32));
                st_sys(reinterpret_cast<uint32_t *>(&E->r.t_end), (uint32_t)t1);
                st_sys(reinterpret_cast<uint32_t *>(&E->r.t_end) + 1, (uint32_t)(t1 >> 32));
                // back to zero for the entry's next request (atomic stores: the
                // next request's parts update these words with atomics)
                st_relaxed(&sy->arrive, 0u);
                st_relaxed(&sy->finished, 0u);
                st_relaxed(&sy->mismatches, 0u);
                st_relaxed(&sy->status, 0u);
                __hip_atomic_store(&sy->t_start, (uint64_t)0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __threadfence();
                st_sys_release(&E->done, 1u);
            }
        }
        __syncthreads();
    }
}

// dist rows of all queries (INT_MAX, 0 at the source), the query counter and
// the launch-wide error/stats words.
__global__ void batch_reset_kernel(uint32_t *dist_all, uint32_t n, uint32_t nq,
                                   const uint32_t *sources, uint32_t *gctl) {
    const uint64_t total = (uint64_t)n * nq;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const uint32_t qi = (uint32_t)(i / n), v = (uint32_t)(i - (uint64_t)qi * n);
        dist_all[i] = v == sources[qi] ? 0u : kUnreached;
    }
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid < kCtlWords) gctl[tid] = 0u;
}

// Hop blocks from the CSR (layout in hop_block.h).  With `packed`, the
// destination word of inline edge k (v -> u) also carries, above kBackShift,
// the slot + 1 of the first edge u -> v in u's block when u is inline too
// (0: u has no such inline edge).
__global__ void build_blocks_kernel(const uint32_t *row, const uint32_t *col, const uint32_t *wgt,
                                    uint32_t n, uint32_t packed, uint4 *blocks) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        const uint32_t s = row[v], deg = row[v + 1] - s;
        uint32_t wd[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) wd[i] = 0;
        wd[0] = deg;
        wd[1] = s;
        if (deg <= kInlineEdges) {
            for (uint32_t k = 0; k < deg; ++k) {
                const uint32_t u = col[s + k];
                uint32_t back = 0;
                if (packed) {
                    const uint32_t su = row[u], du = row[u + 1] - su;
                    if (du <= kInlineEdges)
                        for (uint32_t j = du; j-- > 0;)
                            if (col[su + j] == v) back = j + 1;
                }
                wd[2 + 2 * k] = u | back << kBackShift;
                wd[3 + 2 * k] = wgt[s + k];
            }
        }
        uint4 *p = blocks + 4ull * v;
        p[0] = make_uint4(wd[0], wd[1], wd[2], wd[3]);
        p[1] = make_uint4(wd[4], wd[5], wd[6], wd[7]);
        p[2] = make_uint4(wd[8], wd[9], wd[10], wd[11]);
        p[3] = make_uint4(wd[12], wd[13], wd[14], wd[15]);
    }
}

// Light-edge count for the light/heavy split: for every vertex with more
// than kInlineEdges out-edges (adjacency sorted by weight), word 4 of its
// block = number of edges of weight < delta (binary search).
__global__ void light_count_kernel(const uint32_t *row, const uint32_t *wgt, uint32_t n,
                                   uint32_t delta, uint4 *blocks) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        const uint32_t s = row[v], e = row[v + 1];
        if (e - s <= kInlineEdges) continue;
        uint32_t lo = s, hi = e;
        while (lo < hi) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (wgt[mid] < delta) lo = mid + 1;
            else hi = mid;
        }
        blocks[4ull * v + 1].x = lo - s;
    }
}

// Mark hubs in their hop blocks: word 2 = hub id + 1, word 3 = chunk size.
__global__ void patch_hubs_kernel(const uint2 *hubs, uint32_t nh, uint4 *blocks) {
    const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h < nh) {
        uint4 *p = blocks + 4ull * hubs[h].x;
        uint4 q = p[0];
        q.z = h + 1;
        q.w = hubs[h].y;
        p[0] = q;
    }
}

// Per-query reset: dist = INT_MAX except dist[src] = 0; control block zeroed
// with tail[0] = 1; the seed record (src, 0) in slot 0 of bucket 0, which is
// where the reference's write_through of (src, 0) lands at first_pos 0
// (SS/SSSP/sssp_run.cu:473-486).  Slots are kEmpty after any successful solve.
__global__ void reset_kernel(uint32_t *dist, uint32_t n, uint32_t src, uint32_t *ctl,
                             uint64_t *slots) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = tid; i < n; i += stride) dist[i] = i == src ? 0u : kUnreached;
    for (uint32_t i = tid; i < kCtlWords; i += stride) ctl[i] = i == kTailOff ? 1u : 0u;
    if (tid == 0) slots[0] = make_rec(src, 0u);
}

__global__ void fill_u64_kernel(uint64_t *p, uint64_t count, uint64_t value) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride)
        p[i] = value;
}

__global__ void fill_u32_kernel(uint32_t *p, uint64_t count, uint32_t value) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride)
        p[i] = value;
}

// Converts a reference-layout int32 CSR row array into the solver's uint32 one.
__global__ void copy_u32_kernel(uint32_t *dst, const uint32_t *src, uint64_t count) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride)
        dst[i] = src[i];
}

// ------------------------------------------------------------------ host side

static int grid_for(uint64_t count, int threads) {
    uint64_t g = (count + threads - 1) / threads;
    if (g < 1) g = 1;
    if (g > 65536) g = 65536;
    return (int)g;
}

hipError_t launch_reset(uint32_t *dist, uint32_t n, uint32_t src, uint32_t *ctl, uint64_t *slots,
                        hipStream_t s) {
    const uint64_t work = n > kCtlWords ? n : kCtlWords;
    hipLaunchKernelGGL(reset_kernel, dim3(grid_for(work, 256)), dim3(256), 0, s, dist, n, src, ctl,
                       slots);
    return hipGetLastError();
}

hipError_t launch_fill_u64(uint64_t *p, uint64_t count, uint64_t value, hipStream_t s) {
    hipLaunchKernelGGL(fill_u64_kernel, dim3(grid_for(count, 256)), dim3(256), 0, s, p, count,
                       value);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t *p, uint64_t count, uint32_t value, hipStream_t s) {
    hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(count, 256)), dim3(256), 0, s, p, count,
                       value);
    return hipGetLastError();
}

hipError_t launch_build_blocks(const uint32_t *row, const uint32_t *col, const uint32_t *wgt,
                               uint32_t n, bool packed, uint4 *blocks, hipStream_t s) {
    hipLaunchKernelGGL(build_blocks_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, row, col, wgt,
                       n, packed ? 1u : 0u, blocks);
    return hipGetLastError();
}

hipError_t launch_light_count(const uint32_t *row, const uint32_t *wgt, uint32_t n, uint32_t delta,
                              uint4 *blocks, hipStream_t s) {
    hipLaunchKernelGGL(light_count_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, row, wgt, n, delta,
                       blocks);
    return hipGetLastError();
}

hipError_t launch_patch_hubs(const uint2 *hubs, uint32_t nh, uint4 *blocks, hipStream_t s) {
    if (nh == 0) return hipSuccess;
    hipLaunchKernelGGL(patch_hubs_kernel, dim3(grid_for(nh, 256)), dim3(256), 0, s, hubs, nh, blocks);
    return hipGetLastError();
}

hipError_t launch_copy_u32(uint32_t *dst, const uint32_t *src, uint64_t count, hipStream_t s) {
    hipLaunchKernelGGL(copy_u32_kernel, dim3(grid_for(count, 256)), dim3(256), 0, s, dst, src,
                       count);
    return hipGetLastError();
}

static const void *solver_fn(bool slf, bool hot, bool pair) {
    if (slf && hot)
        return reinterpret_cast<const void *>(&sssp_persistent_kernel<kWavesPerBlock, true, false, true>);
    if (slf && pair)
        return reinterpret_cast<const void *>(
            &sssp_persistent_kernel<kWavesPerBlock, true, false, false, false, true>);
    return slf ? reinterpret_cast<const void *>(&sssp_persistent_kernel<kWavesPerBlock, true, false, false>)
               : reinterpret_cast<const void *>(&sssp_persistent_kernel<kWavesPerBlock, false, false, false>);
}

hipError_t solver_occupancy(bool slf, bool hot, bool pair, int *blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, solver_fn(slf, hot, pair),
                                                        kWavesPerBlock * kWave, 0);
}

template <int kW, bool kExt, uint32_t kCap, bool kPair>
static const void *batch_fn_pair(bool stats) {
    return stats ? reinterpret_cast<const void *>(&sssp_batch_kernel<kW, false, kExt, kCap, true, kPair>)
                 : reinterpret_cast<const void *>(&sssp_batch_kernel<kW, false, kExt, kCap, false, kPair>);
}

template <int kW, bool kExt>
static const void *batch_fn(int cap, bool stats, bool pair) {
    if (cap == 1024)
        return pair ? batch_fn_pair<kW, kExt, 1024, true>(stats) : batch_fn_pair<kW, kExt, 1024, false>(stats);
    return pair ? batch_fn_pair<kW, kExt, kL1Cap, true>(stats) : batch_fn_pair<kW, kExt, kL1Cap, false>(stats);
}

hipError_t batch_occupancy(int waves, int l1_cap, bool ext, bool stats, bool pair, int *blocks_per_cu) {
    const void *fn =
        waves == 16 ? (ext ? batch_fn<16, true>(l1_cap, stats, pair) : batch_fn<16, false>(l1_cap, stats, pair))
        : waves == 8 ? (ext ? batch_fn<8, true>(l1_cap, stats, pair) : batch_fn<8, false>(l1_cap, stats, pair))
                     : (ext ? batch_fn<4, true>(l1_cap, stats, pair) : batch_fn<4, false>(l1_cap, stats, pair));
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, fn, waves * kWave, 0);
}

static bool lane_pairs(const SolverLaunch &L) { return lane_pairs_for(L.pop_max); }

// Kernel arguments from a host launch description.
static SolveArgs make_args(const SolverLaunch &L) {
    SolveArgs a;
    a.row = L.row;
    a.col = L.col;
    a.wgt = L.wgt;
    a.blocks = L.blocks;
    a.hubs = L.hubs;
    a.dist = L.dist;
    a.q.slots = L.slots;
    a.q.ctl = L.ctl;
    a.q.nbuckets = L.nbuckets;
    a.q.window = L.window;
    a.q.cap_log2 = L.cap_log2;
    a.q.delta = L.delta;
    a.q.batch = L.batch;
    a.q.spin_limit = L.spin_limit;
    a.q.strict = L.strict_window;
    a.n = L.n;
    a.lims = (L.min_read & 0xFFu) | (L.pop_max & 0xFFu) << 8 | (L.spill & 0xFFu) << 16 |
             (L.max_sleep & 0xFFu) << 24;
    a.flags = (L.fifo ? 1u : 0u) | (L.hubs && L.split ? 2u : 0u) | (L.stale_check & 3u) << 2 |
              (L.prefetch ? 16u : 0u) | (L.defer_heavy ? 32u : 0u) | (L.light_split ? 64u : 0u) |
              (L.donate_min & 0xFFu) << 8 | (L.csr_stream ? 1u << 16 : 0u) |
              (L.packed ? 1u << 17 : 0u);
    a.timeout_ticks = L.timeout_ticks;
    return a;
}

hipError_t launch_batch_reset(uint32_t *dist_all, uint32_t n, uint32_t nq, const uint32_t *sources,
                              uint32_t *gctl, hipStream_t s) {
    const uint64_t work = (uint64_t)n * nq > kCtlWords ? (uint64_t)n * nq : kCtlWords;
    hipLaunchKernelGGL(batch_reset_kernel, dim3(grid_for(work, 256)), dim3(256), 0, s, dist_all, n,
                       nq, sources, gctl);
    return hipGetLastError();
}

template <int kW, bool kExt, uint32_t kCap, bool kPair>
static void launch_batch_instance(const dim3 &grid, const BatchArgs &b, bool stats, hipStream_t s) {
    if (stats)
        hipLaunchKernelGGL((sssp_batch_kernel<kW, false, kExt, kCap, true, kPair>), grid, dim3(kW * kWave), 0,
                           s, b);
    else
        hipLaunchKernelGGL((sssp_batch_kernel<kW, false, kExt, kCap, false, kPair>), grid, dim3(kW * kWave), 0,
                           s, b);
}

template <int kW, bool kExt>
static void launch_batch_cap(const dim3 &grid, const BatchArgs &b, int cap, bool stats, bool pair,
                             hipStream_t s) {
    if (cap == 1024 && pair) launch_batch_instance<kW, kExt, 1024, true>(grid, b, stats, s);
    else if (cap == 1024) launch_batch_instance<kW, kExt, 1024, false>(grid, b, stats, s);
    else if (pair) launch_batch_instance<kW, kExt, kL1Cap, true>(grid, b, stats, s);
    else launch_batch_instance<kW, kExt, kL1Cap, false>(grid, b, stats, s);
}

template <int kW>
static void launch_batch_ext(const dim3 &grid, const BatchArgs &b, bool ext, int cap, bool stats,
                             bool pair, hipStream_t s) {
    if (ext) launch_batch_cap<kW, true>(grid, b, cap, stats, pair, s);
    else launch_batch_cap<kW, false>(grid, b, cap, stats, pair, s);
}

hipError_t launch_batch(const SolverLaunch &L, const BatchLaunch &B, hipStream_t s) {
    BatchArgs b;
    b.base = make_args(L);
    b.ctl_all = B.ctl_all;
    b.slots_all = B.slots_all;
    b.dist_all = B.dist_all;
    b.sources = B.sources;
    b.nq = B.nq;
    b.gctl = B.gctl;
    b.qstatus = B.qstatus;
    b.qtime = B.qtime;
    const dim3 grid(L.grid_blocks);
    const bool ext = L.light_split || L.defer_heavy || L.strict_window;
    const bool pair = lane_pairs(L);
    if (B.waves == 16) launch_batch_ext<16>(grid, b, ext, B.l1_cap, B.stats, pair, s);
    else if (B.waves == 8) launch_batch_ext<8>(grid, b, ext, B.l1_cap, B.stats, pair, s);
    else launch_batch_ext<4>(grid, b, ext, B.l1_cap, B.stats, pair, s);
    return hipGetLastError();
}

size_t serve_request_bytes() { return sizeof(ServeRequest); }

void serve_request_write(void *table, uint32_t i, const SolverLaunch &L, uint32_t src, bool fresh,
                         const uint32_t *expect, uint32_t wg_first, uint32_t wg_count) {
    ServeRequest &r = static_cast<ServeRequest *>(table)[i];
    r.a = make_args(L);
    r.src = src;
    r.status = 0;
    r.fresh = fresh ? 1u : 0u;
    r.mismatches = 0;
    r.wg_first = wg_first;
    r.wg_count = wg_count;
    r.expect = expect;
    r.t_start = r.t_end = 0;
}

void serve_request_read(const void *table, uint32_t i, uint32_t *status, uint32_t *mismatches,
                        uint64_t *t_start, uint64_t *t_end) {
    const ServeRequest &r = static_cast<const ServeRequest *>(table)[i];
    *status = r.status;
    *mismatches = r.mismatches;
    *t_start = r.t_start;
    *t_end = r.t_end;
}

hipError_t launch_serve(const ServeLaunch &S, hipStream_t s) {
    if (S.nreq == 0) return hipSuccess;
    auto *reqs = static_cast<ServeRequest *>(S.reqs);
    const uint64_t longest = S.max_part_slots > S.max_n ? S.max_part_slots : (uint64_t)S.max_n;
    uint32_t chunks = (uint32_t)((longest + 256 * 64 - 1) / (256 * 64));
    if (chunks > 512) chunks = 512;
    hipLaunchKernelGGL(serve_reset_kernel, dim3(chunks, S.nreq), dim3(256), 0, s, reqs);
    const dim3 grid(S.grid);
    const uint32_t nr = S.nreq;
    if (S.waves == 16 && S.ext)
        hipLaunchKernelGGL((sssp_serve_kernel<16, true>), grid, dim3(16 * kWave), 0, s, reqs, nr);
    else if (S.waves == 16)
        hipLaunchKernelGGL((sssp_serve_kernel<16>), grid, dim3(16 * kWave), 0, s, reqs, nr);
    else if (S.waves == 8 && S.ext)
        hipLaunchKernelGGL((sssp_serve_kernel<8, true>), grid, dim3(8 * kWave), 0, s, reqs, nr);
    else if (S.waves == 8)
        hipLaunchKernelGGL((sssp_serve_kernel<8>), grid, dim3(8 * kWave), 0, s, reqs, nr);
    else if (S.waves == 4 && S.ext)
        hipLaunchKernelGGL((sssp_serve_kernel<4, true>), grid, dim3(4 * kWave), 0, s, reqs, nr);
    else if (S.waves == 4)
        hipLaunchKernelGGL((sssp_serve_kernel<4>), grid, dim3(4 * kWave), 0, s, reqs, nr);
    else if (S.ext)
        hipLaunchKernelGGL((sssp_serve_kernel<1, true>), grid, dim3(kWave), 0, s, reqs, nr);
    else
        hipLaunchKernelGGL((sssp_serve_kernel<1>), grid, dim3(kWave), 0, s, reqs, nr);
    return hipGetLastError();
}

hipError_t launch_serve_verify(const ServeLaunch &S, hipStream_t s) {
    if (S.nreq == 0) return hipSuccess;
    uint32_t chunks = (uint32_t)(((uint64_t)S.max_n + 256 * 16 - 1) / (256 * 16));
    if (chunks > 256) chunks = 256;
    hipLaunchKernelGGL(serve_verify_kernel, dim3(chunks, S.nreq), dim3(256), 0, s,
                       static_cast<ServeRequest *>(S.reqs));
    return hipGetLastError();
}

size_t stream_entry_bytes() { return sizeof(StreamEntry); }
size_t stream_sync_bytes() { return sizeof(StreamSync); }

void stream_entry_write(void *entry, const SolverLaunch &L, uint32_t src, bool fresh, const uint32_t *expect,
                        uint32_t first, uint32_t parts, uint32_t seq) {
    StreamEntry &e = *static_cast<StreamEntry *>(entry);
    serve_request_write(&e.r, 0, L, src, fresh, expect, 0, parts);
    e.first = first;
    e.parts = parts;
    e.seq = seq;
    e.done = 0;
}

void stream_entry_read(const void *entry, uint32_t *done, uint32_t *status, uint32_t *mismatches,
                       uint64_t *t_start, uint64_t *t_end) {
    const volatile StreamEntry &e = *static_cast<const volatile StreamEntry *>(entry);
    *done = e.done;
    if (*done) {
        std::atomic_thread_fence(std::memory_order_acquire);
        *status = e.r.status;
        *mismatches = e.r.mismatches;
        *t_start = e.r.t_start;
        *t_end = e.r.t_end;
    }
}

hipError_t stream_occupancy(int *blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(
        blocks_per_cu, reinterpret_cast<const void *>(&sssp_stream_kernel<1>), kWave, 0);
}

hipError_t launch_stream(const StreamLaunch &S, hipStream_t s) {
    StreamArgs a;
    a.host = static_cast<StreamHostCtl *>(S.host_ctl);
    a.entries = static_cast<StreamEntry *>(S.entries);
    a.ring = S.ring;
    a.ring_size = S.ring_size;
    a.dev = S.dev;
    a.sync = static_cast<StreamSync *>(S.sync);
    a.lifetime_ticks = S.lifetime_ticks;
    hipLaunchKernelGGL((sssp_stream_kernel<1>), dim3(S.grid), dim3(kWave), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_solver(const SolverLaunch &L, hipStream_t s) {
    const SolveArgs a = make_args(L);
    const dim3 grid(L.grid_blocks), block(kWavesPerBlock * kWave);
    const bool hot = L.slf && L.hot;
    const bool ext = L.light_split || L.defer_heavy || L.strict_window;
    const bool pair = L.slf && !hot && lane_pairs(L);
    if (pair && !L.profile && ext)
        hipLaunchKernelGGL((sssp_persistent_kernel<kWavesPerBlock, true, false, false, true, true>), grid, block,
                           0, s, a);
    else if (pair && !L.profile)
        hipLaunchKernelGGL((sssp_persistent_kernel<kWavesPerBlock, true, false, false, false, true>), grid, block,
                           0, s, a);
    else if (pair && ext)
        hipLaunchKernelGGL((sssp_persistent_kernel<kWavesPerBlock, true, true, false, true, true>), grid, block, 0,
                           s, a);
    else if (pair)
        hipLaunchKernelGGL((sssp_persistent_kernel<kWavesPerBlock, true, true, false, false, true>), grid, block,
                           0, s, a);
    else if (L.slf && hot && !L.profile)
        hipLaunchKernelGGL((sssp_persistent_kernel<kWavesPerBlock, true, false, true>), grid, block, 0, s, a);
    else if (L.slf && hot)
        hipLaunchKernelGGL((sssp_persistent_kernel<kWavesPerBlock, true, true, true>), grid, block, 0, s, a);
    else if (L.slf && !L.profile && ext)
        hipLaunchKernelGGL((sssp_persistent_kernel<kWavesPerBlock, true, false, false, true>), grid, block, 0, s, a);
    else if (L.slf && !L.profile)
        hipLaunchKernelGGL((sssp_persistent_kernel<kWavesPerBlock, true, false, false>), grid, block, 0, s, a);
    else if (L.slf && ext)
        hipLaunchKernelGGL((sssp_persistent_kernel<kWavesPerBlock, true, true, false, true>), grid, block, 0, s, a);
    else if (L.slf)
        hipLaunchKernelGGL((sssp_persistent_kernel<kWavesPerBlock, true, true, false>), grid, block, 0, s, a);
    else if (!L.profile && ext)
        hipLaunchKernelGGL((sssp_persistent_kernel<kWavesPerBlock, false, false, false, true>), grid, block, 0, s, a);
    else if (!L.profile)
        hipLaunchKernelGGL((sssp_persistent_kernel<kWavesPerBlock, false, false, false>), grid, block, 0, s, a);
    else if (ext)
        hipLaunchKernelGGL((sssp_persistent_kernel<kWavesPerBlock, false, true, false, true>), grid, block, 0, s, a);
    else
        hipLaunchKernelGGL((sssp_persistent_kernel<kWavesPerBlock, false, true, false>), grid, block, 0, s, a);
    return hipGetLastError();
}

}  // namespace mlmq
