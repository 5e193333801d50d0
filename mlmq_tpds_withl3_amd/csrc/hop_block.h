// Hop blocks: a 64-byte, 64-byte-aligned per-vertex copy of the CSR row.
//
//   word 0      out-degree
//   word 1      CSR start offset (row[v])
//   words 2..15 up to kInlineEdges = 7 out-edges as (destination, weight)
//
// A vertex with at most 7 out-edges (99.5 % of the road-like inputs, whose
// maximum degree is 9) is fully described by its block, so relaxing it needs
// no row-offset or column load: one aligned 64-byte load replaces the
// row -> col dependency chain of CSR.  Heavier vertices keep using the CSR
// arrays through the start offset; their inline words are zero, except for
// hubs (more than kSplitDeg out-edges), whose word 2 holds hub id + 1 and
// word 3 the chunk size their edge list is split into.  The blocks are built
// on the device from the CSR when a graph is loaded; the CSR stays the graph
// of record.
//
// Hub splitting: a wave that pops a hub keeps the first chunk of its edge
// list and pushes one chunk record per further chunk to L2, where idle waves
// take them.  A chunk record has the hub's distance and, in the vertex word,
// kChunkFlag | hub id << kChunkBits | chunk index; the hub table maps hub ids
// back to vertices.  Vertex ids are below 2^31, so the flag bit is free.
//
// Back edges: when every vertex id and weight is below 2^kBackShift (a
// "packed" graph), the destination word of inline edge k of v, to u, also
// holds in bits kBackShift..30 the slot + 1 of u's first inline edge back to
// v (0: none).  A record created by relaxing that edge keeps the word as its
// vertex word, and the expansion of the record skips that slot: an edge back to the vertex p whose expansion (at distance d_p)
// created the record offers d_p + w(p, u) + w(u, p) >= d_p >= dist[p], which
// never lowers dist[p].  On the road-like inputs this is one of about three
// inline edges of every expanded vertex.
#pragma once

#include "device_common.h"

namespace mlmq {

constexpr uint32_t kInlineEdges = 7;
constexpr uint32_t kSplitDeg = 2048;        // hubs: more out-edges than this
constexpr uint32_t kChunkEdges = 1024;      // default chunk of a hub's edge list
constexpr uint32_t kChunkFlag = 0x80000000u;
constexpr uint32_t kChunkBits = 16;         // chunk index bits (chunks per hub < 2^16)
constexpr uint32_t kMaxHubs = 1u << 15;     // hub id bits: 15
constexpr uint32_t kBackShift = 28;         // back-edge slot bits: 28..30

__device__ __host__ inline uint32_t chunk_vertex_word(uint32_t hub, uint32_t ci) {
    return kChunkFlag | (hub << kChunkBits) | ci;
}

struct HopBlock {
    uint4 q[4];
    __device__ __forceinline__ uint32_t deg() const { return q[0].x; }
    __device__ __forceinline__ uint32_t start() const { return q[0].y; }
};

__device__ __forceinline__ HopBlock load_block(const uint4 *blocks, uint32_t v) {
    const uint4 *p = blocks + 4ull * v;
    HopBlock b;
    b.q[0] = p[0];
    b.q[1] = p[1];
    b.q[2] = p[2];
    b.q[3] = p[3];
    return b;
}

__device__ __forceinline__ HopBlock empty_block() {
    HopBlock b;
    b.q[0] = b.q[1] = b.q[2] = b.q[3] = make_uint4(0u, 0u, 0u, 0u);
    return b;
}

}  // namespace mlmq
