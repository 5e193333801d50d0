/*
 * mlmq_graph.h -- host-side graph formats and seeded synthetic inputs for the
 * MLMQ solver (plain C ABI, no device code).
 *
 *   .gr I/O      <- Galois .gr v1 as read by SSSP/csr_graph.cu:185-296 and
 *                   written by SSSP/csr_graph.cu:302-418 (header u64 x4
 *                   {1, 4, n, m}; n u64 row-end offsets; m u32 destinations;
 *                   4 pad bytes when m is odd; m int32 weights), validated as
 *                   scripts/multigpu/supplement_graph.cpp:16-30.
 *   generators   <- no reference counterpart (the reference reads DIMACS road
 *                   graphs); they build the seeded synthetic workloads:
 *                   planar k-nearest-neighbour "road-like" graphs (optionally
 *                   renumbered in Morton order) and symmetrised RMAT graphs.
 *   reordering   <- graph_transform/reorder_gr.cu (BFS renumbering, optional
 *                   level split, cross-edge count); the reorder_gr command
 *                   built beside the library keeps that tool's command line,
 *                   output files and report lines.
 */
#ifndef MLMQ_GRAPH_H
#define MLMQ_GRAPH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mlmq_host_graph mlmq_host_graph;

/* n points uniform in the unit square (seeded), each linked to its k nearest
 * neighbours, symmetrised and de-duplicated; weight = max(1, round(len * s)).
 * weight_scale > 0 gives s directly; weight_scale < 0 picks s so that the
 * longest edge gets weight -weight_scale.  morton != 0 renumbers vertices in
 * Morton (Z) order of their coordinates. */
int mlmq_graph_generate_knn(int32_t n, int32_t k, uint64_t seed, double weight_scale,
                            int32_t morton, mlmq_host_graph **out);
/* RMAT with 2^scale vertices and edge_factor * 2^scale sampled edges using
 * quadrant probabilities (a, b, c, 1-a-b-c); self loops and duplicates
 * removed, symmetrised; weights uniform in [wmin, wmax] per undirected edge. */
int mlmq_graph_generate_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                             uint64_t seed, int32_t wmin, int32_t wmax, mlmq_host_graph **out);
int mlmq_graph_read_gr(const char *path, mlmq_host_graph **out);
int mlmq_graph_from_csr(int32_t n, int64_t m, const int64_t *row, const int32_t *col,
                        const int32_t *w, mlmq_host_graph **out);
int mlmq_graph_write_gr(const mlmq_host_graph *g, const char *path);
int mlmq_graph_dims(const mlmq_host_graph *g, int32_t *n, int64_t *m);
int mlmq_graph_arrays(const mlmq_host_graph *g, const int64_t **row, const int32_t **col,
                      const int32_t **w);
/* Point coordinates of a k-NN graph (2 doubles per vertex, in vertex order). */
int mlmq_graph_coords(const mlmq_host_graph *g, const double **xy);
/* Weight scale actually applied by the k-NN generator (0 for other graphs). */
double mlmq_graph_weight_scale(const mlmq_host_graph *g);
/* Vertex nearest to the centroid of the k-NN points (-1 for other graphs). */
int32_t mlmq_graph_centroid_vertex(const mlmq_host_graph *g);
/* Vertex of largest out-degree (lowest id among ties). */
int32_t mlmq_graph_max_degree_vertex(const mlmq_host_graph *g);
/* Vertex renumbering in breadth-first discovery order; replaces the
 * reorder step of graph_transform/reorder_gr.cu:94-185.  The search starts
 * at `root` (new id 0) and scans each row in CSR order; vertices it does not
 * reach follow in ascending old id (:128-134).  layer_split != 0 then cuts
 * every BFS level (and finally the unreached group), in discovery order,
 * between the lower half [0, n/2) and the upper half [n/2, n), so that the
 * lower half holds min(n/2, ceil(end/2)) of the first `end` positions
 * (:142-156).  perm[new] = old and invperm[old] = new (n entries each,
 * caller-owned, either may be NULL).  *out gets the relabelled graph: rows in
 * new order, each row's edges in their old order with renamed destinations
 * and unchanged weights (:158-177); k-NN coordinates follow their vertices.
 * Returns 0, -1 for a bad argument (root outside [0, n)), -4 out of memory. */
int mlmq_graph_reorder_bfs(const mlmq_host_graph *g, int32_t root, int32_t layer_split,
                           int32_t *perm, int32_t *invperm, mlmq_host_graph **out);
/* Directed edges (u, v) with u and v on different sides of `cut` (ids below
 * cut vs the rest): the partition-quality figure graph_transform/
 * reorder_gr.cu:47-57 reports at cut = n/2.  -1 for a bad argument. */
int64_t mlmq_graph_cross_edges(const mlmq_host_graph *g, int32_t cut);
void mlmq_graph_free(mlmq_host_graph *g);

#ifdef __cplusplus
}
#endif
#endif /* MLMQ_GRAPH_H */
