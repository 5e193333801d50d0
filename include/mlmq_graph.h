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
void mlmq_graph_free(mlmq_host_graph *g);

#ifdef __cplusplus
}
#endif
#endif /* MLMQ_GRAPH_H */
