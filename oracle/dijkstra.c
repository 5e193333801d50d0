/*
 * CPU oracle for single-source shortest paths on a CSR graph.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only the test suite,
 * the smoke entry point and the benchmark's CPU-baseline leg load it.  The
 * product library never links or calls it.
 *
 * What it restates (reference files, relative to the reference root):
 *   - oracle_dijkstra: the CPU path `sssp_sequential`
 *       evidence/l3_supplement_20260921_v2/single_source/SSSP/main.cu:47-117
 *     (binary min-heap of (dist, vertex), lazy deletion through a done[]
 *     array, relax on strict improvement, total_work += out-degree of every
 *     settled vertex), with the 64-bit accumulation, INT_MAX sentinel for
 *     unreachable vertices and the ">= INT_MAX is unrepresentable" guard of
 *       scripts/multigpu/supplement_graph.cpp:32-43.
 *   - oracle_bellman_ford: textbook Bellman-Ford, used only to pin
 *     oracle_dijkstra on tiny graphs (no reference counterpart).
 *   - oracle_validate_gr: the strict Galois .gr v1 validator of
 *       scripts/multigpu/supplement_graph.cpp:11-31
 *     on the layout read by SSSP/csr_graph.cu:185-296 (header u64 x4
 *     {1, 4, n, m}; n u64 row-end offsets; m u32 destinations; 4 pad bytes if
 *     m is odd; m int32 weights).
 *
 * Return codes: 0 ok; negative on error (see ORACLE_E* below).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORACLE_EINVAL  (-1)  /* bad argument / malformed CSR */
#define ORACLE_ENOMEM  (-2)
#define ORACLE_ERANGE  (-3)  /* a finite distance does not fit in int32 */

#define ORACLE_INF ((int64_t)1 << 62)
#define ORACLE_UNREACHED 0x7FFFFFFF

typedef struct { int64_t key; int32_t v; } heap_item;

typedef struct {
    heap_item *a;
    int64_t size, cap;
} min_heap;

static int heap_push(min_heap *h, int64_t key, int32_t v)
{
    if (h->size == h->cap) {
        int64_t nc = h->cap ? 2 * h->cap : 1024;
        heap_item *na = (heap_item *)realloc(h->a, (size_t)nc * sizeof(heap_item));
        if (!na) return ORACLE_ENOMEM;
        h->a = na;
        h->cap = nc;
    }
    int64_t i = h->size++;
    while (i > 0) {
        int64_t p = (i - 1) >> 1;
        if (h->a[p].key <= key) break;
        h->a[i] = h->a[p];
        i = p;
    }
    h->a[i].key = key;
    h->a[i].v = v;
    return 0;
}

static heap_item heap_pop(min_heap *h)
{
    heap_item top = h->a[0];
    heap_item last = h->a[--h->size];
    int64_t i = 0, n = h->size;
    for (;;) {
        int64_t c = 2 * i + 1;
        if (c >= n) break;
        if (c + 1 < n && h->a[c + 1].key < h->a[c].key) c++;
        if (h->a[c].key >= last.key) break;
        h->a[i] = h->a[c];
        i = c;
    }
    if (n > 0) h->a[i] = last;
    return top;
}

static int check_csr(int32_t n, const int64_t *row, const int32_t *col,
                     const int32_t *w, int32_t src)
{
    if (n <= 0 || !row || src < 0 || src >= n || row[0] != 0) return ORACLE_EINVAL;
    for (int32_t u = 0; u < n; ++u)
        if (row[u + 1] < row[u]) return ORACLE_EINVAL;
    int64_t m = row[n];
    if (m > 0 && (!col || !w)) return ORACLE_EINVAL;
    for (int64_t e = 0; e < m; ++e)
        if (col[e] < 0 || col[e] >= n || w[e] < 0) return ORACLE_EINVAL;
    return 0;
}

static int finish(int32_t n, const int64_t *d, int32_t *dist_out)
{
    for (int32_t u = 0; u < n; ++u) {
        if (d[u] == ORACLE_INF) {
            dist_out[u] = ORACLE_UNREACHED;
        } else {
            if (d[u] >= (int64_t)ORACLE_UNREACHED) return ORACLE_ERANGE;
            dist_out[u] = (int32_t)d[u];
        }
    }
    return 0;
}

/*
 * Dijkstra from `src`.  `row` has n+1 entries (0-based CSR starts, row[n] = m),
 * `col`/`w` have m entries.  Writes n distances (INT_MAX = unreachable) and,
 * when `total_work` is non-NULL, the sum of out-degrees of settled vertices
 * (the reference's "CPU total work count", i.e. E_r).
 */
int oracle_dijkstra(int32_t n, const int64_t *row, const int32_t *col,
                    const int32_t *w, int32_t src, int32_t *dist_out,
                    int64_t *total_work, int64_t *reached)
{
    int rc = check_csr(n, row, col, w, src);
    if (rc) return rc;
    int64_t *d = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    unsigned char *done = (unsigned char *)calloc((size_t)n, 1);
    min_heap h = {0, 0, 0};
    if (!d || !done) { free(d); free(done); return ORACLE_ENOMEM; }
    for (int32_t u = 0; u < n; ++u) d[u] = ORACLE_INF;
    d[src] = 0;
    int64_t work = 0, settled = 0;
    rc = heap_push(&h, 0, src);
    while (!rc && h.size > 0) {
        heap_item it = heap_pop(&h);
        int32_t u = it.v;
        if (done[u]) continue;          /* lazy deletion */
        done[u] = 1;
        settled++;
        work += row[u + 1] - row[u];
        int64_t du = d[u];
        for (int64_t e = row[u]; e < row[u + 1]; ++e) {
            int32_t v = col[e];
            int64_t nd = du + (int64_t)w[e];
            if (d[v] > nd) {            /* strict improvement */
                d[v] = nd;
                rc = heap_push(&h, nd, v);
                if (rc) break;
            }
        }
    }
    if (!rc) rc = finish(n, d, dist_out);
    if (!rc && total_work) *total_work = work;
    if (!rc && reached) *reached = settled;
    free(h.a);
    free(done);
    free(d);
    return rc;
}

/*
 * Textbook Bellman-Ford (|V|-1 rounds of relaxing every edge, early exit when
 * a round changes nothing).  O(V*E): for pinning on tiny graphs only.
 */
int oracle_bellman_ford(int32_t n, const int64_t *row, const int32_t *col,
                        const int32_t *w, int32_t src, int32_t *dist_out)
{
    int rc = check_csr(n, row, col, w, src);
    if (rc) return rc;
    int64_t *d = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    if (!d) return ORACLE_ENOMEM;
    for (int32_t u = 0; u < n; ++u) d[u] = ORACLE_INF;
    d[src] = 0;
    for (int32_t round = 0; round + 1 < n || round == 0; ++round) {
        int changed = 0;
        for (int32_t u = 0; u < n; ++u) {
            if (d[u] == ORACLE_INF) continue;
            for (int64_t e = row[u]; e < row[u + 1]; ++e) {
                int64_t nd = d[u] + (int64_t)w[e];
                if (nd < d[col[e]]) { d[col[e]] = nd; changed = 1; }
            }
        }
        if (!changed) break;
    }
    rc = finish(n, d, dist_out);
    free(d);
    return rc;
}

/*
 * Validate an in-memory .gr v1 image of `bytes` bytes.  On success writes
 * n and m.  Mirrors the checks of the reference's strict exporter: version 1,
 * 4-byte edge data, n and m below INT_MAX, no truncation, monotone row ends
 * bounded by m with the last equal to m, destinations in range, weights >= 0.
 */
int oracle_validate_gr(const unsigned char *img, int64_t bytes,
                       int64_t *n_out, int64_t *m_out)
{
    if (!img || bytes < 32) return ORACLE_EINVAL;
    uint64_t h[4];
    memcpy(h, img, sizeof(h));
    if (h[0] != 1 || h[1] != 4 || h[2] >= 0x7FFFFFFFull || h[3] >= 0x7FFFFFFFull)
        return ORACLE_EINVAL;
    uint64_t n = h[2], m = h[3];
    uint64_t need = 32 + 8 * n + 4 * m + 4 * (m & 1) + 4 * m;
    if ((uint64_t)bytes < need) return ORACLE_EINVAL;
    const unsigned char *ends = img + 32;
    const unsigned char *dst = ends + 8 * n;
    const unsigned char *wt = dst + 4 * m + 4 * (m & 1);
    uint64_t prev = 0;
    for (uint64_t u = 0; u < n; ++u) {
        uint64_t e;
        memcpy(&e, ends + 8 * u, 8);
        if (e > m || e < prev) return ORACLE_EINVAL;
        prev = e;
    }
    if (n > 0 && prev != m) return ORACLE_EINVAL;
    if (n == 0 && m != 0) return ORACLE_EINVAL;
    for (uint64_t e = 0; e < m; ++e) {
        uint32_t v;
        int32_t x;
        memcpy(&v, dst + 4 * e, 4);
        memcpy(&x, wt + 4 * e, 4);
        if (v >= n || x < 0) return ORACLE_EINVAL;
    }
    *n_out = (int64_t)n;
    *m_out = (int64_t)m;
    return 0;
}
