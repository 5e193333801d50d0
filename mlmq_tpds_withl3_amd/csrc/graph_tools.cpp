// Host graph formats and seeded synthetic inputs (include/mlmq_graph.h).
#include "../../include/mlmq_graph.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <thread>
#include <vector>

struct mlmq_host_graph {
    int32_t n = 0;
    int64_t m = 0;
    std::vector<int64_t> row;
    std::vector<int32_t> col;
    std::vector<int32_t> w;
    std::vector<double> xy;     // k-NN graphs: 2 coordinates per vertex
    double weight_scale = 0.0;
};

namespace {

constexpr int kOk = 0, kInvalid = -1, kNoMem = -4, kIo = -5;

// splitmix64: seeds and per-item hashes.
inline uint64_t splitmix64(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline uint64_t hash64(uint64_t x) {
    uint64_t s = x;
    return splitmix64(s);
}
// xoshiro256** stream.
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t t = seed;
        for (auto &v : s) v = splitmix64(t);
    }
    static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    inline uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    inline double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

unsigned worker_count() {
    unsigned t = std::thread::hardware_concurrency();
    if (t == 0) t = 1;
    if (t > 16) t = 16;
    return t;
}

template <typename F>
void parallel_for(int64_t count, F f) {
    const unsigned T = worker_count();
    if (count < 4096 || T == 1) {
        f(0, count);
        return;
    }
    std::vector<std::thread> th;
    const int64_t chunk = (count + T - 1) / T;
    for (unsigned t = 0; t < T; ++t) {
        const int64_t lo = t * chunk, hi = std::min<int64_t>(count, lo + chunk);
        if (lo >= hi) break;
        th.emplace_back([=] { f(lo, hi); });
    }
    for (auto &x : th) x.join();
}

inline uint32_t spread_bits(uint32_t x) {  // 16 -> 32 bits, interleave zeros
    x &= 0xFFFF;
    x = (x | (x << 8)) & 0x00FF00FF;
    x = (x | (x << 4)) & 0x0F0F0F0F;
    x = (x | (x << 2)) & 0x33333333;
    x = (x | (x << 1)) & 0x55555555;
    return x;
}

// Undirected pair list (u < v, sorted, unique) -> symmetric CSR with each
// adjacency sorted by destination.
void pairs_to_csr(int32_t n, const std::vector<uint64_t> &pairs, const std::vector<int32_t> &pw,
                  mlmq_host_graph *g) {
    std::vector<int64_t> deg((size_t)n + 1, 0);
    for (uint64_t p : pairs) {
        deg[(uint32_t)(p >> 32)]++;
        deg[(uint32_t)p]++;
    }
    g->row.assign((size_t)n + 1, 0);
    for (int32_t i = 0; i < n; ++i) g->row[i + 1] = g->row[i] + deg[i];
    g->m = g->row[n];
    g->col.assign((size_t)g->m, 0);
    g->w.assign((size_t)g->m, 0);
    std::vector<int64_t> fill(g->row.begin(), g->row.end() - 1);
    // Pairs are sorted by (u, v): the u-side adjacency gets increasing v; the
    // v-side needs a sort afterwards.
    for (size_t i = 0; i < pairs.size(); ++i) {
        const uint32_t u = (uint32_t)(pairs[i] >> 32), v = (uint32_t)pairs[i];
        g->col[fill[u]] = (int32_t)v;
        g->w[fill[u]++] = pw[i];
        g->col[fill[v]] = (int32_t)u;
        g->w[fill[v]++] = pw[i];
    }
    parallel_for(n, [&](int64_t lo, int64_t hi) {
        std::vector<std::pair<int32_t, int32_t>> tmp;
        for (int64_t u = lo; u < hi; ++u) {
            const int64_t b = g->row[u], e = g->row[u + 1];
            tmp.clear();
            for (int64_t j = b; j < e; ++j) tmp.emplace_back(g->col[j], g->w[j]);
            std::sort(tmp.begin(), tmp.end());
            for (int64_t j = b; j < e; ++j) {
                g->col[j] = tmp[j - b].first;
                g->w[j] = tmp[j - b].second;
            }
        }
    });
    g->n = n;
}

// Parallel sort of 64-bit keys: per-thread std::sort then k-way merge passes.
void sort_u64(std::vector<uint64_t> &a) {
    const unsigned T = worker_count();
    if (a.size() < (1u << 20) || T == 1) {
        std::sort(a.begin(), a.end());
        return;
    }
    const size_t chunk = (a.size() + T - 1) / T;
    std::vector<size_t> bounds;
    for (size_t lo = 0; lo < a.size(); lo += chunk) bounds.push_back(lo);
    bounds.push_back(a.size());
    std::vector<std::thread> th;
    for (size_t i = 0; i + 1 < bounds.size(); ++i)
        th.emplace_back([&, i] { std::sort(a.begin() + bounds[i], a.begin() + bounds[i + 1]); });
    for (auto &x : th) x.join();
    while (bounds.size() > 2) {
        std::vector<size_t> nb;
        std::vector<std::thread> mt;
        for (size_t i = 0; i + 1 < bounds.size(); i += 2) {
            nb.push_back(bounds[i]);
            if (i + 2 < bounds.size())
                mt.emplace_back([&, i] {
                    std::inplace_merge(a.begin() + bounds[i], a.begin() + bounds[i + 1],
                                       a.begin() + bounds[i + 2]);
                });
        }
        nb.push_back(a.size());
        for (auto &x : mt) x.join();
        bounds.swap(nb);
    }
}

}  // namespace

extern "C" {

int mlmq_graph_generate_knn(int32_t n, int32_t k, uint64_t seed, double weight_scale, int32_t morton,
                            mlmq_host_graph **out) {
    if (!out || n < 2 || k < 1 || k >= n || weight_scale == 0.0) return kInvalid;
    *out = nullptr;
    mlmq_host_graph *g = new (std::nothrow) mlmq_host_graph();
    if (!g) return kNoMem;
    try {
        // 1. points
        std::vector<double> px((size_t)n), py((size_t)n);
        Rng rng(seed);
        for (int32_t i = 0; i < n; ++i) {
            px[i] = rng.uniform();
            py[i] = rng.uniform();
        }
        // 2. optional Morton renumbering (stable by generation index on ties)
        if (morton) {
            std::vector<uint64_t> key((size_t)n);
            for (int32_t i = 0; i < n; ++i) {
                uint32_t qx = (uint32_t)std::min(65535.0, std::floor(px[i] * 65536.0));
                uint32_t qy = (uint32_t)std::min(65535.0, std::floor(py[i] * 65536.0));
                uint64_t code = (uint64_t)(spread_bits(qx) | (spread_bits(qy) << 1));
                key[i] = (code << 32) | (uint32_t)i;
            }
            sort_u64(key);
            std::vector<double> nx((size_t)n), ny((size_t)n);
            for (int32_t i = 0; i < n; ++i) {
                uint32_t o = (uint32_t)key[i];
                nx[i] = px[o];
                ny[i] = py[o];
            }
            px.swap(nx);
            py.swap(ny);
        }
        // 3. uniform grid, ~2 points per cell
        const int32_t G = std::max<int32_t>(1, (int32_t)std::floor(std::sqrt(n / 2.0)));
        const double h = 1.0 / G;
        auto cell_of = [&](double x) { return std::min<int32_t>(G - 1, (int32_t)(x * G)); };
        std::vector<int64_t> cstart((size_t)G * G + 1, 0);
        std::vector<int32_t> cpts((size_t)n);
        for (int32_t i = 0; i < n; ++i) cstart[(size_t)cell_of(py[i]) * G + cell_of(px[i]) + 1]++;
        for (size_t c = 0; c < (size_t)G * G; ++c) cstart[c + 1] += cstart[c];
        {
            std::vector<int64_t> f(cstart.begin(), cstart.end() - 1);
            for (int32_t i = 0; i < n; ++i) cpts[f[(size_t)cell_of(py[i]) * G + cell_of(px[i])]++] = i;
        }
        // 4. k nearest neighbours (ties by index), ring search over cells
        std::vector<int32_t> nbr((size_t)n * k);
        std::vector<double> nbd((size_t)n * k);
        parallel_for(n, [&](int64_t lo, int64_t hi) {
            std::vector<std::pair<double, int32_t>> best;
            for (int64_t i = lo; i < hi; ++i) {
                best.clear();
                const int32_t cx = cell_of(px[i]), cy = cell_of(py[i]);
                for (int32_t r = 0;; ++r) {
                    for (int32_t yy = cy - r; yy <= cy + r; ++yy) {
                        if (yy < 0 || yy >= G) continue;
                        for (int32_t xx = cx - r; xx <= cx + r; ++xx) {
                            if (xx < 0 || xx >= G) continue;
                            if (std::max(std::abs(xx - cx), std::abs(yy - cy)) != r) continue;
                            const size_t c = (size_t)yy * G + xx;
                            for (int64_t t = cstart[c]; t < cstart[c + 1]; ++t) {
                                const int32_t j = cpts[t];
                                if (j == i) continue;
                                const double dx = px[i] - px[j], dy = py[i] - py[j];
                                best.emplace_back(dx * dx + dy * dy, j);
                            }
                        }
                    }
                    if ((int32_t)best.size() >= k) {
                        std::partial_sort(best.begin(), best.begin() + k, best.end());
                        best.resize(k);
                        const double reach = r * h;
                        if (best[k - 1].first <= reach * reach || r > 2 * G) break;
                    }
                    if (r > 2 * G) break;
                }
                for (int32_t t = 0; t < k; ++t) {
                    nbr[(size_t)i * k + t] = best[t].second;
                    nbd[(size_t)i * k + t] = std::sqrt(best[t].first);
                }
            }
        });
        // 5. symmetrise + dedupe
        std::vector<uint64_t> pairs;
        pairs.reserve((size_t)n * k);
        double max_len = 0;
        for (int32_t i = 0; i < n; ++i)
            for (int32_t t = 0; t < k; ++t) {
                const uint32_t j = (uint32_t)nbr[(size_t)i * k + t];
                const uint32_t a = std::min<uint32_t>(i, j), b = std::max<uint32_t>(i, j);
                pairs.push_back(((uint64_t)a << 32) | b);
                max_len = std::max(max_len, nbd[(size_t)i * k + t]);
            }
        sort_u64(pairs);
        pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
        const double scale = weight_scale > 0 ? weight_scale : (-weight_scale) / max_len;
        std::vector<int32_t> pw(pairs.size());
        parallel_for((int64_t)pairs.size(), [&](int64_t lo, int64_t hi) {
            for (int64_t i = lo; i < hi; ++i) {
                const uint32_t a = (uint32_t)(pairs[i] >> 32), b = (uint32_t)pairs[i];
                const double dx = px[a] - px[b], dy = py[a] - py[b];
                const double wv = std::llround(std::sqrt(dx * dx + dy * dy) * scale);
                pw[i] = (int32_t)std::max(1.0, std::min(wv, 2147483646.0));
            }
        });
        pairs_to_csr(n, pairs, pw, g);
        g->weight_scale = scale;
        g->xy.resize((size_t)2 * n);
        for (int32_t i = 0; i < n; ++i) {
            g->xy[2 * i] = px[i];
            g->xy[2 * i + 1] = py[i];
        }
    } catch (const std::bad_alloc &) {
        delete g;
        return kNoMem;
    }
    *out = g;
    return kOk;
}

int mlmq_graph_generate_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                             uint64_t seed, int32_t wmin, int32_t wmax, mlmq_host_graph **out) {
    if (!out || scale < 1 || scale > 30 || edge_factor < 1 || a < 0 || b < 0 || c < 0 ||
        a + b + c > 1.0 || wmin < 0 || wmax < wmin)
        return kInvalid;
    *out = nullptr;
    mlmq_host_graph *g = new (std::nothrow) mlmq_host_graph();
    if (!g) return kNoMem;
    try {
        const int32_t n = 1 << scale;
        const int64_t ne = (int64_t)edge_factor << scale;
        const uint64_t ta = (uint64_t)(a * 4294967296.0), tab = (uint64_t)((a + b) * 4294967296.0),
                       tabc = (uint64_t)((a + b + c) * 4294967296.0);
        std::vector<uint64_t> pairs((size_t)ne);
        // Edge i draws from its own stream, so the result is independent of threading.
        parallel_for(ne, [&](int64_t lo, int64_t hi) {
            for (int64_t i = lo; i < hi; ++i) {
                Rng r(seed ^ hash64((uint64_t)i * 0x9E3779B97F4A7C15ull + 1));
                uint32_t u = 0, v = 0;
                for (int32_t lvl = 0; lvl < scale; ++lvl) {
                    const uint64_t x = r.next() >> 32;
                    const uint32_t bit = 1u << (scale - 1 - lvl);
                    if (x < ta) {
                    } else if (x < tab) {
                        v |= bit;
                    } else if (x < tabc) {
                        u |= bit;
                    } else {
                        u |= bit;
                        v |= bit;
                    }
                }
                if (u == v) pairs[i] = ~0ull;  // self loop: dropped below
                else pairs[i] = ((uint64_t)std::min(u, v) << 32) | std::max(u, v);
            }
        });
        sort_u64(pairs);
        pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
        if (!pairs.empty() && pairs.back() == ~0ull) pairs.pop_back();
        std::vector<int32_t> pw(pairs.size());
        const uint64_t span = (uint64_t)(wmax - wmin) + 1;
        parallel_for((int64_t)pairs.size(), [&](int64_t lo, int64_t hi) {
            for (int64_t i = lo; i < hi; ++i)
                pw[i] = wmin + (int32_t)(hash64(pairs[i] ^ (seed * 0xD1B54A32D192ED03ull)) % span);
        });
        pairs_to_csr(n, pairs, pw, g);
    } catch (const std::bad_alloc &) {
        delete g;
        return kNoMem;
    }
    *out = g;
    return kOk;
}

int mlmq_graph_read_gr(const char *path, mlmq_host_graph **out) {
    if (!path || !out) return kInvalid;
    *out = nullptr;
    int fd = open(path, O_RDONLY);
    if (fd < 0) return kIo;
    struct stat sb;
    if (fstat(fd, &sb) != 0 || sb.st_size < 32) {
        close(fd);
        return kIo;
    }
    const size_t bytes = (size_t)sb.st_size;
    void *map = mmap(nullptr, bytes, PROT_READ, MAP_PRIVATE, fd, 0);
    close(fd);
    if (map == MAP_FAILED) return kIo;
    const unsigned char *p = (const unsigned char *)map;
    uint64_t h[4];
    std::memcpy(h, p, sizeof h);
    int rc = kOk;
    mlmq_host_graph *g = nullptr;
    if (h[0] != 1 || h[1] != 4 || h[2] == 0 || h[2] >= 0x7FFFFFFFull || h[3] >= 0x7FFFFFFFull) {
        rc = kInvalid;
    } else {
        const uint64_t n = h[2], m = h[3];
        if (bytes < 32 + 8 * n + 4 * m + 4 * (m & 1) + 4 * m) {
            rc = kInvalid;
        } else {
            g = new (std::nothrow) mlmq_host_graph();
            if (!g) rc = kNoMem;
            else {
                g->n = (int32_t)n;
                g->m = (int64_t)m;
                g->row.assign(n + 1, 0);
                const unsigned char *ends = p + 32;
                for (uint64_t u = 0; u < n && rc == kOk; ++u) {
                    uint64_t e;
                    std::memcpy(&e, ends + 8 * u, 8);
                    if (e > m || (int64_t)e < g->row[u]) rc = kInvalid;
                    g->row[u + 1] = (int64_t)e;
                }
                if (rc == kOk && g->row[n] != (int64_t)m) rc = kInvalid;
                if (rc == kOk) {
                    g->col.resize(m);
                    g->w.resize(m);
                    const unsigned char *dst = ends + 8 * n;
                    const unsigned char *wt = dst + 4 * m + 4 * (m & 1);
                    std::memcpy(g->col.data(), dst, 4 * m);
                    std::memcpy(g->w.data(), wt, 4 * m);
                    for (uint64_t e = 0; e < m; ++e)
                        if (g->col[e] < 0 || (uint64_t)g->col[e] >= n || g->w[e] < 0) {
                            rc = kInvalid;
                            break;
                        }
                }
            }
        }
    }
    munmap(map, bytes);
    if (rc != kOk) {
        delete g;
        return rc;
    }
    *out = g;
    return kOk;
}

int mlmq_graph_from_csr(int32_t n, int64_t m, const int64_t *row, const int32_t *col,
                        const int32_t *w, mlmq_host_graph **out) {
    if (!out || n <= 0 || m < 0 || !row || (m > 0 && (!col || !w)) || row[0] != 0 || row[n] != m)
        return kInvalid;
    mlmq_host_graph *g = new (std::nothrow) mlmq_host_graph();
    if (!g) return kNoMem;
    g->n = n;
    g->m = m;
    g->row.assign(row, row + n + 1);
    g->col.assign(col, col + m);
    g->w.assign(w, w + m);
    *out = g;
    return kOk;
}

int mlmq_graph_write_gr(const mlmq_host_graph *g, const char *path) {
    if (!g || !path) return kInvalid;
    FILE *f = std::fopen(path, "wb");
    if (!f) return kIo;
    const uint64_t h[4] = {1, 4, (uint64_t)g->n, (uint64_t)g->m};
    bool ok = std::fwrite(h, sizeof h, 1, f) == 1;
    std::vector<uint64_t> ends((size_t)g->n);
    for (int32_t u = 0; u < g->n; ++u) ends[u] = (uint64_t)g->row[u + 1];
    ok = ok && std::fwrite(ends.data(), 8, ends.size(), f) == ends.size();
    ok = ok && std::fwrite(g->col.data(), 4, g->col.size(), f) == g->col.size();
    if (g->m & 1) {
        const uint32_t pad = 0;
        ok = ok && std::fwrite(&pad, 4, 1, f) == 1;
    }
    ok = ok && std::fwrite(g->w.data(), 4, g->w.size(), f) == g->w.size();
    ok = (std::fclose(f) == 0) && ok;
    return ok ? kOk : kIo;
}

int mlmq_graph_dims(const mlmq_host_graph *g, int32_t *n, int64_t *m) {
    if (!g) return kInvalid;
    if (n) *n = g->n;
    if (m) *m = g->m;
    return kOk;
}

int mlmq_graph_arrays(const mlmq_host_graph *g, const int64_t **row, const int32_t **col,
                      const int32_t **w) {
    if (!g) return kInvalid;
    if (row) *row = g->row.data();
    if (col) *col = g->col.data();
    if (w) *w = g->w.data();
    return kOk;
}

int mlmq_graph_coords(const mlmq_host_graph *g, const double **xy) {
    if (!g || !xy || g->xy.empty()) return kInvalid;
    *xy = g->xy.data();
    return kOk;
}

double mlmq_graph_weight_scale(const mlmq_host_graph *g) { return g ? g->weight_scale : 0.0; }

int32_t mlmq_graph_centroid_vertex(const mlmq_host_graph *g) {
    if (!g || g->xy.empty()) return -1;
    double cx = 0, cy = 0;
    for (int32_t i = 0; i < g->n; ++i) {
        cx += g->xy[2 * i];
        cy += g->xy[2 * i + 1];
    }
    cx /= g->n;
    cy /= g->n;
    int32_t best = 0;
    double bd = 1e300;
    for (int32_t i = 0; i < g->n; ++i) {
        const double dx = g->xy[2 * i] - cx, dy = g->xy[2 * i + 1] - cy;
        const double d = dx * dx + dy * dy;
        if (d < bd) {
            bd = d;
            best = i;
        }
    }
    return best;
}

int32_t mlmq_graph_max_degree_vertex(const mlmq_host_graph *g) {
    if (!g || g->n <= 0) return -1;
    int32_t best = 0;
    int64_t bd = -1;
    for (int32_t i = 0; i < g->n; ++i) {
        const int64_t d = g->row[i + 1] - g->row[i];
        if (d > bd) {
            bd = d;
            best = i;
        }
    }
    return best;
}

int mlmq_graph_reorder_bfs(const mlmq_host_graph *g, int32_t root, int32_t layer_split,
                           int32_t *perm_out, int32_t *invperm_out, mlmq_host_graph **out) {
    if (!g || !out || g->n <= 0 || root < 0 || root >= g->n) return kInvalid;
    *out = nullptr;
    const int32_t n = g->n;
    mlmq_host_graph *r = nullptr;
    try {
        // perm is the BFS queue itself: perm[new] = old in discovery order, and
        // invperm (old -> new, -1 = not yet seen) is the visited set.
        std::vector<int32_t> perm((size_t)n), inv((size_t)n, -1);
        // level_end[i] = one past the last position of BFS level i in perm
        std::vector<int32_t> level_end;
        int32_t tail = 0;
        perm[tail++] = root;
        inv[root] = 0;
        int32_t level_stop = 1;
        for (int32_t head = 0; head < tail; ++head) {
            if (head == level_stop) {
                level_end.push_back(level_stop);
                level_stop = tail;
            }
            const int32_t u = perm[head];
            for (int64_t e = g->row[u]; e < g->row[u + 1]; ++e) {
                const int32_t v = g->col[e];
                if (inv[v] < 0) {
                    inv[v] = tail;
                    perm[tail++] = v;
                }
            }
        }
        level_end.push_back(level_stop);
        // vertices the search did not reach follow in ascending old id
        for (int32_t v = 0; v < n; ++v)
            if (inv[v] < 0) {
                inv[v] = tail;
                perm[tail++] = v;
            }
        if (tail != n) return kInvalid;
        if (tail > level_stop) level_end.push_back(tail);   // the unreached form one group

        if (layer_split) {
            // Each group (a BFS level, then the unreached) is cut in discovery
            // order: its first vertices go to the lower half [0, n/2) so that
            // the lower half holds min(n/2, ceil(end/2)) of the first `end`
            // positions, the rest go to the upper half [n/2, n).
            const int32_t cut = n / 2;
            std::vector<int32_t> split((size_t)n);
            int32_t lo = 0, hi = cut, begin = 0;
            for (const int32_t end : level_end) {
                const int32_t want = std::min<int32_t>(cut, (int32_t)(((int64_t)end + 1) / 2));
                const int32_t take = want - lo;
                for (int32_t i = begin; i < end; ++i)
                    split[(i - begin < take) ? lo++ : hi++] = perm[i];
                begin = end;
            }
            if (lo != cut || hi != n) return kInvalid;
            perm.swap(split);
            for (int32_t i = 0; i < n; ++i) inv[perm[i]] = i;
        }

        r = new mlmq_host_graph();
        r->n = n;
        r->m = g->m;
        r->weight_scale = g->weight_scale;
        r->row.assign((size_t)n + 1, 0);
        for (int32_t i = 0; i < n; ++i)
            r->row[i + 1] = r->row[i] + (g->row[perm[i] + 1] - g->row[perm[i]]);
        r->col.resize((size_t)g->m);
        r->w.resize((size_t)g->m);
        parallel_for(n, [&](int64_t lo, int64_t hi) {
            for (int64_t i = lo; i < hi; ++i) {
                const int32_t old = perm[i];
                int64_t o = r->row[i];
                for (int64_t e = g->row[old]; e < g->row[old + 1]; ++e, ++o) {
                    r->col[o] = inv[g->col[e]];
                    r->w[o] = g->w[e];
                }
            }
        });
        if (!g->xy.empty()) {
            r->xy.resize(g->xy.size());
            for (int32_t i = 0; i < n; ++i) {
                r->xy[2 * (size_t)i] = g->xy[2 * (size_t)perm[i]];
                r->xy[2 * (size_t)i + 1] = g->xy[2 * (size_t)perm[i] + 1];
            }
        }
        if (perm_out) std::memcpy(perm_out, perm.data(), sizeof(int32_t) * (size_t)n);
        if (invperm_out) std::memcpy(invperm_out, inv.data(), sizeof(int32_t) * (size_t)n);
    } catch (const std::bad_alloc &) {
        delete r;
        return kNoMem;
    }
    *out = r;
    return kOk;
}

int64_t mlmq_graph_cross_edges(const mlmq_host_graph *g, int32_t cut) {
    if (!g || cut < 0 || cut > g->n) return -1;
    int64_t c = 0;
    for (int32_t u = 0; u < g->n; ++u)
        for (int64_t e = g->row[u]; e < g->row[u + 1]; ++e)
            c += (u < cut) != (g->col[e] < cut);
    return c;
}

void mlmq_graph_free(mlmq_host_graph *g) { delete g; }

}  // extern "C"
