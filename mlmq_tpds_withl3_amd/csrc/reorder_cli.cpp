// reorder_gr -- BFS vertex renumbering of a .gr graph, with the command line,
// output files and report lines of graph_transform/reorder_gr.cu:59-257:
//   reorder_gr <input.gr> <output.gr> [BFS-root] [--layer-split]
// Writes <output.gr>, <output.gr>.perm.u32 (perm[new] = old) and
// <output.gr>.invperm.u32 (invperm[old] = new), each mapping n little-endian
// uint32 with no header (:32-45).  Existing outputs are never overwritten.
// Exit codes: 0 ok; 2 bad arguments, bad root or an output that exists or
// cannot be created; 255 when the graph cannot be read or written.
// Host-only: built from the graph tools alone, no GPU runtime.
#include <fcntl.h>
#include <unistd.h>

#include <cerrno>
#include <chrono>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mlmq_graph.h"

static double now_ms() {
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

static bool exists(const std::string &p) { return access(p.c_str(), F_OK) == 0; }

// n little-endian uint32 entries; refuses to replace an existing file.
static bool write_mapping(const std::string &path, const std::vector<int32_t> &map) {
    const int fd = open(path.c_str(), O_WRONLY | O_CREAT | O_EXCL, 0644);
    if (fd < 0) {
        std::perror(path.c_str());
        return false;
    }
    FILE *f = fdopen(fd, "wb");
    if (!f) {
        close(fd);
        return false;
    }
    std::vector<unsigned char> buf(4 * map.size());
    for (size_t i = 0; i < map.size(); ++i) {
        const uint32_t v = (uint32_t)map[i];
        buf[4 * i] = (unsigned char)v;
        buf[4 * i + 1] = (unsigned char)(v >> 8);
        buf[4 * i + 2] = (unsigned char)(v >> 16);
        buf[4 * i + 3] = (unsigned char)(v >> 24);
    }
    const bool ok = buf.empty() || std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
    return (std::fclose(f) == 0) && ok;
}

int main(int argc, char **argv) {
    if (argc < 3 || argc > 5) {
        std::printf("Usage: ./reorder_gr <input.gr> <output.gr> [BFS-root] [--layer-split]\n");
        return 255;
    }
    const char *in_path = argv[1];
    const std::string out_path = argv[2];
    const bool layer_split = argc == 5;
    if (layer_split && std::strcmp(argv[4], "--layer-split") != 0) return 2;
    long root = 0;
    if (argc >= 4) {
        char *end = nullptr;
        errno = 0;
        root = std::strtol(argv[3], &end, 10);
        if (errno || end == argv[3] || *end || root < 0 || root > INT_MAX) return 2;
    }
    const double t0 = now_ms();
    mlmq_host_graph *g = nullptr;
    if (mlmq_graph_read_gr(in_path, &g) != 0) {
        std::printf("Read error!\n");
        return 255;
    }
    int32_t n = 0;
    int64_t m = 0;
    mlmq_graph_dims(g, &n, &m);
    if (n <= 0 || root >= n) {
        std::fprintf(stderr, "Invalid graph/root\n");
        mlmq_graph_free(g);
        return 2;
    }
    const double read_ms = now_ms() - t0;

    const double t1 = now_ms();
    const int32_t cut = n / 2;
    const int64_t before = mlmq_graph_cross_edges(g, cut);
    std::vector<int32_t> perm((size_t)n), inv((size_t)n);
    mlmq_host_graph *r = nullptr;
    const int rc = mlmq_graph_reorder_bfs(g, (int32_t)root, layer_split ? 1 : 0, perm.data(),
                                          inv.data(), &r);
    mlmq_graph_free(g);
    if (rc != 0) {
        std::printf("ERROR: reorder failed (%d)\n", rc);
        return 255;
    }
    const int64_t after = mlmq_graph_cross_edges(r, cut);
    const double reorder_ms = now_ms() - t1;

    const double t2 = now_ms();
    const std::string perm_path = out_path + ".perm.u32", inv_path = out_path + ".invperm.u32";
    if (exists(out_path) || exists(perm_path) || exists(inv_path)) {
        std::fprintf(stderr, "Output already exists; use a new path\n");
        mlmq_graph_free(r);
        return 2;
    }
    // Reserve the graph's name first, as the mappings are written before it.
    const int fd = open(out_path.c_str(), O_WRONLY | O_CREAT | O_EXCL, 0644);
    if (fd < 0) {
        std::perror(out_path.c_str());
        mlmq_graph_free(r);
        return 2;
    }
    close(fd);
    if (!write_mapping(perm_path, perm) || !write_mapping(inv_path, inv)) {
        mlmq_graph_free(r);
        return 2;
    }
    if (mlmq_graph_write_gr(r, out_path.c_str()) != 0) {
        std::printf("Write error!\n");
        mlmq_graph_free(r);
        return 255;
    }
    mlmq_graph_free(r);

    const double t3 = now_ms();
    std::printf("REORDER layout=%s root_old=%ld root_new=0 vertices=%d edges=%lld read_ms=%.6f "
                "reorder_ms=%.6f output_ms=%.6f total_ms=%.6f\n",
                layer_split ? "bfs_layer_split" : "bfs", root, n, (long long)m, read_ms,
                reorder_ms, t3 - t2, t3 - t0);
    const double pb = m ? 100.0 * (double)before / (double)m : 0.0;
    const double pa = m ? 100.0 * (double)after / (double)m : 0.0;
    std::printf("cross_edges before: %lld (%.2f%%)\n", (long long)before, pb);
    std::printf("cross_edges after : %lld (%.2f%%)\n", (long long)after, pa);
    if (after > 0)
        std::printf("reduction: %.1fx\n", (double)before / (double)after);
    else
        std::printf("reduction: inf\n");
    return 0;
}
