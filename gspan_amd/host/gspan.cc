// gspan.cc -- fsm() of the reference (gspan.cu:1171-1199) and prune_graph (graph.cc:216-266)
// as thin callers of the C-ABI.
#include "gspan.h"

#include <cstdio>
#include <stdexcept>
#include <string>

#include "gspan_amd.h"

namespace {
int print_pattern(const int* c, int m, int support, void*) {
    for (int i = 0; i < m; ++i)
        std::printf("(%d %d %d %d %d)", c[5 * i], c[5 * i + 1], c[5 * i + 2], c[5 * i + 3],
                    c[5 * i + 4]);
    std::printf(": %d\n", support);
    return 0;
}
}  // namespace

void fsm(const CSRGraph& g, int k, int min_support, int mode, int mni_mode) {
    gs_opts o;
    gs_default_opts(&o);
    o.mode = mode;
    o.mni_mode = mni_mode;
    gs_ctx* ctx = nullptr;
    int rc = gs_create(g.offsets, g.vlabels, reinterpret_cast<const int*>(g.edges), g.nr_nodes,
                       g.nr_edges, &o, &ctx);
    if (rc != GS_OK) throw std::runtime_error("gs_create failed (code " + std::to_string(rc) + ")");
    rc = gs_mine(ctx, k, min_support, print_pattern, nullptr);
    std::string msg = rc == GS_OK ? "" : gs_last_error(ctx);
    gs_destroy(ctx);
    if (rc != GS_OK) throw std::runtime_error("gs_mine failed: " + msg);
}

void fsm(const CSRGraph& g, int k, int min_support) {
    fsm(g, k, min_support, GS_MODE_BATCHED, GS_MNI_GPU);
}

// graph.cc:216-266 on the device: vertices whose label occurs fewer than min_sup times
// are dropped, the rest renumbered in ascending id order with their edges between kept
// vertices in CSR order; prints the reference's size line.
void prune_graph(CSRGraph& pruned, const CSRGraph& g, int min_sup) {
    gs_ctx* ctx = nullptr;
    int rc = gs_create(g.offsets, g.vlabels, reinterpret_cast<const int*>(g.edges), g.nr_nodes,
                       g.nr_edges, nullptr, &ctx);
    if (rc != GS_OK) throw std::runtime_error("gs_create failed (code " + std::to_string(rc) + ")");
    int* offsets = new int[g.nr_nodes + 1];
    int* vlabels = new int[g.nr_nodes > 0 ? g.nr_nodes : 1];
    Edge* edges = new Edge[g.nr_edges > 0 ? g.nr_edges : 1];
    int n = 0, nnz = 0;
    rc = gs_prune_graph(ctx, min_sup, &n, &nnz, offsets, vlabels, reinterpret_cast<int*>(edges));
    const std::string msg = rc == GS_OK ? "" : gs_last_error(ctx);
    gs_destroy(ctx);
    if (rc != GS_OK) {
        delete[] offsets;
        delete[] vlabels;
        delete[] edges;
        throw std::runtime_error("gs_prune_graph failed: " + msg);
    }
    delete[] pruned.offsets;
    delete[] pruned.vlabels;
    delete[] pruned.edges;
    pruned.nr_nodes = n;
    pruned.nr_edges = nnz;
    pruned.offsets = offsets;
    pruned.vlabels = vlabels;
    pruned.edges = edges;
    std::printf("|V|: %d -> %d |E|: %d -> %d\n", g.nr_nodes, pruned.nr_nodes, g.nr_edges,
                pruned.nr_edges);
}
