// ref_harness.cc -- C entry points over the reference's own graph.cc (compiled from
// /root/reference by oracle/Makefile into oracle/_ref/).  TEST INFRASTRUCTURE ONLY:
// used to pin the oracle's restatement of graph.cc:122-159 (pattern graph of a DFS
// code), graph.cc:216-266 (prune_graph) and graph.cc:178-214 (binary load/save).
#include <cstring>
#include "graph.h"

extern "C" {

// CSRGraph(const DFSCode&) (graph.cc:122-159).  dims[0..1] = n, nnz.
int ref_pattern_graph(const int* code5, int m, int* dims, int* offsets, int* vlabels,
                      int* edges3) {
    DFSCode code;
    for (int i = 0; i < m; ++i)
        code.emplace_back(code5[5 * i], code5[5 * i + 1], code5[5 * i + 2], code5[5 * i + 3],
                          code5[5 * i + 4]);
    CSRGraph g(code);
    dims[0] = g.nr_nodes;
    dims[1] = g.nr_edges;
    if (offsets) std::memcpy(offsets, g.offsets, sizeof(int) * (g.nr_nodes + 1));
    if (vlabels) std::memcpy(vlabels, g.vlabels, sizeof(int) * g.nr_nodes);
    if (edges3) std::memcpy(edges3, g.edges, sizeof(Edge) * g.nr_edges);
    return 0;
}

// CSRGraph::load (graph.cc:178-200) followed by prune_graph (graph.cc:216-266); the
// pruned graph is written with CSRGraph::save (graph.cc:202-214).
int ref_prune_file(const char* in_path, int min_sup, const char* out_path) {
    CSRGraph g, pruned;
    g.load(in_path);
    prune_graph(pruned, g, min_sup);
    pruned.save(out_path);
    return 0;
}

}
