// graph.cc -- host implementation of the reference-compatible graph types.
// Behaviour follows graph.cc of haozixu/gSpan (cited per function); the code is this
// project's own.
#include "graph.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

void Graph::add_vertex(int u, int label) { (*this)[u].label = label; }

// graph.cc:32-36: every edge is stored in both directions
void Graph::add_edge(int u, int v, int label) {
    (*this)[u].edges.push_back(Edge{u, v, label});
    (*this)[v].edges.push_back(Edge{v, u, label});
}

// graph.cc:38-64: whitespace-separated tokens; 't' takes two more tokens, 'v' two ints,
// 'e' three ints, anything else is skipped.
Graph Graph::from_txt(const char* filename) {
    FILE* f = std::fopen(filename, "r");
    if (!f) throw std::runtime_error("failed to open file\n");
    std::unique_ptr<FILE, int (*)(FILE*)> closer(f, std::fclose);
    Graph g;
    char tok[256];
    int a, b, w, nr_edges = 0;
    while (std::fscanf(f, "%255s", tok) == 1) {
        if (tok[0] == 't') {
            char dummy[256];
            if (std::fscanf(f, "%255s%d", dummy, &a) != 2) break;
        } else if (tok[0] == 'v') {
            if (std::fscanf(f, "%d%d", &a, &w) != 2) break;
            g.add_vertex(a, w);
        } else if (tok[0] == 'e') {
            if (std::fscanf(f, "%d%d%d", &a, &b, &w) != 3) break;
            g.add_edge(a, b, w);
            ++nr_edges;
        }
    }
    std::printf("|V| = %ld |E| = %d\n", (long)g.size(), nr_edges);
    return g;
}

// graph.cc:66-112: vertex ids renumbered in ascending order; adjacency sorted by (to, label)
CSRGraph::CSRGraph(Graph& g) {
    nr_nodes = (int)g.size();
    std::map<int, int> renum;
    int next = 0;
    long long total = 0;
    for (auto& kv : g) {
        renum[kv.first] = next++;
        total += (long long)kv.second.edges.size();
    }
    nr_edges = (int)total;
    offsets = new int[nr_nodes + 1];
    vlabels = new int[nr_nodes];
    edges = new Edge[nr_edges > 0 ? nr_edges : 1];
    int v = 0, at = 0;
    offsets[0] = 0;
    for (auto& kv : g) {
        std::vector<Edge> adj;
        adj.reserve(kv.second.edges.size());
        for (const Edge& e : kv.second.edges) adj.push_back(Edge{v, renum[e.to], e.label});
        std::sort(adj.begin(), adj.end(), [](const Edge& x, const Edge& y) {
            return x.to != y.to ? x.to < y.to : x.label < y.label;
        });
        for (const Edge& e : adj) edges[at++] = e;
        vlabels[v] = kv.second.label;
        offsets[++v] = at;
    }
}

// graph.cc:122-159: pattern graph of a DFS code; each vertex's adjacency lists the code
// edges touching it in reverse code order; labels come from non-(-1) label fields
CSRGraph::CSRGraph(const DFSCode& code) {
    int vmax = 0;
    for (const auto& e : code) vmax = std::max(vmax, std::max(e.from, e.to));
    nr_nodes = vmax + 1;
    nr_edges = 2 * (int)code.size();
    offsets = new int[nr_nodes + 1];
    vlabels = new int[nr_nodes];
    edges = new Edge[nr_edges > 0 ? nr_edges : 1];
    std::vector<int> deg(nr_nodes, 0);
    for (const auto& e : code) {
        ++deg[e.from];
        ++deg[e.to];
    }
    offsets[0] = 0;
    for (int v = 0; v < nr_nodes; ++v) offsets[v + 1] = offsets[v] + deg[v];
    std::vector<int> fill(offsets + 1, offsets + 1 + nr_nodes);  // write from the end
    for (int v = 0; v < nr_nodes; ++v) vlabels[v] = 0;
    for (const auto& e : code) {
        if (e.from_label != -1) vlabels[e.from] = e.from_label;
        if (e.to_label != -1) vlabels[e.to] = e.to_label;
        edges[--fill[e.from]] = Edge{e.from, e.to, e.edge_label};
        edges[--fill[e.to]] = Edge{e.to, e.from, e.edge_label};
    }
}

CSRGraph::~CSRGraph() {
    delete[] offsets;
    delete[] vlabels;
    delete[] edges;
}

void CSRGraph::dump() const {
    std::printf("vlabels:");
    for (int v = 0; v < nr_nodes; ++v) std::printf(" %d", vlabels[v]);
    std::printf("\noffsets:");
    for (int v = 0; v <= nr_nodes; ++v) std::printf(" %d", offsets[v]);
    std::printf("\nfrom: ");
    for (int i = 0; i < nr_edges; ++i) std::printf(" %d", edges[i].from);
    std::printf("\nto:");
    for (int i = 0; i < nr_edges; ++i) std::printf(" %d", edges[i].to);
    std::printf("\n");
}

// graph.cc:178-200 (same messages)
void CSRGraph::load(const char* filename) {
    FILE* f = std::fopen(filename, "rb");
    if (!f) throw std::runtime_error("failed to open file");
    std::unique_ptr<FILE, int (*)(FILE*)> closer(f, std::fclose);
    if (std::fread(&nr_nodes, sizeof(int), 1, f) != 1) throw std::runtime_error("nr_nodes required");
    if (std::fread(&nr_edges, sizeof(int), 1, f) != 1) throw std::runtime_error("nr_edges required");
    if (nr_nodes < 0 || nr_edges < 0) throw std::runtime_error("bad header");
    delete[] offsets;
    delete[] vlabels;
    delete[] edges;
    offsets = new int[nr_nodes + 1];
    vlabels = new int[nr_nodes > 0 ? nr_nodes : 1];
    edges = new Edge[nr_edges > 0 ? nr_edges : 1];
    if (std::fread(offsets, sizeof(int), nr_nodes + 1, f) != (size_t)nr_nodes + 1)
        throw std::runtime_error("bad offsets");
    if (std::fread(vlabels, sizeof(int), nr_nodes, f) != (size_t)nr_nodes)
        throw std::runtime_error("bad vertex labels");
    if (std::fread(edges, sizeof(Edge), nr_edges, f) != (size_t)nr_edges)
        throw std::runtime_error("bad edges");
}

// graph.cc:202-214
void CSRGraph::save(const char* filename) {
    FILE* f = std::fopen(filename, "wb");
    if (!f) throw std::runtime_error("failed to open file");
    std::unique_ptr<FILE, int (*)(FILE*)> closer(f, std::fclose);
    std::fwrite(&nr_nodes, sizeof(int), 1, f);
    std::fwrite(&nr_edges, sizeof(int), 1, f);
    std::fwrite(offsets, sizeof(int), nr_nodes + 1, f);
    std::fwrite(vlabels, sizeof(int), nr_nodes, f);
    std::fwrite(edges, sizeof(Edge), nr_edges, f);
}

// prune_graph (graph.cc:216-266) runs on the GPU: see gspan.cc.

