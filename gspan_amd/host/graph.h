// graph.h -- data-graph types with the reference's interface (graph.h in haozixu/gSpan),
// so host code written against the reference compiles unchanged:
//   Edge{from,to,label}                      graph.h:7-12
//   Vertex, Graph (std::map<int, Vertex>)    graph.h:14-26
//   PatternEdge, DFSCode                     graph.h:28-49
//   CSRGraph{nr_nodes,nr_edges,offsets,vlabels,edges} + load/save/dump   graph.h:51-65
//   prune_graph                              graph.h:67
// The binary file layout is the reference's (graph.cc:178-214).
#ifndef GSPAN_AMD_GRAPH_H
#define GSPAN_AMD_GRAPH_H

#include <map>
#include <vector>

struct Edge {
    int from, to, label;
};

struct Vertex {
    int label = 0;
    std::vector<Edge> edges;
};

// undirected labelled graph keyed by vertex id
struct Graph : std::map<int, Vertex> {
    void add_vertex(int u, int label);
    void add_edge(int u, int v, int label);  // stored in both directions

    static Graph from_txt(const char* filename);
};

struct PatternEdge {
    int from, to, from_label, edge_label, to_label;

    PatternEdge() {}
    PatternEdge(int u, int v, int ulabel, int elabel, int vlabel)
        : from(u), to(v), from_label(ulabel), edge_label(elabel), to_label(vlabel) {}

    bool is_forward() const { return from < to; }
    bool is_backward() const { return from > to; }
    bool operator==(const PatternEdge& o) const {
        return from == o.from && to == o.to && from_label == o.from_label &&
               edge_label == o.edge_label && to_label == o.to_label;
    }
    bool operator!=(const PatternEdge& o) const { return !(*this == o); }
};

using DFSCode = std::vector<PatternEdge>;

struct CSRGraph {
    int nr_nodes, nr_edges;
    int *offsets, *vlabels;
    Edge* edges;

    CSRGraph() : nr_nodes(0), nr_edges(0), offsets(nullptr), vlabels(nullptr), edges(nullptr) {}
    CSRGraph(Graph&);          // vertices renumbered by id, adjacency sorted by (to, label)
    CSRGraph(const DFSCode&);  // pattern graph of a DFS code
    CSRGraph(const CSRGraph&) = delete;
    CSRGraph& operator=(const CSRGraph&) = delete;
    ~CSRGraph();

    void load(const char* filename);  // binary file, throws std::runtime_error
    void save(const char* filename);
    void dump() const;
};

void prune_graph(CSRGraph& pruned, const CSRGraph& g, int min_sup);

#endif
