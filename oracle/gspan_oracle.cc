// gspan_oracle.cc -- CPU restatement of the reference's single-large-graph gSpan path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only as
// the checker (or the timed CPU baseline), never as the product path.
//
// Every function below restates the behaviour of one reference function and cites
// it (file:line in haozixu/gSpan).  The data structures deliberately follow the
// reference's CPU path (per-thread std::map groupings, bitmaps with atomic OR) so
// that, timed as the CPU baseline, it costs what the reference's OpenMP path costs.
//
// Parity pins (tests/test_oracle_*.py, tests/test_golden.py, tests/test_prune.py):
//   * loaders / CSR build      -> the reference's own convert (oracle/_ref, container only,
//                                 and its outputs committed in tests/golden), hand graphs,
//                                 binary-format round trips
//   * pattern graph            -> the reference's CSRGraph(const DFSCode&) (oracle/_ref and
//                                 tests/golden/pattern_graphs.json)
//   * rules / extension        -> brute-force monomorphism enumeration on small graphs
//   * MNI (both modes)         -> closed forms (K_n, stars) + brute force
//   * canonicality             -> exhaustive minimum-DFS-code enumeration
//   * output order / format    -> parity unpinned: restated from gspan.cu:67-77,
//                                 1142-1168 (the reference's mining code does not build
//                                 here, so no run of it can check the line order)
//
// Build: oracle/Makefile (g++ -O3 -fopenmp -shared).

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include <omp.h>

namespace oracle {

// graph.h:7-12
struct Edge {
    int from, to, label;
};

// graph.h:28-47
struct PatternEdge {
    int from, to, from_label, edge_label, to_label;
    bool is_forward() const { return from < to; }
    bool operator==(const PatternEdge& o) const {
        return from == o.from && to == o.to && from_label == o.from_label &&
               edge_label == o.edge_label && to_label == o.to_label;
    }
    bool operator!=(const PatternEdge& o) const { return !(*this == o); }
};
using DFSCode = std::vector<PatternEdge>;

// graph.h:51-65 (owning vectors instead of raw new[])
struct CSR {
    int n = 0, nnz = 0;
    const int* offsets = nullptr;
    const int* vlabels = nullptr;
    const Edge* edges = nullptr;
};

// gspan.h:18-28
struct Rel {
    int prev, eid;
};
using Rels = std::vector<Rel>;
using Layers = std::vector<const Rels*>;

// gspan.cu:41-50: order of single-edge codes (from_label, edge_label, to_label)
struct InitialLess {
    bool operator()(const PatternEdge& a, const PatternEdge& b) const {
        if (a.from_label != b.from_label) return a.from_label < b.from_label;
        if (a.edge_label != b.edge_label) return a.edge_label < b.edge_label;
        return a.to_label < b.to_label;
    }
};

// gspan.cu:53-65: indices of the forward edges on the path v0 -> v_n, last edge first
static std::vector<int> right_most_path(const DFSCode& c) {
    std::vector<int> rm;
    int last_v = -1;
    for (int i = (int)c.size() - 1; i >= 0; --i) {
        if (c[i].is_forward() && (rm.empty() || c[i].to == last_v)) {
            rm.push_back(i);
            last_v = c[i].from;
        }
    }
    return rm;
}

// gspan.cu:385-394: the m graph edges of a layered row, in code order
static void history_of(const CSR& g, const Layers& L, int idx, std::vector<Edge>& h) {
    h.resize(L.size());
    for (int layer = (int)L.size() - 1; layer >= 0; --layer) {
        const Rel& r = (*L[layer])[idx];
        h[layer] = g.edges[r.eid];
        idx = r.prev;
    }
}

// gspan.cu:396-401: undirected match including the label
static bool has_edge(const std::vector<Edge>& h, const Edge& e) {
    for (const Edge& x : h)
        if (x.label == e.label &&
            ((x.from == e.from && x.to == e.to) || (x.to == e.from && x.from == e.to)))
            return true;
    return false;
}

// gspan.cu:403-408
static bool has_vertex(const std::vector<Edge>& h, int v) {
    for (const Edge& x : h)
        if (x.from == v || x.to == v) return true;
    return false;
}

// gspan.cu:428-445: first neighbour entry of u_n closing a backward edge to e.from
static bool backward_edge(const CSR& g, const Edge& e, int u_n, const std::vector<Edge>& h,
                          int* out) {
    for (int i = g.offsets[u_n]; i < g.offsets[u_n + 1]; ++i) {
        Edge cur{u_n, g.edges[i].to, g.edges[i].label};
        if (has_edge(h, cur)) continue;
        if (cur.to == e.from &&
            (e.label < cur.label ||
             (e.label == cur.label && g.vlabels[e.to] <= g.vlabels[u_n]))) {
            *out = i;
            return true;
        }
    }
    return false;
}

// gspan.cu:447-462
template <class Fn>
static bool first_forward(const CSR& g, int u_n, int min_label, const std::vector<Edge>& h,
                          const Fn& fn) {
    bool any = false;
    for (int i = g.offsets[u_n]; i < g.offsets[u_n + 1]; ++i) {
        int el = g.edges[i].label, to = g.edges[i].to, tl = g.vlabels[to];
        if (tl < min_label || has_vertex(h, to)) continue;
        fn(el, tl, i);
        any = true;
    }
    return any;
}

// gspan.cu:464-483
template <class Fn>
static bool other_forward(const CSR& g, const Edge& e, int min_label,
                          const std::vector<Edge>& h, const Fn& fn) {
    bool any = false;
    int bound = g.vlabels[e.to];
    int u = e.from;
    for (int i = g.offsets[u]; i < g.offsets[u + 1]; ++i) {
        int el = g.edges[i].label, to = g.edges[i].to, tl = g.vlabels[to];
        if (tl < min_label || has_vertex(h, to)) continue;
        if (e.label < el || (e.label == el && bound <= tl)) {
            fn(el, tl, i);
            any = true;
        }
    }
    return any;
}

// gspan.cu:608-609
using ForwardMap = std::map<int, std::map<int, std::map<int, Rels>>>;
using BackwardMap = std::map<int, std::map<int, Rels>>;

// gspan.cu:674-780 (extend_omp): per-thread groupings, serial prefix over threads,
// resize, parallel copy.  Child lists keyed backward[v_i][el], forward[v_i][el][tl].
static long long extend_omp(const CSR& g, const DFSCode& code, const Layers& L,
                            const std::vector<int>& rm, ForwardMap& fwd, BackwardMap& bwd,
                            int nthreads) {
    using BCounts = std::map<int, std::map<int, long long>>;
    using FCounts = std::map<int, std::map<int, std::map<int, long long>>>;
    std::vector<BackwardMap> lb(nthreads);
    std::vector<ForwardMap> lf(nthreads);
    std::vector<BCounts> cb(nthreads);
    std::vector<FCounts> cf(nthreads);

    const int min_label = code[0].from_label;
    const int v_n = code[rm[0]].to;
    const Rels& rows = *L.back();
    const int N = (int)rows.size();

#pragma omp parallel num_threads(nthreads)
    {
        int tid = omp_get_thread_num();
        std::vector<Edge> h;
#pragma omp for schedule(guided)
        for (int idx = 0; idx < N; ++idx) {
            history_of(g, L, idx, h);
            int u_n = h[rm[0]].to;
            for (int i = (int)rm.size() - 1; i >= 1; --i) {
                int k = rm[i], ei;
                int v_i = code[k].from;
                if (backward_edge(g, h[k], u_n, h, &ei)) {
                    int el = g.edges[ei].label;
                    lb[tid][v_i][el].push_back(Rel{idx, ei});
                    cb[tid][v_i][el]++;
                }
            }
            first_forward(g, u_n, min_label, h, [&](int el, int tl, int ei) {
                lf[tid][v_n][el][tl].push_back(Rel{idx, ei});
                cf[tid][v_n][el][tl]++;
            });
            for (size_t i = 0; i < rm.size(); ++i) {
                int k = rm[i];
                int v_i = code[k].from;
                other_forward(g, h[k], min_label, h, [&](int el, int tl, int ei) {
                    lf[tid][v_i][el][tl].push_back(Rel{idx, ei});
                    cf[tid][v_i][el][tl]++;
                });
            }
        }
    }

    // running prefix over threads (gspan.cu:737-746)
    for (int t = 1; t < nthreads; ++t) {
        for (auto& a : cb[t - 1])
            for (auto& b : a.second) cb[t][a.first][b.first] += b.second;
        for (auto& a : cf[t - 1])
            for (auto& b : a.second)
                for (auto& c : b.second) cf[t][a.first][b.first][c.first] += c.second;
    }
    long long total = 0;
    // reserve (gspan.cu:748-755)
    for (auto& a : cb[nthreads - 1])
        for (auto& b : a.second) {
            bwd[a.first][b.first].resize(b.second);
            total += b.second;
        }
    for (auto& a : cf[nthreads - 1])
        for (auto& b : a.second)
            for (auto& c : b.second) {
                fwd[a.first][b.first][c.first].resize(c.second);
                total += c.second;
            }
    // copy (gspan.cu:757-774); the destination maps are fully built above, so the
    // parallel section only reads the map structure.
#pragma omp parallel num_threads(nthreads)
    {
        int t = omp_get_thread_num();
        for (auto& a : lb[t])
            for (auto& b : a.second) {
                Rels& dst = bwd.find(a.first)->second.find(b.first)->second;
                long long off = t > 0 ? cb[t - 1][a.first][b.first] : 0;
                std::copy(b.second.begin(), b.second.end(), dst.begin() + off);
            }
        for (auto& a : lf[t])
            for (auto& b : a.second)
                for (auto& c : b.second) {
                    Rels& dst =
                        fwd.find(a.first)->second.find(b.first)->second.find(c.first)->second;
                    long long off = t > 0 ? cf[t - 1][a.first][b.first][c.first] : 0;
                    std::copy(c.second.begin(), c.second.end(), dst.begin() + off);
                }
    }
    return total;
}

// Bitmap sized as bitmap.cuh:29-32 (bytes rounded up to 16); insert as bitmap.cuh:6-18.
struct Bitmaps {
    int n = 0;
    size_t words = 0;  // 64-bit words per bitmap
    std::vector<std::vector<uint64_t>> bits;
    std::vector<long long> counts;
    void init(int nv, int nvert) {
        n = nvert;
        size_t bytes = (((size_t)(n + 7) / 8) + 15) & ~(size_t)15;
        words = (bytes + 7) / 8;
        if ((int)bits.size() < nv) bits.resize(nv);
        counts.assign(nv, 0);
        for (int v = 0; v < nv; ++v) bits[v].assign(words, 0);
    }
    void insert(int v, uint32_t k) {
        uint64_t bit = 1ull << (k & 63);
        uint64_t old = __atomic_fetch_or(&bits[v][k >> 6], bit, __ATOMIC_RELAXED);
        if (!(old & bit)) __atomic_fetch_add(&counts[v], 1, __ATOMIC_RELAXED);
    }
};

enum MniMode { MNI_GPU = 0, MNI_OMP_LITERAL = 1 };

// MNI support.
//  MNI_GPU:         gspan.cu:165-190 + 282-288 (the shipped path): vertex 0 from layer 0's
//                   `from`, vertex v>0 from the forward layer whose `to` is v.
//  MNI_OMP_LITERAL: gspan.cu:321-374: forward layers insert `to`; every layer with
//                   prev_idx != -1 inserts `from` into code[layer].from (so vertex 0 is
//                   only counted when some later edge leaves it).  The reference's
//                   parallel clear (gspan.cu:342-347) is replaced by a plain clear.
// Both return min over the non-zero set sizes (INT32_MAX when all are empty).
static int support(const CSR& g, const DFSCode& code, const Layers& L, int mode, int nthreads,
                   Bitmaps& bm) {
    int vmax = 0;
    for (auto& e : code) vmax = std::max(vmax, std::max(e.from, e.to));
    const int nv = vmax + 1;
    bm.init(nv, g.n);
    const Rels& last = *L.back();
    const int N = (int)last.size();
    const int m = (int)code.size();
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int row = 0; row < N; ++row) {
        int i = row;
        for (int layer = m - 1; layer >= 0; --layer) {
            const Rel& r = (*L[layer])[i];
            const PatternEdge& e = code[layer];
            const Edge& ge = g.edges[r.eid];
            if (mode == MNI_GPU) {
                if (e.is_forward()) bm.insert(e.to, (uint32_t)ge.to);
                if (layer == 0) bm.insert(0, (uint32_t)ge.from);
            } else {
                if (e.is_forward()) bm.insert(e.to, (uint32_t)ge.to);
                if (r.prev != -1) bm.insert(e.from, (uint32_t)ge.from);
            }
            i = r.prev;
        }
    }
    long long sup = 0x7fffffff;
    for (int v = 0; v < nv; ++v)
        if (bm.counts[v] > 0 && sup > bm.counts[v]) sup = bm.counts[v];
    return (int)sup;
}

// ---------------------------------------------------------------- canonicality
// graph.cc:122-159: pattern graph of a DFS code.  Adjacency lists are filled from the
// end (reverse insertion order); vertex labels come from the non-(-1) label fields.
struct PatternGraph {
    int n = 0, nnz = 0;
    std::vector<int> offsets, vlabels;
    std::vector<Edge> edges;
    CSR view() const {
        CSR c;
        c.n = n;
        c.nnz = nnz;
        c.offsets = offsets.data();
        c.vlabels = vlabels.data();
        c.edges = edges.data();
        return c;
    }
};

static PatternGraph pattern_graph(const DFSCode& code) {
    PatternGraph p;
    std::vector<int> deg(code.size() + 1, 0);
    int vmax = 0;
    for (auto& e : code) {
        vmax = std::max(vmax, std::max(e.from, e.to));
        ++deg[e.from];
        ++deg[e.to];
    }
    p.n = vmax + 1;
    p.nnz = 2 * (int)code.size();
    p.offsets.assign(p.n + 1, 0);
    p.vlabels.assign(p.n, 0);
    p.edges.assign(p.nnz, Edge{0, 0, 0});
    for (int v = 0; v < p.n; ++v) p.offsets[v + 1] = p.offsets[v] + deg[v];
    for (auto& e : code) {
        int u = e.from, v = e.to;
        if (e.from_label != -1) p.vlabels[u] = e.from_label;
        if (e.to_label != -1) p.vlabels[v] = e.to_label;
        int i = p.offsets[u] + (--deg[u]);
        p.edges[i] = Edge{u, v, e.edge_label};
        int j = p.offsets[v] + (--deg[v]);
        p.edges[j] = Edge{v, u, e.edge_label};
    }
    return p;
}

// gspan.h:10-16: pointer-linked projection of a code into the pattern graph.  Stored
// here as (parent index into the previous level, edge index).
struct Proj {
    int prev, eid;
};

struct ProjLevel {
    std::vector<Proj> p;
};

// gspan.cu:376-383: edges of a projection, in code order
static void proj_history(const CSR& g, const std::vector<ProjLevel>& lv, int idx,
                         std::vector<Edge>& h) {
    int m = (int)lv.size();
    h.resize(m);
    for (int l = m - 1; l >= 0; --l) {
        const Proj& p = lv[l].p[idx];
        h[l] = g.edges[p.eid];
        idx = p.prev;
    }
}

// gspan.cu:485-584: greedy growth of the minimum DFS code along all projections,
// compared against the candidate at every step.  Returns true when no further edge
// passes the rules (gspan.cu:583), exactly as the reference does.
static bool projection_is_canonical(const CSR& g, std::vector<ProjLevel>& lv,
                                    const DFSCode& code, DFSCode& minc, bool* overrun) {
    for (;;) {
        std::vector<int> rm = right_most_path(minc);
        const int min_label = minc[0].from_label;
        const int v_n = minc[rm[0]].to;
        const ProjLevel& cur = lv.back();
        const int P = (int)cur.p.size();
        std::vector<Edge> h;

        // backward edges, root side of the rightmost path first (gspan.cu:500-512)
        bool has_b = false;
        int v_i = -1;
        std::map<int, std::vector<Proj>> bmap;
        for (int i = (int)rm.size() - 1; !has_b && i >= 1; --i) {
            int k = rm[i];
            for (int pi = 0; pi < P; ++pi) {
                proj_history(g, lv, pi, h);
                int u_n = h[rm[0]].to;
                int ei;
                if (backward_edge(g, h[k], u_n, h, &ei)) {
                    bmap[g.edges[ei].label].push_back(Proj{pi, ei});
                    v_i = minc[k].from;
                    has_b = true;
                }
            }
        }
        if (has_b) {  // gspan.cu:514-524
            auto it = bmap.begin();
            PatternEdge pe{v_n, v_i, -1, it->first, -1};
            if (minc.size() >= code.size()) {
                *overrun = true;
                return true;
            }
            if (code[minc.size()] != pe) return false;
            minc.push_back(pe);
            lv.push_back(ProjLevel{it->second});
            continue;
        }

        // forward edges (gspan.cu:526-567)
        bool has_f = false;
        std::map<int, std::map<int, std::vector<Proj>>> fmap;
        for (int pi = 0; pi < P; ++pi) {
            proj_history(g, lv, pi, h);
            int u_n = h[rm[0]].to;
            if (first_forward(g, u_n, min_label, h, [&](int el, int tl, int ei) {
                    fmap[el][tl].push_back(Proj{pi, ei});
                }))
                has_f = true;
        }
        if (has_f) v_i = v_n;
        for (size_t i = 0; !has_f && i < rm.size(); ++i) {
            int k = rm[i];
            for (int pi = 0; pi < P; ++pi) {
                proj_history(g, lv, pi, h);
                if (other_forward(g, h[k], min_label, h, [&](int el, int tl, int ei) {
                        fmap[el][tl].push_back(Proj{pi, ei});
                    }))
                    has_f = true;
            }
            if (has_f) {
                v_i = minc[k].from;
                break;
            }
        }
        if (has_f) {  // gspan.cu:569-581
            auto it1 = fmap.begin();
            auto it2 = it1->second.begin();
            PatternEdge pe{v_i, v_n + 1, -1, it1->first, it2->first};
            if (minc.size() >= code.size()) {
                *overrun = true;
                return true;
            }
            if (code[minc.size()] != pe) return false;
            minc.push_back(pe);
            lv.push_back(ProjLevel{it2->second});
            continue;
        }
        return true;
    }
}

// gspan.cu:586-606
static bool is_canonical(const DFSCode& code, bool* overrun) {
    *overrun = false;
    if (code.size() == 1) return true;
    PatternGraph pg = pattern_graph(code);
    CSR g = pg.view();
    std::map<PatternEdge, std::vector<Proj>, InitialLess> init;
    // gspan.cu:410-426 on the pattern graph
    for (int from = 0; from < g.n; ++from) {
        int fl = g.vlabels[from];
        for (int i = g.offsets[from]; i < g.offsets[from + 1]; ++i) {
            int to = g.edges[i].to, el = g.edges[i].label, tl = g.vlabels[to];
            if (fl <= tl) init[PatternEdge{0, 1, fl, el, tl}].push_back(Proj{-1, i});
        }
    }
    auto it = init.begin();
    DFSCode minc{it->first};
    std::vector<ProjLevel> lv{ProjLevel{it->second}};
    return projection_is_canonical(g, lv, code, minc, overrun);
}

// ---------------------------------------------------------------- driver
struct Stats {
    // rows of canonical codes with m edges (level_rows[m]) and child rows extension emits
    // into level m (level_emitted[m]), for the level-size tests
    long long level_rows[9] = {0}, level_emitted[9] = {0};
    long long extended = 0;       // child rows emitted by extension (gspan.cu:708,717,730)
    long long candidates = 0;     // mine() invocations
    long long canonical = 0;      // candidates passing is_canonical
    long long reported = 0;       // frequent patterns reported
    long long overruns = 0;       // canonical walks that ran past the candidate (must be 0)
    double t_extend = 0, t_support = 0, t_canon = 0;
};

typedef int (*report_cb)(const int* code5, int m, int support, void* user);

struct Miner {
    const CSR& g;
    int k, min_sup, mode, nthreads;
    report_cb cb;
    void* user;
    Stats st;
    Bitmaps bm;
    bool stop = false;
    double deadline = 0;  // omp_get_wtime() bound for a bounded sample run (0 = none)

    Miner(const CSR& g_, int k_, int ms, int mode_, int nt, report_cb cb_, void* u)
        : g(g_), k(k_), min_sup(ms), mode(mode_), nthreads(nt), cb(cb_), user(u) {}

    void report(const DFSCode& c, int sup) {  // gspan.cu:67-77
        ++st.reported;
        if (!cb) return;
        std::vector<int> buf(5 * c.size());
        for (size_t i = 0; i < c.size(); ++i) {
            buf[5 * i + 0] = c[i].from;
            buf[5 * i + 1] = c[i].to;
            buf[5 * i + 2] = c[i].from_label;
            buf[5 * i + 3] = c[i].edge_label;
            buf[5 * i + 4] = c[i].to_label;
        }
        if (cb(buf.data(), (int)c.size(), sup, user) != 0) stop = true;
    }

    // gspan.cu:1118-1169
    void mine(DFSCode& code, Layers& L) {
        if (stop) return;
        if (deadline > 0 && omp_get_wtime() > deadline) {
            stop = true;
            return;
        }
        ++st.candidates;
        double t0 = omp_get_wtime();
        bool overrun = false;
        bool canon = is_canonical(code, &overrun);
        st.t_canon += omp_get_wtime() - t0;
        if (overrun) ++st.overruns;
        if (!canon) return;
        ++st.canonical;
        if (code.size() < 9) st.level_rows[code.size()] += (long long)L.back()->size();
        t0 = omp_get_wtime();
        int sup = support(g, code, L, mode, nthreads, bm);
        st.t_support += omp_get_wtime() - t0;
        if (sup < min_sup) return;
        report(code, sup);
        if ((int)code.size() >= k) return;

        BackwardMap bwd;
        ForwardMap fwd;
        std::vector<int> rm = right_most_path(code);
        int v_n = code[rm[0]].to;
        t0 = omp_get_wtime();
        const long long emitted = extend_omp(g, code, L, rm, fwd, bwd, nthreads);
        st.extended += emitted;
        if (code.size() + 1 < 9) st.level_emitted[code.size() + 1] += emitted;
        st.t_extend += omp_get_wtime() - t0;

        for (auto& a : bwd)  // backward children: ascending v_i, then edge label
            for (auto& b : a.second) {
                code.push_back(PatternEdge{v_n, a.first, -1, b.first, -1});
                L.push_back(&b.second);
                mine(code, L);
                L.pop_back();
                code.pop_back();
            }
        for (auto a = fwd.rbegin(); a != fwd.rend(); ++a)  // forward: descending v_i
            for (auto& b : a->second)
                for (auto& c : b.second) {
                    code.push_back(PatternEdge{a->first, v_n + 1, -1, b.first, c.first});
                    L.push_back(&c.second);
                    mine(code, L);
                    L.pop_back();
                    code.pop_back();
                }
    }

    // gspan.cu:1171-1190
    void run() {
        std::map<PatternEdge, Rels, InitialLess> init;
        for (int from = 0; from < g.n; ++from) {
            int fl = g.vlabels[from];
            for (int i = g.offsets[from]; i < g.offsets[from + 1]; ++i) {
                int to = g.edges[i].to, el = g.edges[i].label, tl = g.vlabels[to];
                if (fl <= tl) init[PatternEdge{0, 1, fl, el, tl}].push_back(Rel{-1, i});
            }
        }
        Layers L;
        for (auto& p : init) {
            DFSCode code{p.first};
            L.push_back(&p.second);
            mine(code, L);
            L.pop_back();
            if (stop) break;
        }
    }
};

// ---------------------------------------------------------------- level sizes
// Not a reference function: an estimate of how many rows each DFS level holds when every
// pattern is frequent, for sizing the workloads (graphs too large to mine on the CPU).
// Knuth's random-probe estimator of a tree's level sizes (D. E. Knuth, "Estimating the
// efficiency of backtrack programs", Math. Comp. 29, 1975): a probe starts at a uniformly
// random single-edge row (weight = number of single-edge rows) and at each level moves to
// a uniformly random child row among those whose code is canonical (the rows the miner
// keeps and extends), multiplying its weight by their number.  The mean weight at depth m
// is the number of rows of canonical m-edge codes; the mean of weight x (all children of
// the row, canonical or not) is the number of rows extension emits into depth m + 1.
// Children follow gspan.cu:694-735 exactly (the functions above).
struct ChildRow {
    PatternEdge e;
    int entry;
};

static void children_of_row(const CSR& g, const DFSCode& code, const std::vector<Edge>& h,
                            std::vector<ChildRow>& out) {
    out.clear();
    const std::vector<int> rm = right_most_path(code);
    const int v_n = code[rm[0]].to;
    const int u_n = h[rm[0]].to;
    const int min_label = code[0].from_label;
    for (int i = (int)rm.size() - 1; i >= 1; --i) {  // gspan.cu:701-711
        int idx;
        if (backward_edge(g, h[rm[i]], u_n, h, &idx))
            out.push_back(ChildRow{PatternEdge{v_n, code[rm[i]].from, -1, g.edges[idx].label, -1},
                                   idx});
    }
    first_forward(g, u_n, min_label, h, [&](int el, int tl, int i) {  // gspan.cu:713-720
        out.push_back(ChildRow{PatternEdge{v_n, v_n + 1, -1, el, tl}, i});
    });
    for (size_t i = 0; i < rm.size(); ++i) {  // gspan.cu:722-734
        const int v_i = code[rm[i]].from;
        other_forward(g, h[rm[i]], min_label, h, [&](int el, int tl, int e) {
            out.push_back(ChildRow{PatternEdge{v_i, v_n + 1, -1, el, tl}, e});
        });
    }
}

static std::string code_key(const DFSCode& c) {
    std::string s;
    s.resize(c.size() * sizeof(PatternEdge));
    std::memcpy(&s[0], c.data(), s.size());
    return s;
}

static uint64_t splitmix(uint64_t& st) {
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static DFSCode decode(const int* code5, int m) {
    DFSCode c(m);
    for (int i = 0; i < m; ++i)
        c[i] = PatternEdge{code5[5 * i], code5[5 * i + 1], code5[5 * i + 2], code5[5 * i + 3],
                           code5[5 * i + 4]};
    return c;
}

static CSR make_csr(const int* offsets, const int* vlabels, const int* edges3, int n, int nnz) {
    CSR g;
    g.n = n;
    g.nnz = nnz;
    g.offsets = offsets;
    g.vlabels = vlabels;
    g.edges = reinterpret_cast<const Edge*>(edges3);
    return g;
}

}  // namespace oracle

using namespace oracle;

extern "C" {

// Mining run stopped at the first mine() call after max_seconds (0 = unbounded); returns
// 2 when the bound cut the run short.  Used for bounded CPU-baseline samples.
int oracle_fsm_bounded(const int* offsets, const int* vlabels, const int* edges3, int n, int nnz,
                       int k, int min_sup, int mni_mode, int nthreads, report_cb cb, void* user,
                       long long* stats_out, double max_seconds) {
    try {
        if (nthreads <= 0) nthreads = omp_get_max_threads();
        CSR g = make_csr(offsets, vlabels, edges3, n, nnz);
        Miner mn(g, k, min_sup, mni_mode, nthreads, cb, user);
        if (max_seconds > 0) mn.deadline = omp_get_wtime() + max_seconds;
        mn.run();
        if (stats_out) {
            stats_out[0] = mn.st.extended;
            stats_out[1] = mn.st.candidates;
            stats_out[2] = mn.st.canonical;
            stats_out[3] = mn.st.reported;
            stats_out[4] = mn.st.overruns;
            stats_out[5] = (long long)(mn.st.t_extend * 1e6);
            stats_out[6] = (long long)(mn.st.t_support * 1e6);
            stats_out[7] = (long long)(mn.st.t_canon * 1e6);
        }
        return mn.stop ? 2 : 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "oracle_fsm: %s\n", e.what());
        return 1;
    }
}

// Full mining run (gspan.cu:1171-1199).  stats_out[0..7]: extended, candidates, canonical,
// reported, overruns, t_extend_us, t_support_us, t_canon_us.
int oracle_fsm(const int* offsets, const int* vlabels, const int* edges3, int n, int nnz, int k,
               int min_sup, int mni_mode, int nthreads, report_cb cb, void* user,
               long long* stats_out) {
    return oracle_fsm_bounded(offsets, vlabels, edges3, n, nnz, k, min_sup, mni_mode, nthreads, cb,
                              user, stats_out, 0.0);
}

// is_canonical (gspan.cu:586-606).  Returns 1/0; 2 when the greedy walk ran past the code.
// Exact level sizes: rows of canonical codes per edge count (rows[1..k]) and child rows
// emitted into each level (emitted[2..k]) of a full mining run at min_support 1, where
// every canonical code with a row is reported and (below k) extended.
int oracle_level_counts(const int* offsets, const int* vlabels, const int* edges3, int n,
                        int nnz, int k, int nthreads, long long* rows, long long* emitted) {
    try {
        if (nthreads <= 0) nthreads = omp_get_max_threads();
        CSR g = make_csr(offsets, vlabels, edges3, n, nnz);
        Miner mn(g, k, 1, 0, nthreads, nullptr, nullptr);
        mn.run();
        for (int m = 0; m <= k && m < 9; ++m) {
            rows[m] = mn.st.level_rows[m];
            emitted[m] = mn.st.level_emitted[m];
        }
        return 0;
    } catch (...) {
        return 1;
    }
}

// Knuth estimates of the same sizes from `probes` random probes (see children_of_row):
// rows[m], emitted[m] = mean estimates; rows_sq[m], emitted_sq[m] = mean squared weights
// (for standard errors).  Probes are split over threads with independent generators.
int oracle_level_estimate(const int* offsets, const int* vlabels, const int* edges3, int n,
                          int nnz, int k, long long probes, unsigned long long seed, int nthreads,
                          double* rows, double* rows_sq, double* emitted, double* emitted_sq) {
    try {
        CSR g = make_csr(offsets, vlabels, edges3, n, nnz);
        if (k < 1 || k > 8) return 1;
        long long seeds_total = 0;
        for (int i = 0; i < nnz; ++i)
            if (vlabels[g.edges[i].from] <= vlabels[g.edges[i].to]) ++seeds_total;
        for (int m = 0; m <= k; ++m) rows[m] = rows_sq[m] = emitted[m] = emitted_sq[m] = 0;
        if (seeds_total == 0 || probes <= 0) return 0;
        if (nthreads > 0) omp_set_num_threads(nthreads);
        std::vector<double> acc_r(9, 0), acc_r2(9, 0), acc_e(9, 0), acc_e2(9, 0);
#pragma omp parallel
        {
            std::vector<double> r(9, 0), r2(9, 0), e(9, 0), e2(9, 0);
            std::unordered_map<std::string, bool> canon;
            std::vector<ChildRow> kids, ckids;
            uint64_t st = seed * 0x100000001B3ull + (uint64_t)omp_get_thread_num() * 7919ull + 1;
#pragma omp for schedule(dynamic, 256)
            for (long long pr = 0; pr < probes; ++pr) {
                int entry;
                do {
                    entry = (int)(splitmix(st) % (uint64_t)nnz);
                } while (vlabels[g.edges[entry].from] > vlabels[g.edges[entry].to]);
                const Edge& s0 = g.edges[entry];
                DFSCode code{PatternEdge{0, 1, vlabels[s0.from], s0.label, vlabels[s0.to]}};
                std::vector<Edge> h{s0};
                double w = (double)seeds_total;
                r[1] += w;
                r2[1] += w * w;
                for (int m = 1; m < k; ++m) {
                    children_of_row(g, code, h, kids);
                    const double we = w * (double)kids.size();
                    e[m + 1] += we;
                    e2[m + 1] += we * we;
                    ckids.clear();
                    for (const ChildRow& c : kids) {
                        code.push_back(c.e);
                        const std::string key = code_key(code);
                        auto it = canon.find(key);
                        bool ok;
                        if (it == canon.end()) {
                            bool ov = false;
                            ok = is_canonical(code, &ov);
                            canon.emplace(key, ok);
                        } else {
                            ok = it->second;
                        }
                        code.pop_back();
                        if (ok) ckids.push_back(c);
                    }
                    if (ckids.empty()) break;
                    const ChildRow& pick = ckids[splitmix(st) % (uint64_t)ckids.size()];
                    w *= (double)ckids.size();
                    code.push_back(pick.e);
                    h.push_back(g.edges[pick.entry]);
                    r[m + 1] += w;
                    r2[m + 1] += w * w;
                }
            }
#pragma omp critical
            for (int m = 0; m < 9; ++m) {
                acc_r[m] += r[m];
                acc_r2[m] += r2[m];
                acc_e[m] += e[m];
                acc_e2[m] += e2[m];
            }
        }
        for (int m = 0; m <= k; ++m) {
            rows[m] = acc_r[m] / (double)probes;
            rows_sq[m] = acc_r2[m] / (double)probes;
            emitted[m] = acc_e[m] / (double)probes;
            emitted_sq[m] = acc_e2[m] / (double)probes;
        }
        return 0;
    } catch (...) {
        return 1;
    }
}

int oracle_is_canonical(const int* code5, int m) {
    bool overrun = false;
    bool r = is_canonical(decode(code5, m), &overrun);
    if (overrun) return 2;
    return r ? 1 : 0;
}

// MNI support of one pattern from its layered rows.  layers[l] points at counts[l]
// (prev, eid) pairs.
int oracle_support(const int* offsets, const int* vlabels, const int* edges3, int n, int nnz,
                   const int* code5, int m, const int* const* layers, const long long* counts,
                   int mni_mode, int nthreads) {
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    CSR g = make_csr(offsets, vlabels, edges3, n, nnz);
    DFSCode code = decode(code5, m);
    std::vector<Rels> store(m);
    Layers L(m);
    for (int l = 0; l < m; ++l) {
        store[l].resize(counts[l]);
        std::memcpy(store[l].data(), layers[l], sizeof(Rel) * counts[l]);
        L[l] = &store[l];
    }
    Bitmaps bm;
    return support(g, code, L, mni_mode, nthreads, bm);
}

// One extension step (gspan.cu:674-780) of one pattern.  Children are delivered in the
// reference's recursion order (backward ascending (v_i, el); forward descending v_i,
// ascending el, tl) through child_cb(child_edge5, rows, count, user).  Returns the total
// number of child rows, or -1 on error.
typedef void (*child_cb)(const int* edge5, const int* rows, long long count, void* user);

long long oracle_extend(const int* offsets, const int* vlabels, const int* edges3, int n,
                        int nnz, const int* code5, int m, const int* const* layers,
                        const long long* counts, int nthreads, child_cb cb, void* user) {
    try {
        if (nthreads <= 0) nthreads = omp_get_max_threads();
        CSR g = make_csr(offsets, vlabels, edges3, n, nnz);
        DFSCode code = decode(code5, m);
        std::vector<Rels> store(m);
        Layers L(m);
        for (int l = 0; l < m; ++l) {
            store[l].resize(counts[l]);
            std::memcpy(store[l].data(), layers[l], sizeof(Rel) * counts[l]);
            L[l] = &store[l];
        }
        std::vector<int> rm = right_most_path(code);
        int v_n = code[rm[0]].to;
        ForwardMap fwd;
        BackwardMap bwd;
        long long total = extend_omp(g, code, L, rm, fwd, bwd, nthreads);
        for (auto& a : bwd)
            for (auto& b : a.second) {
                int e5[5] = {v_n, a.first, -1, b.first, -1};
                cb(e5, reinterpret_cast<const int*>(b.second.data()), (long long)b.second.size(),
                   user);
            }
        for (auto a = fwd.rbegin(); a != fwd.rend(); ++a)
            for (auto& b : a->second)
                for (auto& c : b.second) {
                    int e5[5] = {a->first, v_n + 1, -1, b.first, c.first};
                    cb(e5, reinterpret_cast<const int*>(c.second.data()),
                       (long long)c.second.size(), user);
                }
        return total;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "oracle_extend: %s\n", e.what());
        return -1;
    }
}

// Single-edge seeding (gspan.cu:410-426, 1176-1180): codes in InitialEdgeCompare order,
// each with its rows (prev = -1, eid = CSR index) in CSR order.
typedef void (*seed_cb)(const int* edge5, const int* rows, long long count, void* user);

int oracle_seed(const int* offsets, const int* vlabels, const int* edges3, int n, int nnz,
                seed_cb cb, void* user) {
    CSR g = make_csr(offsets, vlabels, edges3, n, nnz);
    std::map<PatternEdge, Rels, InitialLess> init;
    for (int from = 0; from < g.n; ++from) {
        int fl = g.vlabels[from];
        for (int i = g.offsets[from]; i < g.offsets[from + 1]; ++i) {
            int to = g.edges[i].to, el = g.edges[i].label, tl = g.vlabels[to];
            if (fl <= tl) init[PatternEdge{0, 1, fl, el, tl}].push_back(Rel{-1, i});
        }
    }
    for (auto& p : init) {
        int e5[5] = {0, 1, p.first.from_label, p.first.edge_label, p.first.to_label};
        cb(e5, reinterpret_cast<const int*>(p.second.data()), (long long)p.second.size(), user);
    }
    return 0;
}

// Pattern graph of a code (graph.cc:122-159).  Writes n, nnz into dims[0..1] and, when
// the output arrays are non-null, the CSR (offsets n+1, vlabels n, edges 3*nnz).
int oracle_pattern_graph(const int* code5, int m, int* dims, int* offsets, int* vlabels,
                         int* edges3) {
    PatternGraph p = pattern_graph(decode(code5, m));
    dims[0] = p.n;
    dims[1] = p.nnz;
    if (offsets) std::memcpy(offsets, p.offsets.data(), sizeof(int) * (p.n + 1));
    if (vlabels) std::memcpy(vlabels, p.vlabels.data(), sizeof(int) * p.n);
    if (edges3) std::memcpy(edges3, p.edges.data(), sizeof(Edge) * p.nnz);
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------- CLI
// oracle_fsm_cli <graph.bin> <min_support> <k> [omp-literal] : prints the reference's
// pattern lines (gspan.cu:67-77) and its timing line (fsm.cu:40).
#ifdef ORACLE_MAIN
#include <chrono>
static int print_cb(const int* c, int m, int sup, void*) {
    for (int i = 0; i < m; ++i)
        std::printf("(%d %d %d %d %d)", c[5 * i], c[5 * i + 1], c[5 * i + 2], c[5 * i + 3],
                    c[5 * i + 4]);
    std::printf(": %d\n", sup);
    return 0;
}
int main(int argc, char** argv) {
    if (argc < 4) {
        std::printf("usage: %s input_graph min_support k [omp-literal]\n", argv[0]);
        return 0;
    }
    auto t1 = std::chrono::steady_clock::now();
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) {
        std::fprintf(stderr, "failed to open file\n");
        return 1;
    }
    int n, nnz;
    if (std::fread(&n, 4, 1, f) != 1 || std::fread(&nnz, 4, 1, f) != 1) return 1;
    std::vector<int> off(n + 1), vl(n), ed(3 * (size_t)nnz);
    if (std::fread(off.data(), 4, n + 1, f) != (size_t)n + 1) return 1;
    if (std::fread(vl.data(), 4, n, f) != (size_t)n) return 1;
    if (std::fread(ed.data(), 12, nnz, f) != (size_t)nnz) return 1;
    std::fclose(f);
    auto t2 = std::chrono::steady_clock::now();
    int mode = (argc > 4 && std::string(argv[4]) == "omp-literal") ? 1 : 0;
    long long st[8];
    oracle_fsm(off.data(), vl.data(), ed.data(), n, nnz, std::atoi(argv[3]), std::atoi(argv[2]),
               mode, 0, print_cb, nullptr, st);
    auto t3 = std::chrono::steady_clock::now();
    using us = std::chrono::microseconds;
    std::printf("load: %lfs prepare: %lfs fsm: %lfs\n",
                std::chrono::duration_cast<us>(t2 - t1).count() * 1e-6, 0.0,
                std::chrono::duration_cast<us>(t3 - t2).count() * 1e-6);
    std::fprintf(stderr, "extended=%lld candidates=%lld canonical=%lld reported=%lld\n", st[0],
                 st[1], st[2], st[3]);
    return 0;
}
#endif
