// gs_kernels.hip -- gfx950 kernels of the gSpan hot path.
//
//   seed_*        single-edge codes (gspan.cu:410-426, 1176-1180)
//   prep          per parent row: rebuild the vertex images from the layered rows
//                 (gspan.cu:385-394) and size its extension work
//   items_count   rightmost-path extension, pass 1: every (row, candidate entry) pair is
//                 one work item; emitted children are counted per child key
//                 (gspan.cu:694-735, rules gspan.cu:428-483)
//   items_fill    pass 2: children of canonical keys are written grouped by key and their
//                 vertex images are inserted into the MNI sets (gspan.cu:165-190)
//   canon         minimum-DFS-code check, one wave per candidate (gspan.cu:485-606)
//   set_counts    distinct images per (key, pattern vertex)
//
// Work distribution: rows are flattened into items by an exclusive scan of per-row
// work; persistent workgroups steal fixed tiles of items from an atomic counter and
// locate their rows by a search over the scanned offsets staged in LDS.  Output slots
// come from one atomic per (tile, distinct key): keys are aggregated per tile in LDS.
#include <climits>
#include <type_traits>

#include "gs_internal.h"

namespace {

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ unsigned mbcnt64(unsigned long long m) {
    return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

__device__ __forceinline__ uint32_t hash32(uint32_t x) {
    // multiplicative (Fibonacci) hashing followed by a xor-shift to mix the high bits
    x *= 0x9E3779B1u;
    return x ^ (x >> 15);
}

// A row context in registers: eight-lane integer vectors, indexed by compare-and-select
// (a per-lane index into an array would put the array in scratch memory).
typedef int int8v __attribute__((ext_vector_type(8)));

struct RowRegs {
    int8v phi, fcum, base;  // fcum[7] = base[7] = 0
    int p;
};

__device__ __forceinline__ RowRegs load_row(const RowCtx* c) {
    const int4* q = reinterpret_cast<const int4*>(c);
    const int4 a = q[0], b = q[1], d = q[2], e = q[3], f = q[4], h = q[5];
    RowRegs r;
    r.phi = int8v{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    r.fcum = int8v{d.x, d.y, d.z, d.w, e.x, e.y, e.z, 0};
    r.p = e.w;
    r.base = int8v{f.x, f.y, f.z, f.w, h.x, h.y, h.z, 0};
    return r;
}

__device__ __forceinline__ int pick(const int8v& v, int i) {
    int r = v[0];
#pragma unroll
    for (int k = 1; k < 8; ++k) r = (i == k) ? v[k] : r;
    return r;
}

template <int N>
__device__ __forceinline__ int pick(const int (&a)[N], int i) {
    int r = a[0];
#pragma unroll
    for (int k = 1; k < N; ++k) r = (i == k) ? a[k] : r;
    return r;
}

// The same over the first N entries only (i < N): a kernel specialised for patterns of at
// most PM edges selects among PM + 1 vertices or segments instead of all eight.
template <int N>
__device__ __forceinline__ int pickv(const int8v& v, int i) {
    static_assert(N >= 1 && N <= 8, "int8v entries");
    int r = v[0];
#pragma unroll
    for (int k = 1; k < N; ++k) r = (i == k) ? v[k] : r;
    return r;
}

template <int N, int M>
__device__ __forceinline__ int picka(const int (&a)[M], int i) {
    static_assert(N >= 1 && N <= M, "array entries");
    int r = a[0];
#pragma unroll
    for (int k = 1; k < N; ++k) r = (i == k) ? a[k] : r;
    return r;
}

// A set descriptor in registers (four 16-byte loads); its offsets are picked by
// compare-and-select, like a row context's arrays.
struct SetRegs {
    long long off;
    int nv, kinds;
    int8v vo;  // vo[0 .. 7]
    int vo8;   // vo[8]
    int tag;   // the byte a bytemap insert stores
};

__device__ __forceinline__ SetRegs load_set(const SetDesc* p) {
    static_assert(GS_MAXV == 8, "SetRegs holds 9 offsets");
    const int4* q = reinterpret_cast<const int4*>(p);
    const int4 a = q[0], b = q[1], c = q[2], e = q[3];
    SetRegs r;
    r.off = (long long)(((unsigned long long)(unsigned)a.y << 32) | (unsigned)a.x);
    r.nv = a.z;
    r.kinds = a.w;
    r.vo = int8v{b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w};
    r.vo8 = e.x;
    r.tag = e.y;
    return r;
}

// d.vo[i] for i <= NV (a descriptor of at most NV vertex sets)
template <int NV>
__device__ __forceinline__ int vo_at(const SetRegs& d, int i) {
    static_assert(NV >= 1 && NV <= GS_MAXV, "vertex sets");
    if constexpr (NV == GS_MAXV) return i >= GS_MAXV ? d.vo8 : pickv<GS_MAXV>(d.vo, i);
    else return pickv<NV + 1>(d.vo, i);
}

// Descriptor d addresses words [d.off, d.off + d.vo[d.nv]) of a pool of pool_words words,
// for a child of at least `need` and at most NV vertices.
template <int NV>
__device__ __forceinline__ bool set_fits(const SetRegs& d, int need, long long pool_words) {
    if (d.nv < need || d.nv > NV) return false;
    const int all = vo_at<NV>(d, d.nv);
    return d.off >= 0 && all >= 0 && d.off + all <= pool_words;
}

// Inserts the vertex of label-class rank x into vertex set v of descriptor d (x indexes
// the vertices of v's label).  Bits / slots only go from empty to full, so a stale plain
// read can only send us to the atomic, never skip a new key.  A table sized from exact
// counts (>= 2x the rows) never fills; one sized from predicted counts can, and then raises
// *overflow after probing every slot once.  An x past a bitmap or bytemap sets the guard
// bit GS_GUARD_POOL in *gerr and is skipped.
// NV: the descriptor holds at most NV vertex sets (v < NV).
template <int NV>
__device__ __forceinline__ void set_insert(const SetRegs& d, uint32_t* pool, int v, uint32_t x,
                                           int* gerr, int* overflow = nullptr) {
    const int kind = (d.kinds >> (2 * v)) & 3;
    const int v0 = vo_at<NV>(d, v);
    uint32_t* base = pool + d.off + v0;
    const uint32_t words = (uint32_t)(vo_at<NV>(d, v + 1) - v0);
    if (kind == GS_SET_BYTEMAP) {
        if (x < 4u * words) reinterpret_cast<unsigned char*>(base)[x] = (unsigned char)d.tag;
        else atomicOr(gerr, GS_GUARD_POOL);
    } else if (kind == GS_SET_BITMAP) {
        if ((x >> 5) < words) {
            uint32_t* w = base + (x >> 5);
            const uint32_t bit = 1u << (x & 31);
            if (!(*w & bit)) atomicOr(w, bit);
        } else {
            atomicOr(gerr, GS_GUARD_POOL);
        }
    } else {
        const uint32_t mask = words - 1;  // a power of two
        const uint32_t key = x + 1;       // 0 marks an empty slot
        uint32_t h = hash32(x) & mask;
        for (uint32_t probe = 0; probe < words; ++probe) {
            uint32_t cur = __hip_atomic_load(base + h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (cur == key) return;
            if (cur == 0) {
                uint32_t prev = atomicCAS(base + h, 0u, key);
                if (prev == 0 || prev == key) return;
            }
            h = (h + 1) & mask;
        }
        if (overflow) atomicOr(overflow, 1);
    }
}

// Element i of a pattern table: when the pattern is the same for the whole wave (UNI)
// its tables are wave-uniform, read with scalar loads and selected per lane; otherwise
// one vector load.
// (i < N)
template <bool UNI, int N, int M>
__device__ __forceinline__ int pget(const int (&a)[M], int i) {
    if (UNI) return picka<N>(a, i);
    return a[i];
}

// has_edge (gspan.cu:396-401) on a row given by its vertex images: some pattern edge
// with label el joins u and w (a pattern of at most PM edges).
template <bool UNI, int PM>
__device__ __forceinline__ bool row_has_edge(const PatDev& P, const RowRegs& c, int u, int w,
                                             int el) {
    bool hit = false;
#pragma unroll
    for (int j = 0; j < PM; ++j) {
        if (j < P.m && P.cel[j] == el) {
            const int a = pickv<PM + 1>(c.phi, P.cfrom[j]), b = pickv<PM + 1>(c.phi, P.cto[j]);
            hit |= (a == u && b == w) || (a == w && b == u);
        }
    }
    return hit;
}

struct ItemOut {
    bool valid;
    int key;
    int eid;
    int newv;  // label-class rank of the new vertex's image (forward; NEWX) or -1
    int nv;    // vertices of the parent pattern
    int w;     // children the item stands for (1, or a run's children)
    int sv;    // pattern vertex the forward segment starts at (-1: backward)
    long long pbase;  // parent row in its level = pbase + (row in the batch)
};

__device__ __forceinline__ bool in_range(long long x, long long n) {
    return (unsigned long long)x < (unsigned long long)n;
}

__device__ __forceinline__ ItemOut no_item(const PatDev& P) {
    return ItemOut{false, 0, 0, -1, P.nv, 1, -1, P.row_begin - P.brow_begin};
}

// Backward probe q < r-1 of a parent row: rmpath index q+1 (gspan.cu:428-445).  c is the
// row's context in registers.  Vertex ids are checked against the graph before they are
// used as indices (GS_GUARD_GRAPH).
template <bool UNI, int PM>
__device__ __forceinline__ ItemOut eval_back(const DevGraph& g, const PatDev& P, const RowRegs& c,
                                             int q, KeySpace ks, int* gerr) {
    constexpr int NV = PM + 1;  // parent vertices; rightmost path <= PM edges
    ItemOut o = no_item(P);
    const int i = q + 1;
    const int u_n = pickv<NV>(c.phi, P.v_n);
    const int target = pickv<NV>(c.phi, pget<UNI, PM>(P.bw_v, i));
    const int e_el = pget<UNI, PM>(P.bw_el, i);
    const bool le = pget<UNI, PM>(P.bw_le, i) != 0;
    if (!in_range(u_n, g.n)) {
        atomicOr(gerr, GS_GUARD_GRAPH);
        return o;
    }
    int lo = g.offsets[u_n];  // all neighbours of u_n, ordered by (to, label)
    const int end = g.offsets[u_n + 1];
    int hi = end;
    while (lo < hi) {  // first entry with to >= target; adjacency sorted by (to, label)
        const int mid = (lo + hi) >> 1;
        if ((int)g.adj[mid].x < target) lo = mid + 1;
        else hi = mid;
    }
    for (int e = lo; e < end; ++e) {
        const uint2 a = g.adj[e];
        if ((int)a.x != target) break;
        const int el = (int)(a.y >> 16);
        if (row_has_edge<UNI, PM>(P, c, u_n, target, el)) continue;
        if (e_el < el || (e_el == el && le)) {
            o.valid = true;
            o.eid = e;
            o.key = P.key_base + (i - 1) * ks.EL + el;
            break;
        }
    }
    return o;
}

// Segment of forward item (or run) f: fcum is non-decreasing.
template <int PM>
__device__ __forceinline__ int segment_of(const RowRegs& c, int r, int f) {
    int j = 0;
#pragma unroll
    for (int s = 0; s < PM; ++s) j += (s < r && f >= c.fcum[s]) ? 1 : 0;
    return j;
}

// Forward item f of a parent row: one neighbour of u_n (first forward, gspan.cu:447-462) or
// of a rightmost-path source (other forward, gspan.cu:464-483).  NEWX: a child also returns
// the label-class rank of its new vertex (for its MNI set).
template <bool UNI, bool NEWX, int PM>
__device__ __forceinline__ ItemOut eval_fwd(const DevGraph& g, const PatDev& P, const RowRegs& c,
                                            int f, KeySpace ks, int* gerr) {
    constexpr int NV = PM + 1;
    ItemOut o = no_item(P);
    const int r = P.r;
    if (f >= pickv<NV>(c.fcum, r)) return o;  // the empty item of a row with no forward item
    const int j = segment_of<PM>(c, r, f);
    const int el_idx = pickv<NV>(c.base, j) + f;  // in adjl: neighbours by label from min_label
    if (!in_range(el_idx, g.nnz)) {
        atomicOr(gerr, GS_GUARD_GRAPH);
        return o;
    }
    const uint2 a = g.adjl[el_idx];
    const int to = (int)a.x, el = (int)(a.y >> 16), tl = (int)(a.y & 0xffffu);
    if (tl < P.min_label) return o;
    bool seen = false;
#pragma unroll
    for (int v = 0; v < NV; ++v) seen |= v < P.nv && c.phi[v] == to;
    if (seen) return o;
    const int e_el = pget<UNI, NV>(P.seg_el, j);
    if (!(e_el < el || (e_el == el && pget<UNI, NV>(P.seg_tl, j) <= tl))) return o;
    o.valid = true;
    o.eid = g.eidl[el_idx];  // the entry's CSR index
    o.newv = NEWX ? g.adjlx[el_idx] : 0;
    o.sv = pget<UNI, NV>(P.seg_v, j);
    o.key = P.key_base + (r - 1) * ks.EL + (j * ks.EL + el) * ks.VL + tl;
    return o;
}

// Forward run f of a parent row: the neighbours of one segment vertex u with one (label,
// edge label), all children of one key.  The run passes or fails the label and rule tests
// as a whole (gspan.cu:455, 475-477); of its entries, those whose neighbour is one of the
// row's own vertices are not children (gspan.cu:453, 473).  Only the row's vertices of the
// run's label can be among them, each counted once.  In a simple graph u has no self-loop,
// and a row vertex joined to u by a pattern edge is in the run exactly when that edge's
// label is the run's; other row vertices are looked up by binary search, as a run ascends
// by neighbour.  o.w = the run's children.
template <bool UNI, int PM>
__device__ __forceinline__ ItemOut eval_run(const DevGraph& g, const PatDev& P, const RowRegs& c,
                                            int f, KeySpace ks, int* gerr) {
    constexpr int NV = PM + 1;
    ItemOut o = no_item(P);
    const int r = P.r;
    if (f >= pickv<NV>(c.fcum, r)) return o;  // the empty item of a row with no forward run
    const int j = segment_of<PM>(c, r, f);
    const int ri = pickv<NV>(c.base, j) + f;
    if (!in_range(ri, g.nruns)) {
        atomicOr(gerr, GS_GUARD_GRAPH);
        return o;
    }
    const int rk = g.rkey[ri];
    const int tl = rk >> 16, el = rk & 0xffff;
    if (tl < P.min_label) return o;
    const int e_el = pget<UNI, NV>(P.seg_el, j);
    if (!(e_el < el || (e_el == el && pget<UNI, NV>(P.seg_tl, j) <= tl))) return o;
    const int st = g.rstart[ri], en = g.rstart[ri + 1];
    if (st < 0 || en > g.nnz || st >= en) {
        atomicOr(gerr, GS_GUARD_GRAPH);
        return o;
    }
    int excl = 0;
    const int sv = pget<UNI, NV>(P.seg_v, j);
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        if (v >= P.nv || pget<UNI, NV>(P.vlab, v) != tl) continue;
        if (g.simple) {
            if (v == sv) continue;
            int pel = -1;  // label of the pattern edge joining sv and v, if any
#pragma unroll
            for (int k = 0; k < PM; ++k) {
                if (k >= P.m) continue;
                const int a = pget<UNI, PM>(P.cfrom, k), b = pget<UNI, PM>(P.cto, k);
                if ((a == sv && b == v) || (a == v && b == sv)) pel = pget<UNI, PM>(P.cel, k);
            }
            if (pel >= 0) {
                excl += pel == el ? 1 : 0;
                continue;
            }
        }
        const int x = c.phi[v];
        // a self-loop row maps two pattern vertices to one graph vertex: its entries are
        // excluded once (gspan.cu:453, 473 test the neighbour against the row's vertex set)
        bool again = false;
#pragma unroll
        for (int w = 0; w < v; ++w) again |= c.phi[w] == x;
        if (again) continue;
        int lo = st, hi = en;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if ((int)g.adjl[mid].x < x) lo = mid + 1;
            else hi = mid;
        }
        while (lo < en && (int)g.adjl[lo].x == x) {
            ++excl;
            ++lo;
        }
    }
    if (en - st - excl <= 0) return o;
    o.valid = true;
    o.w = en - st - excl;
    o.sv = sv;
    o.key = P.key_base + (r - 1) * ks.EL + (j * ks.EL + el) * ks.VL + tl;
    return o;
}

// ------------------------------------------------------------------ seeds
// Every CSR entry with label(from) <= label(to) is a row of the single-edge code
// (label(from), edge label, label(to)) (gspan.cu:410-426).  Tiles of GS_TILE entries,
// keys aggregated per tile in LDS as in k_items.

constexpr int SEED_SLOTS = 2 * GS_TILE;

__device__ __forceinline__ int seed_lds_slot(int* s_key, int key) {
    unsigned h = ((unsigned)key * 0x9E3779B1u) >> 21;
    for (;;) {
        const int prev = atomicCAS(&s_key[h], -1, key);
        if (prev == -1 || prev == key) return (int)h;
        h = (h + 1) & (SEED_SLOTS - 1);
    }
}

template <bool FILL>
__global__ __launch_bounds__(GS_BLOCK) void k_seed(DevGraph g, int EL, int VL, int* key_count,
                                                   long long* total, const int* key_slot,
                                                   int* cursor, int2* rows_out,
                                                   const SetDesc* sets, uint32_t* pool,
                                                   ItemGuard gd) {
    constexpr int IPT = GS_TILE / GS_BLOCK;
    __shared__ int s_key[SEED_SLOTS];
    __shared__ int s_cnt[SEED_SLOTS];
    __shared__ int s_base[FILL ? SEED_SLOTS : 1];
    for (int h = threadIdx.x; h < SEED_SLOTS; h += GS_BLOCK) {
        s_key[h] = -1;
        s_cnt[h] = 0;
    }
    __syncthreads();
    long long local = 0;
    for (long long t0 = (long long)blockIdx.x * GS_TILE; t0 < g.nnz;
         t0 += (long long)gridDim.x * GS_TILE) {
        int hs[IPT], rank[IPT];
#pragma unroll
        for (int q = 0; q < IPT; ++q) {
            const long long i = t0 + (long long)q * GS_BLOCK + threadIdx.x;
            hs[q] = -1;
            rank[q] = 0;
            if (i >= g.nnz) continue;
            const int fl = g.vlab[g.src[i]];
            const uint2 a = g.adj[i];
            const int tl = (int)(a.y & 0xffffu), el = (int)(a.y >> 16);
            if (fl > tl) continue;  // gspan.cu:421
            const int key = (fl * EL + el) * VL + tl;
            const int agg = FILL ? key_slot[key] : key;
            if (agg < 0) continue;
            if (FILL && !in_range(agg, gd.slots)) {
                atomicOr(gd.err, GS_GUARD_SLOT);
                continue;
            }
            ++local;
            hs[q] = seed_lds_slot(s_key, agg);
            rank[q] = atomicAdd(&s_cnt[hs[q]], 1);
        }
        __syncthreads();
        for (int h = threadIdx.x; h < SEED_SLOTS; h += GS_BLOCK) {
            const int key = s_key[h];
            if (key < 0) continue;
            if (FILL) {
                if (rows_out) s_base[h] = atomicAdd(&cursor[key], s_cnt[h]);
            } else {
                atomicAdd(&key_count[key], s_cnt[h]);
            }
        }
        __syncthreads();
        if (FILL) {
#pragma unroll
            for (int q = 0; q < IPT; ++q) {
                if (hs[q] < 0) continue;
                const long long i = t0 + (long long)q * GS_BLOCK + threadIdx.x;
                const int slot = s_key[hs[q]];
                if (rows_out) {
                    const long long at = (long long)s_base[hs[q]] + rank[q];
                    if (in_range(at, gd.rows_cap)) rows_out[at] = make_int2(-1, (int)i);
                    else atomicOr(gd.err, GS_GUARD_ROW);
                }
                const SetRegs d = load_set(sets + slot);
                if (!set_fits<2>(d, 2, gd.pool_words)) {
                    atomicOr(gd.err, GS_GUARD_POOL);
                    continue;
                }
                set_insert<2>(d, pool, 0, (uint32_t)g.lidx[g.src[i]], gd.err);
                set_insert<2>(d, pool, 1, (uint32_t)g.lidx[g.adj[i].x], gd.err);
            }
        }
        __syncthreads();
        for (int h = threadIdx.x; h < SEED_SLOTS; h += GS_BLOCK) {
            if (s_key[h] >= 0) {
                s_key[h] = -1;
                s_cnt[h] = 0;
            }
        }
        __syncthreads();
    }
    if (!FILL) {  // one atomic per workgroup on the total
        __shared__ long long s_tot[GS_BLOCK / 64];
        for (int off = 32; off > 0; off >>= 1) local += __shfl_xor(local, off);
        if (lane_id() == 0) s_tot[threadIdx.x / 64] = local;
        __syncthreads();
        if (threadIdx.x == 0) {
            long long t = 0;
            for (int w = 0; w < GS_BLOCK / 64; ++w) t += s_tot[w];
            if (t) atomicAdd((unsigned long long*)total, (unsigned long long)t);
        }
    }
}

// ------------------------------------------------------------------ prep

// One parent row: images phi, segment bases and item counts (RowCtx), work items; returns
// the row's items without the label filter.
__device__ __forceinline__ unsigned long long prep_row(const DevGraph& g,
                                                       const PatDev* __restrict__ pats, int npats,
                                                       long long b, const LayerPtrs& L, bool runs,
                                                       RowCtx* __restrict__ out_ctx,
                                                       int* __restrict__ out_work, int* gerr) {
    int lo = 0, hi = npats - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (pats[mid].brow_begin <= b) lo = mid;
        else hi = mid - 1;
    }
    const PatDev& P = pats[lo];
    int8v phi = int8v{-1, -1, -1, -1, -1, -1, -1, -1};
    long long idx = P.row_begin + (b - P.brow_begin);
    bool bad = false;
#pragma unroll
    for (int l = GS_MAXE - 1; l >= 0; --l) {
        if (l >= P.m || bad) continue;
        if (!in_range(idx, L.rows[l])) {
            atomicOr(gerr, GS_GUARD_LAYER);
            bad = true;
            continue;
        }
        const int2 rr = L.layer[l][idx];
        if (!in_range(rr.y, g.nnz)) {
            atomicOr(gerr, GS_GUARD_GRAPH);
            bad = true;
            continue;
        }
        if (P.cfrom[l] < P.cto[l]) {
            const int to = (int)g.adj[rr.y].x, v = P.cto[l];
#pragma unroll
            for (int k = 0; k < GS_MAXV; ++k) phi[k] = (v == k) ? to : phi[k];
        }
        if (l == 0) phi[0] = g.src[rr.y];
        idx = rr.x;
    }
    // a row that failed a guard gets images of vertex 0: its items stay inside the graph
    // and the guard word fails the call
    if (bad) phi = int8v{0, 0, 0, 0, 0, 0, 0, 0};
    int8v fcum = int8v{0, 0, 0, 0, 0, 0, 0, 0}, base = fcum;
    int fc = 0, deg = 0;
#pragma unroll
    for (int j = 0; j < GS_MAXE; ++j) {
        int bj = 0;
        if (j <= P.r) {
            int u = pick(phi, P.seg_v[j]);
            if (!in_range(u, g.n)) {
                atomicOr(gerr, GS_GUARD_GRAPH);
                u = 0;
            }
            const int o0 = g.offsets[u], o1 = g.offsets[u + 1];
            int first;  // first neighbour (by label) with label >= min_label
            if (g.lstart) {
                first = g.lstart[(long long)u * g.lstride + P.min_label];
            } else {
                first = o0;
                int hi = o1;
                while (first < hi) {
                    const int mid = (first + hi) >> 1;
                    if ((int)(g.adjl[mid].y & 0xffffu) < P.min_label) first = mid + 1;
                    else hi = mid;
                }
            }
            if (runs) {  // the runs from `first` on: runid ascends along adjl
                const int rf = first < o1 ? g.runid[first] : 0;
                const int rc = first < o1 ? g.runid[o1 - 1] - rf + 1 : 0;
                bj = rf - fc;
                fc += rc;
            } else {
                bj = first - fc;
                fc += o1 - first;
            }
            deg += o1 - o0;
        }
        base[j] = bj;
        fcum[j] = fc;
    }
    int4* out = reinterpret_cast<int4*>(out_ctx);
    out[0] = make_int4(phi[0], phi[1], phi[2], phi[3]);
    out[1] = make_int4(phi[4], phi[5], phi[6], phi[7]);
    out[2] = make_int4(fcum[0], fcum[1], fcum[2], fcum[3]);
    out[3] = make_int4(fcum[4], fcum[5], fcum[6], lo);
    out[4] = make_int4(base[0], base[1], base[2], base[3]);
    out[5] = make_int4(base[4], base[5], base[6], 0);
    // at least one item per row (an empty one when every neighbour is filtered out), so
    // that a tile of T items spans at most T rows
    *out_work = max(1, (P.r - 1) + fc);
    return (unsigned long long)((P.r - 1) + deg);
}

// partial[blockIdx] (nullable) = the block's items without the label filter; summed by
// k_sum_partials (atomics on one counter from every block would serialise)
__global__ __launch_bounds__(GS_BLOCK) void k_prep(DevGraph g, const PatDev* __restrict__ pats,
                                                   int npats, long long b0, long long R, LayerPtrs L,
                                                   int runs, RowCtx* __restrict__ ctx,
                                                   int* __restrict__ work,
                                                   unsigned long long* __restrict__ partial,
                                                   int* gerr) {
    __shared__ unsigned long long s_w[GS_BLOCK / 64];
    const long long lb = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long full = 0;
    if (lb < R) full = prep_row(g, pats, npats, b0 + lb, L, runs != 0, ctx + lb, work + lb, gerr);
    if (partial) {
        for (int off = 32; off > 0; off >>= 1) full += __shfl_xor(full, off);
        if (lane_id() == 0) s_w[threadIdx.x / 64] = full;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long t = 0;
            for (int w = 0; w < GS_BLOCK / 64; ++w) t += s_w[w];
            partial[blockIdx.x] = t;
        }
    }
}

// *total += sum of partial[0 .. n)
__global__ __launch_bounds__(GS_BLOCK) void k_sum_partials(const unsigned long long* partial,
                                                           long long n,
                                                           unsigned long long* total) {
    __shared__ unsigned long long s_w[GS_BLOCK / 64];
    unsigned long long t = 0;
    for (long long i = threadIdx.x; i < n; i += GS_BLOCK) t += partial[i];
    for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off);
    if (lane_id() == 0) s_w[threadIdx.x / 64] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long s = 0;
        for (int w = 0; w < GS_BLOCK / 64; ++w) s += s_w[w];
        *total += s;
    }
}

// ------------------------------------------------------------------ scan (int32 -> int64)

constexpr int SCAN_PER_THREAD = 16;
constexpr int SCAN_TILE = GS_BLOCK * SCAN_PER_THREAD;

__device__ long long block_exclusive_scan(long long v, long long* total) {
    __shared__ long long s_wave[GS_BLOCK / 64];
    const int lane = lane_id(), wid = threadIdx.x / 64;
    long long x = v;
    for (int off = 1; off < 64; off <<= 1) {
        long long y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == 63) s_wave[wid] = x;
    __syncthreads();
    long long wbase = 0, tot = 0;
    for (int w = 0; w < GS_BLOCK / 64; ++w) {
        if (w < wid) wbase += s_wave[w];
        tot += s_wave[w];
    }
    __syncthreads();
    *total = tot;
    return wbase + x - v;
}

__global__ __launch_bounds__(GS_BLOCK) void k_scan_reduce(const int* in, long long n,
                                                          long long* partial) {
    long long base = (long long)blockIdx.x * SCAN_TILE + (long long)threadIdx.x * SCAN_PER_THREAD;
    long long s = 0;
    for (int i = 0; i < SCAN_PER_THREAD; ++i)
        if (base + i < n) s += in[base + i];
    long long tot;
    block_exclusive_scan(s, &tot);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

__global__ __launch_bounds__(GS_BLOCK) void k_scan_top(long long* partial, long long nb) {
    long long carry = 0;
    for (long long base = 0; base < nb; base += GS_BLOCK) {
        long long i = base + threadIdx.x;
        long long v = i < nb ? partial[i] : 0;
        long long tot;
        long long ex = block_exclusive_scan(v, &tot);
        if (i < nb) partial[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) partial[nb] = carry;
}

__global__ __launch_bounds__(GS_BLOCK) void k_scan_down(const int* in, long long n,
                                                        const long long* partial, long long nb,
                                                        long long* out) {
    long long base = (long long)blockIdx.x * SCAN_TILE + (long long)threadIdx.x * SCAN_PER_THREAD;
    int vals[SCAN_PER_THREAD];
    long long s = 0;
    for (int i = 0; i < SCAN_PER_THREAD; ++i) {
        vals[i] = (base + i < n) ? in[base + i] : 0;
        s += vals[i];
    }
    long long tot;
    long long ex = block_exclusive_scan(s, &tot) + partial[blockIdx.x];
    for (int i = 0; i < SCAN_PER_THREAD; ++i) {
        if (base + i < n) out[base + i] = ex;
        ex += vals[i];
    }
    if (blockIdx.x == nb - 1 && threadIdx.x == 0) out[n] = partial[nb];
}

// ------------------------------------------------------------------ extension items

// Per-tile key aggregation in LDS: an open-addressing table (key -> items in this tile)
// so that a tile issues one global atomic per distinct child key, however hot the key.
// A table has 2*TILE slots, more than the distinct keys of a tile, so probing ends.
constexpr int log2i(int x) { return x <= 1 ? 0 : 1 + log2i(x / 2); }

template <int SLOTS>
__device__ __forceinline__ int lds_slot(int* s_key, int key) {
    unsigned h = ((unsigned)key * 0x9E3779B1u) >> (32 - log2i(SLOTS));
    for (;;) {
        const int prev = atomicCAS(&s_key[h], -1, key);
        if (prev == -1 || prev == key) return (int)h;
        h = (h + 1) & (SLOTS - 1);
    }
}

// Inserts key into an LDS set; true for the one caller that inserted it.
template <int SLOTS>
__device__ __forceinline__ bool lds_first(int* s_set, int key) {
    unsigned h = ((unsigned)key * 0x9E3779B1u) >> (32 - log2i(SLOTS));
    for (;;) {
        const int prev = atomicCAS(&s_set[h], -1, key);
        if (prev == -1) return true;
        if (prev == key) return false;
        h = (h + 1) & (SLOTS - 1);
    }
}

// Inserts key into the workgroup's persistent count table, counting new slots in *fill.
template <int SLOTS>
__device__ __forceinline__ int lds_slot_fill(int* s_key, int key, int* fill) {
    unsigned h = ((unsigned)key * 0x9E3779B1u) >> (32 - log2i(SLOTS));
    for (;;) {
        const int prev = atomicCAS(&s_key[h], -1, key);
        if (prev == -1) {
            atomicAdd(fill, 1);
            return (int)h;
        }
        if (prev == key) return (int)h;
        h = (h + 1) & (SLOTS - 1);
    }
}

// Key spaces up to this size are compacted by one workgroup (k_compact_small).
#ifndef GS_COMPACT_SMALL
#define GS_COMPACT_SMALL (1 << 17)
#endif

// Tiles after which an item kernel flushes its persistent count table (a smaller value can be
// given at build time to exercise the flush).
#ifndef GS_FLUSH_TILES
#define GS_FLUSH_TILES (1 << 14)
#endif

// Adds c > 0 to a child count (used when the launches adding to the counts may process
// more than INT_MAX items; otherwise no count can pass INT_MAX and a plain add is used, as
// the returned value costs a wait).  A count that would pass INT_MAX becomes INT_MIN and stays
// there: every add that sees a negative count (or crosses INT_MAX) stores INT_MIN after
// it, so the last write to the word is such a store whenever the true sum passes INT_MAX
// (a child count is the number of child embeddings; the host reads a negative count as
// "more than INT_MAX").  Below that the count is exact.
__device__ __forceinline__ void count_add(int* count, int c) {
    const int old = atomicAdd(count, c);
    if (old < 0 || old > INT_MAX - c) atomicExch(count, INT_MIN);
}

// Adds a workgroup's count c (64-bit in the run modes) to a child count: a c past INT_MAX
// saturates the count at once; otherwise as count_add (saturate) or a plain add.
template <class CNT>
__device__ __forceinline__ void count_flush(int* count, CNT c, bool saturate) {
    if ((unsigned long long)c > (unsigned long long)INT_MAX) atomicExch(count, INT_MIN);
    else if (saturate) count_add(count, (int)c);
    else atomicAdd(count, (int)c);
}

// First row whose scanned offset is <= item, found by wave 0 with 64-way probes of woff
// (four rounds of one load per lane for 16M rows).  woff is non-decreasing and
// woff[0] = 0 <= item, so the probes that pass form a prefix of the lanes.
__device__ __forceinline__ long long wave_find_row(const long long* __restrict__ woff,
                                                   long long R, long long item) {
    const int lane = lane_id();
    long long lo = 0, hi = R - 1;
    while (lo < hi) {  // uniform over the wave
        const long long step = (hi - lo + 64) / 64;
        const long long p = lo + lane * step;
        const bool ok = p <= hi && woff[p] <= item;
        const unsigned long long m = __ballot(ok);
        const int last = 63 - __clzll((long long)m);  // lane 0 always passes
        lo += last * step;
        hi = min(hi, lo + step - 1);
    }
    return lo;
}

// The same over a staged tile's offsets in LDS: last i in [0, n] with off[i] <= item.
__device__ __forceinline__ int wave_find_lds(const int* off, int n, int item) {
    const int lane = lane_id();
    int lo = 0, hi = n;
    while (lo < hi) {
        const int step = (hi - lo + 64) / 64;
        const int p = lo + lane * step;
        const bool ok = p <= hi && off[p] <= item;
        const unsigned long long m = __ballot(ok);
        const int last = 63 - __clzll((long long)m);
        lo += last * step;
        hi = min(hi, lo + step - 1);
    }
    return lo;
}

// Persistent extension kernel in one of four modes (GS_ITEMS_*).  Items of all parent rows
// (or groups, NEWV) are flattened by the scan of per-row work (woff); workgroups take
// TILE-item tiles from per-XCD queues.  Per tile:
//   phase 1  every thread evaluates its items (row by a search over the tile's staged row
//            offsets, row context, run or adjacency entry), aggregates their child keys in
//            an LDS table and inserts the children's vertex images into their MNI sets
//            right away (the set of a key is known once the key is): the parent's images
//            (SETS, ROWS), the new vertex's image (ROWS, NEWV); meanwhile wave 0 claims
//            the next tile and finds its first row
//   phase 2  one output reservation per distinct key (ROWS); child counts go to the
//            workgroup's persistent count table; the next tile's row offsets are loaded
//   phase 3  (ROWS) child rows at their reserved positions
// So a tile costs three or four barriers, and the claim, row search and offset loads of
// the next tile overlap the work of this one.  NEWV counts nothing: phase 2 only stages.
//
// Child counts (key_count) are summed per workgroup in a persistent LDS table that lives
// across the tiles the workgroup takes (ROWS adds its tile table into it in phase 2, the
// run modes add to it directly in phase 1), and added to global memory when it is more
// than half full, after 2^14 tiles and when the workgroup ends: the counters of a level's
// hot keys see one atomic per workgroup instead of one per tile (device-scope atomics are
// performed at the memory side and serialise on one address).  The run modes count in
// 64 bits (a run adds up to a vertex degree); ROWS adds at most TILE per tile, so its
// 32-bit LDS counts stay below 2^14 * TILE <= 2^24.
template <int MODE, int TILE, int PM>
__global__ __launch_bounds__(GS_BLOCK) void k_items(DevGraph g, const PatDev* __restrict__ pats,
                                                    const RowCtx* __restrict__ ctx,
                                                    const long long* __restrict__ woff,
                                                    long long b0, long long R, KeySpace ks,
                                                    int* key_count,
                                                    const int* __restrict__ key_slot,
                                                    unsigned long long* cursor,
                                                    const int* __restrict__ cap_end,
                                                    int2* rows_out, const SetDesc* __restrict__ sets,
                                                    uint32_t* pool, int* overflow,
                                                    int* tile_counters, int* next_counters,
                                                    ItemGuard gd, int dedup, int saturate) {
    constexpr bool RUNS = MODE == GS_ITEMS_COUNT || MODE == GS_ITEMS_SETS;
    constexpr bool FILL = MODE != GS_ITEMS_COUNT;  // inserts into MNI sets
    constexpr bool ROWS = MODE == GS_ITEMS_ROWS;
    constexpr bool NEWV = MODE == GS_ITEMS_NEWV;
    constexpr bool PARENT = MODE == GS_ITEMS_SETS || ROWS;  // inserts the parent's images
    constexpr bool NEWX = ROWS || NEWV;                      // inserts the new vertex's image
    constexpr bool AGG = !NEWV;                              // aggregates keys per tile
    using cnt_t = typename std::conditional<RUNS, unsigned long long, int>::type;
    constexpr int CV = PM + 2;  // vertices of a child (sets of a child key)
    static_assert(PM >= 1 && PM < GS_MAXE, "parent edges");
    constexpr int LDS_SLOTS = 2 * TILE;
    constexpr int AGG_SLOTS = AGG ? LDS_SLOTS : 1;
    // ROWS keeps a per-tile key table (its output reservations are per tile); the run modes
    // count straight into the persistent table, which leaves LDS for more resident waves
    constexpr bool TTAB = ROWS;
    constexpr int TT_SLOTS = TTAB ? LDS_SLOTS : 1;
    constexpr int IPT = TILE / GS_BLOCK;
    constexpr int STAGE = (TILE + 1 + GS_BLOCK - 1) / GS_BLOCK;  // offsets per thread
    static_assert(TILE % GS_BLOCK == 0 && (TILE & (TILE - 1)) == 0, "tile: power of two");
    static_assert(GS_TILE_QUEUE <= GS_BLOCK, "one thread per queue counter");
    // the queues of the next launch (the previous launch used them and has ended)
    if (blockIdx.x == 0 && threadIdx.x < GS_TILE_QUEUE) next_counters[threadIdx.x] = 0;
    // scanned offsets of the staged tiles' rows relative to the tile's first item, two
    // buffers: the current tile and the next (row 0 may start before the tile: <= 0)
    __shared__ int s_off[2][TILE + 1];
    __shared__ int s_key[TT_SLOTS];
    __shared__ cnt_t s_cnt[TT_SLOTS];
    __shared__ int s_base[ROWS ? LDS_SLOTS : 1];
    // (live slot, row in tile) pairs already seen (ROWS): only their first item inserts the
    // parent's vertex images, which are the same for every item of the pair (a run item is
    // the only one of its (key, row) pair)
    __shared__ int s_pair[ROWS ? LDS_SLOTS : 1];
    // persistent count table: LDS_SLOTS slots, flushed once more than TILE are used, so a
    // tile (at most TILE new keys) always finds room and probing ends
    __shared__ int s_pkey[AGG_SLOTS];
    __shared__ cnt_t s_pcnt[AGG_SLOTS];
    // s_pfill: used slots of the persistent table (low 16 bits, at most LDS_SLOTS) and
    // tiles since its last flush (high bits)
    __shared__ int s_pfill;
    __shared__ long long s_tile[2], s_rlo[2];
    __shared__ int s_nrows[2];
    constexpr int FLUSH_TILES = GS_FLUSH_TILES;
    static_assert(FLUSH_TILES >= 1 && FLUSH_TILES <= (1 << 14), "tile bits of s_pfill");
    if (R <= 0) return;
    if (AGG) {
        for (int h = threadIdx.x; h < LDS_SLOTS; h += GS_BLOCK) {
            if constexpr (TTAB) {
                s_key[h] = -1;
                s_cnt[h] = 0;
                s_pair[h] = -1;
            }
            s_pkey[h] = -1;
            s_pcnt[h] = 0;
        }
    }
    if (threadIdx.x == 0) s_pfill = 0;
    const long long W = woff[R];
    const long long ntiles = (W + TILE - 1) / TILE;
    // XCD-aware stealing: workgroups are dealt to the 8 XCDs round-robin, so workgroup b
    // first drains the b % 8-th contiguous eighth of the tiles (whose rows, patterns and
    // child sets then stay in that XCD's L2), then helps the other eighths.  Tiles are
    // claimed in runs of TB consecutive tiles, about eight claims per workgroup: the queue
    // counters are hot addresses, and the first row of a tile that follows the previous one
    // comes from that tile's staged offsets instead of a search over woff.
    const int x0 = blockIdx.x % GS_XCDS;
    const int TB = (int)max(1LL, min(16LL, ntiles / (8LL * gridDim.x)));
    const bool w0 = threadIdx.x < 64;  // wave 0 claims tiles and finds their first rows
    const int lane = lane_id();
    int xi = 0;                        // queues found empty (lane 0 of wave 0 only)
    long long nxt = 0, claim_end = 0;  // claimed, not yet handed out (lane 0 of wave 0)
    auto claim = [&]() -> long long {  // lane 0 of wave 0
        if (nxt < claim_end) return nxt++;
        while (xi < GS_XCDS) {
            const int x = (x0 + xi) % GS_XCDS;
            const long long beg = ntiles * x / GS_XCDS, end = ntiles * (x + 1) / GS_XCDS;
            const long long t = beg + (long long)atomicAdd(&tile_counters[x * GS_XCD_STRIDE], TB);
            if (t < end) {
                nxt = t + 1;
                claim_end = min(t + TB, end);
                return t;
            }
            ++xi;
        }
        return -1;
    };
    // wave 0: hands tile `tl` (or -1) to buffer b with its first row and row count
    auto publish = [&](int b, long long tl, long long rlo) {
        if (threadIdx.x == 0) {
            s_tile[b] = tl;
            s_rlo[b] = rlo;
            s_nrows[b] = tl < 0 ? 0 : (int)min((long long)TILE, R - rlo);
        }
    };
    // offsets of the rows of buffer b's tile, relative to its first item, clamped so that
    // they fit an int (only comparisons with item offsets below TILE use the large ones)
    auto rel = [&](long long v, long long t0) -> int {
        const long long d = v - t0;
        return (int)(d > (long long)INT_MAX ? (long long)INT_MAX : d);
    };

    // prologue: the first tile
    if (w0) {
        long long t = -1;
        if (threadIdx.x == 0) t = claim();
        t = __shfl(t, 0);
        publish(0, t, t < 0 ? 0 : wave_find_row(woff, R, t * TILE));
    }
    __syncthreads();
    {
        const long long t = s_tile[0];
        if (t >= 0) {
            const long long rlo = s_rlo[0];
            for (int i = threadIdx.x; i <= s_nrows[0]; i += GS_BLOCK)
                s_off[0][i] = rel(woff[rlo + i], t * TILE);
        }
    }
    __syncthreads();

    int buf = 0;
    for (;;) {
        const long long tile = s_tile[buf];
        if (tile < 0) break;
        const int nb = buf ^ 1;
        const long long rlo = s_rlo[buf];
        const int nrows = s_nrows[buf];
        const int* off = s_off[buf];
        const long long t0 = tile * TILE;
        const int tn = (int)min((long long)TILE, W - t0);  // items in this tile
        // wave 0 claims the next tile now; the claim's atomic overlaps the item loads
        long long ntile = -1;
        if (threadIdx.x == 0) ntile = claim();

        // phase 1: evaluate this thread's items, aggregate them per key in LDS, insert the
        // children's vertex images into their sets
        int hs[IPT], rank[IPT], eid[ROWS ? IPT : 1], prow[ROWS ? IPT : 1];
#pragma unroll
        for (int q = 0; q < IPT; ++q) {
            const int it = q * GS_BLOCK + threadIdx.x;  // item offset in the tile
            hs[q] = -1;
            rank[q] = 0;
            ItemOut o{false, 0, 0, -1, 0, 1, -1, 0};
            RowRegs c;
            int p = -1, f = 0, row = 0;
            if (it < tn) {
                int lo = 0, hi = nrows - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (off[mid] <= it) lo = mid;
                    else hi = mid - 1;
                }
                row = lo;
                c = load_row(ctx + rlo + lo);
                p = c.p;
                f = it - off[lo];
            }
            // rows of one pattern are contiguous, so a wave's items mostly share it: then
            // the pattern's tables are read once per wave with scalar loads
            const int p0 = __builtin_amdgcn_readfirstlane(p);
            auto eval = [&](const PatDev& P, auto uni) -> ItemOut {
                constexpr bool U = decltype(uni)::value;
                if constexpr (NEWV) {
                    return eval_fwd<U, true, PM>(g, P, c, f, ks, gd.err);
                } else {
                    if (f < P.r - 1) return eval_back<U, PM>(g, P, c, f, ks, gd.err);
                    if constexpr (RUNS) return eval_run<U, PM>(g, P, c, f - (P.r - 1), ks, gd.err);
                    else return eval_fwd<U, NEWX, PM>(g, P, c, f - (P.r - 1), ks, gd.err);
                }
            };
            if (__all(p == p0 || p < 0)) {
                if (it < tn) o = eval(pats[p0], std::true_type{});
            } else if (it < tn) {
                o = eval(pats[p], std::false_type{});
            }
            // aggregation key: child key (counts) or live slot (fills); a child that is not
            // written (not canonical) has no slot
            bool act = false;
            int agg = -1;
            if (o.valid) {
                if (!in_range(o.key, gd.keys)) {
                    atomicOr(gd.err, GS_GUARD_KEY);
                } else {
                    agg = (FILL && key_slot) ? key_slot[o.key] : o.key;
                    if (FILL && agg >= 0 && !in_range(agg, gd.slots)) atomicOr(gd.err, GS_GUARD_SLOT);
                    else act = agg >= 0;
                }
            }
            // the child's sets: one per child vertex (parent's nv, plus the new vertex of a
            // forward child), indexed by label-class rank
            SetRegs d{};
            if (FILL && act) {
                d = load_set(sets + agg);
                if (!set_fits<CV>(d, o.nv + (o.newv >= 0 ? 1 : 0), gd.pool_words)) {
                    atomicOr(gd.err, GS_GUARD_POOL);
                    act = false;
                }
            }
            // lead: this item inserts the parent's images for its (slot, row) pair;
            // ins: this item inserts images at all
            bool lead = true, ins = false;
            if (ROWS) {
                // one LDS atomic per item on its key's tile slot (grouping the lanes of one
                // key first cost more than it saved once long runs went to the run modes);
                // lead: the pair's first item in the tile
                if (act) {
                    hs[q] = lds_slot<LDS_SLOTS>(s_key, agg);
                    rank[q] = atomicAdd(&s_cnt[hs[q]], 1);
                    lead = !dedup || lds_first<LDS_SLOTS>(s_pair, hs[q] * TILE + row);
                }
                ins = hs[q] >= 0;
            } else if (act) {
                // run modes: each item is its row's only one of its key; weighted counts
                // go straight to the persistent table
                ins = true;
                if (RUNS && key_count) {
                    const int ph = lds_slot_fill<LDS_SLOTS>(s_pkey, agg, &s_pfill);
                    atomicAdd(&s_pcnt[ph], (cnt_t)o.w);
                }
            }
            if (FILL && ins) {
                if (ROWS) {
                    eid[q] = o.eid;
                    prow[q] = (int)(o.pbase + b0 + rlo + row);
                }
                const int nv = o.nv;
                // SETS leaves the image of a forward run's segment vertex to NEWV: it is
                // the group's vertex u, the same for every row of the group, and the group's
                // key has a child exactly when one of NEWV's items for it does
                if (PARENT && lead) {
#pragma unroll
                    for (int v = 0; v < PM + 1; ++v) {
                        if (v >= nv || (MODE == GS_ITEMS_SETS && v == o.sv)) continue;
                        const int x = c.phi[v];
                        if (in_range(x, g.n))
                            set_insert<CV>(d, pool, v, (uint32_t)g.lidx[x], gd.err, overflow);
                        else atomicOr(gd.err, GS_GUARD_GRAPH);
                    }
                }
                if (NEWX && o.newv >= 0)
                    set_insert<CV>(d, pool, nv, (uint32_t)o.newv, gd.err, overflow);
                if (NEWV && o.sv >= 0) {
                    const int x = pickv<PM + 1>(c.phi, o.sv);
                    if (in_range(x, g.n))
                        set_insert<CV>(d, pool, o.sv, (uint32_t)g.lidx[x], gd.err, overflow);
                    else atomicOr(gd.err, GS_GUARD_GRAPH);
                }
            }
        }
        // wave 0: the next tile's first row -- from this tile's offsets when it follows
        // this tile (its first item is t0 + TILE), else by a search over woff
        if (w0) {
            ntile = __shfl(ntile, 0);
            long long nrlo = 0;
            if (ntile == tile + 1) nrlo = rlo + wave_find_lds(off, nrows, TILE);
            else if (ntile >= 0) nrlo = wave_find_row(woff, R, ntile * TILE);
            publish(nb, ntile, nrlo);
        }
        __syncthreads();

        // phase 2: one output reservation per distinct key of the tile; counts go to the
        // workgroup's persistent table; the next tile's row offsets are loaded
        long long stage[STAGE];
        const long long ntl = s_tile[nb], nrlo = s_rlo[nb];
        const int nnr = s_nrows[nb];
#pragma unroll
        for (int k = 0; k < STAGE; ++k) {
            const int i = threadIdx.x + k * GS_BLOCK;
            stage[k] = (ntl >= 0 && i <= nnr) ? woff[nrlo + i] : 0;
        }
        if (AGG && key_count && threadIdx.x == 0) atomicAdd(&s_pfill, 1 << 16);  // a tile counted
        if constexpr (TTAB) {
            for (int h = threadIdx.x; h < LDS_SLOTS; h += GS_BLOCK) {
                const int key = s_key[h];
                if (key < 0) continue;
                if (ROWS && rows_out) {
                    // 64-bit cursors cannot wrap however far a key overruns a predicted
                    // capacity; a reservation that does not fit writes nothing
                    const unsigned long long cnt = (unsigned long long)s_cnt[h];
                    const unsigned long long at = atomicAdd(&cursor[key], cnt);
                    const unsigned long long cap = cap_end ? (unsigned long long)cap_end[key]
                                                           : (unsigned long long)gd.rows_cap;
                    const bool fits =
                        at + cnt <= cap && at + cnt <= (unsigned long long)gd.rows_cap;
                    s_base[h] = fits ? (int)at : -1;
                    if (!fits) atomicOr(cap_end ? overflow : gd.err, cap_end ? 1 : GS_GUARD_ROW);
                }
                if (key_count) {
                    const int ph = lds_slot_fill<LDS_SLOTS>(s_pkey, key, &s_pfill);
                    atomicAdd(&s_pcnt[ph], s_cnt[h]);
                }
            }
        }
        __syncthreads();
        // Flush the persistent table once more than TILE slots are used (a tile adds at
        // most TILE keys, so the next tile still finds room) or after FLUSH_TILES tiles.
        // Uniform: s_pfill was last written before the barrier above.
        if (AGG && key_count && ((s_pfill & 0xffff) > TILE || (s_pfill >> 16) >= FLUSH_TILES)) {
            for (int h = threadIdx.x; h < LDS_SLOTS; h += GS_BLOCK) {
                if (s_pkey[h] >= 0) {
                    count_flush(&key_count[s_pkey[h]], s_pcnt[h], saturate != 0);
                    s_pkey[h] = -1;
                    s_pcnt[h] = 0;
                }
            }
            __syncthreads();
            if (threadIdx.x == 0) s_pfill = 0;  // read again only after later barriers
        }
        // phase 3 (rows): child rows at their reserved positions
        if (ROWS && rows_out) {
#pragma unroll
            for (int q = 0; q < IPT; ++q) {
                if (hs[q] < 0) continue;
                const int base = s_base[hs[q]];
                if (base >= 0) rows_out[base + rank[q]] = make_int2(prow[q], eid[q]);
            }
            __syncthreads();
        }
        if constexpr (TTAB) {
            for (int h = threadIdx.x; h < LDS_SLOTS; h += GS_BLOCK) {
                if (s_key[h] >= 0) {
                    s_key[h] = -1;
                    s_cnt[h] = 0;
                }
                s_pair[h] = -1;
            }
        }
#pragma unroll
        for (int k = 0; k < STAGE; ++k) {
            const int i = threadIdx.x + k * GS_BLOCK;
            if (ntl >= 0 && i <= nnr) s_off[nb][i] = rel(stage[k], ntl * TILE);
        }
        __syncthreads();
        buf = nb;
    }
    // the loop ends after a barrier; the persistent table was last written before it
    if (AGG && key_count) {
        for (int h = threadIdx.x; h < LDS_SLOTS; h += GS_BLOCK) {
            if (s_pkey[h] < 0) continue;
            count_flush(&key_count[s_pkey[h]], s_pcnt[h], saturate != 0);
        }
    }
}

// ------------------------------------------------------------------ groups (NEWV)

constexpr unsigned long long kEmptyKey = ~0ull;

// First candidate neighbour (label >= min_label) of segment vertex u and their count.
__device__ __forceinline__ void seg_candidates(const DevGraph& g, int u, int min_label,
                                               int* first, int* cnt) {
    const int o0 = g.offsets[u], o1 = g.offsets[u + 1];
    int f;
    if (g.lstart) {
        f = g.lstart[(long long)u * g.lstride + min_label];
    } else {
        f = o0;
        int hi = o1;
        while (f < hi) {
            const int mid = (f + hi) >> 1;
            if ((int)(g.adjl[mid].y & 0xffffu) < min_label) f = mid + 1;
            else hi = mid;
        }
    }
    *first = f;
    *cnt = o1 - f;
}

// The pair (row b, segment j) of a chunk: its pattern, segment vertex and candidates.
__device__ __forceinline__ bool pair_of(const DevGraph& g, const PatDev* __restrict__ pats,
                                        const RowCtx* __restrict__ ctx, long long b, int j,
                                        int* p, int* u, int* first, int* cnt, int* gerr) {
    const RowCtx& c = ctx[b];
    *p = c.p;
    const PatDev& P = pats[*p];
    if (j > P.r) return false;
    *u = c.phi[P.seg_v[j]];
    if (!in_range(*u, g.n)) {
        atomicOr(gerr, GS_GUARD_GRAPH);
        return false;
    }
    seg_candidates(g, *u, P.min_label, first, cnt);
    return *cnt > 0;
}

__device__ __forceinline__ unsigned long long group_key(int p, int j, int u) {
    return ((unsigned long long)((unsigned)p * 8u + (unsigned)j) << 32) | (unsigned)u;
}

__device__ __forceinline__ unsigned long long mix64(unsigned long long x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    return x;
}

// pslot[pair] = table slot of the pair's group (-1: no candidates); tcnt = rows per group;
// trep = the row that claimed the slot.  The table has at least twice as many slots as
// pairs, so probing ends.
__global__ void k_group_claim(DevGraph g, const PatDev* __restrict__ pats,
                              const RowCtx* __restrict__ ctx, long long R, int nseg,
                              unsigned long long* tkeys, int* trep, int* tcnt, long long slots,
                              int* pslot, int* gerr) {
    const long long npairs = R * nseg;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long pi = (long long)blockIdx.x * blockDim.x + threadIdx.x; pi < npairs;
         pi += stride) {
        const long long b = pi / nseg;
        const int j = (int)(pi % nseg);
        int p, u, first, cnt;
        int slot = -1;
        if (pair_of(g, pats, ctx, b, j, &p, &u, &first, &cnt, gerr)) {
            const unsigned long long key = group_key(p, j, u);
            const unsigned long long mask = (unsigned long long)slots - 1;
            unsigned long long h = mix64(key) & mask;
            for (long long probe = 0; probe < slots; ++probe) {
                const unsigned long long prev = atomicCAS(&tkeys[h], kEmptyKey, key);
                if (prev == kEmptyKey) {
                    trep[h] = (int)b;
                    slot = (int)h;
                    break;
                }
                if (prev == key) {
                    slot = (int)h;
                    break;
                }
                h = (h + 1) & mask;
            }
            if (slot >= 0) atomicAdd(&tcnt[slot], 1);
            else atomicOr(gerr, GS_GUARD_SLOT);
        }
        pslot[pi] = slot;
    }
}

// pflag[pair] = 1 for the pair that claimed its group's slot
__global__ void k_group_flags(long long npairs, int nseg, const int* __restrict__ pslot,
                              const int* __restrict__ trep, int* pflag) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long pi = (long long)blockIdx.x * blockDim.x + threadIdx.x; pi < npairs;
         pi += stride) {
        const int h = pslot[pi];
        pflag[pi] = (h >= 0 && trep[h] == (int)(pi / nseg)) ? 1 : 0;
    }
}

// tgrp[slot] = group index (the claiming pair's rank among claiming pairs)
__global__ void k_group_index(long long npairs, const int* __restrict__ pslot,
                              const int* __restrict__ pflag, const long long* __restrict__ pidx,
                              int* tgrp) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long pi = (long long)blockIdx.x * blockDim.x + threadIdx.x; pi < npairs;
         pi += stride)
        if (pflag[pi]) tgrp[pslot[pi]] = (int)pidx[pi];
}

// hits[8 g + v]: rows of group g, other than its claiming row, that have the claiming row's
// image of pattern vertex v among their own images
__global__ void k_group_hits(const PatDev* __restrict__ pats, const RowCtx* __restrict__ ctx,
                             long long npairs, int nseg, const int* __restrict__ pslot,
                             const int* __restrict__ pflag, const int* __restrict__ trep,
                             const int* __restrict__ tgrp, int* hits) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long pi = (long long)blockIdx.x * blockDim.x + threadIdx.x; pi < npairs;
         pi += stride) {
        const int h = pslot[pi];
        if (h < 0 || pflag[pi]) continue;
        const RowCtx& me = ctx[pi / nseg];
        const RowCtx& rep = ctx[trep[h]];
        const int nv = pats[me.p].nv;
        const int gi = tgrp[h];
        for (int v = 0; v < GS_MAXV; ++v) {
            if (v >= nv) break;
            const int w = rep.phi[v];
            bool in = false;
            for (int x = 0; x < GS_MAXV; ++x) in |= x < nv && me.phi[x] == w;
            if (in) atomicAdd(&hits[(long long)gi * GS_MAXV + v], 1);
        }
    }
}

// The context and candidate count of every group, from its claiming pair.
__global__ void k_group_ctx(DevGraph g, const PatDev* __restrict__ pats,
                            const RowCtx* __restrict__ ctx, long long npairs, int nseg,
                            const int* __restrict__ pslot, const int* __restrict__ pflag,
                            const long long* __restrict__ pidx, const int* __restrict__ tcnt,
                            const int* __restrict__ hits, RowCtx* gctx, int* gwork, int* gerr) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long pi = (long long)blockIdx.x * blockDim.x + threadIdx.x; pi < npairs;
         pi += stride) {
        if (!pflag[pi]) continue;
        const long long b = pi / nseg;
        const int j = (int)(pi % nseg);
        const long long gi = pidx[pi];
        int p, u, first, cnt;
        if (!pair_of(g, pats, ctx, b, j, &p, &u, &first, &cnt, gerr)) continue;
        const PatDev& P = pats[p];
        const RowCtx& rep = ctx[b];
        const int others = tcnt[pslot[pi]] - 1;
        RowCtx o;
        for (int v = 0; v < GS_MAXV; ++v) {
            const bool common = v < P.nv && (v == P.seg_v[j] || hits[gi * GS_MAXV + v] == others);
            o.phi[v] = common ? rep.phi[v] : -1;
        }
        for (int s2 = 0; s2 < GS_MAXE; ++s2) {
            o.fcum[s2] = s2 < j ? 0 : cnt;  // every item falls in segment j
            o.base[s2] = first;
        }
        o.p = p;
        o.pad = 0;
        gctx[gi] = o;
        gwork[gi] = cnt;
    }
}

// ------------------------------------------------------------------ keys

// Ordered compaction of the keys with a non-zero count: flags, exclusive scan, scatter.
__global__ __launch_bounds__(GS_BLOCK) void k_active_flags(const int* key_count, long long K,
                                                           int* flags) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < K; i += stride)
        flags[i] = key_count[i] != 0 ? 1 : 0;  // negative: more than INT_MAX (count_add)
}

__global__ __launch_bounds__(GS_BLOCK) void k_active_scatter(const int* key_count,
                                                             const long long* pos, long long K,
                                                             int* active, int* active_counts,
                                                             int* n_active) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < K; i += stride) {
        const int c = key_count[i];
        if (c != 0) {
            active[pos[i]] = (int)i;
            active_counts[pos[i]] = c;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *n_active = (int)pos[K];
}

// The same compaction in one workgroup, for key spaces of at most GS_COMPACT_SMALL keys:
// one launch instead of five (flags, three scan stages, scatter), whose launch gaps cost
// more than the work at these sizes.  Each pass takes GS_BLOCK * 16 consecutive keys, 16
// per thread, and places them after the keys of the earlier passes.
__global__ __launch_bounds__(GS_BLOCK) void k_compact_small(const int* key_count, int K,
                                                            int* active, int* active_counts,
                                                            int* n_active, int* clear_word) {
    constexpr int PER = 16;
    long long placed = 0;  // uniform: every thread adds the same pass totals
    for (int base = 0; base < K; base += GS_BLOCK * PER) {
        const int i0 = base + threadIdx.x * PER;
        int c[PER];
        int nz = 0;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            c[q] = i0 + q < K ? key_count[i0 + q] : 0;
            nz += c[q] != 0;  // negative: more than INT_MAX (count_add)
        }
        long long total = 0;
        long long at = placed + block_exclusive_scan(nz, &total);
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            if (c[q] != 0) {
                active[at] = i0 + q;
                active_counts[at] = c[q];
                ++at;
            }
        }
        placed += total;
    }
    if (threadIdx.x == 0) {
        *n_active = (int)placed;
        if (clear_word) *clear_word = 0;
    }
}

__global__ void k_scatter_slots(int* key_slot, const int* keys, int nkeys) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nkeys) key_slot[keys[i]] = i;
}

__global__ void k_fill_i32(int* p, int v, long long n) {
    long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

__global__ void k_fill_u64(unsigned long long* p, unsigned long long v, long long n) {
    long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, long long n) {
    long long stride = (long long)gridDim.x * blockDim.x;
    long long n4 = n / 4;
    uint4* p4 = reinterpret_cast<uint4*>(p);
    uint4 v4 = make_uint4(v, v, v, v);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) p4[i] = v4;
    for (long long i = n4 * 4 + (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        p[i] = v;
}

// child count of a key from k_items' 32-bit count (a negative count: more than INT_MAX)
__device__ __forceinline__ long long child_rows(int c) {
    return c < 0 ? (long long)INT_MAX + 1 : (long long)c;
}

// distinct images per (set, vertex): bytes equal to the set's tag of a bytemap, popcount of
// a bitmap, or occupied hash slots.  One wave per set, sets taken grid-stride (a launch
// holds fewer than 2^32 work-items); each vertex set is read with 16-byte loads.
// counts[8 s + v] for all v < 8 (0 past the set's vertices, and for every vertex of a set
// whose key has fewer than min_rows children when rows is given).
__global__ __launch_bounds__(GS_BLOCK) void k_set_counts(const SetDesc* sets, long long nsets,
                                                         const uint32_t* pool,
                                                         long long pool_words, int* counts,
                                                         int* gerr, const int* rows,
                                                         int min_rows) {
    constexpr int WPB = GS_BLOCK / 64;  // waves per block
    const int lane = lane_id();
    const long long waves = (long long)gridDim.x * WPB;
    for (long long si = (long long)blockIdx.x * WPB + threadIdx.x / 64; si < nsets; si += waves) {
        SetDesc d = sets[si];
        if (rows && child_rows(rows[si]) < (long long)min_rows) d.nv = 0;
        const bool ok = d.nv == 0 || set_fits<GS_MAXV>(load_set(sets + si), 0, pool_words);
        if (!ok && lane == 0) atomicOr(gerr, GS_GUARD_POOL);
        for (int v = 0; v < GS_MAXV; ++v) {
            int c = 0;
            if (ok && v < d.nv) {
                const uint4* base = reinterpret_cast<const uint4*>(pool + d.off + d.vo[v]);
                const int n4 = (d.vo[v + 1] - d.vo[v]) / 4;
                const int kind = set_kind(d, v);
                if (kind == GS_SET_BYTEMAP) {  // bytes equal to the tag
                    const uint32_t t4 = (uint32_t)(d.tag & 0xff) * 0x01010101u;
                    auto eq = [t4](uint32_t w) {  // bytes of w equal to the tag byte
                        const uint32_t z = w ^ t4;
                        return __popc(~(((z & 0x7f7f7f7fu) + 0x7f7f7f7fu) | z | 0x7f7f7f7fu));
                    };
                    for (int i = lane; i < n4; i += 64) {
                        const uint4 w = base[i];
                        c += eq(w.x) + eq(w.y) + eq(w.z) + eq(w.w);
                    }
                } else if (kind == GS_SET_BITMAP) {
                    for (int i = lane; i < n4; i += 64) {
                        const uint4 w = base[i];
                        c += __popc(w.x) + __popc(w.y) + __popc(w.z) + __popc(w.w);
                    }
                } else {
                    for (int i = lane; i < n4; i += 64) {
                        const uint4 w = base[i];
                        c += (w.x != 0) + (w.y != 0) + (w.z != 0) + (w.w != 0);
                    }
                }
                for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
            }
            if (lane == 0) counts[si * GS_MAXV + v] = c;
        }
    }
}

// Support of one pattern straight from its layered rows (gspan.cu:165-190): the images of
// every pattern vertex go into the pattern's sets.
__global__ __launch_bounds__(GS_BLOCK) void k_support_rows(DevGraph g, const PatDev* pat,
                                                           long long N, LayerPtrs L,
                                                           const SetDesc* set, uint32_t* pool,
                                                           long long pool_words, int* gerr) {
    long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= N) return;
    const PatDev& P = *pat;
    int phi[GS_MAXV];
#pragma unroll
    for (int v = 0; v < GS_MAXV; ++v) phi[v] = -1;
    long long idx = b;
    for (int l = P.m - 1; l >= 0; --l) {
        if (!in_range(idx, L.rows[l])) {
            atomicOr(gerr, GS_GUARD_LAYER);
            return;
        }
        int2 rr = L.layer[l][idx];
        if (!in_range(rr.y, g.nnz)) {
            atomicOr(gerr, GS_GUARD_GRAPH);
            return;
        }
        if (P.cfrom[l] < P.cto[l]) phi[P.cto[l]] = (int)g.adj[rr.y].x;
        if (l == 0) phi[0] = g.src[rr.y];
        idx = rr.x;
    }
    const SetRegs d = load_set(set);
    if (!set_fits<GS_MAXV>(d, P.nv, pool_words)) {
        atomicOr(gerr, GS_GUARD_POOL);
        return;
    }
    for (int v = 0; v < P.nv; ++v) {
        if (in_range(phi[v], g.n)) set_insert<GS_MAXV>(d, pool, v, (uint32_t)g.lidx[phi[v]], gerr);
        else atomicOr(gerr, GS_GUARD_GRAPH);
    }
}

// ------------------------------------------------------------------ canonicality
// A candidate code lives in LDS as cf/ct/cel/ctl (ctl = -1 for backward edges) plus the
// label of every pattern vertex.  The pattern graph is the candidate itself
// (graph.cc:122-159): the adjacency of x lists the code edges touching x in reverse code
// order.  A projection packs the image of code vertex c in bits [3c, 3c+3) and the set of
// used pattern edges in bits [24, 32).

struct CanonCode {
    int m;
    int cf[GS_MAXE], ct[GS_MAXE], cel[GS_MAXE], ctl[GS_MAXE];
    int lab[GS_MAXV];
};

__device__ __forceinline__ int pj_phi(uint32_t p, int c) { return (int)((p >> (3 * c)) & 7u); }

// has_edge (gspan.cu:396-401) on a projection: a used pattern edge with label l joins u, w
__device__ __forceinline__ bool pj_has_edge(const CanonCode& C, uint32_t p, int u, int w, int l) {
    uint32_t used = p >> 24;
    while (used) {
        int j = __builtin_ctz(used);
        used &= used - 1;
        if (C.cel[j] == l &&
            ((C.cf[j] == u && C.ct[j] == w) || (C.cf[j] == w && C.ct[j] == u)))
            return true;
    }
    return false;
}

__device__ __forceinline__ bool pj_has_vertex(uint32_t p, int nvm, int w) {
    for (int c = 0; c < nvm; ++c)
        if (pj_phi(p, c) == w) return true;
    return false;
}

__device__ __forceinline__ long long wave_min(long long v) {
    for (int off = 32; off > 0; off >>= 1) {
        long long o = __shfl_xor(v, off);
        v = o < v ? o : v;
    }
    return v;
}

constexpr long long KEY_NONE = 0x7fffffffffffffffll;

__device__ __forceinline__ int wave_excl_sum(int v, int* total) {
    int x = v;
    for (int off = 1; off < 64; off <<= 1) {
        int y = __shfl_up(x, off);
        if (lane_id() >= off) x += y;
    }
    *total = __shfl(x, 63);
    return x - v;
}

// Backward probe of one projection at rightmost-path position k (gspan.cu:428-445 on the
// pattern graph).  Returns the edge label or -1; *je = pattern edge used.
__device__ __forceinline__ int canon_backward(const CanonCode& C, uint32_t p, int mf_k, int mt_k,
                                              int mel_k, int v_n, int* je) {
    const int u_n = pj_phi(p, v_n);
    const int e_from = pj_phi(p, mf_k), e_to = pj_phi(p, mt_k);
    for (int j = C.m - 1; j >= 0; --j) {
        int w;
        if (C.cf[j] == u_n) w = C.ct[j];
        else if (C.ct[j] == u_n) w = C.cf[j];
        else continue;
        const int l = C.cel[j];
        if (pj_has_edge(C, p, u_n, w, l)) continue;
        if (w == e_from && (mel_k < l || (mel_k == l && C.lab[e_to] <= C.lab[u_n]))) {
            *je = j;
            return l;
        }
    }
    return -1;
}

// Returns 1 canonical, 0 not; err bit 0: projection buffer overflow (more than CAP
// projections), bit 1: walk overran.
template <int CAP>
__device__ int canon_wave(const CanonCode& C, uint32_t* bufA, uint32_t* bufB, int* mf, int* mt,
                          int* mel, int* err) {
    const int lane = lane_id();
    const int m = C.m;
    if (m == 1) return 1;  // gspan.cu:592-593
    // initial edge: min (from_label, edge_label, to_label) over entries with lab[from] <= lab[to]
    long long key = KEY_NONE;
    int x = 0, y = 0, jj = 0;
    if (lane < 2 * m) {
        jj = lane >> 1;
        x = (lane & 1) ? C.ct[jj] : C.cf[jj];
        y = (lane & 1) ? C.cf[jj] : C.ct[jj];
        if (C.lab[x] <= C.lab[y])
            key = ((long long)C.lab[x] << 32) | ((long long)C.cel[jj] << 16) | C.lab[y];
    }
    const long long kmin = wave_min(key);
    bool take = lane < 2 * m && key == kmin;
    unsigned long long tm = __ballot(take);
    if (take) bufA[mbcnt64(tm)] = (uint32_t)x | ((uint32_t)y << 3) | (1u << (24 + jj));
    int P = __popcll(tm);
    uint32_t* cur = bufA;
    uint32_t* nxt = bufB;
    int len = 1;
    if (lane == 0) {
        mf[0] = 0;
        mt[0] = 1;
        mel[0] = (int)((kmin >> 16) & 0xffff);
    }
    const int min_label = (int)(kmin >> 32);
    __syncthreads();
    for (;;) {
        // rightmost path of the minimum code so far (gspan.cu:53-65)
        int rm[GS_MAXE];
        int r = 0, last = -1;
        for (int i = len - 1; i >= 0; --i)
            if (mf[i] < mt[i] && (r == 0 || mt[i] == last)) {
                rm[r++] = i;
                last = mf[i];
            }
        const int v_n = mt[rm[0]];
        const int nvm = v_n + 1;

        // backward (gspan.cu:494-524)
        int bi = -1, bl = 0x7fffffff;
        for (int i = r - 1; i >= 1 && bi < 0; --i) {
            const int k = rm[i];
            long long best = KEY_NONE;
            for (int base = 0; base < P; base += 64) {
                int pi = base + lane;
                if (pi < P) {
                    int je;
                    int l = canon_backward(C, cur[pi], mf[k], mt[k], mel[k], v_n, &je);
                    if (l >= 0 && l < best) best = l;
                }
            }
            best = wave_min(best);
            if (best != KEY_NONE) {
                bi = i;
                bl = (int)best;
            }
        }
        if (bi >= 0) {
            const int k = rm[bi];
            if (len >= m) {
                *err |= 2;
                return 1;
            }
            if (!(C.cf[len] == v_n && C.ct[len] == mf[k] && C.cel[len] == bl && C.ctl[len] < 0))
                return 0;
            int np = 0;
            for (int base = 0; base < P; base += 64) {
                int pi = base + lane;
                bool keep = false;
                uint32_t pr = 0;
                if (pi < P) {
                    int je;
                    pr = cur[pi];
                    int l = canon_backward(C, pr, mf[k], mt[k], mel[k], v_n, &je);
                    if (l == bl) {
                        keep = true;
                        pr |= 1u << (24 + je);
                    }
                }
                unsigned long long km = __ballot(keep);
                if (keep) {
                    int at = np + (int)mbcnt64(km);
                    if (at < CAP) nxt[at] = pr;
                }
                np += __popcll(km);
            }
            if (np > CAP) {
                *err |= 1;
                return 1;
            }
            __syncthreads();
            if (lane == 0) {
                mf[len] = v_n;
                mt[len] = mf[k];
                mel[len] = bl;
            }
            ++len;
            uint32_t* t = cur;
            cur = nxt;
            nxt = t;
            P = np;
            __syncthreads();
            continue;
        }

        // forward (gspan.cu:526-581): first forward from v_n, else other forward from the
        // rightmost path, deepest source first
        int seg = -1;
        long long fkey = KEY_NONE;
        for (int s = 0; s <= r && seg < 0; ++s) {
            long long best = KEY_NONE;
            for (int base = 0; base < P; base += 64) {
                int pi = base + lane;
                if (pi >= P) continue;
                const uint32_t pr = cur[pi];
                int u, e_el = 0, bound = 0;
                if (s == 0) {
                    u = pj_phi(pr, v_n);
                } else {
                    const int k = rm[s - 1];
                    u = pj_phi(pr, mf[k]);
                    e_el = mel[k];
                    bound = C.lab[pj_phi(pr, mt[k])];
                }
                for (int j = m - 1; j >= 0; --j) {
                    int w;
                    if (C.cf[j] == u) w = C.ct[j];
                    else if (C.ct[j] == u) w = C.cf[j];
                    else continue;
                    const int l = C.cel[j], tl = C.lab[w];
                    if (tl < min_label || pj_has_vertex(pr, nvm, w)) continue;
                    if (s > 0 && !(e_el < l || (e_el == l && bound <= tl))) continue;
                    const long long kk = ((long long)l << 32) | tl;
                    if (kk < best) best = kk;
                }
            }
            best = wave_min(best);
            if (best != KEY_NONE) {
                seg = s;
                fkey = best;
            }
        }
        if (seg < 0) return 1;  // gspan.cu:583
        const int fel = (int)(fkey >> 32), ftl = (int)(fkey & 0xffffffff);
        const int v_i = seg == 0 ? v_n : mf[rm[seg - 1]];
        if (len >= m) {
            *err |= 2;
            return 1;
        }
        if (!(C.cf[len] == v_i && C.ct[len] == v_n + 1 && C.cel[len] == fel && C.ctl[len] == ftl))
            return 0;
        int np = 0;
        for (int base = 0; base < P; base += 64) {
            int pi = base + lane;
            uint32_t pr = 0;
            int cnt = 0;
            uint32_t outs[GS_MAXE];
            if (pi < P) {
                pr = cur[pi];
                const int u = pj_phi(pr, v_i);
                for (int j = m - 1; j >= 0; --j) {
                    int w;
                    if (C.cf[j] == u) w = C.ct[j];
                    else if (C.ct[j] == u) w = C.cf[j];
                    else continue;
                    const int l = C.cel[j], tl = C.lab[w];
                    if (l != fel || tl != ftl) continue;
                    if (pj_has_vertex(pr, nvm, w)) continue;
                    // the min key already satisfies the segment's rule when it was chosen
                    // from this segment; entries with the same (l, tl) satisfy it too
                    outs[cnt++] = pr | ((uint32_t)w << (3 * (v_n + 1))) | (1u << (24 + j));
                }
            }
            int tot;
            int at = np + wave_excl_sum(cnt, &tot);
            for (int q = 0; q < cnt; ++q)
                if (at + q < CAP) nxt[at + q] = outs[q];
            np += tot;
        }
        if (np > CAP) {
            *err |= 1;
            return 1;
        }
        __syncthreads();
        if (lane == 0) {
            mf[len] = v_i;
            mt[len] = v_n + 1;
            mel[len] = fel;
        }
        ++len;
        uint32_t* t = cur;
        cur = nxt;
        nxt = t;
        P = np;
        __syncthreads();
    }
}

// canon_wave's walk run by one thread, for codes of at most three edges: at most 3!
// injective maps of the code's edges onto the pattern graph's, in 2^3 directions, so at
// most 48 projections at any step (CAPL = 64 leaves room; err bit 0 past it).  The
// reductions over projections become loops; every rule is the one canon_wave applies.
template <int CAPL>
__device__ int canon_seq(const CanonCode& C, uint32_t* bufA, uint32_t* bufB, int* err) {
    const int m = C.m;
    if (m == 1) return 1;  // gspan.cu:592-593
    long long kmin = KEY_NONE;
    for (int e = 0; e < 2 * m; ++e) {
        const int j = e >> 1;
        const int x = (e & 1) ? C.ct[j] : C.cf[j], y = (e & 1) ? C.cf[j] : C.ct[j];
        if (C.lab[x] <= C.lab[y]) {
            const long long key =
                ((long long)C.lab[x] << 32) | ((long long)C.cel[j] << 16) | C.lab[y];
            if (key < kmin) kmin = key;
        }
    }
    int P = 0;
    for (int e = 0; e < 2 * m; ++e) {
        const int j = e >> 1;
        const int x = (e & 1) ? C.ct[j] : C.cf[j], y = (e & 1) ? C.cf[j] : C.ct[j];
        if (C.lab[x] > C.lab[y]) continue;
        const long long key = ((long long)C.lab[x] << 32) | ((long long)C.cel[j] << 16) | C.lab[y];
        if (key == kmin) bufA[P++] = (uint32_t)x | ((uint32_t)y << 3) | (1u << (24 + j));
    }
    uint32_t* cur = bufA;
    uint32_t* nxt = bufB;
    int mf[GS_MAXE], mt[GS_MAXE], mel[GS_MAXE];
    mf[0] = 0;
    mt[0] = 1;
    mel[0] = (int)((kmin >> 16) & 0xffff);
    const int min_label = (int)(kmin >> 32);
    int len = 1;
    for (;;) {
        int rm[GS_MAXE];
        int r = 0, last = -1;
        for (int i = len - 1; i >= 0; --i)
            if (mf[i] < mt[i] && (r == 0 || mt[i] == last)) {
                rm[r++] = i;
                last = mf[i];
            }
        const int v_n = mt[rm[0]];
        const int nvm = v_n + 1;
        int bi = -1, bl = 0x7fffffff;
        for (int i = r - 1; i >= 1 && bi < 0; --i) {
            const int k = rm[i];
            int best = 0x7fffffff;
            for (int pi = 0; pi < P; ++pi) {
                int je;
                const int l = canon_backward(C, cur[pi], mf[k], mt[k], mel[k], v_n, &je);
                if (l >= 0 && l < best) best = l;
            }
            if (best != 0x7fffffff) {
                bi = i;
                bl = best;
            }
        }
        if (bi >= 0) {
            const int k = rm[bi];
            if (len >= m) {
                *err |= 2;
                return 1;
            }
            if (!(C.cf[len] == v_n && C.ct[len] == mf[k] && C.cel[len] == bl && C.ctl[len] < 0))
                return 0;
            int np = 0;
            for (int pi = 0; pi < P; ++pi) {
                int je;
                const uint32_t pr = cur[pi];
                if (canon_backward(C, pr, mf[k], mt[k], mel[k], v_n, &je) != bl) continue;
                if (np == CAPL) {
                    *err |= 1;
                    return 1;
                }
                nxt[np++] = pr | (1u << (24 + je));
            }
            mf[len] = v_n;
            mt[len] = mf[k];
            mel[len] = bl;
            ++len;
            uint32_t* t = cur;
            cur = nxt;
            nxt = t;
            P = np;
            continue;
        }
        int seg = -1;
        long long fkey = KEY_NONE;
        for (int sg = 0; sg <= r && seg < 0; ++sg) {
            long long best = KEY_NONE;
            for (int pi = 0; pi < P; ++pi) {
                const uint32_t pr = cur[pi];
                int u, e_el = 0, bound = 0;
                if (sg == 0) {
                    u = pj_phi(pr, v_n);
                } else {
                    const int k = rm[sg - 1];
                    u = pj_phi(pr, mf[k]);
                    e_el = mel[k];
                    bound = C.lab[pj_phi(pr, mt[k])];
                }
                for (int j = m - 1; j >= 0; --j) {
                    int w;
                    if (C.cf[j] == u) w = C.ct[j];
                    else if (C.ct[j] == u) w = C.cf[j];
                    else continue;
                    const int l = C.cel[j], tl = C.lab[w];
                    if (tl < min_label || pj_has_vertex(pr, nvm, w)) continue;
                    if (sg > 0 && !(e_el < l || (e_el == l && bound <= tl))) continue;
                    const long long kk = ((long long)l << 32) | tl;
                    if (kk < best) best = kk;
                }
            }
            if (best != KEY_NONE) {
                seg = sg;
                fkey = best;
            }
        }
        if (seg < 0) return 1;  // gspan.cu:583
        const int fel = (int)(fkey >> 32), ftl = (int)(fkey & 0xffffffff);
        const int v_i = seg == 0 ? v_n : mf[rm[seg - 1]];
        if (len >= m) {
            *err |= 2;
            return 1;
        }
        if (!(C.cf[len] == v_i && C.ct[len] == v_n + 1 && C.cel[len] == fel && C.ctl[len] == ftl))
            return 0;
        int np = 0;
        for (int pi = 0; pi < P; ++pi) {
            const uint32_t pr = cur[pi];
            const int u = pj_phi(pr, v_i);
            for (int j = m - 1; j >= 0; --j) {
                int w;
                if (C.cf[j] == u) w = C.ct[j];
                else if (C.ct[j] == u) w = C.cf[j];
                else continue;
                if (C.cel[j] != fel || C.lab[w] != ftl || pj_has_vertex(pr, nvm, w)) continue;
                if (np == CAPL) {
                    *err |= 1;
                    return 1;
                }
                nxt[np++] = pr | ((uint32_t)w << (3 * (v_n + 1))) | (1u << (24 + j));
            }
        }
        mf[len] = v_i;
        mt[len] = v_n + 1;
        mel[len] = fel;
        ++len;
        uint32_t* t = cur;
        cur = nxt;
        nxt = t;
        P = np;
    }
}

// is_canonical of a two-edge code (gspan.cu:586-606 with projection_is_canonical,
// gspan.cu:485-584), walked on the code's pattern graph held in registers: at most three
// vertices and two edges, i.e. four directed entries.  The walk starts from every entry
// equal to the least initial entry (label(from) <= label(to), ordered by (from label,
// edge label, to label), gspan.cu:40-50, 410-426).  One edge is then left and the
// backward search needs a rightmost path of two edges, so the minimal code's second edge
// is the least (edge label, to label) first forward edge from an embedding's `to`
// (gspan.cu:447-462), else the least other forward edge from its `from` under the
// edge-label rule (gspan.cu:464-483); with neither the walk accepts (gspan.cu:583).
// Labels are the code's own (vertex v: C.lab[v]).
__device__ __forceinline__ int canon_two(const CanonCode& C) {
    const int L0 = C.lab[0], L1 = C.lab[1], L2 = C.lab[2];
    auto lab = [&](int v) { return v == 0 ? L0 : (v == 1 ? L1 : L2); };
    auto less3 = [](int a, int e, int b, int xa, int xe, int xb) {
        return a < xa || (a == xa && (e < xe || (e == xe && b < xb)));
    };
    int ef[4], et[4], ee[4];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        ef[2 * j] = C.cf[j];
        et[2 * j] = C.ct[j];
        ef[2 * j + 1] = C.ct[j];
        et[2 * j + 1] = C.cf[j];
        ee[2 * j] = ee[2 * j + 1] = C.cel[j];
    }
    int ba = INT_MAX, be = INT_MAX, bb = INT_MAX;  // least initial entry
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int a = lab(ef[i]), b = lab(et[i]);
        if (a <= b && less3(a, ee[i], b, ba, be, bb)) {
            ba = a;
            be = ee[i];
            bb = b;
        }
    }
    const int min_label = ba;
    bool ff = false, of = false;
    int fe = INT_MAX, fl = INT_MAX;  // least (edge label, to label) found
    auto take = [&](int el, int tl) {
        if (el < fe || (el == fe && tl < fl)) {
            fe = el;
            fl = tl;
        }
    };
#pragma unroll
    for (int i = 0; i < 4; ++i) {  // first forward edges from `to` of each embedding
        if (lab(ef[i]) != ba || ee[i] != be || lab(et[i]) != bb) continue;
        const int u = ef[i], v = et[i];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int w = et[k];
            if (ef[k] == v && w != u && w != v && lab(w) >= min_label) {
                ff = true;
                take(ee[k], lab(w));
            }
        }
    }
    if (!ff) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {  // other forward edges from `from` of each embedding
            if (lab(ef[i]) != ba || ee[i] != be || lab(et[i]) != bb) continue;
            const int u = ef[i], v = et[i], eu = ee[i];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int w = et[k], tl = lab(w);
                if (ef[k] == u && w != u && w != v && tl >= min_label &&
                    (eu < ee[k] || (eu == ee[k] && lab(v) <= tl))) {
                    of = true;
                    take(ee[k], tl);
                }
            }
        }
    }
    if (!ff && !of) return 1;
    return C.cf[1] == (ff ? 1 : 0) && C.ct[1] == 2 && C.cel[1] == fe && C.ctl[1] == fl;
}

// Canonicality of the two-edge children of single-edge parents: one thread per active key
// (the child of slot `slot` as decode_child writes it, for a parent of one edge).
// The two-edge code of child key `key` of a batch of single-edge parents.
__device__ __forceinline__ void two_edge_code(const PatDev* __restrict__ pats, int npats,
                                              KeySpace ks, int key, CanonCode& C) {
    const PatDev& P = pats[owner_one_edge(pats, npats, ks, key)];
    const int slot = key - P.key_base;
    C.m = 2;
    C.cf[0] = P.cfrom[0];
    C.ct[0] = P.cto[0];
    C.cel[0] = P.cel[0];
    C.ctl[0] = P.vlab[P.cto[0]];
    C.lab[0] = P.vlab[0];
    C.lab[1] = P.vlab[1];
    C.lab[2] = 0;
    const int nb = (P.r - 1) * ks.EL;
    if (slot < nb) {  // backward: a second edge between the parent's two vertices
        C.cf[1] = P.v_n;
        C.ct[1] = P.cfrom[P.rm[slot / ks.EL + 1]];
        C.cel[1] = slot % ks.EL;
        C.ctl[1] = -1;
    } else {
        const int sl = slot - nb;
        const int j = sl / (ks.EL * ks.VL), el = (sl / ks.VL) % ks.EL, tl = sl % ks.VL;
        C.cf[1] = j == 0 ? P.v_n : P.cfrom[P.rm[j - 1]];
        C.ct[1] = 2;
        C.cel[1] = el;
        C.ctl[1] = tl;
        C.lab[2] = tl;
    }
}

__global__ __launch_bounds__(GS_BLOCK) void k_canon_two(const PatDev* __restrict__ pats,
                                                        int npats, KeySpace ks,
                                                        const int* active_keys,
                                                        const int* n_active, int* flags) {
    const long long na = *n_active;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long a = (long long)blockIdx.x * blockDim.x + threadIdx.x; a < na; a += stride) {
        CanonCode C;
        two_edge_code(pats, npats, ks, active_keys[a], C);
        flags[a] = canon_two(C);
    }
}

// A step over single-edge parents whose children's counts and supports come from the
// two-edge pass, in one launch over its K keys: sel[key] = the child's support when the
// child has rows, its code is canonical and both its child count and its support reach
// min_sup, else 0.  Workgroup b writes part[3 b .. 3 b + 2] = the child rows and the keys
// with children, and the canonical ones among them (what the compaction, canonicality and
// support launches of other steps add up); the host adds the parts (atomics on three
// words from every workgroup would queue on them).
__global__ __launch_bounds__(GS_BLOCK) void k_two_edge_select(
    const PatDev* __restrict__ pats, int npats, KeySpace ks, long long K,
    const int* __restrict__ key_count, const int* __restrict__ ksup, int min_sup, int* sel,
    unsigned long long* __restrict__ part) {
    __shared__ unsigned long long s_t[3][GS_BLOCK / 64];
    unsigned long long rows = 0, canon = 0, cand = 0;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long key = (long long)blockIdx.x * blockDim.x + threadIdx.x; key < K;
         key += stride) {
        const int kc = key_count[key];
        int out = 0;
        if (kc != 0) {
            const long long cr = child_rows(kc);
            rows += (unsigned long long)cr;
            ++cand;
            CanonCode C;
            two_edge_code(pats, npats, ks, (int)key, C);
            if (canon_two(C)) {
                ++canon;
                if (cr >= (long long)min_sup) {
                    const int sp = ksup[key];
                    out = sp >= min_sup ? sp : 0;
                }
            }
        }
        sel[key] = out;
    }
    for (int off = 32; off > 0; off >>= 1) {
        rows += __shfl_xor(rows, off);
        canon += __shfl_xor(canon, off);
        cand += __shfl_xor(cand, off);
    }
    if (lane_id() == 0) {
        s_t[0][threadIdx.x / 64] = rows;
        s_t[1][threadIdx.x / 64] = canon;
        s_t[2][threadIdx.x / 64] = cand;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t[3] = {0, 0, 0};
        for (int w = 0; w < GS_BLOCK / 64; ++w)
            for (int i = 0; i < 3; ++i) t[i] += s_t[i][w];
        part[3 * blockIdx.x] = t[0];
        part[3 * blockIdx.x + 1] = t[2];
        part[3 * blockIdx.x + 2] = t[1];
    }
}

__device__ void decode_child(const PatDev& Pp, int slot, KeySpace ks, CanonCode& C) {
    C.m = Pp.m + 1;
    for (int j = 0; j < Pp.m; ++j) {
        C.cf[j] = Pp.cfrom[j];
        C.ct[j] = Pp.cto[j];
        C.cel[j] = Pp.cel[j];
        C.ctl[j] = Pp.cfrom[j] < Pp.cto[j] ? Pp.vlab[Pp.cto[j]] : -1;
    }
    for (int v = 0; v < GS_MAXV; ++v) C.lab[v] = v < Pp.nv ? Pp.vlab[v] : 0;
    const int nb = (Pp.r - 1) * ks.EL;
    const int m = Pp.m;
    if (slot < nb) {
        const int i = slot / ks.EL + 1, el = slot % ks.EL;
        C.cf[m] = Pp.v_n;
        C.ct[m] = Pp.cfrom[Pp.rm[i]];
        C.cel[m] = el;
        C.ctl[m] = -1;
    } else {
        const int s = slot - nb;
        const int j = s / (ks.EL * ks.VL), el = (s / ks.VL) % ks.EL, tl = s % ks.VL;
        C.cf[m] = j == 0 ? Pp.v_n : Pp.cfrom[Pp.rm[j - 1]];
        C.ct[m] = Pp.v_n + 1;
        C.cel[m] = el;
        C.ctl[m] = tl;
        C.lab[Pp.v_n + 1] = tl;
    }
}

// Canonicality of three-edge children: one thread per active key runs canon_seq.
__global__ __launch_bounds__(GS_BLOCK) void k_canon_three(const PatDev* __restrict__ pats,
                                                          int npats, KeySpace ks,
                                                          const int* active_keys,
                                                          const int* n_active, int* flags,
                                                          int* err) {
    const long long na = *n_active;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long a = (long long)blockIdx.x * blockDim.x + threadIdx.x; a < na; a += stride) {
        const int key = active_keys[a];
        int lo = 0, hi = npats - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (pats[mid].key_base <= key) lo = mid;
            else hi = mid - 1;
        }
        CanonCode C;
        decode_child(pats[lo], key - pats[lo].key_base, ks, C);
        uint32_t bufA[64], bufB[64];
        int e = 0;
        flags[a] = canon_seq<64>(C, bufA, bufB, &e);
        if (e) atomicOr(err, e);
    }
}

template <int CAP>
__global__ __launch_bounds__(64) void k_canon(const PatDev* __restrict__ pats, int npats,
                                              KeySpace ks, const int* active_keys,
                                              const int* n_active, int* flags, int* err) {
    __shared__ uint32_t bufA[CAP], bufB[CAP];
    __shared__ CanonCode C;
    __shared__ int mf[GS_MAXE], mt[GS_MAXE], mel[GS_MAXE];
    const int na = *n_active;
    for (int a = blockIdx.x; a < na; a += gridDim.x) {
        if (threadIdx.x == 0) {
            const int key = active_keys[a];
            int lo = 0, hi = npats - 1;
            while (lo < hi) {
                int mid = (lo + hi + 1) >> 1;
                if (pats[mid].key_base <= key) lo = mid;
                else hi = mid - 1;
            }
            decode_child(pats[lo], key - pats[lo].key_base, ks, C);
        }
        __syncthreads();
        int e = 0;
        int f = canon_wave<CAP>(C, bufA, bufB, mf, mt, mel, &e);
        if (threadIdx.x == 0) {
            flags[a] = f;
            if (e) atomicOr(err, e);
        }
        __syncthreads();
    }
}

// Codes given explicitly: codes[i*(1+5*GS_MAXE)] = m, then (from, to, from_label, edge_label,
// to_label) per edge in label-rank space, -1 where the reference stores no label.
template <int CAP>
__global__ __launch_bounds__(64) void k_canon_codes(const int* codes, int ncodes, int* flags,
                                                    int* err) {
    __shared__ uint32_t bufA[CAP], bufB[CAP];
    __shared__ CanonCode C;
    __shared__ int mf[GS_MAXE], mt[GS_MAXE], mel[GS_MAXE];
    for (int a = blockIdx.x; a < ncodes; a += gridDim.x) {
        if (threadIdx.x == 0) {
            const int* c = codes + (long long)a * (1 + 5 * GS_MAXE);
            C.m = c[0];
            for (int v = 0; v < GS_MAXV; ++v) C.lab[v] = 0;
            for (int j = 0; j < C.m; ++j) {
                const int* e = c + 1 + 5 * j;
                C.cf[j] = e[0];
                C.ct[j] = e[1];
                C.cel[j] = e[3];
                C.ctl[j] = e[0] < e[1] ? e[4] : -1;
                if (e[2] >= 0) C.lab[e[0]] = e[2];
                if (e[4] >= 0) C.lab[e[1]] = e[4];
            }
        }
        __syncthreads();
        int e = 0;
        // C.m is uniform: two- and three-edge codes take the one-thread walks (thread 0)
        int f = 0;
        if (C.m == 2) {
            if (threadIdx.x == 0) f = canon_two(C);
        } else if (C.m == 3) {
            if (threadIdx.x == 0) f = canon_seq<64>(C, bufA, bufB, &e);
        } else {
            f = canon_wave<CAP>(C, bufA, bufB, mf, mt, mel, &e);
        }
        if (threadIdx.x == 0) {
            flags[a] = f;
            if (e) atomicOr(err, e);
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ step tail
// Planning, supports and the frequent children of a step on the device, for steps with
// many candidate keys: the host then sees only the frequent children.

// Pattern owning batch key `key` (pats ordered by key_base).
__device__ __forceinline__ int owner_of(const PatDev* __restrict__ pats, int npats, int key) {
    int lo = 0, hi = npats - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pats[mid].key_base <= key) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Child of active key `key`: its vertex count, the label rank of each of its vertices
// (lab, GS_MAXV entries) and whether (GS_MNI_OMP_LITERAL) its vertex 0 takes part in the
// support -- only when an edge after the first leaves vertex 0, as support_omp inserts an
// edge's source only below layer 0 (gspan.cu:307, 360).
__device__ __forceinline__ void child_shape(const PatDev& P, int slot, KeySpace ks, int* nv,
                                            bool* v0, int* lab) {
    const int nb = (P.r - 1) * ks.EL;
    bool from0 = false;
    for (int j = 1; j < GS_MAXE; ++j)
        if (j < P.m && P.cfrom[j] == 0) from0 = true;
    for (int v = 0; v < GS_MAXV; ++v) lab[v] = v < P.nv ? P.vlab[v] : 0;
    int from;
    if (slot < nb) {
        from = P.v_n;  // backward: from the rightmost vertex
        *nv = P.nv;
    } else {
        const int seg = (slot - nb) / (ks.EL * ks.VL);
        from = seg == 0 ? P.v_n : P.cfrom[P.rm[seg - 1]];
        *nv = P.nv + 1;
        if (P.nv < GS_MAXV) lab[P.nv] = (slot - nb) % ks.VL;  // the new vertex's label
    }
    *v0 = from0 || from == 0;
}

// Per active key a < na: set words and rows of its child (0 unless canonical; no set words
// when the sets were planned for every key beforehand, rows only when rows are written).
__global__ void k_plan_children(const PatDev* __restrict__ pats, int npats, KeySpace ks,
                                const int* __restrict__ akeys, const int* __restrict__ acnt,
                                const int* __restrict__ flags, int na, const int* lcnt,
                                int bytes, int plan_sets, int write_rows, int* set_words,
                                int* rows) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long a = (long long)blockIdx.x * blockDim.x + threadIdx.x; a < na; a += stride) {
        int w = 0, r = 0;
        if (flags[a]) {
            const int key = akeys[a];
            const PatDev& P = pats[owner_of(pats, npats, key)];
            int nv, lab[GS_MAXV];
            bool v0;
            child_shape(P, key - P.key_base, ks, &nv, &v0, lab);
            const long long cr = child_rows(acnt[a]);
            if (plan_sets) {
                SetDesc d;
                w = (int)plan_key_sets(cr, nv, lab, lcnt, bytes != 0, &d);
            }
            if (write_rows) r = (int)min(cr, (long long)INT_MAX);
        }
        set_words[a] = w;
        rows[a] = r;
    }
}

// Set descriptor, slot and row cursor of every active key a < na (slot = a).  Keys that
// are not canonical get an empty descriptor (nv = 0) and no slot.
__global__ void k_build_sets(const PatDev* __restrict__ pats, int npats, KeySpace ks,
                             const int* __restrict__ akeys, const int* __restrict__ acnt,
                             const int* __restrict__ flags, int na, const int* lcnt,
                             int bytes, const long long* __restrict__ set_off,
                             const long long* __restrict__ row_off,
                             const SetDesc* __restrict__ key_sets, SetDesc* sets, int* key_slot,
                             unsigned long long* cursor) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long a = (long long)blockIdx.x * blockDim.x + threadIdx.x; a < na; a += stride) {
        SetDesc d{};  // no sets (nv = 0)
        const int key = akeys[a];
        if (flags[a]) {
            if (key_sets) {  // planned per key beforehand (single pass)
                d = key_sets[key];
            } else {
                const PatDev& P = pats[owner_of(pats, npats, key)];
                int nv, lab[GS_MAXV];
                bool v0;
                child_shape(P, key - P.key_base, ks, &nv, &v0, lab);
                plan_key_sets(child_rows(acnt[a]), nv, lab, lcnt, bytes != 0, &d);
                d.off = set_off[a];
            }
            if (key_slot) key_slot[key] = (int)a;
        }
        sets[a] = d;
        if (cursor) cursor[a] = row_off ? (unsigned long long)row_off[a] : 0ull;
    }
}

// Support of the child of active key a (gspan.cu:282-288 / 367-373: the least non-zero
// count of distinct images over its vertices, vertex 0 only where it takes part);
// sup[a] = the support when it reaches min_sup, else 0.  Totals: [0] += child rows of all
// active keys (a negative count adds 2^31), [1] += canonical keys.  ksup (not null): the
// support of every key, computed with its counts, in place of counts.
__global__ void k_supports(const PatDev* __restrict__ pats, int npats, KeySpace ks,
                           const int* __restrict__ akeys, const int* __restrict__ acnt,
                           const int* __restrict__ flags, int na, const int* __restrict__ counts,
                           const int* __restrict__ ksup, int mni_mode, int min_sup, int* sup,
                           unsigned long long* totals) {
    __shared__ unsigned long long s_t[2][GS_BLOCK / 64];
    unsigned long long rows = 0, canon = 0;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long a = (long long)blockIdx.x * blockDim.x + threadIdx.x; a < na; a += stride) {
        const long long cr = child_rows(acnt[a]);
        rows += (unsigned long long)cr;
        int out = 0;
        if (flags[a]) ++canon;
        // a support never exceeds the child count (each child row adds at most one image
        // to each vertex set)
        if (flags[a] && cr >= (long long)min_sup && ksup) {
            const int s = ksup[akeys[a]];
            out = s >= min_sup ? s : 0;
        } else if (flags[a] && cr >= (long long)min_sup) {
            const int key = akeys[a];
            const PatDev& P = pats[owner_of(pats, npats, key)];
            int nv, lab[GS_MAXV];
            bool v0;
            child_shape(P, key - P.key_base, ks, &nv, &v0, lab);
            if (mni_mode != GS_MNI_OMP_LITERAL) v0 = true;
            long long s = INT_MAX;
            for (int v = 0; v < nv; ++v) {
                if (v == 0 && !v0) continue;
                const int c = counts[a * GS_MAXV + v];
                if (c > 0 && c < s) s = c;
            }
            out = s >= min_sup ? (int)s : 0;
        }
        sup[a] = out;
    }
    for (int off = 32; off > 0; off >>= 1) {
        rows += __shfl_xor(rows, off);
        canon += __shfl_xor(canon, off);
    }
    if (lane_id() == 0) {
        s_t[0][threadIdx.x / 64] = rows;
        s_t[1][threadIdx.x / 64] = canon;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long r = 0, k = 0;
        for (int w = 0; w < GS_BLOCK / 64; ++w) {
            r += s_t[0][w];
            k += s_t[1][w];
        }
        if (r) atomicAdd(&totals[0], r);
        if (k) atomicAdd(&totals[1], k);
    }
}

// Descriptor of every potential child key k < K (see gsk::all_key_sets).
__global__ void k_all_key_sets(const PatDev* __restrict__ pats, int npats, KeySpace ks,
                               long long K, const int* __restrict__ lw,
                               const int* __restrict__ lcum, int kind, int tag, SetDesc* sets) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long key = (long long)blockIdx.x * blockDim.x + threadIdx.x; key < K;
         key += stride) {
        const PatDev& P = pats[owner_of(pats, npats, (int)key)];
        SetDesc d;
        all_key_set(P, (int)(key - P.key_base), ks, lw, lcum, kind, tag, &d);
        sets[key] = d;
    }
}

// The frequent children's keys, counts and first rows, by their active index.
__global__ void k_gather_frequent(const int* __restrict__ fidx, const int* __restrict__ nf,
                                  const int* __restrict__ akeys, const int* __restrict__ acnt,
                                  const long long* __restrict__ row_off, int* fkey, int* fcnt,
                                  long long* frow) {
    const int F = *nf;
    const int stride = gridDim.x * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < F; i += stride) {
        const int a = fidx[i];
        fkey[i] = akeys[a];
        fcnt[i] = acnt[a];
        frow[i] = row_off ? row_off[a] : 0;
    }
}

// A dispatch holds fewer than 2^32 work-items (the grid size of an AQL packet is 32-bit);
// launchers whose grid follows the data check it with this.
inline bool grid_fits(long long blocks, int threads) {
    return blocks > 0 && blocks * (long long)threads < (1LL << 32);
}

inline int grid_for(long long n, int block, int cap) {
    long long g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

}  // namespace

namespace gsk {

hipError_t seed_count(const DevGraph& g, int EL, int VL, int* key_count, long long* total,
                      hipStream_t s) {
    const ItemGuard none{0, 0, 0, 0, nullptr};
    hipLaunchKernelGGL(k_seed<false>, dim3(grid_for(g.nnz, GS_TILE, 4096)), dim3(GS_BLOCK), 0, s,
                       g, EL, VL, key_count, total, nullptr, nullptr, nullptr, nullptr, nullptr,
                       none);
    return hipGetLastError();
}

hipError_t seed_fill(const DevGraph& g, int EL, int VL, const int* key_slot, int* cursor,
                     int2* rows_out, const SetDesc* sets, uint32_t* pool, ItemGuard gd,
                     hipStream_t s) {
    hipLaunchKernelGGL(k_seed<true>, dim3(grid_for(g.nnz, GS_TILE, 4096)), dim3(GS_BLOCK), 0, s,
                       g, EL, VL, nullptr, nullptr, key_slot, cursor, rows_out, sets, pool, gd);
    return hipGetLastError();
}

hipError_t prep(const DevGraph& g, const PatDev* pats, int npats, long long b0, long long R,
                LayerPtrs L, bool runs, RowCtx* ctx, int* work, unsigned long long* partial,
                unsigned long long* work_full, int* guard_err, hipStream_t s) {
    if (R <= 0) return hipSuccess;
    long long grid = (R + GS_BLOCK - 1) / GS_BLOCK;
    if (!grid_fits(grid, GS_BLOCK)) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL(k_prep, dim3((unsigned)grid), dim3(GS_BLOCK), 0, s, g, pats, npats, b0, R,
                       L, runs ? 1 : 0, ctx, work, work_full ? partial : nullptr, guard_err);
    if (work_full)
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(GS_BLOCK), 0, s, partial, grid,
                           work_full);
    return hipGetLastError();
}

long long scan_tmp_elems(long long n) { return (n + SCAN_TILE - 1) / SCAN_TILE + 1; }

hipError_t build_groups(const DevGraph& g, const PatDev* pats, const RowCtx* ctx, long long R,
                        int nseg, unsigned long long* tkeys, int* trep, int* tcnt, int* tgrp,
                        long long slots, int* pslot, int* pflag, long long* pidx,
                        long long* scan_tmp, long long scan_tmp_elems, int* hits, RowCtx* gctx,
                        int* gwork, int* guard_err, hipStream_t s) {
    const long long npairs = R * nseg;
    if (npairs <= 0) return hipSuccess;
    if (nseg < 1 || nseg > GS_MAXE || slots < 2 * npairs || (slots & (slots - 1)) != 0 ||
        slots > (1LL << 31))
        return hipErrorInvalidValue;
    const int gp = grid_for(npairs, 256, 16384), gs = grid_for(slots, 256, 16384);
    hipLaunchKernelGGL(k_fill_u64, dim3(gs), dim3(256), 0, s, tkeys, kEmptyKey, slots);
    hipLaunchKernelGGL(k_fill_i32, dim3(gs), dim3(256), 0, s, tcnt, 0, slots);
    hipLaunchKernelGGL(k_group_claim, dim3(gp), dim3(256), 0, s, g, pats, ctx, R, nseg, tkeys,
                       trep, tcnt, slots, pslot, guard_err);
    hipLaunchKernelGGL(k_group_flags, dim3(gp), dim3(256), 0, s, npairs, nseg, pslot, trep, pflag);
    hipError_t e = scan_exclusive(pflag, pidx, npairs, scan_tmp, scan_tmp_elems, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_group_index, dim3(gp), dim3(256), 0, s, npairs, pslot, pflag, pidx, tgrp);
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(npairs * GS_MAXV, 256, 16384)), dim3(256), 0, s,
                       hits, 0, npairs * GS_MAXV);
    hipLaunchKernelGGL(k_group_hits, dim3(gp), dim3(256), 0, s, pats, ctx, npairs, nseg, pslot,
                       pflag, trep, tgrp, hits);
    hipLaunchKernelGGL(k_group_ctx, dim3(gp), dim3(256), 0, s, g, pats, ctx, npairs, nseg, pslot,
                       pflag, pidx, tcnt, hits, gctx, gwork, guard_err);
    return hipGetLastError();
}

hipError_t scan_exclusive(const int* in, long long* out, long long n, long long* tmp,
                          long long tmp_elems, hipStream_t s) {
    long long nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (nb == 0) {
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, reinterpret_cast<int*>(out), 0, 2);
        return hipGetLastError();
    }
    if (tmp_elems < nb + 1 || !grid_fits(nb, GS_BLOCK)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(GS_BLOCK), 0, s, in, n, tmp);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(GS_BLOCK), 0, s, tmp, nb);
    hipLaunchKernelGGL(k_scan_down, dim3((unsigned)nb), dim3(GS_BLOCK), 0, s, in, n, tmp, nb, out);
    return hipGetLastError();
}

template <int MODE>
hipError_t launch_items(const ItemsLaunch& lc, const DevGraph& g, const PatDev* pats,
                        const RowCtx* ctx, const long long* woff, long long b0, long long R,
                        KeySpace ks, int* key_count, const int* key_slot,
                        unsigned long long* cursor, const int* cap_end, int2* rows_out,
                        const SetDesc* sets, uint32_t* pool, int* overflow, TileQueues& tq,
                        const ItemGuard& gd, int dedup, int saturate, hipStream_t s) {
    int* tile_counters = tq.base + tq.cur * GS_TILE_QUEUE;
    int* next_counters = tq.base + (1 - tq.cur) * GS_TILE_QUEUE;
    if (lc.m < 1 || lc.m >= GS_MAXE) return hipErrorInvalidValue;
    const int pm = items_pm(lc.m);
#define GS_LAUNCH_ITEMS(T, PM)                                                                 \
    hipLaunchKernelGGL((k_items<MODE, T, PM>), dim3(lc.grid), dim3(GS_BLOCK), 0, s, g, pats,  \
                       ctx, woff, b0, R, ks, key_count, key_slot, cursor, cap_end, rows_out,   \
                       sets, pool, overflow, tile_counters, next_counters, gd, dedup, saturate)
#define GS_LAUNCH_PM(T)                                  \
    if (pm == 1) GS_LAUNCH_ITEMS(T, 1);                  \
    else if (pm == 2) GS_LAUNCH_ITEMS(T, 2);             \
    else GS_LAUNCH_ITEMS(T, GS_MAXE - 1)
    switch (lc.tile) {
        case 256: GS_LAUNCH_PM(256); break;
        case 512: GS_LAUNCH_PM(512); break;
        case 1024: GS_LAUNCH_PM(1024); break;
        default: return hipErrorInvalidValue;
    }
#undef GS_LAUNCH_PM
#undef GS_LAUNCH_ITEMS
    const hipError_t e = hipGetLastError();
    if (e == hipSuccess) tq.cur = 1 - tq.cur;
    return e;
}

hipError_t items_count(const ItemsLaunch& lc, const DevGraph& g, const PatDev* pats,
                       const RowCtx* ctx, const long long* woff, long long b0, long long R,
                       KeySpace ks, int* key_count, TileQueues& tq, bool saturate,
                       const ItemGuard& gd, hipStream_t s) {
    if (R <= 0) return hipSuccess;
    return launch_items<GS_ITEMS_COUNT>(lc, g, pats, ctx, woff, b0, R, ks, key_count, nullptr,
                                        nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, tq,
                                        gd, 0, saturate ? 1 : 0, s);
}

hipError_t items_fill(const ItemsLaunch& lc, const DevGraph& g, const PatDev* pats,
                      const RowCtx* ctx, const long long* woff, long long b0, long long R,
                      KeySpace ks, int* key_count, const int* key_slot,
                      unsigned long long* cursor, const int* cap_end, int2* rows_out,
                      const SetDesc* sets, uint32_t* pool, int* overflow, TileQueues& tq,
                      int nlive, bool saturate, const ItemGuard& gd, hipStream_t s,
                      bool by_items) {
    if (R <= 0) return hipSuccess;
    const int dedup = 1;
    (void)nlive;
    if (rows_out || by_items)
        return launch_items<GS_ITEMS_ROWS>(lc, g, pats, ctx, woff, b0, R, ks, key_count, key_slot,
                                           cursor, cap_end, rows_out, sets, pool, overflow, tq,
                                           gd, dedup, saturate ? 1 : 0, s);
    return launch_items<GS_ITEMS_SETS>(lc, g, pats, ctx, woff, b0, R, ks, key_count, key_slot,
                                       cursor, cap_end, nullptr, sets, pool, overflow, tq, gd,
                                       0, saturate ? 1 : 0, s);
}

template <int MODE, int TILE>
hipError_t occupancy_of(int pm, int* per_cu) {
    if (pm == 1)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_items<MODE, TILE, 1>,
                                                            GS_BLOCK, 0);
    if (pm == 2)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_items<MODE, TILE, 2>,
                                                            GS_BLOCK, 0);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_items<MODE, TILE, GS_MAXE - 1>,
                                                        GS_BLOCK, 0);
}

template <int MODE>
hipError_t occupancy_tiles(int tile, int pm, int* per_cu) {
    switch (tile) {
        case 256: return occupancy_of<MODE, 256>(pm, per_cu);
        case 512: return occupancy_of<MODE, 512>(pm, per_cu);
        case 1024: return occupancy_of<MODE, 1024>(pm, per_cu);
        default: return hipErrorInvalidValue;
    }
}

hipError_t items_newv(const ItemsLaunch& lc, const DevGraph& g, const PatDev* pats,
                      const RowCtx* gctx, const long long* gwoff, long long G, KeySpace ks,
                      const int* key_slot, const SetDesc* sets, uint32_t* pool, TileQueues& tq,
                      const ItemGuard& gd, hipStream_t s) {
    if (G <= 0) return hipSuccess;
    return launch_items<GS_ITEMS_NEWV>(lc, g, pats, gctx, gwoff, 0, G, ks, nullptr, key_slot,
                                       nullptr, nullptr, nullptr, sets, pool, nullptr, tq, gd, 0,
                                       0, s);
}

int items_blocks_per_cu(int mode, int tile, int m) {
    int per_cu = 0;
    const int pm = items_pm(m);
    hipError_t e;
    if (mode == GS_ITEMS_COUNT) e = occupancy_tiles<GS_ITEMS_COUNT>(tile, pm, &per_cu);
    else if (mode == GS_ITEMS_SETS) e = occupancy_tiles<GS_ITEMS_SETS>(tile, pm, &per_cu);
    else if (mode == GS_ITEMS_ROWS) e = occupancy_tiles<GS_ITEMS_ROWS>(tile, pm, &per_cu);
    else e = occupancy_tiles<GS_ITEMS_NEWV>(tile, pm, &per_cu);
    if (e != hipSuccess || per_cu <= 0) per_cu = 2;
    return per_cu;
}

hipError_t compact_active_ordered(const int* key_count, long long K, int* active_keys,
                                  int* active_counts, int* n_active, int* clear_word, int* flags,
                                  long long* pos, long long* scan_tmp, long long scan_tmp_elems,
                                  hipStream_t s) {
    if (K <= GS_COMPACT_SMALL) {  // also K == 0: n_active = 0
        hipLaunchKernelGGL(k_compact_small, dim3(1), dim3(GS_BLOCK), 0, s, key_count,
                           (int)(K > 0 ? K : 0), active_keys, active_counts, n_active,
                           clear_word);
        return hipGetLastError();
    }
    if (clear_word) hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, clear_word, 0, 1LL);
    const unsigned nb = (unsigned)grid_for(K, GS_BLOCK, 1 << 20);
    hipLaunchKernelGGL(k_active_flags, dim3(nb), dim3(GS_BLOCK), 0, s, key_count, K, flags);
    hipError_t e = scan_exclusive(flags, pos, K, scan_tmp, scan_tmp_elems, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_active_scatter, dim3(nb), dim3(GS_BLOCK), 0, s, key_count, pos, K,
                       active_keys, active_counts, n_active);
    return hipGetLastError();
}

hipError_t canon(const PatDev* pats, int npats, KeySpace ks, const int* active_keys,
                 const int* n_active, int max_active, int child_m, int* flags, int* err,
                 hipStream_t s) {
    if (max_active <= 0) return hipSuccess;
    if (child_m == 2) {
        hipLaunchKernelGGL(k_canon_two, dim3(grid_for(max_active, GS_BLOCK, 4096)), dim3(GS_BLOCK),
                           0, s, pats, npats, ks, active_keys, n_active, flags);
        return hipGetLastError();
    }
    // one thread per three-edge code pays once the codes outnumber the waves the per-wave
    // walk can run at once; below that a wave per code finishes sooner (1,080 codes: 134 us
    // by threads, about 60 us by waves)
    if (child_m == 3 && max_active >= (1 << 16)) {
        hipLaunchKernelGGL(k_canon_three, dim3(grid_for(max_active, GS_BLOCK, 8192)),
                           dim3(GS_BLOCK), 0, s, pats, npats, ks, active_keys, n_active, flags,
                           err);
        return hipGetLastError();
    }
    int grid = max_active < 16384 ? max_active : 16384;
    if (child_m > 6)
        hipLaunchKernelGGL(k_canon<GS_CANON_CAP7>, dim3(grid), dim3(64), 0, s, pats, npats, ks,
                           active_keys, n_active, flags, err);
    else
        hipLaunchKernelGGL(k_canon<GS_CANON_CAP>, dim3(grid), dim3(64), 0, s, pats, npats, ks,
                           active_keys, n_active, flags, err);
    return hipGetLastError();
}

hipError_t canon_codes(const int* codes, int ncodes, int max_m, int* flags, int* err,
                       hipStream_t s) {
    if (ncodes <= 0) return hipSuccess;
    int grid = ncodes < 16384 ? ncodes : 16384;
    if (max_m > 6)
        hipLaunchKernelGGL(k_canon_codes<GS_CANON_CAP7>, dim3(grid), dim3(64), 0, s, codes, ncodes,
                           flags, err);
    else
        hipLaunchKernelGGL(k_canon_codes<GS_CANON_CAP>, dim3(grid), dim3(64), 0, s, codes, ncodes,
                           flags, err);
    return hipGetLastError();
}

hipError_t fill_i32(int* p, int v, long long n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

hipError_t fill_u32(uint32_t* p, uint32_t v, long long n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n / 4 + 1, 256, 8192)), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

hipError_t scatter_slots(int* key_slot, const int* keys, int nkeys, hipStream_t s) {
    if (nkeys <= 0) return hipSuccess;
    if (!grid_fits((nkeys + 255) / 256, 256)) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL(k_scatter_slots, dim3((nkeys + 255) / 256), dim3(256), 0, s, key_slot, keys,
                       nkeys);
    return hipGetLastError();
}

hipError_t set_counts(const SetDesc* sets, int nsets, const uint32_t* pool,
                      long long pool_words, int* counts, int* guard_err, hipStream_t s,
                      const int* rows, int min_rows) {
    if (nsets <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_set_counts, dim3(grid_for(nsets, GS_BLOCK / 64, 16384)), dim3(GS_BLOCK),
                       0, s, sets, (long long)nsets, pool, pool_words, counts, guard_err, rows,
                       min_rows);
    return hipGetLastError();
}

hipError_t plan_children(const PatDev* pats, int npats, KeySpace ks, const int* akeys,
                         const int* acnt, const int* flags, int na, const int* lcnt, bool bytes,
                         bool plan_sets, bool write_rows, int* set_words, int* rows,
                         hipStream_t s) {
    if (na <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_plan_children, dim3(grid_for(na, 256, 16384)), dim3(256), 0, s, pats,
                       npats, ks, akeys, acnt, flags, na, lcnt, bytes ? 1 : 0, plan_sets ? 1 : 0,
                       write_rows ? 1 : 0, set_words, rows);
    return hipGetLastError();
}

hipError_t build_sets(const PatDev* pats, int npats, KeySpace ks, const int* akeys,
                      const int* acnt, const int* flags, int na, const int* lcnt, bool bytes,
                      const long long* set_off, const long long* row_off,
                      const SetDesc* key_sets, SetDesc* sets, int* key_slot,
                      unsigned long long* cursor, hipStream_t s) {
    if (na <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_build_sets, dim3(grid_for(na, 256, 16384)), dim3(256), 0, s, pats, npats,
                       ks, akeys, acnt, flags, na, lcnt, bytes ? 1 : 0, set_off, row_off, key_sets,
                       sets, key_slot, cursor);
    return hipGetLastError();
}

hipError_t supports(const PatDev* pats, int npats, KeySpace ks, const int* akeys,
                    const int* acnt, const int* flags, int na, const int* counts,
                    const int* ksup, int mni_mode, int min_sup, int* sup,
                    unsigned long long* totals, hipStream_t s) {
    if (na <= 0) return hipSuccess;
    // every block adds its totals with two atomics on the same two words: few blocks
    hipLaunchKernelGGL(k_supports, dim3(grid_for(na, GS_BLOCK, 512)), dim3(GS_BLOCK), 0, s,
                       pats, npats, ks, akeys, acnt, flags, na, counts, ksup, mni_mode, min_sup,
                       sup, totals);
    return hipGetLastError();
}

int two_edge_select_blocks(long long K) { return K <= 0 ? 0 : grid_for(K, GS_BLOCK, 2048); }

hipError_t two_edge_select(const PatDev* pats, int npats, KeySpace ks, long long K,
                           const int* key_count, const int* ksup, int min_sup, int* sel,
                           unsigned long long* part, hipStream_t s) {
    if (K <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_two_edge_select, dim3(two_edge_select_blocks(K)), dim3(GS_BLOCK), 0, s,
                       pats, npats, ks, K, key_count, ksup, min_sup, sel, part);
    return hipGetLastError();
}

hipError_t all_key_sets(const PatDev* pats, int npats, KeySpace ks, long long K, const int* lw,
                        const int* lcum, int kind, int tag, SetDesc* sets, hipStream_t s) {
    if (K <= 0) return hipSuccess;
    if (tag < 1 || tag > 255) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_all_key_sets, dim3(grid_for(K, 256, 16384)), dim3(256), 0, s, pats,
                       npats, ks, K, lw, lcum, kind, tag, sets);
    return hipGetLastError();
}

hipError_t gather_frequent(const int* fidx, const int* nf, int max_f, const int* akeys,
                           const int* acnt, const long long* row_off, int* fkey, int* fcnt,
                           long long* frow, hipStream_t s) {
    if (max_f <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_frequent, dim3(grid_for(max_f, 256, 16384)), dim3(256), 0, s,
                       fidx, nf, akeys, acnt, row_off, fkey, fcnt, frow);
    return hipGetLastError();
}

hipError_t support_rows(const DevGraph& g, const PatDev* pat, long long N, LayerPtrs L,
                        int mni_mode, const SetDesc* set, uint32_t* pool, long long pool_words,
                        int* guard_err, hipStream_t s) {
    (void)mni_mode;
    if (N <= 0) return hipSuccess;
    long long grid = (N + GS_BLOCK - 1) / GS_BLOCK;
    if (!grid_fits(grid, GS_BLOCK)) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL(k_support_rows, dim3((unsigned)grid), dim3(GS_BLOCK), 0, s, g, pat, N, L,
                       set, pool, pool_words, guard_err);
    return hipGetLastError();
}

}  // namespace gsk

// ------------------------------------------------------------------ prune_graph
// graph.cc:216-266: keep the vertices whose label occurs at least min_sup times, renumber
// them in ascending id order, keep the edges between kept vertices in their CSR order.
namespace {

__global__ void k_label_hist(const int* __restrict__ vlab, int n, int* hist) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) atomicAdd(&hist[vlab[v]], 1);
}

__global__ void k_keep(const int* __restrict__ vlab, int n, const int* __restrict__ hist,
                       int min_sup, int* __restrict__ keep) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) keep[v] = hist[vlab[v]] >= min_sup ? 1 : 0;
}

// kept-edge count of every vertex (0 for dropped vertices)
__global__ void k_kept_degree(DevGraph g, const int* __restrict__ keep, int* __restrict__ kdeg) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= g.n) return;
    int c = 0;
    if (keep[v])
        for (int i = g.offsets[v]; i < g.offsets[v + 1]; ++i) c += keep[g.adj[i].x];
    kdeg[v] = c;
}

__global__ void k_prune_scatter(DevGraph g, const int* __restrict__ keep,
                                const long long* __restrict__ newid,
                                const long long* __restrict__ newoff,
                                const int* __restrict__ vl_vals, const int* __restrict__ el_vals,
                                int* __restrict__ out_offsets, int* __restrict__ out_vlabels,
                                int* __restrict__ out_edges) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= g.n || !keep[v]) return;
    const int nv = (int)newid[v];
    out_vlabels[nv] = vl_vals[g.vlab[v]];
    out_offsets[nv + 1] = (int)newoff[v + 1];
    long long at = newoff[v];
    for (int i = g.offsets[v]; i < g.offsets[v + 1]; ++i) {
        const uint2 a = g.adj[i];
        if (!keep[a.x]) continue;
        out_edges[3 * at + 0] = nv;
        out_edges[3 * at + 1] = (int)newid[a.x];
        out_edges[3 * at + 2] = el_vals[a.y >> 16];
        ++at;
    }
}

}  // namespace

namespace gsk {

hipError_t prune_graph(const DevGraph& g, int VL, int min_sup, const int* vl_vals,
                       const int* el_vals, int* hist, int* keep, int* kdeg, long long* newid,
                       long long* newoff, long long* scan_tmp, long long scan_tmp_elems,
                       int* out_offsets, int* out_vlabels, int* out_edges, hipStream_t s) {
    const unsigned nb = (unsigned)((g.n + 255) / 256);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(256), 0, s, hist, 0, (long long)VL);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, out_offsets, 0, 1LL);
    if (g.n == 0) return hipGetLastError();
    hipLaunchKernelGGL(k_label_hist, dim3(nb), dim3(256), 0, s, g.vlab, g.n, hist);
    hipLaunchKernelGGL(k_keep, dim3(nb), dim3(256), 0, s, g.vlab, g.n, hist, min_sup, keep);
    hipError_t e = scan_exclusive(keep, newid, g.n, scan_tmp, scan_tmp_elems, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_kept_degree, dim3(nb), dim3(256), 0, s, g, keep, kdeg);
    e = scan_exclusive(kdeg, newoff, g.n, scan_tmp, scan_tmp_elems, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_prune_scatter, dim3(nb), dim3(256), 0, s, g, keep, newid, newoff, vl_vals,
                       el_vals, out_offsets, out_vlabels, out_edges);
    return hipGetLastError();
}

}  // namespace gsk
