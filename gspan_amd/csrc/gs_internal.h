// gs_internal.h -- types shared by the gfx950 kernels (gs_kernels.hip) and the host
// driver (gs_miner.cpp).  Not part of the public C-ABI (include/gspan_amd.h).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "gspan_amd.h"

// Largest pattern the kernels handle: 7 edges, 8 vertices.  Projections in the
// canonicality kernel pack 8 three-bit vertex ids and an 8-bit used-edge mask into 32 bits.
#define GS_MAXE 7
#define GS_MAXV 8

// Threads per workgroup of the item kernels; items per workgroup tile.
#define GS_BLOCK 256
#define GS_TILE 1024

// Work counters of the two-edge run pass: slot t holds {runs, words} at work[2 t].
#define GS_TE_WORK_SLOTS 64

// XCD-aware tile queues of the item kernels: one counter per XCD, 16 ints apart.
#define GS_XCDS 8
#define GS_XCD_STRIDE 16

// Canonicality kernel: one wave per candidate, projections double-buffered in LDS.  The
// walk of a code with m edges carries at most as many projections as a connected subgraph
// of the pattern has embeddings in it; the most is 7!/(7-i)! for a single-label 7-edge star
// (5040), 720 for 6 edges.  Codes of up to 6 edges use the small buffers, 7-edge codes the
// large ones.
#define GS_CANON_CAP 2048
#define GS_CANON_CAP7 5120

// Child-count predictor: MLP GS_FEATURES -> GS_HIDDEN -> GS_HIDDEN -> 1.
#define GS_FEATURES 32
#define GS_HIDDEN 128
#define GS_MLP_PARAM_COUNT \
    (GS_FEATURES * GS_HIDDEN + GS_HIDDEN + GS_HIDDEN * GS_HIDDEN + GS_HIDDEN + GS_HIDDEN + 1)
static_assert(GS_MLP_PARAMS == GS_MLP_PARAM_COUNT, "parameter count of the public header");

// Set kinds for MNI image sets: one bit per graph vertex (inserted with an atomic OR),
// an open-addressing table (compare-and-swap), or one byte per graph vertex (a plain store
// of 1: every writer stores the same value, so racing stores need no read-modify-write).
#define GS_SET_BITMAP 0
#define GS_SET_HASH 1
#define GS_SET_BYTEMAP 2

// One pattern of a batch, as the extension kernels see it.  Labels are dense ranks
// (order-preserving), so every comparison of the reference's rules is unchanged.
struct PatDev {
    int m;          // edges
    int nv;         // vertices (= v_n + 1)
    int r;          // rightmost-path length
    int v_n;        // rightmost vertex
    int min_label;  // code[0].from_label rank (gspan.cu:618)
    int key_base;   // first child key of this pattern in the batch key space
    int cfrom[GS_MAXE], cto[GS_MAXE], cel[GS_MAXE];
    int vlab[GS_MAXV];  // label rank of every pattern vertex
    int rm[GS_MAXE];    // rightmost path (code edge indices, last forward edge first)
    // Rule constants of the item kernels.  Forward segment j scans the neighbours of
    // pattern vertex seg_v[j] (j = 0: v_n; j >= 1: rmpath[j-1].from); an entry (el, tl)
    // of segment j passes when seg_el[j] < el || (seg_el[j] == el && seg_tl[j] <= tl)
    // (gspan.cu:476-477; seg_el[0] = -1 passes everything).  Backward probe i (1..r-1)
    // looks for an entry of v_n to bw_v[i] = rmpath[i].from that passes
    // bw_el[i] < el || (bw_el[i] == el && bw_le[i]) (gspan.cu:436-440).
    int seg_v[GS_MAXE], seg_el[GS_MAXE], seg_tl[GS_MAXE];
    int bw_v[GS_MAXE], bw_el[GS_MAXE], bw_le[GS_MAXE];
    int support;           // MNI support of the pattern
    int pad0;
    long long row_begin;   // first row in the parent level array
    long long row_count;   // rows of this pattern
    long long brow_begin;  // first row in the batch row space
    long long set_base;    // first pool word of its children's sets (sets of every key)
};

// Per parent row context written by the prep kernel and read by the item kernels.
// phi: image of every pattern vertex; fcum[j]: forward items through segment j
// (segment 0 = neighbours of u_n, segment i+1 = neighbours of rmpath[i].from); the CSR
// entry of forward item f in segment j is base[j] + f.
struct RowCtx {
    int phi[GS_MAXV];
    int fcum[GS_MAXE];
    int p;  // pattern index in the batch
    int base[GS_MAXE];  // index in DevGraph::adjl
    int pad;
};
static_assert(sizeof(RowCtx) == 96, "RowCtx is six 16-byte loads");

struct LayerPtrs {
    const int2* layer[GS_MAXE];  // layer[l] = rows of code edge l: {prev, eid}
    long long rows[GS_MAXE];     // rows in layer[l] (index guard)
};

// Index guards of the kernels.  Every index a kernel derives from data (child keys, slots,
// row reservations, set words, vertex ids, CSR entries, layer rows) is compared with the
// extent of the buffer it addresses; a violation sets a bit of the context's guard word and
// the access is skipped, so an inconsistent launch is reported as GS_ERR_INTERNAL instead
// of touching memory outside its buffers.
#define GS_GUARD_KEY 1      // child key outside the key space
#define GS_GUARD_SLOT 2     // live slot outside the set / cursor arrays
#define GS_GUARD_ROW 4      // row reservation past the row array
#define GS_GUARD_POOL 8     // set words outside the set pool
#define GS_GUARD_GRAPH 16   // vertex id or CSR entry outside the graph
#define GS_GUARD_LAYER 32   // layer row outside its layer
struct ItemGuard {
    long long keys;        // child key space (key_count, and key_slot / cap_end when no slots)
    long long slots;       // entries of sets, cursor and cap_end
    long long rows_cap;    // entries of rows_out
    long long pool_words;  // u32 words of the set pool
    int* err;              // guard word (bits GS_GUARD_*)
};

// MNI image sets of one child key: one set per child vertex.  Every image of child vertex v
// carries v's label, so v's set covers only the graph vertices of that label and is
// indexed by a vertex's rank among them (DevGraph::lidx): a bitmap or bytemap of |V_l|
// entries, or an open-addressing table of 2 x min(rows, |V_l|) slots when that is
// smaller.  Sets are 16-byte aligned in the pool (word offsets are multiples of 4).
struct SetDesc {
    long long off;        // first u32 word of the key's sets in the pool
    int nv;               // vertex sets (vertices of the child pattern); 0 = no sets
    int kinds;            // kind of vertex set v (GS_SET_*) in bits [2v, 2v + 2)
    int vo[GS_MAXV + 1];  // vertex set v: words [off + vo[v], off + vo[v + 1]); vo[nv] = all
    // byte a bytemap insert stores, 1..255: a bytemap holds exactly the bytes equal to it,
    // so a pool reused with a new tag needs no clearing (bytes of older tags read as absent)
    int tag;
    int pad;
};
static_assert(sizeof(SetDesc) == 64, "SetDesc is four 16-byte loads");

__host__ __device__ inline int set_kind(const SetDesc& d, int v) { return (d.kinds >> (2 * v)) & 3; }

// Kind and words of one vertex set: `rows` inserts at most, over `ent` label-class
// vertices.  A hash table of >= 2 * min(rows, ent) slots (16 at least, a power of two) when
// that is smaller than a bitmap, else a bytemap (bytes) or a bitmap; 16-byte padded.
__host__ __device__ inline void plan_vertex_set(long long rows, long long ent, bool bytes,
                                                int* kind, long long* words) {
    const long long bw = ((ent + 31) / 32 + 3) / 4 * 4;
    long long cap = 16;
    const long long want = 2 * (rows < ent ? rows : ent);
    while (cap < want) cap <<= 1;
    if (cap < bw) {
        *kind = GS_SET_HASH;
        *words = cap;
    } else if (bytes) {
        *kind = GS_SET_BYTEMAP;
        *words = (ent + 15) / 16 * 4;
    } else {
        *kind = GS_SET_BITMAP;
        *words = bw;
    }
}

// Sets of a child with nv vertices of label ranks lab[0 .. nv) and `rows` rows; lcnt[l] =
// graph vertices of label rank l.  Returns the key's words (a multiple of 4).
__host__ __device__ inline long long plan_key_sets(long long rows, int nv, const int* lab,
                                                   const int* lcnt, bool bytes, SetDesc* d) {
    long long w = 0;
    d->nv = nv;
    d->kinds = 0;
    for (int v = 0; v <= GS_MAXV; ++v) {
        if (v < nv) {
            int kind;
            long long words;
            plan_vertex_set(rows, lcnt[lab[v]], bytes, &kind, &words);
            d->vo[v] = (int)w;
            d->kinds |= kind << (2 * v);
            w += words;
        } else {
            d->vo[v] = (int)w;
        }
    }
    d->tag = 1;
    d->pad = 0;
    return w;
}

// Batch-wide key decode: key = key_base[p] + slot;
//   slot <  (r-1)*EL                    : backward, rmpath index i = slot/EL + 1, el = slot%EL
//   slot >= (r-1)*EL, s' = slot-(r-1)*EL: forward, j = s'/(EL*VL), el = (s'/VL)%EL, tl = s'%VL
// j = 0: from v_n (first forward); j = i+1: from code[rm[i]].from.
struct KeySpace {
    int EL, VL;
};

// Sets of potential child key `slot` of pattern P when every key of a batch has sets
// (single-pass last level): bitmaps or bytemaps (kind) over the label classes of the
// child's vertices, lw[l] words for label rank l, lcum[l] = lw[0] + ... + lw[l-1].  The
// pattern's keys start at P.set_base: its backward keys (the parent's vertex sets), then
// per (segment, edge label) one key per label rank (the parent's sets plus the new vertex's).
// Batch index of the single-edge parent owning child key `key`: every single-edge pattern
// has the same 2 * EL * VL child keys (no backward ones), laid out in batch order, so the
// index is a division, checked against the pattern's key_base (otherwise searched).
__host__ __device__ inline int owner_one_edge(const PatDev* pats, int npats, KeySpace ks,
                                              int key) {
    const int kp = 2 * ks.EL * ks.VL;
    const int p = key / kp;
    if (p < npats && pats[p].key_base <= key && key - pats[p].key_base < kp) return p;
    int lo = 0, hi = npats - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pats[mid].key_base <= key) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__host__ __device__ inline void all_key_set(const PatDev& P, int slot, KeySpace ks, const int* lw,
                                            const int* lcum, int kind, int tag, SetDesc* d) {
    const int nb = (P.r - 1) * ks.EL;
    const bool fwd = slot >= nb;
    const int tl = fwd ? (slot - nb) % ks.VL : 0;
    d->nv = P.nv + (fwd ? 1 : 0);
    d->kinds = 0;
    int w = 0;
    for (int v = 0; v <= GS_MAXV; ++v) {
        d->vo[v] = w;
        if (v < d->nv) {
            d->kinds |= kind << (2 * v);
            w += lw[v < P.nv ? P.vlab[v] : tl];
        }
    }
    const long long pw = d->vo[P.nv];  // words of the parent's vertex sets
    long long off = P.set_base;
    if (!fwd) {
        off += (long long)slot * pw;
    } else {
        const int grp = (slot - nb) / ks.VL;  // segment * EL + edge label
        off += (long long)nb * pw + (long long)grp * ((long long)ks.VL * pw + lcum[ks.VL]) +
               (long long)tl * pw + lcum[tl];
    }
    d->off = off;
    d->tag = tag;
    d->pad = 0;
}

struct DevGraph {
    int n, nnz;
    const int* offsets;  // n+1
    const int* src;      // nnz: Edge.from
    const uint2* adj;    // nnz: x = to, y = (el_rank << 16) | tl_rank
    const int* vlab;     // n
    // the same entries with each vertex's range ordered by neighbour label (tl_rank), and
    // the CSR index of each: forward segments start at the first neighbour whose label
    // reaches the pattern's min_label, so entries the rules reject by label are never
    // visited
    const uint2* adjl;
    const int* eidl;
    // lstart[v * lstride + l]: first adjl index of v whose neighbour label is >= l
    // (l = 0 .. VL); null when the table would be too large (then binary search)
    const int* lstart;
    int lstride;
    // label classes: lidx[v] = rank of v among the vertices of its label (ascending id),
    // adjlx[i] = lidx[adjl[i].x], lcnt[l] = vertices of label rank l (VL entries)
    const int* lidx;
    const int* adjlx;
    const int* lcnt;
    // runs: maximal ranges of a vertex's adjl entries with one (neighbour label, edge
    // label); run k covers adjl[rstart[k], rstart[k + 1]) (rstart[nruns] = nnz), rkey[k] =
    // (label << 16) | edge label, runid[i] = the run of adjl entry i.  Within a run the
    // entries ascend by neighbour id (the CSR order).
    const int* rstart;
    const int* rkey;
    const int* runid;
    int nruns;
    // 1 when no vertex has a self-loop or two edges to one neighbour (every adjacency
    // ascends strictly by neighbour): the only edge between a row's two vertices joined by a
    // pattern edge is that edge
    int simple;
};

// MLP parameters on the device: w32 holds every parameter in fp32 (W1[32x128], b1, W2[128x128],
// b2, W3[128], b3); w16_all a bf16 copy in the same layout.  Layers read bf16 weights and
// fp32 biases.
struct MlpDev {
    float* w32;
    uint16_t* w16_all;
    const uint16_t* w16[3];
    const float* b[3];
};

// ------------------------------------------------------------ kernel launchers
// All launches go to `stream`; every size argument is checked by the caller.
namespace gsk {

hipError_t seed_count(const DevGraph& g, int EL, int VL, int* key_count, long long* total,
                      hipStream_t s);
hipError_t seed_fill(const DevGraph& g, int EL, int VL, const int* key_slot, int* cursor,
                     int2* rows_out, const SetDesc* sets, uint32_t* pool, ItemGuard gd,
                     hipStream_t s);

// work[row] >= 1 items per row; *work_full (nullable) += the rows' items without the
// label filter, (r - 1) + sum of segment degrees (the reference scans every neighbour);
// partial: ceil(R / GS_BLOCK) elements of scratch for that sum
// runs: a row's forward work counts runs (COUNT / SETS modes), else neighbours
hipError_t prep(const DevGraph& g, const PatDev* pats, int npats, long long b0, long long R,
                LayerPtrs L, bool runs, RowCtx* ctx, int* work, unsigned long long* partial,
                unsigned long long* work_full, int* guard_err, hipStream_t s);
// Groups of a chunk's rows for NEWV (after prep of the chunk, whose contexts give every row's
// images): each (row, forward segment j <= r) pair with candidate neighbours joins the group
// of (pattern, j, segment vertex), found in an open-addressing table of `slots` 64-bit keys
// (a power of two >= 2 x pairs).  A group's context (gctx, ordered by the rows that claimed
// their groups) holds in phi only the vertices that every row of the group has among its
// images (the others -1): a neighbour that one row excludes as its own vertex is still a
// child of another row of the group unless every row excludes it.  gwork[g] = candidate
// neighbours of the group's segment vertex.  Scratch: pairs = R x nseg entries of pslot,
// pflag, pidx (long long, pairs + 1) and the scan's; hits: 8 ints per pair.  The group count
// G is left in pidx[pairs].
hipError_t build_groups(const DevGraph& g, const PatDev* pats, const RowCtx* ctx, long long R,
                        int nseg, unsigned long long* tkeys, int* trep, int* tcnt, int* tgrp,
                        long long slots, int* pslot, int* pflag, long long* pidx,
                        long long* scan_tmp, long long scan_tmp_elems, int* hits, RowCtx* gctx,
                        int* gwork, int* guard_err, hipStream_t s);
hipError_t scan_exclusive(const int* in, long long* out, long long n, long long* tmp,
                          long long tmp_elems, hipStream_t s);
long long scan_tmp_elems(long long n);

// Launch shape of the extension kernels: items per stolen tile (256, 512, 1024) and
// persistent workgroups.
struct ItemsLaunch {
    int tile;
    int grid;
    int m;  // edges of the parent patterns (the kernel is specialised by size class)
};
// Size classes of the extension kernel: parents of 1 edge, 2 edges, up to GS_MAXE - 1.
inline int items_pm(int m) { return m <= 1 ? 1 : m == 2 ? 2 : GS_MAXE - 1; }
inline int items_pm_class(int m) { return m <= 1 ? 0 : m == 2 ? 1 : 2; }
#define GS_PM_CLASSES 3

// Modes of the extension kernel.  COUNT and SETS take one work item per backward probe and
// per forward *run* of a parent row (a run's children share a key, so a run adds its length
// less the row's own vertices among its neighbours): COUNT counts children per key, SETS
// (the last level, no rows) also inserts each parent row's images into the sets of the
// keys it has children in (and counts, in the single-pass last level).  ROWS takes one
// item per candidate neighbour and writes child rows and every image (levels below the
// last).  NEWV takes one item per neighbour of a *group* -- the rows of one pattern whose
// forward segment j starts at the same vertex -- and inserts the new vertex's image once
// per group instead of once per row (the last level).
#define GS_ITEMS_COUNT 0
#define GS_ITEMS_SETS 1
#define GS_ITEMS_ROWS 2
#define GS_ITEMS_NEWV 3
#define GS_ITEMS_MODES 4

// Tile queues of the item kernels: two sets of GS_TILE_QUEUE ints (one counter per XCD,
// GS_XCD_STRIDE apart) used in turn.  A launch takes set `cur` and zeroes the other one for
// the next launch (launches on one stream do not overlap), so no separate reset launch is
// needed; the launchers flip `cur`.  Both sets must be zero before the first launch.
#define GS_TILE_QUEUE (GS_XCDS * GS_XCD_STRIDE)
struct TileQueues {
    int* base;  // 2 * GS_TILE_QUEUE ints
    int cur;
};
hipError_t items_count(const ItemsLaunch& lc, const DevGraph& g, const PatDev* pats,
                       const RowCtx* ctx, const long long* woff, long long b0, long long R,
                       KeySpace ks, int* key_count, TileQueues& tq, bool saturate,
                       const ItemGuard& gd, hipStream_t s);
// saturate: the launches adding to key_count may process more than INT_MAX items in all;
// a count that passes INT_MAX then reads INT_MIN (otherwise counts are exact 32-bit sums).
// key_count (nullable): also count children per key; key_slot == nullptr: slot = key.
// cursor[s]: next row of slot s (64-bit, so an overrun cannot wrap).  cap_end (nullable):
// rows of slot s may only be written below cap_end[s]; a child that does not fit, or a hash
// set that fills up, sets *overflow instead and writes nothing (predicted sizes).
// rows_out == nullptr: sets only -- by runs (GS_ITEMS_SETS: parent images; the new vertices'
// come from items_newv) or, by_items, by neighbours (GS_ITEMS_ROWS writing no rows: every
// image) -- else rows and sets (GS_ITEMS_ROWS).
hipError_t items_fill(const ItemsLaunch& lc, const DevGraph& g, const PatDev* pats,
                      const RowCtx* ctx, const long long* woff, long long b0, long long R,
                      KeySpace ks, int* key_count, const int* key_slot,
                      unsigned long long* cursor, const int* cap_end, int2* rows_out,
                      const SetDesc* sets, uint32_t* pool, int* overflow, TileQueues& tq,
                      int nlive, bool saturate, const ItemGuard& gd, hipStream_t s,
                      bool by_items = false);
// resident workgroups per CU of the extension kernel in a mode (GS_ITEMS_*) at a tile size
// NEWV over G groups (contexts and scanned work from build_groups; trailing groups of no
// work are allowed)
hipError_t items_newv(const ItemsLaunch& lc, const DevGraph& g, const PatDev* pats,
                      const RowCtx* gctx, const long long* gwoff, long long G, KeySpace ks,
                      const int* key_slot, const SetDesc* sets, uint32_t* pool, TileQueues& tq,
                      const ItemGuard& gd, hipStream_t s);
int items_blocks_per_cu(int mode, int tile, int m);

// Keys with a non-zero count, in ascending key order, with their counts (flags / pos: K
// and K+1 elements; scan_tmp as for scan_exclusive over K); *n_active is written and
// *clear_word (nullable) zeroed.
hipError_t compact_active_ordered(const int* key_count, long long K, int* active_keys,
                                  int* active_counts, int* n_active, int* clear_word, int* flags,
                                  long long* pos, long long* scan_tmp, long long scan_tmp_elems,
                                  hipStream_t s);
// child_m: edges of the candidate codes (parent size + 1); max_m: longest code given
hipError_t canon(const PatDev* pats, int npats, KeySpace ks, const int* active_keys,
                 const int* n_active, int max_active, int child_m, int* flags, int* err,
                 hipStream_t s);
hipError_t canon_codes(const int* codes, int ncodes, int max_m, int* flags, int* err,
                       hipStream_t s);

// Device scratch of gs_is_canonical_batch_noalloc (gs_noalloc.hip): looks up the address
// of its static device array on `device` (current) once; called by gs_create.
#define GS_NOALLOC_SCRATCH_BYTES (5 << 19)
hipError_t noalloc_init(int device);
// context fields for gs_noalloc.hip (gs_ctx is defined in gs_miner.cpp)
int ctx_device(const gs_ctx* c);
hipStream_t ctx_stream(const gs_ctx* c);
void ctx_set_error(gs_ctx* c, const char* msg);

// Step tail on the device (steps with many candidate keys), over the active keys a < na
// of an ordered compaction (keys, counts) and their canonicality flags (lcnt: graph
// vertices per label rank, for the label-class sets):
// plan_children: set words (plan_sets) and rows (write_rows) of every canonical child;
// build_sets: its set descriptor (planned from the scanned words, or key_sets[key] when
// the sets were planned per key), slot = a in key_slot, first row in cursor;
// supports: sup[a] = MNI support when >= min_sup, else 0; totals[0] += child rows,
// totals[1] += canonical keys;
// gather_frequent: keys, counts and first rows of the frequent children (fidx, *nf).
hipError_t plan_children(const PatDev* pats, int npats, KeySpace ks, const int* akeys,
                         const int* acnt, const int* flags, int na, const int* lcnt, bool bytes,
                         bool plan_sets, bool write_rows, int* set_words, int* rows,
                         hipStream_t s);
hipError_t build_sets(const PatDev* pats, int npats, KeySpace ks, const int* akeys,
                      const int* acnt, const int* flags, int na, const int* lcnt, bool bytes,
                      const long long* set_off, const long long* row_off,
                      const SetDesc* key_sets, SetDesc* sets, int* key_slot,
                      unsigned long long* cursor, hipStream_t s);
// A step over single-edge parents with the two-edge pass's key_count and ksup: sel[key] =
// the support of a canonical child whose count and support reach min_sup, else 0; part:
// per workgroup (two_edge_select_blocks(K) of them) its child rows, keys with children and
// canonical keys among them, in that order.
int two_edge_select_blocks(long long K);
hipError_t two_edge_select(const PatDev* pats, int npats, KeySpace ks, long long K,
                           const int* key_count, const int* ksup, int min_sup, int* sel,
                           unsigned long long* part, hipStream_t s);
hipError_t supports(const PatDev* pats, int npats, KeySpace ks, const int* akeys,
                    const int* acnt, const int* flags, int na, const int* counts,
                    const int* ksup, int mni_mode, int min_sup, int* sup,
                    unsigned long long* totals, hipStream_t s);
// Sets of every potential child key of a batch (single-pass last level): each key's vertex
// sets are bitmaps or bytemaps over their label classes, lw[l] words for label rank l and
// lcum[l] = lw[0] + ... + lw[l-1] (VL + 1 entries); a pattern's keys start at its set_base:
// its backward keys (parent sets), then per (segment, edge label) one key per label rank.
hipError_t all_key_sets(const PatDev* pats, int npats, KeySpace ks, long long K, const int* lw,
                        const int* lcum, int kind, int tag, SetDesc* sets, hipStream_t s);
hipError_t gather_frequent(const int* fidx, const int* nf, int max_f, const int* akeys,
                           const int* acnt, const long long* row_off, int* fkey, int* fcnt,
                           long long* frow, hipStream_t s);

hipError_t fill_i32(int* p, int v, long long n, hipStream_t s);
hipError_t fill_u32(uint32_t* p, uint32_t v, long long n, hipStream_t s);
hipError_t scatter_slots(int* key_slot, const int* keys, int nkeys, hipStream_t s);
// rows (nullable): the child count of each set's key (k_items' counts); sets of keys with
// fewer than min_rows children are not read (their counts read 0): a support never exceeds
// the child count, so such a key is infrequent at min_sup <= min_rows.
hipError_t set_counts(const SetDesc* sets, int nsets, const uint32_t* pool,
                      long long pool_words, int* counts, int* guard_err, hipStream_t s,
                      const int* rows = nullptr, int min_rows = 0);
// Two-edge last level (gs_twoedge.hip).  te_prepare: run masks m1 / m2 (W words per
// vertex, zeroed here) and the (vertex, run) pairs ordered by class (vertex label * NB +
// run bit), two int4 records per pair and their {class, run length} (pcl), with hist (VL * NB
// ints), cstart (VL * NB + 1: the class starts), rbl ({bit, length} of every run) and scan
// scratch.  te_clear: zeroes the count tables (CNT, CEN: 2 * npats * NB entries, two rows
// per single-edge parent; END: VL * NB * NB entries) and the queue.  te_level: the count
// tables from the pairs, then key_count and the child's support ksup (mni_mode rules;
// INT_MAX when no image count is non-zero) for the batch's K keys; p_of: batch index of
// each single-edge parent (a, e0, b), a <= b, at (a * EL + e0) * VL + b, or -1;
// pmask[l * W + w]: word w of the run bits (l', e0) with (l', e0, l) a batch parent (either
// label order); queue: one int; work[2 t], work[2 t + 1] += runs visited, mask words read
// (the sums over the slots t < GS_TE_WORK_SLOTS are the pass's).  te_gather:
// out[a] = ksup[akeys[a]] for a < *n_active (n_active null: a < max_active).
hipError_t te_prepare(const DevGraph& g, int EL, int NB, int W, uint32_t* m1, uint32_t* m2,
                      int* hist, long long* cstart, long long* scan_tmp, long long scan_elems,
                      int4* rec, int2* pcl, int2* rbl, hipStream_t s);
// per class c < NC: npairs[c] = its pairs, nlen[c] = CSR entries of its runs
hipError_t te_classes(const int2* pcl, long long np, const long long* cstart, long long NC,
                      int* npairs, unsigned long long* nlen, hipStream_t s);
hipError_t te_clear(KeySpace ks, int npats, int* queue, unsigned long long* CNT, int* CEN,
                    int* END, hipStream_t s);
// The single-edge codes of the class sums on the device (flags[i] for code i = (a * EL +
// e0) * VL + b: a <= b, rows, support >= min_sup under mni_mode), then p_of and pmask as
// te_level reads them (p_of[i]: index among the flagged codes in ascending i, or -1) and
// *ncodes = the flagged codes; scan: VL * EL * VL long longs.
hipError_t te_seeds(const int* npairs, const unsigned long long* nlen, int VL, int EL,
                    int mni_mode, int min_sup, int* flags, long long* scan, long long* scan_tmp,
                    long long scan_elems, int* p_of, uint32_t* pmask, int* ncodes,
                    hipStream_t s);
// te_clear for *npats single-edge parents (tables sized for max_pats), count on the device
hipError_t te_clear_dev(KeySpace ks, const int* npats, long long max_pats, int* queue,
                        unsigned long long* CNT, int* CEN, int* END, hipStream_t s);
// te_level's two launches: the tables from the pairs (te_accum), then the keys (te_keys)
hipError_t te_accum(const DevGraph& g, KeySpace ks, const int* p_of, const uint32_t* m1,
                    const uint32_t* m2, const int4* rec, const int2* rbl, const uint32_t* pmask,
                    int* queue, unsigned long long* CNT, int* CEN, int* END,
                    unsigned long long* work, hipStream_t s);
hipError_t te_keys(const PatDev* pats, int npats, KeySpace ks, long long K,
                   const unsigned long long* CNT, const int* CEN, const int* END, int mni_mode,
                   int* key_count, int* ksup, hipStream_t s);
hipError_t te_level(const DevGraph& g, KeySpace ks, const PatDev* pats, int npats, long long K,
                    const int* p_of, const uint32_t* m1, const uint32_t* m2, const int4* rec,
                    const int2* rbl, const uint32_t* pmask, int* queue,
                    unsigned long long* CNT, int* CEN, int* END, int mni_mode, int* key_count,
                    int* ksup, unsigned long long* work, hipStream_t s);
hipError_t te_gather(const int* ksup, const int* akeys, const int* n_active, int max_active,
                     int* out, hipStream_t s);
// The reports of a two-edge last level over every frequent single-edge code (pats, in key
// order), in the reference's order: each code (support ssup[p]) and then its frequent
// children (fkey ascending, supports fsup); codes5 / lens / sups of npats + F reports and
// npats + 2 F edges, original labels from vl_vals / el_vals.
hipError_t te_emit(const PatDev* pats, int npats, KeySpace ks, const int* fkey, const int* fsup,
                   int F, const int* ssup, const int* vl_vals, const int* el_vals, int* codes,
                   int* lens, int* sups, hipStream_t s);
hipError_t support_rows(const DevGraph& g, const PatDev* pat, long long N, LayerPtrs L,
                        int mni_mode, const SetDesc* set, uint32_t* pool,
                        long long pool_words, int* guard_err, hipStream_t s);

// graph.cc:216-266 on the device graph; newid/newoff are exclusive scans (n+1 entries)
hipError_t prune_graph(const DevGraph& g, int VL, int min_sup, const int* vl_vals,
                       const int* el_vals, int* hist, int* keep, int* kdeg, long long* newid,
                       long long* newoff, long long* scan_tmp, long long scan_tmp_elems,
                       int* out_offsets, int* out_vlabels, int* out_edges, hipStream_t s);

// predictor
hipError_t features(const PatDev* pats, int npats, KeySpace ks, long long K, const float* H,
                    const float* vfreq, const float* efreq, const float* dmean, float nverts,
                    float* X, unsigned char* possible, hipStream_t s);
hipError_t log_targets(const int* counts, long long K, float* t, hipStream_t s);
hipError_t mlp_forward(const float* X, long long N, const MlpDev& w, float* h1, float* h2,
                       float* y, hipStream_t s);
hipError_t mlp_backward(const float* X, const float* t, long long N, const MlpDev& w,
                        const float* h1, const float* h2, const float* y, float* dy, float* dh2,
                        float* dh1, float* grads, float* loss, hipStream_t s);
hipError_t mlp_adam(MlpDev& w, float* grads, float* mo, float* ve, int step, float lr,
                    hipStream_t s);
hipError_t mlp_sync_bf16(MlpDev& w, hipStream_t s);

}  // namespace gsk
