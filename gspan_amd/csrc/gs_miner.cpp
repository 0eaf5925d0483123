// gs_miner.cpp -- host side of the MI355X gSpan path: device residency of the graph,
// the level driver, and the C-ABI of include/gspan_amd.h.
//
// Driver (replaces mine/fsm, gspan.cu:1118-1199).  A *batch* is a set of frequent,
// canonical patterns of one size m whose rows live in one device level array.  One
// batch step runs, on the device:
//   prep -> scan -> items_count          child rows per child key (two-pass count,
//                                        gspan.cu:737-755)
//   compact_active -> canon              candidate codes and their canonicality
//   items_fill -> set_counts             child rows grouped by key + MNI image sets
// with two host synchronisations (active keys; set sizes).  GS_MODE_BATCHED feeds every
// frequent child of a step into the next step (one step per DFS level unless the device
// budget forces a split); GS_MODE_PLAIN runs one step per parent pattern in the
// reference's DFS order.  Results are identical; reports are delivered in the reference's
// order (gspan.cu:1142-1168) in both modes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <thread>
#include <vector>

#include "gs_internal.h"
#include "gs_tuner.h"
#include "gspan_amd.h"

namespace {

struct GsError : std::runtime_error {
    int code;
    GsError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHECK(expr)                                                                   \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            throw GsError(GS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// ------------------------------------------------------------------ device memory
// Size-class caching pool: every device buffer of the miner comes from here, so the
// per-batch allocations of the driver do not reach hipMalloc after warm-up.
class DevPool {
   public:
    ~DevPool() { release_all(); }
    void* get(size_t bytes) {
        size_t cls = size_class(bytes);
        auto it = free_.find(cls);
        if (it != free_.end() && !it->second.empty()) {
            void* p = it->second.back();
            it->second.pop_back();
            cached_ -= cls;
            in_use_ += cls;
            peak_ = std::max(peak_, in_use_);
            sizes_[p] = cls;
            return p;
        }
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, cls);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            trim();
            e = hipMalloc(&p, cls);
            if (e != hipSuccess)
                throw GsError(GS_ERR_HIP, std::string("hipMalloc(") + std::to_string(cls) +
                                              "): " + hipGetErrorString(e));
        }
        in_use_ += cls;
        peak_ = std::max(peak_, in_use_);
        sizes_[p] = cls;
        return p;
    }
    void put(void* p) {
        if (!p) return;
        auto it = sizes_.find(p);
        if (it == sizes_.end()) return;
        size_t cls = it->second;
        sizes_.erase(it);
        in_use_ -= cls;
        cached_ += cls;
        free_[cls].push_back(p);
    }
    void trim() {
        for (auto& kv : free_)
            for (void* p : kv.second) (void)hipFree(p);
        free_.clear();
        cached_ = 0;
    }
    void release_all() {
        trim();
        for (auto& kv : sizes_) (void)hipFree(kv.first);
        sizes_.clear();
        in_use_ = 0;
    }
    size_t in_use() const { return in_use_; }
    size_t peak() const { return peak_; }

   private:
    static size_t size_class(size_t b) {
        if (b < 256) b = 256;
        // powers of two up to 1 MiB, then multiples of 1/8 of the next power of two
        size_t p = 256;
        while (p < b && p < (1u << 20)) p <<= 1;
        if (p >= b) return p;
        size_t hb = 1;
        while (hb < b) hb <<= 1;
        size_t step = hb / 8;
        return (b + step - 1) / step * step;
    }
    std::unordered_map<size_t, std::vector<void*>> free_;
    std::unordered_map<void*, size_t> sizes_;
    size_t in_use_ = 0, cached_ = 0, peak_ = 0;
};

class DBuf {
   public:
    DBuf() = default;
    DBuf(DevPool* pool, size_t bytes) : pool_(pool), bytes_(bytes) {
        p_ = bytes ? pool->get(bytes) : nullptr;
    }
    DBuf(DBuf&& o) noexcept : pool_(o.pool_), p_(o.p_), bytes_(o.bytes_) { o.p_ = nullptr; }
    DBuf& operator=(DBuf&& o) noexcept {
        if (this != &o) {
            reset();
            pool_ = o.pool_;
            p_ = o.p_;
            bytes_ = o.bytes_;
            o.p_ = nullptr;
        }
        return *this;
    }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { reset(); }
    void reset() {
        if (p_ && pool_) pool_->put(p_);
        p_ = nullptr;
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(p_);
    }
    size_t bytes() const { return bytes_; }

   private:
    DevPool* pool_ = nullptr;
    void* p_ = nullptr;
    size_t bytes_ = 0;
};

// ------------------------------------------------------------------ host patterns

struct HPat {
    int m = 0;
    int f[GS_MAXE], t[GS_MAXE], el[GS_MAXE], tl[GS_MAXE];  // ranks; tl = -1 for backward
    int fl0 = 0;                                           // from label of edge 0
    long long row_begin = 0, row_count = 0;
    int support = 0;
    int report_id = -1;      // index of this pattern's report, once reported
    int parent_report = -1;  // report of the parent pattern (-1 for single edges)

    int nv() const {
        int mx = 0;
        for (int j = 0; j < m; ++j) mx = std::max(mx, std::max(f[j], t[j]));
        return mx + 1;
    }
    // gspan.cu:53-65
    int rmpath(int* rm) const {
        int r = 0, last = -1;
        for (int i = m - 1; i >= 0; --i)
            if (f[i] < t[i] && (r == 0 || t[i] == last)) {
                rm[r++] = i;
                last = f[i];
            }
        return r;
    }
    void labels(int* lab) const {
        for (int v = 0; v < GS_MAXV; ++v) lab[v] = 0;
        lab[f[0]] = fl0;
        for (int j = 0; j < m; ++j)
            if (f[j] < t[j]) lab[t[j]] = tl[j];
    }
};

struct ChildInfo {
    HPat pat;
    long long count;
    int canonical;
    int support;
    long long row_off;  // in the child row array (-1 when not written)
};

// Position of a pattern in the reference's report order (DFS pre-order of mine()): four
// numbers per code edge, compared lexicographically, a prefix first.  The four numbers of
// an edge are packed into one 64-bit word, 16 bits each, by order-preserving maps (label
// ranks are below 2^16, vertex ids within +-8), so a comparison reads one word per edge.
struct OrderKey {
    int n = 0;  // edges
    uint64_t w[GS_MAXE];
    bool operator<(const OrderKey& o) const {
        const int k = std::min(n, o.n);
        for (int i = 0; i < k; ++i)
            if (w[i] != o.w[i]) return w[i] < o.w[i];
        return n < o.n;
    }
};

// One reported pattern: its last code edge and its parent's report (the parent is reported
// before it), so a report is 40 bytes however long the code; codes are rebuilt along a
// pre-order walk of the report tree (for_each_report).
struct Report {
    int m;
    int e5[5];          // last code edge, original labels, reference placeholder form
    uint64_t edge_key;  // order word of the last code edge (see order_key)
    int parent;         // report of the parent pattern, -1 for a single edge
    int support;
};

// Pinned host memory for the small transfers between host syncs.  A copy from or to
// pageable memory is staged by the runtime and costs tens of microseconds of latency
// each; through pinned memory the copies of one step are queued on the stream and
// complete together at the next sync.  Blocks are kept for reuse; everything taken is
// released by reset() once the stream has been synchronised.
class HostStage {
   public:
    static constexpr size_t kBlock = 1 << 20;
    static constexpr size_t kMaxStaged = 16 << 20;  // larger copies go direct

    HostStage() = default;
    HostStage(const HostStage&) = delete;
    HostStage& operator=(const HostStage&) = delete;
    ~HostStage() {
        for (auto& b : blocks_) (void)hipHostFree(b.p);
    }
    void* take(size_t bytes) {
        bytes = (bytes + 255) & ~size_t(255);
        while (cur_ < blocks_.size() && blocks_[cur_].used + bytes > blocks_[cur_].cap) ++cur_;
        if (cur_ == blocks_.size()) {
            Block b{nullptr, std::max(bytes, kBlock), 0};
            const hipError_t e = hipHostMalloc(&b.p, b.cap, hipHostMallocDefault);
            if (e != hipSuccess)
                throw GsError(GS_ERR_HIP, std::string("hipHostMalloc: ") + hipGetErrorString(e));
            blocks_.push_back(b);
        }
        Block& b = blocks_[cur_];
        void* p = static_cast<char*>(b.p) + b.used;
        b.used += bytes;
        return p;
    }
    void reset() {
        for (auto& b : blocks_) b.used = 0;
        cur_ = 0;
    }

   private:
    struct Block {
        void* p;
        size_t cap, used;
    };
    std::vector<Block> blocks_;
    size_t cur_ = 0;
};

// A device-to-host copy into pinned memory, copied to its destination at the next sync.
struct PendingD2H {
    void* dst;
    const void* staged;
    size_t bytes;
};

}  // namespace

struct gs_ctx {
    gs_opts opts{};
    int device = 0;
    hipStream_t stream = nullptr;
    int cus = 0;
    int n = 0, nnz = 0;
    int max_deg = 0;
    std::vector<int> vl_vals, el_vals;  // rank -> original label
    int VL = 0, EL = 0;
    DevPool pool;
    HostStage stage;
    std::vector<PendingD2H> pending_d2h;
    DBuf d_offsets, d_src, d_adj, d_vlab, d_adjl, d_eidl, d_lstart;
    // label classes: rank of each vertex within its label, the same per adjl entry, and the
    // vertices of each label rank (lcnt on the host too: set planning)
    DBuf d_lidx, d_adjlx, d_lcnt;
    std::vector<int> lcnt;
    // runs of equal (neighbour label, edge label) in adjl
    DBuf d_rstart, d_rkey, d_runid;
    // bytemap pool of the single-pass last level, kept between steps: tagged_words words,
    // bytes of the current step equal `tag`
    DBuf d_tagged;
    long long tagged_words = 0;
    int tag = 0;
    // two-edge last level: run masks M1 / M2 (te_W words per vertex) and the (vertex, run)
    // pairs ordered by class, built on first use (they depend on the graph only)
    // and the first pair of every class (vertex label * VL * EL + run bit; VL * VL * EL + 1
    // entries); built by gs_create
    DBuf d_te_m1, d_te_m2, d_te_pairs, d_te_pcl, d_te_cstart, d_te_rbl;
    bool te_ready = false;
    DBuf d_vl_vals, d_el_vals;  // original label of each rank, on the device
    // A run whose two-edge last level is reached from every frequent single-edge code
    // (flat_ok; flat_seeds codes) builds its reports on the device, in the reference's
    // order (flat): codes, edge counts and supports land in pinned host buffers.
    bool flat = false, flat_ok = false;
    int flat_seeds = 0;
    // one pinned buffer: codes (5 per edge), then edge counts, then supports
    int* flat_buf = nullptr;
    long long flat_cap = 0;
    int *flat_codes = nullptr, *flat_lens = nullptr, *flat_sups = nullptr;  // into flat_buf
    long long flat_reports = 0, flat_edges = 0;
    // The two-edge pass over every frequent single-edge code, launched from the class sums
    // before the host has listed the codes (te_prelaunch): its tables, counters, the codes'
    // count on the device and the slot of its timing event (-1: untimed).  The step over
    // those codes launches the key pass on the tables (on: not taken yet).
    struct TePre {
        bool on = false;
        int ev = -1;
        DBuf cnt, cen, end, q, work, flags, scan, tmp, pof, pmask, ncodes;
    } te_pre;
    int* h_te_np = nullptr;                 // pinned copies of the class sums
    unsigned long long* h_te_nl = nullptr;
    long long h_te_cap = 0;
    hipEvent_t ev_seeds = nullptr;  // the class sums have reached the host
    // neighbours per run as a row scan sees them: a vertex is scanned about once per
    // incident row, so items ~ sum deg^2 and run items ~ sum deg * runs; their ratio.  Long
    // runs make the run modes and per-group new-vertex inserts pay at the last level
    // (measured: 2.15 on the ER workload, where they cost 13%; 2.7 and 5.3 on the R-MAT
    // scale-14 and scale-18 workloads with 100 labels, where they saved 24% and 38%)
    double run_gain = 1.0;
    DevGraph dg{};
    // [0] guard word, [1] n_active, [2] err, [3] overflow, [4..5] int64 total,
    // [6..7] items without the label filter, [8..11] two 64-bit step totals
    DBuf d_misc;
    DBuf d_tiles; // per-XCD tile queues of the item kernels
    LaunchTuner tuner;  // launch shape of the item kernels
    std::string err;
    gs_stats st{};
    size_t budget = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // event pairs recorded around count/fill launches (timing=1, or a launch the tuner
    // measures), resolved at the next host synchronisation so that timing adds no
    // synchronisation of its own
    struct PendingEvent {
        int ev;                 // index in ev_pool
        int kind;               // 0 = count, 1 = fill
        int bucket, cfg;        // tuner bucket / shape, or -1
        const long long* items; // work items of the launch, valid at resolution
    };
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pool;
    std::vector<PendingEvent> ev_pending;
    size_t ev_used = 0;
    // mining state
    int k = 0, min_sup = 0, mni_mode = 0;
    std::vector<Report> reports;
    std::vector<int> report_order;  // reports in the reference's order (order_reports)
    // child-count predictor: label statistics, parameters, Adam state, logged samples
    DBuf d_H, d_vfreq, d_efreq, d_dmean;
    DBuf d_w32, d_w16, d_adam_m, d_adam_v, d_grads;
    MlpDev mlp{};
    int mlp_step = 0;
    bool mlp_trained = false;
    bool logging = false;
    long long log_cap = 0, n_logged = 0;
    DBuf d_logX, d_logT;

    gsk::TileQueues tq{};  // over d_tiles
    // child arrays of finished steps, reused by later ones: a level's array is ~170 bytes
    // per candidate, and fresh memory for it every step costs page faults
    std::vector<std::vector<ChildInfo>> child_bufs;
    // zeroes both tile-queue sets (at the start of every entry point that extends)
    void reset_tile_queues() {
        tq = gsk::TileQueues{d_tiles.as<int>(), 0};
        HIPCHECK(gsk::fill_i32(tq.base, 0, 2 * GS_TILE_QUEUE, stream));
    }
    // index guards of the kernels (GS_GUARD_* bits); read back at every sync
    int* guard_word() const { return d_misc.as<int>(); }
    int guard_host = 0;
    int* n_active() const { return d_misc.as<int>() + 1; }
    int* err_flag() const { return d_misc.as<int>() + 2; }
    int* overflow_flag() const { return d_misc.as<int>() + 3; }
    long long* total64() const { return reinterpret_cast<long long*>(d_misc.as<int>() + 4); }
    // child rows and canonical keys of a step whose tail runs on the device
    unsigned long long* step_totals() const {
        return reinterpret_cast<unsigned long long*>(d_misc.as<int>() + 8);
    }
    // items of a step without the label filter (the algorithmic-bytes figure counts them)
    unsigned long long* work_full64() const {
        return reinterpret_cast<unsigned long long*>(d_misc.as<int>() + 6);
    }
};

namespace {

// Names of the guard bits a kernel raised.
std::string guard_names(int bits) {
    static const char* const names[] = {"child key", "live slot", "row reservation",
                                        "set pool", "graph index", "layer row"};
    std::string s;
    for (int b = 0; b < 6; ++b)
        if (bits & (1 << b)) s += std::string(s.empty() ? "" : ", ") + names[b];
    return s;
}

// Waits for the stream; device-to-host copies issued since the last sync are then in
// their destinations.  The kernels' guard word is read in the same wait: a kernel that
// met an index outside its buffers (and skipped the access) fails the call here.
void sync(gs_ctx* c) {
    const bool guarded = c->d_misc.as<int>() != nullptr;
    if (guarded) {
        void* p = c->stage.take(sizeof(int));
        HIPCHECK(hipMemcpyAsync(p, c->guard_word(), sizeof(int), hipMemcpyDeviceToHost,
                                c->stream));
        c->pending_d2h.push_back(PendingD2H{&c->guard_host, p, sizeof(int)});
    }
    const auto t0 = std::chrono::steady_clock::now();
    HIPCHECK(hipStreamSynchronize(c->stream));
    c->st.t_sync_ms +=
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    for (const PendingD2H& p : c->pending_d2h) std::memcpy(p.dst, p.staged, p.bytes);
    c->pending_d2h.clear();
    c->stage.reset();
    if (guarded && c->guard_host != 0) {
        const int bits = c->guard_host;
        c->guard_host = 0;
        HIPCHECK(gsk::fill_i32(c->guard_word(), 0, 1, c->stream));  // for later calls
        throw GsError(GS_ERR_INTERNAL,
                      "kernel index guard: " + guard_names(bits) + " outside its buffer");
    }
}

// Clears the guard word (at the start of every call that launches kernels).
void clear_guard(gs_ctx* c) {
    HIPCHECK(gsk::fill_i32(c->guard_word(), 0, 1, c->stream));
    c->guard_host = 0;
}

// Host-to-device copy: src may be released as soon as this returns.
template <class T>
void h2d(gs_ctx* c, T* dst, const T* src, size_t count) {
    const size_t bytes = count * sizeof(T);
    if (!bytes) return;
    if (bytes <= HostStage::kMaxStaged) {
        void* p = c->stage.take(bytes);
        std::memcpy(p, src, bytes);
        HIPCHECK(hipMemcpyAsync(dst, p, bytes, hipMemcpyHostToDevice, c->stream));
    } else {
        HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
        HIPCHECK(hipStreamSynchronize(c->stream));  // src belongs to the caller
    }
}
// Device-to-host copy: dst holds the data after the next sync(c).
template <class T>
void d2h(gs_ctx* c, T* dst, const T* src, size_t count) {
    const size_t bytes = count * sizeof(T);
    if (!bytes) return;
    if (bytes <= HostStage::kMaxStaged) {
        void* p = c->stage.take(bytes);
        HIPCHECK(hipMemcpyAsync(p, src, bytes, hipMemcpyDeviceToHost, c->stream));
        c->pending_d2h.push_back(PendingD2H{dst, p, bytes});
    } else {
        HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    }
}

// Distinct labels of count values p[0], p[stride], ... in ascending order (rank -> label),
// and each label's rank: a direct table over the labels' range when it spans at most 2^24
// values (two passes over the values, no sort), else a sort of the values.
struct LabelRanks {
    std::vector<int> vals;
    long long lo = 0;
    std::vector<int> table;  // label - lo -> rank (direct table), else empty
    void build(const int* p, long long count, long long stride) {
        vals.clear();
        table.clear();
        if (count == 0) return;
        int mn = p[0], mx = p[0];
        for (long long i = 1; i < count; ++i) {
            const int v = p[i * stride];
            mn = std::min(mn, v);
            mx = std::max(mx, v);
        }
        lo = mn;
        const long long span = (long long)mx - mn + 1;
        if (span <= (1LL << 24)) {
            table.assign((size_t)span, -1);
            for (long long i = 0; i < count; ++i) table[(size_t)((long long)p[i * stride] - lo)] = 0;
            for (long long k = 0; k < span; ++k)
                if (table[(size_t)k] == 0) {
                    table[(size_t)k] = (int)vals.size();
                    vals.push_back((int)(lo + k));
                }
        } else {
            vals.resize((size_t)count);
            for (long long i = 0; i < count; ++i) vals[(size_t)i] = p[i * stride];
            std::sort(vals.begin(), vals.end());
            vals.erase(std::unique(vals.begin(), vals.end()), vals.end());
        }
    }
    int rank(int label) const {
        if (!table.empty()) return table[(size_t)((long long)label - lo)];
        return int(std::lower_bound(vals.begin(), vals.end(), label) - vals.begin());
    }
};

// fn(begin, end) over [0, n) split among the host's hardware threads (at most 16)
void parallel_ranges(long long n, const std::function<void(long long, long long)>& fn) {
    const int hw = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const int nt = (int)std::min<long long>(hw, std::max(1LL, n / (1 << 16)));
    if (nt <= 1) {
        fn(0, n);
        return;
    }
    std::vector<std::thread> ts;
    for (int t = 0; t < nt; ++t)
        ts.emplace_back(fn, n * t / nt, n * (t + 1) / nt);
    for (auto& th : ts) th.join();
}

int vl_rank(const gs_ctx* c, int label) {
    auto it = std::lower_bound(c->vl_vals.begin(), c->vl_vals.end(), label);
    return int(it - c->vl_vals.begin());
}

long long key_space(int r, int EL, int VL) {
    return (long long)(r - 1) * EL + (long long)(r + 1) * EL * VL;
}

// Pattern of a batch owning child key `key` (patterns ordered by key_base).
int owner_of(const std::vector<PatDev>& pd, int key) {
    const auto it = std::upper_bound(pd.begin(), pd.end(), key,
                                     [](int k, const PatDev& d) { return k < d.key_base; });
    return int(it - pd.begin()) - 1;
}

PatDev to_dev(const HPat& h) {
    PatDev d{};
    d.m = h.m;
    d.nv = h.nv();
    int rm[GS_MAXE];
    d.r = h.rmpath(rm);
    for (int i = 0; i < GS_MAXE; ++i) d.rm[i] = i < d.r ? rm[i] : 0;
    d.v_n = h.t[rm[0]];
    d.min_label = h.fl0;
    for (int j = 0; j < GS_MAXE; ++j) {
        d.cfrom[j] = j < h.m ? h.f[j] : 0;
        d.cto[j] = j < h.m ? h.t[j] : 0;
        d.cel[j] = j < h.m ? h.el[j] : 0;
    }
    h.labels(d.vlab);
    for (int j = 0; j < GS_MAXE; ++j) {
        d.seg_v[j] = 0;
        d.seg_el[j] = -1;
        d.seg_tl[j] = 0;
        d.bw_v[j] = 0;
        d.bw_el[j] = 0;
        d.bw_le[j] = 0;
    }
    d.seg_v[0] = d.v_n;
    for (int j = 1; j <= d.r && j < GS_MAXE; ++j) {
        const int k = rm[j - 1];
        d.seg_v[j] = d.cfrom[k];
        d.seg_el[j] = d.cel[k];
        d.seg_tl[j] = d.vlab[d.cto[k]];
    }
    for (int i = 1; i < d.r; ++i) {
        const int k = rm[i];
        d.bw_v[i] = d.cfrom[k];
        d.bw_el[i] = d.cel[k];
        d.bw_le[i] = d.vlab[d.cto[k]] <= d.vlab[d.v_n] ? 1 : 0;
    }
    d.support = h.support;
    d.row_begin = h.row_begin;
    d.row_count = h.row_count;
    return d;
}

// Child pattern of parent p for a slot of its key space (see KeySpace in gs_internal.h).
void decode_child_into(const HPat& p, const PatDev& d, int slot, int EL, int VL, HPat& c) {
    c = p;
    c.m = p.m + 1;
    c.report_id = -1;
    c.parent_report = p.report_id;
    const int j = p.m;
    const int nb = (d.r - 1) * EL;
    if (slot < nb) {
        const int i = slot / EL + 1;
        c.f[j] = d.v_n;
        c.t[j] = p.f[d.rm[i]];
        c.el[j] = slot % EL;
        c.tl[j] = -1;
    } else {
        const int s = slot - nb;
        const int seg = s / (EL * VL);
        c.f[j] = seg == 0 ? d.v_n : p.f[d.rm[seg - 1]];
        c.t[j] = d.v_n + 1;
        c.el[j] = (s / VL) % EL;
        c.tl[j] = s % VL;
    }
}

HPat decode_child(const HPat& p, const PatDev& d, int slot, int EL, int VL) {
    HPat c;
    decode_child_into(p, d, slot, EL, VL, c);
    return c;
}

// Order of the reference's report lines: seeds by (fl, el, tl); within a node, backward
// children (ascending v_i, el) before forward children (descending v_i, ascending el, tl);
// a pattern precedes its descendants (gspan.cu:1142-1168, 1185).  Each code edge maps to
// one 64-bit word, 16 bits per field by order-preserving maps (label ranks are below
// 2^16, vertex ids within +-8), and a code's position is the sequence of its words.
uint64_t edge_word(const HPat& h, int j) {
    auto put = [](int a, int b, int c, int d) {
        return (uint64_t)a << 48 | (uint64_t)b << 32 | (uint64_t)c << 16 | (uint64_t)d;
    };
    if (j == 0) return put(h.fl0, h.el[0], h.tl[0], 0);  // label ranks
    // backward: (0, to, el, 0); forward: (1, -from, el, tl); vertex ids shifted by 8
    if (h.f[j] > h.t[j]) return put(0, h.t[j] + 8, h.el[j], 0);
    return put(1, 8 - h.f[j], h.el[j], h.tl[j]);
}

OrderKey order_key(const HPat& h) {
    OrderKey k;
    for (int j = 0; j < h.m; ++j) k.w[k.n++] = edge_word(h, j);
    return k;
}

// Reports into the reference's order: the reported patterns form a tree (a pattern's
// parent is reported before it is extended), and the order is the tree's pre-order with
// each node's children -- and the single edges -- sorted by their last edge's word.
// Siblings have distinct last edges, so the order is total.  Linear apart from the sorts
// of sibling groups.
void order_reports(gs_ctx* c) {
    const std::vector<Report>& r = c->reports;
    const int n = (int)r.size();
    // the two fields the order needs, gathered once by all host threads: the passes below
    // then stream 12 bytes per report instead of the whole report
    std::vector<int> par(n);
    std::vector<uint64_t> key(n);
    parallel_ranges(n, [&](long long a, long long b) {
        for (long long i = a; i < b; ++i) {
            par[i] = r[i].parent;
            key[i] = r[i].edge_key;
        }
    });
    std::vector<int> start(n + 2, 0), kids(n);
    for (int i = 0; i < n; ++i) ++start[par[i] + 2];  // parent -1 counts at index 1
    for (int i = 0; i <= n; ++i) start[i + 1] += start[i];
    {
        std::vector<int> at(start.begin(), start.end() - 1);
        for (int i = 0; i < n; ++i) kids[at[par[i] + 1]++] = i;
    }
    // group g (parent g - 1) occupies kids[start[g] .. start[g+1]); children reported in
    // key order mostly arrive in edge order already
    auto by_edge = [&](int a, int b) { return key[a] < key[b]; };
    parallel_ranges((long long)n + 1, [&](long long g0, long long g1) {
        for (long long g = g0; g < g1; ++g) {
            const auto b0 = kids.begin() + start[g], b1 = kids.begin() + start[g + 1];
            if (!std::is_sorted(b0, b1, by_edge)) std::sort(b0, b1, by_edge);
        }
    });
    // Pre-order positions without a walk: a report's parent is reported before it, so
    // subtree sizes add up in one backward pass; then each parent (ascending index, so its
    // own position is set) places its children's subtrees one after another in edge order.
    std::vector<int> size(n, 1), pos(n);
    for (int i = n - 1; i >= 0; --i) {
        const int q = par[i];
        if (q >= i) throw GsError(GS_ERR_INTERNAL, "report before its parent");
        if (q >= 0) size[q] += size[i];
    }
    int at = 0;
    for (int i = start[0]; i < start[1]; ++i) {
        pos[kids[i]] = at;
        at += size[kids[i]];
    }
    if (at != n) throw GsError(GS_ERR_INTERNAL, "report tree is not connected");
    for (int g = 1; g <= n; ++g) {
        int q = pos[g - 1] + 1;
        for (int i = start[g]; i < start[g + 1]; ++i) {
            pos[kids[i]] = q;
            q += size[kids[i]];
        }
    }
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[pos[i]] = i;
    c->report_order.swap(order);
}

// fn(code5, m, support) for every report in the reference's order: the order is a pre-order
// of the report tree, so the codes of a report's ancestors are the prefix of the path built
// so far.
template <class Fn>
void for_each_report(const gs_ctx* c, Fn&& fn) {
    if (c->flat) {
        long long e = 0;
        for (long long r = 0; r < c->flat_reports; ++r) {
            fn(c->flat_codes + 5 * e, c->flat_lens[r], c->flat_sups[r]);
            e += c->flat_lens[r];
        }
        return;
    }
    int path[5 * GS_MAXE];
    for (int i : c->report_order) {
        const Report& r = c->reports[i];
        std::memcpy(path + 5 * (r.m - 1), r.e5, sizeof(r.e5));
        fn(path, r.m, r.support);
    }
}

// Appends the report of h; returns its index.
int add_report(gs_ctx* c, const HPat& h) {
    Report r;
    r.m = h.m;
    const int j = h.m - 1;
    r.e5[0] = h.f[j];
    r.e5[1] = h.t[j];
    r.e5[2] = j == 0 ? c->vl_vals[h.fl0] : -1;
    r.e5[3] = c->el_vals[h.el[j]];
    r.e5[4] = h.tl[j] < 0 ? -1 : c->vl_vals[h.tl[j]];
    r.edge_key = edge_word(h, j);
    r.parent = h.parent_report;
    r.support = h.support;
    c->reports.push_back(r);
    c->st.reported++;
    return (int)c->reports.size() - 1;
}

// Appends the report of the child of parent pattern p (device pattern d) at key slot `slot`
// with its support, straight from the key: the last edge alone is decoded (as in
// decode_child_into), for children that are reported and not extended.
Report child_report(const gs_ctx* c, const HPat& p, const PatDev& d, int slot, int support) {
    const int EL = c->EL, VL = c->VL;
    const int nb = (d.r - 1) * EL;
    int f, t, el, tl;
    if (slot < nb) {
        f = d.v_n;
        t = p.f[d.rm[slot / EL + 1]];
        el = slot % EL;
        tl = -1;
    } else {
        const int s2 = slot - nb, seg = s2 / (EL * VL);
        f = seg == 0 ? d.v_n : p.f[d.rm[seg - 1]];
        t = d.v_n + 1;
        el = (s2 / VL) % EL;
        tl = s2 % VL;
    }
    Report r;
    r.m = p.m + 1;
    r.e5[0] = f;
    r.e5[1] = t;
    r.e5[2] = -1;
    r.e5[3] = c->el_vals[el];
    r.e5[4] = tl < 0 ? -1 : c->vl_vals[tl];
    auto put = [](int a, int b, int cc, int dd) {
        return (uint64_t)a << 48 | (uint64_t)b << 32 | (uint64_t)cc << 16 | (uint64_t)dd;
    };
    r.edge_key = f > t ? put(0, t + 8, el, 0) : put(1, 8 - f, el, tl);  // as edge_word
    r.parent = p.report_id;
    r.support = support;
    return r;
}

void report_child(gs_ctx* c, const HPat& p, const PatDev& d, int slot, int support) {
    c->reports.push_back(child_report(c, p, d, slot, support));
    c->st.reported++;
}

// Support from the per-vertex distinct-image counts (gspan.cu:282-288 / 367-373).
int support_of(const HPat& h, const int* counts, int mni_mode) {
    const int nv = h.nv();
    bool v0 = true;
    if (mni_mode == GS_MNI_OMP_LITERAL) {
        v0 = false;
        for (int j = 1; j < h.m; ++j)
            if (h.f[j] == 0) v0 = true;
    }
    long long sup = INT_MAX;
    for (int v = 0; v < nv; ++v) {
        if (v == 0 && !v0) continue;
        if (counts[v] > 0 && counts[v] < sup) sup = counts[v];
    }
    return (int)sup;
}

// Byte-per-vertex sets unless the caller asked for bits.
bool want_bytes(const gs_ctx* c) { return c->opts.set_kind != GS_SETS_BITS; }

// Sets of a child with `rows` rows and nv vertices of label ranks lab[]: one set per
// vertex over the graph vertices of its label (plan_key_sets), placed at pool_words.
void plan_set(const gs_ctx* c, long long rows, int nv, const int* lab, bool bytes, SetDesc& d,
              long long& pool_words) {
    const long long w = plan_key_sets(rows, nv, lab, c->lcnt.data(), bytes, &d);
    d.off = pool_words;
    pool_words += w;
}

// Rows large enough that every vertex set is a bitmap or bytemap over its label class (sets
// planned before the counts are known).
constexpr long long kAllRows = 1LL << 40;

// Words of a bitmap (bytes = false) or bytemap vertex set over label rank l.
long long class_words(const gs_ctx* c, int l, bool bytes) {
    int kind;
    long long w;
    plan_vertex_set(kAllRows, c->lcnt[l], bytes, &kind, &w);
    return w;
}

// Label statistics for the predictor's features: vertex-label and edge-label frequencies,
// mean degree per vertex label, and H[a][e][b] = mean number of neighbours with edge label e
// and vertex label b of a vertex labelled a.
void init_label_stats(gs_ctx* c, const std::vector<int>& vr, const std::vector<int>& src,
                      const std::vector<uint2>& adj) {
    const int VL = c->VL, EL = c->EL;
    std::vector<double> vcount(VL, 0), ecount(EL, 0), deg(VL, 0);
    for (int v = 0; v < c->n; ++v) vcount[vr[v]] += 1;
    const long long HS = (long long)VL * EL * VL;
    const bool useH = HS <= (1LL << 24);
    std::vector<float> H(useH ? HS : 1, 0.f);
    for (int i = 0; i < c->nnz; ++i) {
        const int a = vr[src[i]], e = (int)(adj[i].y >> 16), b = (int)(adj[i].y & 0xffffu);
        ecount[e] += 1;
        deg[a] += 1;
        if (useH) H[((long long)a * EL + e) * VL + b] += 1.f;
    }
    std::vector<float> vf(VL), ef(EL), dm(VL);
    for (int a = 0; a < VL; ++a) {
        vf[a] = c->n ? (float)(vcount[a] / c->n) : 0.f;
        dm[a] = vcount[a] > 0 ? (float)(deg[a] / vcount[a]) : 0.f;
        if (useH && vcount[a] > 0)
            for (long long q = 0; q < (long long)EL * VL; ++q)
                H[(long long)a * EL * VL + q] /= (float)vcount[a];
    }
    for (int e = 0; e < EL; ++e) ef[e] = c->nnz ? (float)(ecount[e] / c->nnz) : 0.f;
    DevPool* P = &c->pool;
    c->d_H = useH ? DBuf(P, sizeof(float) * HS) : DBuf();
    c->d_vfreq = DBuf(P, sizeof(float) * VL);
    c->d_efreq = DBuf(P, sizeof(float) * EL);
    c->d_dmean = DBuf(P, sizeof(float) * VL);
    if (useH) h2d(c, c->d_H.as<float>(), H.data(), (size_t)HS);
    h2d(c, c->d_vfreq.as<float>(), vf.data(), VL);
    h2d(c, c->d_efreq.as<float>(), ef.data(), EL);
    h2d(c, c->d_dmean.as<float>(), dm.data(), VL);
    sync(c);  // the copies read H, vf, ef and dm, which end with this function
}

void mlp_bind(gs_ctx* c) {
    float* w = c->d_w32.as<float>();
    uint16_t* h = c->d_w16.as<uint16_t>();
    const long long oW1 = 0, ob1 = oW1 + GS_FEATURES * GS_HIDDEN, oW2 = ob1 + GS_HIDDEN,
                    ob2 = oW2 + GS_HIDDEN * GS_HIDDEN, oW3 = ob2 + GS_HIDDEN, ob3 = oW3 + GS_HIDDEN;
    c->mlp.w32 = w;
    c->mlp.w16_all = h;
    c->mlp.w16[0] = h + oW1;
    c->mlp.w16[1] = h + oW2;
    c->mlp.w16[2] = h + oW3;
    c->mlp.b[0] = w + ob1;
    c->mlp.b[1] = w + ob2;
    c->mlp.b[2] = w + ob3;
}

// He-uniform initial weights from a fixed-seed generator (splitmix64), zero biases.
void mlp_init(gs_ctx* c) {
    DevPool* P = &c->pool;
    c->d_w32 = DBuf(P, sizeof(float) * GS_MLP_PARAMS);
    c->d_w16 = DBuf(P, sizeof(uint16_t) * GS_MLP_PARAMS);
    c->d_adam_m = DBuf(P, sizeof(float) * GS_MLP_PARAMS);
    c->d_adam_v = DBuf(P, sizeof(float) * GS_MLP_PARAMS);
    c->d_grads = DBuf(P, sizeof(float) * GS_MLP_PARAMS);
    mlp_bind(c);
    std::vector<float> w(GS_MLP_PARAMS, 0.f);
    unsigned long long st = 0x9E3779B97F4A7C15ull;
    auto rnd = [&]() {
        unsigned long long z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) / 9007199254740992.0;  // [0, 1)
    };
    auto fill = [&](long long off, int fan_in, long long count) {
        const double lim = std::sqrt(6.0 / fan_in);
        for (long long i = 0; i < count; ++i) w[off + i] = (float)((2.0 * rnd() - 1.0) * lim);
    };
    fill(0, GS_FEATURES, GS_FEATURES * GS_HIDDEN);
    fill(GS_FEATURES * GS_HIDDEN + GS_HIDDEN, GS_HIDDEN, GS_HIDDEN * GS_HIDDEN);
    fill(GS_FEATURES * GS_HIDDEN + 2 * GS_HIDDEN + GS_HIDDEN * GS_HIDDEN, GS_HIDDEN, GS_HIDDEN);
    h2d(c, c->d_w32.as<float>(), w.data(), GS_MLP_PARAMS);
    HIPCHECK(gsk::fill_i32(c->d_adam_m.as<int>(), 0, GS_MLP_PARAMS, c->stream));
    HIPCHECK(gsk::fill_i32(c->d_adam_v.as<int>(), 0, GS_MLP_PARAMS, c->stream));
    HIPCHECK(gsk::mlp_sync_bf16(c->mlp, c->stream));
    sync(c);  // the upload reads w, which ends with this function
    c->mlp_step = 0;
    c->mlp_trained = false;
}

// Records the first event of the next timing pair (ev_pool[ev_used]) on the stream.
void ev_begin(gs_ctx* c) {
    if (c->ev_used == c->ev_pool.size()) {
        hipEvent_t a, b;
        HIPCHECK(hipEventCreate(&a));
        HIPCHECK(hipEventCreate(&b));
        c->ev_pool.emplace_back(a, b);
    }
    HIPCHECK(hipEventRecord(c->ev_pool[c->ev_used].first, c->stream));
}

// Folds the event pairs recorded since the last host synchronisation into the stats.
void resolve_timing(gs_ctx* c) {
    for (auto& pk : c->ev_pending) {
        float ms = 0;
        HIPCHECK(hipEventElapsedTime(&ms, c->ev_pool[pk.ev].first, c->ev_pool[pk.ev].second));
        if (c->opts.timing) (pk.kind == 0 ? c->st.t_count_ms : c->st.t_fill_ms) += ms;
        c->tuner.record(pk.bucket, pk.cfg, ms, pk.items ? *pk.items : 0);
    }
    c->ev_pending.clear();
    c->ev_used = 0;
}

// The active keys' arrays (keys, counts and optionally flags, n of them) to the host.
// When the key space is small they are copied whole in the same synchronisation as the
// active count (one host round trip instead of two); otherwise after it.
constexpr long long kSpeculativeKeys = 1 << 16;
// steps with more active keys plan their children and compute supports on the device
constexpr long long kDeviceTailKeys = 1 << 16;
struct ActiveHost {
    std::vector<int> keys, cnt, flag;
};
// limit: when more than `limit` keys are active only the head is read (the step then
// continues on the device).
void fetch_active(gs_ctx* c, long long K, const int* d_active, const int* d_acnt,
                  const int* d_flags, int* head, int nhead, ActiveHost& h,
                  long long limit = LLONG_MAX) {
    const bool spec = K <= kSpeculativeKeys;
    d2h(c, head, c->n_active(), (size_t)nhead);
    if (spec) {
        h.keys.resize(K);
        h.cnt.resize(K);
        d2h(c, h.keys.data(), d_active, (size_t)K);
        d2h(c, h.cnt.data(), d_acnt, (size_t)K);
        if (d_flags) {
            h.flag.resize(K);
            d2h(c, h.flag.data(), d_flags, (size_t)K);
        }
    }
    sync(c);
    const int A = head[0];
    if (spec) {
        h.keys.resize(A);
        h.cnt.resize(A);
        if (d_flags) h.flag.resize(A);
        return;
    }
    if (A > limit) return;
    h.keys.resize(A);
    h.cnt.resize(A);
    d2h(c, h.keys.data(), d_active, (size_t)A);
    d2h(c, h.cnt.data(), d_acnt, (size_t)A);
    if (d_flags) {
        h.flag.resize(A);
        d2h(c, h.flag.data(), d_flags, (size_t)A);
    }
    sync(c);
}

// Keys with a non-zero count, in ascending key order, into active / acnt; *n_active.
// The error word of the canonicality check that follows is zeroed on the way.
void compact_ordered(gs_ctx* c, const int* key_count, long long K, int* active, int* acnt) {
    DevPool* P = &c->pool;
    const long long te = gsk::scan_tmp_elems(K);
    DBuf flags(P, sizeof(int) * std::max(K, 1LL)), pos(P, sizeof(long long) * (K + 1)),
        tmp(P, sizeof(long long) * te);
    HIPCHECK(gsk::compact_active_ordered(key_count, K, active, acnt, c->n_active(),
                                         c->err_flag(), flags.as<int>(), pos.as<long long>(),
                                         tmp.as<long long>(), te, c->stream));
}

// Child count as read back from the device: a negative count means more than INT_MAX
// embeddings (the kernels saturate it); 2^31 stands for it, so a level that would have to
// store those rows fails its 2^31-row check, while a level that stores no rows is exact
// (supports come from the image sets, not the counts).
inline long long child_count(int c) { return c < 0 ? (long long)INT_MAX + 1 : c; }

// Rows of one DFS level: in device memory, or -- when they do not fit the device
// budget -- spilled to pinned host memory that the kernels reach through its device
// mapping (slower, over the host link, but the level can be mined).
class LevelRows {
   public:
    LevelRows() = default;
    LevelRows(const LevelRows&) = delete;
    LevelRows& operator=(const LevelRows&) = delete;
    LevelRows(LevelRows&& o) noexcept { *this = std::move(o); }
    LevelRows& operator=(LevelRows&& o) noexcept {
        if (this != &o) {
            release();
            dev_ = std::move(o.dev_);
            host_ = o.host_;
            ptr_ = o.ptr_;
            o.host_ = nullptr;
            o.ptr_ = nullptr;
        }
        return *this;
    }
    ~LevelRows() { release(); }
    void on_device(DevPool* P, long long rows) {
        release();
        dev_ = DBuf(P, sizeof(int2) * std::max(rows, 1LL));
        ptr_ = dev_.as<int2>();
    }
    void on_host(long long rows) {
        release();
        hipError_t e = hipHostMalloc(&host_, sizeof(int2) * std::max(rows, 1LL),
                                     hipHostMallocDefault);
        if (e != hipSuccess)
            throw GsError(GS_ERR_BUDGET, std::string("spilling level rows to host memory: ") +
                                             hipGetErrorString(e));
        void* d = nullptr;
        e = hipHostGetDevicePointer(&d, host_, 0);
        if (e != hipSuccess) {
            (void)hipHostFree(host_);
            host_ = nullptr;
            throw GsError(GS_ERR_HIP, std::string("hipHostGetDevicePointer: ") +
                                          hipGetErrorString(e));
        }
        ptr_ = static_cast<int2*>(d);
    }
    void adopt(DBuf&& d) {
        release();
        dev_ = std::move(d);
        ptr_ = dev_.as<int2>();
    }
    int2* ptr() const { return ptr_; }
    bool spilled() const { return host_ != nullptr; }

   private:
    // every kernel that read or wrote the rows finished before the step's last sync
    void release() {
        if (host_) (void)hipHostFree(host_);
        host_ = nullptr;
        dev_ = DBuf();
        ptr_ = nullptr;
    }
    DBuf dev_;
    void* host_ = nullptr;
    int2* ptr_ = nullptr;
};

// One DFS level's rows as the kernels read them: {prev, eid} per row.
struct Layer {
    const int2* p = nullptr;
    long long rows = 0;
};

struct StepOut {
    // children with a canonical code (every child when all_live), in key order
    std::vector<ChildInfo> children;
    long long candidates = 0;         // child codes that occurred (active keys)
    long long canonical = 0;          // of them with a canonical code
    LevelRows rows;                   // child rows of the written keys
    long long rows_total = 0;         // rows in `rows`
    long long emitted = 0;
    bool reported_inline = false;     // the frequent children were reported by run_step
};

// error of a canonicality walk that outgrew its projection buffers (never expected: the
// buffers hold the most projections any code of up to 7 edges can have)
const char* const kCanonCapMsg = "canonicality check: more projections than the walk buffers hold";

constexpr int STEP_OK = 0;
constexpr int STEP_SPLIT = 1;

// Row capacity of a child key from the predictor's output y = log1p(predicted count):
// 1.25x the prediction plus 16, at most 2^30.  The clamp is applied in floating point
// before the conversion, so every output -- negative, huge, infinite or NaN -- maps to a
// defined capacity in [16, 2^30].
long long predicted_capacity(float y) {
    constexpr double kMax = (double)(1 << 30);
    const double yd = (double)y;
    if (!(yd == yd)) return (long long)kMax;  // NaN
    const double pred = yd > 40.0 ? kMax : std::max(0.0, std::expm1(yd));
    return (long long)std::min(std::ceil(1.25 * pred) + 16.0, kMax);
}

// The graph and options allow the two-edge last level (gs_twoedge.hip): a simple graph,
// the batched or predicted method, no forced count-then-fill, and count tables of
// VL * (VL * EL)^2 entries (16 bytes each) of at most 1 GiB.
bool two_edge_graph(const gs_ctx* c) {
    const long long nb = (long long)c->VL * c->EL;
    return c->opts.mode != GS_MODE_PLAIN && !c->opts.two_pass &&
           c->opts.two_edge != GS_TWO_EDGE_OFF && c->dg.simple && nb <= 1024 &&
           (long long)c->VL * nb * nb <= (1LL << 26);
}

// The graph's index for the two-edge last level -- run masks, class-ordered (vertex, run)
// pairs and class starts -- built by gs_create, beside the label-ordered adjacency and
// runs it derives from.
void te_build(gs_ctx* c) {
    if (c->te_ready) return;
    DevPool* P = &c->pool;
    const int NB = c->VL * c->EL, W = (NB + 31) / 32;
    const long long NC = (long long)c->VL * NB;
    c->d_te_m1 = DBuf(P, 4 * std::max<long long>((long long)c->n * W, 1));
    c->d_te_m2 = DBuf(P, 4 * std::max<long long>((long long)c->n * W, 1));
    c->d_te_pairs = DBuf(P, 2 * sizeof(int4) * std::max(c->dg.nruns, 1));
    c->d_te_pcl = DBuf(P, sizeof(int2) * std::max(c->dg.nruns, 1));
    c->d_te_cstart = DBuf(P, 8 * (NC + 1));
    c->d_te_rbl = DBuf(P, sizeof(int2) * std::max(c->dg.nruns, 1));
    const long long te = gsk::scan_tmp_elems(NC);
    DBuf hist(P, 4 * NC), tmp(P, 8 * te);
    HIPCHECK(gsk::te_prepare(c->dg, c->EL, NB, W, c->d_te_m1.as<uint32_t>(),
                             c->d_te_m2.as<uint32_t>(), hist.as<int>(),
                             c->d_te_cstart.as<long long>(), tmp.as<long long>(), te,
                             c->d_te_pairs.as<int4>(), c->d_te_pcl.as<int2>(),
                             c->d_te_rbl.as<int2>(), c->stream));
    sync(c);
    c->te_ready = true;
}

// The reports of a run whose two-edge last level was extended in one step from every
// frequent single-edge code: built by gsk::te_emit in the reference's order and copied to
// the context's pinned buffers (report ordering and decoding then cost the host nothing).
void emit_flat(gs_ctx* c, const PatDev* d_pats, const std::vector<HPat>& pats, KeySpace ks,
               const int* d_fkey, const int* d_fsup, int F) {
    DevPool* P = &c->pool;
    const int B = (int)pats.size();
    const long long R = B + (long long)F, E = B + 2LL * F;
    const long long words = 5 * E + 2 * R;  // codes, edge counts, supports: one copy
    if (c->flat_cap < words) {
        if (c->flat_buf) HIPCHECK(hipHostFree(c->flat_buf));
        c->flat_buf = nullptr;
        c->flat_cap = 0;
        const long long n = std::max<long long>(words + words / 2, 1024);
        HIPCHECK(hipHostMalloc((void**)&c->flat_buf, sizeof(int) * n, hipHostMallocDefault));
        c->flat_cap = n;
    }
    c->flat_codes = c->flat_buf;
    c->flat_lens = c->flat_buf + 5 * E;
    c->flat_sups = c->flat_buf + 5 * E + R;
    std::vector<int> ssup(B);
    for (int i = 0; i < B; ++i) ssup[i] = pats[i].support;
    DBuf d_ssup(P, sizeof(int) * std::max(B, 1)), d_out(P, sizeof(int) * std::max(words, 1LL));
    int* d_codes = d_out.as<int>();
    h2d(c, d_ssup.as<int>(), ssup.data(), B);
    HIPCHECK(gsk::te_emit(d_pats, B, ks, d_fkey, d_fsup, F, d_ssup.as<int>(),
                          c->d_vl_vals.as<int>(), c->d_el_vals.as<int>(), d_codes,
                          d_codes + 5 * E, d_codes + 5 * E + R, c->stream));
    HIPCHECK(hipMemcpyAsync(c->flat_buf, d_codes, sizeof(int) * words, hipMemcpyDeviceToHost,
                            c->stream));
    sync(c);
    c->flat = true;
    c->flat_reports = R;
    c->flat_edges = E;
    c->st.reported += F;
}

// One extension step of a batch of size-m patterns.  write_rows: keep child rows (needed
// below the last level); all_live: write rows and sets of non-canonical children too.
int run_step(gs_ctx* c, const std::vector<HPat>& pats, const std::vector<Layer>& layers,
             bool write_rows, bool all_live, StepOut& out) {
    const int B = (int)pats.size();
    if (B == 0) return STEP_OK;
    const int m = pats[0].m;
    if (m + 1 > GS_MAXE)
        throw GsError(GS_ERR_LIMIT, "patterns larger than 7 edges are not supported");
    std::vector<PatDev> pd(B);
    long long R = 0, K = 0;
    for (int i = 0; i < B; ++i) {
        pd[i] = to_dev(pats[i]);
        pd[i].key_base = (int)std::min<long long>(K, INT_MAX);
        pd[i].brow_begin = R;
        K += key_space(pd[i].r, c->EL, c->VL);
        R += pats[i].row_count;
    }
    if (K > (1LL << 30)) {
        if (B > 1) return STEP_SPLIT;
        throw GsError(GS_ERR_LIMIT, "child key space of one pattern exceeds 2^30");
    }
    DevPool* P = &c->pool;
    DBuf d_pats(P, sizeof(PatDev) * B);
    h2d(c, d_pats.as<PatDev>(), pd.data(), B);
    LayerPtrs L{};
    for (int l = 0; l < m; ++l) {
        L.layer[l] = layers[l].p;
        L.rows[l] = layers[l].rows;
    }
    const KeySpace ks{c->EL, c->VL};

    // Two-edge last level: children of single-edge parents take their child rows and image
    // counts from per-vertex neighbour runs (gsk::te_level), with no pass over the parent
    // rows and no image sets.
    const bool two_edge = !write_rows && !all_live && m == 1 && two_edge_graph(c);
    // parent rows per pass (one prep thread per row: launches stay far below 2^32 items)
    const long long chunk =
        c->opts.chunk_rows > 0 ? std::min(c->opts.chunk_rows, 1LL << 26) : (16LL << 20);
    const long long crow = two_edge ? 1 : std::min(R, chunk);
    const int nchunks = (int)((R + chunk - 1) / chunk);
    DBuf d_ctx(P, sizeof(RowCtx) * std::max<long long>(crow, 1));
    DBuf d_work(P, sizeof(int) * std::max<long long>(crow, 1));
    DBuf d_woff(P, sizeof(long long) * (crow + 1));
    DBuf d_partial(P, sizeof(unsigned long long) * ((crow + GS_BLOCK - 1) / GS_BLOCK + 1));
    const long long tmp_elems = gsk::scan_tmp_elems(crow);
    DBuf d_tmp(P, sizeof(long long) * tmp_elems);
    DBuf d_kc(P, sizeof(int) * K);
    HIPCHECK(gsk::fill_i32(d_kc.as<int>(), 0, K, c->stream));
    // Child counts saturate (a returning atomic per flush) only when this step's items --
    // at most 1 + (r - 1) backward probes + (r + 1) forward segments of at most max degree
    // entries per parent row -- may pass INT_MAX.
    double items_bound = 0;
    for (int i = 0; i < B; ++i)
        items_bound += (double)pats[i].row_count *
                       (1.0 + (pd[i].r - 1) + (double)(pd[i].r + 1) * c->max_deg);
    const bool saturate = items_bound > (double)INT_MAX;

    // Single pass at the last level: when label-class sets for every potential child key fit
    // in a quarter of the free budget, the fill pass counts children and fills the MNI sets
    // in one go (no count pass); canonicality is then checked on the keys that occurred.
    // A pattern's potential keys: its backward keys hold the parent's vertex sets; each
    // forward (segment, edge label) holds VL keys, the parent's sets plus one set over each
    // label class.  The descriptors are written on the device (gsk::all_key_sets).
    long long all_words = 0;
    bool single = false, single_bytes = false;
    std::vector<int> lw(c->VL), lcum(c->VL + 1);
    auto class_table = [&](bool by) {
        lcum[0] = 0;
        for (int l = 0; l < c->VL; ++l) {
            lw[l] = (int)class_words(c, l, by);
            lcum[l + 1] = lcum[l] + lw[l];
        }
    };
    auto total_words = [&](bool by) {  // also places each pattern's keys (set_base)
        class_table(by);
        long long tw = 0;
        for (int i = 0; i < B; ++i) {
            long long pw = 0;
            for (int v = 0; v < pd[i].nv; ++v) pw += lw[pd[i].vlab[v]];
            const long long nb = (long long)(pd[i].r - 1) * c->EL;
            const long long nf = (long long)(pd[i].r + 1) * c->EL;
            pd[i].set_base = tw;
            tw += nb * pw + nf * ((long long)c->VL * pw + lcum[c->VL]);
        }
        return tw;
    };
    const long long te_nb = (long long)c->VL * c->EL;
    if (!two_edge && !write_rows && !all_live && !c->opts.two_pass && K < (1 << 21)) {
        const size_t avail = c->budget > c->pool.in_use() ? c->budget - c->pool.in_use() : 0;
        long long tw = 0;
        if (want_bytes(c) && ((tw = total_words(true)) <= c->tagged_words ||
                              (size_t)tw * 4 <= avail / 4)) {
            single = single_bytes = true;
        } else if (c->opts.set_kind != GS_SETS_BYTES &&
                   (size_t)(tw = total_words(false)) * 4 <= avail / 4) {
            single = true;
        }
        if (single) all_words = tw;
    }
    // The bytemaps of every key live in the context's tagged pool: each single-pass step
    // stores a new tag byte (1..255), so the pool is cleared only when it grows or the tag
    // wraps (clearing 25 GB cost 4.9 ms per run on the R-MAT scale-18 workload).  Bitmaps
    // take a pool of their own, cleared.
    DBuf d_all_sets, d_all_pool;
    uint32_t* all_pool = nullptr;
    int all_tag = 1;
    if (single) {
        DBuf d_lw(P, sizeof(int) * c->VL), d_lcum(P, sizeof(int) * (c->VL + 1));
        h2d(c, d_pats.as<PatDev>(), pd.data(), B);  // with set_base
        h2d(c, d_lw.as<int>(), lw.data(), c->VL);
        h2d(c, d_lcum.as<int>(), lcum.data(), c->VL + 1);
        if (single_bytes) {
            if (c->tagged_words < all_words) {
                c->d_tagged = DBuf();
                c->d_tagged = DBuf(P, sizeof(uint32_t) * std::max<long long>(all_words, 4));
                c->tagged_words = all_words;
                HIPCHECK(gsk::fill_u32(c->d_tagged.as<uint32_t>(), 0u, all_words, c->stream));
                c->tag = 0;
            }
            if (++c->tag > 255) {
                HIPCHECK(gsk::fill_u32(c->d_tagged.as<uint32_t>(), 0u, c->tagged_words,
                                       c->stream));
                c->tag = 1;
            }
            all_pool = c->d_tagged.as<uint32_t>();
            all_tag = c->tag;
        } else {
            d_all_pool = DBuf(P, sizeof(uint32_t) * std::max<long long>(all_words, 4));
            HIPCHECK(gsk::fill_u32(d_all_pool.as<uint32_t>(), 0u, all_words, c->stream));
            all_pool = d_all_pool.as<uint32_t>();
        }
        d_all_sets = DBuf(P, sizeof(SetDesc) * K);
        HIPCHECK(gsk::all_key_sets(d_pats.as<PatDev>(), B, ks, K, d_lw.as<int>(),
                                   d_lcum.as<int>(), single_bytes ? GS_SET_BYTEMAP : GS_SET_BITMAP,
                                   all_tag, d_all_sets.as<SetDesc>(), c->stream));
    }
    if (two_edge) single = true;  // one pass, no fill pass below; supports from d_te_sup
    DBuf d_te_sup;                  // support of every key's child
    DBuf d_te_work;                 // work of the run pass (runs visited, mask words read)
    unsigned long long te_work[2 * GS_TE_WORK_SLOTS] = {};

    long long W_total = 0;
    unsigned long long W_full = 0;
    std::vector<long long> Wc_host(nchunks, 0);
    // work items of later launches, read at the step's next synchronisation (the tuner's
    // time per item); a deque keeps the addresses of earlier entries
    std::deque<long long> launch_items;
    auto items_at = [&](const long long* d_total) -> const long long* {
        launch_items.push_back(0);
        d2h(c, &launch_items.back(), d_total, 1);
        return &launch_items.back();
    };
    HIPCHECK(gsk::fill_i32(reinterpret_cast<int*>(c->work_full64()), 0, 2, c->stream));
    // Contexts of a chunk's rows, for the run modes (COUNT, SETS: work counts runs) or the
    // neighbour modes (ROWS); the chunk and mode last prepared are remembered.
    int prepped_ci = -1;
    bool prepped_runs = false;
    auto prep_chunk = [&](int ci, bool first_pass, bool runs) -> long long {
        const long long b0 = (long long)ci * chunk;
        const long long Rc = std::min(chunk, R - b0);
        HIPCHECK(gsk::prep(c->dg, d_pats.as<PatDev>(), B, b0, Rc, L, runs, d_ctx.as<RowCtx>(),
                           d_work.as<int>(), d_partial.as<unsigned long long>(),
                           first_pass ? c->work_full64() : nullptr, c->guard_word(),
                           c->stream));
        HIPCHECK(gsk::scan_exclusive(d_work.as<int>(), d_woff.as<long long>(), Rc,
                                     d_tmp.as<long long>(), tmp_elems, c->stream));
        prepped_ci = ci;
        prepped_runs = runs;
        return Rc;
    };
    // the chunk's contexts in a mode, prepared again only when the last preparation differs
    auto chunk_ready = [&](int ci, bool runs) -> long long {
        if (prepped_ci == ci && prepped_runs == runs) return std::min(chunk, R - (long long)ci * chunk);
        return prep_chunk(ci, false, runs);
    };

    // launch shape from the tuner, and events around the launch when they are wanted
    int t_bucket = -1, t_cfg = 0;
    auto shape = [&](int mode, bool single_pass, long long Rc) {
        return c->tuner.choose(mode, single_pass, m, Rc, &t_bucket, &t_cfg);
    };
    // extents of the buffers a launch addresses, for the kernels' index guards
    auto guard_of = [&](long long slots, long long rows_cap, long long pool_words) {
        return ItemGuard{K, slots, rows_cap, pool_words, c->guard_word()};
    };
    auto time_begin = [&]() {
        if (!c->opts.timing && t_bucket < 0) return;
        ev_begin(c);
    };
    auto time_end = [&](int kind, const long long* items) {
        if (!c->opts.timing && t_bucket < 0) return;
        HIPCHECK(hipEventRecord(c->ev_pool[c->ev_used].second, c->stream));
        c->ev_pending.push_back({(int)c->ev_used, kind, t_bucket, t_cfg, items});
        ++c->ev_used;
    };

    // Last level, after a SETS launch over a chunk: the new vertices' images, inserted once
    // per group of rows (pattern, forward segment, segment vertex) by NEWV.  Rows go to the
    // groups in sub-chunks of at most kGroupPairs (row, segment) pairs: a union of groups
    // over any split of the rows gives the same sets.
    int nseg = 1;
    for (int i = 0; i < B; ++i) nseg = std::max(nseg, pd[i].r + 1);
    // last level by runs and groups, or by every neighbour of every row
    constexpr double kGroupGain = 2.5;
    const bool grouped = c->opts.groups == GS_GROUPS_ON ||
                         (c->opts.groups != GS_GROUPS_OFF && c->run_gain >= kGroupGain);
    auto newv_pass = [&](long long Rc, const int* key_slot, long long nslots, const SetDesc* sets,
                         uint32_t* pool, long long pool_words) {
        constexpr long long kGroupPairs = 1LL << 24;
        const long long sub = std::max(1LL, kGroupPairs / nseg);
        const long long maxp = std::min(Rc, sub) * nseg;
        long long slots = 1;
        while (slots < 2 * maxp) slots <<= 1;
        const long long te = gsk::scan_tmp_elems(maxp);
        DBuf tkeys(P, 8 * slots), trep(P, 4 * slots), tcnt(P, 4 * slots), tgrp(P, 4 * slots);
        DBuf pslot(P, 4 * maxp), pflag(P, 4 * maxp), pidx(P, 8 * (maxp + 1)), ptmp(P, 8 * te);
        DBuf hits(P, 4 * GS_MAXV * maxp), gctx(P, sizeof(RowCtx) * maxp), gwork(P, 4 * maxp);
        DBuf gwoff(P, 8 * (maxp + 1));
        for (long long r0 = 0; r0 < Rc; r0 += sub) {
            const long long Rs = std::min(sub, Rc - r0), np = Rs * nseg;
            HIPCHECK(gsk::fill_i32(gwork.as<int>(), 0, np, c->stream));
            HIPCHECK(gsk::build_groups(c->dg, d_pats.as<PatDev>(), d_ctx.as<RowCtx>() + r0, Rs,
                                       nseg, tkeys.as<unsigned long long>(), trep.as<int>(),
                                       tcnt.as<int>(), tgrp.as<int>(), slots, pslot.as<int>(),
                                       pflag.as<int>(), pidx.as<long long>(),
                                       ptmp.as<long long>(), te, hits.as<int>(),
                                       gctx.as<RowCtx>(), gwork.as<int>(), c->guard_word(),
                                       c->stream));
            HIPCHECK(gsk::scan_exclusive(gwork.as<int>(), gwoff.as<long long>(), np,
                                         ptmp.as<long long>(), te, c->stream));
            const gsk::ItemsLaunch lc = shape(GS_ITEMS_NEWV, false, np);
            time_begin();
            HIPCHECK(gsk::items_newv(lc, c->dg, d_pats.as<PatDev>(), gctx.as<RowCtx>(),
                                     gwoff.as<long long>(), np, ks, key_slot, sets, pool, c->tq,
                                     guard_of(nslots, 0, pool_words), c->stream));
            time_end(1, items_at(gwoff.as<long long>() + np));
            c->st.fill_launches++;
        }
    };

    // ---- learned sizes: one fill pass below the last level from predicted child counts
    if (write_rows && !all_live && c->opts.mode == GS_MODE_PREDICTED && c->mlp_trained &&
        K < (1 << 21)) {
        // predicted counts of every candidate child key
        std::vector<float> y(K);
        std::vector<unsigned char> poss(K);
        {
            DBuf d_X(P, sizeof(float) * K * GS_FEATURES), d_poss(P, K);
            DBuf d_h1(P, sizeof(float) * K * GS_HIDDEN), d_h2(P, sizeof(float) * K * GS_HIDDEN);
            DBuf d_y(P, sizeof(float) * K);
            HIPCHECK(gsk::features(d_pats.as<PatDev>(), B, ks, K, c->d_H.as<float>(),
                                   c->d_vfreq.as<float>(), c->d_efreq.as<float>(),
                                   c->d_dmean.as<float>(), (float)c->n, d_X.as<float>(),
                                   d_poss.as<unsigned char>(), c->stream));
            HIPCHECK(gsk::mlp_forward(d_X.as<float>(), K, c->mlp, d_h1.as<float>(),
                                      d_h2.as<float>(), d_y.as<float>(), c->stream));
            d2h(c, y.data(), d_y.as<float>(), K);
            d2h(c, poss.data(), d_poss.as<unsigned char>(), K);
            sync(c);
        }
        // capacities: 1.25x the prediction plus 16 rows, at most 2^30 (clamped in floating
        // point before any conversion, so a diverged network or foreign weights -- huge,
        // infinite or NaN outputs -- give large but well-defined capacities, and the step
        // then takes the exact path below); keys the rules exclude get none
        std::vector<unsigned long long> row_off(K);
        std::vector<int> cap_end(K);
        std::vector<SetDesc> psets(K);
        long long rows_total = 0, pool_words = 0;
        int p = 0;  // pattern owning `key`: keys ascend, so it only moves forward
        for (long long key = 0; key < K; ++key) {
            while (p + 1 < B && pd[p + 1].key_base <= key) ++p;
            const int slot = (int)(key - pd[p].key_base);
            const int nb_p = (pd[p].r - 1) * c->EL;
            const int nv_c = pd[p].nv + (slot < nb_p ? 0 : 1);
            int lab_c[GS_MAXV];
            for (int v = 0; v < GS_MAXV; ++v) lab_c[v] = pd[p].vlab[v];
            if (slot >= nb_p) lab_c[pd[p].nv] = (slot - nb_p) % c->VL;
            long long cap = 0;
            if (poss[key]) cap = predicted_capacity(y[key]);
            row_off[key] = (unsigned long long)std::min<long long>(rows_total, INT_MAX);
            rows_total += cap;
            cap_end[key] = (int)std::min<long long>(rows_total, INT_MAX);
            SetDesc& d = psets[key];
            if (cap > 0) {
                plan_set(c, cap, nv_c, lab_c, want_bytes(c), d, pool_words);
            } else {
                d = SetDesc{};  // no sets
            }
        }
        const size_t need = (size_t)rows_total * sizeof(int2) + (size_t)pool_words * 4 +
                            (size_t)K * (sizeof(SetDesc) + 16);
        if (rows_total <= INT_MAX && c->pool.in_use() + need <= c->budget) {
            DBuf rows(P, sizeof(int2) * std::max<long long>(rows_total, 1));
            DBuf d_cur(P, sizeof(unsigned long long) * K), d_end(P, sizeof(int) * K);
            DBuf d_psets(P, sizeof(SetDesc) * K);
            DBuf d_ppool(P, sizeof(uint32_t) * std::max<long long>(pool_words, 4));
            h2d(c, d_cur.as<unsigned long long>(), row_off.data(), K);
            h2d(c, d_end.as<int>(), cap_end.data(), K);
            h2d(c, d_psets.as<SetDesc>(), psets.data(), K);
            HIPCHECK(gsk::fill_u32(d_ppool.as<uint32_t>(), 0u, pool_words, c->stream));
            HIPCHECK(gsk::fill_i32(c->overflow_flag(), 0, 1, c->stream));
            long long Wp = 0;
            std::vector<long long> Wc_p(nchunks, 0);
            for (int ci = 0; ci < nchunks; ++ci) {
                const long long b0 = (long long)ci * chunk;
                const long long Rc = prep_chunk(ci, true, false);
                d2h(c, &Wc_p[ci], d_woff.as<long long>() + Rc, 1);
                const gsk::ItemsLaunch lc = shape(GS_ITEMS_ROWS, false, Rc);
                time_begin();
                HIPCHECK(gsk::items_fill(lc, c->dg, d_pats.as<PatDev>(), d_ctx.as<RowCtx>(),
                                         d_woff.as<long long>(), b0, Rc, ks, d_kc.as<int>(),
                                         nullptr, d_cur.as<unsigned long long>(), d_end.as<int>(),
                                         rows.as<int2>(), d_psets.as<SetDesc>(),
                                         d_ppool.as<uint32_t>(), c->overflow_flag(),
                                         c->tq, (int)K, saturate,
                                         guard_of(K, std::max<long long>(rows_total, 1),
                                                  std::max<long long>(pool_words, 4)),
                                         c->stream));
                time_end(1, &Wc_p[ci]);
                c->st.fill_launches++;
            }
            DBuf d_active(P, sizeof(int) * K), d_acnt(P, sizeof(int) * K),
                d_flags(P, sizeof(int) * K);
            compact_ordered(c, d_kc.as<int>(), K, d_active.as<int>(), d_acnt.as<int>());
            HIPCHECK(gsk::canon(d_pats.as<PatDev>(), B, ks, d_active.as<int>(), c->n_active(),
                                (int)K, m + 1, d_flags.as<int>(), c->err_flag(), c->stream));
            int na_err_ov[3] = {0, 0, 0};  // n_active, err, overflow
            d2h(c, &W_full, c->work_full64(), 1);
            ActiveHost ah;
            fetch_active(c, K, d_active.as<int>(), d_acnt.as<int>(), d_flags.as<int>(),
                         na_err_ov, 3, ah);
            resolve_timing(c);
            for (long long w : Wc_p) Wp += w;
            c->st.items += Wp;
            c->st.parent_rows += R;
            if (na_err_ov[1] & 1)
                throw GsError(GS_ERR_LIMIT, kCanonCapMsg);
            if (na_err_ov[1] & 2)
                throw GsError(GS_ERR_INTERNAL, "canonicality walk ran past the code");
            if (na_err_ov[2] == 0) {
                const int A = na_err_ov[0];
                const std::vector<int>& akeys = ah.keys;
                const std::vector<int>& acnt = ah.cnt;
                const std::vector<int>& aflag = ah.flag;
                std::vector<int> order(A);  // keys arrive in key order
                for (int i = 0; i < A; ++i) order[i] = i;
                std::vector<SetDesc> live_sets;
                long long emitted = 0;
                int p = 0;  // pattern owning `key`: keys ascend, so it only moves forward
                for (int oi = 0; oi < A; ++oi) {
                    const int a = order[oi];
                    const int key = akeys[a];
                    while (p + 1 < B && pd[p + 1].key_base <= key) ++p;
                    emitted += child_count(acnt[a]);
                    if (!aflag[a]) continue;  // other children only count as candidates
                    out.children.emplace_back();
                    ChildInfo& ci = out.children.back();
                    decode_child_into(pats[p], pd[p], key - pd[p].key_base, c->EL, c->VL,
                                      ci.pat);
                    ci.count = child_count(acnt[a]);
                    ci.canonical = 1;
                    ci.support = 0;
                    ci.row_off = (long long)row_off[key];
                    live_sets.push_back(psets[key]);
                }
                out.emitted = emitted;
                out.candidates = A;
                out.canonical = (long long)out.children.size();
                const int NL = (int)live_sets.size();
                if (NL > 0) {
                    DBuf d_ls(P, sizeof(SetDesc) * NL), d_counts(P, sizeof(int) * NL * GS_MAXV);
                    h2d(c, d_ls.as<SetDesc>(), live_sets.data(), NL);
                    HIPCHECK(gsk::set_counts(d_ls.as<SetDesc>(), NL, d_ppool.as<uint32_t>(),
                                             std::max<long long>(pool_words, 4),
                                             d_counts.as<int>(), c->guard_word(), c->stream));
                    std::vector<int> counts((size_t)NL * GS_MAXV);
                    d2h(c, counts.data(), d_counts.as<int>(), counts.size());
                    sync(c);
                    int li = 0;
                    for (auto& ch : out.children) {
                        if (!ch.canonical) continue;
                        ch.support = support_of(ch.pat, &counts[(size_t)li * GS_MAXV], c->mni_mode);
                        ch.pat.support = ch.support;
                        ch.pat.row_begin = ch.row_off;
                        ch.pat.row_count = ch.count;
                        ++li;
                    }
                }
                {
                    double bytes = 16.0 * (double)W_full + 8.0 * (double)emitted;
                    for (int i = 0; i < B; ++i)
                        bytes += (double)pats[i].row_count *
                                 (20.0 * m + 8.0 * (pd[i].r + 1) - 16.0 * (pd[i].r - 1));
                    c->st.alg_bytes_fill += bytes;
                }
                out.rows = LevelRows();
                out.rows.adopt(std::move(rows));
                out.rows_total = rows_total;
                c->st.predicted_steps++;
                c->st.batches++;
                return STEP_OK;
            }
            // a child outgrew its predicted capacity: redo this step exactly
            c->st.fallbacks++;
            out.children.clear();
            HIPCHECK(gsk::fill_i32(d_kc.as<int>(), 0, K, c->stream));
            HIPCHECK(gsk::fill_i32(reinterpret_cast<int*>(c->work_full64()), 0, 2, c->stream));
        }
    }

    // ---- pass 1: count children per key (single pass: also fill every key's sets)
    int pre_codes = -1;  // single-edge codes the pre-launched pass found on the device
    gs_ctx::TePre& tp = c->te_pre;
    if (two_edge && tp.on && c->flat_ok && B == c->flat_seeds) {
        // every frequent single-edge code: the tables are being summed already
        // (te_prelaunch); the keys read them
        tp.on = false;
        d_te_sup = DBuf(P, sizeof(int) * K);
        Wc_host[0] = c->dg.nruns;  // (vertex, run) pairs
        HIPCHECK(gsk::te_keys(d_pats.as<PatDev>(), B, ks, K, tp.cnt.as<unsigned long long>(),
                              tp.cen.as<int>(), tp.end.as<int>(), c->mni_mode, d_kc.as<int>(),
                              d_te_sup.as<int>(), c->stream));
        if (tp.ev >= 0 && tp.ev == (int)c->ev_used) time_end(1, &Wc_host[0]);
        c->st.fill_launches++;
        c->st.two_edge_levels++;
        d2h(c, te_work, tp.work.as<unsigned long long>(), 2 * GS_TE_WORK_SLOTS);
        d2h(c, &pre_codes, tp.ncodes.as<int>(), 1);
    } else if (two_edge) {
        const int NB = (int)te_nb;
        te_build(c);
        const int W = (NB + 31) / 32;
        std::vector<int> p_of((size_t)c->VL * c->EL * c->VL, -1);
        std::vector<uint32_t> pmask((size_t)c->VL * W, 0u);
        for (int i = 0; i < B; ++i) {
            const int a = pd[i].vlab[0], e0 = pd[i].cel[0], b = pd[i].vlab[1];
            p_of[((size_t)a * c->EL + e0) * c->VL + b] = i;
            const int ba = a * c->EL + e0, bb = b * c->EL + e0;  // other end seen from b, a
            pmask[(size_t)b * W + (ba >> 5)] |= 1u << (ba & 31);
            pmask[(size_t)a * W + (bb >> 5)] |= 1u << (bb & 31);
        }
        DBuf d_pof(P, sizeof(int) * p_of.size()), d_pmask(P, sizeof(uint32_t) * pmask.size());
        h2d(c, d_pof.as<int>(), p_of.data(), p_of.size());
        h2d(c, d_pmask.as<uint32_t>(), pmask.data(), pmask.size());
        const long long rows = (long long)c->VL * NB * NB, prows = 2LL * B * NB;
        DBuf d_cnt(P, 8 * prows), d_cen(P, 4 * prows), d_end(P, 4 * rows), d_q(P, 4);
        d_te_sup = DBuf(P, sizeof(int) * K);
        d_te_work = DBuf(P, sizeof(te_work));
        HIPCHECK(gsk::fill_i32(d_te_work.as<int>(), 0, 4 * GS_TE_WORK_SLOTS, c->stream));
        Wc_host[0] = c->dg.nruns;  // (vertex, run) pairs
        HIPCHECK(gsk::te_clear(ks, B, d_q.as<int>(), d_cnt.as<unsigned long long>(),
                               d_cen.as<int>(), d_end.as<int>(), c->stream));
        time_begin();
        HIPCHECK(gsk::te_level(c->dg, ks, d_pats.as<PatDev>(), B, K, d_pof.as<int>(),
                               c->d_te_m1.as<uint32_t>(), c->d_te_m2.as<uint32_t>(),
                               c->d_te_pairs.as<int4>(), c->d_te_rbl.as<int2>(),
                               d_pmask.as<uint32_t>(), d_q.as<int>(),
                               d_cnt.as<unsigned long long>(), d_cen.as<int>(), d_end.as<int>(),
                               c->mni_mode, d_kc.as<int>(), d_te_sup.as<int>(),
                               d_te_work.as<unsigned long long>(), c->stream));
        time_end(1, &Wc_host[0]);
        c->st.fill_launches++;
        c->st.two_edge_levels++;
        d2h(c, te_work, d_te_work.as<unsigned long long>(), 2 * GS_TE_WORK_SLOTS);
    }
    for (int ci = 0; ci < (two_edge ? 0 : nchunks); ++ci) {
        const long long b0 = (long long)ci * chunk;
        const bool by_items = single && !grouped;  // every neighbour of every row
        const long long Rc = prep_chunk(ci, true, !by_items);
        d2h(c, &Wc_host[ci], d_woff.as<long long>() + Rc, 1);
        const gsk::ItemsLaunch lc =
            shape(single ? (by_items ? GS_ITEMS_ROWS : GS_ITEMS_SETS) : GS_ITEMS_COUNT, single, Rc);
        time_begin();
        if (single) {
            HIPCHECK(gsk::items_fill(lc, c->dg, d_pats.as<PatDev>(), d_ctx.as<RowCtx>(),
                                     d_woff.as<long long>(), b0, Rc, ks, d_kc.as<int>(), nullptr,
                                     nullptr, nullptr, nullptr, d_all_sets.as<SetDesc>(),
                                     all_pool, c->overflow_flag(),
                                     c->tq, (int)K, saturate,
                                     guard_of(K, K, std::max<long long>(all_words, 4)),
                                     c->stream, by_items));
            time_end(1, &Wc_host[ci]);
            c->st.fill_launches++;
            if (!by_items)
                newv_pass(Rc, nullptr, K, d_all_sets.as<SetDesc>(), all_pool,
                          std::max<long long>(all_words, 4));
        } else {
            HIPCHECK(gsk::items_count(lc, c->dg, d_pats.as<PatDev>(), d_ctx.as<RowCtx>(),
                                      d_woff.as<long long>(), b0, Rc, ks, d_kc.as<int>(),
                                      c->tq, saturate, guard_of(0, 0, 0), c->stream));
            time_end(0, &Wc_host[ci]);
            c->st.count_launches++;
        }
    }
    c->st.parent_rows += R;

    // training samples for the predictor: features and exact counts of every candidate key
    if (c->logging && !single && c->n_logged < c->log_cap) {
        const long long take = std::min(K, c->log_cap - c->n_logged);
        DBuf d_X(P, sizeof(float) * K * GS_FEATURES);
        HIPCHECK(gsk::features(d_pats.as<PatDev>(), B, ks, K, c->d_H.as<float>(),
                               c->d_vfreq.as<float>(), c->d_efreq.as<float>(),
                               c->d_dmean.as<float>(), (float)c->n, d_X.as<float>(), nullptr,
                               c->stream));
        HIPCHECK(hipMemcpyAsync(c->d_logX.as<float>() + c->n_logged * GS_FEATURES,
                                d_X.as<float>(), sizeof(float) * take * GS_FEATURES,
                                hipMemcpyDeviceToDevice, c->stream));
        HIPCHECK(gsk::log_targets(d_kc.as<int>(), take, c->d_logT.as<float>() + c->n_logged,
                                  c->stream));
        c->n_logged += take;
    }

    // algorithmic bytes of the fill launches: per parent row 20*m (layer records + Edge
    // gathers) + sum over scanned vertices of (8 + 16*deg) + 8 per emitted child row
    auto account_bytes = [&](long long emitted) {
        if (two_edge) {  // per pair its 32-byte record; per run visited its 8-byte {bit,
                         // length}; per adjacency or mask word 4 B; per key 4 table reads
                         // (20 B), its count and its support (8 B)
            double runs = 0, words = 0;
            for (int t = 0; t < GS_TE_WORK_SLOTS; ++t) {
                runs += (double)te_work[2 * t];
                words += (double)te_work[2 * t + 1];
            }
            c->st.alg_bytes_fill += 32.0 * (double)c->dg.nruns + 8.0 * runs + 4.0 * words +
                                    28.0 * (double)K;
            return;
        }
        double bytes = 16.0 * (double)W_full;
        for (int i = 0; i < B; ++i)
            bytes += (double)pats[i].row_count *
                     (20.0 * m + 8.0 * (pd[i].r + 1) - 16.0 * (pd[i].r - 1));
        if (!single) c->st.alg_bytes_count += bytes;
        c->st.alg_bytes_fill += bytes + 8.0 * (double)emitted;
    };

    // A k = 2 run's step over every frequent single-edge code: one launch takes each key's
    // canonicality and its frequent support from the two-edge pass's counts and supports,
    // one compaction lists the frequent children, and their reports are built on the device
    // (no host round trip before the frequent count).
    if (two_edge && c->flat_ok && B == c->flat_seeds) {
        DBuf d_sel(P, sizeof(int) * K), d_fkey(P, sizeof(int) * K), d_fsup(P, sizeof(int) * K);
        const int nb = gsk::two_edge_select_blocks(K);
        DBuf d_part(P, sizeof(unsigned long long) * 3 * std::max(nb, 1));
        HIPCHECK(gsk::two_edge_select(d_pats.as<PatDev>(), B, ks, K, d_kc.as<int>(),
                                      d_te_sup.as<int>(), c->min_sup, d_sel.as<int>(),
                                      d_part.as<unsigned long long>(), c->stream));
        compact_ordered(c, d_sel.as<int>(), K, d_fkey.as<int>(), d_fsup.as<int>());
        int F = 0;
        std::vector<unsigned long long> part(3 * (size_t)nb);
        d2h(c, &F, c->n_active(), 1);
        d2h(c, part.data(), d_part.as<unsigned long long>(), part.size());
        sync(c);
        unsigned long long totals[3] = {0, 0, 0};  // child rows, keys with children, canonical
        for (int b = 0; b < nb; ++b)
            for (int i = 0; i < 3; ++i) totals[i] += part[3 * (size_t)b + i];
        resolve_timing(c);
        for (long long w : Wc_host) W_total += w;
        c->st.items += W_total;
        if (pre_codes >= 0 && pre_codes != B)
            throw GsError(GS_ERR_INTERNAL, "single-edge codes of the device and the host differ");
        emit_flat(c, d_pats.as<PatDev>(), pats, ks, d_fkey.as<int>(), d_fsup.as<int>(), F);
        out.candidates = (long long)totals[1];
        out.canonical = (long long)totals[2];
        out.emitted = (long long)totals[0];
        out.reported_inline = true;
        account_bytes(out.emitted);
        c->st.batches++;
        return STEP_OK;
    }

    // ---- candidates and their canonicality
    DBuf d_active(P, sizeof(int) * K), d_acnt(P, sizeof(int) * K), d_flags(P, sizeof(int) * K);
    compact_ordered(c, d_kc.as<int>(), K, d_active.as<int>(), d_acnt.as<int>());
    HIPCHECK(gsk::canon(d_pats.as<PatDev>(), B, ks, d_active.as<int>(), c->n_active(), (int)K,
                        m + 1, d_flags.as<int>(), c->err_flag(), c->stream));
    int na_err[2] = {0, 0};
    d2h(c, &W_full, c->work_full64(), 1);
    ActiveHost ah;
    fetch_active(c, K, d_active.as<int>(), d_acnt.as<int>(), d_flags.as<int>(), na_err, 2, ah,
                 all_live ? LLONG_MAX : kDeviceTailKeys);
    resolve_timing(c);
    for (long long w : Wc_host) W_total += w;
    c->st.items += W_total;
    if (na_err[1] & 1)
        throw GsError(GS_ERR_LIMIT, kCanonCapMsg);
    if (na_err[1] & 2) throw GsError(GS_ERR_INTERNAL, "canonicality walk ran past the code");
    if (pre_codes >= 0 && pre_codes != B)
        throw GsError(GS_ERR_INTERNAL, "single-edge codes of the device and the host differ");
    const int A = na_err[0];
    // ---- many candidate keys: plan, fill, supports and frequent children on the device
    if (!all_live && A > kDeviceTailKeys) {
        const auto t_plan0 = std::chrono::steady_clock::now();
        const bool plan_sets = !single;
        DBuf d_words(P, sizeof(int) * A), d_rows(P, sizeof(int) * A);
        DBuf d_soff(P, sizeof(long long) * (A + 1)), d_roff(P, sizeof(long long) * (A + 1));
        const long long te = gsk::scan_tmp_elems(A);
        DBuf d_stmp(P, sizeof(long long) * te);
        bool bytes = want_bytes(c);
        long long pool_words = 0, rows_total = 0;
        auto plan = [&](bool by) {
            HIPCHECK(gsk::plan_children(d_pats.as<PatDev>(), B, ks, d_active.as<int>(),
                                        d_acnt.as<int>(), d_flags.as<int>(), A,
                                        c->d_lcnt.as<int>(), by,
                                        plan_sets, write_rows, d_words.as<int>(),
                                        d_rows.as<int>(), c->stream));
            HIPCHECK(gsk::scan_exclusive(d_words.as<int>(), d_soff.as<long long>(), A,
                                         d_stmp.as<long long>(), te, c->stream));
            HIPCHECK(gsk::scan_exclusive(d_rows.as<int>(), d_roff.as<long long>(), A,
                                         d_stmp.as<long long>(), te, c->stream));
            d2h(c, &pool_words, d_soff.as<long long>() + A, 1);
            d2h(c, &rows_total, d_roff.as<long long>() + A, 1);
            sync(c);
        };
        auto need_of = [&]() {
            return (size_t)rows_total * sizeof(int2) + (size_t)pool_words * 4 +
                   (size_t)A * (sizeof(SetDesc) + 8 + 4 * GS_MAXV + 24) + (size_t)K * 4;
        };
        // a single-pass last level has neither rows nor sets to place
        if (plan_sets || write_rows) plan(bytes);
        if (plan_sets && bytes && c->opts.set_kind == GS_SETS_AUTO &&
            c->pool.in_use() + need_of() > c->budget) {
            bytes = false;  // bytemaps do not fit: bitmaps take an eighth of the memory
            plan(bytes);
        }
        if (rows_total >= INT_MAX) {
            if (B > 1) return STEP_SPLIT;
            throw GsError(GS_ERR_BUDGET, "child rows of one pattern exceed 2^31");
        }
        bool spill = false;
        if (c->pool.in_use() + need_of() > c->budget) {
            if (B > 1) return STEP_SPLIT;
            spill = write_rows && rows_total > 0;  // one pattern: rows to pinned host memory
        }
        DBuf d_sets(P, sizeof(SetDesc) * A), d_cursor(P, sizeof(unsigned long long) * A);
        DBuf d_slot(P, sizeof(int) * K);
        DBuf d_pool(P, sizeof(uint32_t) * std::max<long long>(plan_sets ? pool_words : 0, 4));
        if (write_rows) {
            if (spill) {
                out.rows.on_host(rows_total);
                c->st.spilled_rows += rows_total;
            } else {
                out.rows.on_device(P, rows_total);
            }
            out.rows_total = rows_total;
        }
        if (!two_edge) {  // the two-edge level's image counts come without sets
            HIPCHECK(gsk::fill_i32(d_slot.as<int>(), -1, K, c->stream));
            if (plan_sets)
                HIPCHECK(gsk::fill_u32(d_pool.as<uint32_t>(), 0u, pool_words, c->stream));
            HIPCHECK(gsk::build_sets(d_pats.as<PatDev>(), B, ks, d_active.as<int>(),
                                     d_acnt.as<int>(), d_flags.as<int>(), A, c->d_lcnt.as<int>(),
                                     bytes, d_soff.as<long long>(),
                                     write_rows ? d_roff.as<long long>() : nullptr,
                                     single ? d_all_sets.as<SetDesc>() : nullptr,
                                     d_sets.as<SetDesc>(), d_slot.as<int>(),
                                     d_cursor.as<unsigned long long>(), c->stream));
        }
        c->st.t_plan_ms += std::chrono::duration<double, std::milli>(
                               std::chrono::steady_clock::now() - t_plan0)
                               .count();
        if (!single) {  // pass 2: child rows and MNI sets
            for (int ci = 0; ci < nchunks; ++ci) {
                const long long b0 = (long long)ci * chunk;
                const bool runs = !write_rows && grouped;
                const long long Rc = chunk_ready(ci, runs);
                const gsk::ItemsLaunch lc = shape(runs ? GS_ITEMS_SETS : GS_ITEMS_ROWS, false, Rc);
                time_begin();
                HIPCHECK(gsk::items_fill(
                    lc, c->dg, d_pats.as<PatDev>(), d_ctx.as<RowCtx>(), d_woff.as<long long>(),
                    b0, Rc, ks, nullptr, d_slot.as<int>(), d_cursor.as<unsigned long long>(),
                    nullptr, write_rows ? out.rows.ptr() : nullptr, d_sets.as<SetDesc>(),
                    d_pool.as<uint32_t>(), c->overflow_flag(), c->tq, A, false,
                    guard_of(A, write_rows ? rows_total : 0, std::max<long long>(pool_words, 4)),
                    c->stream, !runs));
                time_end(1, items_at(d_woff.as<long long>() + Rc));
                c->st.fill_launches++;
                if (runs)
                    newv_pass(Rc, d_slot.as<int>(), A, d_sets.as<SetDesc>(),
                              d_pool.as<uint32_t>(), std::max<long long>(pool_words, 4));
            }
        }
        // supports, then the frequent children in key order
        DBuf d_counts(P, two_edge ? 4 : sizeof(int) * (size_t)A * GS_MAXV);
        DBuf d_sup(P, sizeof(int) * A);
        if (!two_edge)
            HIPCHECK(gsk::set_counts(d_sets.as<SetDesc>(), A,
                                     single ? all_pool : d_pool.as<uint32_t>(),
                                     std::max<long long>(single ? all_words : pool_words, 4),
                                     d_counts.as<int>(), c->guard_word(), c->stream,
                                     d_acnt.as<int>(), c->min_sup));
        HIPCHECK(gsk::fill_i32(reinterpret_cast<int*>(c->step_totals()), 0, 4, c->stream));
        HIPCHECK(gsk::supports(d_pats.as<PatDev>(), B, ks, d_active.as<int>(), d_acnt.as<int>(),
                               d_flags.as<int>(), A, d_counts.as<int>(),
                               two_edge ? d_te_sup.as<int>() : nullptr, c->mni_mode, c->min_sup,
                               d_sup.as<int>(), c->step_totals(), c->stream));
        DBuf d_fidx(P, sizeof(int) * A), d_fsup(P, sizeof(int) * A);
        compact_ordered(c, d_sup.as<int>(), A, d_fidx.as<int>(), d_fsup.as<int>());
        DBuf d_fkey(P, sizeof(int) * A), d_fcnt(P, sizeof(int) * A), d_frow(P, sizeof(long long) * A);
        HIPCHECK(gsk::gather_frequent(d_fidx.as<int>(), c->n_active(), A, d_active.as<int>(),
                                      d_acnt.as<int>(),
                                      write_rows ? d_roff.as<long long>() : nullptr,
                                      d_fkey.as<int>(), d_fcnt.as<int>(), d_frow.as<long long>(),
                                      c->stream));
        int F = 0;
        unsigned long long totals[2] = {0, 0};
        d2h(c, &F, c->n_active(), 1);
        d2h(c, totals, c->step_totals(), 2);
        sync(c);
        resolve_timing(c);
        c->flat_ok = false;  // reports of this run go through the host
        std::vector<int> fkey(F), fcnt(F), fsup(F);
        std::vector<long long> frow(F);
        d2h(c, fkey.data(), d_fkey.as<int>(), F);
        d2h(c, fcnt.data(), d_fcnt.as<int>(), F);
        d2h(c, fsup.data(), d_fsup.as<int>(), F);
        d2h(c, frow.data(), d_frow.as<long long>(), F);
        sync(c);
        const auto t_dec0 = std::chrono::steady_clock::now();
        if (!c->child_bufs.empty()) {
            out.children = std::move(c->child_bufs.back());
            c->child_bufs.pop_back();
            out.children.clear();
        }
        int p = 0;  // pattern owning `key`: keys ascend, so it only moves forward
        if (!write_rows) {  // the last level: frequent children are reported, not extended
            // (every child here is frequent: its support passed min_sup on the device);
            // decoded on the host threads, each range from its own first pattern
            const size_t base = c->reports.size();
            c->reports.resize(base + F);
            parallel_ranges(F, [&](long long i0, long long i1) {
                if (i0 >= i1) return;
                int q = owner_of(pd, fkey[i0]);
                for (long long i = i0; i < i1; ++i) {
                    const int key = fkey[i];
                    while (q + 1 < B && pd[q + 1].key_base <= key) ++q;
                    c->reports[base + i] =
                        child_report(c, pats[q], pd[q], key - pd[q].key_base, fsup[i]);
                }
            });
            c->st.reported += F;
            F = 0;
            out.reported_inline = true;
        }
        out.children.reserve(F);
        for (int i = 0; i < F; ++i) {
            const int key = fkey[i];
            while (p + 1 < B && pd[p + 1].key_base <= key) ++p;
            out.children.emplace_back();
            ChildInfo& ch = out.children.back();
            decode_child_into(pats[p], pd[p], key - pd[p].key_base, c->EL, c->VL, ch.pat);
            ch.count = child_count(fcnt[i]);
            ch.canonical = 1;
            ch.support = fsup[i];
            ch.row_off = write_rows ? frow[i] : -1;
            ch.pat.support = ch.support;
            ch.pat.row_begin = ch.row_off;
            ch.pat.row_count = ch.count;
        }
        out.candidates = A;
        out.canonical = (long long)totals[1];
        out.emitted = (long long)totals[0];
        c->st.t_plan_ms += std::chrono::duration<double, std::milli>(
                               std::chrono::steady_clock::now() - t_dec0)
                               .count();
        account_bytes(out.emitted);
        c->st.batches++;
        return STEP_OK;
    }
    const std::vector<int>& akeys = ah.keys;
    const std::vector<int>& acnt = ah.cnt;
    const std::vector<int>& aflag = ah.flag;
    // keys arrive in key order = (parent, slot) order
    std::vector<int> order(A);
    for (int i = 0; i < A; ++i) order[i] = i;

    // ---- plan child rows and sets of the live keys
    const auto t_plan0 = std::chrono::steady_clock::now();
    std::vector<int> live_keys, live_child;
    std::vector<unsigned long long> cursor;
    std::vector<SetDesc> sets;
    if (!c->child_bufs.empty()) {
        out.children = std::move(c->child_bufs.back());
        c->child_bufs.pop_back();
        out.children.clear();
    }
    long long rows_total = 0, pool_words = 0, emitted = 0, canon_cnt = 0;
    int p = 0;  // pattern owning `key`: keys ascend, so it only moves forward
    for (int oi = 0; oi < A; ++oi) {
        const int a = order[oi];
        const int key = akeys[a];
        while (p + 1 < B && pd[p + 1].key_base <= key) ++p;
        emitted += child_count(acnt[a]);
        canon_cnt += aflag[a] ? 1 : 0;
        if (!(aflag[a] || all_live)) continue;  // other children only count as candidates
        out.children.emplace_back();  // built in place: a ChildInfo is ~170 bytes
        ChildInfo& ci = out.children.back();
        decode_child_into(pats[p], pd[p], key - pd[p].key_base, c->EL, c->VL, ci.pat);
        ci.count = child_count(acnt[a]);
        ci.canonical = aflag[a];
        ci.support = 0;
        ci.row_off = -1;
        live_child.push_back((int)out.children.size() - 1);
        live_keys.push_back(key);
        if (write_rows) {
            ci.row_off = rows_total;
            cursor.push_back((unsigned long long)rows_total);
            rows_total += ci.count;
        } else {
            cursor.push_back(0);
        }
    }
    out.emitted = emitted;
    out.candidates = A;
    out.canonical = canon_cnt;
    const int NL = (int)live_keys.size();
    if (rows_total > INT_MAX) {
        if (B > 1) return STEP_SPLIT;
        throw GsError(GS_ERR_BUDGET, "child rows of one pattern exceed 2^31");
    }
    auto plan_sets = [&](bool bytes) {
        sets.clear();
        pool_words = 0;
        for (int li = 0; li < NL; ++li) {
            SetDesc d;
            const ChildInfo& ch = out.children[live_child[li]];
            if (two_edge) {
                d = SetDesc{};  // no sets: image counts come from the run tables
            } else if (single) {
                const int key = live_keys[li];
                const int p = owner_of(pd, key);
                all_key_set(pd[p], key - pd[p].key_base, KeySpace{c->EL, c->VL}, lw.data(),
                            lcum.data(), single_bytes ? GS_SET_BYTEMAP : GS_SET_BITMAP, all_tag,
                            &d);
            } else {
                int lab[GS_MAXV];
                ch.pat.labels(lab);
                plan_set(c, ch.count, ch.pat.nv(), lab, bytes, d, pool_words);
            }
            sets.push_back(d);
        }
    };
    auto need_of = [&]() {
        return (size_t)rows_total * sizeof(int2) + (size_t)pool_words * 4 +
               (size_t)NL * (sizeof(SetDesc) + 40) + (size_t)K * 4;
    };
    plan_sets(want_bytes(c));
    if (!single && want_bytes(c) && c->opts.set_kind == GS_SETS_AUTO &&
        c->pool.in_use() + need_of() > c->budget)
        plan_sets(false);  // bytemaps do not fit: bitmaps take an eighth of the memory
    const size_t need = need_of();
    bool spill = false;
    if (c->pool.in_use() + need > c->budget) {
        if (B > 1) return STEP_SPLIT;
        // one pattern: its child rows go to pinned host memory; its sets stay on the
        // device (a set is at most a byte per graph vertex; an allocation the device
        // cannot hold fails with GS_ERR_HIP)
        spill = write_rows && rows_total > 0;
    }
    c->st.t_plan_ms +=
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_plan0)
            .count();
    if (NL == 0) {
        c->st.batches++;
        return STEP_OK;
    }

    // ---- pass 2: write child rows and fill the MNI sets (two-pass mode)
    DBuf d_sets(P, sizeof(SetDesc) * NL);
    DBuf d_counts(P, sizeof(int) * NL * GS_MAXV);
    DBuf d_pool;
    h2d(c, d_sets.as<SetDesc>(), sets.data(), NL);
    if (!single) {
        DBuf d_slot(P, sizeof(int) * K);
        DBuf d_live(P, sizeof(int) * NL), d_cursor(P, sizeof(unsigned long long) * NL);
        d_pool = DBuf(P, sizeof(uint32_t) * std::max<long long>(pool_words, 4));
        if (write_rows) {
            if (spill) {
                out.rows.on_host(rows_total);
                c->st.spilled_rows += rows_total;
            } else {
                out.rows.on_device(P, rows_total);
            }
            out.rows_total = rows_total;
        }
        HIPCHECK(gsk::fill_i32(d_slot.as<int>(), -1, K, c->stream));
        h2d(c, d_live.as<int>(), live_keys.data(), NL);
        h2d(c, d_cursor.as<unsigned long long>(), cursor.data(), NL);
        HIPCHECK(gsk::scatter_slots(d_slot.as<int>(), d_live.as<int>(), NL, c->stream));
        HIPCHECK(gsk::fill_u32(d_pool.as<uint32_t>(), 0u, pool_words, c->stream));
        for (int ci = 0; ci < nchunks; ++ci) {
            const long long b0 = (long long)ci * chunk;
            const bool runs = !write_rows && grouped;
            const long long Rc = chunk_ready(ci, runs);
            const gsk::ItemsLaunch lc = shape(runs ? GS_ITEMS_SETS : GS_ITEMS_ROWS, false, Rc);
            time_begin();
            HIPCHECK(gsk::items_fill(lc, c->dg, d_pats.as<PatDev>(), d_ctx.as<RowCtx>(),
                                     d_woff.as<long long>(), b0, Rc, ks, nullptr,
                                     d_slot.as<int>(), d_cursor.as<unsigned long long>(), nullptr,
                                     write_rows ? out.rows.ptr() : nullptr,
                                     d_sets.as<SetDesc>(), d_pool.as<uint32_t>(),
                                     c->overflow_flag(), c->tq, NL, false,
                                     guard_of(NL, write_rows ? rows_total : 0,
                                              std::max<long long>(pool_words, 4)),
                                     c->stream, !runs));
            time_end(1, items_at(d_woff.as<long long>() + Rc));
            c->st.fill_launches++;
            if (runs)
                newv_pass(Rc, d_slot.as<int>(), NL, d_sets.as<SetDesc>(), d_pool.as<uint32_t>(),
                          std::max<long long>(pool_words, 4));
        }
    }
    account_bytes(emitted);
    // Mining reads only the sets of children with at least min_sup rows (a support never
    // exceeds the child count); gs_extend (all_live) wants every child's support.
    std::vector<int> live_rows;
    DBuf d_lrows;
    if (!all_live) {
        live_rows.reserve(NL);
        for (const auto& ch : out.children)
            if (ch.canonical) live_rows.push_back((int)std::min<long long>(ch.count, INT_MAX));
        d_lrows = DBuf(P, sizeof(int) * std::max(NL, 1));
        h2d(c, d_lrows.as<int>(), live_rows.data(), NL);
    }
    DBuf d_lkeys;
    if (two_edge) {
        d_lkeys = DBuf(P, sizeof(int) * std::max(NL, 1));
        h2d(c, d_lkeys.as<int>(), live_keys.data(), NL);
        HIPCHECK(gsk::te_gather(d_te_sup.as<int>(), d_lkeys.as<int>(), nullptr, NL,
                                d_counts.as<int>(), c->stream));  // one support per key
    } else {
        HIPCHECK(gsk::set_counts(d_sets.as<SetDesc>(), NL,
                                 single ? all_pool : d_pool.as<uint32_t>(),
                                 std::max<long long>(single ? all_words : pool_words, 4),
                                 d_counts.as<int>(), c->guard_word(), c->stream,
                                 all_live ? nullptr : d_lrows.as<int>(), c->min_sup));
    }
    std::vector<int> counts((size_t)NL * GS_MAXV);
    d2h(c, counts.data(), d_counts.as<int>(), two_edge ? (size_t)NL : counts.size());
    sync(c);
    resolve_timing(c);
    int li = 0;
    for (auto& ch : out.children) {
        if (!(ch.canonical || all_live)) continue;
        ch.support = !all_live && ch.count < c->min_sup ? 0
                     : two_edge ? counts[li]
                                : support_of(ch.pat, &counts[(size_t)li * GS_MAXV], c->mni_mode);
        ch.pat.support = ch.support;
        ch.pat.row_begin = ch.row_off;
        ch.pat.row_count = ch.count;
        ++li;
    }
    c->st.batches++;
    return STEP_OK;
}

void mine_level(gs_ctx* c, std::vector<HPat>& pats, std::vector<Layer>& layers);

// Extends a batch, reports its frequent canonical children and descends into them.
void expand(gs_ctx* c, std::vector<HPat>& pats, std::vector<Layer>& layers) {
    const int m = pats[0].m;
    const bool last = m + 1 >= c->k;
    StepOut so;
    if (run_step(c, pats, layers, !last, false, so) == STEP_SPLIT) {
        // split by rows into two halves and run them one after the other
        long long tot = 0, half = 0;
        for (auto& p : pats) tot += p.row_count;
        size_t cut = 0;
        while (cut + 1 < pats.size() && half + pats[cut].row_count <= tot / 2)
            half += pats[cut++].row_count;
        if (cut == 0) cut = 1;
        std::vector<HPat> a(pats.begin(), pats.begin() + cut), b(pats.begin() + cut, pats.end());
        expand(c, a, layers);
        expand(c, b, layers);
        return;
    }
    c->st.extended += so.emitted;
    c->st.candidates += so.candidates;
    c->st.canonical += so.canonical;
    std::vector<HPat> next;
    for (auto& ch : so.children) {  // canonical children (the others only count)
        if (so.reported_inline || ch.support < c->min_sup) continue;
        ch.pat.report_id = add_report(c, ch.pat);
        if (!last) next.push_back(ch.pat);
    }
    if (!last && !next.empty()) {
        layers.push_back(Layer{so.rows.ptr(), so.rows_total});
        mine_level(c, next, layers);
        layers.pop_back();
    }
    // kept for reuse: one array per DFS level (at most 7 are alive at once), each of at
    // most 2^18 candidates (~45 MB); larger ones are freed
    if (c->child_bufs.size() < 8 && so.children.capacity() <= (1u << 18))
        c->child_bufs.push_back(std::move(so.children));
}

void mine_level(gs_ctx* c, std::vector<HPat>& pats, std::vector<Layer>& layers) {
    if (pats.empty() || pats[0].m >= c->k) return;
    if (c->opts.mode == GS_MODE_PLAIN) {
        // one parent at a time, in the order the reference visits them
        std::vector<size_t> idx(pats.size());
        for (size_t i = 0; i < idx.size(); ++i) idx[i] = i;
        std::vector<OrderKey> keys(pats.size());
        for (size_t i = 0; i < pats.size(); ++i) keys[i] = order_key(pats[i]);
        std::sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return keys[a] < keys[b]; });
        for (size_t i : idx) {
            std::vector<HPat> one{pats[i]};
            expand(c, one, layers);
        }
    } else {
        expand(c, pats, layers);
    }
}

// Single-edge codes (gspan.cu:1176-1190): counted, sized and filled in two passes over
// the CSR entries; every code of size 1 is canonical (gspan.cu:592-593).
// Mining to k = 2 with the two-edge last level reads no single-edge rows, and each
// single-edge code's rows and images follow from the graph's run classes: code (a, e0, b),
// a <= b, has as many rows as CSR entries from label a to label b over e0 (gspan.cu:410-426),
// the vertices of label a with a run (b, e0) as vertex-0 images and those of label b with a
// run (a, e0) as vertex-1 images.  Codes in key order, as mine_seeds reports them.
// The two-edge pass over every frequent single-edge code, launched from the class sums
// (d_np, d_nl) while the host lists the codes: the codes' flags, batch indexes and end-bit
// masks on the device (gsk::te_seeds), the tables cleared for as many codes as it flags,
// then the run pass (gsk::te_accum).  Skipped when tables for every code a <= b would not
// fit the budget (the step then launches its own pass).
void te_prelaunch(gs_ctx* c, const int* d_np, const unsigned long long* d_nl) {
    const int VL = c->VL, EL = c->EL, NB = VL * EL, W = (NB + 31) / 32;
    const long long NS = (long long)VL * EL * VL;              // single-edge code space
    const long long maxp = (long long)EL * VL * (VL + 1) / 2;  // codes with a <= b
    const long long rows = (long long)VL * NB * NB, prows = 2 * maxp * NB;
    const long long te = gsk::scan_tmp_elems(NS);
    const size_t need = (size_t)(12 * prows + 4 * rows + 16 * NS + 8 * te + 4 * VL * W) + 4096;
    if (c->pool.in_use() + need > c->budget) return;
    DevPool* P = &c->pool;
    gs_ctx::TePre& t = c->te_pre;
    t.flags = DBuf(P, 4 * NS);
    t.scan = DBuf(P, 8 * NS);
    t.tmp = DBuf(P, 8 * te);
    t.pof = DBuf(P, 4 * NS);
    t.pmask = DBuf(P, 4 * (size_t)VL * W);
    t.ncodes = DBuf(P, 4);
    HIPCHECK(gsk::te_seeds(d_np, d_nl, VL, EL, c->mni_mode, c->min_sup, t.flags.as<int>(),
                           t.scan.as<long long>(), t.tmp.as<long long>(), te, t.pof.as<int>(),
                           t.pmask.as<uint32_t>(), t.ncodes.as<int>(), c->stream));
    t.cnt = DBuf(P, 8 * prows);
    t.cen = DBuf(P, 4 * prows);
    t.end = DBuf(P, 4 * rows);
    t.q = DBuf(P, 4);
    t.work = DBuf(P, sizeof(unsigned long long) * 2 * GS_TE_WORK_SLOTS);
    HIPCHECK(gsk::fill_i32(t.work.as<int>(), 0, 4 * GS_TE_WORK_SLOTS, c->stream));
    const KeySpace ks{EL, VL};
    HIPCHECK(gsk::te_clear_dev(ks, t.ncodes.as<int>(), maxp, t.q.as<int>(),
                               t.cnt.as<unsigned long long>(), t.cen.as<int>(),
                               t.end.as<int>(), c->stream));
    t.ev = -1;
    if (c->opts.timing) {
        ev_begin(c);
        t.ev = (int)c->ev_used;
    }
    HIPCHECK(gsk::te_accum(c->dg, ks, t.pof.as<int>(), c->d_te_m1.as<uint32_t>(),
                           c->d_te_m2.as<uint32_t>(), c->d_te_pairs.as<int4>(),
                           c->d_te_rbl.as<int2>(), t.pmask.as<uint32_t>(), t.q.as<int>(),
                           t.cnt.as<unsigned long long>(), t.cen.as<int>(), t.end.as<int>(),
                           t.work.as<unsigned long long>(), c->stream));
    t.on = true;
}

void seeds_from_runs(gs_ctx* c) {
    te_build(c);
    const int VL = c->VL, EL = c->EL, NB = VL * EL;
    const long long NC = (long long)VL * NB;
    DevPool* P = &c->pool;
    DBuf d_np(P, sizeof(int) * NC), d_nl(P, sizeof(long long) * NC);
    HIPCHECK(gsk::te_classes(c->d_te_pcl.as<int2>(), c->dg.nruns,
                             c->d_te_cstart.as<long long>(), NC, d_np.as<int>(),
                             d_nl.as<unsigned long long>(), c->stream));
    if (c->h_te_cap < NC) {
        if (c->h_te_np) HIPCHECK(hipHostFree(c->h_te_np));
        if (c->h_te_nl) HIPCHECK(hipHostFree(c->h_te_nl));
        c->h_te_np = nullptr;
        c->h_te_nl = nullptr;
        c->h_te_cap = 0;
        HIPCHECK(hipHostMalloc((void**)&c->h_te_np, sizeof(int) * NC, hipHostMallocDefault));
        HIPCHECK(hipHostMalloc((void**)&c->h_te_nl, sizeof(long long) * NC,
                               hipHostMallocDefault));
        c->h_te_cap = NC;
    }
    HIPCHECK(hipMemcpyAsync(c->h_te_np, d_np.as<int>(), sizeof(int) * NC, hipMemcpyDeviceToHost,
                            c->stream));
    HIPCHECK(hipMemcpyAsync(c->h_te_nl, d_nl.as<unsigned long long>(), sizeof(long long) * NC,
                            hipMemcpyDeviceToHost, c->stream));
    if (!c->ev_seeds) HIPCHECK(hipEventCreateWithFlags(&c->ev_seeds, hipEventDisableTiming));
    HIPCHECK(hipEventRecord(c->ev_seeds, c->stream));
    // with mining to k = 2 every frequent code is extended in one step: its run pass starts
    // now and overlaps the host's work below
    if (c->k == 2) te_prelaunch(c, d_np.as<int>(), d_nl.as<unsigned long long>());
    HIPCHECK(hipEventSynchronize(c->ev_seeds));
    const int* np = c->h_te_np;
    const unsigned long long* nl = c->h_te_nl;
    auto pairs_of = [&](long long cl) { return (long long)np[cl]; };
    std::vector<HPat> freq;
    long long codes = 0, rows = 0;
    for (int a = 0; a < VL; ++a) {
        for (int e0 = 0; e0 < EL; ++e0) {
            for (int b = a; b < VL; ++b) {
                const long long cl = (long long)a * NB + (long long)b * EL + e0;
                const long long cnt = (long long)nl[cl];
                if (cnt == 0) continue;
                ++codes;
                HPat h;
                h.m = 1;
                h.f[0] = 0;
                h.t[0] = 1;
                h.fl0 = a;
                h.el[0] = e0;
                h.tl[0] = b;
                const long long n1 = pairs_of((long long)b * NB + (long long)a * EL + e0);
                int counts[GS_MAXV] = {(int)pairs_of(cl), (int)n1};
                h.support = support_of(h, counts, c->mni_mode);
                if (h.support < c->min_sup) continue;
                h.row_begin = rows;
                h.row_count = cnt;
                rows += cnt;
                h.report_id = add_report(c, h);
                freq.push_back(h);
            }
        }
    }
    c->st.candidates += codes;
    c->st.canonical += codes;
    if (!freq.empty()) {
        std::vector<Layer> layers{Layer{nullptr, rows}};
        c->flat_ok = true;  // one step over every frequent code may build the reports
        c->flat_seeds = (int)freq.size();
        mine_level(c, freq, layers);
    }
    c->te_pre = gs_ctx::TePre{};  // a pass no step took (the batch was split) is dropped
}

void mine_seeds(gs_ctx* c) {
    if (c->k == 2 && two_edge_graph(c)) return seeds_from_runs(c);
    DevPool* P = &c->pool;
    const long long K = (long long)c->VL * c->EL * c->VL;
    if (K > (1LL << 30)) throw GsError(GS_ERR_LIMIT, "too many labels for the seed key space");
    DBuf d_kc(P, sizeof(int) * K);
    HIPCHECK(gsk::fill_i32(d_kc.as<int>(), 0, K, c->stream));
    HIPCHECK(gsk::fill_i32(reinterpret_cast<int*>(c->total64()), 0, 2, c->stream));
    HIPCHECK(gsk::seed_count(c->dg, c->EL, c->VL, d_kc.as<int>(), c->total64(), c->stream));
    DBuf d_active(P, sizeof(int) * K), d_acnt(P, sizeof(int) * K);
    compact_ordered(c, d_kc.as<int>(), K, d_active.as<int>(), d_acnt.as<int>());
    int A = 0;
    ActiveHost ah;
    fetch_active(c, K, d_active.as<int>(), d_acnt.as<int>(), nullptr, &A, 1, ah);
    const std::vector<int>& akeys = ah.keys;
    const std::vector<int>& acnt = ah.cnt;
    std::vector<int> order(A);  // keys arrive in key order
    for (int i = 0; i < A; ++i) order[i] = i;
    const bool rows = c->k > 1;
    std::vector<HPat> seeds;
    std::vector<int> live, cursor;
    std::vector<SetDesc> sets;
    long long rows_total = 0, pool_words = 0;
    for (int oi = 0; oi < A; ++oi) {
        const int a = order[oi];
        const int key = akeys[a];
        HPat h;
        h.m = 1;
        h.f[0] = 0;
        h.t[0] = 1;
        h.fl0 = key / (c->EL * c->VL);
        h.el[0] = (key / c->VL) % c->EL;
        h.tl[0] = key % c->VL;
        h.row_begin = rows_total;
        h.row_count = acnt[a];
        SetDesc d;
        const int lab[2] = {h.fl0, h.tl[0]};
        plan_set(c, acnt[a], 2, lab, want_bytes(c), d, pool_words);
        sets.push_back(d);
        live.push_back(key);
        cursor.push_back((int)rows_total);
        if (rows) rows_total += acnt[a];
        seeds.push_back(h);
    }
    if (rows_total > INT_MAX) throw GsError(GS_ERR_BUDGET, "single-edge rows exceed 2^31");
    if (want_bytes(c) && c->opts.set_kind == GS_SETS_AUTO &&
        c->pool.in_use() + (size_t)pool_words * 4 + (size_t)rows_total * 8 > c->budget) {
        pool_words = 0;  // bytemaps do not fit: plan bitmaps
        for (size_t i = 0; i < sets.size(); ++i) {
            const int lab[2] = {seeds[i].fl0, seeds[i].tl[0]};
            plan_set(c, seeds[i].row_count, 2, lab, false, sets[i], pool_words);
        }
    }
    const int NL = (int)live.size();
    c->st.candidates += NL;
    c->st.canonical += NL;
    if (NL == 0) return;
    DBuf d_slot(P, sizeof(int) * K), d_live(P, sizeof(int) * NL), d_cursor(P, sizeof(int) * NL);
    DBuf d_sets(P, sizeof(SetDesc) * NL);
    DBuf d_pool(P, sizeof(uint32_t) * std::max<long long>(pool_words, 4));
    DBuf d_counts(P, sizeof(int) * NL * GS_MAXV);
    LevelRows d_rows;
    if (rows) {
        if (c->pool.in_use() + (size_t)rows_total * sizeof(int2) > c->budget) {
            d_rows.on_host(rows_total);  // spilled to pinned host memory
            c->st.spilled_rows += rows_total;
        } else {
            d_rows.on_device(P, rows_total);
        }
    }
    HIPCHECK(gsk::fill_i32(d_slot.as<int>(), -1, K, c->stream));
    h2d(c, d_live.as<int>(), live.data(), NL);
    h2d(c, d_cursor.as<int>(), cursor.data(), NL);
    h2d(c, d_sets.as<SetDesc>(), sets.data(), NL);
    HIPCHECK(gsk::scatter_slots(d_slot.as<int>(), d_live.as<int>(), NL, c->stream));
    HIPCHECK(gsk::fill_u32(d_pool.as<uint32_t>(), 0u, pool_words, c->stream));
    HIPCHECK(gsk::seed_fill(c->dg, c->EL, c->VL, d_slot.as<int>(), d_cursor.as<int>(),
                            rows ? d_rows.ptr() : nullptr, d_sets.as<SetDesc>(),
                            d_pool.as<uint32_t>(),
                            ItemGuard{K, NL, rows_total, std::max<long long>(pool_words, 4),
                                      c->guard_word()},
                            c->stream));
    HIPCHECK(gsk::set_counts(d_sets.as<SetDesc>(), NL, d_pool.as<uint32_t>(),
                             std::max<long long>(pool_words, 4), d_counts.as<int>(),
                             c->guard_word(), c->stream));
    std::vector<int> counts((size_t)NL * GS_MAXV);
    d2h(c, counts.data(), d_counts.as<int>(), counts.size());
    sync(c);
    d_pool.reset();
    std::vector<HPat> freq;
    for (int i = 0; i < NL; ++i) {
        seeds[i].support = support_of(seeds[i], &counts[(size_t)i * GS_MAXV], c->mni_mode);
        if (seeds[i].support < c->min_sup) continue;
        seeds[i].report_id = add_report(c, seeds[i]);
        freq.push_back(seeds[i]);
    }
    if (rows && !freq.empty()) {
        std::vector<Layer> layers{Layer{d_rows.ptr(), rows_total}};
        mine_level(c, freq, layers);
    }
}

void set_error(gs_ctx* c, const std::string& msg) {
    if (c) c->err = msg;
}

// After a failed call: the destinations of its queued device-to-host copies may be gone,
// so the copies are waited for and dropped.
void drop_pending(gs_ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->pending_d2h.clear();
    c->stage.reset();
    // recorded launch timings point at item counts of the failed call's frames
    c->ev_pending.clear();
    c->ev_used = 0;
}

int guard(gs_ctx* c, const std::function<void()>& fn) {
    try {
        fn();
        return GS_OK;
    } catch (const GsError& e) {
        set_error(c, e.what());
        drop_pending(c);
        return e.code;
    } catch (const std::bad_alloc&) {
        set_error(c, "host allocation failed");
        drop_pending(c);
        return GS_ERR_INTERNAL;
    } catch (const std::exception& e) {
        set_error(c, e.what());
        drop_pending(c);
        return GS_ERR_INTERNAL;
    }
}

HPat code_to_hpat(gs_ctx* c, const int* code5, int m) {
    if (m < 1 || m > GS_MAXE) throw GsError(GS_ERR_LIMIT, "pattern size out of range");
    HPat h;
    h.m = m;
    for (int j = 0; j < m; ++j) {
        const int* e = code5 + 5 * j;
        h.f[j] = e[0];
        h.t[j] = e[1];
        if (e[0] < 0 || e[1] < 0 || e[0] >= GS_MAXV || e[1] >= GS_MAXV || e[0] == e[1])
            throw GsError(GS_ERR_INVALID, "bad pattern vertex ids");
        auto ei = std::lower_bound(c->el_vals.begin(), c->el_vals.end(), e[3]);
        if (ei == c->el_vals.end() || *ei != e[3])
            throw GsError(GS_ERR_INVALID, "edge label not present in the graph");
        h.el[j] = int(ei - c->el_vals.begin());
        if (e[0] < e[1]) {
            auto vi = std::lower_bound(c->vl_vals.begin(), c->vl_vals.end(), e[4]);
            if (vi == c->vl_vals.end() || *vi != e[4])
                throw GsError(GS_ERR_INVALID, "vertex label not present in the graph");
            h.tl[j] = int(vi - c->vl_vals.begin());
        } else {
            h.tl[j] = -1;
        }
    }
    auto fi = std::lower_bound(c->vl_vals.begin(), c->vl_vals.end(), code5[2]);
    if (fi == c->vl_vals.end() || *fi != code5[2])
        throw GsError(GS_ERR_INVALID, "vertex label not present in the graph");
    h.fl0 = int(fi - c->vl_vals.begin());
    if (h.f[0] != 0 || h.t[0] != 1) throw GsError(GS_ERR_INVALID, "code must start with (0, 1)");
    // DFS-code shape: forward edges reach the next new vertex from a known one, backward
    // edges join two known vertices (every vertex image is then defined by the rows)
    int nv = 2;
    for (int j = 1; j < m; ++j) {
        if (h.f[j] < h.t[j]) {
            if (h.t[j] != nv || h.f[j] >= nv) throw GsError(GS_ERR_INVALID, "bad forward edge");
            ++nv;
        } else if (h.f[j] >= nv) {
            throw GsError(GS_ERR_INVALID, "bad backward edge");
        }
    }
    return h;
}

std::vector<DBuf> upload_layers(gs_ctx* c, int m, const int* const* layers,
                                const long long* counts, std::vector<Layer>& ptrs) {
    std::vector<DBuf> bufs;
    ptrs.clear();
    for (int l = 0; l < m; ++l) {
        if (counts[l] < 0 || counts[l] > INT_MAX || !layers[l])
            throw GsError(GS_ERR_INVALID, "layer size out of range");
        // every row must point at an existing parent row and CSR entry
        const int* rw = layers[l];
        for (long long i = 0; i < counts[l]; ++i) {
            const int prev = rw[2 * i], eid = rw[2 * i + 1];
            if (eid < 0 || eid >= c->nnz) throw GsError(GS_ERR_INVALID, "edge id out of range");
            if (l > 0 && (prev < 0 || prev >= counts[l - 1]))
                throw GsError(GS_ERR_INVALID, "prev_idx out of range");
        }
        bufs.emplace_back(&c->pool, sizeof(int2) * std::max<long long>(counts[l], 1));
        h2d(c, bufs.back().as<int2>(), reinterpret_cast<const int2*>(layers[l]), counts[l]);
        ptrs.push_back(Layer{bufs.back().as<int2>(), counts[l]});
    }
    return bufs;
}

}  // namespace

namespace gsk {
int ctx_device(const gs_ctx* c) { return c->device; }
hipStream_t ctx_stream(const gs_ctx* c) { return c->stream; }
void ctx_set_error(gs_ctx* c, const char* msg) { c->err = msg; }
}  // namespace gsk

// ================================================================== C-ABI

extern "C" {

const char* gs_version(void) { return "gspan_amd 0.1 gfx950"; }

int gs_default_opts(gs_opts* o) {
    if (!o) return GS_ERR_INVALID;
    std::memset(o, 0, sizeof(*o));
    o->mode = GS_MODE_BATCHED;
    o->mni_mode = GS_MNI_GPU;
    o->autotune = 1;
    return GS_OK;
}

int gs_create(const int* offsets, const int* vlabels, const int* edges3, int n, int nnz,
              const gs_opts* opts, gs_ctx** out) {
    if (!out) return GS_ERR_INVALID;
    *out = nullptr;
    if (n < 0 || nnz < 0 || (n > 0 && (!offsets || !vlabels)) || (nnz > 0 && !edges3) ||
        (n == 0 && nnz != 0))
        return GS_ERR_INVALID;
    std::unique_ptr<gs_ctx> c(new gs_ctx());
    if (opts) c->opts = *opts;
    else gs_default_opts(&c->opts);
    int rc = guard(c.get(), [&]() {
        gs_ctx* x = c.get();
        // graph validation: the CSR invariants CSRGraph(Graph&) establishes (graph.cc:66-112)
        if (n > 0 && offsets[0] != 0) throw GsError(GS_ERR_INVALID, "offsets[0] must be 0");
        if (n > 0 && offsets[n] != nnz) throw GsError(GS_ERR_INVALID, "offsets[n] != nnz");
        for (int v = 0; v < n; ++v) {
            if (offsets[v + 1] < offsets[v]) throw GsError(GS_ERR_INVALID, "offsets not monotone");
            for (int i = offsets[v]; i < offsets[v + 1]; ++i) {
                const int* e = edges3 + 3 * (size_t)i;
                if (e[0] != v) throw GsError(GS_ERR_INVALID, "edge.from does not match its row");
                if (e[1] < 0 || e[1] >= n) throw GsError(GS_ERR_INVALID, "edge.to out of range");
                if (i > offsets[v]) {
                    const int* p = e - 3;
                    if (p[1] > e[1] || (p[1] == e[1] && p[2] > e[2]))
                        throw GsError(GS_ERR_INVALID, "adjacency not sorted by (to, label)");
                }
            }
        }
        x->n = n;
        x->nnz = nnz;
        for (int v = 0; v < n; ++v) x->max_deg = std::max(x->max_deg, offsets[v + 1] - offsets[v]);
        // simple graph: each adjacency ascends strictly by neighbour and skips its vertex
        x->dg.simple = 1;
        for (int v = 0; v < n && x->dg.simple; ++v) {
            for (int i = offsets[v], prev = -1; i < offsets[v + 1]; ++i) {
                const int to = edges3[3 * (long long)i + 1];
                if (to == v || to <= prev) {
                    x->dg.simple = 0;
                    break;
                }
                prev = to;
            }
        }
        LabelRanks vranks, eranks;
        vranks.build(vlabels, n, 1);
        eranks.build(edges3 + 2, nnz, 3);
        x->vl_vals = vranks.vals;
        x->el_vals = eranks.vals;
        if (x->vl_vals.empty()) x->vl_vals.push_back(0);
        if (x->el_vals.empty()) x->el_vals.push_back(0);
        if (x->vl_vals.size() > 65535 || x->el_vals.size() > 65535)
            throw GsError(GS_ERR_LIMIT, "more than 65535 distinct labels");
        x->VL = (int)x->vl_vals.size();
        x->EL = (int)x->el_vals.size();

        HIPCHECK(hipSetDevice(x->opts.device));
        x->device = x->opts.device;
        HIPCHECK(gsk::noalloc_init(x->device));
        HIPCHECK(hipStreamCreateWithFlags(&x->stream, hipStreamNonBlocking));
        HIPCHECK(hipEventCreate(&x->ev0));
        HIPCHECK(hipEventCreate(&x->ev1));
        hipDeviceProp_t prop;
        HIPCHECK(hipGetDeviceProperties(&prop, x->device));
        x->cus = prop.multiProcessorCount;
        if (x->opts.tile != 0 && x->opts.tile != 256 && x->opts.tile != 512 &&
            x->opts.tile != 1024)
            throw GsError(GS_ERR_INVALID, "tile must be 0, 256, 512 or 1024");
        x->tuner.init(x->cus, x->opts.autotune != 0, x->opts.tile);
        size_t fr = 0, tot = 0;
        HIPCHECK(hipMemGetInfo(&fr, &tot));
        x->budget = x->opts.budget_bytes ? (size_t)x->opts.budget_bytes : (size_t)(fr * 0.85);

        std::vector<int> vr(n), src(nnz);
        std::vector<uint2> adj(nnz);
        parallel_ranges(n, [&](long long a, long long b) {
            for (long long v = a; v < b; ++v) vr[v] = vranks.rank(vlabels[v]);
        });
        parallel_ranges(nnz, [&](long long a, long long b) {
            for (long long i = a; i < b; ++i) {
                const int* e = edges3 + 3 * (size_t)i;
                src[i] = e[0];
                adj[i].x = (unsigned)e[1];
                adj[i].y = ((unsigned)eranks.rank(e[2]) << 16) | (unsigned)vr[e[1]];
            }
        });
        // each vertex's entries ordered by (neighbour label, edge label), with their CSR
        // indices: the entries of one child key are then consecutive work items
        std::vector<uint2> adjl(nnz);
        std::vector<int> eidl(nnz);
        auto lab_el = [&](int i) { return ((adj[i].y & 0xffffu) << 16) | (adj[i].y >> 16); };
        parallel_ranges(n, [&](long long va, long long vb) {
            for (long long v = va; v < vb; ++v) {
                const int o0 = offsets[v], o1 = offsets[v + 1];
                for (int i = o0; i < o1; ++i) eidl[i] = i;
                std::stable_sort(eidl.begin() + o0, eidl.begin() + o1,
                                 [&](int a, int b) { return lab_el(a) < lab_el(b); });
                for (int i = o0; i < o1; ++i) adjl[i] = adj[eidl[i]];
            }
        });
        // label classes: each vertex's rank among the vertices of its label (ascending id)
        x->lcnt.assign(x->VL, 0);
        std::vector<int> lidx(n), adjlx(nnz);
        for (int v = 0; v < n; ++v) lidx[v] = x->lcnt[vr[v]]++;
        for (int i = 0; i < nnz; ++i) adjlx[i] = lidx[adjl[i].x];
        // runs: maximal ranges of a vertex's adjl entries with one (label, edge label)
        std::vector<int> rstart, rkey, runid(nnz);
        for (int v = 0; v < n; ++v) {
            const int o0 = offsets[v], o1 = offsets[v + 1];
            for (int i = o0; i < o1; ++i) {
                const int k = (int)(((adjl[i].y & 0xffffu) << 16) | (adjl[i].y >> 16));
                if (i == o0 || k != rkey.back()) {
                    rstart.push_back(i);
                    rkey.push_back(k);
                }
                runid[i] = (int)rkey.size() - 1;
            }
        }
        const int nruns = (int)rkey.size();
        rstart.push_back(nnz);
        {
            double num = 0, den = 0;
            for (int v = 0; v < n; ++v) {
                const int o0 = offsets[v], o1 = offsets[v + 1];
                if (o1 == o0) continue;
                const double d = o1 - o0;
                num += d * d;
                den += d * (runid[o1 - 1] - runid[o0] + 1);
            }
            x->run_gain = den > 0 ? num / den : 1.0;
        }
        // label start table, when it is at most 1 GiB
        const long long lstride = (long long)x->VL + 1;
        std::vector<int> lstart;
        if ((long long)n * lstride * 4 <= (1LL << 30)) {
            lstart.resize((size_t)n * lstride);
            for (int v = 0; v < n; ++v) {
                int at = offsets[v];
                const int o1 = offsets[v + 1];
                for (int l = 0; l <= x->VL; ++l) {
                    while (at < o1 && (int)(adjl[at].y & 0xffffu) < l) ++at;
                    lstart[(size_t)v * lstride + l] = at;
                }
            }
        }
        DevPool* P = &x->pool;
        x->d_offsets = DBuf(P, sizeof(int) * (n + 1));
        const int zero = 0;
        if (n == 0) offsets = &zero;
        x->d_src = DBuf(P, sizeof(int) * std::max(nnz, 1));
        x->d_adj = DBuf(P, sizeof(uint2) * std::max(nnz, 1));
        x->d_adjl = DBuf(P, sizeof(uint2) * std::max(nnz, 1));
        x->d_eidl = DBuf(P, sizeof(int) * std::max(nnz, 1));
        if (!lstart.empty()) x->d_lstart = DBuf(P, sizeof(int) * lstart.size());
        x->d_vlab = DBuf(P, sizeof(int) * std::max(n, 1));
        x->d_lidx = DBuf(P, sizeof(int) * std::max(n, 1));
        x->d_adjlx = DBuf(P, sizeof(int) * std::max(nnz, 1));
        x->d_lcnt = DBuf(P, sizeof(int) * x->VL);
        x->d_rstart = DBuf(P, sizeof(int) * (nruns + 1));
        x->d_rkey = DBuf(P, sizeof(int) * std::max(nruns, 1));
        x->d_runid = DBuf(P, sizeof(int) * std::max(nnz, 1));
        x->d_misc = DBuf(P, 64);
        HIPCHECK(gsk::fill_i32(x->d_misc.as<int>(), 0, 16, x->stream));
        x->d_tiles = DBuf(P, sizeof(int) * 2 * GS_TILE_QUEUE);
        h2d(x, x->d_offsets.as<int>(), offsets, n + 1);
        h2d(x, x->d_src.as<int>(), src.data(), nnz);
        h2d(x, x->d_adj.as<uint2>(), adj.data(), nnz);
        h2d(x, x->d_adjl.as<uint2>(), adjl.data(), nnz);
        h2d(x, x->d_eidl.as<int>(), eidl.data(), nnz);
        if (!lstart.empty()) h2d(x, x->d_lstart.as<int>(), lstart.data(), lstart.size());
        h2d(x, x->d_vlab.as<int>(), vr.data(), n);
        h2d(x, x->d_lidx.as<int>(), lidx.data(), n);
        h2d(x, x->d_adjlx.as<int>(), adjlx.data(), nnz);
        h2d(x, x->d_lcnt.as<int>(), x->lcnt.data(), x->VL);
        h2d(x, x->d_rstart.as<int>(), rstart.data(), nruns + 1);
        h2d(x, x->d_rkey.as<int>(), rkey.data(), nruns);
        h2d(x, x->d_runid.as<int>(), runid.data(), nnz);
        init_label_stats(x, vr, src, adj);
        mlp_init(x);
        sync(x);
        x->dg.n = n;
        x->dg.nnz = nnz;
        x->dg.offsets = x->d_offsets.as<int>();
        x->dg.src = x->d_src.as<int>();
        x->dg.adj = x->d_adj.as<uint2>();
        x->dg.adjl = x->d_adjl.as<uint2>();
        x->dg.eidl = x->d_eidl.as<int>();
        x->dg.lstart = lstart.empty() ? nullptr : x->d_lstart.as<int>();
        x->dg.lstride = (int)lstride;
        x->dg.vlab = x->d_vlab.as<int>();
        x->dg.lidx = x->d_lidx.as<int>();
        x->dg.adjlx = x->d_adjlx.as<int>();
        x->dg.lcnt = x->d_lcnt.as<int>();
        x->dg.rstart = x->d_rstart.as<int>();
        x->dg.rkey = x->d_rkey.as<int>();
        x->dg.runid = x->d_runid.as<int>();
        x->dg.nruns = nruns;
        x->d_vl_vals = DBuf(&x->pool, sizeof(int) * x->vl_vals.size());
        x->d_el_vals = DBuf(&x->pool, sizeof(int) * x->el_vals.size());
        h2d(x, x->d_vl_vals.as<int>(), x->vl_vals.data(), x->vl_vals.size());
        h2d(x, x->d_el_vals.as<int>(), x->el_vals.data(), x->el_vals.size());
        if (two_edge_graph(x)) te_build(x);
        sync(x);
    });
    if (rc != GS_OK) {
        std::fprintf(stderr, "gs_create: %s\n", c->err.c_str());
        gs_destroy(c.release());
        return rc;
    }
    *out = c.release();
    return GS_OK;
}

int gs_mine(gs_ctx* c, int k, int min_support, gs_report_cb cb, void* user) {
    if (!c || k < 1) return GS_ERR_INVALID;
    if (k > GS_MAXE) {
        c->err = "k > 7 is not supported";
        return GS_ERR_LIMIT;
    }
    return guard(c, [&]() {
        HIPCHECK(hipSetDevice(c->device));
        auto t0 = std::chrono::steady_clock::now();
        c->k = k;
        c->min_sup = min_support;
        c->mni_mode = c->opts.mni_mode;
        c->reports.clear();
        c->flat = c->flat_ok = false;
        c->te_pre = gs_ctx::TePre{};
        c->report_order.clear();
        clear_guard(c);
        c->reset_tile_queues();
        mine_seeds(c);
        const auto t_ord = std::chrono::steady_clock::now();
        if (!c->flat) order_reports(c);
        c->st.t_order_ms += std::chrono::duration<double, std::milli>(
                                std::chrono::steady_clock::now() - t_ord)
                                .count();
        auto t1 = std::chrono::steady_clock::now();
        c->st.t_mine_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        c->st.peak_device_bytes = c->pool.peak();
        if (cb) {
            bool stop = false;
            for_each_report(c, [&](const int* code5, int m, int support) {
                if (!stop && cb(code5, m, support, user) != 0) stop = true;
            });
            if (stop) throw GsError(GS_ERR_CALLBACK, "report callback stopped mining");
        }
    });
}

int gs_support(gs_ctx* c, const int* code5, int m, const int* const* layers,
               const long long* counts, int mni_mode, int* support_out) {
    if (!c || !code5 || !layers || !counts || !support_out) return GS_ERR_INVALID;
    return guard(c, [&]() {
        HIPCHECK(hipSetDevice(c->device));
        HPat h = code_to_hpat(c, code5, m);
        std::vector<Layer> ptrs;
        clear_guard(c);
        std::vector<DBuf> bufs = upload_layers(c, m, layers, counts, ptrs);
        PatDev d = to_dev(h);
        LayerPtrs L{};
        for (int l = 0; l < m; ++l) {
            L.layer[l] = ptrs[l].p;
            L.rows[l] = ptrs[l].rows;
        }
        DBuf d_pat(&c->pool, sizeof(PatDev));
        h2d(c, d_pat.as<PatDev>(), &d, 1);
        long long words = 0;
        SetDesc sd;
        int lab[GS_MAXV];
        h.labels(lab);
        plan_set(c, counts[m - 1], h.nv(), lab, want_bytes(c), sd, words);
        DBuf d_set(&c->pool, sizeof(SetDesc)), d_pool(&c->pool, 4 * std::max(words, 4LL));
        DBuf d_counts(&c->pool, sizeof(int) * GS_MAXV);
        h2d(c, d_set.as<SetDesc>(), &sd, 1);
        HIPCHECK(gsk::fill_u32(d_pool.as<uint32_t>(), 0u, words, c->stream));
        HIPCHECK(gsk::support_rows(c->dg, d_pat.as<PatDev>(), counts[m - 1], L, mni_mode,
                                   d_set.as<SetDesc>(), d_pool.as<uint32_t>(),
                                   std::max(words, 4LL), c->guard_word(), c->stream));
        HIPCHECK(gsk::set_counts(d_set.as<SetDesc>(), 1, d_pool.as<uint32_t>(),
                                 std::max(words, 4LL), d_counts.as<int>(), c->guard_word(),
                                 c->stream));
        int cnt[GS_MAXV];
        d2h(c, cnt, d_counts.as<int>(), GS_MAXV);
        sync(c);
        *support_out = support_of(h, cnt, mni_mode);
    });
}

int gs_extend(gs_ctx* c, const int* code5, int m, const int* const* layers,
              const long long* counts, int mni_mode, gs_child_cb cb, void* user,
              long long* total) {
    if (!c || !code5 || !layers || !counts) return GS_ERR_INVALID;
    return guard(c, [&]() {
        HIPCHECK(hipSetDevice(c->device));
        if (m + 1 > GS_MAXE) throw GsError(GS_ERR_LIMIT, "pattern too large to extend");
        HPat h = code_to_hpat(c, code5, m);
        h.row_begin = 0;
        h.row_count = counts[m - 1];
        std::vector<Layer> ptrs;
        clear_guard(c);
        std::vector<DBuf> bufs = upload_layers(c, m, layers, counts, ptrs);
        c->mni_mode = mni_mode;
        c->reset_tile_queues();
        std::vector<HPat> one{h};
        StepOut so;
        if (run_step(c, one, ptrs, true, true, so) != STEP_OK)
            throw GsError(GS_ERR_BUDGET, "extension does not fit the device budget");
        if (total) *total = so.emitted;
        // reference child order: backward ascending (v_i, el), forward descending v_i
        std::vector<size_t> idx(so.children.size());
        for (size_t i = 0; i < idx.size(); ++i) idx[i] = i;
        std::vector<OrderKey> keys(idx.size());
        for (size_t i = 0; i < idx.size(); ++i) keys[i] = order_key(so.children[i].pat);
        std::sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return keys[a] < keys[b]; });
        long long rows_total = 0;
        for (auto& ch : so.children) rows_total += ch.count;
        std::vector<int2> rows((size_t)rows_total);
        d2h(c, rows.data(), so.rows.ptr(), (size_t)rows_total);
        sync(c);
        for (size_t i : idx) {
            const ChildInfo& ch = so.children[i];
            const int j = m;
            int e5[5] = {ch.pat.f[j], ch.pat.t[j], -1, c->el_vals[ch.pat.el[j]],
                         ch.pat.tl[j] < 0 ? -1 : c->vl_vals[ch.pat.tl[j]]};
            if (cb)
                cb(e5, reinterpret_cast<const int*>(rows.data() + ch.row_off), ch.count,
                   ch.canonical, ch.support, user);
        }
    });
}

int gs_is_canonical_batch(gs_ctx* c, const int* codes5, const int* lens, int ncodes,
                          int* flags) {
    if (!c || (ncodes > 0 && (!codes5 || !lens || !flags))) return GS_ERR_INVALID;
    return guard(c, [&]() {
        HIPCHECK(hipSetDevice(c->device));
        // canonicality depends only on the order of labels: rank them per call
        std::vector<int> vls, els;
        long long off = 0;
        for (int i = 0; i < ncodes; ++i) {
            if (lens[i] < 1 || lens[i] > GS_MAXE)
                throw GsError(GS_ERR_LIMIT, "code length out of range");
            for (int j = 0; j < lens[i]; ++j) {
                const int* e = codes5 + 5 * (off + j);
                if (e[2] >= 0) vls.push_back(e[2]);
                if (e[4] >= 0) vls.push_back(e[4]);
                els.push_back(e[3]);
            }
            off += lens[i];
        }
        std::sort(vls.begin(), vls.end());
        vls.erase(std::unique(vls.begin(), vls.end()), vls.end());
        std::sort(els.begin(), els.end());
        els.erase(std::unique(els.begin(), els.end()), els.end());
        auto rk = [](const std::vector<int>& v, int x) {
            return int(std::lower_bound(v.begin(), v.end(), x) - v.begin());
        };
        const int stride = 1 + 5 * GS_MAXE;
        std::vector<int> packed((size_t)ncodes * stride, 0);
        off = 0;
        for (int i = 0; i < ncodes; ++i) {
            int* p = packed.data() + (size_t)i * stride;
            p[0] = lens[i];
            for (int j = 0; j < lens[i]; ++j) {
                const int* e = codes5 + 5 * (off + j);
                if (e[0] < 0 || e[1] < 0 || e[0] >= GS_MAXV || e[1] >= GS_MAXV)
                    throw GsError(GS_ERR_INVALID, "bad pattern vertex ids");
                int* q = p + 1 + 5 * j;
                q[0] = e[0];
                q[1] = e[1];
                q[2] = e[2] >= 0 ? rk(vls, e[2]) : -1;
                q[3] = rk(els, e[3]);
                q[4] = e[4] >= 0 ? rk(vls, e[4]) : -1;
            }
            off += lens[i];
        }
        DBuf d_codes(&c->pool, sizeof(int) * std::max<size_t>(packed.size(), 1));
        DBuf d_flags(&c->pool, sizeof(int) * std::max(ncodes, 1));
        h2d(c, d_codes.as<int>(), packed.data(), packed.size());
        HIPCHECK(gsk::fill_i32(c->err_flag(), 0, 1, c->stream));
        int max_m = 0;
        for (int i = 0; i < ncodes; ++i) max_m = std::max(max_m, lens[i]);
        HIPCHECK(gsk::canon_codes(d_codes.as<int>(), ncodes, max_m, d_flags.as<int>(),
                                  c->err_flag(), c->stream));
        int e = 0;
        d2h(c, flags, d_flags.as<int>(), ncodes);
        d2h(c, &e, c->err_flag(), 1);
        sync(c);
        if (e & 1) throw GsError(GS_ERR_LIMIT, kCanonCapMsg);
        if (e & 2) throw GsError(GS_ERR_INTERNAL, "canonicality walk ran past the code");
    });
}

// samples per network pass: one thread per (sample, hidden unit) in launches of fewer than
// 2^32 work-items
constexpr long long kMaxSamples = 1LL << 24;

int gs_predictor_log(gs_ctx* c, int on, long long capacity) {
    if (!c || capacity < 0 || capacity > kMaxSamples) return GS_ERR_INVALID;
    return guard(c, [&]() {
        HIPCHECK(hipSetDevice(c->device));
        c->logging = on != 0;
        if (on && capacity != c->log_cap) {
            c->d_logX = DBuf(&c->pool, sizeof(float) * std::max<long long>(capacity, 1) * GS_FEATURES);
            c->d_logT = DBuf(&c->pool, sizeof(float) * std::max<long long>(capacity, 1));
            c->log_cap = capacity;
            c->n_logged = 0;
        }
    });
}

int gs_predictor_samples(const gs_ctx* c, long long* count) {
    if (!c || !count) return GS_ERR_INVALID;
    *count = c->n_logged;
    return GS_OK;
}

namespace {
// One forward + backward over n samples already on the device; gradients into d_grads,
// mean squared error into *loss_dev.
void mlp_grad_device(gs_ctx* c, const float* X, const float* T, long long n, DBuf& h1, DBuf& h2,
                     DBuf& y, DBuf& dy, DBuf& dh2, DBuf& dh1, float* loss_dev) {
    HIPCHECK(gsk::fill_i32(c->d_grads.as<int>(), 0, GS_MLP_PARAMS, c->stream));
    HIPCHECK(gsk::fill_i32(reinterpret_cast<int*>(loss_dev), 0, 1, c->stream));
    HIPCHECK(gsk::mlp_forward(X, n, c->mlp, h1.as<float>(), h2.as<float>(), y.as<float>(),
                              c->stream));
    HIPCHECK(gsk::mlp_backward(X, T, n, c->mlp, h1.as<float>(), h2.as<float>(), y.as<float>(),
                               dy.as<float>(), dh2.as<float>(), dh1.as<float>(),
                               c->d_grads.as<float>(), loss_dev, c->stream));
}

struct MlpScratch {
    DBuf h1, h2, y, dy, dh2, dh1, loss;
    MlpScratch(DevPool* P, long long n)
        : h1(P, sizeof(float) * n * GS_HIDDEN), h2(P, sizeof(float) * n * GS_HIDDEN),
          y(P, sizeof(float) * n), dy(P, sizeof(float) * n), dh2(P, sizeof(float) * n * GS_HIDDEN),
          dh1(P, sizeof(float) * n * GS_HIDDEN), loss(P, sizeof(float)) {}
};
}  // namespace

int gs_predictor_train(gs_ctx* c, int epochs, float lr, float* loss) {
    if (!c || epochs < 0 || !(lr > 0)) return GS_ERR_INVALID;
    return guard(c, [&]() {
        HIPCHECK(hipSetDevice(c->device));
        const long long n = c->n_logged;
        if (n == 0) throw GsError(GS_ERR_INVALID, "no logged samples to train on");
        MlpScratch sc(&c->pool, n);
        float l = 0.f;
        for (int e = 0; e < epochs; ++e) {
            mlp_grad_device(c, c->d_logX.as<float>(), c->d_logT.as<float>(), n, sc.h1, sc.h2,
                            sc.y, sc.dy, sc.dh2, sc.dh1, sc.loss.as<float>());
            HIPCHECK(gsk::mlp_adam(c->mlp, c->d_grads.as<float>(), c->d_adam_m.as<float>(),
                                   c->d_adam_v.as<float>(), ++c->mlp_step, lr, c->stream));
        }
        // loss at the final weights
        mlp_grad_device(c, c->d_logX.as<float>(), c->d_logT.as<float>(), n, sc.h1, sc.h2, sc.y,
                        sc.dy, sc.dh2, sc.dh1, sc.loss.as<float>());
        d2h(c, &l, sc.loss.as<float>(), 1);
        sync(c);
        if (loss) *loss = l;
        c->mlp_trained = true;
    });
}

int gs_predictor_get_weights(const gs_ctx* c, float* weights) {
    if (!c || !weights) return GS_ERR_INVALID;
    gs_ctx* m = const_cast<gs_ctx*>(c);
    return guard(m, [&]() {
        d2h(m, weights, c->d_w32.as<float>(), GS_MLP_PARAMS);
        sync(m);
    });
}

int gs_predictor_set_weights(gs_ctx* c, const float* weights) {
    if (!c || !weights) return GS_ERR_INVALID;
    return guard(c, [&]() {
        h2d(c, c->d_w32.as<float>(), weights, GS_MLP_PARAMS);
        HIPCHECK(gsk::mlp_sync_bf16(c->mlp, c->stream));
        HIPCHECK(gsk::fill_i32(c->d_adam_m.as<int>(), 0, GS_MLP_PARAMS, c->stream));
        HIPCHECK(gsk::fill_i32(c->d_adam_v.as<int>(), 0, GS_MLP_PARAMS, c->stream));
        sync(c);
        c->mlp_step = 0;
        c->mlp_trained = true;
    });
}

int gs_mlp_forward(gs_ctx* c, const float* x, long long n, float* y) {
    if (!c || n < 0 || n > kMaxSamples || (n > 0 && (!x || !y))) return GS_ERR_INVALID;
    return guard(c, [&]() {
        if (n == 0) return;
        DBuf dX(&c->pool, sizeof(float) * n * GS_FEATURES);
        MlpScratch sc(&c->pool, n);
        h2d(c, dX.as<float>(), x, (size_t)n * GS_FEATURES);
        HIPCHECK(gsk::mlp_forward(dX.as<float>(), n, c->mlp, sc.h1.as<float>(), sc.h2.as<float>(),
                                  sc.y.as<float>(), c->stream));
        d2h(c, y, sc.y.as<float>(), (size_t)n);
        sync(c);
    });
}

int gs_mlp_gradients(gs_ctx* c, const float* x, const float* t, long long n, float* grads,
                     float* loss) {
    if (!c || n <= 0 || n > kMaxSamples || !x || !t || !grads) return GS_ERR_INVALID;
    return guard(c, [&]() {
        DBuf dX(&c->pool, sizeof(float) * n * GS_FEATURES), dT(&c->pool, sizeof(float) * n);
        MlpScratch sc(&c->pool, n);
        h2d(c, dX.as<float>(), x, (size_t)n * GS_FEATURES);
        h2d(c, dT.as<float>(), t, (size_t)n);
        mlp_grad_device(c, dX.as<float>(), dT.as<float>(), n, sc.h1, sc.h2, sc.y, sc.dy, sc.dh2,
                        sc.dh1, sc.loss.as<float>());
        d2h(c, grads, c->d_grads.as<float>(), GS_MLP_PARAMS);
        float l = 0.f;
        d2h(c, &l, sc.loss.as<float>(), 1);
        sync(c);
        if (loss) *loss = l;
    });
}

int gs_prune_graph(gs_ctx* c, int min_sup, int* n_out, int* nnz_out, int* offsets_out,
                   int* vlabels_out, int* edges3_out) {
    if (!c || !n_out || !nnz_out || !offsets_out || (c->n > 0 && !vlabels_out) ||
        (c->nnz > 0 && !edges3_out))
        return GS_ERR_INVALID;
    return guard(c, [&]() {
        HIPCHECK(hipSetDevice(c->device));
        DevPool* P = &c->pool;
        const long long n = c->n;
        DBuf d_vl(P, sizeof(int) * c->VL), d_el(P, sizeof(int) * c->EL);
        DBuf d_hist(P, sizeof(int) * c->VL), d_keep(P, sizeof(int) * std::max(n, 1LL)),
            d_kdeg(P, sizeof(int) * std::max(n, 1LL));
        DBuf d_newid(P, sizeof(long long) * (n + 1)), d_newoff(P, sizeof(long long) * (n + 1));
        const long long te = gsk::scan_tmp_elems(n);
        DBuf d_tmp(P, sizeof(long long) * te);
        DBuf d_off(P, sizeof(int) * (n + 1)), d_vlab(P, sizeof(int) * std::max(n, 1LL)),
            d_edges(P, sizeof(int) * 3 * std::max(c->nnz, 1));
        h2d(c, d_vl.as<int>(), c->vl_vals.data(), c->VL);
        h2d(c, d_el.as<int>(), c->el_vals.data(), c->EL);
        HIPCHECK(gsk::prune_graph(c->dg, c->VL, min_sup, d_vl.as<int>(), d_el.as<int>(),
                                  d_hist.as<int>(), d_keep.as<int>(), d_kdeg.as<int>(),
                                  d_newid.as<long long>(), d_newoff.as<long long>(),
                                  d_tmp.as<long long>(), te, d_off.as<int>(), d_vlab.as<int>(),
                                  d_edges.as<int>(), c->stream));
        long long sizes[2] = {0, 0};  // kept vertices, kept edges
        if (n > 0) {
            d2h(c, &sizes[0], d_newid.as<long long>() + n, 1);
            d2h(c, &sizes[1], d_newoff.as<long long>() + n, 1);
        }
        sync(c);
        *n_out = (int)sizes[0];
        *nnz_out = (int)sizes[1];
        d2h(c, offsets_out, d_off.as<int>(), (size_t)sizes[0] + 1);
        d2h(c, vlabels_out, d_vlab.as<int>(), (size_t)sizes[0]);
        d2h(c, edges3_out, d_edges.as<int>(), (size_t)sizes[1] * 3);
        sync(c);
    });
}

int gs_tuner_report(const gs_ctx* c, char* buf, int cap, int* len) {
    if (!c || !len || cap < 0 || (cap > 0 && !buf)) return GS_ERR_INVALID;
    const std::string s = c->tuner.report();
    *len = (int)s.size();
    if (cap > 0) {
        const size_t k = std::min(s.size(), (size_t)cap - 1);
        std::memcpy(buf, s.data(), k);
        buf[k] = 0;
    }
    return GS_OK;
}

int gs_result_count(const gs_ctx* c, long long* count, long long* total_edges) {
    if (!c || !count || !total_edges) return GS_ERR_INVALID;
    if (c->flat) {
        *count = c->flat_reports;
        *total_edges = c->flat_edges;
        return GS_OK;
    }
    long long e = 0;
    for (int i : c->report_order) e += c->reports[i].m;
    *count = (long long)c->report_order.size();
    *total_edges = e;
    return GS_OK;
}

int gs_result_copy(const gs_ctx* c, int* codes5, int* lens, int* supports) {
    if (!c || (!c->reports.empty() && (!codes5 || !lens || !supports))) return GS_ERR_INVALID;
    if (c->flat) {
        std::memcpy(codes5, c->flat_codes, sizeof(int) * 5 * c->flat_edges);
        std::memcpy(lens, c->flat_lens, sizeof(int) * c->flat_reports);
        std::memcpy(supports, c->flat_sups, sizeof(int) * c->flat_reports);
        return GS_OK;
    }
    long long at = 0, i = 0;
    for_each_report(c, [&](const int* code5, int m, int support) {
        std::memcpy(codes5 + at, code5, sizeof(int) * 5 * m);
        at += 5LL * m;
        lens[i] = m;
        supports[i] = support;
        ++i;
    });
    return GS_OK;
}

int gs_get_stats(const gs_ctx* c, gs_stats* out) {
    if (!c || !out) return GS_ERR_INVALID;
    *out = c->st;
    return GS_OK;
}

int gs_reset_stats(gs_ctx* c) {
    if (!c) return GS_ERR_INVALID;
    c->st = gs_stats{};
    return GS_OK;
}

const char* gs_last_error(const gs_ctx* c) { return c ? c->err.c_str() : "null context"; }

int gs_destroy(gs_ctx* c) {
    if (!c) return GS_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->d_offsets.reset();
    c->d_src.reset();
    c->d_adj.reset();
    c->d_adjl.reset();
    c->d_eidl.reset();
    c->d_lstart.reset();
    c->d_lidx.reset();
    c->d_adjlx.reset();
    c->d_lcnt.reset();
    c->d_rstart.reset();
    c->d_rkey.reset();
    c->d_runid.reset();
    c->d_tagged.reset();
    c->d_te_m1.reset();
    c->d_te_m2.reset();
    c->d_te_pairs.reset();
    c->d_te_pcl.reset();
    c->d_te_cstart.reset();
    c->d_te_rbl.reset();
    c->d_vl_vals.reset();
    c->d_el_vals.reset();
    if (c->flat_buf) (void)hipHostFree(c->flat_buf);
    c->te_pre = gs_ctx::TePre{};
    if (c->h_te_np) (void)hipHostFree(c->h_te_np);
    if (c->h_te_nl) (void)hipHostFree(c->h_te_nl);
    if (c->ev_seeds) (void)hipEventDestroy(c->ev_seeds);
    c->d_vlab.reset();
    c->d_misc.reset();
    c->d_tiles.reset();
    for (DBuf* b : {&c->d_H, &c->d_vfreq, &c->d_efreq, &c->d_dmean, &c->d_w32, &c->d_w16,
                    &c->d_adam_m, &c->d_adam_v, &c->d_grads, &c->d_logX, &c->d_logT})
        b->reset();
    c->pool.release_all();
    for (auto& e : c->ev_pool) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return GS_OK;
}

}  // extern "C"
