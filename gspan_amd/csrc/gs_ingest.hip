// gs_ingest.hip -- Graph::from_txt (graph.cc:38-63) followed by CSRGraph(Graph&)
// (graph.cc:66-112) on the GPU: the text of a graph file in, the CSR arrays out.
//
// The reference reads whitespace-separated tokens with fscanf: an op token starting with
// 't' is followed by a string and an int, 'v' by vertex id and label, 'e' by two vertex
// ids and an edge label; other tokens are skipped.  Here every line is one record, parsed
// by one thread, which is the layout gSpan files use; a line whose tokens would make the
// reference read across a line boundary (wrong token count, a non-integer argument, or an
// op letter after an unknown first token) is rejected with its line number instead of
// being read differently.
//
// Graph semantics kept: a vertex exists once named by a 'v' or an 'e' record; its label is
// that of its last 'v' record (0 without one); vertices are renumbered in ascending id
// order (std::map order); every 'e' record adds both directions, and each vertex's
// entries are ordered by (to, label).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>

#include "gs_internal.h"
#include "gspan_amd.h"

namespace {

constexpr int IB = 256;                  // threads per block
constexpr int BPT = 16;                  // text bytes per thread
constexpr int CHUNK = IB * BPT;          // text bytes per block
constexpr int REC_SKIP = 0, REC_VERTEX = 1, REC_EDGE = 2, REC_BAD = 3;

__device__ __forceinline__ bool is_space(char ch) {
    return ch == ' ' || ch == '\t' || ch == '\n' || ch == '\r' || ch == '\v' || ch == '\f';
}

// block-wide exclusive sum of one int per thread
__device__ int block_excl(int v, int* total) {
    __shared__ int s_w[IB / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = v;
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == 63) s_w[w] = x;
    __syncthreads();
    int base = 0, tot = 0;
    for (int i = 0; i < IB / 64; ++i) {
        if (i < w) base += s_w[i];
        tot += s_w[i];
    }
    __syncthreads();
    *total = tot;
    return base + x - v;
}

// newlines of each CHUNK followed by another byte (each starts a line)
__global__ __launch_bounds__(IB) void k_nl_count(const char* __restrict__ t, long long len,
                                                 int* __restrict__ cnt) {
    const long long b0 = (long long)blockIdx.x * CHUNK + (long long)threadIdx.x * BPT;
    int c = 0;
    for (int j = 0; j < BPT; ++j) {
        const long long i = b0 + j;
        if (i + 1 < len && t[i] == '\n') ++c;
    }
    int tot;
    block_excl(c, &tot);
    if (threadIdx.x == 0) cnt[blockIdx.x] = tot;
}

// starts[0] = 0; starts[1 + k] = position after the k-th such newline
__global__ __launch_bounds__(IB) void k_nl_write(const char* __restrict__ t, long long len,
                                                 const long long* __restrict__ base,
                                                 long long* __restrict__ starts) {
    const long long b0 = (long long)blockIdx.x * CHUNK + (long long)threadIdx.x * BPT;
    int c = 0;
    for (int j = 0; j < BPT; ++j) {
        const long long i = b0 + j;
        if (i + 1 < len && t[i] == '\n') ++c;
    }
    int tot;
    long long at = base[blockIdx.x] + block_excl(c, &tot);
    for (int j = 0; j < BPT; ++j) {
        const long long i = b0 + j;
        if (i + 1 < len && t[i] == '\n') starts[1 + at++] = i + 1;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) starts[0] = 0;
}

// fscanf("%d") on a whole token: optional sign, digits, within int32
__device__ bool parse_int(const char* t, long long b, long long e, int* out) {
    long long i = b;
    bool neg = false;
    if (i < e && (t[i] == '+' || t[i] == '-')) {
        neg = t[i] == '-';
        ++i;
    }
    if (i >= e) return false;
    long long v = 0;
    for (; i < e; ++i) {
        const char ch = t[i];
        if (ch < '0' || ch > '9') return false;
        v = v * 10 + (ch - '0');
        if (v > 2147483648LL) return false;
    }
    if (neg) v = -v;
    if (v > 2147483647LL || v < -2147483648LL) return false;
    *out = (int)v;
    return true;
}

__global__ __launch_bounds__(IB) void k_parse_lines(const char* __restrict__ t, long long len,
                                                    const long long* __restrict__ starts,
                                                    long long L, int* __restrict__ kind,
                                                    int* __restrict__ A, int* __restrict__ B,
                                                    int* __restrict__ C, int* __restrict__ vflag,
                                                    int* __restrict__ eflag) {
    const long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= L) return;
    const long long s = starts[l];
    const long long e = l + 1 < L ? starts[l + 1] : len;
    // tokens of the line: at most 1 op + 3 arguments are kept; more tokens mark the line
    long long tb[5], te[5];
    int nt = 0;
    bool later_op = false;  // a token after the first starts with t, v or e
    long long i = s;
    while (i < e) {
        while (i < e && is_space(t[i])) ++i;
        if (i >= e) break;
        const long long b = i;
        while (i < e && !is_space(t[i])) ++i;
        if (nt < 5) {
            tb[nt] = b;
            te[nt] = i;
        }
        if (nt > 0 && (t[b] == 't' || t[b] == 'v' || t[b] == 'e')) later_op = true;
        ++nt;
    }
    int k = REC_SKIP, a = 0, bb = 0, c = 0;
    if (nt > 0) {
        const char op = t[tb[0]];
        if (op == 't') {  // "%s%d"
            int dummy;
            k = (nt == 3 && parse_int(t, tb[2], te[2], &dummy)) ? REC_SKIP : REC_BAD;
        } else if (op == 'v') {  // "%d%d"
            k = (nt == 3 && parse_int(t, tb[1], te[1], &a) && parse_int(t, tb[2], te[2], &bb))
                    ? REC_VERTEX
                    : REC_BAD;
        } else if (op == 'e') {  // "%d%d%d"
            k = (nt == 4 && parse_int(t, tb[1], te[1], &a) && parse_int(t, tb[2], te[2], &bb) &&
                 parse_int(t, tb[3], te[3], &c))
                    ? REC_EDGE
                    : REC_BAD;
        } else {
            k = later_op ? REC_BAD : REC_SKIP;  // every token is an ignored op
        }
    }
    kind[l] = k;
    A[l] = a;
    B[l] = bb;
    C[l] = c;
    vflag[l] = k == REC_VERTEX ? 1 : 0;
    eflag[l] = k == REC_EDGE ? 1 : 0;
}

__global__ void k_first_bad(const int* __restrict__ kind, long long L,
                            unsigned long long* __restrict__ first) {
    const long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l < L && kind[l] == REC_BAD) atomicMin(first, (unsigned long long)l);
}

// vertex ids named by the records: V vertex records first, then 2 per edge record
__global__ void k_collect_ids(const int* __restrict__ kind, const int* __restrict__ A,
                              const int* __restrict__ B, const long long* __restrict__ vpos,
                              const long long* __restrict__ epos, long long L, long long V,
                              int* __restrict__ ids) {
    const long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= L) return;
    if (kind[l] == REC_VERTEX) ids[vpos[l]] = A[l];
    else if (kind[l] == REC_EDGE) {
        ids[V + 2 * epos[l]] = A[l];
        ids[V + 2 * epos[l] + 1] = B[l];
    }
}

__global__ void k_uniq_flags(const int* __restrict__ s, long long N, int* __restrict__ f) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) f[i] = (i == 0 || s[i] != s[i - 1]) ? 1 : 0;
}

__global__ void k_uniq_scatter(const int* __restrict__ s, const int* __restrict__ f,
                               const long long* __restrict__ pos, long long N,
                               int* __restrict__ u) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N && f[i]) u[pos[i]] = s[i];
}

__device__ __forceinline__ int rank_of(const int* __restrict__ u, int n, int id) {
    int lo = 0, hi = n;  // first u[x] >= id; id is present
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (u[mid] < id) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// last 'v' record per vertex: max over (line + 1, label) packed in 64 bits
__global__ void k_vertex_labels(const int* __restrict__ kind, const int* __restrict__ A,
                                const int* __restrict__ B, long long L,
                                const int* __restrict__ u, int n,
                                unsigned long long* __restrict__ packed) {
    const long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= L || kind[l] != REC_VERTEX) return;
    const int v = rank_of(u, n, A[l]);
    const unsigned long long p = ((unsigned long long)(l + 1) << 32) |
                                 (unsigned long long)((unsigned)B[l] ^ 0x80000000u);
    atomicMax(&packed[v], p);
}

__global__ void k_vlabels_out(const unsigned long long* __restrict__ packed, int n,
                              int* __restrict__ vlabels) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const unsigned long long p = packed[v];
    vlabels[v] = p ? (int)((unsigned)(p & 0xffffffffu) ^ 0x80000000u) : 0;
}

// both directions of every edge record: entry 2q = (u, v), 2q + 1 = (v, u)
__global__ void k_edge_entries(const int* __restrict__ kind, const int* __restrict__ A,
                               const int* __restrict__ B, const int* __restrict__ C,
                               const long long* __restrict__ epos, long long L,
                               const int* __restrict__ u, int n, int* __restrict__ from,
                               int* __restrict__ to, int* __restrict__ lab,
                               int* __restrict__ idx) {
    const long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= L || kind[l] != REC_EDGE) return;
    const long long q = epos[l];
    const int a = rank_of(u, n, A[l]), b = rank_of(u, n, B[l]);
    from[2 * q] = a;
    to[2 * q] = b;
    from[2 * q + 1] = b;
    to[2 * q + 1] = a;
    lab[2 * q] = lab[2 * q + 1] = C[l];
    idx[2 * q] = (int)(2 * q);
    idx[2 * q + 1] = (int)(2 * q + 1);
}

__global__ void k_pair_keys(const int* __restrict__ from, const int* __restrict__ to,
                            const int* __restrict__ perm, long long N,
                            unsigned long long* __restrict__ key) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int e = perm[i];
    key[i] = ((unsigned long long)(unsigned)from[e] << 32) | (unsigned)to[e];
}

__global__ void k_edges_out(const int* __restrict__ from, const int* __restrict__ to,
                            const int* __restrict__ lab, const int* __restrict__ perm,
                            long long N, int* __restrict__ edges3, int* __restrict__ sfrom) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int e = perm[i];
    edges3[3 * i + 0] = from[e];
    edges3[3 * i + 1] = to[e];
    edges3[3 * i + 2] = lab[e];
    sfrom[i] = from[e];
}

// offsets[v] = first sorted entry with from >= v
__global__ void k_offsets(const int* __restrict__ sfrom, long long N, int n,
                          int* __restrict__ offsets) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v > n) return;
    long long lo = 0, hi = N;
    while (lo < hi) {
        const long long mid = (lo + hi) >> 1;
        if (sfrom[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    offsets[v] = (int)lo;
}

unsigned blocks(long long n, int b) { return (unsigned)std::max(1LL, (n + b - 1) / b); }

struct HipErr {
    hipError_t e;
    const char* what;
};

#define ICHECK(x)                                  \
    do {                                           \
        hipError_t e_ = (x);                       \
        if (e_ != hipSuccess) throw HipErr{e_, #x}; \
    } while (0)

// device buffer owned by the parse
struct Dev {
    void* p = nullptr;
    Dev() = default;
    explicit Dev(size_t bytes) { ICHECK(hipMalloc(&p, std::max<size_t>(bytes, 16))); }
    Dev(const Dev&) = delete;
    Dev& operator=(const Dev&) = delete;
    Dev(Dev&& o) noexcept : p(o.p) { o.p = nullptr; }
    Dev& operator=(Dev&& o) noexcept {
        if (this != &o) {
            if (p) (void)hipFree(p);
            p = o.p;
            o.p = nullptr;
        }
        return *this;
    }
    ~Dev() {
        if (p) (void)hipFree(p);
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

}  // namespace

struct gs_text_graph {
    int device = 0;
    int n = 0, nnz = 0;
    Dev offsets, vlabels, edges3;
};

int gs_text_parse(const char* text, long long len, int device, gs_text_graph** out, int* n_out,
                  int* nnz_out, long long* bad_line) {
    if (!out || !n_out || !nnz_out || len < 0 || (len > 0 && !text)) return GS_ERR_INVALID;
    *out = nullptr;
    if (bad_line) *bad_line = -1;
    hipStream_t s = nullptr;
    try {
        ICHECK(hipSetDevice(device));
        ICHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        std::unique_ptr<gs_text_graph> g(new gs_text_graph());
        g->device = device;
        // text and line starts
        Dev d_text(len);
        if (len) ICHECK(hipMemcpyAsync(d_text.p, text, len, hipMemcpyHostToDevice, s));
        const long long nb = (len + CHUNK - 1) / CHUNK;
        long long L = 0;
        Dev d_starts;
        if (len > 0) {
            Dev d_cnt(sizeof(int) * nb), d_base(sizeof(long long) * (nb + 1));
            const long long te = gsk::scan_tmp_elems(nb);
            Dev d_tmp(sizeof(long long) * te);
            hipLaunchKernelGGL(k_nl_count, dim3((unsigned)nb), dim3(IB), 0, s,
                               d_text.as<const char>(), len, d_cnt.as<int>());
            ICHECK(gsk::scan_exclusive(d_cnt.as<int>(), d_base.as<long long>(), nb,
                                       d_tmp.as<long long>(), te, s));
            long long nl = 0;
            ICHECK(hipMemcpyAsync(&nl, d_base.as<long long>() + nb, sizeof(long long),
                                  hipMemcpyDeviceToHost, s));
            ICHECK(hipStreamSynchronize(s));
            L = nl + 1;
            if (L >= (1LL << 31)) {  // one thread per line in launches of < 2^32 work-items
                (void)hipStreamDestroy(s);
                return GS_ERR_LIMIT;
            }
            d_starts = Dev(sizeof(long long) * L);
            hipLaunchKernelGGL(k_nl_write, dim3((unsigned)nb), dim3(IB), 0, s,
                               d_text.as<const char>(), len, d_base.as<long long>(),
                               d_starts.as<long long>());
        }
        // records
        Dev d_kind(sizeof(int) * L), d_A(sizeof(int) * L), d_B(sizeof(int) * L),
            d_C(sizeof(int) * L), d_vf(sizeof(int) * L), d_ef(sizeof(int) * L);
        Dev d_vpos(sizeof(long long) * (L + 1)), d_epos(sizeof(long long) * (L + 1));
        long long V = 0, E = 0;
        if (L > 0) {
            hipLaunchKernelGGL(k_parse_lines, dim3(blocks(L, IB)), dim3(IB), 0, s,
                               d_text.as<const char>(), len, d_starts.as<long long>(), L,
                               d_kind.as<int>(), d_A.as<int>(), d_B.as<int>(), d_C.as<int>(),
                               d_vf.as<int>(), d_ef.as<int>());
            Dev d_bad(sizeof(unsigned long long));
            const unsigned long long none = ~0ull;
            ICHECK(hipMemcpyAsync(d_bad.p, &none, sizeof(none), hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(k_first_bad, dim3(blocks(L, IB)), dim3(IB), 0, s,
                               d_kind.as<int>(), L, d_bad.as<unsigned long long>());
            const long long te = gsk::scan_tmp_elems(L);
            Dev d_tmp(sizeof(long long) * te);
            ICHECK(gsk::scan_exclusive(d_vf.as<int>(), d_vpos.as<long long>(), L,
                                       d_tmp.as<long long>(), te, s));
            ICHECK(gsk::scan_exclusive(d_ef.as<int>(), d_epos.as<long long>(), L,
                                       d_tmp.as<long long>(), te, s));
            unsigned long long bad = none;
            ICHECK(hipMemcpyAsync(&bad, d_bad.p, sizeof(bad), hipMemcpyDeviceToHost, s));
            ICHECK(hipMemcpyAsync(&V, d_vpos.as<long long>() + L, sizeof(long long),
                                  hipMemcpyDeviceToHost, s));
            ICHECK(hipMemcpyAsync(&E, d_epos.as<long long>() + L, sizeof(long long),
                                  hipMemcpyDeviceToHost, s));
            ICHECK(hipStreamSynchronize(s));
            if (bad != none) {
                if (bad_line) *bad_line = (long long)bad;
                (void)hipStreamDestroy(s);
                return GS_ERR_INVALID;
            }
        }
        if (2 * E > 2147483647LL || V + 2 * E > 2147483647LL) {
            (void)hipStreamDestroy(s);
            return GS_ERR_LIMIT;
        }
        const long long NI = V + 2 * E, N = 2 * E;
        // vertex ids: sort, unique
        int n = 0;
        Dev d_u;
        if (NI > 0) {
            Dev d_ids(sizeof(int) * NI), d_sorted(sizeof(int) * NI);
            hipLaunchKernelGGL(k_collect_ids, dim3(blocks(L, IB)), dim3(IB), 0, s,
                               d_kind.as<int>(), d_A.as<int>(), d_B.as<int>(),
                               d_vpos.as<long long>(), d_epos.as<long long>(), L, V,
                               d_ids.as<int>());
            size_t tb = 0;
            ICHECK(rocprim::radix_sort_keys(nullptr, tb, d_ids.as<int>(), d_sorted.as<int>(),
                                            (size_t)NI, 0, 32, s));
            Dev d_rtmp(tb);
            ICHECK(rocprim::radix_sort_keys(d_rtmp.p, tb, d_ids.as<int>(), d_sorted.as<int>(),
                                            (size_t)NI, 0, 32, s));
            Dev d_f(sizeof(int) * NI), d_pos(sizeof(long long) * (NI + 1));
            const long long te = gsk::scan_tmp_elems(NI);
            Dev d_tmp(sizeof(long long) * te);
            hipLaunchKernelGGL(k_uniq_flags, dim3(blocks(NI, IB)), dim3(IB), 0, s,
                               d_sorted.as<int>(), NI, d_f.as<int>());
            ICHECK(gsk::scan_exclusive(d_f.as<int>(), d_pos.as<long long>(), NI,
                                       d_tmp.as<long long>(), te, s));
            long long nu = 0;
            ICHECK(hipMemcpyAsync(&nu, d_pos.as<long long>() + NI, sizeof(long long),
                                  hipMemcpyDeviceToHost, s));
            ICHECK(hipStreamSynchronize(s));
            n = (int)nu;
            d_u = Dev(sizeof(int) * n);
            hipLaunchKernelGGL(k_uniq_scatter, dim3(blocks(NI, IB)), dim3(IB), 0, s,
                               d_sorted.as<int>(), d_f.as<int>(), d_pos.as<long long>(), NI,
                               d_u.as<int>());
        }
        // labels
        g->vlabels = Dev(sizeof(int) * std::max(n, 1));
        if (n > 0) {
            Dev d_packed(sizeof(unsigned long long) * n);
            ICHECK(hipMemsetAsync(d_packed.p, 0, sizeof(unsigned long long) * n, s));
            hipLaunchKernelGGL(k_vertex_labels, dim3(blocks(L, IB)), dim3(IB), 0, s,
                               d_kind.as<int>(), d_A.as<int>(), d_B.as<int>(), L,
                               d_u.as<int>(), n, d_packed.as<unsigned long long>());
            hipLaunchKernelGGL(k_vlabels_out, dim3(blocks(n, IB)), dim3(IB), 0, s,
                               d_packed.as<unsigned long long>(), n, g->vlabels.as<int>());
            ICHECK(hipStreamSynchronize(s));
        }
        // entries, ordered by (from, to, label): stable radix sorts, label first
        g->edges3 = Dev(sizeof(int) * 3 * std::max(N, 1LL));
        g->offsets = Dev(sizeof(int) * (n + 1));
        Dev d_sfrom(sizeof(int) * std::max(N, 1LL));
        if (N > 0) {
            Dev d_from(sizeof(int) * N), d_to(sizeof(int) * N), d_lab(sizeof(int) * N),
                d_idx(sizeof(int) * N);
            hipLaunchKernelGGL(k_edge_entries, dim3(blocks(L, IB)), dim3(IB), 0, s,
                               d_kind.as<int>(), d_A.as<int>(), d_B.as<int>(), d_C.as<int>(),
                               d_epos.as<long long>(), L, d_u.as<int>(), n, d_from.as<int>(),
                               d_to.as<int>(), d_lab.as<int>(), d_idx.as<int>());
            Dev d_lab2(sizeof(int) * N), d_perm1(sizeof(int) * N);
            size_t tb1 = 0;
            ICHECK(rocprim::radix_sort_pairs(nullptr, tb1, d_lab.as<int>(), d_lab2.as<int>(),
                                             d_idx.as<int>(), d_perm1.as<int>(), (size_t)N, 0,
                                             32, s));
            Dev d_key(sizeof(unsigned long long) * N), d_key2(sizeof(unsigned long long) * N),
                d_perm2(sizeof(int) * N);
            size_t tb2 = 0;
            ICHECK(rocprim::radix_sort_pairs(nullptr, tb2, d_key.as<unsigned long long>(),
                                             d_key2.as<unsigned long long>(), d_perm1.as<int>(),
                                             d_perm2.as<int>(), (size_t)N, 0, 64, s));
            Dev d_rtmp(std::max(tb1, tb2));
            ICHECK(rocprim::radix_sort_pairs(d_rtmp.p, tb1, d_lab.as<int>(), d_lab2.as<int>(),
                                             d_idx.as<int>(), d_perm1.as<int>(), (size_t)N, 0,
                                             32, s));
            hipLaunchKernelGGL(k_pair_keys, dim3(blocks(N, IB)), dim3(IB), 0, s,
                               d_from.as<int>(), d_to.as<int>(), d_perm1.as<int>(), N,
                               d_key.as<unsigned long long>());
            ICHECK(rocprim::radix_sort_pairs(d_rtmp.p, tb2, d_key.as<unsigned long long>(),
                                             d_key2.as<unsigned long long>(), d_perm1.as<int>(),
                                             d_perm2.as<int>(), (size_t)N, 0, 64, s));
            hipLaunchKernelGGL(k_edges_out, dim3(blocks(N, IB)), dim3(IB), 0, s,
                               d_from.as<int>(), d_to.as<int>(), d_lab.as<int>(),
                               d_perm2.as<int>(), N, g->edges3.as<int>(), d_sfrom.as<int>());
            ICHECK(hipStreamSynchronize(s));
        }
        hipLaunchKernelGGL(k_offsets, dim3(blocks((long long)n + 1, IB)), dim3(IB), 0, s,
                           d_sfrom.as<int>(), N, n, g->offsets.as<int>());
        ICHECK(hipGetLastError());
        ICHECK(hipStreamSynchronize(s));
        ICHECK(hipStreamDestroy(s));
        s = nullptr;
        g->n = n;
        g->nnz = (int)N;
        *n_out = n;
        *nnz_out = (int)N;
        *out = g.release();
        return GS_OK;
    } catch (const HipErr&) {
        if (s) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
        return GS_ERR_HIP;
    } catch (const std::bad_alloc&) {
        if (s) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
        return GS_ERR_INTERNAL;
    }
}

int gs_text_copy(const gs_text_graph* g, int* offsets, int* vlabels, int* edges3) {
    if (!g || !offsets || (g->n > 0 && !vlabels) || (g->nnz > 0 && !edges3))
        return GS_ERR_INVALID;
    if (hipSetDevice(g->device) != hipSuccess) return GS_ERR_HIP;
    if (hipMemcpy(offsets, g->offsets.p, sizeof(int) * (g->n + 1), hipMemcpyDeviceToHost) !=
        hipSuccess)
        return GS_ERR_HIP;
    if (g->n > 0 &&
        hipMemcpy(vlabels, g->vlabels.p, sizeof(int) * g->n, hipMemcpyDeviceToHost) != hipSuccess)
        return GS_ERR_HIP;
    if (g->nnz > 0 && hipMemcpy(edges3, g->edges3.p, sizeof(int) * 3 * (size_t)g->nnz,
                                hipMemcpyDeviceToHost) != hipSuccess)
        return GS_ERR_HIP;
    return GS_OK;
}

void gs_text_free(gs_text_graph* g) {
    if (!g) return;
    (void)hipSetDevice(g->device);
    delete g;
}
