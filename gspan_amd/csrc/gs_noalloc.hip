// gs_noalloc.hip -- gs_is_canonical_batch_noalloc: the canonicality check of many codes
// (is_canonical, gspan.cu:586-606) without any device or pinned-host allocation, for
// callers that cannot allocate (code under stream capture, tight serving loops, harnesses
// that forbid hipMalloc).
//
// Its device staging -- the caller's codes as given, their offsets, the packed codes the
// canonicality kernel reads, the flags and an error word -- lives in `scratch`, a
// file-scope device array of GS_NOALLOC_SCRATCH_BYTES that every GPU the library runs on
// holds until the process ends.  Host code never uses the array's name as an address: its
// device address is looked up once per device (noalloc_init, called by gs_create with the
// device current) and kept in g_scratch[device].  Calls on one GPU share that GPU's array
// and leave their contents in it; the error word is cleared by a kernel at the start of
// every call.  Two calls that overlap on one GPU (two host threads, or two contexts on
// different streams) race on the array.
#include <hip/hip_runtime.h>

#include <mutex>

#include "gs_internal.h"

namespace {

// Layout of the scratch array, in 32-bit words.
constexpr long long kErrWord = 0;
constexpr long long kOffs = 64;  // GS_NOALLOC_MAX_CODES + 1 code offsets (in edges)
constexpr long long kRaw = kOffs + ((GS_NOALLOC_MAX_CODES + 1 + 63) / 64) * 64;
constexpr long long kPacked = kRaw + 5LL * GS_MAXE * GS_NOALLOC_MAX_CODES;
constexpr int kPackedStride = 1 + 5 * GS_MAXE;  // k_canon_codes layout: m, then 5 per edge
constexpr long long kFlags = kPacked + (long long)kPackedStride * GS_NOALLOC_MAX_CODES;
constexpr long long kWords = kFlags + GS_NOALLOC_MAX_CODES;
static_assert(kWords * 4 <= GS_NOALLOC_SCRATCH_BYTES, "scratch layout exceeds the array");

__device__ __align__(256) int scratch[GS_NOALLOC_SCRATCH_BYTES / 4];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
int* g_scratch[kMaxDevices] = {};  // device address of `scratch` on each device

__global__ void k_clear_word(int* w) { *w = 0; }

// Dense rank of x among the first n values of v (distinct values below x).
__device__ int dense_rank(const int* v, int n, int x) {
    int r = 0;
    for (int i = 0; i < n; ++i) {
        bool first = v[i] < x;  // count each smaller value once
        for (int j = 0; j < i && first; ++j) first = v[j] != v[i];
        r += first ? 1 : 0;
    }
    return r;
}

// One thread per code: copies code i from the caller's layout (5 ints per edge, codes
// back to back, offs[i] = first edge of code i) into the canonicality kernel's layout with
// vertex and edge labels replaced by their dense ranks within the code.  The canonical walk
// only compares labels of one code with each other, so ranks per code give it the same
// answer as the original labels.  Labels below 0 are the reference's "no label" (-1).
__global__ void k_pack_codes(const int* __restrict__ raw, const int* __restrict__ offs,
                             int ncodes, int* __restrict__ packed) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncodes) return;
    const int m = offs[i + 1] - offs[i];
    const int* c = raw + 5LL * offs[i];
    int vl[2 * GS_MAXE], el[GS_MAXE];
    int nvl = 0;
    for (int j = 0; j < m; ++j) {
        if (c[5 * j + 2] >= 0) vl[nvl++] = c[5 * j + 2];
        if (c[5 * j + 4] >= 0) vl[nvl++] = c[5 * j + 4];
        el[j] = c[5 * j + 3];
    }
    int* p = packed + (long long)i * kPackedStride;
    p[0] = m;
    for (int j = 0; j < m; ++j) {
        const int* e = c + 5 * j;
        int* q = p + 1 + 5 * j;
        q[0] = e[0];
        q[1] = e[1];
        q[2] = e[2] >= 0 ? dense_rank(vl, nvl, e[2]) : -1;
        q[3] = dense_rank(el, m, e[3]);
        q[4] = e[4] >= 0 ? dense_rank(vl, nvl, e[4]) : -1;
    }
}

}  // namespace

namespace gsk {

hipError_t noalloc_init(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[device]) return hipSuccess;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[device] = static_cast<int*>(p);
    return hipSuccess;
}

}  // namespace gsk

extern "C" int gs_is_canonical_batch_noalloc(gs_ctx* c, const int* codes5, const int* lens,
                                             int ncodes, int* flags) {
    if (!c || ncodes < 0 || (ncodes > 0 && (!codes5 || !lens || !flags))) return GS_ERR_INVALID;
    if (ncodes > GS_NOALLOC_MAX_CODES) {
        gsk::ctx_set_error(c, "gs_is_canonical_batch_noalloc: more codes than its scratch "
                              "array holds (GS_NOALLOC_MAX_CODES)");
        return GS_ERR_LIMIT;
    }
    if (ncodes == 0) return GS_OK;
    // offsets of the codes (in edges) and the checks of gs_is_canonical_batch, on the stack
    int offs[GS_NOALLOC_MAX_CODES + 1];
    offs[0] = 0;
    int max_m = 0;
    for (int i = 0; i < ncodes; ++i) {
        if (lens[i] < 1 || lens[i] > GS_MAXE) {
            gsk::ctx_set_error(c, "code length out of range");
            return GS_ERR_LIMIT;
        }
        for (int j = 0; j < lens[i]; ++j) {
            const int* e = codes5 + 5LL * (offs[i] + j);
            if (e[0] < 0 || e[1] < 0 || e[0] >= GS_MAXV || e[1] >= GS_MAXV) {
                gsk::ctx_set_error(c, "bad pattern vertex ids");
                return GS_ERR_INVALID;
            }
        }
        offs[i + 1] = offs[i] + lens[i];
        if (lens[i] > max_m) max_m = lens[i];
    }
    const int device = gsk::ctx_device(c);
    if (hipSetDevice(device) != hipSuccess) {
        gsk::ctx_set_error(c, "hipSetDevice failed");
        return GS_ERR_HIP;
    }
    int* base = device >= 0 && device < kMaxDevices ? g_scratch[device] : nullptr;
    if (!base) {
        gsk::ctx_set_error(c, "scratch array of this device was not looked up");
        return GS_ERR_INTERNAL;
    }
    hipStream_t s = gsk::ctx_stream(c);
    int* err = base + kErrWord;
    int* d_offs = base + kOffs;
    int* d_raw = base + kRaw;
    int* d_packed = base + kPacked;
    int* d_flags = base + kFlags;
    int e_host = 0;
    hipError_t e = hipMemcpyAsync(d_raw, codes5, sizeof(int) * 5 * (size_t)offs[ncodes],
                                  hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_offs, offs, sizeof(int) * (size_t)(ncodes + 1),
                           hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_clear_word, dim3(1), dim3(1), 0, s, err);
        hipLaunchKernelGGL(k_pack_codes, dim3((ncodes + 255) / 256), dim3(256), 0, s, d_raw,
                           d_offs, ncodes, d_packed);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = gsk::canon_codes(d_packed, ncodes, max_m, d_flags, err, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(flags, d_flags, sizeof(int) * (size_t)ncodes, hipMemcpyDeviceToHost,
                           s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(&e_host, err, sizeof(int), hipMemcpyDeviceToHost, s);
    // the copies read the caller's arrays and offs, and write flags and e_host
    const hipError_t es = hipStreamSynchronize(s);
    if (e == hipSuccess) e = es;
    if (e != hipSuccess) {
        gsk::ctx_set_error(c, hipGetErrorString(e));
        return GS_ERR_HIP;
    }
    if (e_host & 1) {
        gsk::ctx_set_error(c, "canonicality check: more projections than the walk buffers hold");
        return GS_ERR_LIMIT;
    }
    if (e_host & 2) {
        gsk::ctx_set_error(c, "canonicality walk ran past the code");
        return GS_ERR_INTERNAL;
    }
    return GS_OK;
}
