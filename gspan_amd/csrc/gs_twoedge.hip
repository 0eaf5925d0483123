// Last level of two-edge children (children of single-edge parents) from per-vertex
// neighbour runs, on simple graphs.
//
// A child of the single-edge parent p = (0, 1, a, e0, b) at forward slot (j, e, l) is the
// path o - s - x: s = seg_v[j] is the segment vertex (label ls), o the parent's other vertex
// (label lo, joined to s by e0), x the new vertex (label l, joined to s by e).  The
// reference enumerates it from every parent row (w, u) (u the image of s, w of o) and every
// neighbour of u outside the row (gspan.cu:447-483, extend_omp gspan.cu:674-780) and takes
// the MNI support from the child's embeddings (gspan.cu:165-190 / 321-374).  With
// R(u, l, e) = the neighbours of u of label l over edge label e (a "run"), and no
// self-loops or repeated neighbours, the same numbers follow from runs alone:
//   child rows       sum over u of label ls:  |R(u, lo, e0)| * (|R(u, l, e)| - [(l, e) = (lo, e0)])
//   images of s      the u of label ls for which both factors are non-zero
//   images of o      the w of label lo with a neighbour u in R(w, ls, e0) whose run
//                    R(u, l, e) holds a vertex other than w
//   images of x      the x of label l with a neighbour u in R(x, ls, e) whose run
//                    R(u, lo, e0) holds a vertex other than x
// Runs are named by bit b = label * EL + edge label (NB = VL * EL bits).  Per vertex u,
// M1(u) marks its runs and M2(u) its runs of two or more entries.  The (vertex, run)
// pairs of the graph, ordered by class c = (vertex label, run bit), fill three tables of NB
// counts per class:
//   CNT[c][b']  sum over the class's pairs (u, b): |run b| * (|run b'| - [b' = b])
//   CEN[c][b']  the class's pairs (u, b) with |run b'| - [b' = b] >= 1
//   END[c][b']  the class's pairs (z, b) for which some vertex u of run b has run b' beyond
//               z: bit b' of the OR over u of M1(u), with M2(u) at b' = (label z, edge
//               label of b), the run of u that holds z
// and a key of the batch reads its child rows and its three image counts from them.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

#include "gs_internal.h"

namespace {

__device__ __forceinline__ int run_bit(int rkey, int EL) {  // rkey = (label << 16) | edge label
    return (rkey >> 16) * EL + (rkey & 0xffff);
}

// M1 / M2 words of every vertex: one thread per vertex over its runs (ascending bits, so
// each word is complete before the next starts); the arrays are zero beforehand.
__global__ void k_te_masks(DevGraph g, int EL, int W, uint32_t* __restrict__ m1,
                           uint32_t* __restrict__ m2) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < g.n; v += stride) {
        const int o0 = g.offsets[v], o1 = g.offsets[v + 1];
        if (o0 == o1) continue;
        const int r0 = g.runid[o0], r1 = g.runid[o1 - 1];
        int cw = -1;
        uint32_t a1 = 0, a2 = 0;
        for (int k = r0; k <= r1; ++k) {
            const int b = run_bit(g.rkey[k], EL);
            const int w = b >> 5;
            if (w != cw) {
                if (cw >= 0) {
                    m1[v * W + cw] = a1;
                    m2[v * W + cw] = a2;
                }
                cw = w;
                a1 = a2 = 0;
            }
            a1 |= 1u << (b & 31);
            if (g.rstart[k + 1] - g.rstart[k] >= 2) a2 |= 1u << (b & 31);
        }
        m1[v * W + cw] = a1;
        m2[v * W + cw] = a2;
    }
}

// pairs and CSR entries per class (hist, clen), then the pairs at their class positions
// (cursor = class starts, advanced); one thread per vertex over its runs
template <bool SCATTER>
__global__ void k_te_pairs(DevGraph g, int EL, int NB, int* __restrict__ hist_or_cursor,
                           unsigned long long* __restrict__ clen,
                           const long long* __restrict__ cstart, int2* __restrict__ pairs) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < g.n; v += stride) {
        const int o0 = g.offsets[v], o1 = g.offsets[v + 1];
        if (o0 == o1) continue;
        const int r0 = g.runid[o0], r1 = g.runid[o1 - 1];
        const long long cbase = (long long)g.vlab[v] * NB;
        for (int k = r0; k <= r1; ++k) {
            const long long c = cbase + run_bit(g.rkey[k], EL);
            if (!SCATTER) {
                atomicAdd(&hist_or_cursor[c], 1);
                atomicAdd(&clen[c], (unsigned long long)(g.rstart[k + 1] - g.rstart[k]));
            } else {
                const long long at = cstart[c] + atomicAdd(&hist_or_cursor[c], 1);
                pairs[at] = make_int2((int)v, k);
            }
        }
    }
}

// The three tables from the class-ordered pairs.  One wave per workgroup takes chunks of
// TE_CHUNK consecutive pairs from a queue and sums them in LDS per class (NB counters of
// each table), adding the non-zero sums to the global tables when the class changes and
// at the end.  p_of[(a * EL + e0) * VL + b]: the batch index of the single-edge parent
// (0, 1, a, e0, b), a <= b, or -1; CNT / CEN only for pairs that are parent runs of one.
// Per pair, the OR over its run's vertices is taken by G = 64 / WP lane groups of WP >= W
// lanes (lane w of a group holds word w).  work[0] += runs visited for CNT / CEN, work[1]
// += mask words read for END.
constexpr int TE_CHUNK = 32;

__global__ __launch_bounds__(64) void k_te_accum(DevGraph g, int EL, int VL, int NB, int W,
                                                 int WP, const uint32_t* __restrict__ m1,
                                                 const uint32_t* __restrict__ m2,
                                                 const int2* __restrict__ pairs, long long npairs,
                                                 const int* __restrict__ p_of, int* queue,
                                                 unsigned long long* __restrict__ CNT,
                                                 int* __restrict__ CEN, int* __restrict__ END,
                                                 unsigned long long* work) {
    extern __shared__ unsigned long long te_smem[];
    unsigned long long* s_cnt = te_smem;
    int* s_cen = reinterpret_cast<int*>(s_cnt + NB);
    int* s_end = s_cen + NB;
    const int lane = threadIdx.x;
    for (int b = lane; b < NB; b += 64) {
        s_cnt[b] = 0;
        s_cen[b] = 0;
        s_end[b] = 0;
    }
    __syncthreads();
    const int G = 64 / WP, grp = lane / WP, wi = lane % WP;
    const long long nchunks = (npairs + TE_CHUNK - 1) / TE_CHUNK;
    long long cls = -1;
    unsigned long long w_runs = 0, w_words = 0;
    auto flush = [&]() {
        __syncthreads();
        const long long row = cls * NB;
        for (int b = lane; b < NB; b += 64) {
            if (s_cnt[b]) atomicAdd(&CNT[row + b], s_cnt[b]);
            if (s_cen[b]) atomicAdd(&CEN[row + b], s_cen[b]);
            if (s_end[b]) atomicAdd(&END[row + b], s_end[b]);
            s_cnt[b] = 0;
            s_cen[b] = 0;
            s_end[b] = 0;
        }
        __syncthreads();
    };
    for (;;) {
        long long ch = 0;
        if (lane == 0) ch = atomicAdd(queue, 1);
        ch = __shfl(ch, 0);
        if (ch >= nchunks) break;
        const long long pe = min(npairs, (ch + 1) * TE_CHUNK);
        for (long long pi = ch * TE_CHUNK; pi < pe; ++pi) {
            const int2 pr = pairs[pi];
            const int x = pr.x, k = pr.y;
            const int lx = g.vlab[x];
            const int rk = g.rkey[k];
            const int tl = rk >> 16, el = rk & 0xffff;
            const int b = tl * EL + el;
            const long long c = (long long)lx * NB + b;
            if (c != cls) {
                if (cls >= 0) flush();
                cls = c;
            }
            const int st = g.rstart[k], len = g.rstart[k + 1] - st;
            // x as the segment vertex of a parent row (x, w), w in run b
            const int lo = lx < tl ? lx : tl, hi = lx < tl ? tl : lx;
            if (p_of[(lo * EL + el) * VL + hi] >= 0) {
                const int o0 = g.offsets[x];
                const int r0 = g.runid[o0], r1 = g.runid[g.offsets[x + 1] - 1];
                for (int q = r0 + lane; q <= r1; q += 64) {
                    const int wq = g.rstart[q + 1] - g.rstart[q] - (q == k ? 1 : 0);
                    if (wq >= 1) {
                        const int bq = run_bit(g.rkey[q], EL);
                        atomicAdd(&s_cnt[bq], (unsigned long long)len * (unsigned long long)wq);
                        atomicAdd(&s_cen[bq], 1);
                    }
                }
                if (lane == 0) w_runs += (unsigned long long)(r1 - r0 + 1);
            }
            // x as an end vertex z: the runs of its neighbours u in run b, beyond z
            const int bs = lx * EL + el;  // the run of u that holds x
            uint32_t acc = 0;
            bool two = false;
            for (int i = grp; i < len; i += G) {
                const int u = (int)g.adjl[st + i].x;
                if (wi < W) acc |= m1[(long long)u * W + wi];
                if (wi == 0) two |= ((m2[(long long)u * W + (bs >> 5)] >> (bs & 31)) & 1u) != 0;
            }
            for (int off = WP; off < 64; off <<= 1) acc |= __shfl_xor(acc, off);
            two = __ballot(two) != 0;
            if (wi == (bs >> 5))
                acc = (acc & ~(1u << (bs & 31))) | (two ? 1u << (bs & 31) : 0u);
            if (lane == 0) w_words += (unsigned long long)len * (unsigned long long)(W + 1);
            for (int bb = lane; bb < NB; bb += 64) {
                const uint32_t word = (uint32_t)__shfl((int)acc, bb >> 5);
                if ((word >> (bb & 31)) & 1u) atomicAdd(&s_end[bb], 1);
            }
        }
    }
    if (cls >= 0) flush();
    if (lane == 0) {
        atomicAdd(&work[0], w_runs);
        atomicAdd(&work[1], w_words);
    }
}

// Every key of the batch (parents of one edge, so no backward slots): its child rows
// (key_count; past INT_MAX: INT_MIN, as the extension kernels' saturated counts) and the
// distinct images of its three vertices (img[8 * key + v]).  Slots the rules reject
// (gspan.cu:455, 475-477) have no children.
__global__ void k_te_keys(const PatDev* __restrict__ pats, int npats, KeySpace ks, long long K,
                          int NB, const unsigned long long* __restrict__ CNT,
                          const int* __restrict__ CEN, const int* __restrict__ END,
                          int* __restrict__ key_count, int* __restrict__ img) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long key = (long long)blockIdx.x * blockDim.x + threadIdx.x; key < K;
         key += stride) {
        int lo = 0, hi = npats - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (pats[mid].key_base <= key) lo = mid;
            else hi = mid - 1;
        }
        const PatDev& P = pats[lo];
        const int slot = (int)(key - P.key_base);
        const int j = slot / (ks.EL * ks.VL), e = (slot / ks.VL) % ks.EL, l = slot % ks.VL;
        int* out = img + 8 * key;
        for (int v = 0; v < 8; ++v) out[v] = 0;
        key_count[key] = 0;
        if (l < P.min_label || !(P.seg_el[j] < e || (P.seg_el[j] == e && P.seg_tl[j] <= l)))
            continue;
        const int sv = P.seg_v[j], ov = 1 - sv;
        const int ls = P.vlab[sv], lo_ = P.vlab[ov], e0 = P.cel[0];
        const long long A = (long long)lo_ * ks.EL + e0, C = (long long)l * ks.EL + e;
        const long long cs = ((long long)ls * NB + A) * NB + C;
        const unsigned long long cnt = CNT[cs];
        key_count[key] = cnt > (unsigned long long)INT_MAX ? INT_MIN : (int)cnt;
        out[sv] = CEN[cs];
        out[ov] = END[((long long)lo_ * NB + ((long long)ls * ks.EL + e0)) * NB + C];
        out[2] = END[((long long)l * NB + ((long long)ls * ks.EL + e)) * NB + A];
    }
}

// counts[8 a + v] = img[8 * akeys[a] + v] for the active keys (n_active null: max_active)
__global__ void k_te_gather(const int* __restrict__ img, const int* __restrict__ akeys,
                            const int* __restrict__ n_active, int max_active,
                            int* __restrict__ counts) {
    const long long na = n_active ? *n_active : max_active;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < 8 * na; i += stride)
        counts[i] = img[8LL * akeys[i / 8] + i % 8];
}

inline int te_grid(long long n, int block, int cap) {
    long long g = (n + block - 1) / block;
    return (int)(g < 1 ? 1 : (g > cap ? cap : g));
}

}  // namespace

namespace gsk {

static_assert(GS_MAXV == 8, "image counts are 8 per key");

hipError_t te_prepare(const DevGraph& g, int EL, int NB, int W, uint32_t* m1, uint32_t* m2,
                      int* hist, long long* cstart, unsigned long long* clen,
                      long long* scan_tmp, long long scan_elems, int2* pairs, hipStream_t s) {
    const long long NC = (long long)(NB / EL) * NB;  // VL * NB classes
    hipError_t e;
    if ((e = fill_u32(m1, 0u, (long long)g.n * W, s)) != hipSuccess) return e;
    if ((e = fill_u32(m2, 0u, (long long)g.n * W, s)) != hipSuccess) return e;
    if ((e = fill_i32(hist, 0, NC, s)) != hipSuccess) return e;
    if ((e = fill_u32(reinterpret_cast<uint32_t*>(clen), 0u, 2 * NC, s)) != hipSuccess) return e;
    const int grid = te_grid(g.n, 256, 8192);
    hipLaunchKernelGGL(k_te_masks, dim3(grid), dim3(256), 0, s, g, EL, W, m1, m2);
    hipLaunchKernelGGL(k_te_pairs<false>, dim3(grid), dim3(256), 0, s, g, EL, NB, hist, clen,
                       (const long long*)nullptr, (int2*)nullptr);
    if ((e = scan_exclusive(hist, cstart, NC, scan_tmp, scan_elems, s)) != hipSuccess) return e;
    if ((e = fill_i32(hist, 0, NC, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_te_pairs<true>, dim3(grid), dim3(256), 0, s, g, EL, NB, hist,
                       (unsigned long long*)nullptr, cstart, pairs);
    return hipGetLastError();
}

hipError_t te_level(const DevGraph& g, KeySpace ks, const PatDev* pats, int npats, long long K,
                    const int* p_of, const uint32_t* m1, const uint32_t* m2, const int2* pairs,
                    int* queue, unsigned long long* CNT, int* CEN, int* END, int* key_count,
                    int* img, unsigned long long* work, hipStream_t s) {
    const int NB = ks.VL * ks.EL, W = (NB + 31) / 32;
    int WP = 1;
    while (WP < W) WP <<= 1;
    const long long rows = (long long)ks.VL * NB * NB;
    hipError_t e;
    if ((e = fill_u32(reinterpret_cast<uint32_t*>(CNT), 0u, 2 * rows, s)) != hipSuccess) return e;
    if ((e = fill_i32(CEN, 0, rows, s)) != hipSuccess) return e;
    if ((e = fill_i32(END, 0, rows, s)) != hipSuccess) return e;
    if ((e = fill_i32(queue, 0, 1, s)) != hipSuccess) return e;
    const long long np = g.nruns;
    const size_t lds = (size_t)NB * 16;
    const long long nchunks = (np + TE_CHUNK - 1) / TE_CHUNK;
    const int grid = (int)(nchunks < 8192 ? (nchunks < 1 ? 1 : nchunks) : 8192);
    hipLaunchKernelGGL(k_te_accum, dim3(grid), dim3(64), lds, s, g, ks.EL, ks.VL, NB, W, WP, m1,
                       m2, pairs, np, p_of, queue, CNT, CEN, END, work);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_te_keys, dim3(te_grid(K, 256, 8192)), dim3(256), 0, s, pats, npats, ks,
                       K, NB, CNT, CEN, END, key_count, img);
    return hipGetLastError();
}

hipError_t te_gather(const int* img, const int* akeys, const int* n_active, int max_active,
                     int* counts, hipStream_t s) {
    if (max_active <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_te_gather, dim3(te_grid(8LL * max_active, 256, 8192)), dim3(256), 0, s,
                       img, akeys, n_active, max_active, counts);
    return hipGetLastError();
}

}  // namespace gsk
