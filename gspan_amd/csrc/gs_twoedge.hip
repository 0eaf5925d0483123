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
// counts per class (CNT and CEN only for the classes that are a parent's segment class --
// class (a, (b, e0)) or (b, (a, e0)) of the batch's single-edge parent p = (a, e0, b),
// a <= b -- at row 2 p, for the class of a, and 2 p + 1, for that of b when b != a):
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

// {bit, length} of every run
__global__ void k_te_runs(DevGraph g, int EL, int2* __restrict__ rbl) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < g.nruns; q += stride)
        rbl[q] = make_int2(run_bit(g.rkey[q], EL), g.rstart[q + 1] - g.rstart[q]);
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

// pairs per class (hist), then the pairs at their class positions (cursor = class starts,
// advanced); one thread per vertex over its runs.  A pair's record is two int4: {vertex,
// run, first entry of the run, run length} and {first run of the vertex, last run of the
// vertex, vertex label, run bit}; pcl holds {class, run length} again, 8 bytes per pair,
// for the per-class sums.
template <bool SCATTER>
__global__ void k_te_pairs(DevGraph g, int EL, int NB, int* __restrict__ hist_or_cursor,
                           const long long* __restrict__ cstart, int4* __restrict__ rec,
                           int2* __restrict__ pcl) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < g.n; v += stride) {
        const int o0 = g.offsets[v], o1 = g.offsets[v + 1];
        if (o0 == o1) continue;
        const int r0 = g.runid[o0], r1 = g.runid[o1 - 1];
        const int lv = g.vlab[v];
        const long long cbase = (long long)lv * NB;
        for (int k = r0; k <= r1; ++k) {
            const int b = run_bit(g.rkey[k], EL);
            const long long c = cbase + b;
            const int st = g.rstart[k], len = g.rstart[k + 1] - st;
            if (!SCATTER) {
                atomicAdd(&hist_or_cursor[c], 1);
            } else {
                const long long at = cstart[c] + atomicAdd(&hist_or_cursor[c], 1);
                rec[2 * at] = make_int4((int)v, k, st, len);
                rec[2 * at + 1] = make_int4(r0, r1, lv, b);
                pcl[at] = make_int2((int)c, len);
            }
        }
    }
}

// The three tables from the class-ordered pairs.  One wave per workgroup takes chunks of
// TE_CHUNK consecutive pairs from a queue (lane i loads the record of the chunk's pair i)
// and sums them per class -- CNT / CEN in LDS, END in registers (lane j holds the
// counts of bits j, j + 64, ...; NT blocks of 64 bits) -- adding the non-zero sums to the
// global tables when the class changes and at the end.  p_of[(a * EL + e0) * VL + b]: the
// batch index of the single-edge parent (0, 1, a, e0, b), a <= b, or -1; CNT / CEN only
// for classes that are parent runs of one.  END rows of other classes keep only the bits
// b' = (label lo, edge label eo) with (lo, eo, run label) a batch parent: pmask[l * W + w],
// word w of those bits for run label l -- the only entries a key reads there.  The wave
// is G = 64 / WP lane groups of WP >= W
// lanes (lane w of a group holds word w of an OR): up to G consecutive pairs of one class
// with short runs go one per group, so their load chains overlap; a pair with a longer run
// has its run's vertices split over the groups.
// work[2 t] += runs visited for CNT / CEN, work[2 t + 1] += words read for END (per run
// vertex: its adjacency entry, two words, and W + 1 mask words), t = workgroup mod
// GS_TE_WORK_SLOTS.
constexpr int TE_CHUNK = 64;
constexpr int TE_SHORT = 16;  // runs up to this length share the wave with other pairs

template <int NT>
__global__ __launch_bounds__(64) void k_te_accum(DevGraph g, int EL, int VL, int NB, int W,
                                                 int WP, const uint32_t* __restrict__ m1,
                                                 const uint32_t* __restrict__ m2,
                                                 const int4* __restrict__ rec,
                                                 const int2* __restrict__ rbl, long long npairs,
                                                 const int* __restrict__ p_of,
                                                 const uint32_t* __restrict__ pmask, int* queue,
                                                 unsigned long long* __restrict__ CNT,
                                                 int* __restrict__ CEN, int* __restrict__ END,
                                                 unsigned long long* work) {
    extern __shared__ unsigned long long te_smem[];
    unsigned long long* s_cnt = te_smem;
    int* s_cen = reinterpret_cast<int*>(s_cnt + NB);
    const int lane = threadIdx.x;
    for (int b = lane; b < NB; b += 64) {
        s_cnt[b] = 0;
        s_cen[b] = 0;
    }
    int endc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) endc[t] = 0;
    __syncthreads();
    const int G = 64 / WP, grp = lane / WP, wi = lane % WP;
    // runs up to this length share the wave with other pairs: one entry per lane of a group
    const int short_run = WP < TE_SHORT ? WP : TE_SHORT;
    const long long nchunks = (npairs + TE_CHUNK - 1) / TE_CHUNK;
    long long cls = -1;
    bool parent = false;  // the current class is a parent run of a batch pattern
    long long prow = 0;   // its CNT / CEN row
    uint32_t keep = 0;    // word wi of the END bits the class's row needs
    bool need_end = false;  // some of them
    unsigned long long w_runs = 0, w_words = 0;
    auto flush = [&]() {
        __syncthreads();
        const long long row = cls * NB;
        if (parent) {
            for (int b = lane; b < NB; b += 64) {
                if (s_cnt[b]) atomicAdd(&CNT[prow + b], s_cnt[b]);
                if (s_cen[b]) atomicAdd(&CEN[prow + b], s_cen[b]);
                s_cnt[b] = 0;
                s_cen[b] = 0;
            }
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int b = lane + 64 * t;
            if (endc[t]) atomicAdd(&END[row + b], endc[t]);
            endc[t] = 0;
        }
        __syncthreads();
    };
    // A workgroup's first chunk is its own index (the grid holds no more workgroups than the
    // device runs at once, so all start together); later ones come from the queue, whose
    // claims then spread over the pass instead of all arriving at its start.
    bool first_chunk = true;
    for (;;) {
        long long ch = 0;
        if (first_chunk) {
            ch = blockIdx.x;
            first_chunk = false;
        } else {
            if (lane == 0) ch = gridDim.x + (long long)atomicAdd(queue, 1);
            ch = __shfl(ch, 0);
        }
        if (ch >= nchunks) break;
        const long long p0 = ch * TE_CHUNK;
        const int cnt = (int)min((long long)TE_CHUNK, npairs - p0);
        int4 ra = make_int4(0, 0, 0, 0), rb = make_int4(0, 0, 0, 0);
        if (lane < cnt) {
            ra = rec[2 * (p0 + lane)];
            rb = rec[2 * (p0 + lane) + 1];
        }
        for (int i = 0; i < cnt;) {
            const int lx = __builtin_amdgcn_readlane(rb.z, i);
            const int b = __builtin_amdgcn_readlane(rb.w, i);
            const int el = b % EL;
            const long long c = (long long)lx * NB + b;
            if (c != cls) {
                if (cls >= 0) flush();
                cls = c;
                const int tl = b / EL;
                const int lo = lx < tl ? lx : tl, hi = lx < tl ? tl : lx;
                const int pi = p_of[(lo * EL + el) * VL + hi];
                parent = pi >= 0;
                prow = (2LL * pi + (lx == lo ? 0 : 1)) * NB;
                keep = parent ? ~0u : (wi < W ? pmask[(long long)tl * W + wi] : 0u);
                need_end = __ballot(keep != 0) != 0;
            }
            // Pairs i .. i + n - 1: of this class, at most G, runs of at most TE_SHORT
            // entries -- lane group q takes pair i + q.  A longer run is taken alone, its
            // entries split over the groups.
            const unsigned long long same =
                __ballot(lane < cnt && rb.z == lx && rb.w == b) >> i;  // bit 0 set
            const unsigned long long longr = __ballot(lane < cnt && ra.w > short_run) >> i;
            const bool alone = (longr & 1ull) != 0;
            int n = 1;
            if (!alone) {
                const unsigned long long z = ~same;
                n = z ? __builtin_ctzll(z) : 64 - i;
                n = n < G ? n : G;
                const unsigned long long lf = longr & (n >= 64 ? ~0ull : (1ull << n) - 1);
                if (lf) n = __builtin_ctzll(lf);
            }
            const bool act = alone || grp < n;
            const int pi = alone ? i : i + (act ? grp : 0);
            const int k = __shfl(ra.y, pi), st = __shfl(ra.z, pi), len = __shfl(ra.w, pi);
            const int q0 = __shfl(rb.x, pi), q1 = __shfl(rb.y, pi);
            // x as an end vertex z: the runs of its neighbours u in run b, beyond z (the
            // loads are issued first, the centre sums below overlap them)
            const int bs = lx * EL + el;  // the run of u that holds x
            uint32_t acc = 0;
            bool two = false;
            const int bw = bs >> 5;
            const uint32_t bbit = 1u << (bs & 31);
            if (need_end && act) {
                if (!alone) {
                    // lane wi of the group loads run entry wi (one round trip), then the
                    // group reads the entries' masks four at a time
                    const int u_own = wi < len ? (int)g.adjl[st + wi].x : 0;
                    if (wi < len) two = (m2[(long long)u_own * W + bw] & bbit) != 0;
                    for (int j = 0; j < len; j += 4) {
                        uint32_t a[4];
#pragma unroll
                        for (int t = 0; t < 4; ++t) {
                            const int u = __shfl(u_own, grp * WP + ((j + t) & (WP - 1)));
                            a[t] = (j + t < len && wi < W) ? m1[(long long)u * W + wi] : 0u;
                        }
                        acc |= a[0] | a[1] | a[2] | a[3];
                    }
                } else {
                    // the run's entries split over the groups, four per group at a time
                    for (int j = grp; j < len; j += 4 * G) {
                        int u[4];
#pragma unroll
                        for (int t = 0; t < 4; ++t)
                            u[t] = j + t * G < len ? (int)g.adjl[st + j + t * G].x : -1;
#pragma unroll
                        for (int t = 0; t < 4; ++t) {
                            if (u[t] < 0) continue;
                            if (wi < W) acc |= m1[(long long)u[t] * W + wi];
                            if (wi == 0) two |= (m2[(long long)u[t] * W + bw] & bbit) != 0;
                        }
                    }
                }
            }
            // x as the segment vertex of a parent row (x, w), w in run b: x's runs, two per
            // lane at a time (both loads in flight before the sums)
            if (parent) {
                const int qs = alone ? 64 : WP;
                for (int qb = q0 + (alone ? lane : wi); act && qb <= q1; qb += 2 * qs) {
                    int2 r[2];
#pragma unroll
                    for (int t = 0; t < 2; ++t)
                        r[t] = qb + t * qs <= q1 ? rbl[qb + t * qs] : make_int2(0, 0);
#pragma unroll
                    for (int t = 0; t < 2; ++t) {
                        const int wq = r[t].y - (qb + t * qs == k ? 1 : 0);
                        if (wq >= 1) {
                            atomicAdd(&s_cnt[r[t].x],
                                      (unsigned long long)len * (unsigned long long)wq);
                            atomicAdd(&s_cen[r[t].x], 1);
                        }
                    }
                }
            }
            const bool lead = alone ? lane == 0 : (act && wi == 0);  // counts the pair's work
            if (lead) {
                if (parent) w_runs += (unsigned long long)(q1 - q0 + 1);
                if (need_end) w_words += (unsigned long long)len * (unsigned long long)(W + 3);
            }
            if (alone) {
                for (int off = WP; off < 64; off <<= 1) acc |= __shfl_xor(acc, off);
                two = __ballot(two) != 0;
            } else {  // any lane of the group
                const unsigned long long gm = WP == 64 ? ~0ull : ((1ull << WP) - 1) << (grp * WP);
                two = (__ballot(two) & gm) != 0;
            }
            if (wi == bw) acc = (acc & ~bbit) | (two ? bbit : 0u);
            acc &= keep;
            for (int q = 0; q < n; ++q) {
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    const int bb = lane + 64 * t;
                    const uint32_t word = (uint32_t)__shfl((int)acc, (q * WP + (bb >> 5)) & 63);
                    endc[t] += (bb < NB && ((word >> (bb & 31)) & 1u)) ? 1 : 0;
                }
            }
            i += n;
        }
    }
    if (cls >= 0) flush();
    for (int off = 32; off > 0; off >>= 1) {
        w_runs += __shfl_xor(w_runs, off);
        w_words += __shfl_xor(w_words, off);
    }
    if (lane == 0) {  // spread over slots: every wave ends at about the same time
        const int t = (int)(blockIdx.x % GS_TE_WORK_SLOTS);
        atomicAdd(&work[2 * t], w_runs);
        atomicAdd(&work[2 * t + 1], w_words);
    }
}

// Per class: its pairs (npairs, from the class starts) and the CSR entries of its runs
// (nlen, zero beforehand).  Each wave takes TE_CLS_SPAN consecutive (class-ordered) pairs,
// all loaded at once, and sums them by class 64 at a time -- a segmented scan over its
// lanes -- carrying the class open at the last lane into the next 64: one atomic per class
// and span, not per 64 pairs (a large class's sums would queue on one address).
constexpr int TE_CLS_SPAN = 512;
__global__ __launch_bounds__(256) void k_te_classes(const int2* __restrict__ pcl, long long np,
                                                    const long long* __restrict__ cstart,
                                                    long long NC, int* __restrict__ npairs,
                                                    unsigned long long* __restrict__ nlen) {
    const int lane = threadIdx.x & 63;
    const long long gstride = (long long)gridDim.x * blockDim.x;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < NC; c += gstride)
        npairs[c] = (int)(cstart[c + 1] - cstart[c]);
    const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long waves = gstride >> 6;
    const long long nspans = (np + TE_CLS_SPAN - 1) / TE_CLS_SPAN;
    for (long long sp = wave; sp < nspans; sp += waves) {
        const long long p0 = sp * TE_CLS_SPAN, p1 = min(np, p0 + TE_CLS_SPAN);
        int2 v[TE_CLS_SPAN / 64];
#pragma unroll
        for (int t = 0; t < TE_CLS_SPAN / 64; ++t) {
            const long long i = p0 + 64 * t + lane;
            v[t] = i < p1 ? pcl[i] : make_int2(-1, 0);
        }
        int ccls = -1;
        unsigned long long carry = 0;
#pragma unroll
        for (int t = 0; t < TE_CLS_SPAN / 64; ++t) {
            const int c0 = __shfl(v[t].x, 0);
            if (ccls >= 0 && c0 != ccls && lane == 0) atomicAdd(&nlen[ccls], carry);
            unsigned long long sum = (unsigned long long)v[t].y;
            for (int off = 1; off < 64; off <<= 1) {
                const unsigned long long up = __shfl_up(sum, off);
                const int uc = __shfl_up(v[t].x, off);
                if (lane >= off && uc == v[t].x) sum += up;
            }
            if (ccls >= 0 && v[t].x == ccls) sum += carry;
            const int nc = __shfl_down(v[t].x, 1);
            if (v[t].x >= 0 && lane != 63 && nc != v[t].x) atomicAdd(&nlen[v[t].x], sum);
            ccls = __shfl(v[t].x, 63);
            carry = __shfl(sum, 63);
        }
        if (ccls >= 0 && lane == 0) atomicAdd(&nlen[ccls], carry);
    }
}

// Single-edge code i = (a * EL + e0) * VL + b: flags[i] = 1 when a <= b, the code has rows
// (CSR entries from label a to label b over e0, nlen of class (a, (b, e0))) and its support
// reaches min_sup: the least non-zero count of its vertices' images -- the pairs of classes
// (a, (b, e0)) and (b, (a, e0)) -- vertex 0 left out under GS_MNI_OMP_LITERAL, as no later
// edge leaves it (gspan.cu:282-288 / 367-373; seeding gspan.cu:410-426).
__global__ void k_te_seed_flags(const int* __restrict__ npairs,
                                const unsigned long long* __restrict__ nlen, int VL, int EL,
                                int mni_mode, int min_sup, int* __restrict__ flags) {
    const long long n = (long long)VL * EL * VL, NB = (long long)VL * EL;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int a = (int)(i / ((long long)EL * VL)), e0 = (int)((i / VL) % EL),
                  b = (int)(i % VL);
        int f = 0;
        if (a <= b) {
            const long long cl = a * NB + (long long)b * EL + e0;
            if (nlen[cl] > 0) {
                const int n0 = npairs[cl];
                const int n1 = npairs[(long long)b * NB + (long long)a * EL + e0];
                int s = INT_MAX;
                if (mni_mode != GS_MNI_OMP_LITERAL && n0 > 0) s = n0;
                if (n1 > 0 && n1 < s) s = n1;
                f = s >= min_sup ? 1 : 0;
            }
        }
        flags[i] = f;
    }
}

// p_of[i] = the index of flagged code i among the flagged codes (scan: exclusive sums of the
// flags), or -1; the pmask bits of every flagged code (a, e0, b): run bit (a, e0) of label b
// and (b, e0) of label a; *ncodes = the flagged codes.
__global__ void k_te_pof(const int* __restrict__ flags, const long long* __restrict__ scan,
                         int VL, int EL, int W, int* __restrict__ p_of, uint32_t* pmask,
                         int* ncodes) {
    const long long n = (long long)VL * EL * VL;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int f = flags[i];
        p_of[i] = f ? (int)scan[i] : -1;
        if (i == n - 1) *ncodes = (int)(scan[i] + f);
        if (f) {
            const int a = (int)(i / ((long long)EL * VL)), e0 = (int)((i / VL) % EL),
                      b = (int)(i % VL);
            const int ba = a * EL + e0, bb = b * EL + e0;
            atomicOr(&pmask[(long long)b * W + (ba >> 5)], 1u << (ba & 31));
            atomicOr(&pmask[(long long)a * W + (bb >> 5)], 1u << (bb & 31));
        }
    }
}

// p[0 .. min(*nrows * row_words, max_words)) = 0
__global__ void k_te_zero_rows(uint32_t* __restrict__ p, const int* __restrict__ nrows,
                               long long row_words, long long max_words) {
    long long n = (long long)*nrows * row_words;
    n = n < max_words ? n : max_words;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        p[i] = 0u;
}

// Every key of the batch (parents of one edge, so no backward slots): its child rows
// (key_count; past INT_MAX: INT_MIN, as the extension kernels' saturated counts) and its
// support (ksup): the least non-zero count of distinct images over the child's three
// vertices, vertex 0 only where it takes part (GS_MNI_OMP_LITERAL: when the new edge leaves
// it), INT_MAX when none is non-zero (gspan.cu:282-288 / 367-373).  Slots the rules reject
// (gspan.cu:455, 475-477) have no children.
__global__ void k_te_keys(const PatDev* __restrict__ pats, int npats, KeySpace ks, long long K,
                          int NB, int mni_mode, const unsigned long long* __restrict__ CNT,
                          const int* __restrict__ CEN, const int* __restrict__ END,
                          int* __restrict__ key_count, int* __restrict__ ksup) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long key = (long long)blockIdx.x * blockDim.x + threadIdx.x; key < K;
         key += stride) {
        const int lo = owner_one_edge(pats, npats, ks, (int)key);
        const PatDev& P = pats[lo];
        const int slot = (int)(key - P.key_base);
        const int j = slot / (ks.EL * ks.VL), e = (slot / ks.VL) % ks.EL, l = slot % ks.VL;
        if (l < P.min_label || !(P.seg_el[j] < e || (P.seg_el[j] == e && P.seg_tl[j] <= l))) {
            key_count[key] = 0;
            ksup[key] = INT_MAX;
            continue;
        }
        const int sv = P.seg_v[j], ov = 1 - sv;
        const int ls = P.vlab[sv], lo_ = P.vlab[ov], e0 = P.cel[0];
        const long long A = (long long)lo_ * ks.EL + e0, C = (long long)l * ks.EL + e;
        // the segment class's CNT / CEN row: 2 lo for the parent's smaller label
        const int lmin = P.vlab[0] < P.vlab[1] ? P.vlab[0] : P.vlab[1];
        const long long cs = (2LL * lo + (ls == lmin ? 0 : 1)) * NB + C;
        const unsigned long long cnt = CNT[cs];
        key_count[key] = cnt > (unsigned long long)INT_MAX ? INT_MIN : (int)cnt;
        int im[3];
        im[sv] = CEN[cs];
        im[ov] = END[((long long)lo_ * NB + ((long long)ls * ks.EL + e0)) * NB + C];
        im[2] = END[((long long)l * NB + ((long long)ls * ks.EL + e)) * NB + A];
        const bool v0 = mni_mode != GS_MNI_OMP_LITERAL || sv == 0;
        int sup = INT_MAX;
        for (int v = v0 ? 0 : 1; v < 3; ++v)
            if (im[v] > 0 && im[v] < sup) sup = im[v];
        ksup[key] = sup;
    }
}

// out[a] = ksup[akeys[a]] for the active keys (n_active null: max_active)
__global__ void k_te_gather(const int* __restrict__ ksup, const int* __restrict__ akeys,
                            const int* __restrict__ n_active, int max_active,
                            int* __restrict__ out) {
    const long long na = n_active ? *n_active : max_active;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < na; i += stride)
        out[i] = ksup[akeys[i]];
}

// Reports of a two-edge last level reached from every frequent single-edge code (the
// batch, in key order = their report order), in the reference's order: each single-edge
// code, then its frequent children in key order -- within a parent, key order (segment of
// the rightmost vertex first, then edge label, then label) is the order of their last edges.
// fkey / fsup: the F frequent children, keys ascending, and their supports; ssup: the
// codes' supports; vl_vals / el_vals: original label of each rank.  Report r of the output
// has lens[r] edges at codes[5 * (first edge of r)]; single-edge code p is report
// p + C_p with its edge at p + 2 C_p, C_p = children of the codes before it.
__device__ __forceinline__ int te_lower_bound(const int* __restrict__ a, int n, int v) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__global__ void k_te_emit(const PatDev* __restrict__ pats, int npats, KeySpace ks,
                          const int* __restrict__ fkey, const int* __restrict__ fsup, int F,
                          const int* __restrict__ ssup, const int* __restrict__ vl_vals,
                          const int* __restrict__ el_vals, int* __restrict__ codes,
                          int* __restrict__ lens, int* __restrict__ sups) {
    const long long total = (long long)npats + F;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
        int p, i = -1;
        if (t < npats) {
            p = (int)t;
        } else {
            i = (int)(t - npats);
            p = owner_one_edge(pats, npats, ks, fkey[i]);
        }
        const PatDev& P = pats[p];
        const long long cp = te_lower_bound(fkey, F, P.key_base);
        long long r = p + cp, e = p + 2 * cp;
        if (i >= 0) {
            const long long j = i - cp;
            r += 1 + j;
            e += 1 + 2 * j;
        }
        int* out = codes + 5 * e;
        out[0] = 0;
        out[1] = 1;
        out[2] = vl_vals[P.vlab[0]];
        out[3] = el_vals[P.cel[0]];
        out[4] = vl_vals[P.vlab[1]];
        lens[r] = i < 0 ? 1 : 2;
        sups[r] = i < 0 ? ssup[p] : fsup[i];
        if (i >= 0) {  // the child's edge, as its key names it (no backward slots)
            const int s2 = fkey[i] - P.key_base - (P.r - 1) * ks.EL;
            const int seg = s2 / (ks.EL * ks.VL);
            out[5] = seg == 0 ? P.v_n : P.cfrom[P.rm[seg - 1]];
            out[6] = P.v_n + 1;
            out[7] = -1;
            out[8] = el_vals[(s2 / ks.VL) % ks.EL];
            out[9] = vl_vals[s2 % ks.VL];
        }
    }
}

inline int te_grid(long long n, int block, int cap) {
    long long g = (n + block - 1) / block;
    return (int)(g < 1 ? 1 : (g > cap ? cap : g));
}

}  // namespace

namespace gsk {

hipError_t te_prepare(const DevGraph& g, int EL, int NB, int W, uint32_t* m1, uint32_t* m2,
                      int* hist, long long* cstart, long long* scan_tmp, long long scan_elems,
                      int4* rec, int2* pcl, int2* rbl, hipStream_t s) {
    const long long NC = (long long)(NB / EL) * NB;  // VL * NB classes
    hipError_t e;
    if ((e = fill_u32(m1, 0u, (long long)g.n * W, s)) != hipSuccess) return e;
    if ((e = fill_u32(m2, 0u, (long long)g.n * W, s)) != hipSuccess) return e;
    if ((e = fill_i32(hist, 0, NC, s)) != hipSuccess) return e;
    const int grid = te_grid(g.n, 256, 8192);
    hipLaunchKernelGGL(k_te_masks, dim3(grid), dim3(256), 0, s, g, EL, W, m1, m2);
    hipLaunchKernelGGL(k_te_runs, dim3(te_grid(g.nruns, 256, 8192)), dim3(256), 0, s, g, EL, rbl);
    hipLaunchKernelGGL(k_te_pairs<false>, dim3(grid), dim3(256), 0, s, g, EL, NB, hist,
                       (const long long*)nullptr, (int4*)nullptr, (int2*)nullptr);
    if ((e = scan_exclusive(hist, cstart, NC, scan_tmp, scan_elems, s)) != hipSuccess) return e;
    if ((e = fill_i32(hist, 0, NC, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_te_pairs<true>, dim3(grid), dim3(256), 0, s, g, EL, NB, hist, cstart,
                       rec, pcl);
    return hipGetLastError();
}

hipError_t te_clear(KeySpace ks, int npats, int* queue, unsigned long long* CNT, int* CEN,
                    int* END, hipStream_t s) {
    const long long NB = (long long)ks.VL * ks.EL, rows = ks.VL * NB * NB, prows = 2LL * npats * NB;
    hipError_t e;
    if ((e = fill_u32(reinterpret_cast<uint32_t*>(CNT), 0u, 2 * prows, s)) != hipSuccess) return e;
    if ((e = fill_i32(CEN, 0, prows, s)) != hipSuccess) return e;
    if ((e = fill_i32(END, 0, rows, s)) != hipSuccess) return e;
    return fill_i32(queue, 0, 1, s);
}

hipError_t te_seeds(const int* npairs, const unsigned long long* nlen, int VL, int EL,
                    int mni_mode, int min_sup, int* flags, long long* scan, long long* scan_tmp,
                    long long scan_elems, int* p_of, uint32_t* pmask, int* ncodes,
                    hipStream_t s) {
    const long long n = (long long)VL * EL * VL;
    const int W = (VL * EL + 31) / 32;
    hipError_t e;
    if ((e = fill_u32(pmask, 0u, (long long)VL * W, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_te_seed_flags, dim3(te_grid(n, 256, 8192)), dim3(256), 0, s, npairs,
                       nlen, VL, EL, mni_mode, min_sup, flags);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = scan_exclusive(flags, scan, n, scan_tmp, scan_elems, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_te_pof, dim3(te_grid(n, 256, 8192)), dim3(256), 0, s, flags, scan, VL,
                       EL, W, p_of, pmask, ncodes);
    return hipGetLastError();
}

hipError_t te_clear_dev(KeySpace ks, const int* npats, long long max_pats, int* queue,
                        unsigned long long* CNT, int* CEN, int* END, hipStream_t s) {
    const long long NB = (long long)ks.VL * ks.EL, rows = ks.VL * NB * NB;
    const long long cnt_words = 4 * NB * max_pats, cen_words = 2 * NB * max_pats;
    hipLaunchKernelGGL(k_te_zero_rows, dim3(te_grid(cnt_words, 256, 8192)), dim3(256), 0, s,
                       reinterpret_cast<uint32_t*>(CNT), npats, 4 * NB, cnt_words);
    hipError_t e;
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_te_zero_rows, dim3(te_grid(cen_words, 256, 8192)), dim3(256), 0, s,
                       reinterpret_cast<uint32_t*>(CEN), npats, 2 * NB, cen_words);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = fill_i32(END, 0, rows, s)) != hipSuccess) return e;
    return fill_i32(queue, 0, 1, s);
}

hipError_t te_level(const DevGraph& g, KeySpace ks, const PatDev* pats, int npats, long long K,
                    const int* p_of, const uint32_t* m1, const uint32_t* m2, const int4* rec,
                    const int2* rbl, const uint32_t* pmask, int* queue,
                    unsigned long long* CNT, int* CEN, int* END, int mni_mode, int* key_count,
                    int* ksup, unsigned long long* work, hipStream_t s) {
    hipError_t e = te_accum(g, ks, p_of, m1, m2, rec, rbl, pmask, queue, CNT, CEN, END, work, s);
    if (e != hipSuccess) return e;
    return te_keys(pats, npats, ks, K, CNT, CEN, END, mni_mode, key_count, ksup, s);
}

hipError_t te_accum(const DevGraph& g, KeySpace ks, const int* p_of, const uint32_t* m1,
                    const uint32_t* m2, const int4* rec, const int2* rbl, const uint32_t* pmask,
                    int* queue, unsigned long long* CNT, int* CEN, int* END,
                    unsigned long long* work, hipStream_t s) {
    const int NB = ks.VL * ks.EL, W = (NB + 31) / 32;
    int WP = 1;
    while (WP < W) WP <<= 1;
    const long long np = g.nruns;
    const size_t lds = (size_t)NB * 12;
    const long long nchunks = (np + TE_CHUNK - 1) / TE_CHUNK;
    const int nt = (NB + 63) / 64;
    // as many workgroups (one wave each) as the device holds at once, at most 8192
    int dev = 0, cus = 0, per = 0;
    hipError_t e;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) !=
        hipSuccess)
        return e;
#define GS_TE_OCC(T) \
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_te_accum<T>, 64, lds)
    if (nt <= 1) GS_TE_OCC(1);
    else if (nt <= 2) GS_TE_OCC(2);
    else if (nt <= 4) GS_TE_OCC(4);
    else if (nt <= 8) GS_TE_OCC(8);
    else if (nt <= 16) GS_TE_OCC(16);
    else return hipErrorInvalidValue;
#undef GS_TE_OCC
    if (e != hipSuccess) return e;
    long long cap = (long long)cus * per;
    cap = cap < 1 ? 1 : (cap > 8192 ? 8192 : cap);
    const int grid = (int)(nchunks < cap ? (nchunks < 1 ? 1 : nchunks) : cap);
#define GS_TE_ACCUM(T)                                                                        \
    hipLaunchKernelGGL(k_te_accum<T>, dim3(grid), dim3(64), lds, s, g, ks.EL, ks.VL, NB, W, WP, \
                       m1, m2, rec, rbl, np, p_of, pmask, queue, CNT, CEN, END, work)
    if (nt <= 1) GS_TE_ACCUM(1);
    else if (nt <= 2) GS_TE_ACCUM(2);
    else if (nt <= 4) GS_TE_ACCUM(4);
    else if (nt <= 8) GS_TE_ACCUM(8);
    else if (nt <= 16) GS_TE_ACCUM(16);
    else return hipErrorInvalidValue;
#undef GS_TE_ACCUM
    return hipGetLastError();
}

hipError_t te_keys(const PatDev* pats, int npats, KeySpace ks, long long K,
                   const unsigned long long* CNT, const int* CEN, const int* END, int mni_mode,
                   int* key_count, int* ksup, hipStream_t s) {
    if (K <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_te_keys, dim3(te_grid(K, 256, 8192)), dim3(256), 0, s, pats, npats, ks,
                       K, ks.VL * ks.EL, mni_mode, CNT, CEN, END, key_count, ksup);
    return hipGetLastError();
}

hipError_t te_classes(const int2* pcl, long long np, const long long* cstart, long long NC,
                      int* npairs, unsigned long long* nlen, hipStream_t s) {
    if (NC <= 0) return hipSuccess;
    hipError_t e;
    if ((e = fill_u32(reinterpret_cast<uint32_t*>(nlen), 0u, 2 * NC, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_te_classes, dim3(te_grid(np > NC ? np : NC, 256, 4096)), dim3(256), 0, s,
                       pcl, np, cstart, NC, npairs, nlen);
    return hipGetLastError();
}

hipError_t te_emit(const PatDev* pats, int npats, KeySpace ks, const int* fkey, const int* fsup,
                   int F, const int* ssup, const int* vl_vals, const int* el_vals, int* codes,
                   int* lens, int* sups, hipStream_t s) {
    const long long total = (long long)npats + F;
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_te_emit, dim3(te_grid(total, 256, 8192)), dim3(256), 0, s, pats, npats,
                       ks, fkey, fsup, F, ssup, vl_vals, el_vals, codes, lens, sups);
    return hipGetLastError();
}

hipError_t te_gather(const int* ksup, const int* akeys, const int* n_active, int max_active,
                     int* out, hipStream_t s) {
    if (max_active <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_te_gather, dim3(te_grid(max_active, 256, 8192)), dim3(256), 0, s, ksup,
                       akeys, n_active, max_active, out);
    return hipGetLastError();
}

}  // namespace gsk
