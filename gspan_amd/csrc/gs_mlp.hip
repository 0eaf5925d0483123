// gs_mlp.hip -- the child-count predictor: features of candidate child keys and an MLP
// 32 -> 128 -> 128 -> 1 (ReLU, bf16 weights, fp32 accumulation) with its training
// kernels (forward, backward, Adam).  The work is tiny next to mining, so the layers are
// plain LDS-free fp32 FMA kernels: one thread per output element, weights read as bf16.
#include "gs_internal.h"

namespace {

__device__ __forceinline__ float bf16_to_f32(uint16_t h) {
    return __uint_as_float((uint32_t)h << 16);
}

// round to nearest even (finite inputs; the weights stay finite under Adam)
__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// Y[n, m] = act(b[m] + sum_k X[n, k] * W[k, m]),  W row-major [K][M] in bf16
__global__ void k_dense_fwd(const float* __restrict__ X, long long N, int K,
                            const uint16_t* __restrict__ W, const float* __restrict__ b, int M,
                            int relu, float* __restrict__ Y) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N * M) return;
    const long long n = i / M;
    const int m = (int)(i % M);
    const float* x = X + n * K;
    float acc = b[m];
    for (int k = 0; k < K; ++k) acc = fmaf(x[k], bf16_to_f32(W[(long long)k * M + m]), acc);
    Y[i] = relu ? fmaxf(acc, 0.f) : acc;
}

// dX[n, k] = (sum_m dY[n, m] * W[k, m]) * (act[n, k] > 0)   (act = the layer's input, the
// previous layer's ReLU output; null for the network input)
__global__ void k_dense_bwd_data(const float* __restrict__ dY, long long N, int M,
                                 const uint16_t* __restrict__ W, int K,
                                 const float* __restrict__ act, float* __restrict__ dX) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N * K) return;
    const long long n = i / K;
    const int k = (int)(i % K);
    const float* dy = dY + n * M;
    float acc = 0.f;
    for (int m = 0; m < M; ++m) acc = fmaf(dy[m], bf16_to_f32(W[(long long)k * M + m]), acc);
    if (act && act[i] <= 0.f) acc = 0.f;
    dX[i] = acc;
}

// dW[k, m] += sum_n X[n, k] * dY[n, m];  db[m] += sum_n dY[n, m]  over a chunk of samples
// per workgroup, accumulated with one atomic per (workgroup, weight)
constexpr int WCHUNK = 256;
__global__ void k_dense_bwd_weight(const float* __restrict__ X, const float* __restrict__ dY,
                                   long long N, int K, int M, float* dW, float* db) {
    const long long n0 = (long long)blockIdx.x * WCHUNK;
    const long long n1 = min(n0 + WCHUNK, N);
    for (int e = threadIdx.x; e < K * M + M; e += blockDim.x) {
        float acc = 0.f;
        if (e < K * M) {
            const int k = e / M, m = e % M;
            for (long long n = n0; n < n1; ++n) acc = fmaf(X[n * K + k], dY[n * M + m], acc);
            atomicAdd(&dW[e], acc);
        } else {
            const int m = e - K * M;
            for (long long n = n0; n < n1; ++n) acc += dY[n * M + m];
            atomicAdd(&db[m], acc);
        }
    }
}

// dy = 2 (y - t) / N;  loss += (y - t)^2 / N
__global__ void k_mse(const float* __restrict__ y, const float* __restrict__ t, long long N,
                      float* __restrict__ dy, float* loss) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (i < N) {
        const float d = y[i] - t[i];
        dy[i] = 2.f * d / (float)N;
        l = d * d / (float)N;
    }
    for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off);
    if (__lane_id() == 0 && l != 0.f) atomicAdd(loss, l);
}

__global__ void k_adam(float* __restrict__ w, float* __restrict__ mo, float* __restrict__ ve,
                       const float* __restrict__ g, int n, float lr, float beta1, float beta2,
                       float eps, float c1, float c2, uint16_t* __restrict__ w16) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float m = beta1 * mo[i] + (1.f - beta1) * gi;
    const float v = beta2 * ve[i] + (1.f - beta2) * gi * gi;
    mo[i] = m;
    ve[i] = v;
    const float wi = w[i] - lr * (m / c1) / (sqrtf(v / c2) + eps);
    w[i] = wi;
    w16[i] = f32_to_bf16(wi);
}

__global__ void k_to_bf16(const float* __restrict__ w, int n, uint16_t* __restrict__ w16) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) w16[i] = f32_to_bf16(w[i]);
}

// Features of every candidate child key of a batch (see FEATURES in gs_internal.h).
__global__ void k_features(const PatDev* __restrict__ pats, int npats, KeySpace ks, long long K,
                           const float* __restrict__ H, const float* __restrict__ vfreq,
                           const float* __restrict__ efreq, const float* __restrict__ dmean,
                           float nverts, float* __restrict__ X,
                           unsigned char* __restrict__ possible) {
    const long long key = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (key >= K) return;
    int lo = 0, hi = npats - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (pats[mid].key_base <= key) lo = mid;
        else hi = mid - 1;
    }
    const PatDev& P = pats[lo];
    const int slot = (int)(key - P.key_base);
    const int nb = (P.r - 1) * ks.EL;
    int back, pos, el, src_lab, dst_lab;
    bool ok;
    if (slot < nb) {  // backward (gspan.cu:436-438)
        back = 1;
        const int i = slot / ks.EL + 1;
        el = slot % ks.EL;
        const int k = P.rm[i];
        pos = i;
        src_lab = P.vlab[P.v_n];
        dst_lab = P.vlab[P.cfrom[k]];
        ok = P.cel[k] < el || (P.cel[k] == el && P.vlab[P.cto[k]] <= P.vlab[P.v_n]);
    } else {  // forward (gspan.cu:455, 475-477)
        back = 0;
        const int s = slot - nb;
        const int j = s / (ks.EL * ks.VL);
        el = (s / ks.VL) % ks.EL;
        dst_lab = s % ks.VL;
        pos = j;
        const int src = j == 0 ? P.v_n : P.cfrom[P.rm[j - 1]];
        src_lab = P.vlab[src];
        ok = dst_lab >= P.min_label;
        if (j > 0) {
            const int k = P.rm[j - 1];
            ok = ok && (P.cel[k] < el || (P.cel[k] == el && P.vlab[P.cto[k]] <= dst_lab));
        }
    }
    const float rows = (float)P.row_count;
    const float h = H ? H[((long long)src_lab * ks.EL + el) * ks.VL + dst_lab] : 0.f;
    // expected children: forward = rows x neighbours of the source with (el, dst label);
    // backward = rows x chance that the specific closing vertex is such a neighbour
    const float fdst = vfreq[dst_lab];
    const float naive = back ? rows * fminf(1.f, h / fmaxf(fdst * nverts, 1.f)) : rows * h;
    const float okf = ok ? 1.f : 0.f;
    float* x = X + key * GS_FEATURES;
    const float f[GS_FEATURES] = {
        log1pf(rows), log1pf(naive), okf, (float)back, pos / 7.f, P.m / 7.f, P.nv / 8.f,
        P.r / 7.f, log1pf(h), vfreq[src_lab], fdst, efreq[el], log1pf(dmean[src_lab]),
        log1pf(dmean[dst_lab]), src_lab == dst_lab ? 1.f : 0.f, okf * log1pf(rows),
        okf * log1pf(naive), log1pf(naive) - log1pf(rows), log1pf((float)P.support),
        P.min_label == src_lab ? 1.f : 0.f, P.min_label == dst_lab ? 1.f : 0.f,
        okf * (float)back, okf * log1pf(h), src_lab / 64.f, dst_lab / 64.f, el / 8.f,
        okf * vfreq[dst_lab], okf * log1pf(dmean[src_lab]), 0.f, 0.f, 0.f, 1.f};
#pragma unroll
    for (int q = 0; q < GS_FEATURES; ++q) x[q] = f[q];
    if (possible) possible[key] = ok ? 1 : 0;
}

// target = log1p(count) for logged samples; predicted count = expm1(y)
__global__ void k_log_targets(const int* __restrict__ counts, long long K, float* __restrict__ t) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < K) t[i] = log1pf(counts[i] < 0 ? 2147483648.0f : (float)counts[i]);  // < 0: > INT_MAX
}

inline unsigned blocks(long long n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

namespace gsk {

hipError_t mlp_forward(const float* X, long long N, const MlpDev& w, float* h1, float* h2,
                       float* y, hipStream_t s) {
    if (N <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dense_fwd, dim3(blocks(N * GS_HIDDEN, 256)), dim3(256), 0, s, X, N,
                       GS_FEATURES, w.w16[0], w.b[0], GS_HIDDEN, 1, h1);
    hipLaunchKernelGGL(k_dense_fwd, dim3(blocks(N * GS_HIDDEN, 256)), dim3(256), 0, s, h1, N,
                       GS_HIDDEN, w.w16[1], w.b[1], GS_HIDDEN, 1, h2);
    hipLaunchKernelGGL(k_dense_fwd, dim3(blocks(N, 256)), dim3(256), 0, s, h2, N, GS_HIDDEN,
                       w.w16[2], w.b[2], 1, 0, y);
    return hipGetLastError();
}

hipError_t mlp_backward(const float* X, const float* t, long long N, const MlpDev& w,
                        const float* h1, const float* h2, const float* y, float* dy, float* dh2,
                        float* dh1, float* grads, float* loss, hipStream_t s) {
    if (N <= 0) return hipSuccess;
    float* gW1 = grads;
    float* gb1 = gW1 + GS_FEATURES * GS_HIDDEN;
    float* gW2 = gb1 + GS_HIDDEN;
    float* gb2 = gW2 + GS_HIDDEN * GS_HIDDEN;
    float* gW3 = gb2 + GS_HIDDEN;
    float* gb3 = gW3 + GS_HIDDEN;
    hipLaunchKernelGGL(k_mse, dim3(blocks(N, 256)), dim3(256), 0, s, y, t, N, dy, loss);
    // layer 3 (128 -> 1, linear)
    hipLaunchKernelGGL(k_dense_bwd_weight, dim3(blocks(N, WCHUNK)), dim3(256), 0, s, h2, dy, N,
                       GS_HIDDEN, 1, gW3, gb3);
    hipLaunchKernelGGL(k_dense_bwd_data, dim3(blocks(N * GS_HIDDEN, 256)), dim3(256), 0, s, dy,
                       N, 1, w.w16[2], GS_HIDDEN, h2, dh2);
    // layer 2 (128 -> 128, ReLU): dh2 already carries the ReLU mask of h2
    hipLaunchKernelGGL(k_dense_bwd_weight, dim3(blocks(N, WCHUNK)), dim3(256), 0, s, h1, dh2, N,
                       GS_HIDDEN, GS_HIDDEN, gW2, gb2);
    hipLaunchKernelGGL(k_dense_bwd_data, dim3(blocks(N * GS_HIDDEN, 256)), dim3(256), 0, s, dh2,
                       N, GS_HIDDEN, w.w16[1], GS_HIDDEN, h1, dh1);
    // layer 1 (32 -> 128, ReLU)
    hipLaunchKernelGGL(k_dense_bwd_weight, dim3(blocks(N, WCHUNK)), dim3(256), 0, s, X, dh1, N,
                       GS_FEATURES, GS_HIDDEN, gW1, gb1);
    return hipGetLastError();
}

hipError_t mlp_adam(MlpDev& w, float* grads, float* mo, float* ve, int step, float lr,
                    hipStream_t s) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3(blocks(GS_MLP_PARAMS, 256)), dim3(256), 0, s, w.w32,
                       mo, ve, grads, GS_MLP_PARAMS, lr, b1, b2, eps, c1, c2, w.w16_all);
    return hipGetLastError();
}

hipError_t mlp_sync_bf16(MlpDev& w, hipStream_t s) {
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks(GS_MLP_PARAMS, 256)), dim3(256), 0, s, w.w32,
                       GS_MLP_PARAMS, w.w16_all);
    return hipGetLastError();
}

hipError_t features(const PatDev* pats, int npats, KeySpace ks, long long K, const float* H,
                    const float* vfreq, const float* efreq, const float* dmean, float nverts,
                    float* X, unsigned char* possible, hipStream_t s) {
    if (K <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_features, dim3(blocks(K, 256)), dim3(256), 0, s, pats, npats, ks, K, H,
                       vfreq, efreq, dmean, nverts, X, possible);
    return hipGetLastError();
}

hipError_t log_targets(const int* counts, long long K, float* t, hipStream_t s) {
    if (K <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_log_targets, dim3(blocks(K, 256)), dim3(256), 0, s, counts, K, t);
    return hipGetLastError();
}

}  // namespace gsk
