// Built-in objective kernels (the IProblemEvaluator::evaluate role,
// NI/problem/problem_evaluator.cuh:10-31).
//
// Quadratic-center: f_c(x) = sum_j (x_j - c_cj)^2.  HBM-bound: one wave reads a
// row once and produces all m objectives.  Fixed summation order (shared with
// the oracle): lane l accumulates j = l, l+64, ... ascending with one fma per
// element, then a pairwise tree over lanes (offsets 32, 16, ..., 1).
#include "batch_internal.hpp"
#include "common.hpp"
#include "moea_internal.hpp"

namespace moea {

namespace {

// One wave computes the m objectives of one row.
template <int M>
__device__ __forceinline__ void quadratic_row(const float* __restrict__ x, int d, const float* __restrict__ c,
                                     float* __restrict__ f, int ldf, float* __restrict__ cv, int row,
                                     int lane)
{
    const float* xr = x + static_cast<size_t>(row) * d;
    float acc[M];
#pragma unroll
    for (int k = 0; k < M; ++k) acc[k] = 0.0f;
    for (int j = lane; j < d; j += 64) {
        const float v = xr[j];
#pragma unroll
        for (int k = 0; k < M; ++k) {
            const float t = v - c[static_cast<size_t>(k) * d + j];
            acc[k] = fmaf(t, t, acc[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < M; ++k) {
        float s = acc[k];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) s = s + __shfl_xor(s, o);
        acc[k] = s;
    }
    if (lane < M) {
        float out = acc[0];
#pragma unroll
        for (int k = 1; k < M; ++k) if (lane == k) out = acc[k];
        f[static_cast<size_t>(lane) * ldf + row] = out;
    }
    if (lane == 0 && cv) cv[row] = 0.0f;
}

template <int M>
__global__ __launch_bounds__(256) void quadratic_kernel(const float* __restrict__ x, int n, int d,
                                                        const float* __restrict__ c,
                                                        float* __restrict__ f, int ldf,
                                                        float* __restrict__ cv)
{
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= n) return;
    quadratic_row<M>(x, d, c, f, ldf, cv, row, static_cast<int>(threadIdx.x & 63));
}

// Batched: the instances with M objectives, 4 rows per block; `which` 0 is
// the parent slot, 1 the offspring slot.
template <int M>
__global__ __launch_bounds__(256) void b_quadratic_kernel(const BInst* __restrict__ insts, BMap mp, int which)
{
    int local;
    const int i = bmap_find(mp, static_cast<int>(blockIdx.x), local);
    const BInst& I = insts[i];
    const int row = local * 4 + static_cast<int>(threadIdx.x >> 6);
    if (row >= I.n) return;
    const int slot = (I.cur + which) % 3;
    quadratic_row<M>(I.x[slot], I.d, I.centers, I.f[slot], I.n, I.cv[slot], row,
                     static_cast<int>(threadIdx.x & 63));
}

// MLP hidden layer H = tanh(X W1^T): (n, d) x (hidden, d)^T -> (n, hidden).
// 64 x 64 output tile per 256-thread block, 16-wide k slices staged in LDS,
// each thread owns a 4 x 4 register sub-tile.
constexpr int kGT = 64, kGK = 16;

__device__ __forceinline__ void mlp_hidden_tile(const float* __restrict__ x, int n, int d,
                                                const float* __restrict__ w1, int hidden,
                                                float* __restrict__ h, int r0, int c0)
{
    __shared__ float s_x[kGK][kGT + 1];
    __shared__ float s_w[kGK][kGT + 1];
    const int tr = threadIdx.x / 16, tc = threadIdx.x % 16;
    float acc[4][4] = {};
    for (int k0 = 0; k0 < d; k0 += kGK) {
        for (int e = threadIdx.x; e < kGT * kGK; e += 256) {
            const int rr = e / kGK, kk = e % kGK;
            const int gr = r0 + rr, gk = k0 + kk;
            s_x[kk][rr] = (gr < n && gk < d) ? x[static_cast<size_t>(gr) * d + gk] : 0.0f;
            const int gc = c0 + rr;
            s_w[kk][rr] = (gc < hidden && gk < d) ? w1[static_cast<size_t>(gc) * d + gk] : 0.0f;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kGK; ++kk) {
            float a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = s_x[kk][tr * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = s_w[kk][tc * 4 + j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int gr = r0 + tr * 4 + i, gc = c0 + tc * 4 + j;
            if (gr < n && gc < hidden) h[static_cast<size_t>(gr) * hidden + gc] = tanhf(acc[i][j]);
        }
}

__global__ __launch_bounds__(256) void mlp_hidden_kernel(const float* __restrict__ x, int n, int d,
                                                         const float* __restrict__ w1, int hidden,
                                                         float* __restrict__ h)
{
    mlp_hidden_tile(x, n, d, w1, hidden, h, static_cast<int>(blockIdx.x) * kGT,
                    static_cast<int>(blockIdx.y) * kGT);
}

// Output head: y = W2 h, f_c = y_c^2 + 0.01 * ||x - 0.5||^2; one wave per row.
template <int M>
__device__ __forceinline__ void mlp_head_row(const float* __restrict__ x, int d, const float* __restrict__ h,
                                             int hidden, const float* __restrict__ w2, float* __restrict__ f,
                                             int ldf, float* __restrict__ cv, int row, int lane)
{
    float y[M];
#pragma unroll
    for (int k = 0; k < M; ++k) y[k] = 0.0f;
    for (int j = lane; j < hidden; j += 64) {
        const float hv = h[static_cast<size_t>(row) * hidden + j];
#pragma unroll
        for (int k = 0; k < M; ++k) y[k] = fmaf(w2[static_cast<size_t>(k) * hidden + j], hv, y[k]);
    }
    float s = 0.0f;
    for (int j = lane; j < d; j += 64) {
        const float t = x[static_cast<size_t>(row) * d + j] - 0.5f;
        s = fmaf(t, t, s);
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        s = s + __shfl_xor(s, o);
#pragma unroll
        for (int k = 0; k < M; ++k) y[k] = y[k] + __shfl_xor(y[k], o);
    }
    if (lane < M) {
        float yk = y[0];
#pragma unroll
        for (int k = 1; k < M; ++k) if (lane == k) yk = y[k];
        f[static_cast<size_t>(lane) * ldf + row] = fmaf(yk, yk, 0.01f * s);
    }
    if (lane == 0 && cv) cv[row] = 0.0f;
}

template <int M>
__global__ __launch_bounds__(256) void mlp_head_kernel(const float* __restrict__ x, int n, int d,
                                                       const float* __restrict__ h, int hidden,
                                                       const float* __restrict__ w2,
                                                       float* __restrict__ f, int ldf,
                                                       float* __restrict__ cv)
{
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= n) return;
    mlp_head_row<M>(x, d, h, hidden, w2, f, ldf, cv, row, static_cast<int>(threadIdx.x & 63));
}

// Batched MLP: hidden-layer tiles, then the head rows of the instances with M objectives.
__global__ __launch_bounds__(256) void b_mlp_hidden_kernel(const BInst* __restrict__ insts, BMap mp, int which)
{
    int local;
    const int i = bmap_find(mp, static_cast<int>(blockIdx.x), local);
    const BInst& I = insts[i];
    const int slot = (I.cur + which) % 3;
    const int nr = (I.n + kGT - 1) / kGT;
    mlp_hidden_tile(I.x[slot], I.n, I.d, I.w1, I.hidden, I.hwork, (local % nr) * kGT, (local / nr) * kGT);
}

template <int M>
__global__ __launch_bounds__(256) void b_mlp_head_kernel(const BInst* __restrict__ insts, BMap mp, int which)
{
    int local;
    const int i = bmap_find(mp, static_cast<int>(blockIdx.x), local);
    const BInst& I = insts[i];
    const int row = local * 4 + static_cast<int>(threadIdx.x >> 6);
    if (row >= I.n) return;
    const int slot = (I.cur + which) % 3;
    mlp_head_row<M>(I.x[slot], I.d, I.hwork, I.hidden, I.w2, I.f[slot], I.n, I.cv[slot], row,
                    static_cast<int>(threadIdx.x & 63));
}

// DTLZ1 / DTLZ2 (Deb, Thiele, Laumanns, Zitzler 2005) with the reference's
// protections: dtlz1_func_kernel / dtlz2_func_kernel
// (problem/dtlz/dtlz_kernels.cu:28-112, 118-190).  g sums over the last
// k = d - m + 1 variables; DTLZ1 clamps x_j to [0, 1] in the products and
// 1 - x to >= 0; DTLZ2 uses max(0, cos).  The angle arguments are formed in
// double and rounded to float, as the reference's M_PI expressions are.  One
// wave per row; lanes own the objectives.
template <int M, int KIND>
__device__ __forceinline__ void dtlz_rows(const float* __restrict__ x, int n, int d, float* __restrict__ f, int ldf,
                                 float* __restrict__ cv)
{
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const float* xr = x + static_cast<size_t>(row) * d;
    float s = 0.0f;
    for (int j = M - 1 + lane; j < d; j += 64) {
        const float v = xr[j] - 0.5f;
        if (KIND == 1) s = s + (v * v - cosf(static_cast<float>(20.0 * M_PI * static_cast<double>(v))));
        else s = s + v * v;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) s = s + __shfl_xor(s, o);
    if (lane < M) {
        const int k = d - (M - 1);
        float out = (KIND == 1) ? 0.5f * (1.0f + 100.0f * (static_cast<float>(k) + s)) : 1.0f + s;
        const int prod = M - 1 - lane;
        for (int j = 0; j < prod; ++j) {
            const float xj = xr[j];
            if (KIND == 1) out *= fminf(fmaxf(xj, 0.0f), 1.0f);
            else out *= fmaxf(0.0f, cosf(static_cast<float>(static_cast<double>(xj) * M_PI / 2.0)));
        }
        if (lane > 0) {
            const float xi = xr[M - 1 - lane];
            if (KIND == 1) out *= fmaxf(0.0f, 1.0f - xi);
            else out *= sinf(static_cast<float>(static_cast<double>(xi) * M_PI / 2.0));
        }
        f[static_cast<size_t>(lane) * ldf + row] = out;
    }
    if (lane == 0 && cv) cv[row] = 0.0f;
}

template <int M>
__global__ __launch_bounds__(256) void dtlz1_kernel(const float* __restrict__ x, int n, int d,
                                                    float* __restrict__ f, int ldf, float* __restrict__ cv)
{
    dtlz_rows<M, 1>(x, n, d, f, ldf, cv);
}

template <int M>
__global__ __launch_bounds__(256) void dtlz2_kernel(const float* __restrict__ x, int n, int d,
                                                    float* __restrict__ f, int ldf, float* __restrict__ cv)
{
    dtlz_rows<M, 2>(x, n, d, f, ldf, cv);
}

}  // namespace

int eval_dtlz(int kind, const float* x, int n, int d, int m, float* f, int ldf, float* cv, hipStream_t stream)
{
    if (!x || !f || n <= 0 || m <= 1 || m > kMaxM || d < m || ldf < n || (kind != 1 && kind != 2))
        return fail(kInvalidArgument, "eval_dtlz: invalid arguments (DTLZ1/2 need 2 <= m <= d)");
    const dim3 grid(ceil_div(n, 4)), block(256);
    if (kind == 1) {
        MOEA_LAUNCH_M(m, dtlz1_kernel, grid, block, 0, stream, x, n, d, f, ldf, cv);
    } else {
        MOEA_LAUNCH_M(m, dtlz2_kernel, grid, block, 0, stream, x, n, d, f, ldf, cv);
    }
    MOEA_HIP(hipGetLastError());
    return kOk;
}

int eval_mlp(const float* x, int n, int d, const float* w1, int hidden, const float* w2, int m,
             float* f, int ldf, float* cv, float* hwork, hipStream_t stream)
{
    if (!x || !w1 || !w2 || !f || !hwork || n <= 0 || d <= 0 || hidden <= 0 || m <= 0 || m > kMaxM ||
        ldf < n)
        return fail(kInvalidArgument, "eval_mlp: invalid arguments");
    hipLaunchKernelGGL(mlp_hidden_kernel, dim3(ceil_div(n, kGT), ceil_div(hidden, kGT)), dim3(256), 0,
                       stream, x, n, d, w1, hidden, hwork);
    MOEA_HIP(hipGetLastError());
    const dim3 grid(ceil_div(n, 4)), block(256);
    switch (m) {
#define MOEA_H(MM) case MM: hipLaunchKernelGGL(mlp_head_kernel<MM>, grid, block, 0, stream, x, n, d, hwork, hidden, w2, f, ldf, cv); break;
        MOEA_H(1) MOEA_H(2) MOEA_H(3) MOEA_H(4) MOEA_H(5) MOEA_H(6) MOEA_H(7) MOEA_H(8) MOEA_H(9) MOEA_H(10)
#undef MOEA_H
        default: return fail(kUnsupported, "eval_mlp: unsupported objective count");
    }
    MOEA_HIP(hipGetLastError());
    return kOk;
}

int eval_quadratic(const float* x, int n, int d, const float* centers, int m, float* f, int ldf,
                   float* cv, hipStream_t stream)
{
    if (!x || !centers || !f || n <= 0 || d <= 0 || m <= 0 || m > kMaxM || ldf < n)
        return fail(kInvalidArgument, "eval_quadratic: invalid arguments");
    const dim3 grid(ceil_div(n, 4)), block(256);
    switch (m) {
#define MOEA_Q(MM) case MM: hipLaunchKernelGGL(quadratic_kernel<MM>, grid, block, 0, stream, x, n, d, centers, f, ldf, cv); break;
        MOEA_Q(1) MOEA_Q(2) MOEA_Q(3) MOEA_Q(4) MOEA_Q(5) MOEA_Q(6) MOEA_Q(7) MOEA_Q(8) MOEA_Q(9) MOEA_Q(10)
#undef MOEA_Q
        default: return fail(kUnsupported, "eval_quadratic: unsupported objective count");
    }
    MOEA_HIP(hipGetLastError());
    return kOk;
}

int batch_evaluate(const BatchLaunch& L, const BEvalGrids& G, int which)
{
    for (int m = 1; m <= kMaxM; ++m) {
        if (G.quad_rows4[m].blocks <= 0) continue;
        const dim3 grid(G.quad_rows4[m].blocks), block(256);
        MOEA_LAUNCH_M(m, b_quadratic_kernel, grid, block, 0, L.stream, L.insts, G.quad_rows4[m].map, which);
        MOEA_HIP(hipGetLastError());
    }
    if (G.mlp_tiles.blocks > 0) {
        hipLaunchKernelGGL(b_mlp_hidden_kernel, dim3(G.mlp_tiles.blocks), dim3(256), 0, L.stream, L.insts,
                           G.mlp_tiles.map, which);
        MOEA_HIP(hipGetLastError());
        for (int m = 1; m <= kMaxM; ++m) {
            if (G.mlp_rows4[m].blocks <= 0) continue;
            const dim3 grid(G.mlp_rows4[m].blocks), block(256);
            MOEA_LAUNCH_M(m, b_mlp_head_kernel, grid, block, 0, L.stream, L.insts, G.mlp_rows4[m].map, which);
            MOEA_HIP(hipGetLastError());
        }
    }
    return kOk;
}

}  // namespace moea
