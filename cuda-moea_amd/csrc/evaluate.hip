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
    // each lane sums its elements j = lane, lane + 64, ... in that order; four
    // of them loaded at a time
    int j = lane;
    for (; j + 192 < d; j += 256) {
        float v[4], cc[M][4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = xr[j + 64 * u];
#pragma unroll
        for (int k = 0; k < M; ++k)
#pragma unroll
            for (int u = 0; u < 4; ++u) cc[k][u] = c[static_cast<size_t>(k) * d + j + 64 * u];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int k = 0; k < M; ++k) {
                const float t = v[u] - cc[k][u];
                acc[k] = fmaf(t, t, acc[k]);
            }
    }
    for (; j < d; j += 64) {
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

// MLP hidden layer H = tanh(X W1^T): (n, d) x (hidden, d)^T -> (n, hidden),
// on the FP32 matrix cores (v_mfma_f32_16x16x4_f32).  A 256-thread block
// owns a 64 x 128 tile of H; its four waves each own 32 x 64 (2 x 4 MFMA
// tiles).  K advances in slices of 32 staged in LDS (double-buffered, next
// slice prefetched into registers during the current one's MFMAs).  Both
// operands are K-contiguous, so a lane reads four consecutive k of its row
// with one 16-byte LDS read; lane group g = lane / 16 takes k = 4g + s in MFMA
// step s (the MFMA sums over its four lane groups), which covers every k of
// the slice once.  Accumulation is FP32 throughout; the summation order is
// fixed by the tiling, so a row of H does not depend on anything else in the
// launch (the batched launch reproduces the single-run values bit for bit).
typedef float f32x4 __attribute__((ext_vector_type(4)));
constexpr int kHM = 64, kHN = 128, kHK = 32, kHS = kHK + 4;   // tile rows, cols, k slice, LDS row stride

// (row tile, col tile) of tile index t: the col tiles of one row tile sit
// 8 indices apart, so a round-robin dispatch over the 8 XCDs puts them on one
// XCD and its L2 serves the shared X rows.  Returns false past the last row tile.
__device__ __forceinline__ bool mlp_tile_of(int t, int nrt, int nct, int& rt, int& ct)
{
    const int xcd = t & 7, q = t >> 3;
    ct = q % nct;
    rt = (q / nct) * 8 + xcd;
    return rt < nrt;
}

}  // namespace

int mlp_tile_count(int n, int hidden)
{
    const int nrt = (n + kHM - 1) / kHM, nct = (hidden + kHN - 1) / kHN;
    return ((nrt + 7) / 8) * 8 * nct;
}

namespace {

struct MlpSmem {
    float a[2][kHM * kHS];
    float b[2][kHN * kHS];
};

// Loads are issued unconditionally from clamped (in-bounds) addresses and the
// out-of-range elements are zeroed when the registers are stored to LDS, so
// no load sits behind a branch (a branch join would wait for it).
// VEC: ld % 4 == 0 (one 16-byte load); otherwise four scalar loads.
template <bool VEC>
__device__ __forceinline__ float4 mlp_raw4(const float* __restrict__ base, int row, int nrows, int ld, int k)
{
    const float* p = base + static_cast<size_t>(min(row, nrows - 1)) * ld;
    if (VEC) return *reinterpret_cast<const float4*>(p + min(k, ld - 4));
    float4 v;
    v.x = p[min(k, ld - 1)];
    v.y = p[min(k + 1, ld - 1)];
    v.z = p[min(k + 2, ld - 1)];
    v.w = p[min(k + 3, ld - 1)];
    return v;
}

__device__ __forceinline__ float4 mlp_mask4(float4 v, int row, int nrows, int ld, int k)
{
    const bool r = row < nrows;
    v.x = (r && k < ld) ? v.x : 0.0f;
    v.y = (r && k + 1 < ld) ? v.y : 0.0f;
    v.z = (r && k + 2 < ld) ? v.z : 0.0f;
    v.w = (r && k + 3 < ld) ? v.w : 0.0f;
    return v;
}

// The blocks of column tile 0 also sum (x - 0.5)^2 over each row of their
// row tile from the x slices they stage (each of the 8 loader lanes of a row
// sums its 4 columns of every slice, then a fixed xor reduction over the 8)
// and store it in rowsq[row]; the head reads it there instead of reading x
// again.
template <bool VEC>
__device__ __forceinline__ void mlp_hidden_tile(const float* __restrict__ x, int n, int d,
                                                const float* __restrict__ w1, int hidden,
                                                float* __restrict__ h, int r0, int c0, MlpSmem& sm,
                                                float* __restrict__ rowsq)
{
    const int tid = static_cast<int>(threadIdx.x), lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int g = lane >> 4, i = lane & 15;
    // loader: 8 float4 per 32-float row slice; A 64 rows (2 per thread), B 128 rows (4 per thread)
    const int lrow = tid >> 3, lk = (tid & 7) * 4;
    // two register sets: slices s + 1 and s + 2 are in flight while slice s is in LDS
    float4 pa[2][2], pb[2][4];
    auto load = [&](int set, int k0) {
#pragma unroll
        for (int j = 0; j < 2; ++j) pa[set][j] = mlp_raw4<VEC>(x, r0 + lrow + 32 * j, n, d, k0 + lk);
#pragma unroll
        for (int j = 0; j < 4; ++j) pb[set][j] = mlp_raw4<VEC>(w1, c0 + lrow + 32 * j, hidden, d, k0 + lk);
    };
    const bool sq = (c0 == 0);
    float sqs[2] = {0.0f, 0.0f};
    auto store = [&](int set, int buf, int k0) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const float4 v = mlp_mask4(pa[set][j], r0 + lrow + 32 * j, n, d, k0 + lk);
            *reinterpret_cast<float4*>(&sm.a[buf][(lrow + 32 * j) * kHS + lk]) = v;
            if (sq) {
                const float t0 = v.x - 0.5f, t1 = v.y - 0.5f, t2 = v.z - 0.5f, t3 = v.w - 0.5f;
                const int kk = k0 + lk;
                sqs[j] = (kk < d) ? fmaf(t0, t0, sqs[j]) : sqs[j];
                sqs[j] = (kk + 1 < d) ? fmaf(t1, t1, sqs[j]) : sqs[j];
                sqs[j] = (kk + 2 < d) ? fmaf(t2, t2, sqs[j]) : sqs[j];
                sqs[j] = (kk + 3 < d) ? fmaf(t3, t3, sqs[j]) : sqs[j];
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
            *reinterpret_cast<float4*>(&sm.b[buf][(lrow + 32 * j) * kHS + lk]) =
                mlp_mask4(pb[set][j], c0 + lrow + 32 * j, hidden, d, k0 + lk);
    };
    f32x4 acc[2][4];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[t][u] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    auto compute = [&](int buf) {
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            float4 a[2], b[4];
#pragma unroll
            for (int t = 0; t < 2; ++t)
                a[t] = *reinterpret_cast<const float4*>(&sm.a[buf][(wm * 32 + t * 16 + i) * kHS + half * 16 + 4 * g]);
#pragma unroll
            for (int u = 0; u < 4; ++u)
                b[u] = *reinterpret_cast<const float4*>(&sm.b[buf][(wn * 64 + u * 16 + i) * kHS + half * 16 + 4 * g]);
#pragma unroll
            for (int st = 0; st < 4; ++st)
#pragma unroll
                for (int t = 0; t < 2; ++t)
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[t][st], b[u][st], acc[t][u], 0, 0, 0);
        }
    };
    // Loads and stores are unconditional (a slice past d loads clamped
    // addresses and stores zeros), so every path through the loop has the
    // same loads in flight and the waits stay exact.
    load(0, 0);
    store(0, 0, 0);
    load(1, kHK);
    load(0, 2 * kHK);
    __syncthreads();
    // slice s sits in LDS buffer s % 2; slice s + 1 in register set (s + 1) % 2
    for (int k0 = 0;; k0 += 2 * kHK) {
        compute(0);                                   // slice k0
        store(1, 1, k0 + kHK);                        // slice k0 + 32
        load(1, k0 + 3 * kHK);
        __syncthreads();
        if (k0 + kHK >= d) break;
        compute(1);                                   // slice k0 + 32
        store(0, 0, k0 + 2 * kHK);                    // slice k0 + 64
        load(0, k0 + 4 * kHK);
        __syncthreads();
        if (k0 + 2 * kHK >= d) break;
    }
    if (sq) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            float v = sqs[j];
            v = v + __shfl_xor(v, 1);
            v = v + __shfl_xor(v, 2);
            v = v + __shfl_xor(v, 4);
            const int row = r0 + lrow + 32 * j;
            if ((tid & 7) == 0 && row < n) rowsq[row] = v;
        }
    }
    // D of a 16 x 16 MFMA tile: lane (g, i) holds rows 4g .. 4g+3 of column i
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = r0 + wm * 32 + t * 16 + 4 * g + r;
                const int col = c0 + wn * 64 + u * 16 + i;
                if (row < n && col < hidden) h[static_cast<size_t>(row) * hidden + col] = tanhf(acc[t][u][r]);
            }
}

__global__ __launch_bounds__(256) void mlp_hidden_kernel(const float* __restrict__ x, int n, int d,
                                                         const float* __restrict__ w1, int hidden,
                                                         float* __restrict__ h, float* __restrict__ rowsq)
{
    __shared__ __align__(16) MlpSmem sm;
    int rt, ct;
    if (!mlp_tile_of(static_cast<int>(blockIdx.x), (n + kHM - 1) / kHM, (hidden + kHN - 1) / kHN, rt, ct)) return;
    if ((d & 3) == 0) mlp_hidden_tile<true>(x, n, d, w1, hidden, h, rt * kHM, ct * kHN, sm, rowsq);
    else mlp_hidden_tile<false>(x, n, d, w1, hidden, h, rt * kHM, ct * kHN, sm, rowsq);
}

// Output head: y = W2 h, f_c = y_c^2 + 0.01 * ||x - 0.5||^2; one wave per row.
// ||x - 0.5||^2 comes from the hidden-layer launch, parked in f[0][row]
// (read before any objective of the row is written).
template <int M>
__device__ __forceinline__ void mlp_head_row(const float* __restrict__ h, int hidden, const float* __restrict__ w2,
                                             float* __restrict__ f, int ldf, float* __restrict__ cv, int row,
                                             int lane)
{
    const float s = f[row];
    float y[M];
#pragma unroll
    for (int k = 0; k < M; ++k) y[k] = 0.0f;
    for (int j = lane; j < hidden; j += 64) {
        const float hv = h[static_cast<size_t>(row) * hidden + j];
#pragma unroll
        for (int k = 0; k < M; ++k) y[k] = fmaf(w2[static_cast<size_t>(k) * hidden + j], hv, y[k]);
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
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
__global__ __launch_bounds__(256) void mlp_head_kernel(int n, const float* __restrict__ h, int hidden,
                                                       const float* __restrict__ w2, float* __restrict__ f, int ldf,
                                                       float* __restrict__ cv)
{
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= n) return;
    mlp_head_row<M>(h, hidden, w2, f, ldf, cv, row, static_cast<int>(threadIdx.x & 63));
}

// Batched MLP: hidden-layer tiles, then the head rows of the instances with M objectives.
__global__ __launch_bounds__(256) void b_mlp_hidden_kernel(const BInst* __restrict__ insts, BMap mp, int which)
{
    int local;
    const int i = bmap_find(mp, static_cast<int>(blockIdx.x), local);
    const BInst& I = insts[i];
    const int slot = (I.cur + which) % 3;
    __shared__ __align__(16) MlpSmem sm;
    int rt, ct;
    if (!mlp_tile_of(local, (I.n + kHM - 1) / kHM, (I.hidden + kHN - 1) / kHN, rt, ct)) return;
    if ((I.d & 3) == 0)
        mlp_hidden_tile<true>(I.x[slot], I.n, I.d, I.w1, I.hidden, I.hwork, rt * kHM, ct * kHN, sm, I.f[slot]);
    else
        mlp_hidden_tile<false>(I.x[slot], I.n, I.d, I.w1, I.hidden, I.hwork, rt * kHM, ct * kHN, sm, I.f[slot]);
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
    mlp_head_row<M>(I.hwork, I.hidden, I.w2, I.f[slot], I.n, I.cv[slot], row,
                    static_cast<int>(threadIdx.x & 63));
}

// The DTLZ suite (Deb, Thiele, Laumanns, Zitzler 2005) with the reference's
// protections (problem/dtlz/dtlz_kernels.cu:20-22, 28-717).  KIND is the
// problem number; 8 is ConvexDTLZ2.  g runs over the last k = d - m + 1
// variables:
//   DTLZ1, 3: 100 (k + sum((x - 1/2)^2 - cos(20 pi (x - 1/2))))
//   DTLZ2, 4, 5, ConvexDTLZ2: sum (x - 1/2)^2;  DTLZ6: sum x^0.1
//   DTLZ7: 1 + 9/k sum x
// DTLZ1 clamps x_j to [0, 1] in its products and 1 - x to >= 0; the
// spherical kinds use max(0, cos); DTLZ4 maps x to x^100; DTLZ5/6 map
// x_j, j >= 1, to theta = (1 + 2 g x_j) / (2 (1 + g)), clamped to [0, 1] in
// DTLZ5; ConvexDTLZ2 raises the first m - 1 objectives to the 4th power and
// the last to the 2nd.  The angle arguments are formed in double and rounded
// to float, as the reference's M_PI expressions are.  One wave per row;
// lanes own the objectives.
template <int KIND>
__device__ __forceinline__ float dtlz_angle(const float* __restrict__ xr, int j, float g)
{
    float t = xr[j];
    if constexpr (KIND == 4) t = powf(t, 100.0f);
    if constexpr (KIND == 5 || KIND == 6) {
        if (j > 0) {
            t = (1.0f + 2.0f * g * t) / (2.0f * (1.0f + g));
            if constexpr (KIND == 5) t = fminf(fmaxf(t, 0.0f), 1.0f);
        }
    }
    return static_cast<float>(static_cast<double>(t) * M_PI / 2.0);
}

// Constraint violation of the constrained variants from the row's objectives
// fv (compute_cv, dtlz_kernels.cu:824-1044, 1097-1165).  CKIND 1: C1-DTLZ1,
// 2: C1-DTLZ3, 3: C2-DTLZ2, 4: C2-ConvexDTLZ2, 5: C3-DTLZ1, 6: C3-DTLZ4;
// R is the radius the reference picks for the objective count.
template <int M, int CKIND>
__device__ __forceinline__ float dtlz_cv(const float (&fv)[M], float R)
{
    if constexpr (CKIND == 1) {
        float c = fv[M - 1] / 0.6f;
        for (int i = 0; i < M - 1; ++i) c += fv[i] / 0.5f;
        return fmaxf(c - 1.0f, 0.0f);
    } else if constexpr (CKIND == 2) {
        float s2 = 0.0f;
        for (int i = 0; i < M; ++i) s2 += fv[i] * fv[i];
        return fmaxf((s2 - 16.0f) * (R * R - s2), 0.0f);
    } else if constexpr (CKIND == 3) {
        const float inv_sqrt_m = 1.0f / sqrtf(static_cast<float>(M));
        const float inv_r2 = 1.0f / (R * R);
        float s2 = 0.0f;
        for (int i = 0; i < M; ++i) s2 = fmaf(fv[i], fv[i], s2);
        float lo = INFINITY;
        for (int i = 0; i < M; ++i) {
            const float diff = fv[i] - 1.0f;
            lo = fminf(lo, fmaf(diff, diff, s2 - fv[i] * fv[i]) * inv_r2 - 1.0f);
        }
        float c2 = 0.0f;
        for (int i = 0; i < M; ++i) {
            const float df = fv[i] - inv_sqrt_m;
            c2 = fmaf(df, df, c2);
        }
        return fmaxf(fminf(lo, c2 * inv_r2 - 1.0f), 0.0f);
    } else if constexpr (CKIND == 4) {
        float sum = 0.0f, sum_sq = 0.0f;
        for (int i = 0; i < M; ++i) {
            sum += fv[i];
            sum_sq += fv[i] * fv[i];
        }
        const float mean = sum * (1.0f / static_cast<float>(M));
        const float variance = sum_sq - static_cast<float>(M) * mean * mean;
        return fmaxf(1.0f - variance * (1.0f / (R * R)), 0.0f);
    } else if constexpr (CKIND == 5) {
        float s1 = 0.0f;
        for (int i = 0; i < M; ++i) s1 += fv[i];
        s1 *= 2.0f;
        float c = 0.0f;
        for (int i = 0; i < M; ++i) c += fmaxf(fv[i] + 1.0f - s1, 0.0f);
        return c;
    } else {
        float s2 = 0.0f;
        for (int i = 0; i < M; ++i) s2 += fv[i] * fv[i];
        float c = 0.0f;
        for (int i = 0; i < M; ++i) c += fmaxf(0.75f * (fv[i] * fv[i]) - s2 + 1.0f, 0.0f);
        return c;
    }
}

template <int M, int KIND, int CKIND = 0>
__device__ __forceinline__ void dtlz_rows(const float* __restrict__ x, int n, int d, float* __restrict__ f, int ldf,
                                 float* __restrict__ cv, float R = 0.0f)
{
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const float* xr = x + static_cast<size_t>(row) * d;
    float s = 0.0f;
    for (int j = M - 1 + lane; j < d; j += 64) {
        if constexpr (KIND == 6) {
            s = s + powf(xr[j], 0.1f);
        } else if constexpr (KIND == 7) {
            s = s + xr[j];
        } else {
            const float v = xr[j] - 0.5f;
            if constexpr (KIND == 1 || KIND == 3) s = s + (v * v - cosf_det(static_cast<float>(20.0 * M_PI * static_cast<double>(v))));
            else s = s + v * v;
        }
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) s = s + __shfl_xor(s, o);
    float out = 0.0f;
    if (lane < M) {
        const int k = d - (M - 1);
        if constexpr (KIND == 7) {
            const float g = 1.0f + (9.0f / static_cast<float>(k)) * s;
            if (lane < M - 1) {
                out = xr[lane];
            } else {
                float h = 0.0f;
                for (int i = 0; i < M - 1; ++i) {
                    const float fi = xr[i];
                    h += (fi / (1.0f + g)) * (1.0f + sinf(static_cast<float>(3.0 * M_PI * static_cast<double>(fi))));
                }
                out = (1.0f + g) * (static_cast<float>(M) - h);
            }
        } else if constexpr (KIND == 1) {
            out = 0.5f * (1.0f + 100.0f * (static_cast<float>(k) + s));
            for (int j = 0; j < M - 1 - lane; ++j) out *= fminf(fmaxf(xr[j], 0.0f), 1.0f);
            if (lane > 0) out *= fmaxf(0.0f, 1.0f - xr[M - 1 - lane]);
        } else {
            const float g = (KIND == 3) ? 100.0f * (static_cast<float>(k) + s) : s;
            out = 1.0f + g;
            for (int j = 0; j < M - 1 - lane; ++j) out *= fmaxf(0.0f, cosf(dtlz_angle<KIND>(xr, j, g)));
            if (lane > 0) out *= sinf(dtlz_angle<KIND>(xr, M - 1 - lane, g));
            if constexpr (KIND == 8) {
                out = out * out;
                if (lane < M - 1) out = out * out;
            }
        }
        f[static_cast<size_t>(lane) * ldf + row] = out;
    }
    if constexpr (CKIND != 0) {
        float fv[M];
#pragma unroll
        for (int i = 0; i < M; ++i) fv[i] = __shfl(out, i);
        if (lane == 0) cv[row] = dtlz_cv<M, CKIND>(fv, R);
    } else {
        if (lane == 0 && cv) cv[row] = 0.0f;
    }
}

template <int M> __global__ __launch_bounds__(256) void dtlz1_kernel(const float* __restrict__ x, int n, int d, float* __restrict__ f, int ldf, float* __restrict__ cv) { dtlz_rows<M, 1>(x, n, d, f, ldf, cv); }
template <int M> __global__ __launch_bounds__(256) void dtlz2_kernel(const float* __restrict__ x, int n, int d, float* __restrict__ f, int ldf, float* __restrict__ cv) { dtlz_rows<M, 2>(x, n, d, f, ldf, cv); }
template <int M> __global__ __launch_bounds__(256) void dtlz3_kernel(const float* __restrict__ x, int n, int d, float* __restrict__ f, int ldf, float* __restrict__ cv) { dtlz_rows<M, 3>(x, n, d, f, ldf, cv); }
template <int M> __global__ __launch_bounds__(256) void dtlz4_kernel(const float* __restrict__ x, int n, int d, float* __restrict__ f, int ldf, float* __restrict__ cv) { dtlz_rows<M, 4>(x, n, d, f, ldf, cv); }
template <int M> __global__ __launch_bounds__(256) void dtlz5_kernel(const float* __restrict__ x, int n, int d, float* __restrict__ f, int ldf, float* __restrict__ cv) { dtlz_rows<M, 5>(x, n, d, f, ldf, cv); }
template <int M> __global__ __launch_bounds__(256) void dtlz6_kernel(const float* __restrict__ x, int n, int d, float* __restrict__ f, int ldf, float* __restrict__ cv) { dtlz_rows<M, 6>(x, n, d, f, ldf, cv); }
template <int M> __global__ __launch_bounds__(256) void dtlz7_kernel(const float* __restrict__ x, int n, int d, float* __restrict__ f, int ldf, float* __restrict__ cv) { dtlz_rows<M, 7>(x, n, d, f, ldf, cv); }
template <int M> __global__ __launch_bounds__(256) void convex_dtlz2_kernel(const float* __restrict__ x, int n, int d, float* __restrict__ f, int ldf, float* __restrict__ cv) { dtlz_rows<M, 8>(x, n, d, f, ldf, cv); }
// constrained variants: objectives of their base problem, then the violation
template <int M> __global__ __launch_bounds__(256) void c1_dtlz1_kernel(const float* __restrict__ x, int n, int d, float* __restrict__ f, int ldf, float* __restrict__ cv, float R) { dtlz_rows<M, 1, 1>(x, n, d, f, ldf, cv, R); }
template <int M> __global__ __launch_bounds__(256) void c1_dtlz3_kernel(const float* __restrict__ x, int n, int d, float* __restrict__ f, int ldf, float* __restrict__ cv, float R) { dtlz_rows<M, 3, 2>(x, n, d, f, ldf, cv, R); }
template <int M> __global__ __launch_bounds__(256) void c2_dtlz2_kernel(const float* __restrict__ x, int n, int d, float* __restrict__ f, int ldf, float* __restrict__ cv, float R) { dtlz_rows<M, 2, 3>(x, n, d, f, ldf, cv, R); }
template <int M> __global__ __launch_bounds__(256) void c2_convex_dtlz2_kernel(const float* __restrict__ x, int n, int d, float* __restrict__ f, int ldf, float* __restrict__ cv, float R) { dtlz_rows<M, 8, 4>(x, n, d, f, ldf, cv, R); }
template <int M> __global__ __launch_bounds__(256) void c3_dtlz1_kernel(const float* __restrict__ x, int n, int d, float* __restrict__ f, int ldf, float* __restrict__ cv, float R) { dtlz_rows<M, 1, 5>(x, n, d, f, ldf, cv, R); }
template <int M> __global__ __launch_bounds__(256) void c3_dtlz4_kernel(const float* __restrict__ x, int n, int d, float* __restrict__ f, int ldf, float* __restrict__ cv, float R) { dtlz_rows<M, 4, 6>(x, n, d, f, ldf, cv, R); }

// CSDP, the car side-impact design problem (csdp_func_kernel / csdp_cv_kernel,
// dtlz_kernels.cu:719-738, 1046-1085): 7 variables, 3 objectives and the sum of
// 10 constraint violations; one thread per row.  The objectives' constants are
// double, as the reference writes them, and rounded to float once.
__global__ __launch_bounds__(256) void csdp_kernel(const float* __restrict__ x, int n, float* __restrict__ f, int ldf,
                                                   float* __restrict__ cv)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* r = x + static_cast<size_t>(i) * 7;
    const float x1 = r[0], x2 = r[1], x3 = r[2], x4 = r[3], x5 = r[4], x6 = r[5], x7 = r[6];
    const float F = static_cast<float>(4.72 - 0.5 * x4 - 0.19 * x2 * x3);
    const float vmbp = static_cast<float>(10.58 - 0.674 * x1 * x2 - 0.67275 * x2);
    const float vfd = static_cast<float>(16.45 - 0.489 * x3 * x7 - 0.843 * x5 * x6);
    f[i] = static_cast<float>(1.98 + 4.9 * x1 + 6.67 * x2 + 6.98 * x3 + 4.01 * x4 + 1.78 * x5 + 0.00001 * x6 +
                              2.73 * x7);
    f[static_cast<size_t>(ldf) + i] = F;
    f[2 * static_cast<size_t>(ldf) + i] = 0.5f * (vmbp + vfd);
    const float Ff = 4.72f - 0.5f * x4 - 0.19f * x2 * x3;
    const float vmbpf = 10.58f - 0.674f * x1 * x2 - 0.67275f * x2;
    const float vfdf = 16.45f - 0.489f * x3 * x7 - 0.843f * x5 * x6;
    const float g[10] = {
        1.16f - 0.3717f * x2 * x4 - 0.0092928f * x3 - 1.0f,
        0.261f - 0.0159f * x1 * x2 - 0.06486f * x1 - 0.019f * x2 * x7 + 0.0144f * x3 * x5 + 0.0154464f * x6 - 0.32f,
        0.214f + 0.00817f * x5 - 0.045195f * x1 - 0.0135168f * x1 + 0.03099f * x2 * x6 - 0.018f * x2 * x7 +
            0.007176f * x3 + 0.023232f * x3 - 0.00364f * x5 * x6 - 0.018f * x2 * x2 - 0.32f,
        0.74f - 0.61f * x2 - 0.031296f * x3 - 0.031872f * x7 + 0.227f * x2 * x2 - 0.32f,
        28.98f + 3.818f * x3 - 4.2f * x1 * x2 + 1.27296f * x6 - 2.68065f * x7 - 32.0f,
        33.86f + 2.95f * x3 - 5.057f * x1 * x2 - 3.795f * x2 - 3.4431f * x7 + 1.45728f - 32.0f,
        46.36f - 9.9f * x2 - 4.4505f * x1 - 32.0f,
        Ff - 4.0f,
        vmbpf - 9.9f,
        vfdf - 15.7f,
    };
    float total = 0.0f;
#pragma unroll
    for (int j = 0; j < 10; ++j) total += fmaxf(g[j], 0.0f);
    cv[i] = total;
}

// The radius of C1-DTLZ3, C2-DTLZ2 and C2-ConvexDTLZ2 by objective count
// (compute_cv, dtlz_kernels.cu:1110-1158); 0 for a count the reference rejects.
float dtlz_radius(int kind, int m)
{
    if (kind == 10) {
        if (m == 2) return 6.0f;
        if (m == 3) return 5.0f;
        if (m == 5) return 12.5f;
        return (m == 8 || m == 10 || m == 15) ? 15.0f : 0.0f;
    }
    if (kind == 11) {
        if (m == 2) return 0.15f;
        if (m == 3) return 0.40f;
        return (m == 5 || m == 8 || m == 10 || m == 15) ? 1.0f : 0.0f;
    }
    if (kind == 12) {
        if (m == 3) return 0.20f;
        if (m == 5) return 0.225f;
        if (m == 8 || m == 10) return 0.26f;
        return m == 15 ? 0.27f : 0.0f;
    }
    return 1.0f;
}

}  // namespace

int eval_dtlz(int kind, const float* x, int n, int d, int m, float* f, int ldf, float* cv, hipStream_t stream)
{
    if (kind == 15) {
        if (!x || !f || !cv || n <= 0 || ldf < n) return fail(kInvalidArgument, "eval_dtlz: invalid arguments");
        if (d != 7 || m != 3) return fail(kInvalidArgument, "CSDP needs 7 variables and 3 objectives");
        hipLaunchKernelGGL(csdp_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, stream, x, n, f, ldf, cv);
        MOEA_HIP(hipGetLastError());
        return kOk;
    }
    if (!x || !f || n <= 0 || m <= 1 || m > kMaxM || d < m || ldf < n || kind < 1 || kind > 14)
        return fail(kInvalidArgument, "eval_dtlz: invalid arguments (kinds 1-15, 2 <= m <= d)");
    if (kind > 8 && !cv) return fail(kInvalidArgument, "eval_dtlz: the constrained variants need a cv array");
    const float R = dtlz_radius(kind, m);
    if (R == 0.0f) return fail(kInvalidArgument, "eval_dtlz: objective count not defined for this constrained variant");
    const dim3 grid(ceil_div(n, 4)), block(256);
    switch (kind) {
    case 9: MOEA_LAUNCH_M(m, c1_dtlz1_kernel, grid, block, 0, stream, x, n, d, f, ldf, cv, R); break;
    case 10: MOEA_LAUNCH_M(m, c1_dtlz3_kernel, grid, block, 0, stream, x, n, d, f, ldf, cv, R); break;
    case 11: MOEA_LAUNCH_M(m, c2_dtlz2_kernel, grid, block, 0, stream, x, n, d, f, ldf, cv, R); break;
    case 12: MOEA_LAUNCH_M(m, c2_convex_dtlz2_kernel, grid, block, 0, stream, x, n, d, f, ldf, cv, R); break;
    case 13: MOEA_LAUNCH_M(m, c3_dtlz1_kernel, grid, block, 0, stream, x, n, d, f, ldf, cv, R); break;
    case 14: MOEA_LAUNCH_M(m, c3_dtlz4_kernel, grid, block, 0, stream, x, n, d, f, ldf, cv, R); break;
    case 1: MOEA_LAUNCH_M(m, dtlz1_kernel, grid, block, 0, stream, x, n, d, f, ldf, cv); break;
    case 2: MOEA_LAUNCH_M(m, dtlz2_kernel, grid, block, 0, stream, x, n, d, f, ldf, cv); break;
    case 3: MOEA_LAUNCH_M(m, dtlz3_kernel, grid, block, 0, stream, x, n, d, f, ldf, cv); break;
    case 4: MOEA_LAUNCH_M(m, dtlz4_kernel, grid, block, 0, stream, x, n, d, f, ldf, cv); break;
    case 5: MOEA_LAUNCH_M(m, dtlz5_kernel, grid, block, 0, stream, x, n, d, f, ldf, cv); break;
    case 6: MOEA_LAUNCH_M(m, dtlz6_kernel, grid, block, 0, stream, x, n, d, f, ldf, cv); break;
    case 7: MOEA_LAUNCH_M(m, dtlz7_kernel, grid, block, 0, stream, x, n, d, f, ldf, cv); break;
    default: MOEA_LAUNCH_M(m, convex_dtlz2_kernel, grid, block, 0, stream, x, n, d, f, ldf, cv); break;
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
    hipLaunchKernelGGL(mlp_hidden_kernel, dim3(mlp_tile_count(n, hidden)), dim3(256), 0, stream, x, n, d, w1,
                       hidden, hwork, f);
    MOEA_HIP(hipGetLastError());
    const dim3 grid(ceil_div(n, 4)), block(256);
    switch (m) {
#define MOEA_H(MM) case MM: hipLaunchKernelGGL(mlp_head_kernel<MM>, grid, block, 0, stream, n, hwork, hidden, w2, f, ldf, cv); break;
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
