// NSGA-III environmental selection on CDNA4.
//
// Replaces NSGA3EnvironmentSelector::select
// (NS/selection/nsga3/nsga3_environment_selector.cu:289-538).  Every size that
// the reference reads back to the host (pair count, front sizes, SVD status,
// N_prior / N_last) stays on the device: kernels are launched for the largest
// possible extent and read the live sizes from a small metadata array, so one
// generation is a fixed launch sequence that a HIP graph can capture.
//
// Merged index space: s in [0, n) is parent s, s in [n, 2n) is offspring s-n.
// Variables are never merged; the survivor gather reads them in place.
#include <cmath>

#include "batch_internal.hpp"
#include "common.hpp"
#include "moea_internal.hpp"

namespace moea {

namespace {

// meta layout (int32, device)
enum Meta : int { kMaxFront = 0, kNPrior = 1, kNLast = 2, kIll = 3, kNRem = 4, kDevError = 6, kMetaLen = 8 };

constexpr int kScanThreads = 1024;

// Exclusive scan of one int per thread across a 1024-thread block.
__device__ __forceinline__ int block_exclusive_scan(int v, int* s_wave, int& total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int incl = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(incl, o);
        if (lane >= o) incl += t;
    }
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    if (wave == 0) {
        const int nw = blockDim.x >> 6;
        int w = lane < nw ? s_wave[lane] : 0;
        int wi = w;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_up(wi, o);
            if (lane >= o) wi += t;
        }
        if (lane < nw) s_wave[lane] = wi - w;
        if (lane == nw - 1) s_wave[63] = wi;
    }
    __syncthreads();
    const int res = s_wave[wave] + incl - v;
    total = s_wave[63];
    __syncthreads();
    return res;
}

// f (m, 2n) and cv (2n) of the merged population (nsga3_population_update.cu:7-115
// copies variables too; here only the O(M n) objective columns move).
// omin (optional, 256-thread blocks): per block and objective, the fminf of
// its offspring rows, for the running ideal point (norm_prep finishes it).
// keys / vals (optional): the rows' sort keys and indices, for the sort of
// the merged rows that follows.
__global__ __launch_bounds__(256) void merge_objectives(const float* __restrict__ pf, const float* __restrict__ of,
                                                        const float* __restrict__ pcv, const float* __restrict__ ocv,
                                                        int n, int m, float* __restrict__ fm, float* __restrict__ cvm,
                                                        float* __restrict__ omin, uint64_t* __restrict__ keys,
                                                        int32_t* __restrict__ vals)
{
    __shared__ float s_min[4][kMaxM];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nm = 2 * n;
    const bool valid = i < nm;
    const bool par = i < n;
    const int s = par ? i : i - n;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    float f0 = 0.0f;
    bool any_nan = false;
    for (int k = 0; k < m; ++k) {
        float v = INFINITY;
        if (valid) {
            v = par ? pf[static_cast<size_t>(k) * n + s] : of[static_cast<size_t>(k) * n + s];
            fm[static_cast<size_t>(k) * nm + i] = v;
            if (k == 0) f0 = v;
            any_nan = any_nan || isnan(v);
        }
        float z = (valid && !par) ? v : INFINITY;
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) z = fminf(z, __shfl_xor(z, o));
        if (lane == 0) s_min[wave][k] = z;
    }
    if (valid) {
        const float c = par ? pcv[s] : ocv[s];
        cvm[i] = c;
        if (keys) {
            keys[i] = nds_key_from(f0, any_nan, c);
            vals[i] = i;
        }
    }
    if (!omin) return;
    __syncthreads();
    if (static_cast<int>(threadIdx.x) < m) {
        const int k = threadIdx.x;
        omin[blockIdx.x * m + k] = fminf(fminf(s_min[0][k], s_min[1][k]), fminf(s_min[2][k], s_min[3][k]));
    }
}

// The caller's stats words: the selection's meta words, except that the
// device-error word only ever turns on (it stays set over later selections
// until the caller clears it).
__device__ __forceinline__ void copy_stats(const int32_t* __restrict__ meta, int32_t* __restrict__ stats)
{
    for (int j = 0; j < kMetaLen; ++j)
        if (j != kDevError) stats[j] = meta[j];
    if (meta[kDevError] != 0) stats[kDevError] = 1;
}

// Merged row s (front fr) is a prior-front survivor: fr < max_front.  A
// negative max_front means the sort failed on the device (its peel gave up
// waiting); then the n parents survive unchanged, so every later launch still
// reads valid indices, and the selection's device-error word is set.
__device__ __forceinline__ bool prior_survivor(int fr, int mf, int s, int n)
{
    return mf < 0 ? (s < n) : (fr >= 0 && fr < mf);
}

// prior = {front < max_front}, last = {front == max_front}, each ascending
// (nds_frontmask_kernel + copy_if, non_dominated_sort.cu:1580-1629);
// gidx = [prior; last].  Two passes over blocks of kScanThreads entries:
// per-block counts, then each block adds the counts of the blocks before it,
// scans its own flags and scatters.
__global__ __launch_bounds__(kScanThreads) void front_counts(const int32_t* __restrict__ fronts, int nm,
                                                             const int32_t* __restrict__ meta,
                                                             int32_t* __restrict__ counts)
{
    __shared__ int s_wave[64];
    const int mf = meta[kMaxFront];
    const int i = blockIdx.x * kScanThreads + threadIdx.x;
    const int fr = i < nm ? fronts[i] : -1;
    int tp, tl;
    block_exclusive_scan((fr >= 0 && fr < mf) ? 1 : 0, s_wave, tp);
    block_exclusive_scan((fr == mf) ? 1 : 0, s_wave, tl);
    if (threadIdx.x == 0) {
        counts[2 * blockIdx.x] = tp;
        counts[2 * blockIdx.x + 1] = tl;
    }
}

// It also writes the prior-front survivor flags (gather_selected_kernel's
// prior copy, niching.cu:303-312) that the niching completes.
__global__ __launch_bounds__(kScanThreads) void front_scatter(const int32_t* __restrict__ fronts, int nm,
                                                              int n, const int32_t* __restrict__ counts,
                                                              int nblk, int32_t* __restrict__ gidx,
                                                              int32_t* __restrict__ meta,
                                                              uint8_t* __restrict__ flag)
{
    __shared__ int s_wave[64];
    const int mf = meta[kMaxFront];
    const int b = blockIdx.x;
    int bp = 0, bl = 0, ap = 0, al = 0;   // before this block, all blocks
    for (int j = threadIdx.x; j < nblk; j += kScanThreads) {
        const int cp = counts[2 * j], cl = counts[2 * j + 1];
        ap += cp;
        al += cl;
        if (j < b) { bp += cp; bl += cl; }
    }
    int t;
    block_exclusive_scan(bp, s_wave, bp);
    block_exclusive_scan(bl, s_wave, bl);
    block_exclusive_scan(ap, s_wave, ap);
    block_exclusive_scan(al, s_wave, al);
    const int i = b * kScanThreads + threadIdx.x;
    const int fr = i < nm ? fronts[i] : -1;
    if (i < nm) flag[i] = prior_survivor(fr, mf, i, n) ? 1 : 0;
    if (mf < 0) {
        // the sort failed on the device (its peel gave up waiting): keep the
        // parents, so every later launch reads valid indices, and flag it
        if (i < n) gidx[i] = i;
        if (b == 0 && threadIdx.x == 0) {
            meta[kNPrior] = n;
            meta[kNLast] = 0;
            meta[kNRem] = 0;
            meta[kIll] = -1;
            meta[kDevError] = 1;
        }
        return;
    }
    const int fp = (fr >= 0 && fr < mf) ? 1 : 0, fl = (fr == mf) ? 1 : 0;
    const int op = block_exclusive_scan(fp, s_wave, t);
    const int ol = block_exclusive_scan(fl, s_wave, t);
    if (fp) gidx[bp + op] = i;
    if (fl) gidx[ap + bl + ol] = i;
    if (b == 0 && threadIdx.x == 0) {
        meta[kNPrior] = ap;
        meta[kNLast] = al;
        meta[kNRem] = n - ap;
        meta[kIll] = -1;
        meta[kDevError] = 0;
    }
}

// Running ideal point: ideal = fminf(ideal, min over offspring) (ideal_point.cu:25-165).
__device__ __forceinline__ void ideal_body(const float* __restrict__ f, int ldf, int col0, int count, int k,
                                  float* __restrict__ ideal, int init)
{
    __shared__ float s_min[1024 / 64];
    float z = INFINITY;
    for (int i = threadIdx.x; i < count; i += blockDim.x)
        z = fminf(z, f[static_cast<size_t>(k) * ldf + col0 + i]);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) z = fminf(z, __shfl_xor(z, o));
    if ((threadIdx.x & 63) == 0) s_min[threadIdx.x >> 6] = z;
    __syncthreads();
    if (threadIdx.x == 0) {
        float r = INFINITY;
        for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) r = fminf(r, s_min[w]);
        ideal[k] = init ? r : fminf(ideal[k], r);
    }
}

__global__ void ideal_update(const float* __restrict__ f, int ldf, int col0, int count, int m,
                             float* __restrict__ ideal, int init)
{
    ideal_body(f, ldf, col0, count, static_cast<int>(blockIdx.x), ideal, init);
}

// Translate, clamp and reduce the extreme-point / row-max candidates per block
// (translated_ndsfit_kernel, compute_asf_legacy_kernel, find_intercepts_legacy_kernel:
// normalization.cu:370-444, 647-729, 802-856).
template <int M>
__device__ __forceinline__ void norm_prep_body(const float* __restrict__ fm, int nm, const int32_t* __restrict__ gidx,
                                      const float* __restrict__ ideal, const int32_t* __restrict__ meta,
                                      float* __restrict__ fn, float* __restrict__ asf_val,
                                      int32_t* __restrict__ asf_idx, float* __restrict__ row_max, int bx)
{
    __shared__ float s_val[4][M];
    __shared__ int s_idx[4][M];
    __shared__ float s_max[4][M];
    const int n_nds = meta[kNPrior] + meta[kNLast];
    if (meta[kNRem] <= 0) return;
    if (bx * 256 >= n_nds) return;
    const int i = bx * 256 + threadIdx.x;
    const bool valid = i < n_nds;
    float v[M];
    if (valid) {
        const int s = gidx[i];
#pragma unroll
        for (int k = 0; k < M; ++k) {
            float t = fm[static_cast<size_t>(k) * nm + s] - ideal[k];
            if (isfinite(t)) t = fmaxf(t, 0.0f);
            v[k] = t;
            fn[static_cast<size_t>(k) * nm + i] = t;
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < M; ++j) {
        float best = INFINITY;
        int bi = -1;
        if (valid) {
            float mx = -INFINITY;
#pragma unroll
            for (int k = 0; k < M; ++k) {
                const float t = (v[k] != v[k]) ? INFINITY : v[k];
                mx = fmaxf(mx, t * ((k == j) ? 1.0f : 1e6f));
            }
            if (mx < best) { best = mx; bi = i; }
        }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            const float ov = __shfl_xor(best, o);
            const int oi = __shfl_xor(bi, o);
            if (ov < best || (ov == best && oi < bi)) { best = ov; bi = oi; }
        }
        float rm = -INFINITY;
        if (valid) rm = (v[j] != v[j]) ? -INFINITY : v[j];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) rm = fmaxf(rm, __shfl_xor(rm, o));
        if (lane == 0) { s_val[wave][j] = best; s_idx[wave][j] = bi; s_max[wave][j] = rm; }
    }
    __syncthreads();
    if (threadIdx.x < M) {
        const int j = threadIdx.x;
        float best = INFINITY;
        int bi = -1;
        float rm = -INFINITY;
        for (int w = 0; w < 4; ++w) {
            const float ov = s_val[w][j];
            const int oi = s_idx[w][j];
            if (ov < best || (ov == best && oi < bi)) { best = ov; bi = oi; }
            rm = fmaxf(rm, s_max[w][j]);
        }
        asf_val[bx * M + j] = best;
        asf_idx[bx * M + j] = bi;
        row_max[bx * M + j] = rm;
    }
}

template <int M>
__global__ __launch_bounds__(256) void norm_prep(const float* __restrict__ fm, int nm,
                                                 const int32_t* __restrict__ gidx,
                                                 const int32_t* __restrict__ meta,
                                                 float* __restrict__ fn, float* __restrict__ asf_val,
                                                 int32_t* __restrict__ asf_idx,
                                                 float* __restrict__ row_max, int32_t* __restrict__ rho, int k,
                                                 const float* __restrict__ omin, int nbo, float* ideal_io)
{
    // the niche counts start at zero (the association kernels add the prior front)
    for (int j = static_cast<int>(blockIdx.x * 256 + threadIdx.x); j < k; j += static_cast<int>(gridDim.x * 256))
        rho[j] = 0;
    // running ideal: fminf(ideal, min over the offspring) from merge_objectives'
    // per-block minima, reduced in one fixed order by every block; block 0
    // stores it (a block reading the stored value gets the same result, since
    // fminf(fminf(a, r), r) = fminf(a, r))
    __shared__ float s_ideal[M];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int j = wave; j < M; j += 4) {
        float z = INFINITY;
        for (int b = lane; b < nbo; b += 64) z = fminf(z, omin[b * M + j]);
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) z = fminf(z, __shfl_xor(z, o));
        if (lane == 0) s_ideal[j] = fminf(ideal_io[j], z);
    }
    __syncthreads();
    if (blockIdx.x == 0 && threadIdx.x < M) ideal_io[threadIdx.x] = s_ideal[threadIdx.x];
    norm_prep_body<M>(fm, nm, gidx, s_ideal, meta, fn, asf_val, asf_idx, row_max, static_cast<int>(blockIdx.x));
}

// Extreme points, FP64 hyperplane or fallback, sanitize (normalization.cu:453-540,
// 572-583, 1593-1636).  256 threads reduce the per-block candidates, then one
// thread solves.  Extreme point j is the (value, index)-smallest candidate with
// a finite ASF value (-1 if none) -- an order-independent reduction.
constexpr int kSolveThreads = 256;
// Lane i's value of v, uniform over the wave (two scalar lane reads).
__device__ __forceinline__ double readlane_f64(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), lane);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), lane);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

// LDS of the one-wave hyperplane solve: A and V row-major (kMaxM x kMaxM),
// three sum rows, and the column weights.
struct SolveLds {
    double a[kMaxM * kMaxM];
    double v[kMaxM * kMaxM];
    double sums[3][16];
    double w[kMaxM];
    int skip[kMaxM];
};

// hyperplane_fp64 (common.hpp) on one wave, bit-identical to it: the same
// rotation sequence, every element update the same expression, every sum
// accumulated in the same order.  Lane i < m updates row i of A and V (in
// LDS); each rotation's three column sums are taken by lanes 0, 1, 2, each
// summing its row of products in order, and broadcast by lane reads; the
// rotation parameters are then computed redundantly by every lane.  The
// singular values and weights are computed by lane j for column j.  The code
// is a compact loop: a fully unrolled 10 x 10 sweep does not fit the
// instruction cache once other kernels' code shares it, and then runs several
// times slower.  Returns 1 for the row-max fallback (the scalar solve's
// conditions); lane r < m gets a_out[r] in `a_lane`.
__device__ __forceinline__ int hyperplane_fp64_wave(int m, SolveLds& L, double& a_lane)
{
    const int lane = static_cast<int>(threadIdx.x & 63);
    const bool own = lane < m;
    double* ra = L.a + lane * kMaxM;
    double* rv = L.v + lane * kMaxM;
    bool bad = false;
    if (own)
        for (int j = 0; j < m; ++j) {
            bad = bad || !isfinite(ra[j]);
            rv[j] = (lane == j) ? 1.0 : 0.0;
        }
    if (__ballot(bad)) return 1;
    // Two equal rows (the same individual extreme on two axes, as a rule at
    // M >= 8) make the result the fallback whatever the sweeps do: rotations
    // act on rows alike, so the rows stay equal, the columns lie in an
    // (m-1)-dimensional subspace, and m of them cannot be orthogonal unless one
    // is (numerically) zero -- the sweeps then either never meet the 1e-15
    // test (60 sweeps, fallback) or end with sigma_min / sigma_max far below
    // 1e-6 (fallback).  Taking the fallback at once saves the 60 sweeps.
    bool dup = false;
    if (own)
        for (int i = 0; i < lane; ++i) {
            bool eq = true;
            for (int j = 0; j < m; ++j) eq = eq && (L.a[i * kMaxM + j] == ra[j]);
            dup = dup || eq;
        }
    if (__ballot(dup)) return 1;
    int converged = 0;
    for (int sweep = 0; sweep < 60 && !converged; ++sweep) {
        int rotated = 0;
        for (int p = 0; p < m - 1; ++p) {
            for (int q = p + 1; q < m; ++q) {
                const double ap = own ? ra[p] : 0.0, aq = own ? ra[q] : 0.0;
                if (own) {
                    L.sums[0][lane] = ap * ap;
                    L.sums[1][lane] = aq * aq;
                    L.sums[2][lane] = ap * aq;
                }
                __builtin_amdgcn_wave_barrier();
                double acc = 0.0;
                if (lane < 3) {
                    double t[kMaxM];
#pragma unroll
                    for (int i = 0; i < kMaxM; ++i) t[i] = L.sums[lane][i];
#pragma unroll
                    for (int i = 0; i < kMaxM; ++i)
                        if (i < m) acc = acc + t[i];
                }
                const double alpha = readlane_f64(acc, 0), beta = readlane_f64(acc, 1),
                             gamma = readlane_f64(acc, 2);
                __builtin_amdgcn_wave_barrier();
                if (gamma == 0.0 || fabs(gamma) <= 1e-15 * sqrt(alpha * beta)) continue;
                rotated = 1;
                const double zeta = (beta - alpha) / (2.0 * gamma);
                const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / sqrt(1.0 + t * t);
                const double sn = c * t;
                if (own) {
                    ra[p] = c * ap - sn * aq;
                    ra[q] = sn * ap + c * aq;
                    const double vp = rv[p], vq = rv[q];
                    rv[p] = c * vp - sn * vq;
                    rv[q] = sn * vp + c * vq;
                }
            }
        }
        if (!rotated) converged = 1;
    }
    if (!converged) return 1;
    // lane j: singular value of column j, then its weight
    double sig = 0.0;
    if (own) {
        double ss = 0.0;
        for (int i = 0; i < m; ++i) {
            const double e = L.a[i * kMaxM + lane];
            ss = ss + e * e;
        }
        sig = sqrt(ss);
    }
    const bool finite_bad = own && !isfinite(sig);
    if (__ballot(finite_bad)) return 1;
    double smax = 0.0, smin = static_cast<double>(INFINITY);
    for (int j = 0; j < m; ++j) {
        const double sj = readlane_f64(sig, j);
        if (sj > smax) smax = sj;
        if (sj < smin) smin = sj;
    }
    if (smin < 1e-10) return 1;
    if (smax / smin > 1e6) return 1;
    if (own) {
        const bool use = sig > 1e-5;
        double w = 0.0;
        if (use) {
            double usum = 0.0;
            for (int i = 0; i < m; ++i) usum = usum + L.a[i * kMaxM + lane] / sig;
            w = usum / sig;
        }
        L.w[lane] = w;
        L.skip[lane] = use ? 0 : 1;
    }
    __builtin_amdgcn_wave_barrier();
    double out = 0.0;
    if (own)
        for (int j = 0; j < m; ++j)
            if (!L.skip[j]) out = out + rv[j] * L.w[j];
    a_lane = out;
    return 0;
}

// Wave 0 of the solve block: extreme-point matrix rows, the solve, the
// row-max fallback and the float intercepts (lane k writes a_out[k]).
// Optional outputs: out_ext (merged-row index of each axis's extreme point,
// through gidx; -1 when none) and out_a (the intercepts before the float
// rounding and sanitizing).
__device__ __forceinline__ void norm_solve_tail(const float* __restrict__ fn, int nm, int m, int n_nds,
                                                const int (&ext)[kMaxM], const float (&rmax)[kMaxM],
                                                SolveLds& L, int32_t* __restrict__ meta, float* __restrict__ a_out,
                                                const int32_t* __restrict__ gidx, int32_t* __restrict__ out_ext,
                                                double* __restrict__ out_a)
{
    const int lane = static_cast<int>(threadIdx.x & 63);
    int my_ext = -1;
    float my_rmax = 0.0f;
#pragma unroll
    for (int i = 0; i < kMaxM; ++i)
        if (lane == i) { my_ext = ext[i]; my_rmax = rmax[i]; }
    if (lane < m)
        for (int k = 0; k < m; ++k)
            L.a[lane * kMaxM + k] = (my_ext >= 0 && my_ext < n_nds)
                                        ? static_cast<double>(fn[static_cast<size_t>(k) * nm + my_ext])
                                        : static_cast<double>(INFINITY);
    __builtin_amdgcn_wave_barrier();
    double a = 0.0;
    const int ill = hyperplane_fp64_wave(m, L, a);
    if (lane >= m) return;
    if (ill) a = (my_rmax > 1e-10f) ? 1.0 / static_cast<double>(fmaxf(my_rmax, 1e-6f)) : 1.0;
    if (out_a) out_a[lane] = a;
    if (out_ext) out_ext[lane] = (my_ext >= 0 && my_ext < n_nds) ? gidx[my_ext] : -1;
    float af = static_cast<float>(a);
    if (!isfinite(af) || af <= 0.0f) af = 1.0f;
    a_out[lane] = af;
    if (lane == 0) meta[kIll] = ill;
}

__device__ __forceinline__ void norm_solve_body(const float* __restrict__ fn, int nm, int m,
                                       const float* __restrict__ asf_val, const int32_t* __restrict__ asf_idx,
                                       const float* __restrict__ row_max, int32_t* __restrict__ meta,
                                       float* __restrict__ a_out, const int32_t* __restrict__ gidx = nullptr,
                                       int32_t* __restrict__ out_ext = nullptr,
                                       double* __restrict__ out_a = nullptr)
{
    __shared__ float s_best[kSolveThreads / 64][kMaxM];
    __shared__ int s_bidx[kSolveThreads / 64][kMaxM];
    __shared__ float s_rmax[kSolveThreads / 64][kMaxM];
    if (meta[kNRem] <= 0) return;
    const int n_nds = meta[kNPrior] + meta[kNLast];
    const int nb = (n_nds + 255) / 256;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // every objective's candidates in one pass over the blocks (each objective
    // still reduced in ascending block order per thread)
    float best[kMaxM], rm[kMaxM];
    int bi[kMaxM];
#pragma unroll
    for (int j = 0; j < kMaxM; ++j) { best[j] = INFINITY; rm[j] = -INFINITY; bi[j] = -1; }
    for (int b = threadIdx.x; b < nb; b += kSolveThreads) {
#pragma unroll
        for (int j = 0; j < kMaxM; ++j) {
            if (j >= m) break;
            float v = asf_val[b * m + j];
            const int idx = asf_idx[b * m + j];
            v = (v != v) ? INFINITY : v;
            if (idx >= 0 && v < INFINITY && (bi[j] < 0 || v < best[j] || (v == best[j] && idx < bi[j]))) {
                best[j] = v;
                bi[j] = idx;
            }
            rm[j] = fmaxf(rm[j], row_max[b * m + j]);
        }
    }
#pragma unroll
    for (int j = 0; j < kMaxM; ++j) {
        if (j >= m) break;
        float bj = best[j], rj = rm[j];
        int ij = bi[j];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            const float ov = __shfl_xor(bj, o);
            const int oi = __shfl_xor(ij, o);
            if (oi >= 0 && ov < INFINITY && (ij < 0 || ov < bj || (ov == bj && oi < ij))) { bj = ov; ij = oi; }
            rj = fmaxf(rj, __shfl_xor(rj, o));
        }
        if (lane == 0) { s_best[wave][j] = bj; s_bidx[wave][j] = ij; s_rmax[wave][j] = rj; }
    }
    __syncthreads();
    if (wave != 0) return;
    int ext[kMaxM];
    float rmax[kMaxM];
#pragma unroll
    for (int j = 0; j < kMaxM; ++j) {
        float best = INFINITY, rm = -INFINITY;
        int bi = -1;
        if (j < m) {
            for (int w = 0; w < kSolveThreads / 64; ++w) {
                const float ov = s_best[w][j];
                const int oi = s_bidx[w][j];
                if (oi >= 0 && ov < INFINITY && (bi < 0 || ov < best || (ov == best && oi < bi))) { best = ov; bi = oi; }
                rm = fmaxf(rm, s_rmax[w][j]);
            }
        }
        ext[j] = bi;
        rmax[j] = rm;
    }
    __shared__ SolveLds s_solve;
    if (m >= 1 && m <= kMaxM)
        norm_solve_tail(fn, nm, m, n_nds, ext, rmax, s_solve, meta, a_out, gidx, out_ext, out_a);
}

__global__ __launch_bounds__(kSolveThreads) void norm_solve(const float* __restrict__ fn, int nm, int m,
                                                            const float* __restrict__ asf_val,
                                                            const int32_t* __restrict__ asf_idx,
                                                            const float* __restrict__ row_max,
                                                            int32_t* __restrict__ meta,
                                                            float* __restrict__ a_out,
                                                            const int32_t* __restrict__ gidx,
                                                            int32_t* __restrict__ out_ext,
                                                            double* __restrict__ out_a)
{
    norm_solve_body(fn, nm, m, asf_val, asf_idx, row_max, meta, a_out, gidx, out_ext, out_a);
}

// Perpendicular-distance association, one individual per thread against a
// chunk of directions staged in LDS (association.cu:85-188): dot by an fmaf
// chain, residual x - dot*w as one fma, d2 by an fmaf chain, first minimum
// wins.  Grid: (ceil(2n/256), ceil(K/kAssocChunk)).
constexpr int kAssocChunk = 1024;

template <int M>
__device__ __forceinline__ void assoc_partial_body(const float* __restrict__ fn, int nm, const float* __restrict__ a,
                                          const float* __restrict__ w, int k, const int32_t* __restrict__ meta,
                                          float* __restrict__ part_d2, int32_t* __restrict__ part_idx, int bx,
                                          int by)
{
    __shared__ float s_w[M][kAssocChunk];
    if (meta[kNRem] <= 0) return;
    const int n_nds = meta[kNPrior] + meta[kNLast];
    if (bx * 256 >= n_nds) return;
    const int j0 = by * kAssocChunk;
    const int jn = min(kAssocChunk, k - j0);
    for (int e = threadIdx.x; e < M * kAssocChunk; e += 256) {
        const int c = e / kAssocChunk, j = e % kAssocChunk;
        s_w[c][j] = (j < jn) ? w[static_cast<size_t>(c) * k + j0 + j] : 0.0f;
    }
    __syncthreads();
    const int i = bx * 256 + threadIdx.x;
    if (i >= n_nds) return;
    float x[M];
#pragma unroll
    for (int c = 0; c < M; ++c) x[c] = fn[static_cast<size_t>(c) * nm + i] * a[c];
    float best = INFINITY;
    int bj = -1;
    for (int j = 0; j < jn; ++j) {
        float dot = 0.0f;
#pragma unroll
        for (int c = 0; c < M; ++c) dot = fmaf(x[c], s_w[c][j], dot);
        float d2 = 0.0f;
#pragma unroll
        for (int c = 0; c < M; ++c) {
            const float r = fmaf(-dot, s_w[c][j], x[c]);
            d2 = fmaf(r, r, d2);
        }
        if (d2 < best) { best = d2; bj = j0 + j; }
    }
    part_d2[static_cast<size_t>(by) * nm + i] = best;
    part_idx[static_cast<size_t>(by) * nm + i] = bj;
}

template <int M>
__global__ __launch_bounds__(256) void associate_partial(const float* __restrict__ fn, int nm,
                                                         const float* __restrict__ a,
                                                         const float* __restrict__ w, int k,
                                                         const int32_t* __restrict__ meta,
                                                         float* __restrict__ part_d2,
                                                         int32_t* __restrict__ part_idx)
{
    assoc_partial_body<M>(fn, nm, a, w, k, meta, part_d2, part_idx, static_cast<int>(blockIdx.x),
                          static_cast<int>(blockIdx.y));
}

// Reduce the chunk partials in chunk order (association_phase2_kernel, :397-447).
// rho (optional): prior-front individuals add one to their direction's count
// (prior_histogram_kernel, niching.cu:21-33).
__device__ __forceinline__ void assoc_final_body(const float* __restrict__ part_d2, const int32_t* __restrict__ part_idx,
                                        int nchunks, int nm, const int32_t* __restrict__ meta,
                                        float* __restrict__ dist, int32_t* __restrict__ ref, int i,
                                        int32_t* __restrict__ rho = nullptr)
{
    if (meta[kNRem] <= 0) return;
    const int n_nds = meta[kNPrior] + meta[kNLast];
    if (i >= n_nds) return;
    float best = INFINITY;
    int bj = -1;
    for (int c = 0; c < nchunks; ++c) {
        const float d = part_d2[static_cast<size_t>(c) * nm + i];
        const int idx = part_idx[static_cast<size_t>(c) * nm + i];
        if (d < best || (d == best && idx >= 0 && idx < bj)) { best = d; bj = idx; }
    }
    dist[i] = (best != INFINITY) ? sqrtf(best) : INFINITY;
    ref[i] = bj;
    if (rho && bj >= 0 && i < meta[kNPrior]) atomicAdd(&rho[bj], 1);
}

// ------------------------------------------------------ pruned association --
// Exact association with pruning, for many directions.  Directions are cut
// into blocks of 64 consecutive indices; each block gets a cone (unit center c
// and half-angle alpha covering every direction of the block).  For an
// individual x the kernel's d2 for direction w (dot by an fmaf chain,
// r = fmaf(-dot, w, x), d2 by an fmaf chain) satisfies
//     d2 >= (1 - 1e-6) |x|^2 max(0, sin(theta) - 1e-6 |w|^2)^2,
// theta the angle between x and w: the perpendicular residual of the exact
// product is at least |x| sin(theta) (|w|^2 (2 - |w|^2) <= 1), the dot
// product's error is at most gamma_M |x| |w| and every later rounding is
// relative (gamma_10 < 6.1e-7; all terms of the sums are non-negative).
// With cur the best d2 found so far, every w with sin(theta) > S, where
//     S = sqrt(max(cur, 2^-80) / ((1 - 1e-6) |x|^2)) + 1e-6 max|w|^2,
// has d2 > cur: it is neither a minimum nor a tie (the floor keeps the bound
// where every float rounding of d2 is relative, not subnormal).  Every direction of a block
// lies within alpha of its center, whose angle to x is theta_c, so if
//     |cos(theta_c)| < cos(alpha + asin S) = cos(alpha) sqrt(1 - S^2) - sin(alpha) S
// (alpha + asin S < pi/2) all its angles lie in (asin S, pi - asin S) and the
// block is skipped; the others are evaluated with the kernel's own arithmetic.
// S is rounded up, sqrt(1 - S^2) down, cos(alpha) down and sin(alpha) up (in
// the float table too), and the compare keeps a margin of 2e-6 over the error
// of cos(theta_c), a float fmaf chain over x / |x| (formed in double, rounded
// to float: 2^-24 relative, or below 2^-149 absolute) and the table's float
// center (within 2^-24 of the unit center): with the chain's gamma_M, at most
// (M + 2) 2^-24 <= 7.2e-7 for M <= 10, with no overflow or underflow of |x|.
// The result (smallest d2, then smallest index, non-finite d2 never chosen) is
// therefore the brute-force result bit for bit.  One wave per individual: its
// lanes compute cos(theta_c) for every block, evaluate the block with the
// nearest center first, set the threshold from that block's minimum, then
// evaluate every block that passes it.
constexpr int kConeMaxBlocks = 1024;   // directions up to 65536
constexpr int kConeMinK = 2048;        // below this the brute-force kernel is cheaper
constexpr int kConeAutoMaxM = 3;

__device__ __forceinline__ double wave_sum_f64(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Cone of direction block b: center c (M components), cos alpha, sin alpha,
// and the largest |w|^2 of the block, computed in double; the angle terms are
// rounded outwards, then stored as floats rounded outwards again (cos alpha
// down, sin alpha and |w|^2 up).  Layout: M + 3 rows of ceil(k / 64) floats,
// row r of block b at r * ceil(k / 64) + b, so a wave's loads are contiguous.
template <int M>
__device__ __forceinline__ void assoc_cone_body(const float* __restrict__ w, int k, const int32_t* __restrict__ perm,
                                                float* __restrict__ wp, float* __restrict__ cones, int b)
{
    const int lane = static_cast<int>(threadIdx.x & 63);
    const bool valid = b * 64 + lane < k;
    const int j = valid ? perm[b * 64 + lane] : 0;
    double u[M];
    double nn = 0.0;
#pragma unroll
    for (int c = 0; c < M; ++c) {
        const float wc = valid ? w[static_cast<size_t>(c) * k + j] : 0.0f;
        if (valid) wp[static_cast<size_t>(c) * k + b * 64 + lane] = wc;   // directions in block order
        u[c] = static_cast<double>(wc);
        nn += u[c] * u[c];
    }
    const bool unit = valid && nn > 0.0 && nn == nn;
    const double inv = unit ? 1.0 / sqrt(nn) : 0.0;
    double cen[M];
    double cn = 0.0;
#pragma unroll
    for (int c = 0; c < M; ++c) {
        u[c] *= inv;
        cen[c] = wave_sum_f64(u[c]);
        cn += cen[c] * cen[c];
    }
    cn = sqrt(cn);
#pragma unroll
    for (int c = 0; c < M; ++c) cen[c] = (cn > 0.0) ? cen[c] / cn : (c == 0 ? 1.0 : 0.0);
    double cj = 1.0;                               // lanes past k do not widen the cone
    if (valid) {
        if (unit) {
            cj = 0.0;
#pragma unroll
            for (int c = 0; c < M; ++c) cj += cen[c] * u[c];
        } else {
            cj = -1.0;                             // zero or non-finite direction: no pruning
        }
    }
    double w2 = valid ? nn : 0.0;
    if (!(w2 == w2)) w2 = INFINITY;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        cj = fmin(cj, __shfl_xor(cj, o));
        w2 = fmax(w2, __shfl_xor(w2, o));
    }
    double ca = cj - 1e-12;
    if (ca < -1.0) ca = -1.0;
    const double sa = sqrt(fmax(0.0, 1.0 - ca * ca)) + 1e-12;
    if (lane == 0) {
        const size_t nbk = static_cast<size_t>((k + 63) / 64);
#pragma unroll
        for (int c = 0; c < M; ++c) cones[c * nbk + b] = static_cast<float>(cen[c]);
        cones[M * nbk + b] = __double2float_rd(ca);
        cones[(M + 1) * nbk + b] = __double2float_ru(sa);
        cones[(M + 2) * nbk + b] = __double2float_ru(w2);
    }
}

template <int M>
__global__ __launch_bounds__(256) void assoc_cones(const float* __restrict__ w, int k, const int32_t* __restrict__ perm,
                                                   float* __restrict__ wp, float* __restrict__ cones)
{
    const int b = static_cast<int>(blockIdx.x) * 4 + static_cast<int>(threadIdx.x >> 6);
    if (b >= (k + 63) / 64) return;
    assoc_cone_body<M>(w, k, perm, wp, cones, b);
}

// Spatial order of the directions for the cone blocks: a Morton key of the
// first M - 1 coordinates of each unit direction (the last is implied on the
// simplex), so 64 consecutive keys cover a compact patch.  Any order gives
// the same association; this one gives tight cones.
constexpr int kMortonBits = 60;

__host__ __device__ __forceinline__ int morton_coords(int m) { return m > 1 ? m - 1 : 1; }

template <int M>
__global__ void assoc_dir_keys(const float* __restrict__ w, int k, uint64_t* __restrict__ keys,
                               int32_t* __restrict__ vals)
{
    const int j = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (j >= k) return;
    constexpr int C = (M > 1) ? M - 1 : 1;
    constexpr int bits = (kMortonBits / C) < 16 ? (kMortonBits / C) : 16;
    double u[M];
    double nn = 0.0;
#pragma unroll
    for (int c = 0; c < M; ++c) {
        u[c] = static_cast<double>(w[static_cast<size_t>(c) * k + j]);
        nn += u[c] * u[c];
    }
    const double inv = (nn > 0.0 && nn == nn) ? 1.0 / sqrt(nn) : 0.0;
    uint32_t q[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        double v = u[c] * inv;
        v = (v == v) ? fmin(1.0, fmax(0.0, v)) : 0.0;   // negative components fold to 0: order only
        q[c] = static_cast<uint32_t>(v * static_cast<double>((1u << bits) - 1) + 0.5);
    }
    uint64_t key = 0;
    for (int bit = bits - 1; bit >= 0; --bit)
#pragma unroll
        for (int c = 0; c < C; ++c) key = (key << 1) | ((q[c] >> bit) & 1u);
    keys[j] = key;
    vals[j] = j;
}

template <int M>
__device__ __forceinline__ void assoc_pruned_body(const float* __restrict__ fn, int nm, const float* __restrict__ a,
                                                  const float* __restrict__ wp, int k,
                                                  const int32_t* __restrict__ perm,
                                                  const float* __restrict__ cones,
                                                  const int32_t* __restrict__ meta, float* __restrict__ dist,
                                                  int32_t* __restrict__ ref, int i, uint16_t* __restrict__ list,
                                                  int32_t* __restrict__ rho)
{
    constexpr int T = kConeMaxBlocks / 64;   // blocks per lane: lane, lane + 64, ...
    const int lane = static_cast<int>(threadIdx.x & 63);
    const int nbk = (k + 63) / 64;
    float x[M];
#pragma unroll
    for (int c = 0; c < M; ++c) x[c] = fn[static_cast<size_t>(c) * nm + i] * a[c];
    double xx = 0.0;
#pragma unroll
    for (int c = 0; c < M; ++c) xx += static_cast<double>(x[c]) * static_cast<double>(x[c]);
    const double xn = sqrt(xx);
    const bool finite_x = xx < INFINITY;                    // false for an inf or nan coordinate
    const double inv_xn = (xn > 0.0 && finite_x) ? 1.0 / xn : 0.0;
    float u[M];                                             // x / |x|
#pragma unroll
    for (int c = 0; c < M; ++c) u[c] = static_cast<float>(static_cast<double>(x[c]) * inv_xn);
    float cth[T];                                           // cos(theta_c) of blocks lane, lane + 64, ...
    float best_cos = -3.0f, w2max = 0.0f;
    int best_b = 0;
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int b = lane + 64 * t;
        cth[t] = 0.0f;
        if (b < nbk) {
            float ct = 0.0f;
#pragma unroll
            for (int c = 0; c < M; ++c) ct = fmaf(u[c], cones[c * nbk + b], ct);
            cth[t] = ct;
            w2max = fmaxf(w2max, cones[(M + 2) * nbk + b]);
            if (ct > best_cos) { best_cos = ct; best_b = b; }
        }
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const float oc = __shfl_xor(best_cos, o);
        const int ob = __shfl_xor(best_b, o);
        if (oc > best_cos || (oc == best_cos && ob < best_b)) { best_cos = oc; best_b = ob; }
        w2max = fmaxf(w2max, __shfl_xor(w2max, o));
    }
    const int first = wave_uniform(best_b);
    float bd = INFINITY;
    int bq = -1;                 // block-order position of the lane's best; its original index decides ties
    // the kernel's arithmetic for the direction at block-order position q
    auto consider = [&](const float (&wj)[M], int q) {
        float dot = 0.0f;
#pragma unroll
        for (int c = 0; c < M; ++c) dot = fmaf(x[c], wj[c], dot);
        float d2 = 0.0f;
#pragma unroll
        for (int c = 0; c < M; ++c) {
            const float r = fmaf(-dot, wj[c], x[c]);
            d2 = fmaf(r, r, d2);
        }
        if (d2 < bd) {
            bd = d2;
            bq = q;
        } else if (d2 == bd && d2 < INFINITY && perm[q] < perm[bq]) {
            bq = q;
        }
    };
    {
        const int q = first * 64 + lane;
        if (q < k) {
            float wj[M];
#pragma unroll
            for (int c = 0; c < M; ++c) wj[c] = wp[static_cast<size_t>(c) * k + q];
            consider(wj, q);
        }
    }
    float cur = bd;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) cur = fminf(cur, __shfl_xor(cur, o));
    // the skip threshold from the nearest block's minimum
    bool prune = finite_x && xx > 0.0 && cur < INFINITY && w2max < INFINITY;
    double S = 1.0, C = 0.0;
    if (prune) {
        // below 2^-80 the float d2 may carry subnormal (absolute) rounding
        const double floor_cur = fmax(static_cast<double>(cur), 0x1p-80);
        const double s0 = sqrt(floor_cur / ((1.0 - 1e-6) * xx)) * (1.0 + 1e-12);
        S = s0 + 1e-6 * static_cast<double>(w2max) * (1.0 + 1e-12);
        prune = S < 1.0;
        if (prune) C = sqrt(1.0 - S * S) * (1.0 - 1e-12);
    }
    // candidates: the blocks that fail the skip test, compacted in ascending
    // order into the wave's list
    int nc = 0;
#pragma unroll
    for (int t = 0; t < T; ++t) {
        if (64 * t >= nbk) break;
        const int b = lane + 64 * t;
        bool skip = false;
        if (prune && b < nbk) {
            const double ca = cones[M * nbk + b], sa = cones[(M + 1) * nbk + b];
            skip = ca >= 0.0 && static_cast<double>(fabsf(cth[t])) < ca * C - sa * S - 2e-6;
        }
        const bool cand = b < nbk && b != first && !skip;
        const uint64_t mk = ballot(cand);
        if (cand)
            list[nc + static_cast<int>(__builtin_amdgcn_mbcnt_hi(
                          static_cast<uint32_t>(mk >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mk), 0u)))] =
                static_cast<uint16_t>(b);
        nc += __builtin_popcountll(mk);
    }
    __builtin_amdgcn_wave_barrier();
    // evaluate them four at a time, their loads in flight together
    constexpr int kBatch = 4;
    for (int base = 0; base < nc; base += kBatch) {
        float wv[kBatch][M];
        int qv[kBatch];
#pragma unroll
        for (int v = 0; v < kBatch; ++v) {
            const int q = (base + v < nc) ? static_cast<int>(list[base + v]) * 64 + lane : k;
            qv[v] = q;
#pragma unroll
            for (int c = 0; c < M; ++c) wv[v][c] = q < k ? wp[static_cast<size_t>(c) * k + q] : 0.0f;
        }
#pragma unroll
        for (int v = 0; v < kBatch; ++v)
            if (qv[v] < k) consider(wv[v], qv[v]);
    }
    int bj = bq >= 0 ? perm[bq] : -1;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const float od = __shfl_xor(bd, o);
        const int oj = __shfl_xor(bj, o);
        if (od < bd || (od == bd && od < INFINITY && oj < bj)) { bd = od; bj = oj; }
    }
    if (lane == 0) {
        dist[i] = (bd != INFINITY) ? sqrtf(bd) : INFINITY;
        ref[i] = bj;
        if (rho && bj >= 0 && i < meta[kNPrior]) atomicAdd(&rho[bj], 1);
    }
}

template <int M>
__global__ __launch_bounds__(256) void associate_pruned(const float* __restrict__ fn, int nm,
                                                        const float* __restrict__ a, const float* __restrict__ wp,
                                                        int k, const int32_t* __restrict__ perm,
                                                        const float* __restrict__ cones,
                                                        const int32_t* __restrict__ meta, float* __restrict__ dist,
                                                        int32_t* __restrict__ ref, int32_t* __restrict__ rho)
{
    __shared__ __align__(8) uint16_t s_list[4][kConeMaxBlocks];   // candidate blocks per wave
    if (meta[kNRem] <= 0) return;
    const int n_nds = meta[kNPrior] + meta[kNLast];
    const int wave = static_cast<int>(threadIdx.x >> 6);
    const int i = static_cast<int>(blockIdx.x) * 4 + wave;
    if (i >= n_nds) return;
    assoc_pruned_body<M>(fn, nm, a, wp, k, perm, cones, meta, dist, ref, i, s_list[wave], rho);
}

__global__ void associate_final(const float* __restrict__ part_d2, const int32_t* __restrict__ part_idx,
                                int nchunks, int nm, const int32_t* __restrict__ meta,
                                float* __restrict__ dist, int32_t* __restrict__ ref, int32_t* __restrict__ rho)
{
    assoc_final_body(part_d2, part_idx, nchunks, nm, meta, dist, ref,
                     static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x), rho);
}

// Buffers of the cone preparation: direction keys and permutation, the sort
// scratch and the cones.
struct ConeBufs {
    uint64_t* kin; uint64_t* kout; int32_t* vin; int32_t* perm; void* tmp; size_t tmp_bytes; float* cones;
    float* wp;
};

__host__ __forceinline__ int morton_key_bits(int m)
{
    const int c = morton_coords(m);
    const int bits = (kMortonBits / c) < 16 ? (kMortonBits / c) : 16;
    return bits * c;
}

ConeBufs carve_cones(Carve& c, int m, int k)
{
    ConeBufs B{};
    B.kin = c.take<uint64_t>(k);
    B.kout = c.take<uint64_t>(k);
    B.vin = c.take<int32_t>(k);
    B.perm = c.take<int32_t>(k);
    B.tmp_bytes = sort_pairs_temp_bytes(k, morton_key_bits(m));
    B.tmp = c.take_bytes(B.tmp_bytes);
    B.cones = c.take<float>(static_cast<size_t>(ceil_div(k, 64)) * (m + 3));
    B.wp = c.take<float>(static_cast<size_t>(m) * k);
    return B;
}

int cones_prepare(const float* w, int k, int m, const ConeBufs& B, hipStream_t st)
{
    const dim3 b256(256);
    MOEA_LAUNCH_M(m, assoc_dir_keys, dim3(ceil_div(k, 256)), b256, 0, st, w, k, B.kin, B.vin);
    MOEA_HIP(hipGetLastError());
    MOEA_CHECK(sort_pairs(B.tmp, B.tmp_bytes, B.kin, B.kout, B.vin, B.perm, k, st, morton_key_bits(m)));
    MOEA_LAUNCH_M(m, assoc_cones, dim3(ceil_div(ceil_div(k, 64), 4)), b256, 0, st, w, k, B.perm, B.wp, B.cones);
    MOEA_HIP(hipGetLastError());
    return kOk;
}


// rho_prior (prior_histogram_kernel, niching.cu:21-33) is cleared by norm_prep
// and filled by the association kernels as they write the prior front's
// directions.  Individuals with no association (ref -1, from non-finite
// normalized objectives) count in no niche; the reference would index rho[-1]
// there.

// Phase-2 keys: ref << 32 | bits(dist), value = last-front local index
// (build_lf_asrpts_asdist_key_kernel, niching.cu:68-92).  The high word is
// the direction, with an unassociated row (ref -1, whose u32 sorts after every
// direction) written as k and slots past N_last as k + 1 with an all-ones low
// word: the same order in 32 + bitlen(k + 1) key bits.
__device__ __forceinline__ uint32_t ref_word(int r, int k) { return r < 0 ? static_cast<uint32_t>(k) : static_cast<uint32_t>(r); }

// hi_base: bits above the key (a batch index); key arrays are the instance's view.
__device__ __forceinline__ void niche_keys1_body(const int32_t* __restrict__ ref, const float* __restrict__ dist,
                                        const int32_t* __restrict__ meta, int k, uint64_t hi_base,
                                        uint64_t* __restrict__ key, int32_t* __restrict__ val, int t)
{
    const int n_last = (meta[kNRem] > 0) ? meta[kNLast] : 0;
    const int n_prior = meta[kNPrior];
    uint64_t kk = (static_cast<uint64_t>(k + 1) << 32) | 0xFFFFFFFFull;
    if (t < n_last) {
        const uint32_t r = ref_word(ref[n_prior + t], k);
        kk = (static_cast<uint64_t>(r) << 32) | __float_as_uint(dist[n_prior + t]);
    }
    key[t] = hi_base | kk;
    val[t] = t;
}

__global__ void niche_keys1(const int32_t* __restrict__ ref, const float* __restrict__ dist,
                            const int32_t* __restrict__ meta, int nm, int k, uint64_t* __restrict__ key,
                            int32_t* __restrict__ val)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nm) niche_keys1_body(ref, dist, meta, k, 0ull, key, val, t);
}

// First position of the run of equal high words containing t (keys sorted).
__device__ __forceinline__ int segment_head(const uint64_t* __restrict__ key, int t)
{
    const uint32_t hi = static_cast<uint32_t>(key[t] >> 32);
    int lo = 0, up = t;
    while (lo < up) {
        const int mid = (lo + up) >> 1;
        if (static_cast<uint32_t>(key[mid] >> 32) < hi) lo = mid + 1;
        else up = mid;
    }
    return lo;
}

__device__ __forceinline__ uint32_t niche_draw(const uint32_t* injected, int which, int t, uint64_t seed,
                                      uint32_t gen)
{
    return injected ? injected[t] : splitmix32(niche_base(gen, which) + static_cast<uint64_t>(t), seed);
}

// Phase-4 rank keys: ref << 32 | (rho == 0 && local == 0 ? 0 : draw_a | 1)
// (compute_rank_keys_kernel, niching.cu:114-139).
// word_mask: the direction field of the high word (below any batch index).
__device__ __forceinline__ void niche_keys2_body(const uint64_t* __restrict__ key1s, const int32_t* __restrict__ meta,
                                        const int32_t* __restrict__ rho, int k, uint32_t word_mask,
                                        uint64_t hi_base, const uint32_t* __restrict__ inj_a, uint64_t seed,
                                        uint32_t gen, uint64_t* __restrict__ key2, int t)
{
    const int n_last = (meta[kNRem] > 0) ? meta[kNLast] : 0;
    uint64_t kk = (static_cast<uint64_t>(k + 1) << 32) | 0xFFFFFFFFull;
    if (t < n_last) {
        const uint32_t r = static_cast<uint32_t>(key1s[t] >> 32) & word_mask;
        const int local = t - segment_head(key1s, t);
        const int ri = static_cast<int>(r);
        const int rh = (ri >= 0 && ri < k) ? rho[ri] : 0;
        const uint32_t low = (rh == 0 && local == 0) ? 0u : (niche_draw(inj_a, 0, t, seed, gen) | 1u);
        kk = (static_cast<uint64_t>(r) << 32) | low;
    }
    key2[t] = hi_base | kk;
}

__global__ void niche_keys2(const uint64_t* __restrict__ key1s, const int32_t* __restrict__ meta,
                            const int32_t* __restrict__ rho, int k, int nm,
                            const uint32_t* __restrict__ inj_a, uint64_t seed, uint32_t gen,
                            uint64_t* __restrict__ key2)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nm) niche_keys2_body(key1s, meta, rho, k, 0xFFFFFFFFu, 0ull, inj_a, seed, gen, key2, t);
}

// Phase-5 keys: (rho + rank) << 32 | draw_b (compute_niching_keys_kernel, niching.cu:157-178).
// rho + rank < N_prior + N_last <= nm, so slots past N_last take the high word
// nm: 32 + bitlen(nm) key bits.
__device__ __forceinline__ void niche_keys3_body(const uint64_t* __restrict__ key2s, const int32_t* __restrict__ meta,
                                        const int32_t* __restrict__ rho, int k, int nm, uint32_t word_mask,
                                        uint64_t hi_base, const uint32_t* __restrict__ inj_b, uint64_t seed,
                                        uint32_t gen, uint64_t* __restrict__ key3, int t)
{
    const int n_last = (meta[kNRem] > 0) ? meta[kNLast] : 0;
    uint64_t kk = (static_cast<uint64_t>(nm) << 32) | 0xFFFFFFFFull;
    if (t < n_last) {
        const int ri = static_cast<int>(static_cast<uint32_t>(key2s[t] >> 32) & word_mask);
        const int rank = t - segment_head(key2s, t);
        const int rh = (ri >= 0 && ri < k) ? rho[ri] : 0;
        kk = (static_cast<uint64_t>(static_cast<uint32_t>(rh + rank)) << 32) |
             niche_draw(inj_b, 1, t, seed, gen);
    }
    key3[t] = hi_base | kk;
}

__global__ void niche_keys3(const uint64_t* __restrict__ key2s, const int32_t* __restrict__ meta,
                            const int32_t* __restrict__ rho, int k, int nm,
                            const uint32_t* __restrict__ inj_b, uint64_t seed, uint32_t gen,
                            uint64_t* __restrict__ key3)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nm) niche_keys3_body(key2s, meta, rho, k, nm, 0xFFFFFFFFu, 0ull, inj_b, seed, gen, key3, t);
}

// Survivor flags of the first N_rem niching picks, beside the prior-front
// flags front_scatter wrote (gather_selected_kernel, niching.cu:192-206).
__global__ void mark_selected(const int32_t* __restrict__ val3s, const int32_t* __restrict__ gidx,
                              const int32_t* __restrict__ meta, uint8_t* __restrict__ flag)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int n_rem = meta[kNRem];
    if (n_rem <= 0 || meta[kNLast] <= 0) return;
    if (t >= n_rem) return;
    flag[gidx[meta[kNPrior] + val3s[t]]] = 1;
}

// ------------------------------------------------------ one-block niching --
// The three niching sorts and the survivor marks in one workgroup.  Only the
// last front's n_last members are sorted (the device-wide path sorts all 2N
// slots, the others as sentinels after them, so the members' order is the
// same), in LDS when they fit, else in the workspace's key/value buffers.
// Each sort is stable LSD radix with 8-bit digits.  Wave w owns the w-th of
// 16 contiguous chunks: it counts its chunk's digits, one prefix over
// (digit, wave) gives every wave its first position per digit, and the wave
// then walks its chunk 64 members at a time, each member going to its digit's
// next position plus its rank among the step's lanes with that digit (eight
// ballots).  Four barriers a pass whatever n_last is; a pass whose digit is one
// value for every member moves nothing and is skipped.  Keys and draws are
// those of niche_keys1..3.
constexpr int kNicheThreads = 1024;
constexpr int kNicheWaves = kNicheThreads / 64;
constexpr int kNicheLdsCap = 5888;   // members sorted in LDS

struct NicheShared {
    uint64_t key[2][kNicheLdsCap];
    int32_t val[2][kNicheLdsCap];
    int32_t row[kNicheWaves][256];   // per wave: digit counts, then next positions
    int32_t wsum[kNicheWaves];
    int32_t skip;
};

struct NicheLdsBufs {
    NicheShared* S;
    __device__ __forceinline__ uint64_t* key(int b) const { return b ? S->key[1] : S->key[0]; }
    __device__ __forceinline__ int32_t* val(int b) const { return b ? S->val[1] : S->val[0]; }
};
struct NicheGlobalBufs {
    uint64_t* k0;
    uint64_t* k1;
    int32_t* v0;
    int32_t* v1;
    __device__ __forceinline__ uint64_t* key(int b) const { return b ? k1 : k0; }
    __device__ __forceinline__ int32_t* val(int b) const { return b ? v1 : v0; }
};

// Lanes of the wave whose digit equals this lane's (valid lanes only).
__device__ __forceinline__ uint64_t digit_peers(uint32_t dg, bool valid)
{
    uint64_t peer = ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const bool bit = ((dg >> b) & 1u) != 0;
        const uint64_t m = ballot(bit);
        peer &= bit ? m : ~m;
    }
    return peer;
}

// Stable sort of the n pairs in buffer src by key bits [0, nbits); returns the
// buffer that holds the result.  pass: a counter unique to each pass of the
// kernel (tags the skip flag).
template <class Bufs>
__device__ int niche_sort(const Bufs& B, int src, int n, int nbits, NicheShared& S, int& pass)
{
    const int tid = static_cast<int>(threadIdx.x);
    const int lane = tid & 63, w = tid >> 6;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int chunk = (n + kNicheWaves - 1) / kNicheWaves;
    const int lo = min(n, w * chunk), hi = min(n, lo + chunk);
    int32_t* row = S.row[w];
    for (int shift = 0; shift < nbits; shift += 8, ++pass) {
#pragma unroll
        for (int j = 0; j < 4; ++j) row[lane * 4 + j] = 0;
        for (int i0 = lo; i0 < hi; i0 += 64) {          // the chunk's digit counts
            const int i = i0 + lane;
            const bool valid = i < hi;
            const uint32_t dg = valid ? static_cast<uint32_t>(B.key(src)[i] >> shift) & 255u : 0u;
            const uint64_t peer = digit_peers(dg, valid);
            if (valid && (peer & below) == 0) row[dg] += __popcll(peer);
        }
        __syncthreads();
        // thread t: digit d = t / 4 and waves 4q .. 4q + 3 (q = t % 4); wave w holds digits 16w .. 16w + 15
        const int d = tid >> 2, q = tid & 3;
        int c[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) c[j] = S.row[4 * q + j][d];
        const int sum = c[0] + c[1] + c[2] + c[3];
        int incl = sum;                                 // over the 4 threads of the digit
        {
            const int u1 = __shfl_up(incl, 1, 4);
            if (q >= 1) incl += u1;
            const int u2 = __shfl_up(incl, 2, 4);
            if (q >= 2) incl += u2;
        }
        const int before_q = incl - sum;
        const int total = __shfl(incl, 3, 4);           // the digit's count
        int dscan = total;                              // over the wave's 16 digits (lanes 4j .. 4j + 3)
#pragma unroll
        for (int o = 4; o < 64; o <<= 1) {
            const int u = __shfl_up(dscan, o);
            if (lane >= o) dscan += u;
        }
        if (lane == 63) S.wsum[w] = dscan;
        if (total == n && q == 0) S.skip = pass;        // every member has digit d
        __syncthreads();
        int base = dscan - total;                       // digits of earlier waves
        for (int v = 0; v < w; ++v) base += S.wsum[v];
        int at = base + before_q;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            S.row[4 * q + j][d] = at;
            at += c[j];
        }
        __syncthreads();
        if (S.skip == pass) continue;
        const int dst = src ^ 1;
        for (int i0 = lo; i0 < hi; i0 += 64) {
            const int i = i0 + lane;
            const bool valid = i < hi;
            uint64_t kv = 0;
            int32_t vv = 0;
            uint32_t dg = 0;
            if (valid) {
                kv = B.key(src)[i];
                vv = B.val(src)[i];
                dg = static_cast<uint32_t>(kv >> shift) & 255u;
            }
            const uint64_t peer = digit_peers(dg, valid);
            const int rank = __popcll(peer & below);
            if (valid) {
                const int pos = row[dg] + rank;
                B.key(dst)[pos] = kv;
                B.val(dst)[pos] = vv;
            }
            if (valid && rank == 0) row[dg] += __popcll(peer);
        }
        __syncthreads();
        src = dst;
    }
    return src;
}

// First position of the run of equal high words containing t (keys sorted).
__device__ __forceinline__ int run_head(const uint64_t* key, int t)
{
    const uint32_t hi = static_cast<uint32_t>(key[t] >> 32);
    int lo = 0, up = t;
    while (lo < up) {
        const int mid = (lo + up) >> 1;
        if (static_cast<uint32_t>(key[mid] >> 32) < hi) lo = mid + 1;
        else up = mid;
    }
    return lo;
}

struct NicheArgs {
    const int32_t* ref;
    const float* dist;
    const int32_t* rho;
    int k, bits_ref, bits_rank;
    const uint32_t* inj_a;
    const uint32_t* inj_b;
    uint64_t seed;
    uint32_t gen;
    const int32_t* gidx;
    uint8_t* flag;
    int bucket;   // 0: always the radix sorts
};

template <class Bufs>
__device__ void niche_body(const Bufs& B, const NicheArgs& A, NicheShared& S, int n_prior, int n_last, int n_rem)
{
    const int tid = static_cast<int>(threadIdx.x);
    const int k = A.k;
    int pass = 0;
    // phase 2: (direction, distance), value = last-front index
    for (int t = tid; t < n_last; t += kNicheThreads) {
        const uint32_t r = ref_word(A.ref[n_prior + t], k);
        B.key(0)[t] = (static_cast<uint64_t>(r) << 32) | __float_as_uint(A.dist[n_prior + t]);
        B.val(0)[t] = t;
    }
    __syncthreads();
    int src = niche_sort(B, 0, n_last, A.bits_ref, S, pass);
    // phase 4: (direction, 0 for the nearest member of an empty niche, else draw_a | 1)
    for (int t = tid; t < n_last; t += kNicheThreads) {
        const uint32_t r = static_cast<uint32_t>(B.key(src)[t] >> 32);
        const int local = t - run_head(B.key(src), t);
        const int ri = static_cast<int>(r);
        const int rh = (ri >= 0 && ri < k) ? A.rho[ri] : 0;
        const uint32_t low = (rh == 0 && local == 0) ? 0u : (niche_draw(A.inj_a, 0, t, A.seed, A.gen) | 1u);
        B.key(src ^ 1)[t] = (static_cast<uint64_t>(r) << 32) | low;
        B.val(src ^ 1)[t] = B.val(src)[t];
    }
    __syncthreads();
    src = niche_sort(B, src ^ 1, n_last, A.bits_ref, S, pass);
    // phase 5: (rho + rank in the niche, draw_b)
    for (int t = tid; t < n_last; t += kNicheThreads) {
        const int ri = static_cast<int>(static_cast<uint32_t>(B.key(src)[t] >> 32));
        const int rank = t - run_head(B.key(src), t);
        const int rh = (ri >= 0 && ri < k) ? A.rho[ri] : 0;
        B.key(src ^ 1)[t] = (static_cast<uint64_t>(static_cast<uint32_t>(rh + rank)) << 32) |
                            niche_draw(A.inj_b, 1, t, A.seed, A.gen);
        B.val(src ^ 1)[t] = B.val(src)[t];
    }
    __syncthreads();
    src = niche_sort(B, src ^ 1, n_last, A.bits_rank, S, pass);
    const int take = n_rem < n_last ? n_rem : n_last;
    for (int t = tid; t < take; t += kNicheThreads) A.flag[A.gidx[n_prior + B.val(src)[t]]] = 1;
}

// Direction buckets: the usual case.  Sort 1 and sort 2 keep every niche's
// members together in direction order, so a member's position is its niche's
// start plus its rank inside the niche; niches are small, so the ranks are
// counted directly: by (distance bits, last-front index) for sort 1, by
// (phase-4 low word, sort-1 position) for sort 2.  Sort 3 only decides which
// n_rem members come first: its keys (rho + rank, draw_b, sort-2 position) are
// distinct, so an 8-round radix select of the n_rem-th smallest finds the
// chosen set exactly.  A niche of more than kBucketMaxRun members (counting
// would be quadratic) or more than 65535 members in all (16-bit positions)
// takes the radix sorts instead.
constexpr int kNicheLdsBytes = 162816;
constexpr int kBucketMaxRun = 128;
constexpr int kBucketSlotBytes = 10;   // LDS: t, r / p2, p1 / lev: 2 B each; a0 / db: 4 B
constexpr int kBucketRegs = 10;        // members a thread holds through the sort-2 step (LDS case)

struct BucketSlots {
    uint16_t* t;     // last-front index
    uint16_t* r;     // direction word (k: unassociated)
    uint32_t* a0;    // distance bits, then the phase-4 low word
    uint16_t* p1;    // sort-1 position
    uint16_t* lev;   // rho + rank in the niche
    uint32_t* db;    // draw_b
    uint16_t* p2;    // sort-2 position
};

// LDS: four arrays of `cap` members; the sort-2 step replaces a0 by db, p1 by
// lev and r by p2 (its results wait in registers until every member's inputs
// are read).
__device__ __forceinline__ BucketSlots bucket_slots(unsigned char* base, int cap)
{
    BucketSlots v;
    v.a0 = reinterpret_cast<uint32_t*>(base);
    v.db = v.a0;
    v.t = reinterpret_cast<uint16_t*>(v.a0 + cap);
    v.r = v.t + cap;
    v.p2 = v.r;
    v.p1 = v.r + cap;
    v.lev = v.p1;
    return v;
}

// Workspace (members past the LDS capacity): a0 and db in k0, the 16-bit
// arrays in k1 and v0, each of 2n slots (8, 8 and 4 bytes a slot).
__device__ __forceinline__ BucketSlots bucket_slots_global(uint64_t* k0, uint64_t* k1, int32_t* v0, int nm)
{
    BucketSlots v;
    v.a0 = reinterpret_cast<uint32_t*>(k0);
    v.db = v.a0 + nm;
    v.t = reinterpret_cast<uint16_t*>(k1);
    v.r = v.t + nm;
    v.p1 = v.r + nm;
    v.lev = v.p1 + nm;
    v.p2 = reinterpret_cast<uint16_t*>(v0);
    return v;
}

__device__ __forceinline__ int half_get(const uint32_t* table, int r)
{
    const uint32_t w = table[r >> 1];
    return static_cast<int>((r & 1) ? (w >> 16) : (w & 0xFFFFu));
}

// Slots and ranks, then the selection; table holds each niche's start (16-bit).
// kShared: V aliases the sort-2 outputs onto its inputs (n_last <= 1024 kBucketRegs).
template <bool kShared>
__device__ __forceinline__ void bucket_finish(const NicheArgs& A, uint32_t* table, int32_t* misc,
                                              const BucketSlots& V, int n_prior, int n_last, int n_rem)
{
    const int tid = static_cast<int>(threadIdx.x);
    const int k = A.k;
    // place each member in its niche (order inside a niche does not matter:
    // the ranks below break ties explicitly); afterwards table holds the ends
    for (int t = tid; t < n_last; t += kNicheThreads) {
        const int r = static_cast<int>(ref_word(A.ref[n_prior + t], k));
        const uint32_t old = atomicAdd(&table[r >> 1], (r & 1) ? 0x10000u : 1u);
        const int slot = static_cast<int>((r & 1) ? (old >> 16) : (old & 0xFFFFu));
        V.t[slot] = static_cast<uint16_t>(t);
        V.r[slot] = static_cast<uint16_t>(r);
        V.a0[slot] = __float_as_uint(A.dist[n_prior + t]);
    }
    __syncthreads();
    // sort 1: rank by (distance bits, t) inside the niche
    for (int s = tid; s < n_last; s += kNicheThreads) {
        const int r = V.r[s];
        const int st = r ? half_get(table, r - 1) : 0, en = half_get(table, r);
        const uint32_t d = V.a0[s];
        const int t = V.t[s];
        int rank = 0;
        for (int u = st; u < en; ++u) {
            const uint32_t du = V.a0[u];
            rank += (du < d || (du == d && static_cast<int>(V.t[u]) < t)) ? 1 : 0;
        }
        V.p1[s] = static_cast<uint16_t>(st + rank);
    }
    __syncthreads();
    // phase 4 low word
    for (int s = tid; s < n_last; s += kNicheThreads) {
        const int r = V.r[s];
        const int st = r ? half_get(table, r - 1) : 0;
        const int p1 = V.p1[s];
        const int rh = (r < k) ? A.rho[r] : 0;
        V.a0[s] = (rh == 0 && p1 == st) ? 0u : (niche_draw(A.inj_a, 0, p1, A.seed, A.gen) | 1u);
    }
    __syncthreads();
    // sort 2: rank by (low, p1) inside the niche; phase-5 level and draw
    auto sort2 = [&](int s, int& lev, uint32_t& db, int& p2) {
        const int r = V.r[s];
        const int st = r ? half_get(table, r - 1) : 0, en = half_get(table, r);
        const uint32_t lo = V.a0[s];
        const int p1 = V.p1[s];
        int rank = 0;
        for (int u = st; u < en; ++u) {
            const uint32_t lu = V.a0[u];
            rank += (lu < lo || (lu == lo && static_cast<int>(V.p1[u]) < p1)) ? 1 : 0;
        }
        const int rh = (r < k) ? A.rho[r] : 0;
        p2 = st + rank;
        lev = rh + rank;
        db = niche_draw(A.inj_b, 1, p2, A.seed, A.gen);
    };
    if (kShared) {
        int lev[kBucketRegs], p2[kBucketRegs];
        uint32_t db[kBucketRegs];
#pragma unroll
        for (int j = 0; j < kBucketRegs; ++j) {
            const int s = tid + j * kNicheThreads;
            if (s < n_last) sort2(s, lev[j], db[j], p2[j]);
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kBucketRegs; ++j) {
            const int s = tid + j * kNicheThreads;
            if (s < n_last) {
                V.lev[s] = static_cast<uint16_t>(lev[j]);
                V.db[s] = db[j];
                V.p2[s] = static_cast<uint16_t>(p2[j]);
            }
        }
    } else {
        for (int s = tid; s < n_last; s += kNicheThreads) {
            int lev, p2;
            uint32_t db;
            sort2(s, lev, db, p2);
            V.lev[s] = static_cast<uint16_t>(lev);
            V.db[s] = db;
            V.p2[s] = static_cast<uint16_t>(p2);
        }
    }
    __syncthreads();
    // sort 3 as a selection: the n_rem-th smallest of the distinct keys
    // (lev << 48 | draw_b << 16 | p2), 8 bits a round from the first byte in
    // which the keys differ, until one candidate is left
    const int take = n_rem < n_last ? n_rem : n_last;
    int32_t* hist = misc;              // 256
    int32_t* pick = misc + 256;        // digit, count below, its count
    uint64_t* ext = reinterpret_cast<uint64_t*>(misc + 264);   // per wave min, max
    uint64_t kmin = ~0ull, kmax = 0;
    for (int s = tid; s < n_last; s += kNicheThreads) {
        const uint64_t key = (static_cast<uint64_t>(V.lev[s]) << 48) | (static_cast<uint64_t>(V.db[s]) << 16) |
                             static_cast<uint64_t>(V.p2[s]);
        kmin = key < kmin ? key : kmin;
        kmax = key > kmax ? key : kmax;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const uint64_t a = __shfl_xor(kmin, o), b = __shfl_xor(kmax, o);
        kmin = a < kmin ? a : kmin;
        kmax = b > kmax ? b : kmax;
    }
    if ((tid & 63) == 0) {
        ext[2 * (tid >> 6)] = kmin;
        ext[2 * (tid >> 6) + 1] = kmax;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kNicheWaves; ++j) {
        const uint64_t a = ext[2 * j], b = ext[2 * j + 1];
        kmin = a < kmin ? a : kmin;
        kmax = b > kmax ? b : kmax;
    }
    const uint64_t diff = kmin ^ kmax;
    const int top = diff ? (63 - __builtin_clzll(diff)) / 8 * 8 : -8;   // highest byte that differs
    uint64_t mask = (top >= 56) ? 0ull : (~0ull << (top + 8));
    uint64_t prefix = kmin & mask;
    int want = take;                   // rank of the target among the candidates (1-based)
    for (int sh = top; sh >= 0; sh -= 8) {
        if (tid < 256) hist[tid] = 0;
        __syncthreads();
        for (int s = tid; s < n_last; s += kNicheThreads) {
            const uint64_t key = (static_cast<uint64_t>(V.lev[s]) << 48) | (static_cast<uint64_t>(V.db[s]) << 16) |
                                 static_cast<uint64_t>(V.p2[s]);
            if ((key & mask) == prefix) atomicAdd(&hist[(key >> sh) & 255u], 1);
        }
        __syncthreads();
        if (tid < 64) {
            int c[4];
            int sum = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                c[j] = hist[tid * 4 + j];
                sum += c[j];
            }
            int incl = sum;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int u = __shfl_up(incl, o);
                if (tid >= o) incl += u;
            }
            int before = incl - sum;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (before < want && want <= before + c[j]) {
                    pick[0] = tid * 4 + j;
                    pick[1] = before;
                    pick[2] = c[j];
                }
                before += c[j];
            }
        }
        __syncthreads();
        prefix |= static_cast<uint64_t>(pick[0]) << sh;
        mask |= 255ull << sh;
        want -= pick[1];
        const int left = pick[2];
        __syncthreads();
        if (left == 1) break;          // the target is the one candidate left
    }
    // chosen: every key whose masked value does not exceed the target's (with
    // one candidate left, or the full mask, exactly the `take` smallest)
    for (int s = tid; s < n_last; s += kNicheThreads) {
        const uint64_t key = (static_cast<uint64_t>(V.lev[s]) << 48) | (static_cast<uint64_t>(V.db[s]) << 16) |
                             static_cast<uint64_t>(V.p2[s]);
        if ((key & mask) <= prefix) A.flag[A.gidx[n_prior + V.t[s]]] = 1;
    }
}

// The last front's niching (rho must hold the prior-front counts).
__device__ __forceinline__ void niche_last(const NicheArgs& A, const int32_t* __restrict__ meta, int nm,
                                           uint64_t* gk0, uint64_t* gk1, int32_t* gv0, int32_t* gv1, int lds_cap,
                                           unsigned char* s_raw, int* s_wave)
{
    const int n_rem = meta[kNRem];
    const int n_last = meta[kNLast];
    const int n_prior = meta[kNPrior];
    const int tid = static_cast<int>(threadIdx.x);
    const int k = A.k;
    // bucket table: k + 1 niches, 16-bit counts, two a word, wpt consecutive
    // words a thread (a multiple of 4); then misc, then slots
    const int nwords = (k + 2) / 2;                     // niches 0 .. k
    const int wpt = ((nwords + kNicheThreads - 1) / kNicheThreads + 3) & ~3;
    const int words = wpt * kNicheThreads;
    const int misc_ints = 512;
    const int fixed = 4 * words + 4 * misc_ints;
    int cap_b = (fixed < kNicheLdsBytes) ? ((kNicheLdsBytes - fixed) / kBucketSlotBytes) & ~15 : 0;
    cap_b = min(cap_b, kBucketRegs * kNicheThreads);
    if (A.bucket && fixed <= kNicheLdsBytes && n_last <= 65535 && n_prior <= 65535 - kBucketMaxRun) {
        uint32_t* table = reinterpret_cast<uint32_t*>(s_raw);
        int32_t* misc = reinterpret_cast<int32_t*>(s_raw + 4 * words);
        for (int j = tid; j < words; j += kNicheThreads) table[j] = 0;
        __syncthreads();
        for (int t = tid; t < n_last; t += kNicheThreads) {
            const int r = static_cast<int>(ref_word(A.ref[n_prior + t], k));
            atomicAdd(&table[r >> 1], (r & 1) ? 0x10000u : 1u);
        }
        __syncthreads();
        // niche starts in place: thread t owns words [t wpt, (t + 1) wpt)
        const int lane = tid & 63, w = tid >> 6;
        uint4* mine = reinterpret_cast<uint4*>(table + tid * wpt);
        int sum = 0, run_max = 0;
        for (int j = 0; j < wpt / 4; ++j) {
            const uint4 v = mine[j];
            const uint32_t e[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int c0 = static_cast<int>(e[i] & 0xFFFFu), c1 = static_cast<int>(e[i] >> 16);
                sum += c0 + c1;
                run_max = max(run_max, max(c0, c1));
            }
        }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) run_max = max(run_max, __shfl_xor(run_max, o));
        if (lane == 0) misc[w] = run_max;
        int total;
        int at = block_exclusive_scan(sum, s_wave, total);   // its barriers also publish misc
        int big = 0;
#pragma unroll
        for (int j = 0; j < kNicheWaves; ++j) big = max(big, misc[j]);
        for (int j = 0; j < wpt / 4; ++j) {
            const uint4 v = mine[j];
            uint32_t e[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int c0 = static_cast<int>(e[i] & 0xFFFFu), c1 = static_cast<int>(e[i] >> 16);
                e[i] = static_cast<uint32_t>(at) | (static_cast<uint32_t>(at + c0) << 16);
                at += c0 + c1;
            }
            mine[j] = make_uint4(e[0], e[1], e[2], e[3]);
        }
        __syncthreads();
        if (big <= kBucketMaxRun) {
            if (n_last <= cap_b && n_last <= lds_cap)
                bucket_finish<true>(A, table, misc, bucket_slots(s_raw + fixed, cap_b), n_prior, n_last, n_rem);
            else
                bucket_finish<false>(A, table, misc, bucket_slots_global(gk0, gk1, gv0, nm), n_prior, n_last, n_rem);
            return;
        }
    }
    NicheShared& S = *reinterpret_cast<NicheShared*>(s_raw);
    if (tid == 0) S.skip = -1;
    __syncthreads();
    if (n_last <= lds_cap && n_last <= kNicheLdsCap) niche_body(NicheLdsBufs{&S}, A, S, n_prior, n_last, n_rem);
    else niche_body(NicheGlobalBufs{gk0, gk1, gv0, gv1}, A, S, n_prior, n_last, n_rem);
}

// Survivor flags (the prior fronts, then the niching picks), the niching, and
// the survivors in ascending merged order (gather_selected_kernel,
// niching.cu:192-206, prior copy :303-312; thrust::sort of the selection,
// nsga3_population_update.cu:269), in one workgroup.
// fronts == nullptr: the prior-front flags are already written (front_scatter).
__device__ __forceinline__ void niche_select_body(const NicheArgs& A, int32_t* __restrict__ meta, int nm, int n,
                                                  const int32_t* __restrict__ fronts, int32_t* __restrict__ chosen,
                                                  uint64_t* gk0, uint64_t* gk1, int32_t* gv0, int32_t* gv1,
                                                  int lds_cap, unsigned char* s_raw, int* s_wave)
{
    const int tid = static_cast<int>(threadIdx.x);
    const int lane = tid & 63, w = tid >> 6;
    const int mf = meta[kMaxFront];
    // prior-front flags, four individuals a thread (nm = 2n is a multiple of 4)
    const int nq = nm / 4;
    uint32_t* fw = reinterpret_cast<uint32_t*>(A.flag);
    constexpr int kRows = 16;          // loads in flight per thread
    for (int q0 = tid; fronts && q0 < nq; q0 += kRows * kNicheThreads) {
        int4 fr[kRows];
#pragma unroll
        for (int u = 0; u < kRows; ++u) {
            const int q = q0 + u * kNicheThreads;
            fr[u] = (q < nq) ? reinterpret_cast<const int4*>(fronts)[q] : make_int4(-1, -1, -1, -1);
        }
#pragma unroll
        for (int u = 0; u < kRows; ++u) {
            const int q = q0 + u * kNicheThreads;
            const int4 f = fr[u];
            if (q < nq)
                fw[q] = (prior_survivor(f.x, mf, 4 * q, n) ? 1u : 0u) |
                        (prior_survivor(f.y, mf, 4 * q + 1, n) ? 0x100u : 0u) |
                        (prior_survivor(f.z, mf, 4 * q + 2, n) ? 0x10000u : 0u) |
                        (prior_survivor(f.w, mf, 4 * q + 3, n) ? 0x1000000u : 0u);
        }
    }
    __syncthreads();
    if (meta[kNRem] > 0 && meta[kNLast] > 0) niche_last(A, meta, nm, gk0, gk1, gv0, gv1, lds_cap, s_raw, s_wave);
    __syncthreads();
    // survivors in ascending order: wave w takes a run of flag words in rows of
    // 64 consecutive words, one word (four flags) a lane
    const int per_w = ((nq + kNicheWaves - 1) / kNicheWaves + 63) / 64 * 64;
    const int q0 = min(nq, w * per_w), q1 = min(nq, q0 + per_w);
    int cnt = 0;
    for (int qr = q0; qr < q1; qr += kRows * 64) {
        uint32_t v[kRows];
#pragma unroll
        for (int u = 0; u < kRows; ++u) {
            const int q = qr + u * 64 + lane;
            v[u] = (q < q1) ? fw[q] : 0u;
        }
#pragma unroll
        for (int u = 0; u < kRows; ++u)
#pragma unroll
            for (int j = 0; j < 4; ++j) cnt += __popcll(ballot(((v[u] >> (8 * j)) & 1u) != 0));
    }
    if (lane == 0) s_wave[w] = cnt;
    __syncthreads();
    int base = 0, total = 0;
#pragma unroll
    for (int j = 0; j < kNicheWaves; ++j) {
        if (j < w) base += s_wave[j];
        total += s_wave[j];
    }
    for (int qb = q0; qb < q1; qb += kRows * 64) {
        uint32_t vv[kRows];
#pragma unroll
        for (int u = 0; u < kRows; ++u) {
            const int q = qb + u * 64 + lane;
            vv[u] = (q < q1) ? (fw[q] & 0x01010101u) : 0u;
        }
        const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
        for (int u = 0; u < kRows; ++u) {
            const int q = qb + u * 64 + lane;
            const uint32_t v = vv[u];
            int o = base, row = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint64_t b = ballot(((v >> (8 * j)) & 1u) != 0);
                o += __popcll(b & below);
                row += __popcll(b);
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if ((v >> (8 * j)) & 1u) {
                    if (o < n) chosen[o] = 4 * q + j;
                    ++o;
                }
            }
            base += row;
        }
    }
    if (tid == 0) meta[5] = total;
}

// stats (optional): the selection's meta words, as write_stats.
__global__ __launch_bounds__(kNicheThreads) void niche_block(NicheArgs A, int32_t* __restrict__ meta, int nm, int n,
                                                             const int32_t* __restrict__ fronts,
                                                             int32_t* __restrict__ chosen, uint64_t* gk0,
                                                             uint64_t* gk1, int32_t* gv0, int32_t* gv1, int lds_cap,
                                                             int32_t* __restrict__ stats)
{
    __shared__ __align__(16) unsigned char s_raw[kNicheLdsBytes];
    __shared__ int s_wave[64];
    niche_select_body(A, meta, nm, n, fronts, chosen, gk0, gk1, gv0, gv1, lds_cap, s_raw, s_wave);
    if (stats && threadIdx.x == 0) copy_stats(meta, stats);
}

// Chosen indices in ascending merged order (thrust::sort of the selection,
// nsga3_population_update.cu:269): the same two passes as front_counts /
// front_scatter over the survivor flags.
__global__ __launch_bounds__(kScanThreads) void flag_counts(const uint8_t* __restrict__ flag, int nm,
                                                            int32_t* __restrict__ counts)
{
    __shared__ int s_wave[64];
    const int i = blockIdx.x * kScanThreads + threadIdx.x;
    int tot;
    block_exclusive_scan(i < nm ? static_cast<int>(flag[i]) : 0, s_wave, tot);
    if (threadIdx.x == 0) counts[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kScanThreads) void flag_scatter(const uint8_t* __restrict__ flag, int nm,
                                                             int n, const int32_t* __restrict__ counts,
                                                             int nblk, int32_t* __restrict__ chosen,
                                                             int32_t* __restrict__ meta)
{
    __shared__ int s_wave[64];
    const int b = blockIdx.x;
    int before = 0, all = 0;
    for (int j = threadIdx.x; j < nblk; j += kScanThreads) {
        const int c = counts[j];
        all += c;
        if (j < b) before += c;
    }
    block_exclusive_scan(before, s_wave, before);
    block_exclusive_scan(all, s_wave, all);
    const int i = b * kScanThreads + threadIdx.x;
    const int f = (i < nm) ? static_cast<int>(flag[i]) : 0;
    int t;
    const int o = before + block_exclusive_scan(f, s_wave, t);
    if (f && o < n) chosen[o] = i;
    if (b == 0 && threadIdx.x == 0) meta[5] = all;
}

// Survivor gather (pop_update_*_kernel + fit_and_fronts_update_kernel,
// nsga3_population_update.cu:119-225); one wave per row.
__device__ __forceinline__ void gather_row(const int32_t* __restrict__ chosen, int n, int d, int m,
                                  const float* __restrict__ px, const float* __restrict__ ox,
                                  const float* __restrict__ fm, const float* __restrict__ cvm,
                                  const int32_t* __restrict__ fronts_mix, float* __restrict__ nx,
                                  float* __restrict__ nf, float* __restrict__ ncv,
                                  int32_t* __restrict__ nfronts, int row, int lane)
{
    int s = chosen[row];
    if (s < 0 || s >= 2 * n) s = 0;  // unreachable: exactly n survivors are flagged
    const float* src = (s < n) ? px + static_cast<size_t>(s) * d : ox + static_cast<size_t>(s - n) * d;
    float* dst = nx + static_cast<size_t>(row) * d;
    if ((d & 3) == 0) {
        // four 16-byte loads in flight per lane before their stores
        const float4* s4 = reinterpret_cast<const float4*>(src);
        float4* d4 = reinterpret_cast<float4*>(dst);
        const int d4n = d >> 2;
        int j = lane;
        for (; j + 192 < d4n; j += 256) {
            const float4 a = s4[j], b = s4[j + 64], c = s4[j + 128], e = s4[j + 192];
            d4[j] = a;
            d4[j + 64] = b;
            d4[j + 128] = c;
            d4[j + 192] = e;
        }
        for (; j < d4n; j += 64) d4[j] = s4[j];
    } else {
        for (int j = lane; j < d; j += 64) dst[j] = src[j];
    }
    const int nm = 2 * n;
    if (lane < m) nf[static_cast<size_t>(lane) * n + row] = fm[static_cast<size_t>(lane) * nm + s];
    if (lane == 0) {
        ncv[row] = cvm[s];
        nfronts[row] = fronts_mix[s];
    }
}

__global__ __launch_bounds__(256) void gather_next(const int32_t* __restrict__ chosen, int n, int d,
                                                   int m, const float* __restrict__ px,
                                                   const float* __restrict__ ox,
                                                   const float* __restrict__ fm,
                                                   const float* __restrict__ cvm,
                                                   const int32_t* __restrict__ fronts_mix,
                                                   float* __restrict__ nx, float* __restrict__ nf,
                                                   float* __restrict__ ncv,
                                                   int32_t* __restrict__ nfronts)
{
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= n) return;
    gather_row(chosen, n, d, m, px, ox, fm, cvm, fronts_mix, nx, nf, ncv, nfronts, row,
               static_cast<int>(threadIdx.x & 63));
}

__global__ void write_stats(const int32_t* __restrict__ meta, int32_t* __restrict__ stats)
{
    if (threadIdx.x == 0) copy_stats(meta, stats);
}

// quantize_cv / cv_quantize_kernel (selection/nsga3/kernels/cv_quantization.cu:8-113):
// the bin of each violation, 0 when feasible.
__global__ void quantize_cv_kernel(const float* __restrict__ cv, float* __restrict__ out, int n, int bins,
                                   float clip, float inv_clip, float log_alpha, float inv_clip_log, float eps)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float raw = cv[i];
    const float u = fminf(fmaxf(raw, 0.0f), clip);
    const float x = (log_alpha > 0.0f) ? log1pf(log_alpha * u) * inv_clip_log : u * inv_clip;
    int bin = 1 + static_cast<int>(floorf(x * static_cast<float>(bins)));
    bin = max(1, min(bins, bin));
    const bool infeasible = (eps == 0.0f) ? (raw != 0.0f) : (raw > eps);
    out[i] = infeasible ? static_cast<float>(bin) : 0.0f;
}

struct Layout {
    float* fm; float* cvm; float* cvq; int32_t* fronts; int32_t* meta; int32_t* gidx; float* fn;
    float* asf_val; int32_t* asf_idx; float* row_max; float* a; float* omin;
    float* part_d2; int32_t* part_idx; float* dist; int32_t* ref; int32_t* rho; ConeBufs cones;
    uint64_t* k0; uint64_t* k1; int32_t* v0; int32_t* v1; void* sort_tmp; size_t sort_bytes;
    uint8_t* flag; int32_t* chosen; int32_t* counts; void* nds_ws; size_t nds_bytes;
};

Layout carve(void* ws, size_t cap, int n, int m, int k, Carve& c)
{
    const int nm = 2 * n;
    const int nb = ceil_div(nm, 256);
    const int nchunks = ceil_div(k, kAssocChunk);
    Layout L{};
    L.fm = c.take<float>(static_cast<size_t>(m) * nm);
    L.cvm = c.take<float>(nm);
    L.cvq = c.take<float>(nm);
    L.fronts = c.take<int32_t>(nm);
    L.meta = c.take<int32_t>(kMetaLen);
    L.gidx = c.take<int32_t>(nm);
    L.fn = c.take<float>(static_cast<size_t>(m) * nm);
    L.asf_val = c.take<float>(static_cast<size_t>(nb) * m);
    L.asf_idx = c.take<int32_t>(static_cast<size_t>(nb) * m);
    L.row_max = c.take<float>(static_cast<size_t>(nb) * m);
    L.a = c.take<float>(kMaxM);
    L.omin = c.take<float>(static_cast<size_t>(nb) * m);
    L.part_d2 = c.take<float>(static_cast<size_t>(nchunks) * nm);
    L.part_idx = c.take<int32_t>(static_cast<size_t>(nchunks) * nm);
    L.dist = c.take<float>(nm);
    L.ref = c.take<int32_t>(nm);
    L.rho = c.take<int32_t>(k);
    L.cones = carve_cones(c, m, k);
    L.k0 = c.take<uint64_t>(nm);
    L.k1 = c.take<uint64_t>(nm);
    L.v0 = c.take<int32_t>(nm);
    L.v1 = c.take<int32_t>(nm);
    L.sort_bytes = sort_pairs_temp_bytes(nm);
    L.sort_tmp = c.take_bytes(L.sort_bytes);
    L.flag = c.take<uint8_t>(nm);
    L.chosen = c.take<int32_t>(n);
    L.counts = c.take<int32_t>(2 * static_cast<size_t>(ceil_div(nm, kScanThreads)));
    L.nds_bytes = nds_workspace_bytes(nm);
    L.nds_ws = c.take_bytes(L.nds_bytes);
    (void)ws; (void)cap;
    return L;
}

}  // namespace

// Automatic pruning: at most 3 objectives (the cones of higher-dimensional
// direction sets leave too many candidate blocks to beat the tiled kernel).
bool assoc_prunable(int k, int m)
{
    return m <= kConeAutoMaxM && k >= kConeMinK && ceil_div(k, 64) <= kConeMaxBlocks;
}

size_t cone_bytes(int m, int k)
{
    Carve c(nullptr, 0);
    carve_cones(c, m, k);
    return c.used;
}

void cone_views(void* buf, int m, int k, float** wp, int32_t** perm, float** cones)
{
    Carve c(buf, ~size_t{0});
    const ConeBufs B = carve_cones(c, m, k);
    *wp = B.wp;
    *perm = B.perm;
    *cones = B.cones;
}

int cones_build(const float* w, int k, int m, void* buf, hipStream_t st)
{
    if (!w || !buf || m <= 0 || m > kMaxM || !assoc_prunable(k, m))
        return fail(kInvalidArgument, "cones_build: invalid");
    Carve c(buf, ~size_t{0});
    return cones_prepare(w, k, m, carve_cones(c, m, k), st);
}

bool cv_quant_valid(const moea_cv_quantization& q)
{
    // validate_cv_quant_params (cv_quantization.cu:93-112)
    if (q.bins <= 0) return false;
    if (!(q.clip_upper > 0.0f) || !std::isfinite(q.clip_upper)) return false;
    if (q.log_alpha < 0.0f || !std::isfinite(q.log_alpha)) return false;
    if (q.feasibility_epsilon < 0.0f || !std::isfinite(q.feasibility_epsilon)) return false;
    if (q.log_alpha > 0.0f) {
        const float clip_log = std::log1p(q.log_alpha * q.clip_upper);
        if (!(clip_log > 0.0f) || !std::isfinite(clip_log)) return false;
    }
    return true;
}

int quantize_cv(const float* cv, float* out, int n, const moea_cv_quantization& q, hipStream_t st)
{
    if (!cv || !out || n <= 0) return fail(kInvalidArgument, "quantize_cv: invalid arguments");
    if (!cv_quant_valid(q)) return fail(kInvalidArgument, "Invalid NSGA-III CV quantization configuration");
    const float inv_clip = 1.0f / q.clip_upper;
    const float inv_clip_log = (q.log_alpha > 0.0f) ? 1.0f / std::log1p(q.log_alpha * q.clip_upper) : 0.0f;
    hipLaunchKernelGGL(quantize_cv_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, st, cv, out, n, q.bins,
                       q.clip_upper, inv_clip, q.log_alpha, inv_clip_log, q.feasibility_epsilon);
    MOEA_HIP(hipGetLastError());
    return kOk;
}

size_t nsga3_workspace_bytes(int n, int m, int k)
{
    Carve c(nullptr, 0);
    carve(nullptr, 0, n, m, k, c);
    return c.used;
}

namespace {

__global__ void assoc_meta(int32_t* __restrict__ meta, int n)
{
    if (threadIdx.x == 0) {
        meta[kMaxFront] = 0;
        meta[kNPrior] = n;
        meta[kNLast] = 0;
        meta[kNRem] = 1;
    }
}

struct AssocLayout {
    int32_t* meta; ConeBufs cones; float* part_d2; int32_t* part_idx;
};

AssocLayout carve_assoc(Carve& c, int n, int m, int k)
{
    AssocLayout L{};
    L.meta = c.take<int32_t>(kMetaLen);
    L.cones = carve_cones(c, m, k);
    L.part_d2 = c.take<float>(static_cast<size_t>(ceil_div(k, kAssocChunk)) * n);
    L.part_idx = c.take<int32_t>(static_cast<size_t>(ceil_div(k, kAssocChunk)) * n);
    return L;
}

}  // namespace

// The brute-force partials are strided by the leading dimension: size by ld.
size_t associate_workspace_bytes(int ld, int m, int k)
{
    if (ld <= 0 || m <= 0 || m > kMaxM || k <= 0) return 0;
    Carve c(nullptr, 0);
    carve_assoc(c, ld, m, k);
    return c.used;
}

int associate(const float* fn, int ld, int n, int m, const float* a, const float* w, int k, float* dist,
              int32_t* ref, int method, void* ws, size_t ws_bytes, hipStream_t st)
{
    if (!fn || !a || !w || !dist || !ref || n <= 0 || m <= 0 || m > kMaxM || k <= 0 || ld < n ||
        method < 0 || method > 2)
        return fail(kInvalidArgument, "associate: invalid arguments");
    const size_t need = associate_workspace_bytes(ld, m, k);
    if (!ws || ws_bytes < need) return fail(kWorkspaceTooSmall, "associate: workspace too small");
    Carve c(ws, ws_bytes);
    const AssocLayout L = carve_assoc(c, ld, m, k);
    const int nbk = ceil_div(k, 64);
    const bool prunable = nbk <= kConeMaxBlocks;
    if (method == 2 && !prunable) return fail(kUnsupported, "associate: too many directions for the pruned path");
    const bool pruned = (method == 2) || (method == 0 && assoc_prunable(k, m));
    hipLaunchKernelGGL(assoc_meta, dim3(1), dim3(64), 0, st, L.meta, n);
    const dim3 b256(256);
    if (pruned) {
        MOEA_CHECK(cones_prepare(w, k, m, L.cones, st));
        MOEA_LAUNCH_M(m, associate_pruned, dim3(ceil_div(n, 4)), b256, 0, st, fn, ld, a, L.cones.wp, k,
                      L.cones.perm, L.cones.cones, L.meta, dist, ref, nullptr);
    } else {
        const int nchunks = ceil_div(k, kAssocChunk);
        MOEA_LAUNCH_M(m, associate_partial, dim3(ceil_div(n, 256), nchunks), b256, 0, st, fn, ld, a, w, k, L.meta,
                      L.part_d2, L.part_idx);
        hipLaunchKernelGGL(associate_final, dim3(ceil_div(n, 256)), b256, 0, st, L.part_d2, L.part_idx, nchunks, ld,
                           L.meta, dist, ref, nullptr);
    }
    MOEA_HIP(hipGetLastError());
    return kOk;
}

int ideal_point(const float* f, int ldf, int n, int m, float* ideal, hipStream_t stream)
{
    if (!f || !ideal || n <= 0 || m <= 0 || m > kMaxM) return fail(kInvalidArgument, "ideal_point: invalid arguments");
    hipLaunchKernelGGL(ideal_update, dim3(m), dim3(1024), 0, stream, f, ldf, 0, n, m, ideal, 1);
    MOEA_HIP(hipGetLastError());
    return kOk;
}

// Niching path (moea_niching): 0 automatic (one block), 1 one block, 2
// device-wide sorts, 3 one block with its radix sorts only.
bool niching_valid(const moea_niching& c) { return c.mode >= 0 && c.mode <= 3 && c.lds_cap >= 0; }

namespace {
bool niching_one_block(const moea_niching& c) { return c.mode != 2; }
int niching_buckets(const moea_niching& c) { return c.mode == 3 ? 0 : 1; }
// each path also stops at its own LDS capacity
int niching_lds_cap(const moea_niching& c) { return c.lds_cap > 0 ? c.lds_cap : (1 << 30); }
constexpr int kBatchNicheMaxNm = 8192;
}  // namespace

int nsga3_select(const moea_nsga3_select_args& A)
{
    const int n = A.n, m = A.m, d = A.d, k = A.k;
    if (n <= 0 || (n & 1) || m <= 0 || m > kMaxM || d <= 0 || k <= 0)
        return fail(kInvalidArgument, "nsga3_select: invalid sizes (n must be even and positive)");
    if (!niching_valid(A.niching)) return fail(kInvalidArgument, "nsga3_select: invalid niching mode or capacity");
    if (!A.parents.x || !A.parents.f || !A.parents.cv || !A.offspring.x || !A.offspring.f ||
        !A.offspring.cv || !A.next.x || !A.next.f || !A.next.cv || !A.next_fronts ||
        !A.directions || !A.ideal || !A.workspace)
        return fail(kInvalidArgument, "nsga3_select: null pointer");
    if (A.next.x == A.parents.x || A.next.x == A.offspring.x)
        return fail(kInvalidArgument, "nsga3_select: next must not alias parents or offspring");
    if (2 * n > nds_max_n()) return fail(kUnsupported, "nsga3_select: population too large");
    Carve c(A.workspace, A.workspace_bytes);
    const Layout L = carve(A.workspace, A.workspace_bytes, n, m, k, c);
    if (!c.ok()) return fail(kWorkspaceTooSmall, "nsga3_select: workspace too small");
    hipStream_t st = static_cast<hipStream_t>(A.stream);
    const int nm = 2 * n;
    const int nb = ceil_div(nm, 256);
    const int nchunks = ceil_div(k, kAssocChunk);
    const uint32_t gen = static_cast<uint32_t>(A.generation);
    const moea_draws* dr = A.draws;

    // without quantized violations the merge also writes the sort keys
    const bool keys_in_merge = A.cv_quant.bins <= 0;
    uint64_t* sort_keys = nullptr;
    int32_t* sort_vals = nullptr;
    if (keys_in_merge) nds_key_buffers(L.nds_ws, nm, &sort_keys, &sort_vals);
    hipLaunchKernelGGL(merge_objectives, dim3(nb), dim3(256), 0, st, A.parents.f, A.offspring.f,
                       A.parents.cv, A.offspring.cv, n, m, L.fm, L.cvm, L.omin, sort_keys, sort_vals);
    MOEA_HIP(hipGetLastError());
    const float* cv_sort = L.cvm;
    if (A.cv_quant.bins > 0) {
        // the sort compares bins; the survivors keep their violations
        MOEA_CHECK(quantize_cv(L.cvm, L.cvq, nm, A.cv_quant, st));
        cv_sort = L.cvq;
    }
    MOEA_CHECK(nds_run(L.fm, nm, cv_sort, nm, m, n, L.fronts, L.meta + kMaxFront, L.nds_ws, L.nds_bytes, st,
                       A.stats ? reinterpret_cast<uint32_t*>(A.stats + kDevError) : nullptr, keys_in_merge));
    stage_begin(kStSelectPrep, st);
    const int nsb = ceil_div(nm, kScanThreads);
    hipLaunchKernelGGL(front_counts, dim3(nsb), dim3(kScanThreads), 0, st, L.fronts, nm, L.meta, L.counts);
    hipLaunchKernelGGL(front_scatter, dim3(nsb), dim3(kScanThreads), 0, st, L.fronts, nm, n, L.counts, nsb,
                       L.gidx, L.meta, L.flag);
    const dim3 g_nb(nb), b256(256), g_assoc(nb, nchunks);
    MOEA_LAUNCH_M(m, norm_prep, g_nb, b256, 0, st, L.fm, nm, L.gidx, L.meta, L.fn, L.asf_val, L.asf_idx,
                  L.row_max, L.rho, k, L.omin, nb, A.ideal);
    hipLaunchKernelGGL(norm_solve, dim3(1), dim3(kSolveThreads), 0, st, L.fn, nm, m, L.asf_val,
                       L.asf_idx, L.row_max, L.meta, L.a, L.gidx, A.out_extremes, A.out_intercepts);
    stage_end(kStSelectPrep, st);
    stage_begin(kStAssociation, st);
    if (assoc_prunable(k, m)) {
        if (!A.directions_ready) MOEA_CHECK(cones_prepare(A.directions, k, m, L.cones, st));
        MOEA_LAUNCH_M(m, associate_pruned, dim3(ceil_div(nm, 4)), b256, 0, st, L.fn, nm, L.a, L.cones.wp, k,
                      L.cones.perm, L.cones.cones, L.meta, L.dist, L.ref, L.rho);
    } else {
        MOEA_LAUNCH_M(m, associate_partial, g_assoc, b256, 0, st, L.fn, nm, L.a, A.directions, k,
                      L.meta, L.part_d2, L.part_idx);
        hipLaunchKernelGGL(associate_final, dim3(nb), dim3(256), 0, st, L.part_d2, L.part_idx, nchunks,
                           nm, L.meta, L.dist, L.ref, L.rho);
    }
    stage_end(kStAssociation, st);
    stage_begin(kStNiching, st);
    const int bits_ref = 32 + (32 - __builtin_clz(static_cast<unsigned>(k + 1)));
    const int bits_rank = 32 + (32 - __builtin_clz(static_cast<unsigned>(nm)));
    if (niching_one_block(A.niching)) {
        NicheArgs na{L.ref, L.dist, L.rho, k, bits_ref, bits_rank, dr ? dr->niche_a : nullptr,
                     dr ? dr->niche_b : nullptr, A.seed, gen, L.gidx, L.flag, niching_buckets(A.niching)};
        hipLaunchKernelGGL(niche_block, dim3(1), dim3(kNicheThreads), 0, st, na, L.meta, nm, n, nullptr, L.chosen,
                           L.k0, L.k1, L.v0, L.v1, niching_lds_cap(A.niching), A.stats);
    } else {
        hipLaunchKernelGGL(niche_keys1, dim3(nb), dim3(256), 0, st, L.ref, L.dist, L.meta, nm, k, L.k0, L.v0);
        MOEA_HIP(hipGetLastError());
        MOEA_CHECK(sort_pairs(L.sort_tmp, L.sort_bytes, L.k0, L.k1, L.v0, L.v1, nm, st, bits_ref));
        hipLaunchKernelGGL(niche_keys2, dim3(nb), dim3(256), 0, st, L.k1, L.meta, L.rho, k, nm,
                           dr ? dr->niche_a : nullptr, A.seed, gen, L.k0);
        MOEA_HIP(hipGetLastError());
        MOEA_CHECK(sort_pairs(L.sort_tmp, L.sort_bytes, L.k0, L.k1, L.v1, L.v0, nm, st, bits_ref));
        hipLaunchKernelGGL(niche_keys3, dim3(nb), dim3(256), 0, st, L.k1, L.meta, L.rho, k, nm,
                           dr ? dr->niche_b : nullptr, A.seed, gen, L.k0);
        MOEA_HIP(hipGetLastError());
        MOEA_CHECK(sort_pairs(L.sort_tmp, L.sort_bytes, L.k0, L.k1, L.v0, L.v1, nm, st, bits_rank));
        hipLaunchKernelGGL(mark_selected, dim3(ceil_div(n, 256)), dim3(256), 0, st, L.v1, L.gidx, L.meta,
                           L.flag);
        hipLaunchKernelGGL(flag_counts, dim3(nsb), dim3(kScanThreads), 0, st, L.flag, nm, L.counts);
        hipLaunchKernelGGL(flag_scatter, dim3(nsb), dim3(kScanThreads), 0, st, L.flag, nm, n, L.counts, nsb,
                           L.chosen, L.meta);
    }
    stage_end(kStNiching, st);
    stage_begin(kStGather, st);
    hipLaunchKernelGGL(gather_next, dim3(ceil_div(n, 4)), dim3(256), 0, st, L.chosen, n, d, m,
                       A.parents.x, A.offspring.x, L.fm, L.cvm, L.fronts, A.next.x, A.next.f,
                       A.next.cv, A.next_fronts);
    stage_end(kStGather, st);
    if (A.stats && !niching_one_block(A.niching)) hipLaunchKernelGGL(write_stats, dim3(1), dim3(64), 0, st, L.meta, A.stats);
    MOEA_HIP(hipGetLastError());
    return kOk;
}

namespace {

// ---------------------------------------------------------------- batched --
// The selection of every instance of a batch, each through the same bodies as
// the single-population kernels.  Sort keys carry the instance index above
// their own bits, so each global stable sort orders every instance's entries
// exactly as that instance's own sort would.

__global__ void b_merge(const BInst* __restrict__ insts, BMap mp)
{
    int local;
    const int i = bmap_find(mp, static_cast<int>(blockIdx.x), local);
    const BInst& I = insts[i];
    const int cur = I.cur, off = (I.cur + 1) % 3;
    const int t = local * 256 + static_cast<int>(threadIdx.x);
    if (t >= I.nm) return;
    const bool par = t < I.n;
    const int s = par ? t : t - I.n;
    const float* pf = par ? I.f[cur] : I.f[off];
    for (int k = 0; k < I.m; ++k) I.fm[static_cast<size_t>(k) * I.nm + t] = pf[static_cast<size_t>(k) * I.n + s];
    I.cvm[t] = par ? I.cv[cur][s] : I.cv[off][s];
}

// prior / last split of one instance per workgroup: a counting pass, then a
// scatter pass, each over chunks of kScanThreads entries.
__global__ __launch_bounds__(kScanThreads) void b_split(const BInst* __restrict__ insts)
{
    __shared__ int s_wave[64];
    const BInst& I = insts[blockIdx.x];
    if (I.state != 1) return;
    const int mf = I.meta[kMaxFront];
    if (mf < 0) {
        // the sort failed on the device: keep the parents (prior_survivor)
        for (int t = threadIdx.x; t < I.n; t += kScanThreads) I.gidx[t] = t;
        if (threadIdx.x == 0) {
            I.meta[kNPrior] = I.n;
            I.meta[kNLast] = 0;
            I.meta[kNRem] = 0;
            I.meta[kIll] = -1;
            I.meta[kDevError] = 1;
        }
        return;
    }
    int tp = 0, tl = 0;
    for (int base = 0; base < I.nm; base += kScanThreads) {
        const int t = base + static_cast<int>(threadIdx.x);
        const int fr = t < I.nm ? I.fronts_mix[t] : -1;
        int cp, cl;
        block_exclusive_scan((fr >= 0 && fr < mf) ? 1 : 0, s_wave, cp);
        block_exclusive_scan((fr == mf) ? 1 : 0, s_wave, cl);
        tp += cp;
        tl += cl;
    }
    int op = 0, ol = 0;
    for (int base = 0; base < I.nm; base += kScanThreads) {
        const int t = base + static_cast<int>(threadIdx.x);
        const int fr = t < I.nm ? I.fronts_mix[t] : -1;
        const int fp = (fr >= 0 && fr < mf) ? 1 : 0, fl = (fr == mf) ? 1 : 0;
        int cp, cl;
        const int ep = block_exclusive_scan(fp, s_wave, cp);
        const int el = block_exclusive_scan(fl, s_wave, cl);
        if (fp) I.gidx[op + ep] = t;
        if (fl) I.gidx[tp + ol + el] = t;
        op += cp;
        ol += cl;
    }
    if (threadIdx.x == 0) {
        I.meta[kNPrior] = tp;
        I.meta[kNLast] = tl;
        I.meta[kNRem] = I.n - tp;
        I.meta[kIll] = -1;
        I.meta[kDevError] = 0;
    }
}

// mode 0: ideal point of the n parents; mode 1: update with the offspring
// columns of the merged objectives.
__global__ void b_ideal(const BInst* __restrict__ insts, BMap mp, int mode)
{
    int k;
    const int i = bmap_find(mp, static_cast<int>(blockIdx.x), k);
    const BInst& I = insts[i];
    if (mode == 0) ideal_body(I.f[I.cur], I.n, 0, I.n, k, I.ideal, 1);
    else ideal_body(I.fm, I.nm, I.n, I.n, k, I.ideal, 0);
}

template <int M>
__global__ __launch_bounds__(256) void b_norm_prep(const BInst* __restrict__ insts, BMap mp)
{
    int local;
    const int i = bmap_find(mp, static_cast<int>(blockIdx.x), local);
    const BInst& I = insts[i];
    norm_prep_body<M>(I.fm, I.nm, I.gidx, I.ideal, I.meta, I.fn, I.asf_val, I.asf_idx, I.row_max, local);
}

__global__ __launch_bounds__(kSolveThreads) void b_norm_solve(const BInst* __restrict__ insts)
{
    const BInst& I = insts[blockIdx.x];
    if (I.state != 1) return;
    norm_solve_body(I.fn, I.nm, I.m, I.asf_val, I.asf_idx, I.row_max, I.meta, I.a);
}

template <int M>
__global__ __launch_bounds__(256) void b_assoc_partial(const BInst* __restrict__ insts, BMap mp)
{
    int local;
    const int i = bmap_find(mp, static_cast<int>(blockIdx.x), local);
    const BInst& I = insts[i];
    const int nbx = (I.nm + 255) / 256;
    assoc_partial_body<M>(I.fn, I.nm, I.a, I.dirs, I.k, I.meta, I.part_d2, I.part_idx, local % nbx, local / nbx);
}

// One wave per individual of the instances with cone tables.
template <int M>
__global__ __launch_bounds__(256) void b_assoc_pruned(const BInst* __restrict__ insts, BMap mp)
{
    __shared__ __align__(8) uint16_t s_list[4][kConeMaxBlocks];
    int local;
    const int i = bmap_find(mp, static_cast<int>(blockIdx.x), local);
    const BInst& I = insts[i];
    if (I.meta[kNRem] <= 0) return;
    const int n_nds = I.meta[kNPrior] + I.meta[kNLast];
    const int wave = static_cast<int>(threadIdx.x >> 6);
    const int p = local * 4 + wave;
    if (p >= n_nds) return;
    assoc_pruned_body<M>(I.fn, I.nm, I.a, I.wp, I.k, I.cperm, I.cones, I.meta, I.dist, I.ref, p, s_list[wave],
                         nullptr);
}

__global__ void b_assoc_final(const BInst* __restrict__ insts, BMap mp)
{
    int local;
    const int i = bmap_find(mp, static_cast<int>(blockIdx.x), local);
    const BInst& I = insts[i];
    if (I.cones) return;                       // associated by b_assoc_pruned
    const int t = local * 256 + static_cast<int>(threadIdx.x);
    if (t < I.nm) assoc_final_body(I.part_d2, I.part_idx, I.nchunks, I.nm, I.meta, I.dist, I.ref, t);
}

// rho_prior of one instance per workgroup (rho_clear + rho_hist).
__global__ __launch_bounds__(1024) void b_rho(const BInst* __restrict__ insts)
{
    const BInst& I = insts[blockIdx.x];
    if (I.state != 1) return;
    for (int j = threadIdx.x; j < I.k; j += blockDim.x) I.rho[j] = 0;
    __syncthreads();
    if (I.meta[kNRem] <= 0) return;
    const int n_prior = I.meta[kNPrior];
    for (int t = threadIdx.x; t < n_prior; t += blockDim.x) {
        const int r = I.ref[t];
        if (r >= 0 && r < I.k) atomicAdd(&I.rho[r], 1);
    }
}

__global__ void b_niche_keys1(const BInst* __restrict__ insts, BMap mp, int ref_bits, uint64_t* __restrict__ keys,
                              int32_t* __restrict__ vals)
{
    int local;
    const int i = bmap_find(mp, static_cast<int>(blockIdx.x), local);
    const BInst& I = insts[i];
    const int t = local * 256 + static_cast<int>(threadIdx.x);
    if (t >= I.nm) return;
    niche_keys1_body(I.ref, I.dist, I.meta, I.k, static_cast<uint64_t>(i) << (32 + ref_bits), keys + I.key_off,
                     vals + I.key_off, t);
}

__global__ void b_niche_keys2(const BInst* __restrict__ insts, BMap mp, int ref_bits,
                              const uint64_t* __restrict__ key1s, uint64_t* __restrict__ key2)
{
    int local;
    const int i = bmap_find(mp, static_cast<int>(blockIdx.x), local);
    const BInst& I = insts[i];
    const int t = local * 256 + static_cast<int>(threadIdx.x);
    if (t >= I.nm) return;
    niche_keys2_body(key1s + I.key_off, I.meta, I.rho, I.k, (1u << ref_bits) - 1u,
                     static_cast<uint64_t>(i) << (32 + ref_bits), nullptr, I.seed, static_cast<uint32_t>(I.gen),
                     key2 + I.key_off, t);
}

__global__ void b_niche_keys3(const BInst* __restrict__ insts, BMap mp, int ref_bits, int rank_bits,
                              const uint64_t* __restrict__ key2s, uint64_t* __restrict__ key3)
{
    int local;
    const int i = bmap_find(mp, static_cast<int>(blockIdx.x), local);
    const BInst& I = insts[i];
    const int t = local * 256 + static_cast<int>(threadIdx.x);
    if (t >= I.nm) return;
    niche_keys3_body(key2s + I.key_off, I.meta, I.rho, I.k, I.nm, (1u << ref_bits) - 1u,
                     static_cast<uint64_t>(i) << (32 + rank_bits), nullptr, I.seed, static_cast<uint32_t>(I.gen),
                     key3 + I.key_off, t);
}

// The niching, survivor flags and their ascending compaction of each running
// instance, one workgroup each (niche_select_body); the instance's slots of the
// batched sort buffers are its workspace.
__global__ __launch_bounds__(kNicheThreads) void b_niche(const BInst* __restrict__ insts, uint64_t* keys_in,
                                                         uint64_t* keys_out, int32_t* vals_in, int32_t* vals_out,
                                                         int bucket, int lds_cap)
{
    __shared__ __align__(16) unsigned char s_raw[kNicheLdsBytes];
    __shared__ int s_wave[64];
    const BInst& I = insts[blockIdx.x];
    if (I.state != 1) return;
    const int bits_ref = 32 + (32 - __clz(I.k + 1));
    const int bits_rank = 32 + (32 - __clz(I.nm));
    const NicheArgs A{I.ref, I.dist, I.rho, I.k, bits_ref, bits_rank, nullptr, nullptr, I.seed,
                      static_cast<uint32_t>(I.gen), I.gidx, I.flag, bucket};
    niche_select_body(A, I.meta, I.nm, I.n, I.fronts_mix, I.chosen, keys_in + I.key_off, keys_out + I.key_off,
                      vals_in + I.key_off, vals_out + I.key_off, lds_cap, s_raw, s_wave);
}

// Survivor flags and their ascending compaction, one instance per workgroup.
__global__ __launch_bounds__(kScanThreads) void b_survivors(const BInst* __restrict__ insts,
                                                            const int32_t* __restrict__ val3s)
{
    __shared__ int s_wave[64];
    const BInst& I = insts[blockIdx.x];
    if (I.state != 1) return;
    const int mf = I.meta[kMaxFront];
    for (int t = threadIdx.x; t < I.nm; t += kScanThreads) I.flag[t] = prior_survivor(I.fronts_mix[t], mf, t, I.n) ? 1 : 0;
    __syncthreads();
    const int n_rem = I.meta[kNRem];
    if (n_rem > 0 && I.meta[kNLast] > 0) {
        const int n_prior = I.meta[kNPrior];
        for (int t = threadIdx.x; t < n_rem; t += kScanThreads) I.flag[I.gidx[n_prior + val3s[I.key_off + t]]] = 1;
    }
    __syncthreads();
    int o = 0;
    for (int base = 0; base < I.nm; base += kScanThreads) {
        const int t = base + static_cast<int>(threadIdx.x);
        const int f = t < I.nm ? static_cast<int>(I.flag[t]) : 0;
        int c;
        const int e = block_exclusive_scan(f, s_wave, c);
        if (f && o + e < I.n) I.chosen[o + e] = t;
        o += c;
    }
    if (threadIdx.x == 0) I.meta[5] = o;
}

__global__ __launch_bounds__(256) void b_gather(const BInst* __restrict__ insts, BMap mp)
{
    int local;
    const int i = bmap_find(mp, static_cast<int>(blockIdx.x), local);
    const BInst& I = insts[i];
    const int cur = I.cur, off = (I.cur + 1) % 3, nxt = (I.cur + 2) % 3, fnext = 1 - I.fcur;
    const int row = local * 4 + static_cast<int>(threadIdx.x >> 6);
    if (row >= I.n) return;
    gather_row(I.chosen, I.n, I.d, I.m, I.x[cur], I.x[off], I.fm, I.cvm, I.fronts_mix, I.x[nxt], I.f[nxt],
               I.cv[nxt], I.fronts[fnext], row, static_cast<int>(threadIdx.x & 63));
}

// After a generation (rotate = 1) the parents move to slot cur + 2 and the
// fronts ping-pong; instances in `from_state` then take `to_state`.
__global__ void b_advance(BInst* __restrict__ insts, int count, int from_state, int to_state, int rotate)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    BInst& I = insts[i];
    if (I.state != from_state) return;
    if (rotate) {
        I.cur = (I.cur + 2) % 3;
        I.fcur = 1 - I.fcur;
        I.gen = I.gen + 1;
    }
    I.state = to_state;
}

}  // namespace

int batch_ideal(const BatchLaunch& L, const BSelGrids& G, int mode)
{
    if (G.objectives.blocks <= 0) return kOk;
    hipLaunchKernelGGL(b_ideal, dim3(G.objectives.blocks), dim3(1024), 0, L.stream, L.insts, G.objectives.map,
                       mode);
    MOEA_HIP(hipGetLastError());
    return kOk;
}

int batch_advance(const BatchLaunch& L, int from_state, int to_state, int rotate)
{
    hipLaunchKernelGGL(b_advance, dim3(ceil_div(L.count, 256)), dim3(256), 0, L.stream, L.insts, L.count,
                       from_state, to_state, rotate);
    MOEA_HIP(hipGetLastError());
    return kOk;
}

int batch_select(const BatchLaunch& L, const BSelGrids& G, const BNdsGrids& NG, const BNdsScratch& S)
{
    hipStream_t st = L.stream;
    const int total = static_cast<int>(G.total_merged);
    if (G.merged256.blocks <= 0) return kOk;
    stage_begin(kStSelectPrep, st);
    hipLaunchKernelGGL(b_merge, dim3(G.merged256.blocks), dim3(256), 0, st, L.insts, G.merged256.map);
    MOEA_HIP(hipGetLastError());
    stage_end(kStSelectPrep, st);
    MOEA_CHECK(batch_nds(L, NG, S, 1, 1));
    stage_begin(kStSelectPrep, st);
    hipLaunchKernelGGL(b_split, dim3(L.count), dim3(kScanThreads), 0, st, L.insts);
    MOEA_CHECK(batch_ideal(L, G, 1));
    for (int m = 1; m <= kMaxM; ++m) {
        if (G.norm_prep[m].blocks <= 0) continue;
        const dim3 grid(G.norm_prep[m].blocks), block(256);
        MOEA_LAUNCH_M(m, b_norm_prep, grid, block, 0, st, L.insts, G.norm_prep[m].map);
    }
    hipLaunchKernelGGL(b_norm_solve, dim3(L.count), dim3(kSolveThreads), 0, st, L.insts);
    MOEA_HIP(hipGetLastError());
    stage_end(kStSelectPrep, st);
    stage_begin(kStAssociation, st);
    for (int m = 1; m <= kMaxM; ++m) {
        if (G.assoc[m].blocks <= 0) continue;
        const dim3 grid(G.assoc[m].blocks), block(256);
        MOEA_LAUNCH_M(m, b_assoc_partial, grid, block, 0, st, L.insts, G.assoc[m].map);
    }
    for (int m = 1; m <= kMaxM; ++m) {
        if (G.assoc_pruned[m].blocks <= 0) continue;
        const dim3 grid(G.assoc_pruned[m].blocks), block(256);
        MOEA_LAUNCH_M(m, b_assoc_pruned, grid, block, 0, st, L.insts, G.assoc_pruned[m].map);
    }
    hipLaunchKernelGGL(b_assoc_final, dim3(G.merged256.blocks), dim3(256), 0, st, L.insts, G.merged256.map);
    MOEA_HIP(hipGetLastError());
    stage_end(kStAssociation, st);
    stage_begin(kStNiching, st);
    hipLaunchKernelGGL(b_rho, dim3(L.count), dim3(1024), 0, st, L.insts);
    MOEA_HIP(hipGetLastError());
    // one workgroup per instance while every instance is small (2n <= 8192);
    // larger runs (many objectives put most of 2n in the last front) sort
    // faster device-wide
    if (niching_one_block(L.niching) && (L.niching.mode == 1 || G.max_nm <= kBatchNicheMaxNm)) {
        hipLaunchKernelGGL(b_niche, dim3(L.count), dim3(kNicheThreads), 0, st, L.insts, S.keys_in, S.keys_out,
                           S.vals_in, S.vals_out, niching_buckets(L.niching), niching_lds_cap(L.niching));
        MOEA_HIP(hipGetLastError());
        stage_end(kStNiching, st);
        stage_begin(kStGather, st);
        hipLaunchKernelGGL(b_gather, dim3(G.rows4.blocks), dim3(256), 0, st, L.insts, G.rows4.map);
        MOEA_HIP(hipGetLastError());
        stage_end(kStGather, st);
        return kOk;
    }
    const int bits_ref = 32 + G.ref_bits + S.inst_bits;
    const int bits_rank = 32 + G.rank_bits + S.inst_bits;
    hipLaunchKernelGGL(b_niche_keys1, dim3(G.merged256.blocks), dim3(256), 0, st, L.insts, G.merged256.map,
                       G.ref_bits, S.keys_in, S.vals_in);
    MOEA_HIP(hipGetLastError());
    MOEA_CHECK(sort_pairs(S.sort_temp, S.sort_bytes, S.keys_in, S.keys_out, S.vals_in, S.vals_out, total, st,
                          bits_ref));
    hipLaunchKernelGGL(b_niche_keys2, dim3(G.merged256.blocks), dim3(256), 0, st, L.insts, G.merged256.map,
                       G.ref_bits, S.keys_out, S.keys_in);
    MOEA_HIP(hipGetLastError());
    MOEA_CHECK(sort_pairs(S.sort_temp, S.sort_bytes, S.keys_in, S.keys_out, S.vals_out, S.vals_in, total, st,
                          bits_ref));
    hipLaunchKernelGGL(b_niche_keys3, dim3(G.merged256.blocks), dim3(256), 0, st, L.insts, G.merged256.map,
                       G.ref_bits, G.rank_bits, S.keys_out, S.keys_in);
    MOEA_HIP(hipGetLastError());
    MOEA_CHECK(sort_pairs(S.sort_temp, S.sort_bytes, S.keys_in, S.keys_out, S.vals_in, S.vals_out, total, st,
                          bits_rank));
    hipLaunchKernelGGL(b_survivors, dim3(L.count), dim3(kScanThreads), 0, st, L.insts, S.vals_out);
    MOEA_HIP(hipGetLastError());
    stage_end(kStNiching, st);
    stage_begin(kStGather, st);
    hipLaunchKernelGGL(b_gather, dim3(G.rows4.blocks), dim3(256), 0, st, L.insts, G.rows4.map);
    MOEA_HIP(hipGetLastError());
    stage_end(kStGather, st);
    return kOk;
}

}  // namespace moea
