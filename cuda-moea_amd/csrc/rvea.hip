// RVEA environmental selection, reference-vector adaptation and gamma on CDNA4.
//
// Replaces RVEAEnvironmentSelector::select (NS/selection/rvea/rvea_environment_selector.cu:163-287)
// and AdaptiveRVEADirections::update (NS/reference/reference_directions.cu:132-162).
// The reference reads the winner count back to the host (elitism_selection.cu:255-257);
// here the count stays on the device: the truncation sort always runs over all
// 2n slots with non-winners keyed last, and the gather reads the live count.
//
// Arithmetic notes: the partition's rsqrtf (reference_partition.cu:77-78) is
// evaluated as 1 / sqrtf, and acosf as the deterministic acosf_det, because
// CUDA's approximate last bits cannot be reproduced.
#include "common.hpp"
#include "moea_internal.hpp"

namespace moea {

namespace {

constexpr float kSentinel = 1e15f;   // rvea_population_update.cu:176-177
constexpr float kMetricMax = 1e30f;  // elitism_selection.cu:11
constexpr int kPartChunk = 1024;
constexpr int kScan = 1024;

enum RMeta : int { kRWinners = 0, kRActive = 1, kRLen = 8 };

__global__ void rv_merge(const float* __restrict__ pf, const float* __restrict__ of,
                         const float* __restrict__ pcv, const float* __restrict__ ocv, int n, int m,
                         float* __restrict__ fm, float* __restrict__ cvm)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nm = 2 * n;
    if (i >= nm) return;
    const bool par = i < n;
    const int s = par ? i : i - n;
    for (int k = 0; k < m; ++k)
        fm[static_cast<size_t>(k) * nm + i] = par ? pf[static_cast<size_t>(k) * n + s] : of[static_cast<size_t>(k) * n + s];
    cvm[i] = par ? pcv[s] : ocv[s];
}

// z_min over all 2n columns (fminf, NaN ignored) then f' = max(f - z_min, 1e-32)
// (objective_translation.cu:15-111).  One block per objective.
__global__ void rv_translate(const float* __restrict__ fm, int nm, float* __restrict__ tf)
{
    __shared__ float s_min[16];
    const int k = blockIdx.x;
    float z = INFINITY;
    for (int i = threadIdx.x; i < nm; i += blockDim.x) z = fminf(z, fm[static_cast<size_t>(k) * nm + i]);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) z = fminf(z, __shfl_xor(z, o));
    if ((threadIdx.x & 63) == 0) s_min[threadIdx.x >> 6] = z;
    __syncthreads();
    float zm = INFINITY;
    for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) zm = fminf(zm, s_min[w]);
    for (int i = threadIdx.x; i < nm; i += blockDim.x)
        tf[static_cast<size_t>(k) * nm + i] = fmaxf(fm[static_cast<size_t>(k) * nm + i] - zm, 1e-32f);
}

// Cosine partition partials over a chunk of directions (partition_phase1_kernel,
// reference_partition.cu:47-161): first maximum wins.
template <int M>
__global__ __launch_bounds__(256) void rv_partition(const float* __restrict__ tf, int nm,
                                                    const float* __restrict__ v, int k,
                                                    float* __restrict__ part_cos,
                                                    int32_t* __restrict__ part_idx)
{
    __shared__ float s_v[M][kPartChunk];
    const int j0 = blockIdx.y * kPartChunk;
    const int jn = min(kPartChunk, k - j0);
    for (int e = threadIdx.x; e < M * kPartChunk; e += 256) {
        const int c = e / kPartChunk, j = e % kPartChunk;
        s_v[c][j] = (j < jn) ? v[static_cast<size_t>(c) * k + j0 + j] : 0.0f;
    }
    __syncthreads();
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nm) return;
    float x[M];
    float ss = 0.0f;
#pragma unroll
    for (int c = 0; c < M; ++c) {
        x[c] = tf[static_cast<size_t>(c) * nm + i];
        ss = fmaf(x[c], x[c], ss);
    }
    const float inv = 1.0f / sqrtf(ss + 1e-12f);
    float best = -INFINITY;
    int bj = -1;
    for (int j = 0; j < jn; ++j) {
        float dot = 0.0f;
#pragma unroll
        for (int c = 0; c < M; ++c) dot = fmaf(x[c], s_v[c][j], dot);
        const float cs = dot * inv;
        if (cs > best) { best = cs; bj = j0 + j; }
    }
    part_cos[static_cast<size_t>(blockIdx.y) * nm + i] = best;
    part_idx[static_cast<size_t>(blockIdx.y) * nm + i] = bj;
}

// Chunk reduction, theta, dead-slot marking, APD (reference_partition.cu:236-335,
// angle_penalized_distance.cu:44-97), and the elitism key
// ref << 32 | infeasible << 31 | metric bits (elitism_selection.cu:60-96).
template <int M>
__global__ void rv_apd_keys(const float* __restrict__ tf, int nm, int n, int active_host,
                            const int32_t* __restrict__ active_dev, int k,
                            const float* __restrict__ part_cos, const int32_t* __restrict__ part_idx,
                            int nchunks, const float* __restrict__ gamma, float penalty,
                            const float* __restrict__ cvm, uint64_t* __restrict__ key,
                            int32_t* __restrict__ val)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nm) return;
    float best = -INFINITY;
    int bj = -1;
    for (int c = 0; c < nchunks; ++c) {
        const float cs = part_cos[static_cast<size_t>(c) * nm + i];
        const int idx = part_idx[static_cast<size_t>(c) * nm + i];
        if (cs > best || (cs == best && idx >= 0 && idx < bj)) { best = cs; bj = idx; }
    }
    const float theta = (best > -INFINITY) ? acosf_det(fminf(fmaxf(best, 0.0f), 1.0f)) : INFINITY;
    const int active = active_dev ? active_dev[0] : active_host;
    int ref = bj;
    if (i >= active && i < n) ref = k;
    float ns = 0.0f;
#pragma unroll
    for (int c = 0; c < M; ++c) {
        const float t = tf[static_cast<size_t>(c) * nm + i];
        ns = fmaf(t, t, ns);
    }
    const float nrm = sqrtf(ns);
    float apd;
    if (nrm > 1e-12f) {
        if (ref < 0 || ref >= k) {
            apd = kMetricMax;
        } else {
            const float th = fminf(fmaxf(theta, 0.0f), 3.14159265358979323846f);
            const float gm = fmaxf(gamma[ref], 1e-6f);
            const float a = (1.0f + penalty * th / gm) * nrm;
            apd = (!isfinite(a) || a < 0.0f) ? kMetricMax : a;
        }
    } else {
        apd = 0.0f;
    }
    const float cv = cvm[i];
    const bool feas = cv <= 1e-8f;
    float met = feas ? apd : cv;
    if (!(isfinite(met) && met >= 0.0f)) met = kMetricMax;
    const uint32_t mb = __float_as_uint(met) & 0x7FFFFFFFu;
    key[i] = (static_cast<uint64_t>(static_cast<uint32_t>(ref)) << 32) |
             (static_cast<uint64_t>(feas ? 0u : 1u) << 31) | mb;
    val[i] = i;
}

// Segment winners (extract_segment_winners_kernel, elitism_selection.cu:105-125)
// get the truncation key infeasible << 63 | metric << 32 | index
// (build_winner_trunc_key_kernel, :130-160); everything else the all-ones key.
__global__ void rv_winner_keys(const uint64_t* __restrict__ ks, const int32_t* __restrict__ vs, int nm,
                               int k, const float* __restrict__ cvm, uint64_t* __restrict__ tkey,
                               int32_t* __restrict__ tval, uint8_t* __restrict__ is_winner)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nm) return;
    const uint32_t hi = static_cast<uint32_t>(ks[t] >> 32);
    const int r = static_cast<int>(hi);
    const bool head = (t == 0) || (static_cast<uint32_t>(ks[t - 1] >> 32) != hi);
    const bool win = head && r >= 0 && r < k;
    const int g = vs[t];
    uint64_t kk = ~0ull;
    if (win) {
        // metric bits and feasibility flag are the low 32 bits of the elitism key
        const uint32_t low = static_cast<uint32_t>(ks[t]);
        const uint64_t flag = (low >> 31) & 1u;
        const uint64_t mb = low & 0x7FFFFFFFu;
        kk = (flag << 63) | (mb << 32) | static_cast<uint32_t>(g);
    }
    (void)cvm;
    tkey[t] = kk;
    tval[t] = g;
    is_winner[t] = win ? 1 : 0;
}

__device__ inline int block_scan_excl(int v, int* s_wave, int& total)
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

// Winner count, survivor flags and the ascending survivor list
// (generate_newpop_gidx, elitism_selection.cu:272-419; sort at
// rvea_population_update.cu:281).  Single 1024-thread block.
__global__ __launch_bounds__(kScan) void rv_choose(const uint8_t* __restrict__ is_winner, int nm, int n,
                                                   const int32_t* __restrict__ tvals_sorted,
                                                   const int32_t* __restrict__ elit_vals_sorted,
                                                   uint8_t* __restrict__ flag,
                                                   int32_t* __restrict__ chosen,
                                                   int32_t* __restrict__ meta)
{
    __shared__ int s_wave[64];
    const int chunk = (nm + kScan - 1) / kScan;
    const int b = threadIdx.x * chunk, e = min(nm, b + chunk);
    int c = 0;
    for (int i = b; i < e; ++i) c += is_winner[i];
    int winners;
    block_scan_excl(c, s_wave, winners);
    for (int i = b; i < e; ++i) flag[i] = 0;
    __syncthreads();
    const int take = (winners == 0) ? n : min(winners, n);
    const int32_t* src = (winners == 0) ? elit_vals_sorted : tvals_sorted;
    for (int t = threadIdx.x; t < take; t += kScan) flag[src[t]] = 1;
    __syncthreads();
    int f = 0;
    for (int i = b; i < e; ++i) f += flag[i];
    int tot;
    int o = block_scan_excl(f, s_wave, tot);
    for (int i = b; i < e; ++i)
        if (flag[i] && o < n) chosen[o++] = i;
    if (threadIdx.x == 0) {
        meta[kRWinners] = winners;
        meta[kRActive] = take;
    }
}

// Survivor gather; dead rows get NaN variables and 1e15 objectives / cv
// (fv_cv_update_kernel + write_sentinels_kernel, rvea_population_update.cu:137-205).
__global__ __launch_bounds__(256) void rv_gather(const int32_t* __restrict__ chosen,
                                                 const int32_t* __restrict__ meta, int n, int d, int m,
                                                 const float* __restrict__ px, const float* __restrict__ ox,
                                                 const float* __restrict__ fm, const float* __restrict__ cvm,
                                                 float* __restrict__ nx, float* __restrict__ nf,
                                                 float* __restrict__ ncv, int32_t* __restrict__ out_active)
{
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const int active = meta[kRActive];
    const int nm = 2 * n;
    float* dst = nx + static_cast<size_t>(row) * d;
    if (row < active) {
        const int s = chosen[row];
        const float* src = (s < n) ? px + static_cast<size_t>(s) * d : ox + static_cast<size_t>(s - n) * d;
        for (int j = lane; j < d; j += 64) dst[j] = src[j];
        if (lane < m) nf[static_cast<size_t>(lane) * n + row] = fm[static_cast<size_t>(lane) * nm + s];
        if (lane == 0) ncv[row] = cvm[s];
    } else {
        for (int j = lane; j < d; j += 64) dst[j] = __builtin_nanf("");
        if (lane < m) nf[static_cast<size_t>(lane) * n + row] = kSentinel;
        if (lane == 0) ncv[row] = kSentinel;
    }
    if (row == 0 && lane == 0 && out_active) out_active[0] = active;
}

// gamma_j = acos(max_{q != j} v_j . v_q), pi when there is no other direction
// (gamma_phase1/2_kernel, angle_penalized_distance.cu:196-362).
template <int M>
__global__ void rv_gamma(const float* __restrict__ v, int k, float* __restrict__ gamma)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= k) return;
    float q[M];
#pragma unroll
    for (int c = 0; c < M; ++c) q[c] = v[static_cast<size_t>(c) * k + j];
    float mx = -INFINITY;
    for (int o = 0; o < k; ++o) {
        if (o == j) continue;
        float dot = 0.0f;
#pragma unroll
        for (int c = 0; c < M; ++c) dot = fmaf(q[c], v[static_cast<size_t>(c) * k + o], dot);
        mx = fmaxf(mx, dot);
    }
    gamma[j] = (mx > -INFINITY) ? acosf_det(fminf(fmaxf(mx, 0.0f), 1.0f)) : 3.14159265358979323846f;
}

// z range over the active columns, then V = normalize(V0 * range)
// (rvea_reference_adaptation.cu:19-232).
__global__ void rv_range(const float* __restrict__ f, int ldf, const int32_t* __restrict__ n_active,
                         float* __restrict__ range)
{
    __shared__ float s_lo[16], s_hi[16];
    const int k = blockIdx.x;
    const int na = n_active[0];
    float lo = INFINITY, hi = -INFINITY;
    for (int i = threadIdx.x; i < na; i += blockDim.x) {
        const float x = f[static_cast<size_t>(k) * ldf + i];
        lo = fminf(lo, x);
        hi = fmaxf(hi, x);
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) { lo = fminf(lo, __shfl_xor(lo, o)); hi = fmaxf(hi, __shfl_xor(hi, o)); }
    if ((threadIdx.x & 63) == 0) { s_lo[threadIdx.x >> 6] = lo; s_hi[threadIdx.x >> 6] = hi; }
    __syncthreads();
    if (threadIdx.x == 0) {
        float l = INFINITY, h = -INFINITY;
        for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) { l = fminf(l, s_lo[w]); h = fmaxf(h, s_hi[w]); }
        float r = h - l;
        if (!isfinite(r) || r < 1e-12f) r = 1e-12f;
        range[k] = r;
    }
}

template <int M>
__global__ void rv_adapt(const float* __restrict__ v0, const float* __restrict__ range, int k,
                         float* __restrict__ vout)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= k) return;
    float base[M], sc[M];
    float bn = 0.0f, sn = 0.0f;
#pragma unroll
    for (int c = 0; c < M; ++c) {
        float b = v0[static_cast<size_t>(c) * k + j];
        if (!isfinite(b)) b = 0.0f;
        base[c] = b;
        bn = fmaf(b, b, bn);
        float s = b * range[c];
        if (!isfinite(s)) s = 0.0f;
        sc[c] = s;
        sn = fmaf(s, s, sn);
    }
    const bool use_s = isfinite(sn) && sn > 1e-12f;
    const bool use_b = isfinite(bn) && bn > 1e-12f;
    const float is = use_s ? 1.0f / sqrtf(sn) : 0.0f;
    const float ib = use_b ? 1.0f / sqrtf(bn) : 0.0f;
#pragma unroll
    for (int c = 0; c < M; ++c) {
        float o;
        if (use_s) o = sc[c] * is;
        else if (use_b) o = base[c] * ib;
        else o = (c == 0) ? 1.0f : 0.0f;
        vout[static_cast<size_t>(c) * k + j] = o;
    }
}

struct RLayout {
    float* fm; float* cvm; float* tf; float* part_cos; int32_t* part_idx;
    uint64_t* k0; uint64_t* k1; int32_t* v0; int32_t* v1;
    uint64_t* t0; uint64_t* t1; int32_t* tv0; int32_t* tv1; uint8_t* win; uint8_t* flag;
    int32_t* chosen; int32_t* meta; void* sort_tmp; size_t sort_bytes;
};

RLayout rcarve(int n, int m, int k, Carve& c)
{
    const int nm = 2 * n;
    const int nchunks = ceil_div(k, kPartChunk);
    RLayout L{};
    L.fm = c.take<float>(static_cast<size_t>(m) * nm);
    L.cvm = c.take<float>(nm);
    L.tf = c.take<float>(static_cast<size_t>(m) * nm);
    L.part_cos = c.take<float>(static_cast<size_t>(nchunks) * nm);
    L.part_idx = c.take<int32_t>(static_cast<size_t>(nchunks) * nm);
    L.k0 = c.take<uint64_t>(nm); L.k1 = c.take<uint64_t>(nm);
    L.v0 = c.take<int32_t>(nm); L.v1 = c.take<int32_t>(nm);
    L.t0 = c.take<uint64_t>(nm); L.t1 = c.take<uint64_t>(nm);
    L.tv0 = c.take<int32_t>(nm); L.tv1 = c.take<int32_t>(nm);
    L.win = c.take<uint8_t>(nm); L.flag = c.take<uint8_t>(nm);
    L.chosen = c.take<int32_t>(n);
    L.meta = c.take<int32_t>(kRLen);
    L.sort_bytes = sort_pairs_temp_bytes(nm);
    L.sort_tmp = c.take_bytes(L.sort_bytes);
    return L;
}

}  // namespace

size_t rvea_workspace_bytes(int n, int m, int k)
{
    Carve c(nullptr, 0);
    rcarve(n, m, k, c);
    return c.used;
}

int rvea_select(const moea_rvea_select_args& A)
{
    const int n = A.n, m = A.m, d = A.d, k = A.k;
    if (n <= 0 || m <= 0 || m > kMaxM || d <= 0 || k <= 0 || A.active <= 0 || A.active > n)
        return fail(kInvalidArgument, "rvea_select: invalid sizes or active count");
    if (!A.parents.x || !A.parents.f || !A.parents.cv || !A.offspring.x || !A.offspring.f ||
        !A.offspring.cv || !A.next.x || !A.next.f || !A.next.cv || !A.directions || !A.gamma ||
        !A.workspace)
        return fail(kInvalidArgument, "rvea_select: null pointer");
    if (A.next.x == A.parents.x || A.next.x == A.offspring.x)
        return fail(kInvalidArgument, "rvea_select: next must not alias parents or offspring");
    Carve c(A.workspace, A.workspace_bytes);
    const RLayout L = rcarve(n, m, k, c);
    if (!c.ok()) return fail(kWorkspaceTooSmall, "rvea_select: workspace too small");
    hipStream_t st = static_cast<hipStream_t>(A.stream);
    const int nm = 2 * n;
    const int nb = ceil_div(nm, 256);
    const int nchunks = ceil_div(k, kPartChunk);
    hipLaunchKernelGGL(rv_merge, dim3(nb), dim3(256), 0, st, A.parents.f, A.offspring.f, A.parents.cv,
                       A.offspring.cv, n, m, L.fm, L.cvm);
    hipLaunchKernelGGL(rv_translate, dim3(m), dim3(1024), 0, st, L.fm, nm, L.tf);
    const dim3 g_part(nb, nchunks), g_nb(nb), b256(256);
    MOEA_LAUNCH_M(m, rv_partition, g_part, b256, 0, st, L.tf, nm, A.directions, k, L.part_cos,
                  L.part_idx);
    MOEA_LAUNCH_M(m, rv_apd_keys, g_nb, b256, 0, st, L.tf, nm, n, A.active, A.active_device, k, L.part_cos,
                  L.part_idx, nchunks, A.gamma, A.penalty, L.cvm, L.k0, L.v0);
    MOEA_HIP(hipGetLastError());
    MOEA_CHECK(sort_pairs(L.sort_tmp, L.sort_bytes, L.k0, L.k1, L.v0, L.v1, nm, st));
    hipLaunchKernelGGL(rv_winner_keys, dim3(nb), dim3(256), 0, st, L.k1, L.v1, nm, k, L.cvm, L.t0, L.tv0,
                       L.win);
    MOEA_HIP(hipGetLastError());
    MOEA_CHECK(sort_pairs(L.sort_tmp, L.sort_bytes, L.t0, L.t1, L.tv0, L.tv1, nm, st));
    hipLaunchKernelGGL(rv_choose, dim3(1), dim3(kScan), 0, st, L.win, nm, n, L.tv1, L.v1, L.flag,
                       L.chosen, L.meta);
    hipLaunchKernelGGL(rv_gather, dim3(ceil_div(n, 4)), dim3(256), 0, st, L.chosen, L.meta, n, d, m,
                       A.parents.x, A.offspring.x, L.fm, L.cvm, A.next.x, A.next.f, A.next.cv,
                       A.out_active);
    MOEA_HIP(hipGetLastError());
    return kOk;
}

int rvea_gamma(const float* v, int m, int k, float* gamma, hipStream_t stream)
{
    if (!v || !gamma || m <= 0 || m > kMaxM || k <= 0) return fail(kInvalidArgument, "rvea_gamma: invalid arguments");
    const dim3 g(ceil_div(k, 256)), b(256);
    MOEA_LAUNCH_M(m, rv_gamma, g, b, 0, stream, v, k, gamma);
    MOEA_HIP(hipGetLastError());
    return kOk;
}

int rvea_adapt(const float* v0, const float* f, int ldf, const int32_t* n_active, int m, int k,
               float* v_out, float* scratch, hipStream_t stream)
{
    if (!v0 || !f || !n_active || !v_out || !scratch || m <= 0 || m > kMaxM || k <= 0)
        return fail(kInvalidArgument, "rvea_adapt: invalid arguments");
    hipLaunchKernelGGL(rv_range, dim3(m), dim3(1024), 0, stream, f, ldf, n_active, scratch);
    const dim3 g(ceil_div(k, 256)), b(256);
    MOEA_LAUNCH_M(m, rv_adapt, g, b, 0, stream, v0, scratch, k, v_out);
    MOEA_HIP(hipGetLastError());
    return kOk;
}

}  // namespace moea
