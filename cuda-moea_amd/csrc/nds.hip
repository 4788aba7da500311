// Constrained non-dominated sorting on CDNA4.
//
// Replaces the reference's run_ndsort / run_ndsort_init
// (NS/selection/nsga3/kernels/non_dominated_sort.cu:1407-1574, 1333-1405):
// same dominance rule (constraint_dominatee_bitmask_kernel, :409-459), same
// stopping rule (fronts are peeled until at least stop_at individuals are
// assigned; the rest keep front -1; :824-840, 1182-1204), but with no pair
// buffer, no host round trip per front and no capacity limit.
//
// Data layout of the dominator bitmask B (u64 words):
//   W  = ceil(n / 64) words per row (bit q%64 of word q/64 = "q dominates p"),
//   Wc = ceil(W / 8) chunks of 8 words per row,
//   word (p, w) lives at (((p/64) * Wc + w/8) * 64 + p%64) * 8 + w%8.
// A 64-row x 8-word chunk is 4 KiB contiguous, so the tile kernel writes whole
// 64-byte row segments and the peel kernel reads them back coalesced.
#include "common.hpp"
#include "moea_internal.hpp"

namespace moea {

namespace {

constexpr int kTB = 8;                   // tiles per block edge (8 x 8 tiles of 64 x 64)
constexpr int kDomThreads = 256;         // 4 waves

__host__ __device__ inline size_t bword_index(int p, int w, int Wc)
{
    return ((static_cast<size_t>(p >> 6) * Wc + (w >> 3)) * 64 + (p & 63)) * 8 + (w & 7);
}

// Per 64-row block: 1 when every row is feasible (cv == 0) and has only finite
// objectives, so the compare-only fast path applies.
__global__ void nds_block_flags(const float* __restrict__ f, int ldf, const float* __restrict__ cv,
                                int n, int m, uint8_t* __restrict__ simple)
{
    const int blk = blockIdx.x;
    const int p = blk * 64 + threadIdx.x;
    bool ok = true;
    if (p < n) {
        ok = (cv[p] == 0.0f);
        for (int k = 0; k < m; ++k) ok = ok && isfinite(f[static_cast<size_t>(k) * ldf + p]);
    }
    const uint64_t bad = ballot(!ok);
    if (threadIdx.x == 0) simple[blk] = (bad == 0) ? 1 : 0;
}

// One wave processes one 64 x 64 tile (rows p in block pb, lanes q in block qb)
// and returns, through the out-parameters:
//   row_word (lane i): bits q of "q dominates p = pb*64+i"      -> word (p, qb)
//   col_word (lane q): bits j of "p = pb*64+j dominates q"        -> word (q, pb)
// The reversed-direction word is built bit-serially with acc = 2*acc + bit
// (one v_addc per row) and bit-reversed at the end.
template <int M, bool SIMPLE>
__device__ inline void dom_tile(const float* __restrict__ f, int ldf, const float* __restrict__ cv,
                                int n, int pb, int qb, bool need_col, uint64_t& row_word,
                                uint64_t& col_word)
{
    const int lane = lane_id();
    const int q = qb * 64 + lane;
    const bool qvalid = q < n;
    float qv[M];
#pragma unroll
    for (int k = 0; k < M; ++k) qv[k] = qvalid ? f[static_cast<size_t>(k) * ldf + q] : 0.0f;
    const float cq = qvalid ? cv[q] : 0.0f;
    const bool qf = (cq == 0.0f);
    const uint64_t valid = ballot(qvalid);

    uint32_t rw_lo = 0, rw_hi = 0;   // row words assembled with writelane
    uint32_t acc_lo = 0, acc_hi = 0; // bit-serial column word
    const int p0 = pb * 64;
    const int pcount = min(64, n - p0);
    static_for([&](auto JC) {
        constexpr int j = decltype(JC)::value;
        uint64_t q_dom_p = 0;
        bool pdq_lane = false;
        if (j < pcount) {
            const int p = p0 + j;
            float pv[M];
#pragma unroll
            for (int k = 0; k < M; ++k) pv[k] = f[static_cast<size_t>(k) * ldf + p];
            if (SIMPLE) {
                bool lt = false, gt = false;
#pragma unroll
                for (int k = 0; k < M; ++k) {
                    lt = lt | (qv[k] < pv[k]);
                    gt = gt | (qv[k] > pv[k]);
                }
                q_dom_p = ballot(lt && !gt) & valid;
                pdq_lane = gt && !lt && qvalid;
            } else {
                const float cp = cv[p];
                const bool pf = (cp == 0.0f);
                bool le_qp = true, lt_qp = false, le_pq = true, lt_pq = false;
#pragma unroll
                for (int k = 0; k < M; ++k) {
                    le_qp = le_qp & (qv[k] <= pv[k]);
                    lt_qp = lt_qp | (qv[k] < pv[k]);
                    le_pq = le_pq & (pv[k] <= qv[k]);
                    lt_pq = lt_pq | (pv[k] < qv[k]);
                }
                const bool qdp = (qf && !pf) || (!qf && !pf && (cq < cp)) ||
                                 (qf && pf && le_qp && lt_qp);
                const bool pdq = (pf && !qf) || (!pf && !qf && (cp < cq)) ||
                                 (pf && qf && le_pq && lt_pq);
                q_dom_p = ballot(qdp && q != p) & valid;
                pdq_lane = pdq && q != p && qvalid;
            }
        }
        rw_lo = writelane<j>(rw_lo, static_cast<uint32_t>(q_dom_p));
        rw_hi = writelane<j>(rw_hi, static_cast<uint32_t>(q_dom_p >> 32));
        if (need_col) {
            const uint32_t bit = pdq_lane ? 1u : 0u;
            if (j < 32) acc_lo = acc_lo + acc_lo + bit;
            else acc_hi = acc_hi + acc_hi + bit;
        }
    }, std::make_integer_sequence<int, 64>{});
    row_word = (static_cast<uint64_t>(rw_hi) << 32) | rw_lo;
    col_word = (static_cast<uint64_t>(__builtin_bitreverse32(acc_hi)) << 32) |
               __builtin_bitreverse32(acc_lo);
}

// Grid: one workgroup per 8 x 8 block of tiles (a, b) with a <= b in the
// upper triangle of the block grid.  Results are staged in LDS and written as
// full 8-word row chunks in both directions.
template <int M>
__global__ __launch_bounds__(kDomThreads) void nds_dominance(
    const float* __restrict__ f, int ldf, const float* __restrict__ cv, int n, int W, int Wc,
    const uint8_t* __restrict__ simple, uint64_t* __restrict__ B)
{
    __shared__ uint64_t s_row[kTB][kTB][64];   // [pb_local][qb_local][lane]  -> rows of P blocks
    __shared__ uint64_t s_col[kTB][kTB][64];   // [qb_local][pb_local][lane]  -> rows of Q blocks

    // decode (a, b), a <= b, from the linear block id of an upper triangle
    const int nb = Wc;  // block-grid edge in units of 8 tiles
    int rem = blockIdx.x, a = 0;
    while (rem >= nb - a) { rem -= nb - a; ++a; }
    const int b = a + rem;
    const bool diag = (a == b);

    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    for (int t = wave; t < kTB * kTB; t += kDomThreads / 64) {
        const int pl = t / kTB, ql = t % kTB;
        const int pb = a * kTB + pl, qb = b * kTB + ql;
        uint64_t rw = 0, cw = 0;
        const bool in_range = pb < W && qb < W;
        const bool needed = in_range && (!diag || pl <= ql);
        if (needed) {
            const bool need_col = !(diag && pl == ql);
            const bool s = simple[pb] && simple[qb];
            if (s) dom_tile<M, true>(f, ldf, cv, n, pb, qb, need_col, rw, cw);
            else dom_tile<M, false>(f, ldf, cv, n, pb, qb, need_col, rw, cw);
        }
        if (!diag || pl <= ql) s_row[pl][ql][lane] = rw;
        if (!diag || pl < ql) s_col[ql][pl][lane] = cw;
    }
    __syncthreads();
    // write: rows of P blocks, chunk b; rows of Q blocks, chunk a
    for (int e = threadIdx.x; e < kTB * 64 * 8; e += kDomThreads) {
        const int pl = e / 512, r = (e / 8) % 64, w = e % 8;
        const int pb = a * kTB + pl;
        if (pb < W && b * kTB + w < W) {
            uint64_t v;
            if (!diag || pl <= w) v = s_row[pl][w][r];
            else v = s_col[pl][w][r];
            B[((static_cast<size_t>(pb) * Wc + b) * 64 + r) * 8 + w] = v;
        }
        if (!diag) {
            const int qb = b * kTB + pl;
            if (qb < W && a * kTB + w < W)
                B[((static_cast<size_t>(qb) * Wc + a) * 64 + r) * 8 + w] = s_col[pl][w][r];
        }
    }
}

// Peel state (device): [0..2] barrier, [3] total assigned, [4] max_front,
// [5] error.  A0 / A1: assigned-set bitsets (W words each), double-buffered
// across rounds.
struct PeelArgs {
    const uint64_t* B;
    int n, W, Wc, stop_at;
    int32_t* fronts;      // (n,) out
    int32_t* cursor;      // (n,) scratch
    uint64_t* A0;
    uint64_t* A1;
    uint32_t* state;      // barrier + counters (see above)
    int32_t* out_max_front;  // (1,) out
};

constexpr int kPeelThreads = 256;
constexpr int kMaxWordsLds = 4096;  // n <= 262144

__global__ __launch_bounds__(kPeelThreads) void nds_peel(PeelArgs P)
{
    __shared__ uint64_t s_A[kMaxWordsLds];
    __shared__ int s_count;
    const int nblk = (P.n + 63) / 64;
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int waves_total = gridDim.x * (kPeelThreads / 64);
    const int gw = blockIdx.x * (kPeelThreads / 64) + wave;
    uint32_t phase = 0;

    // init owned rows
    for (int rb = gw; rb < nblk; rb += waves_total) {
        const int p = rb * 64 + lane;
        if (p < P.n) { P.fronts[p] = -1; P.cursor[p] = 0; }
        if (lane == 0) { st_relaxed_u64(&P.A0[rb], 0ull); st_relaxed_u64(&P.A1[rb], 0ull); }
    }
    if (!grid_barrier(P.state, gridDim.x, phase)) return;

    int total_prev = 0;
    for (int k = 0;; ++k) {
        const uint64_t* Acur = (k & 1) ? P.A1 : P.A0;
        uint64_t* Anext = (k & 1) ? P.A0 : P.A1;
        for (int w = threadIdx.x; w < P.W; w += kPeelThreads) s_A[w] = ld_relaxed_u64(&Acur[w]);
        if (threadIdx.x == 0) s_count = 0;
        __syncthreads();

        int my_new = 0;
        for (int rb = gw; rb < nblk; rb += waves_total) {
            const int p = rb * 64 + lane;
            const bool live = (p < P.n) && (P.fronts[p] < 0);
            // phase A: is the row still blocked at its cursor word?
            int cur = live ? P.cursor[p] : P.W;
            bool unblocked = false;
            if (live) {
                const uint64_t word = P.B[bword_index(p, cur, P.Wc)];
                unblocked = (word & ~s_A[cur]) == 0;
            }
            uint64_t todo = ballot(unblocked);
            uint64_t ready = 0;
            // phase B: scan unblocked rows forward, 64 words per step, wave-cooperative
            while (todo) {
                const int r = __builtin_ctzll(todo);
                todo &= todo - 1;
                const int pr = rb * 64 + r;
                int c = __shfl(cur, r);
                while (true) {
                    const int w = c + lane;
                    bool blk = false;
                    if (w < P.W) blk = (P.B[bword_index(pr, w, P.Wc)] & ~s_A[w]) != 0;
                    const uint64_t bm = ballot(blk);
                    if (bm) { c = c + __builtin_ctzll(bm); break; }
                    c += 64;
                    if (c >= P.W) { c = P.W; break; }
                }
                if (lane == r) cur = c;
                if (c >= P.W) ready |= 1ull << r;
            }
            if (live) P.cursor[p] = cur;
            if ((ready >> lane) & 1ull) P.fronts[p] = k;
            if (lane == 0) {
                st_relaxed_u64(&Anext[rb], s_A[rb] | ready);
                my_new += __builtin_popcountll(ready);
            }
        }
        if (lane == 0 && my_new) atomicAdd(&s_count, my_new);
        __syncthreads();
        if (threadIdx.x == 0 && s_count)
            __hip_atomic_fetch_add(&P.state[3], static_cast<uint32_t>(s_count), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        if (!grid_barrier(P.state, gridDim.x, phase)) return;
        const int total = static_cast<int>(ld_acquire_u32(&P.state[3]));
        if (total >= P.stop_at || total == total_prev || total >= P.n) {
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                P.out_max_front[0] = (total == total_prev) ? k - 1 : k;
                P.state[4] = static_cast<uint32_t>(k);
            }
            return;
        }
        total_prev = total;
    }
}

__global__ void zero_u32(uint32_t* p, int count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0u;
}

}  // namespace

size_t nds_workspace_bytes(int n)
{
    const int W = ceil_div(n, 64);
    const int Wc = ceil_div(W, 8);
    size_t b = 0;
    b += align_up(static_cast<size_t>(W) * Wc * 8 * 64 * sizeof(uint64_t), 256);  // B
    b += align_up(static_cast<size_t>(W), 256);                                   // flags
    b += align_up(static_cast<size_t>(n) * sizeof(int32_t), 256);                 // cursor
    b += 2 * align_up(static_cast<size_t>(W) * sizeof(uint64_t), 256);            // A0, A1
    b += 256;                                                                      // state
    return b;
}

int nds_max_n() { return kMaxWordsLds * 64; }

int nds_run(const float* f, int ldf, const float* cv, int n, int m, int stop_at, int32_t* fronts,
            int32_t* max_front, void* ws, size_t ws_bytes, hipStream_t stream)
{
    if (n <= 0 || m <= 0 || m > kMaxM || !f || !cv || !fronts || !max_front)
        return fail(kInvalidArgument, "nds: invalid arguments");
    if (n > nds_max_n()) return fail(kUnsupported, "nds: population larger than supported");
    if (ws_bytes < nds_workspace_bytes(n)) return fail(kWorkspaceTooSmall, "nds: workspace too small");
    const int W = ceil_div(n, 64);
    const int Wc = ceil_div(W, 8);
    char* base = static_cast<char*>(ws);
    uint64_t* B = reinterpret_cast<uint64_t*>(base);
    base += align_up(static_cast<size_t>(W) * Wc * 8 * 64 * sizeof(uint64_t), 256);
    uint8_t* flags = reinterpret_cast<uint8_t*>(base);
    base += align_up(static_cast<size_t>(W), 256);
    int32_t* cursor = reinterpret_cast<int32_t*>(base);
    base += align_up(static_cast<size_t>(n) * sizeof(int32_t), 256);
    uint64_t* A0 = reinterpret_cast<uint64_t*>(base);
    base += align_up(static_cast<size_t>(W) * sizeof(uint64_t), 256);
    uint64_t* A1 = reinterpret_cast<uint64_t*>(base);
    base += align_up(static_cast<size_t>(W) * sizeof(uint64_t), 256);
    uint32_t* state = reinterpret_cast<uint32_t*>(base);

    hipLaunchKernelGGL(nds_block_flags, dim3(W), dim3(64), 0, stream, f, ldf, cv, n, m, flags);
    MOEA_HIP(hipGetLastError());
    const int tri = Wc * (Wc + 1) / 2;
    stage_begin(kStDominance, stream);
#define MOEA_DOM_CASE(MM)                                                                   \
    case MM:                                                                               \
        hipLaunchKernelGGL(nds_dominance<MM>, dim3(tri), dim3(kDomThreads), 0, stream, f, ldf, \
                           cv, n, W, Wc, flags, B);                                        \
        break;
    switch (m) {
        MOEA_DOM_CASE(1) MOEA_DOM_CASE(2) MOEA_DOM_CASE(3) MOEA_DOM_CASE(4) MOEA_DOM_CASE(5)
        MOEA_DOM_CASE(6) MOEA_DOM_CASE(7) MOEA_DOM_CASE(8) MOEA_DOM_CASE(9) MOEA_DOM_CASE(10)
        default: return fail(kUnsupported, "nds: unsupported objective count");
    }
#undef MOEA_DOM_CASE
    MOEA_HIP(hipGetLastError());
    stage_end(kStDominance, stream);

    hipLaunchKernelGGL(zero_u32, dim3(1), dim3(64), 0, stream, state, 64);
    MOEA_HIP(hipGetLastError());
    PeelArgs args{B, n, W, Wc, stop_at, fronts, cursor, A0, A1, state, max_front};
    const int nblk = W;
    int grid = peel_grid_size(nblk);
    void* kargs[] = {&args};
    stage_begin(kStPeel, stream);
    MOEA_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(nds_peel), dim3(grid),
                                        dim3(kPeelThreads), kargs, 0, stream));
    stage_end(kStPeel, stream);
    return kOk;
}

int peel_grid_size(int nblk)
{
    static int max_blocks = 0;
    if (max_blocks == 0) {
        int dev = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess) return 1;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return 1;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(nds_peel),
                                                        kPeelThreads, 0) != hipSuccess)
            per_cu = 1;
        max_blocks = cus * (per_cu > 0 ? per_cu : 1);
    }
    const int want = ceil_div(nblk, kPeelThreads / 64);
    return want < max_blocks ? (want > 0 ? want : 1) : max_blocks;
}

}  // namespace moea
