// Constrained non-dominated sorting on CDNA4.
//
// Replaces the reference's run_ndsort / run_ndsort_init
// (NS/selection/nsga3/kernels/non_dominated_sort.cu:1407-1574, 1333-1405):
// same dominance rule (constraint_dominatee_bitmask_kernel, :409-459), same
// stopping rule (fronts are peeled until at least stop_at individuals are
// assigned; the rest keep front -1; :824-840, 1182-1204), but with no pair
// buffer, no host round trip per front and no capacity limit.
//
// Pipeline (all on one stream, no host synchronisation):
//  1. nds_keys + radix sort: rows are ordered by the key
//       segment 0  feasible, no NaN objective    -> f0 (with -0 read as +0)
//       segment 1  feasible, some NaN objective  -> (input order)
//       segment 2  infeasible                    -> cv (NaN last)
//     In this order a later row can dominate an earlier row only when both
//     are in segment 0 with equal f0: a feasible row is never dominated by an
//     infeasible one, a NaN objective compares false both ways, infeasible
//     rows dominate by strictly smaller cv, and q <= p in every objective with
//     f0(q) > f0(p) is impossible.
//  2. nds_block_prep: objectives and cv in sorted order; per 64-row block b,
//     kind[b] (all rows in segment 0?) and hi[b], the last block holding a row
//     whose key equals the key of block b's last row (b itself unless an f0
//     tie runs past the block).  Dominators of a row of block b all lie in
//     blocks <= hi[b].
//  3. nds_dominance: the dominator bitmask B in sorted order, words (p, w) for
//     w <= hi[p/64].  Tiles below the diagonal whose blocks are all segment 0
//     and not tied need only "q_k <= p_k for k >= 1" (f0 is already strictly
//     smaller for every q of the earlier block): M - 1 compares per pair.
//  4. nds_peel: front peeling with per-row cursors on B in one cooperative
//     launch; fronts are written back in input order.
//
// Data layout of B (u64 words, sorted row index p):
//   W  = ceil(n / 64) words per row (bit q%64 of word q/64 = "q dominates p"),
//   Wc = ceil(W / 8) chunks of 8 words per row,
//   word (p, w) lives at (((p/64) * Wc + w/8) * 64 + p%64) * 8 + w%8.
// A 64-row x 8-word chunk is 4 KiB contiguous: the dominance kernel writes it
// as one 64-byte segment per lane, the peel reads a row's words 8 at a time.
#include <algorithm>
#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "batch_internal.hpp"
#include "common.hpp"
#include "moea_internal.hpp"

namespace moea {

namespace {

constexpr int kRowEndMinM = 4;           // fewest objectives for which the sort computes row ends
constexpr int kDomWaves = 4;             // waves (work units) per dominance workgroup
constexpr int kDomThreads = 64 * kDomWaves;
constexpr int kKeyBits = 34;             // segment (2 bits) above a 32-bit order key
constexpr uint8_t kPlain = 0;            // every row of the block is in segment 0
constexpr uint8_t kMixed = 1;

__host__ __device__ __forceinline__ size_t bword_index(int p, int w, int Wc)
{
    return ((static_cast<size_t>(p >> 6) * Wc + (w >> 3)) * 64 + (p & 63)) * 8 + (w & 7);
}

// Wave-uniform reads of data the kernel never writes, through the constant
// address space so they compile to scalar loads even where alias analysis
// cannot prove the pointer read-only (device functions, pointers from tables).
template <typename T>
__device__ __forceinline__ T ld_uniform(const T* p, size_t i)
{
    using const_ptr = const __attribute__((address_space(4))) T*;
    return ((const_ptr)p)[i];   // address-space conversion (a C-style cast performs it)
}

__device__ __forceinline__ uint64_t nds_key_of(const float* __restrict__ f, int ldf, const float* __restrict__ cv,
                                      int s, int m)
{
    const float c = cv[s];
    bool nan = false;
    if (c == 0.0f)
        for (int k = 0; k < m; ++k) nan = nan || isnan(f[static_cast<size_t>(k) * ldf + s]);
    return nds_key_from(f[s], nan, c);
}

__global__ void nds_keys(const float* __restrict__ f, int ldf, const float* __restrict__ cv, int n,
                         int m, uint64_t* __restrict__ keys, int32_t* __restrict__ vals)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    keys[s] = nds_key_of(f, ldf, cv, s, m);
    vals[s] = s;
}

// Sorted copies of the objectives (m rows of npad) and cv of row t; padding
// rows are 0.
__device__ __forceinline__ void gather_sorted_row(const float* __restrict__ f, int ldf, const float* __restrict__ cv,
                                                  int n, int m, int npad, const int32_t* __restrict__ perm,
                                                  float* __restrict__ fs, float* __restrict__ cvs, int t)
{
    if (t >= npad) return;
    if (t < n) {
        const int s = perm[t];
        for (int k = 0; k < m; ++k) fs[static_cast<size_t>(k) * npad + t] = f[static_cast<size_t>(k) * ldf + s];
        cvs[t] = cv[s];
    } else {
        for (int k = 0; k < m; ++k) fs[static_cast<size_t>(k) * npad + t] = 0.0f;
        cvs[t] = 0.0f;
    }
}

// One wave per 64-row block: kind[b] and hi[b] (see the header comment).  The
// segment is bits 32-33 of the key (higher bits may carry a batch index).
__device__ __forceinline__ void block_info_one(const uint64_t* __restrict__ keys, int n, int b,
                                      uint8_t* __restrict__ kind, int32_t* __restrict__ hi)
{
    const int t = b * 64 + static_cast<int>(threadIdx.x);
    const bool other = (t < n) && ((keys[t] >> 32) & 3u) != 0;
    const uint64_t any_other = ballot(other);
    if (threadIdx.x != 0) return;
    kind[b] = any_other ? kMixed : kPlain;
    const int last = min(n, b * 64 + 64) - 1;
    const uint64_t kl = keys[last];
    int end = last;
    if (((kl >> 32) & 3u) == 0) {
        // first position after `last` whose key differs (keys ascend)
        int lo = last + 1, hi_ = n;
        while (lo < hi_) {
            const int mid = (lo + hi_) >> 1;
            if (keys[mid] == kl) lo = mid + 1;
            else hi_ = mid;
        }
        end = lo - 1;
    }
    hi[b] = end >> 6;
}

// One 64-thread block per 64-row block after the sort: the block's sorted
// objectives and cv (gather_sorted_row) and its kind and tie bound
// (block_info_one).  zero (optional, `words` of them): cleared on the way, for
// a later launch; mf_init (optional): set to -1, the value the peel replaces
// with its result; rend (optional, one per padded row): set to 0, the row ends
// the dominance kernel raises.
__global__ __launch_bounds__(64) void nds_block_prep(const float* __restrict__ f, int ldf,
                                                     const float* __restrict__ cv, int m, int npad,
                                                     const int32_t* __restrict__ perm, float* __restrict__ fs,
                                                     float* __restrict__ cvs, const uint64_t* __restrict__ keys, int n,
                                                     uint8_t* __restrict__ kind, int32_t* __restrict__ hi,
                                                     uint32_t* __restrict__ zero, int words,
                                                     int32_t* __restrict__ mf_init, int32_t* __restrict__ rend)
{
    gather_sorted_row(f, ldf, cv, n, m, npad, perm, fs, cvs, static_cast<int>(blockIdx.x * 64 + threadIdx.x));
    for (int j = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x); j < words;
         j += static_cast<int>(gridDim.x * blockDim.x))
        zero[j] = 0u;
    if (mf_init && blockIdx.x == 0 && threadIdx.x == 0) mf_init[0] = -1;
    if (rend) rend[blockIdx.x * 64 + threadIdx.x] = 0;
    block_info_one(keys, n, static_cast<int>(blockIdx.x), kind, hi);
}

// Tile in which every earlier row q (block qb, wave-uniform) has f0 strictly
// below every row p (lane): q dominates p iff q_k <= p_k for k = 1..M-1.
// Returns the row word of lane p over the 64 rows q of the block.
template <int M>
__device__ __forceinline__ uint64_t tile_fast(const float* __restrict__ fs, int npad, int q0,
                                     const float (&pv)[M])
{
    constexpr int G = (M <= 5) ? 16 : 8;   // rows q per group of scalar loads
    uint32_t acc_hi = 0, acc_lo = 0;
    static_for([&](auto GC) {
        constexpr int g0 = 64 - G * (decltype(GC)::value + 1);   // groups in descending order
        float qv[M > 1 ? M - 1 : 1][G];
#pragma unroll
        for (int k = 1; k < M; ++k)
#pragma unroll
            for (int i = 0; i < G; ++i) qv[k - 1][i] = ld_uniform(fs, static_cast<size_t>(k) * npad + q0 + g0 + i);
        static_for([&](auto IC) {
            constexpr int i = G - 1 - decltype(IC)::value;
            constexpr int j = g0 + i;                  // descending: bit j lands at position j
            uint64_t dom = ~0ull;
#pragma unroll
            for (int k = 1; k < M; ++k) dom &= mask_not_gt(qv[k - 1][i], pv[k]);
            if (j >= 32) acc_hi = shift_in(acc_hi, dom);
            else acc_lo = shift_in(acc_lo, dom);
        }, std::make_integer_sequence<int, G>{});
    }, std::make_integer_sequence<int, 64 / G>{});
    return (static_cast<uint64_t>(acc_hi) << 32) | acc_lo;
}

// Any tile with the full constrained rule.  Row word of lane p over the rows q
// of block qb (q < n, q != p); when `backward` is set, the words "p dominates
// q" of the rows q are stored at word pb of their rows (tied tiles only).
template <int M>
__device__ __forceinline__ uint64_t tile_general(const float* __restrict__ fs, const float* __restrict__ cvs,
                                        int npad, int n, int qb, int pb, const float (&pv)[M],
                                        float cp, bool backward, uint64_t* __restrict__ B, int Wc,
                                        int32_t* __restrict__ rend)
{
    const int lane = lane_id();
    const int p = pb * 64 + lane;
    const bool pvalid = p < n;
    const bool pf = (cp == 0.0f);
    const int q0 = qb * 64;
    uint32_t acc_hi = 0, acc_lo = 0;
    uint32_t col_lo = 0, col_hi = 0;
    static_for([&](auto JC) {
        constexpr int j = 63 - decltype(JC)::value;
        const int q = q0 + j;
        uint64_t qdp_m = 0, pdq_m = 0;
        if (q < n) {
            const float cq = ld_uniform(cvs, q);
            const bool qf = (cq == 0.0f);
            bool le_qp = true, lt_qp = false, le_pq = true, lt_pq = false;
#pragma unroll
            for (int k = 0; k < M; ++k) {
                const float qk = ld_uniform(fs, static_cast<size_t>(k) * npad + q);
                le_qp = le_qp && (qk <= pv[k]);
                lt_qp = lt_qp || (qk < pv[k]);
                le_pq = le_pq && (pv[k] <= qk);
                lt_pq = lt_pq || (pv[k] < qk);
            }
            const bool qdp = (qf && !pf) || (!qf && !pf && (cq < cp)) || (qf && pf && le_qp && lt_qp);
            const bool pdq = (pf && !qf) || (!pf && !qf && (cp < cq)) || (pf && qf && le_pq && lt_pq);
            qdp_m = ballot(qdp && q != p);
            pdq_m = ballot(pdq && pvalid && q != p);
        }
        if (j >= 32) acc_hi = shift_in(acc_hi, qdp_m);
        else acc_lo = shift_in(acc_lo, qdp_m);
        if (backward) {
            col_lo = writelane<j>(col_lo, static_cast<uint32_t>(pdq_m));
            col_hi = writelane<j>(col_hi, static_cast<uint32_t>(pdq_m >> 32));
        }
    }, std::make_integer_sequence<int, 64>{});
    if (backward) {
        const int q = q0 + lane;
        if (q < n) {
            B[bword_index(q, pb, Wc)] = (static_cast<uint64_t>(col_hi) << 32) | col_lo;
            if (rend && (col_hi | col_lo)) atomicMax(&rend[q], pb + 1);
        }
    }
    return (static_cast<uint64_t>(acc_hi) << 32) | acc_lo;
}

// The same, out of line.  Inlined, the 64 unrolled steps of tile_general raise
// the SGPR count of the dominance kernel past what fits without spilling, and
// the spill code runs inside the fast tiles' loop; out of line, the call costs
// more than the spills where diagonal tiles are common (a few row blocks).
// The row's objectives go by value, in registers (by reference they would be
// stored to scratch in every unit).  Up to kOutlineMaxM objectives: from 7 the
// out-of-line form lowers the occupancy and the inlined one measured faster
// (M = 8, n = 65536: 0.97 against 1.00 ms).
constexpr int kOutlineMaxM = 6;
template <int M>
struct RowObjectives {
    float v[M];
};

template <int M>
__device__ __attribute__((noinline)) uint64_t tile_general_call(const float* __restrict__ fs,
                                                                const float* __restrict__ cvs, int npad, int n, int qb,
                                                                int pb, RowObjectives<M> pv, float cp, bool backward,
                                                                uint64_t* __restrict__ B, int Wc,
                                                                int32_t* __restrict__ rend)
{
    return tile_general<M>(fs, cvs, npad, n, qb, pb, pv.v, cp, backward, B, Wc, rend);
}

template <int M>
__device__ __forceinline__ RowObjectives<M> row_pv(const float (&pv)[M])
{
    RowObjectives<M> r;
#pragma unroll
    for (int k = 0; k < M; ++k) r.v[k] = pv[k];
    return r;
}

// Work unit u -> (row block pb, chunk qc), qc <= pb / 8.  Units are grouped by
// a = pb / 8: group a holds 8 row blocks x (a + 1) chunks.
__device__ __forceinline__ void decode_unit(long long u, int& pb, int& qc)
{
    long long a = static_cast<long long>((sqrt(1.0 + static_cast<double>(u)) - 1.0) * 0.5);
    while (4 * a * (a + 1) > u) --a;
    while (4 * (a + 1) * (a + 2) <= u) ++a;
    const long long rem = u - 4 * a * (a + 1);
    pb = static_cast<int>(8 * a + rem / (a + 1));
    qc = static_cast<int>(rem % (a + 1));
}

// One wave per unit (pb, qc): the 8 words (p, 8qc .. 8qc+7) of the 64 rows of
// block pb, accumulated in registers and written as one 64-byte segment per
// lane.  In the diagonal chunk only words <= pb are written; words past pb
// belong to tied tiles, whose backward direction the wave of the later block
// stores.  rend (optional): each row's end, one past its last non-zero word,
// raised with atomicMax from every unit that holds a non-zero word of it.
template <int M, bool kOutline = false>
__device__ __forceinline__ void dom_unit(const float* __restrict__ fs, const float* __restrict__ cvs, int npad,
                                int n, int W, int Wc, const uint8_t* __restrict__ kind,
                                const int32_t* __restrict__ hi, uint64_t* __restrict__ B, long long u,
                                int32_t* __restrict__ rend = nullptr)
{
    int pb, qc;
    decode_unit(u, pb, qc);
    pb = wave_uniform(pb);
    qc = wave_uniform(qc);
    if (pb >= W) return;
    const int lane = lane_id();
    const int p = pb * 64 + lane;
    float pv[M];
#pragma unroll
    for (int k = 0; k < M; ++k) pv[k] = fs[static_cast<size_t>(k) * npad + p];
    const float cp = cvs[p];
    const bool p_plain = ld_uniform(kind, pb) == kPlain;

    uint64_t w0 = 0, w1 = 0, w2 = 0, w3 = 0, w4 = 0, w5 = 0, w6 = 0, w7 = 0;
#pragma unroll 1
    for (int t = 0; t < 8; ++t) {
        const int qb = qc * 8 + t;
        uint64_t word = 0;
        if (qb < pb) {
            const bool tied = ld_uniform(hi, qb) >= pb;
            if (!tied && p_plain && ld_uniform(kind, qb) == kPlain) word = tile_fast<M>(fs, npad, qb * 64, pv);
            else if constexpr (kOutline && M <= kOutlineMaxM) word = tile_general_call<M>(fs, cvs, npad, n, qb, pb, row_pv(pv), cp, tied, B, Wc, rend);
            else word = tile_general<M>(fs, cvs, npad, n, qb, pb, pv, cp, tied, B, Wc, rend);
        } else if (qb == pb) {
            if constexpr (kOutline && M <= kOutlineMaxM) word = tile_general_call<M>(fs, cvs, npad, n, qb, pb, row_pv(pv), cp, false, B, Wc, rend);
            else word = tile_general<M>(fs, cvs, npad, n, qb, pb, pv, cp, false, B, Wc, rend);
        }
        // rotate: after 8 steps w0..w7 hold tiles 0..7
        w0 = w1; w1 = w2; w2 = w3; w3 = w4; w4 = w5; w5 = w6; w6 = w7; w7 = word;
    }
    uint64_t* dst = B + ((static_cast<size_t>(pb) * Wc + qc) * 64 + lane) * 8;
    if (qc * 8 + 7 < pb) {
        ulonglong2* d2 = reinterpret_cast<ulonglong2*>(dst);
        d2[0] = make_ulonglong2(w0, w1);
        d2[1] = make_ulonglong2(w2, w3);
        d2[2] = make_ulonglong2(w4, w5);
        d2[3] = make_ulonglong2(w6, w7);
    } else {
        const int last = pb - qc * 8;   // words 0..last of this chunk
        dst[0] = w0;
        if (last >= 1) dst[1] = w1;
        if (last >= 2) dst[2] = w2;
        if (last >= 3) dst[3] = w3;
        if (last >= 4) dst[4] = w4;
        if (last >= 5) dst[5] = w5;
        if (last >= 6) dst[6] = w6;
        if (last >= 7) dst[7] = w7;
    }
    if (rend && p < n) {
        const int top = w7 ? 8 : w6 ? 7 : w5 ? 6 : w4 ? 5 : w3 ? 4 : w2 ? 3 : w1 ? 2 : w0 ? 1 : 0;
        if (top) atomicMax(&rend[p], qc * 8 + top);
    }
}

template <int M>
__global__ __launch_bounds__(kDomThreads) void nds_dominance(
    const float* __restrict__ fs, const float* __restrict__ cvs, int npad, int n, int W, int Wc,
    const uint8_t* __restrict__ kind, const int32_t* __restrict__ hi, uint64_t* __restrict__ B,
    long long units, int32_t* __restrict__ rend)
{
    const int wave = wave_uniform(static_cast<int>(threadIdx.x >> 6));
    const long long u = static_cast<long long>(blockIdx.x) * kDomWaves + wave;
    if (u >= units) return;
    dom_unit<M>(fs, cvs, npad, n, W, Wc, kind, hi, B, u, rend);
}

// The same for many row blocks, where nearly every tile is a fast one: the
// general tiles out of line keep the fast loop free of spill code.
template <int M>
__global__ __launch_bounds__(kDomThreads) void nds_dominance_wide(
    const float* __restrict__ fs, const float* __restrict__ cvs, int npad, int n, int W, int Wc,
    const uint8_t* __restrict__ kind, const int32_t* __restrict__ hi, uint64_t* __restrict__ B,
    long long units, int32_t* __restrict__ rend)
{
    const int wave = wave_uniform(static_cast<int>(threadIdx.x >> 6));
    const long long u = static_cast<long long>(blockIdx.x) * kDomWaves + wave;
    if (u >= units) return;
    dom_unit<M, true>(fs, cvs, npad, n, W, Wc, kind, hi, B, u, rend);
}

// Peel state (device), u32 words: [0, kBarWords) group barrier,
// [kPeelTotal] total assigned, [kPeelTotal + 1] last round, [kPeelTotal + 2]
// spare output word, [kPeelTotal + 4 + (k & 1)] the total as of the barrier
// ending round k (snapshot read by every block).  A0 / A1: assigned-set
// bitsets in sorted order (W words each), double-buffered across rounds.
//
// Workgroup g owns row blocks g, g + G, g + 2G, ... (G workgroups, spread over
// the whole sort order).  Each owned row keeps, in LDS for the whole launch,
// its blocking word c (the first word of its dominator row that held an
// unassigned dominator when last looked at) and rem, the dominators of word c
// that were still unassigned then.  Round k:
//   A  rem &= ~A[c] against the set assigned before round k.  A row with rem
//      left stays blocked without touching HBM.  A row whose rem emptied reads
//      on from c: kCheckChunks chunks from the one holding c, in one batch
//      (four lanes a row over all 16 waves when the workgroup owns four row
//      blocks or fewer, else a lane per row).  Rows whose dominator words run
//      out are front k; rows still clear with words left go to the
//      workgroup's task queue in LDS as (row, c).
//   B  the workgroup's 16 waves scan the queued rows in sweeps of one round
//      trip per row (8 lanes and 64 words a row in the first sweep, twice as
//      many in each later one while lanes are free), each until it meets its
//      next blocking word (the row keeps (c, rem)) or runs out of words
//      (front k).
// Rows, cursors, fronts and tasks never leave their workgroup (the fronts go
// out once, when the peel stops); only the assigned set and the total cross
// workgroups, at the one grid barrier per round.  Over all rounds a row reads
// each of its words about once, plus what a batch reads past its blocking
// word.
//
// Everything the workgroups exchange inside the launch (barrier words,
// totals, the assigned sets) is accessed with device-scope atomics, which the
// hardware keeps coherent across the XCDs' L2 caches; nothing else written in
// the launch is read in it.  The barrier therefore orders only those
// accesses (coherent_barrier) and does not write back or invalidate the L2.
struct PeelArgs {
    const uint64_t* B;
    const int32_t* hi;       // (W,) last word block of each row block
    const int32_t* perm;     // (n,) sorted position -> input index
    int n, W, Wc, stop_at;
    int blocks_per_group;    // ceil(W / group size)
    int32_t* fronts;         // (n,) out, input order
    uint64_t* A0;
    uint64_t* A1;
    uint32_t* state;
    int32_t* out_max_front;  // (1,) out: last front, kPeelFailed when a bounded wait gave up
    const int32_t* rend;     // (n,) optional: one past each row's last non-zero word (sorted order)
    uint32_t spin_limit;     // polls before a barrier wait gives up
    uint32_t* error;         // optional sticky flag: set to 1 when a bounded wait gave up
};

constexpr int32_t kPeelFailed = -2;
constexpr int kPeelTotal = kBarWords;
constexpr int kPeelStateWords = kBarWords + 16;
static_assert(kPeelStateWords <= kBatchPeelStateWords, "batched peel state too small");
constexpr int kPeelThreads = 1024;
constexpr int kPeelWaves = kPeelThreads / 64;
constexpr int kCheckChunks = 2;     // phase A: chunks read in one batch from the cursor's chunk
constexpr int kScanLoads = 8;       // phase B: words per lane in flight (8 per lane of a row)
constexpr int kMinLanes = 8;        // phase B: fewest lanes per row (64 words a trip)

// LDS layout of the peel (dynamic): s_A (W u64), s_next (bpg u64), s_task
// (2 x 64 bpg u64), s_rem (64 bpg u64), s_cur, s_front and s_end (64 bpg int
// each).
__host__ __device__ __forceinline__ size_t peel_lds_bytes(int W, int bpg)
{
    const size_t rows = static_cast<size_t>(bpg) * 64;
    return static_cast<size_t>(W) * 8 + static_cast<size_t>(bpg) * 8 + rows * 36;
}

__device__ __forceinline__ uint64_t chunk_word(const ulonglong2 (&v)[4], int i)
{
    const ulonglong2 h = v[i >> 1];
    return (i & 1) ? h.y : h.x;
}

__device__ __forceinline__ const uint64_t* peel_rowbase(const uint64_t* __restrict__ B, int Wc, int rb, int lane)
{
    return B + (static_cast<size_t>(rb) * Wc * 64 + lane) * 8;
}

__device__ __forceinline__ void peel_load_chunk(const uint64_t* rowbase, int c, ulonglong2 (&v)[4])
{
    const ulonglong2* src = reinterpret_cast<const ulonglong2*>(rowbase + static_cast<size_t>(c) * 512);
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = src[i];
}

// Phase A read for the lane's row (live, rem emptied): the kCheckChunks
// chunks from the one holding cur, loaded in one batch (one HBM round trip).
// ready: no dominator word left; queued: clear so far with words left;
// otherwise blocked at the returned cur with the unassigned dominators of
// that word in rem.
__device__ __forceinline__ void peel_check(const uint64_t* __restrict__ B, int Wc, int W, int rb, int lane, int end,
                                           const uint64_t* s_A, int& cur, bool& ready, bool& queued, uint64_t& rem)
{
    const uint64_t* rowbase = peel_rowbase(B, Wc, rb, lane);
    const int c = cur >> 3;
    ulonglong2 v[kCheckChunks][4];
#pragma unroll
    for (int a = 0; a < kCheckChunks; ++a)
        if ((c + a) * 8 < end) peel_load_chunk(rowbase, c + a, v[a]);
    // the assigned-set words of the same span, read while the row's words
    // are in flight (indices clamped to the set; words past stop are unused)
    uint64_t am[kCheckChunks * 8];
#pragma unroll
    for (int j = 0; j < kCheckChunks * 8; ++j) am[j] = s_A[min(c * 8 + j, W - 1)];
    const int stop = min((c + kCheckChunks) * 8, end);
    int blocked = stop;
#pragma unroll
    for (int a = kCheckChunks - 1; a >= 0; --a) {
#pragma unroll
        for (int i = 7; i >= 0; --i) {
            const int ww = (c + a) * 8 + i;
            const uint64_t x = chunk_word(v[a], i) & ~am[a * 8 + i];
            if (ww >= cur && ww < stop && x != 0) { blocked = ww; rem = x; }
        }
    }
    cur = blocked;
    ready = (cur >= end);
    queued = !ready && (cur == stop);
}

// Phase B, one trip: groups of L lanes (L = 8, 16, 32 or 64) of a wave, one
// queued row each.  Lane sl of group gi reads words base + iL + sl
// (i < kScanLoads, base = c rounded down to a multiple of 8) of its row: 8L
// words a trip, each 8-word run one 64-byte segment; words before c or from
// end on count as clear.  Returns true when the row meets a word holding an
// unassigned dominator (its index in c, its unassigned bits in rem, the same
// in every lane of the group); otherwise c moves past the words read (to end
// when they run out).  Every lane of the wave calls it; lanes of a group
// without a row pass has = false.
__device__ __forceinline__ bool peel_trip(const uint64_t* __restrict__ B, int Wc, int W, int pr, int& c, int end,
                                          const uint64_t* s_A, int lane, int L, bool has, uint64_t& rem)
{
    const int sl = lane & (L - 1);
    const int shift = lane - sl;
    const uint64_t gmask = (L == 64) ? ~0ull : ((1ull << L) - 1);
    const int base = c & ~7;
    uint64_t x[kScanLoads];
#pragma unroll
    for (int i = 0; i < kScanLoads; ++i) {
        const int w = base + i * L + sl;
        x[i] = (has && w >= c && w < end) ? B[bword_index(pr, w, Wc)] : 0ull;
    }
    // the assigned-set words, read while the row's words are in flight
    uint64_t am[kScanLoads];
#pragma unroll
    for (int i = 0; i < kScanLoads; ++i) am[i] = s_A[min(base + i * L + sl, W - 1)];
#pragma unroll
    for (int i = 0; i < kScanLoads; ++i) {
        const int w = base + i * L + sl;
        x[i] = (has && w >= c && w < end) ? (x[i] & ~am[i]) : 0ull;
    }
    // first blocking word of this lane's group: the lowest (i, sl)
    int first = -1;
    uint64_t hit = 0;
#pragma unroll
    for (int i = kScanLoads - 1; i >= 0; --i) {
        const uint64_t gm = (ballot(x[i] != 0) >> shift) & gmask;
        if (gm) {
            first = i * L + __builtin_ctzll(gm);
            hit = x[i];
        }
    }
    // the blocking word's bits, from the lane that read it
    const int src = shift + (first & (L - 1));
    const uint32_t lo = static_cast<uint32_t>(__shfl(static_cast<int>(hit), src));
    const uint32_t hi = static_cast<uint32_t>(__shfl(static_cast<int>(hit >> 32), src));
    if (first >= 0) {
        c = base + first;
        rem = (static_cast<uint64_t>(hi) << 32) | lo;
        return true;
    }
    c = min(base + kScanLoads * L, end);
    return false;
}

// Records a failed bounded wait: the run's front count becomes kPeelFailed
// and the optional sticky flag is set.
__device__ __forceinline__ void peel_fail(const PeelArgs& P)
{
    if (threadIdx.x != 0) return;
    __hip_atomic_exchange(P.out_max_front, kPeelFailed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (P.error) __hip_atomic_fetch_or(P.error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One peel over the G blocks of a group; g is this block's rank.  Returns
// false when a barrier gave up (the launch is broken; the block exits after
// recording the failure).
__device__ __forceinline__ bool peel_body(const PeelArgs& P, int g, int G, uint64_t* s_dyn)
{
    const int nblk = P.W;
    const int bpg = P.blocks_per_group;
    uint64_t* s_A = s_dyn;                                          // (W) assigned before this round
    uint64_t* s_next = s_A + nblk;                                  // (bpg) rows of own blocks assigned this round
    const int rows = bpg * 64;
    uint64_t* s_task = s_next + bpg;                                // (2 x 64 bpg) queued (cursor << 32 | row)
    uint64_t* s_rem = s_task + 2 * static_cast<size_t>(rows);       // (64 bpg) unassigned bits of the blocking word
    int* s_cur = reinterpret_cast<int*>(s_rem + rows);              // (64 bpg) blocking word
    int* s_front = s_cur + rows;                                    // (64 bpg) front of each owned row
    int* s_end = s_front + rows;                                    // (64 bpg) words 0 .. end-1 can hold dominators
    __shared__ int s_nq[3];         // queue lengths, rotating over the sweeps of phase B
    __shared__ int s_count;
    const int lane = lane_id();
    const int lw = wave_uniform(static_cast<int>(threadIdx.x >> 6));
    uint32_t phase = 0;

    for (int j = lw; j < bpg; j += kPeelWaves) {
        const int rb = g + j * G;
        s_cur[j * 64 + lane] = 0;
        s_rem[j * 64 + lane] = 0ull;
        s_front[j * 64 + lane] = -1;
        if (lane == 0) s_next[j] = 0ull;
        if (rb >= nblk) continue;
        const int p = rb * 64 + lane;
        s_end[j * 64 + lane] = (P.rend && p < P.n) ? P.rend[p] : P.hi[rb] + 1;
        if (lane == 0) { st_relaxed_u64(&P.A0[rb], 0ull); st_relaxed_u64(&P.A1[rb], 0ull); }
    }
    if (!coherent_barrier(P.state, g, G, phase, P.spin_limit)) { peel_fail(P); return false; }

    int total_prev = 0;
    for (int k = 0;; ++k) {
        const uint64_t* Acur = (k & 1) ? P.A1 : P.A0;
        uint64_t* Anext = (k & 1) ? P.A0 : P.A1;
        for (int w = threadIdx.x; w < nblk; w += kPeelThreads) s_A[w] = ld_relaxed_u64(&Acur[w]);
        if (threadIdx.x == 0) { s_count = 0; s_nq[0] = 0; s_nq[1] = 0; }
        __syncthreads();

        // ---- phase A
        int my_new = 0;
        if (bpg * 4 <= kPeelWaves) {
            // four lanes a row: wave lw takes rows 16 (lw & 3) .. 16 (lw & 3) + 15
            // of row block lw / 4, and lane q = lane & 3 of a row reads words
            // 4q .. 4q + 3 of the two chunks from the one holding its cursor
            const int j = lw >> 2;
            const int rb = g + j * G;
            if (j < bpg && rb < nblk) {
                const int q = lane & 3;
                const int r = (lw & 3) * 16 + (lane >> 2);
                const uint64_t amask = s_A[rb];
                const int p = rb * 64 + r;
                const int slot = j * 64 + r;
                const int end = s_end[slot];
                const bool live = (p < P.n) && !((amask >> r) & 1ull);
                int cur = end;
                bool ready = false, queued = false, check = false;
                uint64_t rem = 0;
                if (live) {
                    cur = s_cur[slot];
                    if (cur >= end) {
                        ready = true;
                    } else {
                        // still blocked while dominators of its blocking word are unassigned
                        rem = s_rem[slot] & ~s_A[cur];
                        check = (rem == 0);
                    }
                }
                if (ballot(check)) {
                    const int c = cur >> 3;
                    const int cc = c + (q >> 1);
                    const int w0 = cc * 8 + (q & 1) * 4;
                    ulonglong2 v0 = make_ulonglong2(0ull, 0ull), v1 = v0;
                    if (check && cc * 8 < end) {
                        const ulonglong2* src = reinterpret_cast<const ulonglong2*>(
                            peel_rowbase(P.B, P.Wc, rb, r) + static_cast<size_t>(cc) * 512 + (q & 1) * 4);
                        v0 = src[0];
                        v1 = src[1];
                    }
                    // the assigned-set words, read while the row's words are in flight
                    uint64_t am[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i) am[i] = s_A[min(w0 + i, P.W - 1)];
                    const int stop = min((c + kCheckChunks) * 8, end);
                    const uint64_t wv[4] = {v0.x, v0.y, v1.x, v1.y};
                    int fw = -1;
                    uint64_t fx = 0;
#pragma unroll
                    for (int i = 3; i >= 0; --i) {
                        const int w = w0 + i;
                        const uint64_t x = wv[i] & ~am[i];
                        if (check && w >= cur && w < stop && x != 0) { fw = w; fx = x; }
                    }
                    // the row's first blocking word: from its lowest lane that has one
                    const uint32_t gm = static_cast<uint32_t>(ballot(fw >= 0) >> (lane & ~3)) & 0xFu;
                    const int src = (lane & ~3) + (gm ? __builtin_ctz(gm) : 0);
                    const int bw = __shfl(fw, src);
                    const uint32_t lo = static_cast<uint32_t>(__shfl(static_cast<int>(fx), src));
                    const uint32_t hi = static_cast<uint32_t>(__shfl(static_cast<int>(fx >> 32), src));
                    if (check) {
                        if (gm) {
                            cur = bw;
                            rem = (static_cast<uint64_t>(hi) << 32) | lo;
                        } else {
                            cur = stop;
                            ready = (cur >= end);
                            queued = !ready;
                        }
                    }
                }
                const bool lead = (q == 0);
                if (live && !ready && lead) {
                    s_cur[slot] = cur;
                    s_rem[slot] = rem;
                }
                const uint64_t ready_m = ballot(ready && lead);
                const uint64_t task_m = ballot(queued && lead);
                if (ready && lead) s_front[slot] = k;
                if (task_m) {
                    int base = 0;
                    if (lane == 0) base = atomicAdd(&s_nq[0], __builtin_popcountll(task_m));
                    base = __builtin_amdgcn_readfirstlane(base);
                    if (queued && lead) {
                        const int below = static_cast<int>(__builtin_amdgcn_mbcnt_hi(
                            static_cast<uint32_t>(task_m >> 32),
                            __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(task_m), 0u)));
                        s_task[base + below] = (static_cast<uint64_t>(cur) << 32) | static_cast<uint32_t>(p);
                    }
                }
                if (ready_m) {
                    // the rows' bits sit at lanes 0, 4, ..., 60: bits 0 .. 15 of the wave's rows
                    uint64_t rbits = 0;
#pragma unroll
                    for (int t = 0; t < 16; ++t) rbits |= ((ready_m >> (4 * t)) & 1ull) << t;
                    if (lane == 0) {
                        atomicOr(reinterpret_cast<unsigned long long*>(&s_next[j]),
                                 static_cast<unsigned long long>(rbits << ((lw & 3) * 16)));
                        my_new += __builtin_popcountll(rbits);
                    }
                }
            }
        } else {
            // lane per row
            for (int j = lw; j < bpg; j += kPeelWaves) {
                const int rb = g + j * G;
                if (rb >= nblk) break;
                const uint64_t amask = s_A[rb];
                const int p = rb * 64 + lane;
                const int slot = j * 64 + lane;
                const int end = s_end[slot];
                const bool live = (p < P.n) && !((amask >> lane) & 1ull);
                int cur = end;
                bool ready = false, queued = false;
                if (live) {
                    cur = s_cur[slot];
                    if (cur >= end) {
                        ready = true;
                    } else {
                        // still blocked while dominators of its blocking word are unassigned
                        uint64_t rem = s_rem[slot] & ~s_A[cur];
                        if (rem == 0) {
                            peel_check(P.B, P.Wc, P.W, rb, lane, end, s_A, cur, ready, queued, rem);
                            s_cur[slot] = cur;
                        }
                        s_rem[slot] = rem;
                    }
                }
                const uint64_t ready_m = ballot(ready);
                const uint64_t task_m = ballot(queued);
                if (ready) s_front[slot] = k;
                if (task_m) {
                    int base = 0;
                    if (lane == 0) base = atomicAdd(&s_nq[0], __builtin_popcountll(task_m));
                    base = __builtin_amdgcn_readfirstlane(base);
                    if (queued) {
                        const int below = static_cast<int>(__builtin_amdgcn_mbcnt_hi(
                            static_cast<uint32_t>(task_m >> 32),
                            __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(task_m), 0u)));
                        s_task[base + below] = (static_cast<uint64_t>(cur) << 32) | static_cast<uint32_t>(p);
                    }
                }
                if (lane == 0) {
                    s_next[j] = ready_m;
                    my_new += __builtin_popcountll(ready_m);
                }
            }
        }

        // ---- phase B: the queued rows, in sweeps of one trip per row, each
        // row until it blocks or runs out of words.  Sweep s takes queue s % 2
        // (length s_nq[s % 3]) and queues the rows still clear in the other.
        // A row gets 8 lanes (64 words a trip) in the first sweep and twice as
        // many in each later one, up to a wave, as far as the queue leaves
        // lanes free (8 lanes for more than 64 rows, a wave for 16 or fewer):
        // long scans take few trips, and a row reads at most about twice the
        // words it needs.
        for (int sw = 0;; ++sw) {
            __syncthreads();
            const int nq = s_nq[sw % 3];
            if (nq == 0) break;
            if (threadIdx.x == 0) s_nq[(sw + 2) % 3] = 0;
            int L = 64;
            while (L > kMinLanes && nq > kPeelWaves * (64 / L)) L >>= 1;
            L = min(L, kMinLanes << min(sw, 3));
            const int gpw = 64 / L;
            const uint64_t* q_in = s_task + (sw & 1) * static_cast<size_t>(rows);
            uint64_t* q_out = s_task + ((sw + 1) & 1) * static_cast<size_t>(rows);
            for (int t0 = lw * gpw; t0 < nq; t0 += kPeelWaves * gpw) {
                const int t = t0 + lane / L;
                const bool has = t < nq;
                const uint64_t rec = has ? q_in[t] : 0ull;
                const int pr = static_cast<int>(static_cast<uint32_t>(rec));
                const int rb = pr >> 6;
                const int slot = ((rb - g) / G) * 64 + (pr & 63);
                const int end = has ? s_end[slot] : 0;
                int c = static_cast<int>(rec >> 32);
                uint64_t rem = 0;
                const bool blocked = peel_trip(P.B, P.Wc, P.W, pr, c, end, s_A, lane, L, has, rem);
                if (has && (lane & (L - 1)) == 0) {
                    if (blocked) {
                        s_cur[slot] = c;
                        s_rem[slot] = rem;
                    } else if (c >= end) {
                        s_front[slot] = k;
                        atomicOr(reinterpret_cast<unsigned long long*>(&s_next[(rb - g) / G]),
                                 static_cast<unsigned long long>(1ull << (pr & 63)));
                        ++my_new;
                    } else {
                        q_out[atomicAdd(&s_nq[(sw + 1) % 3], 1)] =
                            (static_cast<uint64_t>(c) << 32) | static_cast<uint32_t>(pr);
                    }
                }
            }
        }
        if (my_new) atomicAdd(&s_count, my_new);
        __syncthreads();
        for (int j = threadIdx.x; j < bpg; j += kPeelThreads) {
            const int rb = g + j * G;
            if (rb < nblk) st_relaxed_u64(&Anext[rb], s_A[rb] | s_next[j]);
            s_next[j] = 0ull;
        }
        if (threadIdx.x == 0 && s_count)
            __hip_atomic_fetch_add(&P.state[kPeelTotal], static_cast<uint32_t>(s_count), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        // the running total is read from the snapshot block 0 takes inside
        // the barrier: a block that reads late must not see counts that
        // faster blocks already add for the next round, or the blocks could
        // disagree about stopping and strand each other at the next barrier
        const int snap = kPeelTotal + 4 + (k & 1);
        if (!coherent_barrier(P.state, g, G, phase, P.spin_limit, kPeelTotal, snap)) { peel_fail(P); return false; }
        const int total = static_cast<int>(ld_relaxed_u32(&P.state[snap]));
        if (total >= P.stop_at || total == total_prev || total >= P.n) {
            if (g == 0 && threadIdx.x == 0) {
                // the launch before the peel set the output to -1; a block
                // that gave up in this last barrier may already have marked
                // the run failed, and that mark stays
                int32_t expected = -1;
                __hip_atomic_compare_exchange_strong(P.out_max_front, &expected,
                                                     (total == total_prev) ? k - 1 : k, __ATOMIC_RELAXED,
                                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            // the owned rows' fronts (-1: not assigned), in input order
            for (int j = lw; j < bpg; j += kPeelWaves) {
                const int p = (g + j * G) * 64 + lane;
                if (p < P.n) P.fronts[P.perm[p]] = s_front[j * 64 + lane];
            }
            return true;
        }
        total_prev = total;
    }
}

__global__ __launch_bounds__(kPeelThreads) void nds_peel(PeelArgs P)
{
    extern __shared__ uint64_t s_dyn[];
    peel_body(P, static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x), s_dyn);
}

// ---------------------------------------------------------------- batched --
// The NDS of every instance of a batch: mode 0 ranks the n parents (slot cur)
// into fronts[fcur], mode 1 the 2n merged rows into fronts_mix and meta[0].  Keys carry the instance index above the 34 NDS bits, so one
// stable sort orders every instance's rows exactly as its own sort would.
struct BRows {
    const float* f;
    const float* cv;
    int rows;
    long long key_off;
    int32_t* fronts;
};

__device__ __forceinline__ BRows brows(const BInst& I, int mode)
{
    BRows r;
    if (mode == 0) {
        r.f = I.f[I.cur]; r.cv = I.cv[I.cur]; r.rows = I.n; r.key_off = I.key_off_n; r.fronts = I.fronts[I.fcur];
    } else {
        r.f = I.fm; r.cv = I.cvm; r.rows = I.nm; r.key_off = I.key_off; r.fronts = I.fronts_mix;
    }
    return r;
}

__global__ void b_nds_keys(const BInst* __restrict__ insts, BMap mp, int mode, uint64_t* __restrict__ keys,
                           int32_t* __restrict__ vals)
{
    int local;
    const int i = bmap_find(mp, static_cast<int>(blockIdx.x), local);
    const BInst& I = insts[i];
    const BRows R = brows(I, mode);
    const int s = local * 256 + static_cast<int>(threadIdx.x);
    if (s >= R.rows) return;
    keys[R.key_off + s] = nds_key_of(R.f, R.rows, R.cv, s, I.m) | (static_cast<uint64_t>(i) << kKeyBits);
    vals[R.key_off + s] = s;
}

__global__ void b_nds_gather(const BInst* __restrict__ insts, BMap mp, int mode,
                             const int32_t* __restrict__ vals_sorted)
{
    int local;
    const int i = bmap_find(mp, static_cast<int>(blockIdx.x), local);
    const BInst& I = insts[i];
    const BRows R = brows(I, mode);
    const int npad = ((R.rows + 63) / 64) * 64;
    const int t = local * 256 + static_cast<int>(threadIdx.x);
    if (t >= npad) return;
    if (t < R.rows) {
        const int s = vals_sorted[R.key_off + t];
        I.perm[t] = s;
        for (int k = 0; k < I.m; ++k)
            I.fs[static_cast<size_t>(k) * npad + t] = R.f[static_cast<size_t>(k) * R.rows + s];
        I.cvs[t] = R.cv[s];
    } else {
        for (int k = 0; k < I.m; ++k) I.fs[static_cast<size_t>(k) * npad + t] = 0.0f;
        I.cvs[t] = 0.0f;
    }
}

__global__ void b_nds_block_info(const BInst* __restrict__ insts, BMap mp, int mode,
                                 const uint64_t* __restrict__ keys_sorted)
{
    int local;
    const int i = bmap_find(mp, static_cast<int>(blockIdx.x), local);
    const BInst& I = insts[i];
    const BRows R = brows(I, mode);
    block_info_one(keys_sorted + R.key_off, R.rows, local, I.kind, I.hi);
    // row ends start at 0; the dominance kernel raises them (four or more objectives)
    const int row = local * 64 + static_cast<int>(threadIdx.x);
    if (I.m >= kRowEndMinM && row < R.rows) I.rend[row] = 0;
}

template <int M>
__global__ __launch_bounds__(kDomThreads) void b_nds_dominance(const BInst* __restrict__ insts, BMap mp,
                                                              int mode)
{
    int local;
    const int i = bmap_find(mp, static_cast<int>(blockIdx.x), local);
    const BInst& I = insts[i];
    const BRows R = brows(I, mode);
    const int W = (R.rows + 63) / 64, Wc = (W + 7) / 8;
    const long long units = 4ll * Wc * (Wc + 1);
    const long long u = static_cast<long long>(local) * kDomWaves + wave_uniform(static_cast<int>(threadIdx.x >> 6));
    if (u >= units) return;
    dom_unit<M>(I.fs, I.cvs, W * 64, R.rows, W, Wc, I.kind, I.hi, I.B, u, M >= kRowEndMinM ? I.rend : nullptr);
}

// Where a batched peel leaves the index of its last front: the selection's
// meta word for the merged rows, a spare peel-state word for the parents.
__device__ __forceinline__ int32_t* peel_max_front(const BInst& I, int mode)
{
    return mode == 1 ? I.meta : reinterpret_cast<int32_t*>(I.pstate + kPeelTotal + 2);
}

// Front peeling of every instance of a batch in one cooperative launch.  The
// blocks form groups; each group peels its instances one after another
// (peel_body, with the group barrier on the instance's own state words), so
// a large instance gets many blocks and small ones share single blocks.
// Jobs of block w: jobs[first[w] .. first[w + 1]) as (instance, rank, group size).
__global__ __launch_bounds__(kPeelThreads) void b_nds_peel_groups(const BInst* __restrict__ insts, int mode,
                                                                 const int* __restrict__ first,
                                                                 const int* __restrict__ jobs, uint32_t spin_limit,
                                                                 uint32_t* error)
{
    extern __shared__ uint64_t s_dyn[];
    const int w = static_cast<int>(blockIdx.x);
    const int j1 = first[w + 1];
    for (int j = first[w]; j < j1; ++j) {
        const int i = jobs[3 * j], g = jobs[3 * j + 1], G = jobs[3 * j + 2];
        const BInst& I = insts[i];
        const BRows R = brows(I, mode);
        const int W = (R.rows + 63) / 64, Wc = (W + 7) / 8;
        PeelArgs P{I.B, I.hi, I.perm, R.rows, W, Wc, mode == 1 ? I.n : R.rows, (W + G - 1) / G, R.fronts,
                   I.A, I.A + W, I.pstate, peel_max_front(I, mode), I.m >= kRowEndMinM ? I.rend : nullptr,
                   spin_limit, error};
        __syncthreads();
        if (!peel_body(P, g, G, s_dyn)) return;
    }
}

// Clears the peel state words of the instances in `state` (barrier flags and
// totals start at zero) and sets their front-count outputs to -1.
__global__ void b_peel_reset(const BInst* __restrict__ insts, int mode, int state)
{
    const BInst& I = insts[blockIdx.x];
    if (I.state != state) return;
    for (int t = threadIdx.x; t < kPeelStateWords; t += blockDim.x) I.pstate[t] = 0u;
    __syncthreads();
    if (threadIdx.x == 0) peel_max_front(I, mode)[0] = -1;
}


constexpr int kMaxWords = 4096;   // n <= 262144: the assigned set fits in 32 KiB of LDS
constexpr int kWideBlocks = 64;   // row blocks (n >= 4033) from which dominance uses nds_dominance_wide

// Polls before a peel barrier gives up (kSpinLimit unless a test lowers it
// with moea_debug_set_spin_limit to make the bounded wait give up on purpose).
std::atomic<uint32_t> g_spin_limit{kSpinLimit};

uint32_t peel_spin_limit() { return g_spin_limit.load(std::memory_order_relaxed); }

struct NdsLayout {
    uint64_t* B;
    uint64_t* keys_in;
    uint64_t* keys_out;
    int32_t* vals_in;
    int32_t* perm;
    void* sort_temp;
    size_t sort_bytes;
    float* fs;
    float* cvs;
    int32_t* hi;
    uint8_t* kind;
    uint64_t* A0;
    uint64_t* A1;
    uint32_t* state;
    int32_t* rend;
};

size_t carve_nds(void* ws, int n, NdsLayout& L)
{
    const int W = ceil_div(n, 64);
    const int Wc = ceil_div(W, 8);
    const int npad = W * 64;
    Carve c(ws, ~size_t{0});
    L.B = c.take<uint64_t>(static_cast<size_t>(W) * Wc * 8 * 64);
    L.keys_in = c.take<uint64_t>(n);
    L.keys_out = c.take<uint64_t>(n);
    L.vals_in = c.take<int32_t>(n);
    L.perm = c.take<int32_t>(n);
    L.sort_bytes = sort_pairs_temp_bytes(n, kKeyBits);
    L.sort_temp = c.take_bytes(L.sort_bytes);
    L.fs = c.take<float>(static_cast<size_t>(kMaxM) * npad);
    L.cvs = c.take<float>(npad);
    L.hi = c.take<int32_t>(W);
    L.kind = c.take<uint8_t>(W);
    L.A0 = c.take<uint64_t>(W);
    L.A1 = c.take<uint64_t>(W);
    L.state = c.take<uint32_t>(kPeelStateWords);
    L.rend = c.take<int32_t>(npad);
    return c.used;
}

// Cooperative grid of the peel: at most one 1024-thread workgroup per CU and
// about 4 row blocks per workgroup; 0 if the kernel cannot be resident.
int peel_grid_size(int nblk)
{
    static int cus = 0;
    if (cus == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    }
    int grid = ceil_div(nblk, 4);
    if (grid > cus) grid = cus;
    if (grid > kBarMaxBlocks) grid = kBarMaxBlocks;
    if (grid < 1) grid = 1;
    const int bpg = ceil_div(nblk, grid);
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(nds_peel),
                                                    kPeelThreads, peel_lds_bytes(nblk, bpg)) != hipSuccess ||
        per_cu < 1)
        return 0;
    return grid;
}

}  // namespace

size_t nds_workspace_bytes(int n)
{
    NdsLayout L;
    return carve_nds(nullptr, n, L);
}

int nds_max_n() { return kMaxWords * 64; }

void nds_key_buffers(void* ws, int n, uint64_t** keys, int32_t** vals)
{
    NdsLayout L;
    carve_nds(ws, n, L);
    *keys = L.keys_in;
    *vals = L.vals_in;
}

int nds_run(const float* f, int ldf, const float* cv, int n, int m, int stop_at, int32_t* fronts,
            int32_t* max_front, void* ws, size_t ws_bytes, hipStream_t stream, uint32_t* error, bool keys_ready)
{
    if (n <= 0 || m <= 0 || m > kMaxM || !f || !cv || !fronts || !max_front || ldf < n)
        return fail(kInvalidArgument, "nds: invalid arguments");
    if (n > nds_max_n()) return fail(kUnsupported, "nds: population larger than supported");
    if (!ws || ws_bytes < nds_workspace_bytes(n)) return fail(kWorkspaceTooSmall, "nds: workspace too small");
    const int W = ceil_div(n, 64);
    const int Wc = ceil_div(W, 8);
    const int npad = W * 64;
    NdsLayout L;
    carve_nds(ws, n, L);

    stage_begin(kStNdsSort, stream);
    const int tb = 256;
    if (!keys_ready) {
        hipLaunchKernelGGL(nds_keys, dim3(ceil_div(n, tb)), dim3(tb), 0, stream, f, ldf, cv, n, m,
                           L.keys_in, L.vals_in);
        MOEA_HIP(hipGetLastError());
    }
    MOEA_CHECK(sort_pairs(L.sort_temp, L.sort_bytes, L.keys_in, L.keys_out, L.vals_in, L.perm, n,
                          stream, kKeyBits));
    // row ends pay for their atomics when many rows have no dominator near
    // their own block, which is the case for many objectives
    int32_t* rend = m >= kRowEndMinM ? L.rend : nullptr;
    // also clears the peel's barrier state (not read before the peel)
    hipLaunchKernelGGL(nds_block_prep, dim3(W), dim3(64), 0, stream, f, ldf, cv, m, npad, L.perm, L.fs, L.cvs,
                       L.keys_out, n, L.kind, L.hi, L.state, kPeelStateWords, max_front, rend);
    MOEA_HIP(hipGetLastError());
    stage_end(kStNdsSort, stream);

    stage_begin(kStDominance, stream);
    const long long units = 4ll * Wc * (Wc + 1);
    const dim3 dgrid(static_cast<unsigned>((units + kDomWaves - 1) / kDomWaves));
    const dim3 dblock(kDomThreads);
    if (W >= kWideBlocks) {
        MOEA_LAUNCH_M(m, nds_dominance_wide, dgrid, dblock, 0, stream, L.fs, L.cvs, npad, n, W, Wc, L.kind,
                      L.hi, L.B, units, rend);
    } else {
        MOEA_LAUNCH_M(m, nds_dominance, dgrid, dblock, 0, stream, L.fs, L.cvs, npad, n, W, Wc, L.kind,
                      L.hi, L.B, units, rend);
    }
    MOEA_HIP(hipGetLastError());
    stage_end(kStDominance, stream);

    const int grid = peel_grid_size(W);
    if (grid <= 0) return fail(kUnsupported, "nds: peel kernel cannot be resident on this device");
    const int bpg = ceil_div(W, grid);
    const size_t lds = peel_lds_bytes(W, bpg);
    PeelArgs args{L.B, L.hi, L.perm, n, W, Wc, stop_at, bpg, fronts, L.A0, L.A1, L.state, max_front, rend,
                  peel_spin_limit(), error};
    void* kargs[] = {&args};
    stage_begin(kStPeel, stream);
    MOEA_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(nds_peel), dim3(grid),
                                        dim3(kPeelThreads), kargs, static_cast<unsigned>(lds), stream));
    MOEA_HIP(hipGetLastError());
    stage_end(kStPeel, stream);
    return kOk;
}

// ------------------------------------------------------ allocation-free --
// moea_nds_rank_static: the same sort with its workspace in this file-scope
// device array instead of a caller's buffer, for callers that cannot allocate
// (graph capture, serving loops).  Every GPU holds its own copy until the
// process ends; host code reaches it only through the address that
// hipGetSymbolAddress reports for the current device (the array's name is a
// host-side stand-in, not a device address).  Nothing in it needs clearing
// between calls: nds_block_prep zeroes the peel state and the peel
// initializes cursors, fronts and the assigned sets.  Calls that overlap on
// one GPU share the array and race.
constexpr size_t kStaticScratchBytes = size_t{16} << 20;
constexpr int kStaticMaxN = 8192;
constexpr int kStaticMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[kStaticMaxDevices] = {};
}  // namespace

int nds_run_static(const float* f, int ldf, const float* cv, int n, int m, int stop_at, int32_t* fronts,
                   int32_t* max_front, hipStream_t stream)
{
    if (n > kStaticMaxN)
        return fail(kInvalidArgument, "moea_nds_rank_static: n = " + std::to_string(n) + " exceeds the " +
                                          std::to_string(kStaticMaxN) + " rows its static workspace holds");
    if (n > 0 && nds_workspace_bytes(n) > kStaticScratchBytes)
        return fail(kInvalidArgument, "moea_nds_rank_static: workspace for n = " + std::to_string(n) +
                                          " exceeds the static scratch array");
    int dev = 0;
    MOEA_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kStaticMaxDevices) return fail(kUnsupported, "moea_nds_rank_static: device index too large");
    void* ptr = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (!g_scratch_ptr[dev]) {
            void* p = nullptr;
            const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
            if (!p) return fail(kHipError, "hipGetSymbolAddress(scratch) returned a null address");
            g_scratch_ptr[dev] = p;
        }
        ptr = g_scratch_ptr[dev];
    }
    return nds_run(f, ldf, cv, n, m, stop_at, fronts, max_front, ptr, kStaticScratchBytes, stream);
}

int nds_static_max_n() { return kStaticMaxN; }

void set_peel_spin_limit(uint32_t limit) { g_spin_limit.store(limit == 0 ? kSpinLimit : limit); }

int batch_peel_capacity()
{
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(b_nds_peel_groups),
                                                    kPeelThreads, kBatchPeelMaxLds) != hipSuccess)
        return 0;
    const int cap = cus * per_cu;
    return cap < kBarMaxBlocks ? cap : kBarMaxBlocks;
}

// Groups for the batched peel.  An instance of W row blocks wants one block
// per `rows_per_block` row blocks (16: one row block per wave; 32: two) and
// needs at least one per 32 (the per-block task queue is sized for 32).  When
// the wanted blocks exceed the capacity, the largest instances still get
// groups of their own (scaled down) and the rest join the least-loaded group
// large enough for them; a group's blocks run its instances in the same
// order, so no two groups wait on each other.  The plan decides only which
// workgroup peels which rows: the fronts do not depend on it.
int batch_peel_plan(const std::vector<int>& inst, const std::vector<int>& W, int capacity, int rows_per_block,
                    std::vector<int>& first, std::vector<int>& jobs, size_t& lds)
{
    const int count = static_cast<int>(inst.size());
    first.assign(1, 0);
    jobs.clear();
    lds = 0;
    if (count == 0) return 0;
    if (capacity < 1) return fail(kUnsupported, "batched peel cannot be resident on this device");
    if (rows_per_block < 1 || rows_per_block > 32) return fail(kInvalidArgument, "batched peel: rows per block out of range");
    std::vector<int> want(count), need(count), order(count);
    long long demand = 0;
    for (int i = 0; i < count; ++i) {
        want[i] = std::max(1, std::min(ceil_div(W[i], rows_per_block), capacity));
        need[i] = std::max(1, ceil_div(W[i], 32));
        if (need[i] > capacity) return fail(kUnsupported, "batched peel: population too large for one device");
        demand += want[i];
        order[i] = i;
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return W[a] > W[b]; });
    struct Group { int size; double load; std::vector<int> members; };
    std::vector<Group> groups;
    int used = 0;
    for (int oi = 0; oi < count; ++oi) {
        const int i = order[oi];
        int size = want[i];
        if (demand > capacity)
            size = std::max(need[i], static_cast<int>(static_cast<double>(want[i]) * capacity / demand));
        const double cost = static_cast<double>(W[i]) * W[i];
        if (used + size <= capacity) {
            groups.push_back({size, cost / size, {i}});
            used += size;
            continue;
        }
        int best = -1;
        for (int gi = 0; gi < static_cast<int>(groups.size()); ++gi)
            if (groups[gi].size >= need[i] && (best < 0 || groups[gi].load < groups[best].load)) best = gi;
        if (best < 0) return fail(kUnsupported, "batched peel: no group can hold an instance");
        groups[best].members.push_back(i);
        groups[best].load += cost / groups[best].size;
    }
    for (const Group& gr : groups) {
        for (int r = 0; r < gr.size; ++r) {
            for (int i : gr.members) {
                jobs.push_back(inst[i]);
                jobs.push_back(r);
                jobs.push_back(gr.size);
            }
            first.push_back(static_cast<int>(jobs.size() / 3));
        }
        for (int i : gr.members) lds = std::max(lds, peel_lds_bytes(W[i], ceil_div(W[i], gr.size)));
    }
    if (lds > static_cast<size_t>(kBatchPeelMaxLds)) return fail(kUnsupported, "batched peel: LDS bound exceeded");
    return kOk;
}

size_t batch_nds_sort_bytes(long long total_rows, int inst_bits)
{
    return sort_pairs_temp_bytes(static_cast<int>(total_rows), kKeyBits + inst_bits);
}

int batch_nds(const BatchLaunch& L, const BNdsGrids& G, const BNdsScratch& S, int mode, int state)
{
    hipStream_t st = L.stream;
    if (G.rows256.blocks <= 0) return kOk;
    stage_begin(kStNdsSort, st);
    hipLaunchKernelGGL(b_nds_keys, dim3(G.rows256.blocks), dim3(256), 0, st, L.insts, G.rows256.map, mode,
                       S.keys_in, S.vals_in);
    MOEA_HIP(hipGetLastError());
    MOEA_CHECK(sort_pairs(S.sort_temp, S.sort_bytes, S.keys_in, S.keys_out, S.vals_in, S.vals_out,
                          static_cast<int>(G.total_rows), st, kKeyBits + S.inst_bits));
    hipLaunchKernelGGL(b_nds_gather, dim3(G.npad256.blocks), dim3(256), 0, st, L.insts, G.npad256.map, mode,
                       S.vals_out);
    MOEA_HIP(hipGetLastError());
    hipLaunchKernelGGL(b_nds_block_info, dim3(G.blocks64.blocks), dim3(64), 0, st, L.insts, G.blocks64.map,
                       mode, S.keys_out);
    MOEA_HIP(hipGetLastError());
    stage_end(kStNdsSort, st);
    stage_begin(kStDominance, st);
    for (int m = 1; m <= kMaxM; ++m) {
        if (G.dom[m].blocks <= 0) continue;
        const dim3 grid(G.dom[m].blocks), block(kDomThreads);
        MOEA_LAUNCH_M(m, b_nds_dominance, grid, block, 0, st, L.insts, G.dom[m].map, mode);
        MOEA_HIP(hipGetLastError());
    }
    stage_end(kStDominance, st);
    stage_begin(kStPeel, st);
    hipLaunchKernelGGL(b_peel_reset, dim3(L.count), dim3(256), 0, st, L.insts, mode, state);
    MOEA_HIP(hipGetLastError());
    int mode_arg = mode;
    const BInst* insts_arg = L.insts;
    const int* first_arg = G.peel_first;
    const int* jobs_arg = G.peel_jobs;
    uint32_t spin_arg = peel_spin_limit();
    uint32_t* error_arg = S.error;
    void* kargs[] = {&insts_arg, &mode_arg, &first_arg, &jobs_arg, &spin_arg, &error_arg};
    MOEA_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(b_nds_peel_groups), dim3(G.peel_wgs),
                                        dim3(kPeelThreads), kargs, static_cast<unsigned>(G.peel_lds), st));
    MOEA_HIP(hipGetLastError());
    stage_end(kStPeel, st);
    return kOk;
}

}  // namespace moea
