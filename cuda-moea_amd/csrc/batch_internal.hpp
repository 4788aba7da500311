// Batched engine: many independent NSGA-III populations of uneven N / M / D
// advanced by one launch sequence per generation.
//
// Every per-instance buffer lives in one caller-provided workspace.  BInst is
// the device-side record of one instance; a BMap sends each block of a batched
// launch to (instance, block within the instance) through a prefix table of
// per-instance block counts, so no block straddles two instances and each
// instance computes exactly what a single-population launch would.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "common.hpp"
#include "../../include/moea_hip.h"

namespace moea {

constexpr int kBatchPeelStateWords = 1024 + 64;   // per instance: peel group barrier flags and totals
constexpr int kBatchPeelMaxLds = 80 * 1024;        // LDS bound of one batched peel block

struct BInst {
    // lifecycle: 0 free, 1 running, 2 admitted (initialization pending)
    int state;
    int cur, fcur, gen;           // parent slot of x/f/cv, parent fronts slot, generation
    // sizes
    int n, m, d, k;
    int nm, W, Wc, npad;          // NDS over the 2n merged rows
    int max_gen;
    int nchunks;                  // association chunks of 1024 directions
    FastDiv ddiv;                 // division by d (reproduction element indices)
    uint64_t seed;
    float sbx_eta0, sbx_eta1, sbx_prob, sbx_copy, pm_eta0, pm_eta1, pm_prob;
    // problem (caller-owned device arrays)
    int problem;                  // MOEA_PROBLEM_QUADRATIC or MOEA_PROBLEM_MLP
    const float* w1;              // mlp: (hidden, d)
    const float* w2;              // mlp: (m, hidden)
    int hidden;
    float* hwork;                 // mlp: (n, hidden) scratch
    const float* centers;         // (m, d)
    const float* bounds;          // (2d)
    const float* dirs;            // (m, k)
    // populations: three rotating slots
    float* x[3];
    float* f[3];
    float* cv[3];
    int32_t* fronts[2];           // parent fronts, ping-pong
    int32_t* winners;             // (n) mating winners
    float* ideal;                 // (m)
    // selection scratch (2n merged rows)
    float* fm;                    // (m, nm)
    float* cvm;                   // (nm)
    int32_t* fronts_mix;          // (nm)
    int32_t* meta;                // (8) max_front, n_prior, n_last, ill, n_rem, chosen count
    int32_t* gidx;                // (nm)
    float* fn;                    // (m, nm)
    float* asf_val;               // (ceil(nm / 256), m) extreme-point candidates per block
    int32_t* asf_idx;             // (ceil(nm / 256), m)
    float* row_max;               // (ceil(nm / 256), m)
    float* a;                     // (kMaxM)
    float* part_d2;               // (nchunks, nm)
    int32_t* part_idx;            // (nchunks, nm)
    float* dist;                  // (nm)
    int32_t* ref;                 // (nm)
    int32_t* rho;                 // (k)
    uint8_t* flag;                // (nm)
    int32_t* chosen;              // (n)
    // pruned association (K >= 2048): cone tables built at admission, else null
    float* wp;                    // (m, k) directions in block order
    int32_t* cperm;               // (k) block position -> direction index
    float* cones;                 // m + 3 rows of ceil(k / 64)
    void* cone_buf;               // the tables and their build scratch
    // NDS scratch (2n merged rows, or n for the initial parents)
    uint64_t* B;                  // (W * Wc * 512) dominator bitmask, sorted order
    float* fs;                    // (m, npad) sorted objectives
    float* cvs;                   // (npad)
    int32_t* perm;                // (nm) sorted position -> local row
    int32_t* hi;                  // (W)
    uint8_t* kind;                // (W)
    int32_t* rend;                // (nm) sorted row -> 1 + last word holding a dominator (0: none)
    uint64_t* A;                  // (2 W) assigned sets of the peel (this round, next round)
    uint32_t* pstate;             // (kBatchPeelStateWords) peel barrier and totals
    long long key_off;            // first of this instance's nm entries in the batch key arrays
    long long key_off_n;          // first of this instance's n entries (NDS of the parents)
};

// Block map: block b of a batched launch belongs to instance i with
// start[i] <= b < start[i + 1]; its local block index is b - start[i].
struct BMap {
    const int* start;             // (count + 1) prefix of per-instance block counts
    int count;
};

// Instance of block b (wave-uniform; binary search over the prefix).
__device__ inline int bmap_find(const BMap& mp, int b, int& local)
{
    int lo = 0, hi = mp.count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (mp.start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    local = b - mp.start[lo];
    return lo;
}

// A batched launch: its block map and total block count.
struct BGrid {
    BMap map;
    int blocks;
};

// Batched stage launches (defined next to their single-population kernels).
struct BatchLaunch {
    BInst* insts;                 // device instance table
    int count;
    hipStream_t stream;
    moea_niching niching;         // niching path of the batch's selections
};

// Grids of one batched NDS call; mode 0 ranks the n parents of a slot, mode 1
// the 2n merged rows.
struct BNdsGrids {
    BGrid rows256;                // nds keys: rows per instance, 256 per block
    BGrid npad256;                // sorted gather: 64 W per instance, 256 per block
    BGrid blocks64;               // block info: W one-wave blocks per instance
    BGrid dom[kMaxM + 1];         // dominance units, 4 per block, of the instances with m objectives
    long long total_rows;         // sum of rows (sort length)
    // peel: cooperative launch of peel_wgs blocks; block w runs
    // peel_jobs[peel_first[w] .. peel_first[w + 1]) (instance, rank, group size)
    const int* peel_first;
    const int* peel_jobs;
    int peel_wgs;
    size_t peel_lds;
};

// Grids of the batched selection stages.
struct BSelGrids {
    BGrid merged256;              // 2n entries per instance, 256 per block
    BGrid rows4;                  // n rows per instance, one wave per row
    BGrid objectives;             // one 1024-thread block per objective per instance
    BGrid norm_prep[kMaxM + 1];   // ceil(2n / 256) blocks, instances with m objectives
    BGrid assoc[kMaxM + 1];       // ceil(2n / 256) x association chunks, instances with m objectives
    BGrid assoc_pruned[kMaxM + 1];// ceil(2n / 4), instances with m objectives and cone tables
    int ref_bits;                 // bits of a direction index (with the two sentinels)
    int rank_bits;                // bits of rho + rank (with the sentinel)
    long long total_merged;       // sum of 2n (sort length)
    int max_nm;                   // largest 2n of the state's instances
};

struct BNdsScratch {
    uint64_t* keys_in;
    uint64_t* keys_out;
    int32_t* vals_in;
    int32_t* vals_out;
    void* sort_temp;
    size_t sort_bytes;
    int inst_bits;                // bits of the instance index above the 34-bit NDS key
    uint32_t* error;              // (1) sticky: set when a peel's bounded wait gave up
};

// Evaluation grids: per objective count, the quadratic rows and the MLP
// hidden-layer tiles and head rows of the instances of that kind.
struct BEvalGrids {
    BGrid quad_rows4[kMaxM + 1];
    BGrid mlp_tiles;              // mlp_tile_count(n, hidden) tile blocks per MLP instance
    BGrid mlp_rows4[kMaxM + 1];
};

// Batched stage launchers (each defined next to its single-population kernels).
// `state` selects the instances a launch serves (1 running, 2 initializing);
// the block maps hold blocks for those instances only.  `which` picks the
// population slot: 0 the parents (slot cur), 1 the offspring (slot cur + 1).
int batch_init_population(const BatchLaunch& L, const BMap& mp, int nblocks);                 // reproduce.hip
int batch_reproduce(const BatchLaunch& L, const BMap& rows256, int nrows_blocks, const BMap& sbx256,
                    int nsbx_blocks);                                                         // reproduce.hip
int batch_evaluate(const BatchLaunch& L, const BEvalGrids& G, int which);                     // evaluate.hip
int mlp_tile_count(int n, int hidden);            // blocks of one MLP hidden-layer launch       // evaluate.hip
size_t batch_nds_sort_bytes(long long total_rows, int inst_bits);                             // nds.hip
int batch_nds(const BatchLaunch& L, const BNdsGrids& G, const BNdsScratch& S, int mode, int state);  // nds.hip
int batch_peel_capacity();                                                                    // nds.hip
int batch_peel_plan(const std::vector<int>& inst, const std::vector<int>& W, int capacity, int rows_per_block,
                    std::vector<int>& first, std::vector<int>& jobs, size_t& lds);           // nds.hip
int batch_ideal(const BatchLaunch& L, const BSelGrids& G, int mode);                          // nsga3.hip
int batch_select(const BatchLaunch& L, const BSelGrids& G, const BNdsGrids& NG, const BNdsScratch& S);  // nsga3.hip
int batch_advance(const BatchLaunch& L, int from_state, int to_state, int rotate);              // nsga3.hip

}  // namespace moea
