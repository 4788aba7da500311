/*
 * moea_oracle.c -- CPU restatement of the reference's NSGA-III / RVEA generation.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library or the
 * Python package) may link, load or call this file.  Only tests/, the smoke
 * check and bench.py's cpu_baseline leg use it, as the checker.
 *
 * Each function cites the reference file:line it restates.  Paths are relative
 * to the reference tree; NS = python/csrc/native/src.
 *
 * Arithmetic contract (kept identical in the HIP library):
 *   - compiled with -ffp-contract=off; every fused multiply-add is an explicit
 *     fmaf()/fma() call placed where the reference's nvcc build contracts one
 *     (nvcc defaults to -fmad=true, CMakeLists.txt:115-123 sets no override);
 *   - sqrtf / float division are IEEE correctly rounded;
 *   - powf / acosf are replaced by deterministic functions, because CUDA's
 *     powf/acosf last-bit behaviour cannot be reproduced: the SBX spread
 *     factor and the mutation root (exponent <= 1) by orc_powf_root, in float32
 *     with a fixed operation sequence; the mutation power (exponent > 1, whose
 *     results reach far below 2^-126) by orc_powf, and acosf by orc_acosf,
 *     which evaluate in double and round once to float;
 *   - rsqrtf (RVEA partition and adaptation) is replaced by 1/sqrtf.
 *
 * Layouts follow the reference (NI/core/types.cuh:21-51): objectives (M, n)
 * objective-major, variables (n, D) row-major, constraints (n,), bounds
 * interleaved [lb0, ub0, lb1, ub1, ...], directions (M, K).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_API __attribute__((visibility("default")))

/* ======================================================================== */
/* Counter-based draws                                                       */
/* ======================================================================== */

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11).  Multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key
 * increments 0x9E3779B9 / 0xBB67AE85, ten rounds. */
ORC_API void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2],
                               uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n1 = lo1;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        const uint32_t n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* SplitMix64 finalizer over (counter ^ seed), low 32 bits: the reference's
 * niching draw, NS/core/cuda/cuda_random.cu:1606-1614 (mix constants from
 * Vigna's splitmix64.c). */
ORC_API uint32_t orc_splitmix32(uint64_t counter, uint64_t seed)
{
    uint64_t z = counter ^ seed;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z = z ^ (z >> 31);
    return (uint32_t)z;
}

ORC_API uint64_t orc_splitmix64_mix(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* 24-bit uniform in [0, 1): never 1.0, so the reference's reverse_bounds
 * (cuda_random.cu:57-59, 1.0 -> 0.0) is a no-op on these draws. */
static inline float u01(uint32_t u) { return (float)(u >> 8) * 5.9604644775390625e-8f; }
ORC_API float orc_u01(uint32_t u) { return u01(u); }

/* Draw streams of the in-kernel generator (same numbering in the library). */
enum { ST_INIT = 1, ST_MATE = 2, ST_SBX = 3, ST_PM = 4 };

static inline void philox_at(uint64_t seed, uint64_t index, uint32_t gen, uint32_t stream,
                             uint32_t out[4])
{
    const uint32_t ctr[4] = {(uint32_t)index, (uint32_t)(index >> 32), gen, stream};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    orc_philox4x32_10(ctr, key, out);
}

/* Niching draw counters: stream a of generation g starts at g<<33, stream b
 * at (g<<33) | (1<<32); each consumes at most 2^32 counters. */
static inline uint64_t niche_base(uint32_t gen, int which)
{
    return ((uint64_t)gen << 33) | ((uint64_t)(which & 1) << 32);
}

/* ======================================================================== */
/* Deterministic double-evaluated elementary functions                       */
/* ======================================================================== */

static const double LN2_HI = 6.93147180369123816490e-01;
static const double LN2_LO = 1.90821492927058770002e-10;
static const double INV_LN2 = 1.44269504088896338700e+00;

/* log2(x) for finite x > 0: x = 2^e * m, m in [sqrt(1/2), sqrt(2)),
 * ln m = 2 atanh(s), s = (m-1)/(m+1), odd series to s^23. */
ORC_API double orc_log2_d(double x)
{
    int e;
    double m = frexp(x, &e);            /* m in [0.5, 1) */
    if (m < 0.70710678118654752440) { m = m * 2.0; e -= 1; }
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    double p = 1.0 / 23.0;
    p = fma(p, s2, 1.0 / 21.0);
    p = fma(p, s2, 1.0 / 19.0);
    p = fma(p, s2, 1.0 / 17.0);
    p = fma(p, s2, 1.0 / 15.0);
    p = fma(p, s2, 1.0 / 13.0);
    p = fma(p, s2, 1.0 / 11.0);
    p = fma(p, s2, 1.0 / 9.0);
    p = fma(p, s2, 1.0 / 7.0);
    p = fma(p, s2, 1.0 / 5.0);
    p = fma(p, s2, 1.0 / 3.0);
    p = fma(p, s2, 1.0);
    const double ln_m = 2.0 * s * p;
    return (double)e + ln_m * INV_LN2;
}

/* 2^t: t = k + r, |r| <= 1/2, e^(r ln2) by Taylor series to degree 14. */
ORC_API double orc_exp2_d(double t)
{
    if (t > 1100.0) return INFINITY;
    if (t < -1100.0) return 0.0;
    const double k = nearbyint(t);
    const double r = t - k;
    const double y = fma(r, LN2_HI, r * LN2_LO);
    double p = 1.0 / 87178291200.0;           /* 1/14! */
    p = fma(p, y, 1.0 / 6227020800.0);
    p = fma(p, y, 1.0 / 479001600.0);
    p = fma(p, y, 1.0 / 39916800.0);
    p = fma(p, y, 1.0 / 3628800.0);
    p = fma(p, y, 1.0 / 362880.0);
    p = fma(p, y, 1.0 / 40320.0);
    p = fma(p, y, 1.0 / 5040.0);
    p = fma(p, y, 1.0 / 720.0);
    p = fma(p, y, 1.0 / 120.0);
    p = fma(p, y, 1.0 / 24.0);
    p = fma(p, y, 1.0 / 6.0);
    p = fma(p, y, 0.5);
    p = fma(p, y, 1.0);
    p = fma(p, y, 1.0);
    return ldexp(p, (int)k);
}

/* x^y for the SBX / PM call sites (sbx_crossover.cu:47-49,
 * polynomial_mutation.cu:38-52): x >= 0 finite, y > 0 finite. */
ORC_API float orc_powf(float x, float y)
{
    if (x != x || y != y) return NAN;
    if (x < 0.0f) return NAN;
    if (x == 0.0f) return y > 0.0f ? 0.0f : INFINITY;
    if (x == 1.0f || y == 0.0f) return 1.0f;
    if (isinf(x)) return y > 0.0f ? INFINITY : 0.0f;
    return (float)orc_exp2_d((double)y * orc_log2_d((double)x));
}

/* x^y for 0 <= x < 2^24, 0 < y <= 1 in float32 arithmetic: the SBX spread
 * factor (sbx_crossover.cu:47-49) and the polynomial-mutation root
 * (polynomial_mutation.cu:43,49).  x = 2^e m, m in [sqrt(1/2), sqrt(2)];
 * log2 m = (2 / ln 2) atanh((m - 1) / (m + 1)) by the odd atanh series to
 * s^11 / 11 (Abramowitz & Stegun 4.6.22), the quotient and the (2 / ln 2)
 * product carried to double-float precision; z = y (e + log2 m) as hi + lo;
 * 2^z = 2^rint(hi) e^w with e^w by its Taylor polynomial to w^8 / 8!
 * (A&S 4.2.1).  Within 1.05 ulp of the exact power on these domains. */
static inline uint32_t f2u(float x) { uint32_t u; memcpy(&u, &x, sizeof u); return u; }
static inline float u2f(uint32_t u) { float x; memcpy(&x, &u, sizeof x); return x; }

ORC_API float orc_powf_root(float x, float y)
{
    if (!(x > 0.0f)) return x == 0.0f ? 0.0f : NAN;
    if (x == 1.0f) return 1.0f;
    if (x == INFINITY) return INFINITY;
    int e = 0;
    if (x < 1.17549435e-38f) { x = x * 16777216.0f; e = -24; }
    const uint32_t b = f2u(x);
    e += (int)((b >> 23) & 0xffu) - 127;
    float m = u2f((b & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421356f) { m = m * 0.5f; e += 1; }
    const float num = m - 1.0f;
    const float den = m + 1.0f;
    const float den_lo = m - (den - 1.0f);
    const float rden = 1.0f / den;
    const float s = num * rden;
    float rem = fmaf(-s, den, num);
    rem = fmaf(-s, den_lo, rem);
    const float s_lo = rem * rden;
    const float s2 = s * s;
    float p = 1.0f / 11.0f;
    p = fmaf(p, s2, 1.0f / 9.0f);
    p = fmaf(p, s2, 1.0f / 7.0f);
    p = fmaf(p, s2, 1.0f / 5.0f);
    p = fmaf(p, s2, 1.0f / 3.0f);
    p = p * s2;
    const float c_hi = 2.88539004f, c_lo = 3.851926e-08f;
    const float tail = fmaf(s, p, s_lo);
    const float lm = s * c_hi;
    const float lm_err = fmaf(s, c_hi, -lm);
    const float lm_lo = fmaf(tail, c_hi, fmaf(s, c_lo, lm_err));
    const float ef = (float)e;
    const float hi0 = y * ef;
    const float hi0_err = fmaf(y, ef, -hi0);
    const float t1 = y * lm;
    const float t1_err = fmaf(y, lm, -t1);
    const float hi = hi0 + t1;
    const float hi_err = (hi0 - hi) + t1;
    const float lo = fmaf(y, lm_lo, hi0_err + t1_err) + hi_err;
    const float k = rintf(hi);
    const float r = (hi - k) + lo;
    const float l_hi = 0.693147182f, l_lo = -1.9046542e-09f;
    const float w = fmaf(r, l_hi, r * l_lo);
    float q = 1.0f / 40320.0f;
    q = fmaf(q, w, 1.0f / 5040.0f);
    q = fmaf(q, w, 1.0f / 720.0f);
    q = fmaf(q, w, 1.0f / 120.0f);
    q = fmaf(q, w, 1.0f / 24.0f);
    q = fmaf(q, w, 1.0f / 6.0f);
    q = fmaf(q, w, 0.5f);
    q = fmaf(q, w, 1.0f);
    q = fmaf(q, w, 1.0f);
    return ldexpf(q, (int)k);
}

/* orc_powf_root over arrays (tests). */
ORC_API void orc_powf_root_array(const float* x, const float* y, float* out, long long n)
{
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < n; ++i) out[i] = orc_powf_root(x[i], y[i]);
}

/* atan(t) for t in [0, 1]: reduce by the nearest c in {0, 1/4, 1/2, 3/4, 1},
 * atan t = atan c + atan((t - c) / (1 + t c)), odd Taylor series to r^19. */
static const double ATAN_C[5] = {
    0.0,
    2.44978663126864143473e-01,
    4.63647609000806093515e-01,
    6.43501108793284370968e-01,
    7.85398163397448278999e-01,
};
static double atan01_d(double t)
{
    const int i = (int)nearbyint(t * 4.0);
    const double c = 0.25 * (double)i;
    const double r = (t - c) / fma(t, c, 1.0);
    const double r2 = r * r;
    double p = -1.0 / 19.0;
    p = fma(p, r2, 1.0 / 17.0);
    p = fma(p, r2, -1.0 / 15.0);
    p = fma(p, r2, 1.0 / 13.0);
    p = fma(p, r2, -1.0 / 11.0);
    p = fma(p, r2, 1.0 / 9.0);
    p = fma(p, r2, -1.0 / 7.0);
    p = fma(p, r2, 1.0 / 5.0);
    p = fma(p, r2, -1.0 / 3.0);
    p = fma(p, r2, 1.0);
    return ATAN_C[i] + r * p;
}

/* acos(x) for x in [0, 1] (the reference clamps the cosine to [0, 1] before
 * acosf: reference_partition.cu:262-263, angle_penalized_distance.cu:353-354):
 * acos x = 2 atan(sqrt((1-x)/(1+x))). */
ORC_API float orc_acosf(float xf)
{
    if (xf != xf) return NAN;
    double x = (double)xf;
    if (x >= 1.0) return 0.0f;
    if (x <= 0.0) x = 0.0;
    const double t = sqrt((1.0 - x) / (1.0 + x));
    return (float)(2.0 * atan01_d(t));
}

/* ======================================================================== */
/* Reference directions                                                      */
/* ======================================================================== */

static long long comb_ll(int n, int k)
{
    if (k > n) return 0;
    if (k == 0 || k == n) return 1;
    if (k > n - k) k = n - k;
    long long r = 1;
    for (int i = 1; i <= k; ++i) { r *= (n - k + i); r /= i; }
    return r;
}

/* One Das-Dennis layer in std::prev_permutation order of a mask whose first
 * M-1 bits are set (nsga3_reference_generator.cu:28-60).  Writes raw points
 * (M, count) at stride ld starting at column col0. */
static int das_dennis_layer(int h, int m, float* out, int ld, int col0)
{
    const int len = h + m - 1;
    const long long count = comb_ll(len, m - 1);
    if (!out) return (int)count;
    /* positions of the set bits, lexicographically smallest first: the order
     * prev_permutation visits starting from 1..10..0 */
    int* pos = (int*)malloc(sizeof(int) * (size_t)(m > 1 ? m - 1 : 1));
    for (int i = 0; i < m - 1; ++i) pos[i] = i;
    for (long long p = 0; p < count; ++p) {
        int prev = -1;
        for (int dim = 0; dim < m; ++dim) {
            const int next = (dim < m - 1) ? pos[dim] : len;
            out[(size_t)dim * ld + col0 + p] = (float)(next - prev - 1) / (float)h;
            prev = next;
        }
        /* advance to the next combination in lexicographic order of set-bit
         * positions (= prev_permutation order of the bool mask) */
        int i = m - 2;
        while (i >= 0 && pos[i] == len - (m - 1) + i) --i;
        if (i < 0) break;
        pos[i] += 1;
        for (int j = i + 1; j < m - 1; ++j) pos[j] = pos[j - 1] + 1;
    }
    free(pos);
    return (int)count;
}

/* Das-Dennis set sized from a requested count (nsga3_reference_generator.cu:62-119,
 * rvea_reference_generator.cu:78-129) or from explicit partitions
 * (nsga3_reference_generator.cu:182-221).  kind: 0 = NSGA-III by count,
 * 1 = NSGA-III by partitions, 2 = RVEA by count (values clamped >= 1e-6 before
 * normalizing, rvea_reference_generator.cu:197-199).
 * Returns K; out (M, K) receives the L2-normalized directions when non-NULL;
 * raw (M, K) receives the unnormalized points when non-NULL. */
ORC_API int orc_das_dennis(int n_req, int m, int partitions, int kind, float* out, float* raw)
{
    if (m <= 0) return -1;
    int h1, h2 = 0, n1, n2 = 0;
    if (kind == 1) {
        if (partitions <= 0) return -1;
        h1 = partitions;
        n1 = (int)comb_ll(h1 + m - 1, m - 1);
    } else {
        h1 = 0;
        while (comb_ll(h1 + m - 1, m - 1) <= n_req) ++h1;
        --h1;
        if (h1 <= 0) return -1;
        n1 = (int)comb_ll(h1 + m - 1, m - 1);
        if (h1 < m) {
            h2 = 1;
            while (comb_ll(h1 + m - 1, m - 1) + comb_ll(h2 + m - 1, m - 1) <= n_req) ++h2;
            --h2;
            if (h2 > 0) n2 = (int)comb_ll(h2 + m - 1, m - 1);
        }
    }
    const int k = n1 + n2;
    if (!out && !raw) return k;
    float* pts = (float*)malloc(sizeof(float) * (size_t)m * k);
    das_dennis_layer(h1, m, pts, k, 0);
    if (n2 > 0) {
        float* tmp = (float*)malloc(sizeof(float) * (size_t)m * n2);
        das_dennis_layer(h2, m, tmp, n2, 0);
        const float centroid = 1.0f / (float)m;
        for (int p = 0; p < n2; ++p)
            for (int d = 0; d < m; ++d)
                pts[(size_t)d * k + n1 + p] = 0.5f * (tmp[(size_t)d * n2 + p] + centroid);
        free(tmp);
    }
    if (kind == 2)
        for (size_t i = 0; i < (size_t)m * k; ++i) if (pts[i] < 1e-6f) pts[i] = 1e-6f;
    if (raw) memcpy(raw, pts, sizeof(float) * (size_t)m * k);
    if (out) {
        for (int p = 0; p < k; ++p) {
            float ss = 0.0f;
            for (int d = 0; d < m; ++d) { const float v = pts[(size_t)d * k + p]; ss = ss + v * v; }
            if (kind == 1) {
                /* by-partitions path multiplies by the reciprocal
                 * (nsga3_reference_generator.cu:199-213) */
                const float inv = 1.0f / sqrtf(ss);
                for (int d = 0; d < m; ++d) out[(size_t)d * k + p] = pts[(size_t)d * k + p] * inv;
            } else {
                const float nrm = sqrtf(ss);
                for (int d = 0; d < m; ++d) out[(size_t)d * k + p] = pts[(size_t)d * k + p] / nrm;
            }
        }
    }
    free(pts);
    return k;
}

/* UserDefinedDirections normalization (reference_directions.cu:218-237). */
ORC_API int orc_normalize_directions(float* v, int m, int k)
{
    for (int p = 0; p < k; ++p) {
        float ss = 0.0f;
        for (int d = 0; d < m; ++d) { const float x = v[(size_t)d * k + p]; ss = ss + x * x; }
        if (ss <= 0.0f) return -1;
        const float inv = 1.0f / sqrtf(ss);
        for (int d = 0; d < m; ++d) v[(size_t)d * k + p] = v[(size_t)d * k + p] * inv;
    }
    return 0;
}

/* ======================================================================== */
/* Constrained dominance and front peeling                                   */
/* ======================================================================== */

/* q dominates p (non_dominated_sort.cu:409-459).  Feasible means cv == 0.0f
 * exactly (cuda_globals.cuh:46). */
static inline int dominates(const float* f, int ld, int m, const float* cv, int q, int p)
{
    if (q == p) return 0;
    const float cq = cv[q], cp = cv[p];
    const int qf = (cq == 0.0f), pf = (cp == 0.0f);
    if (qf && !pf) return 1;
    if (!qf && !pf) return cq < cp;
    if (!qf && pf) return 0;
    int le = 1, lt = 0;
    for (int k = 0; k < m; ++k) {
        const float qv = f[(size_t)k * ld + q], pv = f[(size_t)k * ld + p];
        le &= (qv <= pv);
        lt |= (qv < pv);
    }
    return le && lt;
}

ORC_API int orc_dominates(const float* f, int ld, int m, const float* cv, int q, int p)
{
    return dominates(f, ld, m, cv, q, p);
}

/* Non-dominated sorting with the reference's stopping rule: fronts are peeled
 * until the number of assigned individuals reaches stop_at; later individuals
 * keep front -1 (non_dominated_sort.cu:824-840, 1182-1204).  Returns max_front,
 * the index of the last peeled front. */
ORC_API int orc_nds(const float* f, const float* cv, int n, int m, int stop_at, int32_t* fronts)
{
    const int words = (n + 63) / 64;
    uint64_t* dom = (uint64_t*)calloc((size_t)n * words, sizeof(uint64_t)); /* row p: dominators of p */
#pragma omp parallel for schedule(dynamic, 64)
    for (int p = 0; p < n; ++p)
        for (int q = 0; q < n; ++q)
            if (dominates(f, n, m, cv, q, p))
                dom[(size_t)p * words + (q >> 6)] |= 1ULL << (q & 63);
    uint64_t* assigned = (uint64_t*)calloc((size_t)words, sizeof(uint64_t));
    int* cursor = (int*)calloc((size_t)n, sizeof(int));
    for (int i = 0; i < n; ++i) fronts[i] = -1;
    int total = 0, front = 0, max_front = 0;
    while (total < n) {
        int added = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : added)
        for (int p = 0; p < n; ++p) {
            if (fronts[p] >= 0) continue;
            int w = cursor[p];
            const uint64_t* row = dom + (size_t)p * words;
            while (w < words && (row[w] & ~assigned[w]) == 0) ++w;
            cursor[p] = w;
            if (w == words) { fronts[p] = front; ++added; }
        }
        if (added == 0) break; /* unreachable for a strict partial order */
        for (int p = 0; p < n; ++p)
            if (fronts[p] == front) assigned[p >> 6] |= 1ULL << (p & 63);
        total += added;
        max_front = front;
        if (total >= stop_at) break;
        ++front;
    }
    free(cursor);
    free(assigned);
    free(dom);
    return max_front;
}

/* ======================================================================== */
/* NSGA-III selection                                                        */
/* ======================================================================== */

/* Row minimum with fminf semantics (NaN ignored), ideal_point.cu:25-85. */
static void row_min(const float* f, int ld, int col0, int n, int m, float* out)
{
    for (int k = 0; k < m; ++k) {
        float z = INFINITY;
        for (int i = 0; i < n; ++i) z = fminf(z, f[(size_t)k * ld + col0 + i]);
        out[k] = z;
    }
}
ORC_API void orc_ideal(const float* f, int ld, int col0, int n, int m, float* out)
{
    row_min(f, ld, col0, n, m, out);
}

/* One-sided Jacobi SVD of the M x M matrix A (row i = extreme point i), then
 * a = pinv(A) * 1 with singular values <= 1e-5 dropped.  Reproduces the
 * reference's cusolverDnSgesvd + cublasSgemv pipeline (normalization.cu:1302-1394)
 * in FP64.  Returns 1 when the reference would take its fallback: non-finite
 * input or singular values, sigma_min < 1e-10, sigma_max / sigma_min > 1e6
 * (normalization.cu:739-771), or no convergence (normalization.cu:1340-1348). */
ORC_API int orc_hyperplane_fp64(const double* a_in, int m, double* a_out, double* sigma_out)
{
    double A[100], V[100];
    if (m > 10) return 1;
    for (int i = 0; i < m * m; ++i) {
        A[i] = a_in[i];
        if (!isfinite(A[i])) return 1;
    }
    for (int i = 0; i < m; ++i) for (int j = 0; j < m; ++j) V[i * m + j] = (i == j) ? 1.0 : 0.0;
    int converged = 0;
    for (int sweep = 0; sweep < 60 && !converged; ++sweep) {
        int rotated = 0;
        for (int p = 0; p < m - 1; ++p) {
            for (int q = p + 1; q < m; ++q) {
                double alpha = 0.0, beta = 0.0, gamma = 0.0;
                for (int i = 0; i < m; ++i) {
                    const double ap = A[i * m + p], aq = A[i * m + q];
                    alpha = alpha + ap * ap;
                    beta = beta + aq * aq;
                    gamma = gamma + ap * aq;
                }
                if (gamma == 0.0 || fabs(gamma) <= 1e-15 * sqrt(alpha * beta)) continue;
                rotated = 1;
                const double zeta = (beta - alpha) / (2.0 * gamma);
                const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / sqrt(1.0 + t * t);
                const double s = c * t;
                for (int i = 0; i < m; ++i) {
                    const double ap = A[i * m + p], aq = A[i * m + q];
                    A[i * m + p] = c * ap - s * aq;
                    A[i * m + q] = s * ap + c * aq;
                    const double vp = V[i * m + p], vq = V[i * m + q];
                    V[i * m + p] = c * vp - s * vq;
                    V[i * m + q] = s * vp + c * vq;
                }
            }
        }
        if (!rotated) converged = 1;
    }
    if (!converged) return 1;
    double sig[10];
    double smax = 0.0, smin = INFINITY;
    for (int j = 0; j < m; ++j) {
        double ss = 0.0;
        for (int i = 0; i < m; ++i) ss = ss + A[i * m + j] * A[i * m + j];
        sig[j] = sqrt(ss);
        if (!isfinite(sig[j])) return 1;
        if (sig[j] > smax) smax = sig[j];
        if (sig[j] < smin) smin = sig[j];
    }
    if (sigma_out) for (int j = 0; j < m; ++j) sigma_out[j] = sig[j];
    if (smin < 1e-10) return 1;
    if (smax / smin > 1e6) return 1;
    /* a = sum_j v_j * (u_j . 1) / sigma_j, u_j = A[:, j] / sigma_j */
    for (int r = 0; r < m; ++r) a_out[r] = 0.0;
    for (int j = 0; j < m; ++j) {
        if (!(sig[j] > 1e-5)) continue;
        double usum = 0.0;
        for (int i = 0; i < m; ++i) usum = usum + A[i * m + j] / sig[j];
        const double w = usum / sig[j];
        for (int r = 0; r < m; ++r) a_out[r] = a_out[r] + V[r * m + j] * w;
    }
    return 0;
}

/* Extreme points, intercepts and normalization, legacy path
 * (normalization.cu:1593-1636).  fn is (M, n) and is normalized in place.
 * Outputs: a_f32 (M,) reciprocal intercepts applied, extreme (M,) indices,
 * return value = 1 when the row-max fallback was used. */
ORC_API int orc_normalize(float* fn, int n, int m, const float* ideal, float* a_f32,
                          int32_t* extreme, double* a_f64)
{
    /* (i) translate, clamp finite values at 0 (normalization.cu:370-395) */
    for (int k = 0; k < m; ++k)
        for (int i = 0; i < n; ++i) {
            float v = fn[(size_t)k * n + i] - ideal[k];
            if (isfinite(v)) v = fmaxf(v, 0.0f);
            fn[(size_t)k * n + i] = v;
        }
    /* (ii) ASF extreme points: ASF_j = max_k f'_k * (k == j ? 1 : 1e6), NaN -> inf,
     * smallest index on ties, -1 when every value is +inf (normalization.cu:647-729,
     * 453-509) */
    int ext[10];
    for (int j = 0; j < m; ++j) {
        float best = INFINITY;
        int bi = -1;
        for (int i = 0; i < n; ++i) {
            float mx = -INFINITY;
            for (int k = 0; k < m; ++k) {
                float v = fn[(size_t)k * n + i];
                if (v != v) v = INFINITY;
                const float w = (k == j) ? 1.0f : 1e6f;
                mx = fmaxf(mx, v * w);
            }
            if (mx < best) { best = mx; bi = i; }
        }
        ext[j] = bi;
        if (extreme) extreme[j] = bi;
    }
    /* (iii) E[i][k] = f'_k(e_i); rows are extreme points (normalization.cu:521-540
     * builds it row-major by objective and cuSOLVER reads it column-major) */
    double E[100] = {0.0}, a[10] = {0.0};
    if (m > 10) return -1;
    for (int i = 0; i < m; ++i)
        for (int k = 0; k < m; ++k)
            E[i * m + k] = (ext[i] >= 0 && ext[i] < n) ? (double)fn[(size_t)k * n + ext[i]]
                                                       : (double)INFINITY;
    const int ill = orc_hyperplane_fp64(E, m, a, NULL);
    if (ill) {
        /* (iv) fallback: 1 / max(rowmax, 1e-6) when rowmax > 1e-10, else 1;
         * NaN excluded from the max (normalization.cu:802-856) */
        for (int k = 0; k < m; ++k) {
            float mx = -INFINITY;
            for (int i = 0; i < n; ++i) {
                float v = fn[(size_t)k * n + i];
                if (v != v) v = -INFINITY;
                mx = fmaxf(mx, v);
            }
            a[k] = (mx > 1e-10f) ? 1.0 / (double)fmaxf(mx, 1e-6f) : 1.0;
        }
    }
    for (int k = 0; k < m; ++k) {
        if (a_f64) a_f64[k] = a[k];
        float af = (float)a[k];
        /* (v) sanitize (normalization.cu:572-583) */
        if (!isfinite(af) || af <= 0.0f) af = 1.0f;
        a_f32[k] = af;
    }
    /* (vi) f'' = f' * a (normalization.cu:598-612) */
    for (int k = 0; k < m; ++k)
        for (int i = 0; i < n; ++i) fn[(size_t)k * n + i] = fn[(size_t)k * n + i] * a_f32[k];
    return ill;
}

/* Perpendicular-distance association (association.cu:85-188, 397-447):
 * dot by an fmaf chain from 0, residual x_m - dot*w_m contracted to one fma
 * by nvcc, d2 by an fmaf chain, argmin with the smallest index on ties,
 * dist = sqrtf(d2), ref = -1 / dist = inf when no d2 compares below +inf. */
ORC_API void orc_associate(const float* fn, int n, int m, const float* w, int k,
                           float* dist, int32_t* ref)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        float x[16];
        for (int c = 0; c < m; ++c) x[c] = fn[(size_t)c * n + i];
        float best = INFINITY;
        int bj = -1;
        for (int j = 0; j < k; ++j) {
            float dot = 0.0f;
            for (int c = 0; c < m; ++c) dot = fmaf(x[c], w[(size_t)c * k + j], dot);
            float d2 = 0.0f;
            for (int c = 0; c < m; ++c) {
                const float r = fmaf(-dot, w[(size_t)c * k + j], x[c]);
                d2 = fmaf(r, r, d2);
            }
            if (d2 < best) { best = d2; bj = j; }
        }
        dist[i] = (best != INFINITY) ? sqrtf(best) : INFINITY;
        ref[i] = bj;
    }
}

/* Stable sort of (key, value) pairs by key (LSD radix, 16-bit digits); stable
 * ordering equals CUB DeviceRadixSort::SortPairs (cuda_ref_segment_ops.cu:51-83). */
static void stable_sort_u64(uint64_t* key, int32_t* val, int n)
{
    uint64_t* k2 = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(n > 0 ? n : 1));
    int32_t* v2 = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    size_t* cnt = (size_t*)malloc(sizeof(size_t) * 65536);
    for (int pass = 0; pass < 4; ++pass) {
        const int sh = pass * 16;
        memset(cnt, 0, sizeof(size_t) * 65536);
        for (int i = 0; i < n; ++i) cnt[(key[i] >> sh) & 0xFFFF]++;
        size_t s = 0;
        for (int b = 0; b < 65536; ++b) { const size_t c = cnt[b]; cnt[b] = s; s += c; }
        for (int i = 0; i < n; ++i) {
            const size_t d = cnt[(key[i] >> sh) & 0xFFFF]++;
            k2[d] = key[i];
            v2[d] = val[i];
        }
        memcpy(key, k2, sizeof(uint64_t) * (size_t)n);
        memcpy(val, v2, sizeof(int32_t) * (size_t)n);
    }
    free(cnt);
    free(v2);
    free(k2);
}

static int cmp_i32(const void* a, const void* b)
{
    const int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    return (x > y) - (x < y);
}

/* Niching (niching.cu:250-400 with cuda_ref_segment_ops.cu:17-134).
 * Inputs are in N_nds space: [0, n_prior) prior fronts, [n_prior, n_nds) last
 * front.  ua, ub: the two u32 draw arrays (length n_last), consumed by sorted
 * position.  Writes n_rem selected N_nds-local indices (in selection order) to
 * out_local.  Association refs of -1 (no finite distance) are counted in no
 * niche and read a niche count of 0; the reference indexes out of bounds there. */
ORC_API void orc_niching(const int32_t* ref, const float* dist, int n_prior, int n_last,
                         int k, int n_rem, const uint32_t* ua, const uint32_t* ub,
                         int32_t* out_local, int32_t* rho_out)
{
    int32_t* rho = (int32_t*)calloc((size_t)k, sizeof(int32_t));
    for (int i = 0; i < n_prior; ++i)
        if (ref[i] >= 0 && ref[i] < k) rho[ref[i]]++;
    if (rho_out) memcpy(rho_out, rho, sizeof(int32_t) * (size_t)k);
    if (n_rem <= 0 || n_last <= 0) { free(rho); return; }
    const int L = n_last;
    uint64_t* key = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)L);
    int32_t* idx = (int32_t*)malloc(sizeof(int32_t) * (size_t)L);
    int32_t* rank = (int32_t*)malloc(sizeof(int32_t) * (size_t)L);
    uint32_t dbits;
    /* phase 2: key = ref << 32 | bits(dist), value = local index */
    for (int t = 0; t < L; ++t) {
        memcpy(&dbits, &dist[n_prior + t], 4);
        key[t] = ((uint64_t)(uint32_t)ref[n_prior + t] << 32) | dbits;
        idx[t] = t;
    }
    stable_sort_u64(key, idx, L);
    /* phase 3: local index within each ref segment */
#define SEG_RANK()                                                            \
    do {                                                                      \
        int head = 0;                                                         \
        for (int t = 0; t < L; ++t) {                                         \
            if (t == 0 || (uint32_t)(key[t] >> 32) != (uint32_t)(key[t - 1] >> 32)) head = t; \
            rank[t] = t - head;                                               \
        }                                                                     \
    } while (0)
    SEG_RANK();
    /* phase 4: rank key = ref << 32 | (rho == 0 && local == 0 ? 0 : ua | 1) */
    for (int t = 0; t < L; ++t) {
        const int r = (int)(int32_t)(uint32_t)(key[t] >> 32);
        const int rh = (r >= 0 && r < k) ? rho[r] : 0;
        const uint32_t low = (rh == 0 && rank[t] == 0) ? 0u : (ua[t] | 1u);
        key[t] = ((uint64_t)(uint32_t)r << 32) | low;
    }
    stable_sort_u64(key, idx, L);
    SEG_RANK();
#undef SEG_RANK
    /* phase 5: key = (rho + rank) << 32 | ub */
    for (int t = 0; t < L; ++t) {
        const int r = (int)(int32_t)(uint32_t)(key[t] >> 32);
        const int rh = (r >= 0 && r < k) ? rho[r] : 0;
        key[t] = ((uint64_t)(uint32_t)(rh + rank[t]) << 32) | ub[t];
    }
    stable_sort_u64(key, idx, L);
    /* phase 6: first n_rem */
    for (int t = 0; t < n_rem; ++t) out_local[t] = n_prior + idx[t];
    free(rank);
    free(idx);
    free(key);
}

/* NSGA-III environmental selection (nsga3_environment_selector.cu:289-538) on a
 * merged population: fm (M, 2n) and cvm (2n,) hold parents in columns [0, n) and
 * offspring in [n, 2n).  ideal_last (M,) is the running ideal point and is
 * updated.  ua / ub (length >= n_last) are the niching draws.
 * Outputs: chosen (n,) merged indices sorted ascending; fronts_mix (2n,);
 * stats = {max_front, n_prior, n_last, ill_conditioned}. */
ORC_API int orc_nsga3_select(const float* fm, const float* cvm, int n, int m,
                             const float* w, int k, float* ideal_last,
                             const uint32_t* ua, const uint32_t* ub,
                             int32_t* chosen, int32_t* fronts_mix, int32_t* stats,
                             float* a_out, int32_t* ext_out, double* a64_out)
{
    const int nm = 2 * n;
    const int max_front = orc_nds(fm, cvm, nm, m, n, fronts_mix);
    int n_prior = 0, n_last = 0;
    for (int i = 0; i < nm; ++i) {
        if (fronts_mix[i] >= 0 && fronts_mix[i] < max_front) ++n_prior;
        else if (fronts_mix[i] == max_front) ++n_last;
    }
    const int n_nds = n_prior + n_last, n_rem = n - n_prior;
    int32_t* gidx = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n_nds > 0 ? n_nds : 1));
    int p = 0, l = n_prior;
    for (int i = 0; i < nm; ++i) {
        if (fronts_mix[i] >= 0 && fronts_mix[i] < max_front) gidx[p++] = i;
        else if (fronts_mix[i] == max_front) gidx[l++] = i;
    }
    /* ideal point: running min over the initial parents and every offspring
     * batch (nsga3_environment_selector.cu:418-429, ideal_point.cu:137-165) */
    float zoff[16];
    row_min(fm, nm, n, n, m, zoff);
    for (int c = 0; c < m; ++c) ideal_last[c] = fminf(ideal_last[c], zoff[c]);
    int ill = -1;
    for (int i = 0; i < n_prior; ++i) chosen[i] = gidx[i];
    if (n_rem > 0) {
        float* fn = (float*)malloc(sizeof(float) * (size_t)m * (n_nds > 0 ? n_nds : 1));
        for (int c = 0; c < m; ++c)
            for (int i = 0; i < n_nds; ++i) fn[(size_t)c * n_nds + i] = fm[(size_t)c * nm + gidx[i]];
        float a[16];
        int32_t ext[16];
        double a64[16];
        ill = orc_normalize(fn, n_nds, m, ideal_last, a, ext, a64);
        if (a_out) memcpy(a_out, a, sizeof(float) * (size_t)m);
        if (a64_out) memcpy(a64_out, a64, sizeof(double) * (size_t)m);
        if (ext_out)
            for (int c = 0; c < m; ++c) ext_out[c] = (ext[c] >= 0 && ext[c] < n_nds) ? gidx[ext[c]] : -1;
        float* dist = (float*)malloc(sizeof(float) * (size_t)n_nds);
        int32_t* ref = (int32_t*)malloc(sizeof(int32_t) * (size_t)n_nds);
        orc_associate(fn, n_nds, m, w, k, dist, ref);
        int32_t* loc = (int32_t*)malloc(sizeof(int32_t) * (size_t)n_rem);
        orc_niching(ref, dist, n_prior, n_last, k, n_rem, ua, ub, loc, NULL);
        for (int t = 0; t < n_rem; ++t) chosen[n_prior + t] = gidx[loc[t]];
        free(loc);
        free(ref);
        free(dist);
        free(fn);
    }
    qsort(chosen, (size_t)n, sizeof(int32_t), cmp_i32);
    if (stats) { stats[0] = max_front; stats[1] = n_prior; stats[2] = n_last; stats[3] = ill; }
    free(gidx);
    return max_front;
}

/* ======================================================================== */
/* Reproduction                                                              */
/* ======================================================================== */

/* Linear parameter annealing (sbx_crossover.cu:12-18, polynomial_mutation.cu:12-18). */
ORC_API float orc_annealed(float begin, float end, int gen, int total)
{
    float progress = 1.0f;
    if (total > 0) {
        progress = (float)gen / (float)total;
        if (progress < 0.0f) progress = 0.0f;
        if (progress > 1.0f) progress = 1.0f;
    }
    return begin + (end - begin) * progress;
}

/* Binary tournament (mating_selection.cu:11-35). */
ORC_API void orc_tournament(const int32_t* ca, const int32_t* cb, int count,
                            const int32_t* rank, const float* cv, int32_t* winners)
{
    for (int i = 0; i < count; ++i) {
        const int a = ca[i], b = cb[i];
        const float cva = cv[a], cvb = cv[b];
        int w;
        if (cva == 0.0f && cvb != 0.0f) w = a;
        else if (cva != 0.0f && cvb == 0.0f) w = b;
        else if (cva != 0.0f && cvb != 0.0f) w = (cva <= cvb) ? a : b;
        else w = (rank[a] <= rank[b]) ? a : b;
        winners[i] = w;
    }
}

/* SBX for one (pair, d) element (sbx_crossover.cu:20-70).  In
 * midpoint +/- delta both operands are products; the NVPTX combiner fuses the
 * first one, 0.5f * (x1 + x2), whose product is exact, so the result equals
 * the unfused sum written here. */
static inline void sbx_one(float x1, float x2, float ub_, float up, int sign, float uc,
                           float lo, float hi, float eta, float prob, float copy_prob,
                           float* ca, float* cb)
{
    if (up >= prob || uc < copy_prob) { *ca = x1; *cb = x2; return; }
    const float power = 1.0f / (1.0f + eta);
    float beta = (ub_ < 0.5f) ? orc_powf_root(2.0f * ub_, power)
                              : orc_powf_root(1.0f / (2.0f - 2.0f * ub_), power);
    beta = beta * (float)(2 * sign - 1);
    const float mid = 0.5f * (x1 + x2);
    const float delta = 0.5f * beta * (x1 - x2);
    *ca = fminf(fmaxf(mid + delta, lo), hi);
    *cb = fminf(fmaxf(mid - delta, lo), hi);
}

/* Polynomial mutation of one element (polynomial_mutation.cu:20-60).  The
 * value_mut sums have a product on both sides; the fused one (the first,
 * 2u or 2(1-u)) is exact, so they evaluate as the unfused sums below.  The
 * final value + delta * range is fused. */
static inline float pm_one(float value, float ug, float ud, float lo, float hi, float eta,
                           float rate)
{
    if (ug >= rate) return value;
    const float range = hi - lo;
    const float normalized = (value - lo) / range;
    const float exponent = eta + 1.0f;
    float delta;
    if (ud <= 0.5f) {
        const float xy = 1.0f - normalized;
        const float vm = 2.0f * ud + (1.0f - 2.0f * ud) * orc_powf(xy, exponent);
        delta = orc_powf_root(vm, 1.0f / exponent) - 1.0f;
    } else {
        const float xy = normalized;
        const float vm = 2.0f * (1.0f - ud) + 2.0f * (ud - 0.5f) * orc_powf(xy, exponent);
        delta = 1.0f - orc_powf_root(vm, 1.0f / exponent);
    }
    return fminf(fmaxf(fmaf(delta, range, value), lo), hi);
}

/* SBX over all pairs with injected draws (sbx_crossover.cu:123-190); draw
 * arrays are indexed pair * D + d; odd n_off copies the last parent. */
ORC_API void orc_sbx(const float* parents, const int32_t* idx, const float* bounds,
                     const float* u_beta, const float* u_p, const int32_t* sign,
                     const float* u_copy, float* off, float eta, float prob, float copy_prob,
                     int n_off, int d)
{
    const int pairs = n_off / 2;
#pragma omp parallel for schedule(static)
    for (int pr = 0; pr < pairs; ++pr) {
        const float* x1 = parents + (size_t)idx[2 * pr] * d;
        const float* x2 = parents + (size_t)idx[2 * pr + 1] * d;
        for (int j = 0; j < d; ++j) {
            const size_t e = (size_t)pr * d + j;
            sbx_one(x1[j], x2[j], u_beta[e], u_p[e], sign[e], u_copy[e], bounds[2 * j],
                    bounds[2 * j + 1], eta, prob, copy_prob, &off[(size_t)(2 * pr) * d + j],
                    &off[(size_t)(2 * pr + 1) * d + j]);
        }
    }
    if (n_off % 2)
        for (int j = 0; j < d; ++j)
            off[(size_t)(n_off - 1) * d + j] = parents[(size_t)idx[n_off - 1] * d + j];
}

/* PM over all elements with injected draws indexed row * D + d
 * (polynomial_mutation.cu:95-124). */
ORC_API void orc_pm(const float* u_gate, const float* u_delta, float* off, const float* bounds,
                    float eta, float prob, int n, int d)
{
    const float rate = prob / (float)d;
#pragma omp parallel for schedule(static)
    for (int r = 0; r < n; ++r)
        for (int j = 0; j < d; ++j) {
            const size_t e = (size_t)r * d + j;
            off[e] = pm_one(off[e], u_gate[e], u_delta[e], bounds[2 * j], bounds[2 * j + 1],
                            eta, rate);
        }
}

/* ======================================================================== */
/* Built-in objectives                                                       */
/* ======================================================================== */

/* Quadratic-center objectives f_c(x) = sum_d (x_d - c_cd)^2 with the library's
 * fixed summation order: 64 strided partial sums (partial j takes d = j, j+64,
 * ... ascending, each step fmaf(t, t, acc)), then a pairwise tree
 * acc[j] += acc[j + s] for s = 32, 16, ..., 1.  The reference has no such
 * kernel; evaluation goes through IProblemEvaluator::evaluate
 * (NI/problem/problem_evaluator.cuh:10-31). */
ORC_API void orc_eval_quadratic(const float* x, int n, int d, const float* centers, int m,
                                float* f, int ldf, float* cv)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        for (int c = 0; c < m; ++c) {
            float acc[64];
            for (int j = 0; j < 64; ++j) acc[j] = 0.0f;
            for (int j = 0; j < d; ++j) {
                const float t = x[(size_t)i * d + j] - centers[(size_t)c * d + j];
                acc[j & 63] = fmaf(t, t, acc[j & 63]);
            }
            for (int s = 32; s >= 1; s >>= 1)
                for (int j = 0; j < s; ++j) acc[j] = acc[j] + acc[j + s];
            f[(size_t)c * ldf + i] = acc[0];
        }
        if (cv) cv[i] = 0.0f;
    }
}

/* cos t in double with one fixed operation sequence (the library's cos_d):
 * t = k (pi / 2) + r with k = rint(t 2 / pi) and pi / 2 as a two-term sum, then
 * the Taylor series of sin r (to r^17) or cos r (to r^18) by k's quadrant. */
static double orc_cos_d(double t)
{
    const double k = nearbyint(t * 6.36619772367581382433e-01);
    double r = fma(-k, 1.57079632679489655800e+00, t);
    r = fma(-k, 6.12323399573676603587e-17, r);
    const double r2 = r * r;
    const int q = (int)k & 3;
    double v;
    if (q & 1) {
        double p = 1.0 / 355687428096000.0;
        p = fma(p, -r2, 1.0 / 1307674368000.0);
        p = fma(p, -r2, 1.0 / 6227020800.0);
        p = fma(p, -r2, 1.0 / 39916800.0);
        p = fma(p, -r2, 1.0 / 362880.0);
        p = fma(p, -r2, 1.0 / 5040.0);
        p = fma(p, -r2, 1.0 / 120.0);
        p = fma(p, -r2, 1.0 / 6.0);
        v = fma(-r * r2, p, r);
    } else {
        double p = 1.0 / 6402373705728000.0;
        p = fma(p, -r2, 1.0 / 20922789888000.0);
        p = fma(p, -r2, 1.0 / 87178291200.0);
        p = fma(p, -r2, 1.0 / 479001600.0);
        p = fma(p, -r2, 1.0 / 3628800.0);
        p = fma(p, -r2, 1.0 / 40320.0);
        p = fma(p, -r2, 1.0 / 720.0);
        p = fma(p, -r2, 1.0 / 24.0);
        p = fma(p, -r2, 0.5);
        v = fma(-r2, p, 1.0);
    }
    return (q == 1 || q == 2) ? -v : v;
}

ORC_API float orc_cosf_det(float x) { return (float)orc_cos_d((double)x); }

/* DTLZ1 (problem = 0) or C1-DTLZ1 (problem = 1) of n rows of x (n, d) into f
 * (m, ldf) and cv (dtlz1_func_kernel, dtlz_kernels.cu:17-80, and the C1 rule
 * of compute_cv, :824-860), in the library's summation order: 64 strided
 * partial sums of g (partial j takes i = m-1+j, m-1+j+64, ... ascending), a
 * pairwise tree of the partials, cosf of the distance term as orc_cosf_det.
 * cv = 0 for DTLZ1. */
ORC_API void orc_eval_dtlz1(const float* x, int n, int d, int m, float* f, int ldf, float* cv, int problem)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        const float* xr = x + (size_t)i * d;
        float acc[64];
        for (int j = 0; j < 64; ++j) acc[j] = 0.0f;
        for (int j = 0; j < 64; ++j)
            for (int t = m - 1 + j; t < d; t += 64) {
                const float v = xr[t] - 0.5f;
                acc[j] = acc[j] + (v * v - orc_cosf_det((float)(20.0 * 3.14159265358979323846 * (double)v)));
            }
        for (int sft = 32; sft >= 1; sft >>= 1)
            for (int j = 0; j < sft; ++j) acc[j] = acc[j] + acc[j + sft];
        const float s = acc[0];
        const int k = d - (m - 1);
        float fv[16];
        for (int c = 0; c < m; ++c) {
            float out = 0.5f * (1.0f + 100.0f * ((float)k + s));
            for (int j = 0; j < m - 1 - c; ++j) out *= fminf(fmaxf(xr[j], 0.0f), 1.0f);
            if (c > 0) out *= fmaxf(0.0f, 1.0f - xr[m - 1 - c]);
            f[(size_t)c * ldf + i] = out;
            fv[c] = out;
        }
        if (problem == 1) {
            float c = fv[m - 1] / 0.6f;
            for (int j = 0; j < m - 1; ++j) c += fv[j] / 0.5f;
            cv[i] = fmaxf(c - 1.0f, 0.0f);
        } else if (cv) {
            cv[i] = 0.0f;
        }
    }
}

/* ======================================================================== */
/* In-kernel draw generator, restated for the CPU                            */
/* ======================================================================== */

ORC_API void orc_draw_init(uint64_t seed, int n, int d, const float* bounds, float* x)
{
#pragma omp parallel for schedule(static)
    for (int r = 0; r < n; ++r)
        for (int j = 0; j < d; ++j) {
            uint32_t o[4];
            const uint64_t e = (uint64_t)r * d + j;
            philox_at(seed, e, 0, ST_INIT, o);
            const float lo = bounds[2 * j], hi = bounds[2 * j + 1];
            /* population.cu:12-24: lower + u * (upper - lower), contracted */
            x[e] = fmaf(u01(o[0]), hi - lo, lo);
        }
}

ORC_API void orc_draw_mating(uint64_t seed, uint32_t gen, int count, int active, int32_t* ca,
                             int32_t* cb)
{
    for (int i = 0; i < count; ++i) {
        uint32_t o[4];
        philox_at(seed, (uint64_t)i, gen, ST_MATE, o);
        ca[i] = (int32_t)(o[0] % (uint32_t)active);
        if (cb) cb[i] = (int32_t)(o[1] % (uint32_t)active);
    }
}

/* In-kernel draw layout of one (pair, variable) element e = pair * d + j (the
 * device's sbx_pm_elem): one Philox block of stream ST_SBX gives u_beta,
 * u_prob, u_copy (top 24 bits of words 0..2), the SBX direction (bit 7 of
 * word 3), child a's mutation gate (top 24 bits of word 3) and child b's gate
 * (the low bytes of words 0..2, most significant first).  Mutation deltas come
 * from a block of stream ST_PM (child a word 0, child b word 1).  The device
 * draws that block only when a gate opens; the values it uses are the same.
 * The tail row of an odd population takes child a's draws at pairs * d + j. */
static float gate_b(const uint32_t o[4])
{
    return (float)(((o[0] & 0xFFu) << 16) | ((o[1] & 0xFFu) << 8) | (o[2] & 0xFFu)) * 5.9604644775390625e-8f;
}

ORC_API void orc_draw_sbx(uint64_t seed, uint32_t gen, int pairs, int d, float* ub, float* up,
                          int32_t* sign, float* uc)
{
#pragma omp parallel for schedule(static)
    for (int pr = 0; pr < pairs; ++pr)
        for (int j = 0; j < d; ++j) {
            uint32_t o[4];
            const uint64_t e = (uint64_t)pr * d + j;
            philox_at(seed, e, gen, ST_SBX, o);
            ub[e] = u01(o[0]);
            up[e] = u01(o[1]);
            uc[e] = u01(o[2]);
            sign[e] = (int32_t)((o[3] >> 7) & 1u);
        }
}

ORC_API void orc_draw_pm(uint64_t seed, uint32_t gen, int n, int d, float* ug, float* ud)
{
#pragma omp parallel for schedule(static)
    for (int r = 0; r < n; ++r)
        for (int j = 0; j < d; ++j) {
            uint32_t os[4], op[4];
            const uint64_t e = (uint64_t)(r >> 1) * d + j;
            philox_at(seed, e, gen, ST_SBX, os);
            philox_at(seed, e, gen, ST_PM, op);
            const size_t i = (size_t)r * d + j;
            if ((r & 1) == 0) {
                ug[i] = u01(os[3]);
                ud[i] = u01(op[0]);
            } else {
                ug[i] = gate_b(os);
                ud[i] = u01(op[1]);
            }
        }
}

ORC_API void orc_draw_niche(uint64_t seed, uint32_t gen, int count, uint32_t* ua, uint32_t* ub)
{
    for (int t = 0; t < count; ++t) {
        ua[t] = orc_splitmix32(niche_base(gen, 0) + (uint64_t)t, seed);
        ub[t] = orc_splitmix32(niche_base(gen, 1) + (uint64_t)t, seed);
    }
}

/* ======================================================================== */
/* Full NSGA-III run (generation loop order of NS/algorithm/algorithm.cu:219-286) */
/* ======================================================================== */

/* NSGA-III with tournament mating, SBX, PM and quadratic-center objectives,
 * draws from the in-kernel generator.  x (n, d) receives the initial population
 * unless init_x is given; on return x/f/cv/fronts hold the final population.
 * Returns the wall time in ms of the generation loop (after the initial
 * evaluation, the reference's timing boundary algorithm.cu:313-329). */
/* Objectives of the run loops: 0 quadratic centers, 1 DTLZ1, 2 C1-DTLZ1 whose
 * constraints count from the generation g with (g + 1) / gens >= ratio
 * (MOEAStdTestEvaluator::is_cv_active, dtlz_kernels.cu:1300-1309; ratio <= 0:
 * always, the initial population included). */
typedef struct OrcProblem { int kind; const float* centers; float ratio; } OrcProblem;

static int orc_cv_active(const OrcProblem* pb, int g, int gens)
{
    if (pb->kind != 2 || !(pb->ratio > 0.0f)) return 1;
    if (g < 0) return 0;
    return (float)(g + 1) / (float)(gens > 0 ? gens : 1) >= pb->ratio;
}

static void orc_evaluate(const OrcProblem* pb, const float* x, int n, int d, int m, float* f, int ldf,
                         float* cv, int g, int gens)
{
    if (pb->kind == 0) {
        orc_eval_quadratic(x, n, d, pb->centers, m, f, ldf, cv);
        return;
    }
    orc_eval_dtlz1(x, n, d, m, f, ldf, cv, pb->kind == 2 ? 1 : 0);
    if (!orc_cv_active(pb, g, gens))
        for (int i = 0; i < n; ++i) cv[i] = 0.0f;
}

static double nsga3_run_impl(int n, int m, int d, int k_req, int partitions, int gens,
                             uint64_t seed, const float* bounds, const OrcProblem* pb,
                             float sbx_eta0, float sbx_eta1, float sbx_prob, float sbx_copy,
                             float pm_eta0, float pm_eta1, float pm_prob,
                             const float* init_x, float* x, float* f, float* cv,
                             int32_t* fronts, int gen_count_limit);

ORC_API double orc_nsga3_run(int n, int m, int d, int k_req, int partitions, int gens,
                             uint64_t seed, const float* bounds, const float* centers,
                             float sbx_eta0, float sbx_eta1, float sbx_prob, float sbx_copy,
                             float pm_eta0, float pm_eta1, float pm_prob,
                             const float* init_x, float* x, float* f, float* cv,
                             int32_t* fronts, int gen_count_limit)
{
    const OrcProblem pb = {0, centers, 0.0f};
    return nsga3_run_impl(n, m, d, k_req, partitions, gens, seed, bounds, &pb, sbx_eta0, sbx_eta1, sbx_prob,
                          sbx_copy, pm_eta0, pm_eta1, pm_prob, init_x, x, f, cv, fronts, gen_count_limit);
}

/* NSGA-III on DTLZ1 (problem 1) or C1-DTLZ1 (problem 2) with the constraint
 * activation ratio: before the switch the violations are 0; at the switch
 * generation (0 when they count from the start) the parents are evaluated
 * again and their fronts re-sorted before reproduction
 * (DTLZProblem::prepare_parent, dtlz_kernels.cu:1311-1394; algorithm.cu:230-241). */
ORC_API double orc_nsga3_run_dtlz1(int n, int m, int d, int k_req, int partitions, int gens,
                                   uint64_t seed, const float* bounds, int problem, float ratio,
                                   float sbx_eta0, float sbx_eta1, float sbx_prob, float sbx_copy,
                                   float pm_eta0, float pm_eta1, float pm_prob,
                                   float* x, float* f, float* cv, int32_t* fronts, int gen_count_limit)
{
    const OrcProblem pb = {problem, NULL, ratio};
    return nsga3_run_impl(n, m, d, k_req, partitions, gens, seed, bounds, &pb, sbx_eta0, sbx_eta1, sbx_prob,
                          sbx_copy, pm_eta0, pm_eta1, pm_prob, NULL, x, f, cv, fronts, gen_count_limit);
}

static double nsga3_run_impl(int n, int m, int d, int k_req, int partitions, int gens,
                             uint64_t seed, const float* bounds, const OrcProblem* pb,
                             float sbx_eta0, float sbx_eta1, float sbx_prob, float sbx_copy,
                             float pm_eta0, float pm_eta1, float pm_prob,
                             const float* init_x, float* x, float* f, float* cv,
                             int32_t* fronts, int gen_count_limit)
{
    const int k = orc_das_dennis(k_req, m, partitions, partitions > 0 ? 1 : 0, NULL, NULL);
    float* w = (float*)malloc(sizeof(float) * (size_t)m * k);
    orc_das_dennis(k_req, m, partitions, partitions > 0 ? 1 : 0, w, NULL);
    const int nm = 2 * n;
    float* xm = (float*)malloc(sizeof(float) * (size_t)nm * d);
    float* fm = (float*)malloc(sizeof(float) * (size_t)m * nm);
    float* cvm = (float*)malloc(sizeof(float) * (size_t)nm);
    int32_t* fr_mix = (int32_t*)malloc(sizeof(int32_t) * (size_t)nm);
    int32_t* ca = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int32_t* cb = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int32_t* win = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    const int pairs = n / 2;
    float* ubeta = (float*)malloc(sizeof(float) * (size_t)(pairs > 0 ? pairs : 1) * d);
    float* up = (float*)malloc(sizeof(float) * (size_t)(pairs > 0 ? pairs : 1) * d);
    int32_t* sg = (int32_t*)malloc(sizeof(int32_t) * (size_t)(pairs > 0 ? pairs : 1) * d);
    float* uc = (float*)malloc(sizeof(float) * (size_t)(pairs > 0 ? pairs : 1) * d);
    float* ug = (float*)malloc(sizeof(float) * (size_t)n * d);
    float* ud = (float*)malloc(sizeof(float) * (size_t)n * d);
    uint32_t* ua = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)nm);
    uint32_t* ubn = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)nm);
    int32_t* chosen = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    float* off = xm + (size_t)n * d;

    if (init_x) memcpy(x, init_x, sizeof(float) * (size_t)n * d);
    else orc_draw_init(seed, n, d, bounds, x);
    orc_evaluate(pb, x, n, d, m, f, n, cv, -1, gens);
    /* prepare: parent fronts over N with stop at N and the initial ideal point
     * (nsga3_environment_selector.cu:236-287) */
    orc_nds(f, cv, n, m, n, fronts);
    float ideal[16];
    row_min(f, n, 0, n, m, ideal);

    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    const int g_end = (gen_count_limit > 0 && gen_count_limit < gens) ? gen_count_limit : gens;
    for (int g = 0; g < g_end; ++g) {
        const int sw = pb->kind == 2 && (g == 0 ? orc_cv_active(pb, 0, gens)
                                                : (!orc_cv_active(pb, g - 1, gens) && orc_cv_active(pb, g, gens)));
        if (sw) {
            orc_evaluate(pb, x, n, d, m, f, n, cv, g, gens);
            orc_nds(f, cv, n, m, n, fronts);
        }
        const float eta_c = orc_annealed(sbx_eta0, sbx_eta1, g, gens);
        const float eta_m = orc_annealed(pm_eta0, pm_eta1, g, gens);
        orc_draw_mating(seed, (uint32_t)g, n, n, ca, cb);
        orc_tournament(ca, cb, n, fronts, cv, win);
        orc_draw_sbx(seed, (uint32_t)g, pairs, d, ubeta, up, sg, uc);
        orc_sbx(x, win, bounds, ubeta, up, sg, uc, off, eta_c, sbx_prob, sbx_copy, n, d);
        orc_draw_pm(seed, (uint32_t)g, n, d, ug, ud);
        orc_pm(ug, ud, off, bounds, eta_m, pm_prob, n, d);
        /* merged objectives: parents then offspring */
        for (int c = 0; c < m; ++c) memcpy(fm + (size_t)c * nm, f + (size_t)c * n, sizeof(float) * n);
        memcpy(cvm, cv, sizeof(float) * n);
        orc_evaluate(pb, off, n, d, m, fm + n, nm, cvm + n, g, gens);
        memcpy(xm, x, sizeof(float) * (size_t)n * d);
        orc_draw_niche(seed, (uint32_t)g, nm, ua, ubn);
        orc_nsga3_select(fm, cvm, n, m, w, k, ideal, ua, ubn, chosen, fr_mix, NULL, NULL, NULL, NULL);
        for (int i = 0; i < n; ++i) {
            const int s = chosen[i];
            memcpy(x + (size_t)i * d, xm + (size_t)s * d, sizeof(float) * d);
            for (int c = 0; c < m; ++c) f[(size_t)c * n + i] = fm[(size_t)c * nm + s];
            cv[i] = cvm[s];
            fronts[i] = fr_mix[s];
        }
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    free(chosen); free(ubn); free(ua); free(ud); free(ug); free(uc); free(sg); free(up);
    free(ubeta); free(win); free(cb); free(ca); free(fr_mix); free(cvm); free(fm); free(xm);
    free(w);
    return (double)(t1.tv_sec - t0.tv_sec) * 1e3 + (double)(t1.tv_nsec - t0.tv_nsec) * 1e-6;
}

/* ======================================================================== */
/* RVEA                                                                      */
/* ======================================================================== */

/* gamma_j = acos(max_{k != j} v_j . v_k), fmaf dot chain; pi when K == 1
 * (angle_penalized_distance.cu:196-362). */
ORC_API void orc_rvea_gamma(const float* v, int m, int k, float* gamma)
{
    for (int j = 0; j < k; ++j) {
        float mx = -INFINITY;
        for (int q = 0; q < k; ++q) {
            if (q == j) continue;
            float dot = 0.0f;
            for (int c = 0; c < m; ++c) dot = fmaf(v[(size_t)c * k + j], v[(size_t)c * k + q], dot);
            mx = fmaxf(mx, dot);
        }
        if (mx > -INFINITY) {
            const float cl = fminf(fmaxf(mx, 0.0f), 1.0f);
            gamma[j] = orc_acosf(cl);
        } else {
            gamma[j] = 3.14159265358979323846f;
        }
    }
}

/* APD penalty scale M * (t / G)^alpha evaluated on the host
 * (angle_penalized_distance.cu:159-184); t = generation + 1
 * (rvea_environment_selector.cu:239). */
ORC_API float orc_rvea_penalty(int m, int t, int max_gen, float alpha)
{
    float total = (float)max_gen;
    if (total < 1.0f) total = 1.0f;
    float progress = (float)t / total;
    if (progress < 0.0f) progress = 0.0f;
    if (progress > 1.0f) progress = 1.0f;
    return (float)m * (float)pow((double)progress, (double)alpha);
}

/* RVEA environmental selection (rvea_environment_selector.cu:163-287) on a merged
 * population fm (M, 2n), cvm (2n,).  active = live parent slots [0, active).
 * Writes the selected merged indices sorted ascending to selected (the first
 * return-value entries) and returns the new active count.  Optional outputs:
 * ref_out, theta_out, apd_out (2n,). */
ORC_API int orc_rvea_select(const float* fm, const float* cvm, int n, int m, const float* v,
                            const float* gamma, int k, int active, float penalty,
                            int32_t* selected, int32_t* ref_out, float* theta_out,
                            float* apd_out)
{
    const int nm = 2 * n;
    float zmin[16];
    row_min(fm, nm, 0, nm, m, zmin);
    float* tf = (float*)malloc(sizeof(float) * (size_t)m * nm);
    for (int c = 0; c < m; ++c)
        for (int i = 0; i < nm; ++i) {
            /* objective_translation.cu:87-100 */
            tf[(size_t)c * nm + i] = fmaxf(fm[(size_t)c * nm + i] - zmin[c], 1e-32f);
        }
    int32_t* ref = (int32_t*)malloc(sizeof(int32_t) * (size_t)nm);
    float* theta = (float*)malloc(sizeof(float) * (size_t)nm);
    float* apd = (float*)malloc(sizeof(float) * (size_t)nm);
    for (int i = 0; i < nm; ++i) {
        /* reference_partition.cu:47-161, 236-270 */
        float x[16];
        float ss = 0.0f;
        for (int c = 0; c < m; ++c) { x[c] = tf[(size_t)c * nm + i]; ss = fmaf(x[c], x[c], ss); }
        const float inv = 1.0f / sqrtf(ss + 1e-12f);
        float best = -INFINITY;
        int bj = -1;
        for (int j = 0; j < k; ++j) {
            float dot = 0.0f;
            for (int c = 0; c < m; ++c) dot = fmaf(x[c], v[(size_t)c * k + j], dot);
            const float cs = dot * inv;
            if (cs > best) { best = cs; bj = j; }
        }
        if (best > -INFINITY) theta[i] = orc_acosf(fminf(fmaxf(best, 0.0f), 1.0f));
        else theta[i] = INFINITY;
        ref[i] = bj;
    }
    /* dead parent slots [active, n) point at K (reference_partition.cu:320-335) */
    for (int i = active; i < n; ++i) ref[i] = k;
    for (int i = 0; i < nm; ++i) {
        /* angle_penalized_distance.cu:44-97 */
        float ns = 0.0f;
        for (int c = 0; c < m; ++c) { const float t = tf[(size_t)c * nm + i]; ns = fmaf(t, t, ns); }
        const float nrm = sqrtf(ns);
        if (nrm > 1e-12f) {
            const int r = ref[i];
            if (r < 0 || r >= k) { apd[i] = 1e30f; continue; }
            const float th = fminf(fmaxf(theta[i], 0.0f), 3.14159265358979323846f);
            const float gm = fmaxf(gamma[r], 1e-6f);
            const float val = (1.0f + penalty * th / gm) * nrm;
            apd[i] = (!isfinite(val) || val < 0.0f) ? 1e30f : val;
        } else {
            apd[i] = 0.0f;
        }
    }
    /* elitism (elitism_selection.cu:60-125, 194-271): key = ref << 32 |
     * infeasible << 31 | metric bits; first entry of each valid ref segment wins */
    uint64_t* key = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)nm);
    int32_t* idx = (int32_t*)malloc(sizeof(int32_t) * (size_t)nm);
    for (int i = 0; i < nm; ++i) {
        const int feas = cvm[i] <= 1e-8f;
        float met = feas ? apd[i] : cvm[i];
        if (!(isfinite(met) && met >= 0.0f)) met = 1e30f;
        uint32_t mb;
        memcpy(&mb, &met, 4);
        mb &= 0x7FFFFFFFu;
        key[i] = ((uint64_t)(uint32_t)ref[i] << 32) | ((uint64_t)(feas ? 0u : 1u) << 31) | mb;
        idx[i] = i;
    }
    stable_sort_u64(key, idx, nm);
    int32_t* winners = (int32_t*)malloc(sizeof(int32_t) * (size_t)nm);
    int nw = 0;
    for (int t = 0; t < nm; ++t) {
        const int r = (int)(int32_t)(uint32_t)(key[t] >> 32);
        const int head = (t == 0) || ((uint32_t)(key[t] >> 32) != (uint32_t)(key[t - 1] >> 32));
        if (head && r >= 0 && r < k) winners[nw++] = idx[t];
    }
    int n_active;
    if (nw == 0) {
        for (int t = 0; t < n; ++t) selected[t] = idx[t];
        n_active = n;
    } else if (nw <= n) {
        for (int t = 0; t < nw; ++t) selected[t] = winners[t];
        n_active = nw;
    } else {
        /* truncation key: infeasible << 63 | metric << 32 | index
         * (elitism_selection.cu:130-160) */
        uint64_t* tk = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)nw);
        int32_t* ti = (int32_t*)malloc(sizeof(int32_t) * (size_t)nw);
        for (int t = 0; t < nw; ++t) {
            const int g = winners[t];
            const int feas = cvm[g] <= 1e-8f;
            float met = feas ? apd[g] : cvm[g];
            if (!(isfinite(met) && met >= 0.0f)) met = 1e30f;
            uint32_t mb;
            memcpy(&mb, &met, 4);
            mb &= 0x7FFFFFFFu;
            tk[t] = ((uint64_t)(feas ? 0u : 1u) << 63) | ((uint64_t)mb << 32) | (uint32_t)g;
            ti[t] = t;
        }
        stable_sort_u64(tk, ti, nw);
        for (int t = 0; t < n; ++t) selected[t] = winners[ti[t]];
        free(ti);
        free(tk);
        n_active = n;
    }
    qsort(selected, (size_t)n_active, sizeof(int32_t), cmp_i32);
    if (ref_out) memcpy(ref_out, ref, sizeof(int32_t) * (size_t)nm);
    if (theta_out) memcpy(theta_out, theta, sizeof(float) * (size_t)nm);
    if (apd_out) memcpy(apd_out, apd, sizeof(float) * (size_t)nm);
    free(winners); free(idx); free(key); free(apd); free(theta); free(ref); free(tf);
    return n_active;
}

/* Reference-vector adaptation V = normalize(V0 * (zmax - zmin)) over the active
 * columns (rvea_reference_adaptation.cu:19-232). */
ORC_API void orc_rvea_adapt(const float* v0, const float* f, int n_active, int ld, int m, int k,
                            float* v_out)
{
    float rng[16];
    for (int c = 0; c < m; ++c) {
        float lo = INFINITY, hi = -INFINITY;
        for (int i = 0; i < n_active; ++i) {
            lo = fminf(lo, f[(size_t)c * ld + i]);
            hi = fmaxf(hi, f[(size_t)c * ld + i]);
        }
        float r = hi - lo;
        if (!isfinite(r) || r < 1e-12f) r = 1e-12f;
        rng[c] = r;
    }
    for (int j = 0; j < k; ++j) {
        float base[16], sc[16];
        float bn = 0.0f, sn = 0.0f;
        for (int c = 0; c < m; ++c) {
            float b = v0[(size_t)c * k + j];
            if (!isfinite(b)) b = 0.0f;
            base[c] = b;
            bn = fmaf(b, b, bn);
            float s = b * rng[c];
            if (!isfinite(s)) s = 0.0f;
            sc[c] = s;
            sn = fmaf(s, s, sn);
        }
        const int use_s = isfinite(sn) && sn > 1e-12f;
        const int use_b = isfinite(bn) && bn > 1e-12f;
        const float is = use_s ? 1.0f / sqrtf(sn) : 0.0f;
        const float ib = use_b ? 1.0f / sqrtf(bn) : 0.0f;
        for (int c = 0; c < m; ++c) {
            float o;
            if (use_s) o = sc[c] * is;
            else if (use_b) o = base[c] * ib;
            else o = (c == 0) ? 1.0f : 0.0f;
            v_out[(size_t)c * k + j] = o;
        }
    }
}

/* A full RVEA run (EvolutionaryAlgorithm::step, algorithm.cu:219-286, with the
 * built-in RVEA strategies) on quadratic-center objectives, with the
 * in-kernel draw generator: random mating over the live parents
 * (mating_selection.cu:39-62, `lb + curand() % active`), SBX + PM as for
 * NSGA-III, RVEA selection with t = generation + 1 in the penalty
 * (rvea_environment_selector.cu:236-239), the next population's dead rows
 * [active, n) holding NaN variables and 1e15 objectives and violations
 * (rvea_population_update.cu:137-205), then the direction adaptation when
 * generation + 1 is a multiple of max(1, round(frequency x gens)) over the
 * live rows of the next population (reference_directions.cu:132-162) and a
 * new gamma.  v0: the initial (m, k) directions.  Outputs x (n, d), f (m, n),
 * cv (n), v (m, k) of the last generation; returns the live count. */
ORC_API int orc_rvea_run(int n, int m, int d, const float* v0, int k, int gens, uint64_t seed,
                         const float* bounds, const float* centers, float alpha, float frequency,
                         float sbx_eta0, float sbx_eta1, float sbx_prob, float sbx_copy,
                         float pm_eta0, float pm_eta1, float pm_prob,
                         float* x, float* f, float* cv, float* v, int gen_count_limit)
{
    const int nm = 2 * n;
    float* gamma = (float*)malloc(sizeof(float) * (size_t)k);
    float* xm = (float*)malloc(sizeof(float) * (size_t)nm * d);
    float* fm = (float*)malloc(sizeof(float) * (size_t)m * nm);
    float* cvm = (float*)malloc(sizeof(float) * (size_t)nm);
    int32_t* win = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    const int pairs = n / 2;
    const size_t pd = (size_t)(pairs > 0 ? pairs : 1) * d;
    float* ubeta = (float*)malloc(sizeof(float) * pd);
    float* up = (float*)malloc(sizeof(float) * pd);
    int32_t* sg = (int32_t*)malloc(sizeof(int32_t) * pd);
    float* uc = (float*)malloc(sizeof(float) * pd);
    float* ug = (float*)malloc(sizeof(float) * (size_t)n * d);
    float* ud = (float*)malloc(sizeof(float) * (size_t)n * d);
    int32_t* sel = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    float* off = xm + (size_t)n * d;

    memcpy(v, v0, sizeof(float) * (size_t)m * k);
    orc_rvea_gamma(v, m, k, gamma);
    orc_draw_init(seed, n, d, bounds, x);
    orc_eval_quadratic(x, n, d, centers, m, f, n, cv);
    int active = n;
    int interval = (int)lround((double)frequency * gens);
    if (interval < 1) interval = 1;
    const int g_end = (gen_count_limit > 0 && gen_count_limit < gens) ? gen_count_limit : gens;
    for (int g = 0; g < g_end; ++g) {
        const float eta_c = orc_annealed(sbx_eta0, sbx_eta1, g, gens);
        const float eta_m = orc_annealed(pm_eta0, pm_eta1, g, gens);
        orc_draw_mating(seed, (uint32_t)g, n, active, win, NULL);
        orc_draw_sbx(seed, (uint32_t)g, pairs, d, ubeta, up, sg, uc);
        orc_sbx(x, win, bounds, ubeta, up, sg, uc, off, eta_c, sbx_prob, sbx_copy, n, d);
        orc_draw_pm(seed, (uint32_t)g, n, d, ug, ud);
        orc_pm(ug, ud, off, bounds, eta_m, pm_prob, n, d);
        for (int c = 0; c < m; ++c) memcpy(fm + (size_t)c * nm, f + (size_t)c * n, sizeof(float) * n);
        memcpy(cvm, cv, sizeof(float) * n);
        orc_eval_quadratic(off, n, d, centers, m, fm + n, nm, cvm + n);
        memcpy(xm, x, sizeof(float) * (size_t)n * d);
        const float pen = orc_rvea_penalty(m, g + 1, gens, alpha);
        const int na = orc_rvea_select(fm, cvm, n, m, v, gamma, k, active, pen, sel, NULL, NULL, NULL);
        for (int i = 0; i < n; ++i) {
            if (i < na) {
                const int s = sel[i];
                memcpy(x + (size_t)i * d, xm + (size_t)s * d, sizeof(float) * d);
                for (int c = 0; c < m; ++c) f[(size_t)c * n + i] = fm[(size_t)c * nm + s];
                cv[i] = cvm[s];
            } else {
                for (int j = 0; j < d; ++j) x[(size_t)i * d + j] = NAN;
                for (int c = 0; c < m; ++c) f[(size_t)c * n + i] = 1e15f;
                cv[i] = 1e15f;
            }
        }
        active = na;
        if ((g + 1) % interval == 0) {
            orc_rvea_adapt(v0, f, active, n, m, k, v);
            orc_rvea_gamma(v, m, k, gamma);
        }
    }
    free(sel); free(ud); free(ug); free(uc); free(sg); free(up); free(ubeta); free(win);
    free(cvm); free(fm); free(xm); free(gamma);
    return active;
}

ORC_API int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

ORC_API void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
