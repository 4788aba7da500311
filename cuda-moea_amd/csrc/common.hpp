// Shared device/host helpers of the MI355X MOEA engine.
//
// Arithmetic contract: every translation unit is compiled with
// -ffp-contract=off, so each fused multiply-add is an explicit fmaf()/fma()
// placed where the reference's nvcc build contracts one (nvcc's default
// -fmad=true, CMakeLists.txt:115-123).  fp32 division and sqrt are IEEE
// correctly rounded (-fhip-fp32-correctly-rounded-divide-sqrt).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <utility>

namespace moea {

// ----------------------------------------------------------------------------
// Errors: C-ABI entry points return a status and record a message.
// ----------------------------------------------------------------------------
enum Status : int {
    kOk = 0,
    kInvalidArgument = -1,
    kHipError = -2,
    kWorkspaceTooSmall = -3,
    kUnsupported = -4,
    kDeviceFault = -5,
};

void set_last_error(const std::string& msg);
int fail(int status, const std::string& msg);
int hip_fail(hipError_t err, const char* what);

#define MOEA_HIP(call)                                                        \
    do {                                                                      \
        hipError_t _e = (call);                                               \
        if (_e != hipSuccess) return ::moea::hip_fail(_e, #call);             \
    } while (0)

#define MOEA_CHECK(status_expr)                                               \
    do {                                                                      \
        int _s = (status_expr);                                               \
        if (_s != ::moea::kOk) return _s;                                     \
    } while (0)

// Launch kern<M> for the runtime objective count m (grid/block must be
// single tokens, e.g. dim3 variables).
#define MOEA_LAUNCH_M(m, kern, grid, block, shmem, stream, ...)                                    \
    switch (m) {                                                                                \
        case 1: hipLaunchKernelGGL(kern<1>, grid, block, shmem, stream, __VA_ARGS__); break;     \
        case 2: hipLaunchKernelGGL(kern<2>, grid, block, shmem, stream, __VA_ARGS__); break;     \
        case 3: hipLaunchKernelGGL(kern<3>, grid, block, shmem, stream, __VA_ARGS__); break;     \
        case 4: hipLaunchKernelGGL(kern<4>, grid, block, shmem, stream, __VA_ARGS__); break;     \
        case 5: hipLaunchKernelGGL(kern<5>, grid, block, shmem, stream, __VA_ARGS__); break;     \
        case 6: hipLaunchKernelGGL(kern<6>, grid, block, shmem, stream, __VA_ARGS__); break;     \
        case 7: hipLaunchKernelGGL(kern<7>, grid, block, shmem, stream, __VA_ARGS__); break;     \
        case 8: hipLaunchKernelGGL(kern<8>, grid, block, shmem, stream, __VA_ARGS__); break;     \
        case 9: hipLaunchKernelGGL(kern<9>, grid, block, shmem, stream, __VA_ARGS__); break;     \
        case 10: hipLaunchKernelGGL(kern<10>, grid, block, shmem, stream, __VA_ARGS__); break;   \
        default: return ::moea::fail(::moea::kUnsupported, "unsupported objective count");      \
    }

constexpr int kMaxM = 10;          // objectives supported (reference MAX_M = 8)
constexpr int kWave = 64;          // CDNA wavefront width

inline int ceil_div(long long a, long long b) { return static_cast<int>((a + b - 1) / b); }
inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// ----------------------------------------------------------------------------
// Counter-based draws.  Philox4x32-10 (Salmon et al., SC'11) keyed by the run
// seed; the counter is (index lo, index hi, generation, stream).
// ----------------------------------------------------------------------------
enum DrawStream : uint32_t { kStInit = 1, kStMate = 2, kStSbx = 3, kStPm = 4 };

struct U4 { uint32_t x, y, z, w; };

__host__ __device__ inline U4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                            uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c0;
        const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c2;
        const uint32_t n0 = static_cast<uint32_t>(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n1 = static_cast<uint32_t>(p1);
        const uint32_t n2 = static_cast<uint32_t>(p0 >> 32) ^ c3 ^ k1;
        const uint32_t n3 = static_cast<uint32_t>(p0);
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return {c0, c1, c2, c3};
}

__host__ __device__ inline U4 draw4(uint64_t seed, uint64_t index, uint32_t gen, uint32_t stream)
{
    return philox4x32_10(static_cast<uint32_t>(index), static_cast<uint32_t>(index >> 32), gen,
                         stream, static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32));
}

// 24-bit uniform in [0, 1); never 1.0.
__host__ __device__ inline float u01(uint32_t u) { return static_cast<float>(u >> 8) * 5.9604644775390625e-8f; }

// In-kernel draws of one (pair, variable) element e = pair * d + j.  One
// Philox block of stream kStSbx carries the SBX draws and both children's
// mutation gates: u_beta, u_prob, u_copy from the top 24 bits of words x, y, z
// (u01), the gate of child a from the top 24 bits of w, the gate of child b
// from the low bytes of x, y, z, and the SBX direction from bit 7 of w.  The
// mutation deltas come from a second block (stream kStPm: child a word x,
// child b word y), generated only when a gate opens; with the usual rate 1/D
// that is a small fraction of the elements, so most elements cost one Philox
// block instead of two.  The tail row of an odd population takes child a's
// draws at index pairs * d + j.
__host__ __device__ inline float gate_b(const U4& r)
{
    return static_cast<float>(((r.x & 0xFFu) << 16) | ((r.y & 0xFFu) << 8) | (r.z & 0xFFu)) *
           5.9604644775390625e-8f;
}
__host__ __device__ inline int sbx_direction(const U4& r) { return static_cast<int>((r.w >> 7) & 1u); }

// The reference's niching draw: SplitMix64 finalizer of (counter ^ seed), low
// 32 bits (NS/core/cuda/cuda_random.cu:1606-1614).
__host__ __device__ inline uint32_t splitmix32(uint64_t counter, uint64_t seed)
{
    uint64_t z = counter ^ seed;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z = z ^ (z >> 31);
    return static_cast<uint32_t>(z);
}

// Niching counters: stream a of generation g starts at g<<33, stream b at
// (g<<33) | (1<<32).
__host__ __device__ inline uint64_t niche_base(uint32_t gen, int which)
{
    return (static_cast<uint64_t>(gen) << 33) | (static_cast<uint64_t>(which & 1) << 32);
}

// ----------------------------------------------------------------------------
// Division by an invariant divisor d >= 1 for dividends below 2^31: one
// high multiply and a shift (Granlund and Montgomery, "Division by invariant
// integers using multiplication", PLDI 1994): with l = ceil(log2 d) and
// p = 31 + l, m = ceil(2^p / d) < 2^32 and floor(x / d) = (x * m) >> p.
// ----------------------------------------------------------------------------
struct FastDiv {
    uint32_t d, mul, shr;
};

__host__ __device__ inline FastDiv fast_div_make(uint32_t d)
{
    FastDiv f{d, 0u, 0u};
    if (d <= 1) return f;
    uint32_t l = 0;
    while ((1ull << l) < d) ++l;
    const uint32_t p = 31 + l;
    f.mul = static_cast<uint32_t>(((1ull << p) + d - 1) / d);
    f.shr = p - 32;
    return f;
}

__host__ __device__ inline uint32_t fast_div(uint32_t x, const FastDiv& f)
{
    if (f.d <= 1) return x;
#if defined(__HIP_DEVICE_COMPILE__)
    return __umulhi(x, f.mul) >> f.shr;
#else
    return static_cast<uint32_t>((static_cast<uint64_t>(x) * f.mul) >> 32) >> f.shr;
#endif
}

// ----------------------------------------------------------------------------
// Deterministic pow / acos: evaluated in double with a fixed operation
// sequence and rounded once, identical on host and device.  They stand in for
// CUDA's powf / acosf, whose last-bit behaviour cannot be reproduced.
// ----------------------------------------------------------------------------
__host__ __device__ inline double log2_d(double x)
{
    int e;
    double m = frexp(x, &e);
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
    return static_cast<double>(e) + ln_m * 1.44269504088896338700e+00;
}

__host__ __device__ inline double exp2_d(double t)
{
    if (t > 1100.0) return static_cast<double>(INFINITY);
    if (t < -1100.0) return 0.0;
    const double k = nearbyint(t);
    const double r = t - k;
    const double y = fma(r, 6.93147180369123816490e-01, r * 1.90821492927058770002e-10);
    double p = 1.0 / 87178291200.0;
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
    return ldexp(p, static_cast<int>(k));
}

__host__ __device__ inline float powf_det(float x, float y)
{
    if (x != x || y != y) return NAN;
    if (x < 0.0f) return NAN;
    if (x == 0.0f) return y > 0.0f ? 0.0f : INFINITY;
    if (x == 1.0f || y == 0.0f) return 1.0f;
    if (isinf(x)) return y > 0.0f ? INFINITY : 0.0f;
    return static_cast<float>(exp2_d(static_cast<double>(y) * log2_d(static_cast<double>(x))));
}

// x^y for 0 <= x < 2^24 and 0 < y <= 1 (the SBX spread factor
// (2u)^(1/(1+eta)) or (2 - 2u)^(-1/(1+eta)), and the polynomial-mutation root
// vm^(1/(eta+1))) in float32 arithmetic only, one fixed operation sequence,
// within 1.05 ulp of the exact value over those domains:
//   x = 2^e m with m in [sqrt(1/2), sqrt(2)];
//   log2 m = (2 / ln 2) atanh(s), s = (m - 1) / (m + 1), by the odd series
//     atanh s = s + s^3/3 + ... + s^11/11 (Abramowitz & Stegun 4.6.22), with
//     s = (m - 1) r with r = 1 / (m + 1) one IEEE reciprocal, its low part
//     from the fused remainder times r, and the (2 / ln 2) product carried
//     to double-float precision;
//   z = y (e + log2 m) as an unevaluated sum hi + lo;
//   2^z = 2^k e^w, k = rint(hi), w = ((hi - k) + lo) ln 2, e^w by its Taylor
//     polynomial to w^8 / 8! (A&S 4.2.1), |w| <= 0.35.
// Identical on host and device (fmaf, IEEE division, no contraction).
__host__ __device__ inline float powf_root(float x, float y)
{
    if (!(x > 0.0f)) return x == 0.0f ? 0.0f : NAN;
    if (x == 1.0f) return 1.0f;
    if (x == INFINITY) return INFINITY;
    int e = 0;
    if (x < 1.17549435e-38f) { x = x * 16777216.0f; e = -24; }      // subnormal: scale by 2^24
    const uint32_t b = __builtin_bit_cast(uint32_t, x);
    e += static_cast<int>((b >> 23) & 0xffu) - 127;
    float m = __builtin_bit_cast(float, (b & 0x007fffffu) | 0x3f800000u);   // [1, 2)
    if (m > 1.41421356f) { m = m * 0.5f; e += 1; }
    const float num = m - 1.0f;                           // exact
    const float den = m + 1.0f;
    const float den_lo = m - (den - 1.0f);                // den + den_lo = m + 1 exactly
    const float rden = 1.0f / den;                        // one IEEE reciprocal for s and s_lo
    const float s = num * rden;
    float rem = fmaf(-s, den, num);                       // remainder, to a rounding of 2^-24 of itself
    rem = fmaf(-s, den_lo, rem);
    const float s_lo = rem * rden;                        // s + s_lo = (m - 1) / (m + 1) to ~2^-45
    const float s2 = s * s;
    float p = 1.0f / 11.0f;
    p = fmaf(p, s2, 1.0f / 9.0f);
    p = fmaf(p, s2, 1.0f / 7.0f);
    p = fmaf(p, s2, 1.0f / 5.0f);
    p = fmaf(p, s2, 1.0f / 3.0f);
    p = p * s2;                                           // atanh(s) / s - 1
    const float kC_hi = 2.88539004f, kC_lo = 3.851926e-08f;           // 2 / ln 2
    const float tail = fmaf(s, p, s_lo);
    const float lm = s * kC_hi;
    const float lm_err = fmaf(s, kC_hi, -lm);
    const float lm_lo = fmaf(tail, kC_hi, fmaf(s, kC_lo, lm_err));   // log2 m = lm + lm_lo
    const float ef = static_cast<float>(e);
    const float hi0 = y * ef;
    const float hi0_err = fmaf(y, ef, -hi0);
    const float t1 = y * lm;
    const float t1_err = fmaf(y, lm, -t1);
    const float hi = hi0 + t1;
    const float hi_err = (hi0 - hi) + t1;                 // |hi0| >= 2 |t1| or hi0 = 0
    const float lo = fmaf(y, lm_lo, hi0_err + t1_err) + hi_err;
    const float k = rintf(hi);
    const float r = (hi - k) + lo;
    const float kL_hi = 0.693147182f, kL_lo = -1.9046542e-09f;        // ln 2
    const float w = fmaf(r, kL_hi, r * kL_lo);
    float q = 1.0f / 40320.0f;
    q = fmaf(q, w, 1.0f / 5040.0f);
    q = fmaf(q, w, 1.0f / 720.0f);
    q = fmaf(q, w, 1.0f / 120.0f);
    q = fmaf(q, w, 1.0f / 24.0f);
    q = fmaf(q, w, 1.0f / 6.0f);
    q = fmaf(q, w, 0.5f);
    q = fmaf(q, w, 1.0f);
    q = fmaf(q, w, 1.0f);
    return ldexpf(q, static_cast<int>(k));
}

// cos(t) in double for |t| <= 2^20 with one fixed operation sequence:
// t = k (pi / 2) + r, k = rint(t 2 / pi), pi / 2 as a two-term sum, then the
// Taylor series of sin r or cos r (|r| <= pi / 4, to r^17 / r^18) by the
// quadrant of k.
__host__ __device__ inline double cos_d(double t)
{
    const double k = nearbyint(t * 6.36619772367581382433e-01);
    double r = fma(-k, 1.57079632679489655800e+00, t);
    r = fma(-k, 6.12323399573676603587e-17, r);
    const double r2 = r * r;
    const int q = static_cast<int>(k) & 3;
    double v;
    if (q & 1) {                 // +-sin r
        double p = 1.0 / 355687428096000.0;                    // 1/17!
        p = fma(p, -r2, 1.0 / 1307674368000.0);                // 1/15!
        p = fma(p, -r2, 1.0 / 6227020800.0);                   // 1/13!
        p = fma(p, -r2, 1.0 / 39916800.0);                     // 1/11!
        p = fma(p, -r2, 1.0 / 362880.0);                       // 1/9!
        p = fma(p, -r2, 1.0 / 5040.0);                         // 1/7!
        p = fma(p, -r2, 1.0 / 120.0);                          // 1/5!
        p = fma(p, -r2, 1.0 / 6.0);                            // 1/3!
        v = fma(-r * r2, p, r);
    } else {                     // +-cos r
        double p = 1.0 / 6402373705728000.0;                   // 1/18!
        p = fma(p, -r2, 1.0 / 20922789888000.0);               // 1/16!
        p = fma(p, -r2, 1.0 / 87178291200.0);                  // 1/14!
        p = fma(p, -r2, 1.0 / 479001600.0);                    // 1/12!
        p = fma(p, -r2, 1.0 / 3628800.0);                      // 1/10!
        p = fma(p, -r2, 1.0 / 40320.0);                        // 1/8!
        p = fma(p, -r2, 1.0 / 720.0);                          // 1/6!
        p = fma(p, -r2, 1.0 / 24.0);                           // 1/4!
        p = fma(p, -r2, 0.5);                                  // 1/2!
        v = fma(-r2, p, 1.0);
    }
    // quadrants: 0 cos r, 1 -sin r, 2 -cos r, 3 sin r
    return (q == 1 || q == 2) ? -v : v;
}

// cosf of the DTLZ1 / DTLZ3 distance terms, rounded once from double.
__host__ __device__ inline float cosf_det(float x) { return static_cast<float>(cos_d(static_cast<double>(x))); }

__host__ __device__ inline double atan01_d(double t)
{
    const int i = static_cast<int>(nearbyint(t * 4.0));
    double c0;
    switch (i) {
        case 0: c0 = 0.0; break;
        case 1: c0 = 2.44978663126864143473e-01; break;
        case 2: c0 = 4.63647609000806093515e-01; break;
        case 3: c0 = 6.43501108793284370968e-01; break;
        default: c0 = 7.85398163397448278999e-01; break;
    }
    const double c = 0.25 * static_cast<double>(i);
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
    return c0 + r * p;
}

// acos on [0, 1] (callers clamp the cosine first, as the reference does).
__host__ __device__ inline float acosf_det(float xf)
{
    if (xf != xf) return NAN;
    double x = static_cast<double>(xf);
    if (x >= 1.0) return 0.0f;
    if (x <= 0.0) x = 0.0;
    const double t = sqrt((1.0 - x) / (1.0 + x));
    return static_cast<float>(2.0 * atan01_d(t));
}

// ----------------------------------------------------------------------------
// M x M hyperplane solve in FP64.  One-sided Jacobi SVD of A (row i = extreme
// point i), then a = pinv(A) * 1 with singular values <= 1e-5 dropped: the
// reference's cusolverDnSgesvd + cublasSgemv pipeline
// (selection/nsga3/kernels/normalization.cu:1302-1394) evaluated in double.
// Returns 1 when the reference takes its row-max fallback: non-finite input or
// singular values, sigma_min < 1e-10, sigma_max / sigma_min > 1e6
// (normalization.cu:739-771) or no convergence (normalization.cu:1340-1348).
// The operation sequence matches the oracle's orc_hyperplane_fp64.
// ----------------------------------------------------------------------------
__host__ __device__ inline int hyperplane_fp64(const double* a_in, int m, double* a_out)
{
    double A[kMaxM * kMaxM], V[kMaxM * kMaxM];
    if (m > kMaxM) return 1;
    for (int i = 0; i < m * m; ++i) {
        A[i] = a_in[i];
        if (!isfinite(A[i])) return 1;
    }
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) V[i * m + j] = (i == j) ? 1.0 : 0.0;
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
    double sig[kMaxM];
    double smax = 0.0, smin = static_cast<double>(INFINITY);
    for (int j = 0; j < m; ++j) {
        double ss = 0.0;
        for (int i = 0; i < m; ++i) ss = ss + A[i * m + j] * A[i * m + j];
        sig[j] = sqrt(ss);
        if (!isfinite(sig[j])) return 1;
        if (sig[j] > smax) smax = sig[j];
        if (sig[j] < smin) smin = sig[j];
    }
    if (smin < 1e-10) return 1;
    if (smax / smin > 1e6) return 1;
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

// ----------------------------------------------------------------------------
// Device-scope atomics and a bounded grid barrier for cooperative launches.
// ----------------------------------------------------------------------------
__device__ inline uint32_t ld_acquire_u32(const uint32_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint32_t ld_relaxed_u32(const uint32_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline int32_t ld_relaxed_i32(const int32_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint64_t ld_relaxed_u64(const uint64_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_relaxed_u64(uint64_t* p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Barrier of co-resident blocks for cooperative launches (at most
// kBarMaxBlocks blocks) whose only data shared inside the launch is accessed
// with device-scope atomics (coherent across the XCDs' L2 caches without
// cache maintenance).  Each block publishes its arrival with a device-scope
// store to its own flag word; every thread of block 0 polls a share of the
// flags, then block 0 stores the generation word that the other blocks poll.
// Ordering comes from waiting for each block's outstanding memory operations
// (s_waitcnt 0) before its flag or generation store, and from loads issued
// only after the polled value arrived; no L2 write-back or invalidate is
// issued, so data the blocks read from the L2 stays cached across rounds.
// Words of `state`, all zero at launch:
constexpr int kBarMaxBlocks = 1024;
constexpr int kBarGen = kBarMaxBlocks;        // generation
constexpr int kBarError = kBarGen + 16;       // set when a spin gives up
constexpr int kBarWords = kBarError + 16;     // first word free for the caller
// The spin is bounded: after `spin_limit` polls a thread records an error and
// the barrier returns false, so a broken launch ends instead of hanging.
constexpr uint32_t kSpinLimit = 1u << 24;

// Every memory operation this wave issued has completed (device-scope atomic
// stores are then visible to every block); also a compiler barrier.
__device__ inline void drain_memory()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ inline void st_relaxed_u32(uint32_t* p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Barrier of the `nblocks` co-resident blocks that share `state`; `rank` is
// this block's index among them (0 .. nblocks-1).  With snap_src >= 0, block
// 0 copies state[snap_src] to state[snap_dst] after every block has arrived
// and before any leaves: a value accumulated before the barrier, read by
// every block afterwards from snap_dst, is then the same for all of them even
// if fast blocks start adding to snap_src for the next round.
__device__ inline bool coherent_barrier(uint32_t* state, uint32_t rank, uint32_t nblocks, uint32_t& phase,
                                        uint32_t spin_limit, int snap_src = -1, int snap_dst = -1)
{
    __shared__ int s_ok;
    const uint32_t target = phase + 1;
    if (threadIdx.x == 0) s_ok = 1;
    drain_memory();
    __syncthreads();
    if (threadIdx.x == 0) st_relaxed_u32(&state[rank], target);
    if (rank == 0) {
        for (uint32_t b = threadIdx.x; b < nblocks; b += blockDim.x) {
            uint32_t spins = 0;
            while (ld_relaxed_u32(&state[b]) != target) {
                if (++spins > spin_limit) {
                    st_relaxed_u32(&state[kBarError], 1u);
                    s_ok = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            if (snap_src >= 0) st_relaxed_u32(&state[snap_dst], ld_relaxed_u32(&state[snap_src]));
            drain_memory();
            st_relaxed_u32(&state[kBarGen], target);
        }
    } else if (threadIdx.x == 0) {
        uint32_t spins = 0;
        while (ld_relaxed_u32(&state[kBarGen]) != target) {
            if (++spins > spin_limit) {
                st_relaxed_u32(&state[kBarError], 1u);
                s_ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    if (threadIdx.x == 0 && ld_relaxed_u32(&state[kBarError]) != 0) s_ok = 0;
    __syncthreads();
    phase = target;
    return s_ok != 0;
}

// ----------------------------------------------------------------------------
// Sort keys of the non-dominated sort
// ----------------------------------------------------------------------------
// Monotone map float -> u32 (numeric order for non-NaN values; -0 maps like +0).
__device__ __forceinline__ uint32_t order_key(float x)
{
    const uint32_t u = __float_as_uint(x == 0.0f ? 0.0f : x);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// A row's 34-bit key from its first objective, whether any objective is NaN,
// and its violation: segment 0 (feasible, no NaN) by f0, segment 1 (feasible
// with a NaN objective) all equal, segment 2 (infeasible) by cv, NaN last.
__device__ __forceinline__ uint64_t nds_key_from(float f0, bool any_nan, float c)
{
    if (c == 0.0f) return any_nan ? (1ull << 32) : order_key(f0);
    return (2ull << 32) | (isnan(c) ? 0xFFFFFFFFu : order_key(c));
}

// ----------------------------------------------------------------------------
// Wave helpers
// ----------------------------------------------------------------------------
__device__ inline uint64_t ballot(bool pred) { return __builtin_amdgcn_ballot_w64(pred); }

// v_writelane_b32 with a compile-time lane: lane J of the result takes the
// wave-uniform `value` (an inline-constant lane keeps the instruction within
// the constant-bus limit).
template <int J>
__device__ inline uint32_t writelane(uint32_t old, uint32_t value)
{
    uint32_t r = old;
    const uint32_t v = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(value)));
    asm volatile("v_writelane_b32 %0, %1, %2" : "+v"(r) : "s"(v), "n"(J));
    return r;
}

// Compile-time loop: f(std::integral_constant<int, J>) for J in Js.
template <typename F, int... Js>
__device__ inline void static_for(F&& f, std::integer_sequence<int, Js...>)
{
    (f(std::integral_constant<int, Js>{}), ...);
}

__device__ inline int lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// Marks a value the whole wave agrees on as uniform (kept in an SGPR), so
// loads indexed by it become scalar loads.
__device__ inline int wave_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Lane mask of !(s > v) for a wave-uniform s: one v_cmp_ngt_f32 writing an
// SGPR pair (true for unordered operands).
__device__ inline uint64_t mask_not_gt(float s, float v)
{
    const float su = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(s)));
    uint64_t m;
    asm("v_cmp_ngt_f32_e64 %0, %1, %2" : "=s"(m) : "s"(su), "v"(v));
    return m;
}

// Per lane: acc * 2 + (bit `lane` of mask).  One v_addc_co_u32 with the
// wave-uniform lane mask as carry-in, i.e. a bit-serial shift of a ballot
// into a per-lane word.
__device__ inline uint32_t shift_in(uint32_t acc, uint64_t mask)
{
    const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(mask)));
    const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(mask >> 32)));
    const uint64_t m = (static_cast<uint64_t>(hi) << 32) | lo;
    uint32_t r;
    uint64_t carry_out;
    asm("v_addc_co_u32 %0, %1, %2, %2, %3" : "=v"(r), "=s"(carry_out) : "v"(acc), "s"(m));
    return r;
}

}  // namespace moea
