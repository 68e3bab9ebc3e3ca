// plis_numerics.h -- gfx950 device numerics for the RDME tau-leap step.
//
// Counter-based Philox4x32-10 (Salmon et al., SC'11), uniform doubles from
// 32-bit words, exp / log / sqrt / log k! written from published methods with
// every fused multiply-add explicit (the library is built with
// -ffp-contract=off), and a Poisson sampler: sequential-search inversion for
// mean < 10 (Devroye 1986, ch. X.3) and PTRS transformed rejection for
// mean >= 10 (Hoermann 1993).  The host CPU restatement applies the same
// operations in the same order, so device and host draws agree bit for bit.
//
// These replace the curand calls of the reference (curand_init
// PLISCuda/src/init.cu:3-13, curand_poisson leap.cu:27,51, curand_uniform
// tau.cu:193, leap.cu:89,182, ssa.cu:64,90).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace plis {

// ---- constants (60-digit decimal evaluation, rounded to double) ----
constexpr double kLn2Hi = 0x1.62e42fefa3800p-1;  // ln2, 42 significant bits
constexpr double kLn2Lo = 0x1.ef35793c76730p-45;
constexpr double kInvLn2 = 0x1.71547652b82fep+0;
constexpr double kSqrt2 = 0x1.6a09e667f3bcdp+0;
constexpr double kHalfLog2Pi = 0x1.d67f1c864beb5p-1;
constexpr double kC12 = 0x1.5555555555555p-4;    // 1/12
constexpr double kC360 = 0x1.6c16c16c16c17p-9;   // 1/360
constexpr double kC1260 = 0x1.a01a01a01a01ap-11; // 1/1260
constexpr double kPoissonInvMax = 10.0;

// Philox stream purposes (counter word 3, bits 7..9).
enum Purpose : uint32_t { kTau2 = 1, kCrit = 2, kReact = 3, kHop = 4, kExtra = 5 };
constexpr uint32_t kSelfId = 0xFFFFFFFFu;
constexpr uint32_t kMaxAttempt = 63u;

__device__ __forceinline__ uint32_t ctr_word3(uint32_t real, uint32_t attempt, uint32_t purpose,
                                              uint32_t block) {
    return (real << 16) | (attempt << 10) | (purpose << 7) | block;
}

struct U4 { uint32_t x, y, z, w; };

__device__ __forceinline__ U4 philox10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                       uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;   // one v_mad_u64_u32 each
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        c0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        c2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return U4{c0, c1, c2, c3};
}

__device__ __forceinline__ uint32_t pick(const U4& o, uint32_t i) {
    return i == 0 ? o.x : (i == 1 ? o.y : (i == 2 ? o.z : o.w));
}

// Uniforms of the leap draws.  Draw w of a counter (hop w = 2s + direction
// of edge {c0 < c1}, direction 0 for the hop c0 -> c1; reaction w = j of
// subvolume c0 with c1 = kSelfId) takes the 32-bit word
//     (coarse byte << 24) | (fine bits & 0xFFFFFF):
// the coarse byte is byte w & 15 of the Philox block numbered w >> 4 (16
// draws per block: bytes 0..3 of word 0, then of words 1, 2, 3), the fine
// bits word w & 3 of the block numbered kFineBlock + (w >> 2).  A draw
// whose coarse byte already proves it 0 (coarse_zero*) never needs its fine
// block, so the common no-event draw costs a sixteenth of a Philox call.
constexpr uint32_t kFineBlock = 64u;

__device__ __forceinline__ uint32_t coarse_byte(const U4& o, uint32_t w) {
    return (pick(o, (w >> 2) & 3u) >> (8u * (w & 3u))) & 0xFFu;
}

__device__ __forceinline__ uint32_t draw_word(uint32_t hb, const U4& fine, uint32_t w) {
    return (hb << 24) | (pick(fine, w & 3u) & 0xFFFFFFu);
}

// Every word with top byte hb is at most (hb + 1) 2^24 - 1, so its uniform
// is below (hb + 1) 2^-8 (exact in double): when that bound passes the
// zero shortcut of poisson() (double mean), the draw is 0 whatever its
// fine bits.
__device__ __forceinline__ bool coarse_zero(double lam, uint32_t hb) {
    return !(lam > 0.0) || (double)(hb + 1u) * 0x1p-8 <= (1.0 - lam) - 0x1p-50;
}

// The same decision for a mean lf formed in single precision (within
// 2^-21 relative of the double mean lam, as hop_mean_f forms it), as one
// compare per draw: the draw is 0 when (float)hb <= coarse_bound(lf), with
// g = (255 - 2^-12) - lf (256 + 2^-10) in float for lf > 0, 256 otherwise.
// Then 255 - hb > lf 256 (1 + 2^-18)(1 - 2^-24) (g carries at most 2^-17
// of rounding, less than the 2^-12 taken off), so (hb + 1) 2^-8 < 1 - lam
// <= e^-lam: every word with that top byte gives 0.  A mean above 1 is
// never settled here.
__device__ __forceinline__ float coarse_bound(float lf) {
    return lf > 0.0f ? 0x1.fdffe0p+7f - lf * 0x1.00004p+8f : 256.0f;
}

__device__ __forceinline__ double uniform01(uint32_t w) { return ((double)w + 0.5) * 0x1p-32; }

// exp(x): Cody-Waite reduction r = x - n ln2 (two fma steps), degree-13
// Taylor polynomial of e^r in Horner form, scale by 2^n built from bits.
__device__ __forceinline__ double dexp(double x) {
    if (x < -708.0) return 0.0;
    if (x > 709.0) return __builtin_huge_val();
    const double n = __builtin_rint(x * kInvLn2);
    double r = __builtin_fma(-n, kLn2Hi, x);
    r = __builtin_fma(-n, kLn2Lo, r);
    double p = 0x1.6124613a86d09p-33;            // 1/13!
    p = __builtin_fma(p, r, 0x1.1eed8eff8d898p-29);  // 1/12!
    p = __builtin_fma(p, r, 0x1.ae64567f544e4p-26);  // 1/11!
    p = __builtin_fma(p, r, 0x1.27e4fb7789f5cp-22);  // 1/10!
    p = __builtin_fma(p, r, 0x1.71de3a556c734p-19);  // 1/9!
    p = __builtin_fma(p, r, 0x1.a01a01a01a01ap-16);  // 1/8!
    p = __builtin_fma(p, r, 0x1.a01a01a01a01ap-13);  // 1/7!
    p = __builtin_fma(p, r, 0x1.6c16c16c16c17p-10);  // 1/6!
    p = __builtin_fma(p, r, 0x1.1111111111111p-7);   // 1/5!
    p = __builtin_fma(p, r, 0x1.5555555555555p-5);   // 1/4!
    p = __builtin_fma(p, r, 0x1.5555555555555p-3);   // 1/3!
    p = __builtin_fma(p, r, 0x1.0000000000000p-1);   // 1/2!
    p = __builtin_fma(p, r, 1.0);
    p = __builtin_fma(p, r, 1.0);
    const long long ni = (long long)n;
    return p * __longlong_as_double((ni + 1023) << 52);
}

// log(x) for normal x > 0: x = m 2^e with m in (sqrt2/2, sqrt2],
// log m = 2 atanh(s), s = (m-1)/(m+1), odd series through s^21.
__device__ __attribute__((noinline)) double dlog(double x) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    int e = (int)((b >> 52) & 0x7FF) - 1023;
    double m = __longlong_as_double((long long)((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull));
    if (m > kSqrt2) { m = m * 0.5; e += 1; }
    const double s = (m - 1.0) / (m + 1.0);
    const double z = s * s;
    double P = 0x1.8618618618618p-5;                 // 1/21
    P = __builtin_fma(P, z, 0x1.af286bca1af28p-5);   // 1/19
    P = __builtin_fma(P, z, 0x1.e1e1e1e1e1e1ep-5);   // 1/17
    P = __builtin_fma(P, z, 0x1.1111111111111p-4);   // 1/15
    P = __builtin_fma(P, z, 0x1.3b13b13b13b14p-4);   // 1/13
    P = __builtin_fma(P, z, 0x1.745d1745d1746p-4);   // 1/11
    P = __builtin_fma(P, z, 0x1.c71c71c71c71cp-4);   // 1/9
    P = __builtin_fma(P, z, 0x1.2492492492492p-3);   // 1/7
    P = __builtin_fma(P, z, 0x1.999999999999ap-3);   // 1/5
    P = __builtin_fma(P, z, 0x1.5555555555555p-2);   // 1/3
    P = __builtin_fma(P, z, 1.0);
    const double t = (2.0 * s) * P;
    const double ed = (double)e;
    return __builtin_fma(ed, kLn2Hi, __builtin_fma(ed, kLn2Lo, t));
}

// sqrt by six Newton-Heron steps from an exponent-halving first guess.
__device__ __forceinline__ double dsqrt(double x) {
    if (!(x > 0.0)) return 0.0;
    double y = __longlong_as_double(
        (long long)(((unsigned long long)__double_as_longlong(x) >> 1) + (0x3FF0000000000000ull >> 1)));
#pragma unroll
    for (int i = 0; i < 6; ++i) y = 0.5 * (y + x / y);
    return y;
}

// log k!: exact table below 10, Stirling series through 1/z^5 above.
__device__ __forceinline__ double dlogfact(long long k) {
    if (k < 10) {
        switch ((int)k) {
            case 2: return 0x1.62e42fefa39efp-1;
            case 3: return 0x1.cab0bfa2a2002p+0;
            case 4: return 0x1.96ca77c922cf9p+1;
            case 5: return 0x1.326643c4479c9p+2;
            case 6: return 0x1.a51273acf01cap+2;
            case 7: return 0x1.10ce1f32dcc30p+3;
            case 8: return 0x1.5358e82fcb70dp+3;
            case 9: return 0x1.99a8921a7f7cfp+3;
            default: return 0.0;
        }
    }
    const double z = (double)k + 1.0;
    const double z2 = z * z;
    return (((double)k + 0.5) * dlog(z) - z) + kHalfLog2Pi + (kC12 - (kC360 - kC1260 / z2) / z2) / z;
}

// Draw identity: main counter (c0..c3 with purpose/block), key, word in block.
struct DrawId {
    uint32_t c0, c1, c2, c3, k0, k1, word;
};

// q-th uniform of the draw's PTRS stream: purpose replaced by kExtra, key0
// perturbed by an odd multiplier of (4*call + word + 1).
__device__ __forceinline__ double extra_uniform(const DrawId& id, uint32_t q) {
    const uint32_t c3 = (id.c3 & ~(7u << 7)) | (kExtra << 7);
    const uint32_t k0 = id.k0 ^ (0x85EBCA6Bu * (4u * (q >> 2) + id.word + 1u));
    const U4 o = philox10(id.c0, id.c1, id.c2, c3, k0, id.k1);
    return uniform01(pick(o, q & 3u));
}

__device__ __attribute__((noinline)) long long poisson_ptrs(double lam, DrawId id) {
    const double slam = dsqrt(lam);
    const double loglam = dlog(lam);
    const double b = 0.931 + 2.53 * slam;
    const double a = -0.059 + 0.02483 * b;
    const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
    const double vr = 0.9277 - 3.6224 / (b - 2.0);
    const double loginvalpha = dlog(invalpha);
    const uint32_t c3 = (id.c3 & ~(7u << 7)) | (kExtra << 7);
    // iteration it uses uniforms 2it, 2it+1 of the draw's stream; one Philox
    // call (4 words) covers two iterations
#pragma unroll 1
    for (uint32_t call = 0; call < 50000u; ++call) {
        const uint32_t k0 = id.k0 ^ (0x85EBCA6Bu * (4u * call + id.word + 1u));
        const U4 o = philox10(id.c0, id.c1, id.c2, c3, k0, id.k1);
#pragma unroll 1
        for (uint32_t h = 0; h < 2u; ++h) {
            const double U = uniform01(h ? o.z : o.x) - 0.5;
            const double V = uniform01(h ? o.w : o.y);
            const double us = 0.5 - __builtin_fabs(U);
            const double kd = __builtin_floor((2.0 * a / us + b) * U + lam + 0.43);
            if (us >= 0.07 && V <= vr) return (long long)kd;
            if (kd < 0.0 || (us < 0.013 && V > us)) continue;
            const long long k = (long long)kd;
            const double lhs = (dlog(V) + loginvalpha) - dlog(a / (us * us) + b);
            const double rhs = (-lam + kd * loglam) - dlogfact(k);
            if (lhs <= rhs) return k;
        }
    }
    return (long long)__builtin_floor(lam);
}

// Slow path of the sampler: full sequential-search inversion (mean < 10) or
// PTRS (mean >= 10).  Kept out of line so the common no-event case below
// stays a handful of instructions at every call site.
__device__ __forceinline__ long long inversion_search(double lam, uint32_t u0) {
    const double u = uniform01(u0);
    double p = dexp(-lam);
    double F = p;
    long long k = 0;
#pragma unroll 1
    while (u > F && k < 1000) {
        k += 1;
        p = (p * lam) / (double)k;
        F = F + p;
    }
    return k;
}

__device__ __attribute__((noinline)) long long poisson_slow(double lam, uint32_t u0, DrawId id) {
    if (lam < kPoissonInvMax) return inversion_search(lam, u0);
    return poisson_ptrs(lam, id);
}

// ---- single-precision squeeze of the inversion search ----
// F_k is the double CDF that poisson_slow accumulates, G_k the float one
// below, evaluated at the float-rounded mean.  For every float mean in
// (0, 10) and k <= kSqueezeK, |G_k - F_k| stays below 2^-20 (checked over
// all such floats on the host with the same operations); rounding the mean
// moves F_k by at most 2^-21 (|dF_k/dlam| <= 1), rounding u by at most
// 2^-25, and forming G -+ delta by 2^-25 more.  With delta = 2^-15 >> the
// sum, u_f <= G_k - delta proves u <= F_k and u_f > G_k + delta proves
// u > F_k: the search result is found without double arithmetic, and is
// the one the double search returns.  u within delta of some G_k (odds
// about 6e-5 per comparison) or k > kSqueezeK leaves it to the double search.
constexpr float kSqueeze = 0x1p-15f;
constexpr int kSqueezeK = 15;
__constant__ float kRecipF[kSqueezeK + 1] = {
    0.0f,           0x1p+0f,        0x1p-1f,        0x1.555556p-2f, 0x1p-2f,        0x1.99999ap-3f,
    0x1.555556p-3f, 0x1.24924ap-3f, 0x1p-3f,        0x1.c71c72p-4f, 0x1.99999ap-4f, 0x1.745d18p-4f,
    0x1.555556p-4f, 0x1.3b13b2p-4f, 0x1.24924ap-4f, 0x1.111112p-4f};

// e^-l for 0 <= l < 10 in single precision: n = rint(l / ln2), r = n ln2 - l
// in two fma steps (ln2 split in 16 + 24 bits; n ln2_hi is exact), degree-7
// Taylor polynomial of e^r in Horner form (|r| <= ln2/2), scaled by 2^-n.
__device__ __forceinline__ float fexp_neg(float l) {
    const float n = __builtin_rintf(l * 0x1.715476p+0f);
    float r = __builtin_fmaf(n, 0x1.62e4p-1f, -l);
    r = __builtin_fmaf(n, 0x1.7f7d1cp-20f, r);
    float p = 0x1.a01a02p-13f;                 // 1/7!
    p = __builtin_fmaf(p, r, 0x1.6c16c2p-10f);  // 1/6!
    p = __builtin_fmaf(p, r, 0x1.111112p-7f);   // 1/5!
    p = __builtin_fmaf(p, r, 0x1.555556p-5f);   // 1/4!
    p = __builtin_fmaf(p, r, 0x1.555556p-3f);   // 1/3!
    p = __builtin_fmaf(p, r, 0.5f);
    p = __builtin_fmaf(p, r, 1.0f);
    p = __builtin_fmaf(p, r, 1.0f);
    return p * __int_as_float((127 - (int)n) << 23);
}

// The inversion result k for mean l and uniform u (both float), or -1 when
// the squeeze cannot decide it.
__device__ __forceinline__ int poisson_squeeze(float l, float u) {
    float p = fexp_neg(l);
    float G = p;
#pragma unroll 1
    for (int k = 0;; ++k) {
        if (u <= G - kSqueeze) return k;
        if (!(u > G + kSqueeze) || k == kSqueezeK) return -1;
        p = (p * l) * kRecipF[k + 1];
        G = G + p;
    }
}

// Poisson(lam) from the packed word u0 (inversion) or the draw's PTRS stream.
// Fast exit: 1 - lam <= e^-lam, and the computed exp is within 2^-51 of
// e^-lam, so u <= (1 - lam) - 2^-50 implies u <= F_0 and the inversion
// returns 0 -- the same value the full search gives, without the exp.  The
// rest of the inversion range goes through the float squeeze first.
__device__ __forceinline__ long long poisson(double lam, uint32_t u0, const DrawId& id) {
    if (!(lam > 0.0)) return 0;
    if (lam < kPoissonInvMax) {
        const double u = uniform01(u0);
        if (u <= (1.0 - lam) - 0x1p-50) return 0;
        const int k = poisson_squeeze((float)lam, (float)u);
        if (k >= 0) return k;
    }
    return poisson_slow(lam, u0, id);
}

// Leap draw w of counter (c0, c1, step, c3base | block) with mean lam and
// coarse byte hb, complete sampler: 0 when the coarse byte decides it, else
// the fine block's bits complete the word (PTRS draws its extra uniforms
// from the fine block's stream).
__device__ __forceinline__ long long leap_draw(double lam, uint32_t hb, uint32_t c0, uint32_t c1, uint32_t step,
                                               uint32_t c3base, uint32_t k0, uint32_t k1, uint32_t w) {
    if (coarse_zero(lam, hb)) return 0;
    const uint32_t fb = c3base | (kFineBlock + (w >> 2));
    const U4 f = philox10(c0, c1, step, fb, k0, k1);
    return poisson(lam, draw_word(hb, f, w), DrawId{c0, c1, step, fb, k0, k1, w & 3u});
}

// Largest single-precision mean the float squeeze takes: a float mean
// below it is within 8 2^-24 (relative) of a double mean below 10, so the
// double sampler would use the inversion search too.
constexpr float kSqueezeMaxF = 9.9999f;

// The fast part of a draw whose mean lf was formed in single precision
// (float operands and products, within 8 2^-24 lf of the double mean lam):
// 0 for lf <= 0; 0 when u_f <= (1 - lf) - 2^-19, which implies u <= e^-lam
// (u_f is within 2^-25 + 2^-33 of u, the margin covers both roundings and
// the mean's error for lf <= 1); the float squeeze below kSqueezeMaxF (its
// guard band 2^-15 covers the mean's error, <= 5e-6 there); -1 otherwise.
// A decided value is the one poisson() returns from the double mean.
__device__ __forceinline__ int poisson_fast_f(float lf, uint32_t u0) {
    if (!(lf > 0.0f)) return 0;
    const float u = __uint2float_rn(u0) * 0x1p-32f;
    if (u <= (1.0f - lf) - 0x1p-19f) return 0;
    if (!(lf < kSqueezeMaxF)) return -1;
    return poisson_squeeze(lf, u);
}

// The zero shortcuts alone: 0 when the draw is certainly 0 (the value
// poisson() returns), -1 when it may not be.  poisson_zero_f takes a float
// mean (as poisson_fast_f), poisson_zero a double one (as poisson()).
__device__ __forceinline__ int poisson_zero_f(float lf, uint32_t u0) {
    if (!(lf > 0.0f)) return 0;
    const float u = __uint2float_rn(u0) * 0x1p-32f;
    return u <= (1.0f - lf) - 0x1p-19f ? 0 : -1;
}
__device__ __forceinline__ int poisson_zero(double lam, uint32_t u0) {
    if (!(lam > 0.0)) return 0;
    return uniform01(u0) <= (1.0 - lam) - 0x1p-50 ? 0 : -1;
}

// The fast part of poisson() for kernels that defer the rare hard cases:
// the zero shortcut and the float squeeze settle the draw (k >= 0, the value
// poisson() returns); -1 leaves it to the double search (inversion_search)
// or, for mean >= 10, to PTRS.
__device__ __forceinline__ int poisson_fast(double lam, uint32_t u0) {
    if (!(lam > 0.0)) return 0;
    if (!(lam < kPoissonInvMax)) return -1;
    const double u = uniform01(u0);
    if (u <= (1.0 - lam) - 0x1p-50) return 0;
    return poisson_squeeze((float)lam, (float)u);
}

}  // namespace plis
