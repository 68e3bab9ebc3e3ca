// plis_kernels.hip -- gfx950 kernels for the RDME tau-leap + SSA step.
//
// One global step of one realization is two persistent kernels:
//   phase A (tau): per subvolume FP64 mass-action propensities, Cao tau' with
//     hops treated as first-order channels, the critical set and an
//     exponential tau'' clock; a wave-shuffle min/argmin per 64-subvolume
//     tile written as a tile partial; a one-block-per-realization decide
//     kernel reduces the partials and decides the step (tau, critical event).  Replaces compute_rates
//     (PLISCuda/src/rates.cu:82-96), compute_taus (tau.cu:197-251) and the
//     host thrust::min_element + tau[i] round trips (driver.cu:267-277).
//   phase B (leap): per subvolume Poisson firings of non-critical reactions,
//     gather-form diffusion (each subvolume regenerates its in-neighbours'
//     hop counts from the same counter-based Philox draw the sender used, so
//     there are no atomics on counts), the critical event of the argmin
//     subvolume, and the negative-population check; a one-block-per-
//     realization commit kernel accepts the step or leaps back (tau' halved,
//     new attempt index).  Replaces
//     leap_step (leap.cu:3-201), ssa_step (ssa.cu:56-110), check_state +
//     none_of + revert (leap.cu:203-212, driver.cu:292-323).
//
// Wavefront-per-tile: lane i of a wave owns subvolume v0+i of its tile;
// species-major counts make every per-species load and store one coalesced
// 256-byte access.  Blocks are dispatched round-robin over the 8 XCDs, so
// each XCD walks a contiguous eighth of the tile list (its L2 holds the
// neighbouring tiles it gathers from).
#include <type_traits>

#include "plis_numerics.h"
#include "plis_types.h"

namespace plis {

constexpr double kInf = __builtin_huge_val();

// The network tables are read-only for every kernel: reading them through
// the constant address space lets the compiler use scalar loads (scalar
// cache, no vector-memory round trip) wherever the index is wave-uniform.
template <class T>
using cptr = const T __attribute__((address_space(4)))*;

// Per-kernel read-only data (tile table, realization info and control)
// through the same address space.
template <class T>
__device__ __forceinline__ cptr<T> ro(const T* p) {
    return (cptr<T>)p;
}

// A whole read-only struct, word by word (merged into wide scalar loads).
template <class T>
__device__ __forceinline__ T ld_ro(const T* p) {
    static_assert(sizeof(T) % 4 == 0, "struct of 32-bit words");
    T v;
    uint32_t* d = reinterpret_cast<uint32_t*>(&v);
    const cptr<uint32_t> w = (cptr<uint32_t>)p;
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) d[i] = w[i];
    return v;
}

struct NetView {
    cptr<int32_t> nu, rs, rr, nrs, order, hor, mult;
    cptr<double> k, d;
    int S, R, C;
};

// MC = false: the network has one compartment, known at compile time, so
// every compartment lookup folds to 0 and the rate constants and diffusion
// constants become wave-uniform (scalar) loads.
template <bool MC = true>
__device__ __forceinline__ NetView net_view(const KernelArgs& a) {
    NetView n;
    n.nu = (cptr<int32_t>)(a.net_i + a.nh.off_nu);
    n.rs = (cptr<int32_t>)(a.net_i + a.nh.off_rs);
    n.rr = (cptr<int32_t>)(a.net_i + a.nh.off_rr);
    n.nrs = (cptr<int32_t>)(a.net_i + a.nh.off_nrs);
    n.order = (cptr<int32_t>)(a.net_i + a.nh.off_order);
    n.hor = (cptr<int32_t>)(a.net_i + a.nh.off_hor);
    n.mult = (cptr<int32_t>)(a.net_i + a.nh.off_mult);
    n.k = (cptr<double>)(a.net_d + a.nh.off_k);
    n.d = (cptr<double>)(a.net_d + a.nh.off_d);
    n.S = a.nh.S;
    n.R = a.nh.R;
    n.C = MC ? a.nh.C : 1;
    return n;
}

__device__ __forceinline__ uint32_t wave_uniform(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

// Orders one wavefront's LDS writes before its later LDS reads.
__device__ __forceinline__ void lds_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}


// XCD-aware persistent tile walk (grid is a multiple of 8 blocks).
__device__ __forceinline__ uint32_t tile_at(uint32_t it, uint32_t ntiles, uint32_t wpb) {
    const uint32_t G = gridDim.x;
    const uint32_t b = blockIdx.x;
    const uint32_t wv = wave_uniform(threadIdx.x >> 6);
    const uint32_t xcd = b & 7u;
    const uint32_t nbx = (G - xcd + 7u) >> 3;
    const uint32_t lb = b >> 3;
    const uint32_t cb = (uint32_t)(((uint64_t)ntiles * xcd) >> 3);
    const uint32_t ce = (uint32_t)(((uint64_t)ntiles * (xcd + 1u)) >> 3);
    const uint32_t t = cb + lb * wpb + wv + it * nbx * wpb;
    return t < ce ? t : 0xFFFFFFFFu;
}

// an interleaved edge entry's high word for weight 1.0 (a padding entry)
constexpr uint64_t kOneW = (uint64_t)0x3F800000u << 32;

__device__ __forceinline__ double binom_d(uint32_t x, int r) {
    if (x < (uint32_t)r) return 0.0;  // r >= 1
    const double xd = (double)x;
    if (r == 1) return xd;
    if (r == 2) return (xd * (xd - 1.0)) * 0.5;
    return ((xd * (xd - 1.0)) * (xd - 2.0)) / 6.0;
}

// Cao's critical test of one reactant, floor(x / r) < n_c, as the
// equivalent x < n_c r (integers, r >= 1, n_c >= 0): no division; n_c r is
// formed in 64 bits from scalars, one 64-bit compare per lane.
__device__ __forceinline__ bool below_nc(uint32_t x, int r, int nc) {
    return (uint64_t)x < (uint64_t)(uint32_t)nc * (uint64_t)(uint32_t)r;
}

// A hop is active (non-critical) when d > 0 and x >= max(n_c, 1).
__device__ __forceinline__ bool hop_active(double d, uint32_t x, int nc) {
    return d > 0.0 && x >= (uint32_t)(nc > 1 ? nc : 1);
}

__device__ __forceinline__ double g_cao(int hor, int mult, uint32_t xs) {
    const double xd = (double)xs;
    if (hor == 1) return 1.0;
    if (hor == 2) return mult == 2 ? 2.0 + 1.0 / (xd - 1.0) : 2.0;
    if (mult == 3) return (3.0 + 1.0 / (xd - 1.0)) + 2.0 / (xd - 2.0);
    if (mult == 2) return 1.5 * (2.0 + 1.0 / (xd - 1.0));
    return 3.0;
}

// eps x / g(x): for g = 1 and g = 2 (first order; second order with one
// molecule of the species) the quotient is formed without a divide -- x / 1
// and x / 2 are exact, so the result is the same IEEE value.
__device__ __forceinline__ double eps_x_over_g(double eps, int hor, int mult, uint32_t xs) {
    const double ex = eps * (double)xs;
    if (hor == 1) return ex;
    if (hor == 2 && mult != 2) return ex * 0.5;
    return ex / g_cao(hor, mult, xs);
}

// Reaction propensities a_j = k_j prod_s C(x_s, r_sj) and the critical flags
// (Cao: L_j = min_s floor(x_s / r_sj) < n_c, i.e. some reactant has
// floor(x_s / r_sj) < n_c) at one subvolume.
template <int MR, int MS>
__device__ __forceinline__ void react_props(const NetView& n, const uint32_t (&x)[MS], int c, int nc,
                                            double (&av)[MR], uint32_t (&crit)[(MR + 31) / 32]) {
#pragma unroll
    for (int i = 0; i < (MR + 31) / 32; ++i) crit[i] = 0u;
#pragma unroll
    for (int j = 0; j < MR; ++j) {
        if (j >= n.R) break;
        double aj = n.k[c * n.R + j];
        bool low = false;
        const int nr = n.nrs[j];
        for (int i = 0; i < nr; ++i) {
            const int s = n.rs[j * 3 + i], r = n.rr[j * 3 + i];
            uint32_t xs = 0;
#pragma unroll
            for (int q = 0; q < MS; ++q)
                if (q == s) xs = x[q];
            aj = aj * binom_d(xs, r);
            low = low || below_nc(xs, r, nc);
        }
        av[j] = aj;
        if (n.order[j] > 0 && low) crit[j >> 5] |= 1u << (j & 31);
    }
}

template <int MR>
__device__ __forceinline__ bool is_crit(const uint32_t (&crit)[(MR + 31) / 32], int j) {
    return (crit[j >> 5] >> (j & 31)) & 1u;
}

// Post-leap leap condition in propensity form (Cao, Gillespie & Petzold
// 2006, eq. 5, with their one-molecule floor; checked after the leap as in
// Anderson 2008).  A non-critical channel j with a_j(x) > 0 breaks it when
//     |a_j(x') - a_j(x)| > theta max(eps, phi_j) a_j(x),
// phi_j = max over reactants i of r_ij / x_i (one molecule's relative
// effect); evaluated without divides as D > (theta a_j) eps and
// D x_i > (theta a_j) r_ij for every reactant i.  An active hop channel of
// species s breaks it when |x'_s - x_s| > theta max(eps x_s, 1).  Bounds
// carry a relative slack of 2^-30 so that a change meeting one exactly (a
// one-molecule change at the floor) holds whatever the rounding of the
// products.  A negative or overflowing count breaks it too.  av / crit:
// propensities and critical flags of x.
constexpr double kLcSlack = 1.0 + 0x1p-30;
// av_new: a_j(x') when the caller has it (same operations, same values),
// else nullptr and the products are formed here.
template <int MS, int MR>
__device__ __forceinline__ uint32_t leap_condition_broken(const NetView& n, int c, int nc, double eps, double theta,
                                                          const uint32_t (&x)[MS], const long long (&nx)[MS],
                                                          const double (&av)[MR],
                                                          const uint32_t (&crit)[(MR + 31) / 32],
                                                          const double* av_new = nullptr) {
    uint32_t xn[MS];
    bool neg = false;
#pragma unroll
    for (int s = 0; s < MS; ++s) {
        neg = neg || (s < n.S && (nx[s] < 0 || nx[s] > 0xFFFFFFFFll));
        xn[s] = (uint32_t)nx[s];
    }
    if (neg) return 1u;
    uint32_t broken = 0u;
#pragma unroll
    for (int j = 0; j < MR; ++j) {
        if (j >= n.R) break;
        const int nr = n.nrs[j];
        if (is_crit<MR>(crit, j) || !(av[j] > 0.0) || nr == 0) continue;
        double aj;
        if (av_new) {
            aj = av_new[j];
        } else {
            aj = n.k[c * n.R + j];
            for (int i = 0; i < nr; ++i) {
                const int sp = n.rs[j * 3 + i], r = n.rr[j * 3 + i];
                uint32_t xs = 0;
#pragma unroll
                for (int q = 0; q < MS; ++q)
                    if (q == sp) xs = xn[q];
                aj = aj * binom_d(xs, r);
            }
        }
        const double D = __builtin_fabs(aj - av[j]);
        const double t = (theta * av[j]) * kLcSlack;
        bool over = D > t * eps;
        for (int i = 0; i < nr; ++i) {
            const int sp = n.rs[j * 3 + i], r = n.rr[j * 3 + i];
            uint32_t xs = 0;
#pragma unroll
            for (int q = 0; q < MS; ++q)
                if (q == sp) xs = x[q];
            over = over && D * (double)xs > t * (double)r;
        }
        if (over) broken = 1u;
    }
#pragma unroll
    for (int s = 0; s < MS; ++s) {
        if (s >= n.S) break;
        if (!hop_active(n.d[c * n.S + s], x[s], nc)) continue;
        if (__builtin_fabs((double)xn[s] - (double)x[s]) > (theta * fmax(eps * (double)x[s], 1.0)) * kLcSlack)
            broken = 1u;
    }
    return broken;
}

// ---------------------------------------------------------------------------
// per-realization step decision (tau = min(tau', tau''), clamp to t_end)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void decide(RealCtl& c) {
    double tau = fmin(c.tau_p, c.tau2);
    c.crit = (c.tau2 <= c.tau_p) ? 1 : 0;
    c.last = 0;
    if (c.t + tau >= c.t_end) {
        tau = c.t_end - c.t;
        c.crit = 0;
        c.last = 1;
    }
    c.tau = tau;
}

// Critical event at v*: a critical reaction or a critical hop, chosen with
// probability proportional to its propensity; a hop's edge is chosen in
// proportion to its weight.  Run by a whole wavefront (every value is
// wave-uniform).  (The reference picks a uniform slot among 6
// including self-padding, PLISCuda/src/leap.cu:180-194, ssa.cu:88-106.)
template <int MS, int MR>
__device__ void choose_event(const KernelArgs& a, const NetView& n, const RealInfo& ri, RealCtl& c,
                             const uint32_t* X) {
    const uint32_t v = c.vstar;
    const uint32_t N = ri.N;
    uint32_t x[MS];
#pragma unroll
    for (int s = 0; s < MS; ++s) x[s] = s < n.S ? X[ri.x_off + (uint64_t)s * N + v] : 0u;
    const int cc = n.C > 1 ? a.comp[ri.comp_off + v] : 0;
    double av[MR];
    uint32_t crit[(MR + 31) / 32];
    react_props<MR, MS>(n, x, cc, a.n_critical, av, crit);
    const uint32_t rb = a.rp[ri.rp_off + v], re = a.rp[ri.rp_off + v + 1];
    const bool weighted = (ri.flags & 2u) != 0;
    const double W = a.wsum[ri.rp_off + v];
    double a0c = 0.0;
#pragma unroll
    for (int j = 0; j < MR; ++j)
        if (j < n.R && is_crit<MR>(crit, j)) a0c += av[j];
#pragma unroll
    for (int s = 0; s < MS; ++s) {
        if (s >= n.S) break;
        const double d = n.d[cc * n.S + s];
        if (d > 0.0 && !hop_active(d, x[s], a.n_critical)) a0c += (d * (double)x[s]) * W;
    }
    const U4 o = philox10(v, kSelfId, c.step, ctr_word3(ri.id, 0u, kCrit, 0u), a.seed_lo, a.seed_hi);
    const double target = uniform01(o.x) * a0c;
    double acc = 0.0;
    int kind = 0, idx = -1;
    for (int j = 0; j < n.R; ++j) {
        if (!is_crit<MR>(crit, j) || !(av[j] > 0.0)) continue;
        acc += av[j];
        idx = j;
        if (target < acc) { kind = 1; break; }
    }
    int lastkind = idx >= 0 ? 1 : 0;
    if (kind == 0) {
        for (int s = 0; s < n.S; ++s) {
            const double d = n.d[cc * n.S + s];
            if (!(d > 0.0) || hop_active(d, x[s], a.n_critical)) continue;
            const double ao = (d * (double)x[s]) * W;
            if (!(ao > 0.0)) continue;
            acc += ao;
            idx = s;
            lastkind = 2;
            if (target < acc) { kind = 2; break; }
        }
    }
    if (kind == 0) kind = lastkind;
    c.ev_kind = kind;
    c.ev_idx = idx;
    c.ev_dst = v;
    if (kind == 2) {
        // the wave loads up to 64 edges of v's row at once; the weights are
        // then summed one edge at a time in CSR order
        const uint32_t lane = threadIdx.x & 63u;
        const double tw = uniform01(o.y) * W;
        double accw = 0.0;
        uint32_t dst = v;
        bool found = false;
        for (uint32_t q0 = rb; q0 < re && !found; q0 += 64u) {
            const uint32_t q = q0 + lane;
            const float wl = q < re ? (weighted ? a.w[ri.e_off + q] : 1.0f) : 0.0f;
            const uint32_t cl = q < re ? a.col[ri.e_off + q] : v;
            const uint32_t cnt = min(64u, re - q0);
            for (uint32_t i = 0; i < cnt; ++i) {
                accw += (double)__shfl(wl, (int)i);
                dst = (uint32_t)__shfl((int)cl, (int)i);
                if (tw < accw) {
                    found = true;
                    break;
                }
            }
        }
        c.ev_dst = dst;
    }
}

// ---------------------------------------------------------------------------
// phase A: tau' / tau'' per subvolume, min / argmin per realization
// ---------------------------------------------------------------------------
// Cao's tau' of one subvolume is assembled in three pieces, in this order:
// tau_own (reactions and out-hops), tau_in_edge once per in-neighbour in
// in-CSR order, tau_bound.  The FP64 sums therefore keep one order whether
// the in-edges are walked by the lane itself or staged through LDS.
template <int MS>
struct TauAcc {
    uint32_t x[MS];
    double mu[MS], sig[MS];
    double a0c;
};

template <int MS, int MR>
__device__ __forceinline__ void tau_own(const KernelArgs& a, const NetView& n, const RealInfo& ri,
                                        const uint32_t* X, uint32_t v, int c, TauAcc<MS>& t) {
    const uint32_t N = ri.N;
    const int nc = a.n_critical;
#pragma unroll
    for (int s = 0; s < MS; ++s) t.x[s] = s < n.S ? X[ri.x_off + (uint64_t)s * N + v] : 0u;
    double av[MR];
    uint32_t crit[(MR + 31) / 32];
    react_props<MR, MS>(n, t.x, c, nc, av, crit);
    const double W = a.wsum[ri.rp_off + v];
    t.a0c = 0.0;
#pragma unroll
    for (int j = 0; j < MR; ++j)
        if (j < n.R && is_crit<MR>(crit, j)) t.a0c += av[j];
#pragma unroll
    for (int s = 0; s < MS; ++s) {
        t.mu[s] = 0.0;
        t.sig[s] = 0.0;
    }
#pragma unroll
    for (int j = 0; j < MR; ++j) {
        if (j >= n.R) break;
        if (is_crit<MR>(crit, j)) continue;
#pragma unroll
        for (int s = 0; s < MS; ++s) {
            if (s >= n.S) break;
            const int nu = n.nu[s * n.R + j];
            if (nu == 0) continue;
            t.mu[s] += (double)nu * av[j];
            t.sig[s] += (double)(nu * nu) * av[j];
        }
    }
#pragma unroll
    for (int s = 0; s < MS; ++s) {
        if (s >= n.S) break;
        const double d = n.d[c * n.S + s];
        if (d > 0.0 && !hop_active(d, t.x[s], nc)) t.a0c += (d * (double)t.x[s]) * W;
        if (hop_active(d, t.x[s], nc)) {
            const double aout = (d * (double)t.x[s]) * W;
            t.mu[s] -= aout;
            t.sig[s] += aout;
        }
    }
}

// In-hops (nu = +1) from one in-neighbour, species whose hops are non-critical:
// du / xu are the neighbour's diffusion constants and counts.  A critical or
// absent hop (count below max(n_c, 1), or d = 0) adds +0 instead of being
// skipped: the sums of the active terms are the same values (only a -0 sum
// would turn into +0, and tau' uses |mu|), without a branch per term.
template <int MS>
__device__ __forceinline__ void tau_in_edge(int S, const double (&du)[MS], const uint32_t (&xu)[MS], double wq,
                                            int nc, TauAcc<MS>& t) {
    const uint32_t nc1 = nc > 1 ? (uint32_t)nc : 1u;
#pragma unroll
    for (int s = 0; s < MS; ++s) {
        // the small instantiations sum every species of the bound (the sums
        // of species S.. are never read; xu and du hold valid padding values
        // for them), so the loop has no runtime exit per species
        if constexpr (MS > 5)
            if (s >= S) break;
        uint32_t xe = xu[s] >= nc1 ? xu[s] : 0u;
        asm("" : "+v"(xe));  // select the count, not its conversion (one select instead of two)
        const double ain = (du[s] * (double)xe) * wq;
        t.mu[s] += ain;
        t.sig[s] += ain;
    }
}

template <int MS>
__device__ __forceinline__ double tau_bound(const KernelArgs& a, const NetView& n, const TauAcc<MS>& t) {
    double tp = kInf;
    const double eps = a.epsilon;
#pragma unroll
    for (int s = 0; s < MS; ++s) {
        if (s >= n.S) break;
        const int hor = n.hor[s];
        if (hor == 0) continue;
        const double mm = fmax(eps_x_over_g(eps, hor, n.mult[s], t.x[s]), 1.0);
        const double t1 = mm / __builtin_fabs(t.mu[s]);
        const double t2 = (mm * mm) / t.sig[s];
        tp = fmin(tp, fmin(t1, t2));
    }
    return tp;
}

// tau' and the critical propensity of one subvolume, the lane walking its
// own in-edges.
template <int MS, int MR>
__device__ __forceinline__ void tau_subvolume(const KernelArgs& a, const NetView& n, const RealInfo& ri,
                                              const uint32_t* X, uint32_t v, double& tp_out,
                                              double& a0c_out) {
    const uint32_t N = ri.N;
    const int c = n.C > 1 ? a.comp[ri.comp_off + v] : 0;
    TauAcc<MS> t;
    tau_own<MS, MR>(a, n, ri, X, v, c, t);
    const bool sym = (ri.flags & 1u) != 0;
    // in-edges: neighbour id and weight in one 8-byte load
    const uint64_t* icw = sym ? a.cw + ri.e_off : a.icw + ri.ie_off;
    const uint32_t* irp = sym ? a.rp + ri.rp_off : a.irp + ri.irp_off;
    const uint32_t ib = irp[v], ie = irp[v + 1];
    // one compartment: the neighbours' diffusion constants are the tile's
    double du[MS];
#pragma unroll
    for (int s = 0; s < MS; ++s) du[s] = n.d[s < n.S ? s : 0];
    // software pipeline: in-edge q's counts were requested in iteration q-1,
    // in-edge q+1's id and weight in iteration q-1 too, so each iteration
    // waits on nothing issued in it
    const uint64_t cw0 = ib < ie ? icw[ib] : (uint64_t)v | kOneW;
    const uint64_t cw1 = ib + 1u < ie ? icw[ib + 1u] : (uint64_t)v | kOneW;
    uint32_t u_cur = (uint32_t)cw0;
    float w_cur = __uint_as_float((uint32_t)(cw0 >> 32));
    uint32_t u_next = (uint32_t)cw1;
    float w_next = __uint_as_float((uint32_t)(cw1 >> 32));
    uint32_t xu[MS];
#pragma unroll
    for (int s = 0; s < MS; ++s) xu[s] = X[ri.x_off + (uint64_t)(s < n.S ? s : n.S - 1) * N + u_cur];
    for (uint32_t q = ib; q < ie; ++q) {
        const uint32_t u = u_cur;
        const double wq = (double)w_cur;
        uint32_t xc[MS];
#pragma unroll
        for (int s = 0; s < MS; ++s) xc[s] = xu[s];
        u_cur = u_next;
        w_cur = w_next;
        if (q + 1u < ie) {
#pragma unroll
            for (int s = 0; s < MS; ++s) xu[s] = X[ri.x_off + (uint64_t)(s < n.S ? s : n.S - 1) * N + u_cur];
            if (q + 2u < ie) {
                const uint64_t cw2 = icw[q + 2u];
                u_next = (uint32_t)cw2;
                w_next = __uint_as_float((uint32_t)(cw2 >> 32));
            }
        }
        if (n.C > 1) {
            const int cu = a.comp[ri.comp_off + u];
#pragma unroll
            for (int s = 0; s < MS; ++s) du[s] = n.d[cu * n.S + (s < n.S ? s : 0)];
        }
        tau_in_edge<MS>(n.S, du, xc, wq, a.n_critical, t);
    }
    tp_out = tau_bound<MS>(a, n, t);
    a0c_out = t.a0c;
}

__device__ __forceinline__ void reduce_partA(double& tp, double& t2, uint32_t& v2) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double otp = __shfl_xor(tp, off);
        const double ot2 = __shfl_xor(t2, off);
        const uint32_t ov = (uint32_t)__shfl_xor((int)v2, off);
        tp = fmin(tp, otp);
        if (ot2 < t2 || (ot2 == t2 && ov < v2)) {
            t2 = ot2;
            v2 = ov;
        }
    }
}

template <int MS, int MR>
__device__ void finalize_A(const KernelArgs& a, const NetView& n, const RealInfo& ri, uint32_t r,
                           double tp, double t2, uint32_t v2, RealCtl c) {
    const bool lead = (threadIdx.x & 63u) == 0u;
    if (tp == kInf && t2 == kInf) {
        c.status = kAbsorbed;
        c.done = 1u;
        if (lead) {
            a.ctl[r] = c;
            atomicAdd(a.ndone, 1u);
        }
        return;
    }
    c.tau_p = tp;
    c.tau2 = t2;
    c.vstar = v2;
    c.learned = (a.learned_on && !c.fallback) ? 1u : 0u;
    c.ev_kind = 0;
    c.ev_idx = -1;
    c.ev_dst = v2;
    decide(c);
    // the event is needed only if tau'' <= tau' (a leap-back halves tau' and
    // can lift the clamp to t_end, but never makes a larger tau'' fire)
    if (t2 < kInf && t2 <= tp) choose_event<MS, MR>(a, n, ri, c, c.cur ? a.x1 : a.x0);
    if (lead) a.ctl[r] = c;
}

// Cao's tau' / tau'' of one tile's subvolumes into the tile partial (the
// debug outputs, when given, get every subvolume's values instead).
// (tp, t2, v2) <- min tau', argmin tau'' (lowest subvolume on ties) of
// itself and (otp, ot2, ov)
__device__ __forceinline__ void merge_partA(double& tp, double& t2, uint32_t& v2, double otp, double ot2,
                                            uint32_t ov) {
    tp = fmin(tp, otp);
    if (ot2 < t2 || (ot2 == t2 && ov < v2)) {
        t2 = ot2;
        v2 = ov;
    }
}

// spec: the counts the leap pass just wrote (the buffer that becomes current
// if the step is accepted) and the next step's counter.
template <int MS, int MR>
__device__ __forceinline__ void cao_tau_tile(const KernelArgs& a, const NetView& n, const RealInfo& ri,
                                             const cptr<RealCtl> cp, const TileEnt& te, uint32_t t, bool dbg,
                                             bool spec = false) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t cur = wave_uniform(cp->cur) ^ (spec ? 1u : 0u);
    const uint32_t* X = cur ? a.x1 : a.x0;
    const uint32_t v = te.v0 + lane;
    double tp = kInf, t2 = kInf, a0c = 0.0;
    uint32_t v2 = v;
    if (v < ri.N) {
        tau_subvolume<MS, MR>(a, n, ri, X, v, tp, a0c);
        if (a0c > 0.0) {
            const uint32_t step = wave_uniform(cp->step) + (spec ? 1u : 0u);
            const U4 o = philox10(v, kSelfId, step, ctr_word3(ri.id, 0u, kTau2, 0u), a.seed_lo, a.seed_hi);
            t2 = -dlog(uniform01(o.x)) / a0c;
        }
        if (dbg) {
            const uint64_t base = ri.x_off / (uint64_t)n.S;
            a.dbg_tp[base + v] = tp;
            a.dbg_a0c[base + v] = a0c;
            a.dbg_t2[base + v] = t2;
        }
    }
    if (dbg) return;
    reduce_partA(tp, t2, v2);
    if (lane == 0) a.partA_tile[t] = PartA{tp, t2, v2, 0u};
}

// The one-compartment tau pass of 3 to 5 species fits in 64 VGPRs without
// spills; asking for 8 waves per SIMD (70 VGPRs and 7 waves otherwise)
// hides more of its gathers' latency (Brusselator mesh 77.6 -> 74.8 us; the
// one-species instantiation measured slower with it, 11.3 -> 12.4 us).
template <int MS, int MR, bool MC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(!MC && MS >= 3 && MS <= 5 ? 8 : 1)))
void phase_a_kernel(KernelArgs a) {
    for (uint32_t it = 0;; ++it) {
        const uint32_t t = tile_at(it, a.ntiles, a.waves_per_block);
        if (t == 0xFFFFFFFFu) break;
        // the network tables are read through pointers made opaque per tile,
        // so their scalar loads stay inside the tile's work instead of being
        // hoisted out of the tile loop into (spilled) scalar registers
        KernelArgs al = a;
        asm volatile("" : "+s"(al.net_i), "+s"(al.net_d));
        const NetView n = net_view<MC>(al);
        const TileEnt te = ld_ro(a.tiles + t);
        const uint32_t r = wave_uniform(te.r);
        const RealInfo ri = ld_ro(a.info + r);
        const cptr<RealCtl> cp = ro(a.ctl + r);
        const uint32_t done = wave_uniform(cp->done), retry = wave_uniform(cp->retry);
        const bool dbg = a.dbg_tp != nullptr;
        // (in the folded order a retried leap has just been made, so only
        // realizations that did not leap are skipped)
        const bool spec = a.tau_spec != 0u && !dbg;
        if (!dbg && (done || (retry && !spec))) continue;
        cao_tau_tile<MS, MR>(a, n, ri, cp, te, t, dbg, spec);
    }
}

// One block per realization: min tau' / argmin tau'' over the realization's
// tile partials (written by phase A; the kernel boundary makes them visible),
// then the step decision.
// Threads of decide_kernel and commit_decide_kernel: 4 wavefronts.  With
// 1,024 tile partials per chunk block that is 4 loads per thread, one
// round; 16-wavefront blocks, one load per thread, measured slower (the
// barrier and the 16-way reduction outweigh the round they save:
// commit_decide_kernel 12.8 -> 9.8 us on the Brusselator mesh).
template <int MS>
constexpr uint32_t decide_threads() { return 256u; }

// decide_kernel runs decide_chunks blocks per realization (one per 1,024
// tile partials of the largest realization, at most kDecideChunks): each
// reduces one chunk of the realization's tile partials (the reads spread
// over as many CUs), publishes its partial and counts itself in; the last
// block of the realization reduces the chunk partials and decides the
// step.  Release fence before the count, acquire fence after it (agent
// scope: the blocks run on different XCDs); the last block clears the
// counter for the next step (the next launch is ordered after this
// kernel).  With one block per realization it decides directly.
constexpr uint32_t kDecideChunks = 16;

template <int MS>
__device__ __forceinline__ void reduce_block_partA(double& tp, double& t2, uint32_t& v2, double* s_tp, double* s_t2,
                                                   uint32_t* s_v2) {
    reduce_partA(tp, t2, v2);
    const uint32_t wv = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    if (lane == 0) {
        s_tp[wv] = tp;
        s_t2[wv] = t2;
        s_v2[wv] = v2;
    }
    __syncthreads();
    if (wv == 0) {  // the first wavefront, every lane alike
        for (uint32_t w = 1; w < (blockDim.x >> 6); ++w) {
            tp = fmin(tp, s_tp[w]);
            if (s_t2[w] < t2 || (s_t2[w] == t2 && s_v2[w] < v2)) {
                t2 = s_t2[w];
                v2 = s_v2[w];
            }
        }
    }
}

template <int MS, int MR>
__global__ __launch_bounds__(decide_threads<MS>()) void decide_kernel(KernelArgs a) {
    const uint32_t nch = a.decide_chunks;
    const uint32_t r = blockIdx.x / nch, chunk = blockIdx.x % nch;
    if (blockIdx.x == 0 && threadIdx.x == 0) *a.redo_count = 0u;  // the coming leap pass's redo list
    if (r >= a.nreal) return;
    // the realization's control block and info in one round of loads
    const RealInfo ri = ld_ro(a.info + r);
    const RealCtl c0 = ld_ro(a.ctl + r);
    if (c0.done || c0.retry) return;  // the same in every block of the realization
    __shared__ double s_tp[16], s_t2[16];
    __shared__ uint32_t s_v2[16];
    __shared__ uint32_t s_last;
    const uint32_t per = (ri.ntiles + nch - 1u) / nch;
    const uint32_t qb = min(chunk * per, ri.ntiles), qe = min(qb + per, ri.ntiles);
    double tp = kInf, t2 = kInf;
    uint32_t v2 = 0xFFFFFFFFu;
    // 8 independent loads in flight per thread
    for (uint32_t q0 = qb; q0 < qe; q0 += 8u * blockDim.x) {
        PartA p[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const uint32_t q = q0 + (uint32_t)i * blockDim.x + threadIdx.x;
            p[i] = q < qe ? a.partA_tile[ri.tile_begin + q] : PartA{kInf, kInf, 0xFFFFFFFFu, 0u};
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            tp = fmin(tp, p[i].tp);
            if (p[i].t2 < t2 || (p[i].t2 == t2 && p[i].v2 < v2)) {
                t2 = p[i].t2;
                v2 = p[i].v2;
            }
        }
    }
    reduce_block_partA<MS>(tp, t2, v2, s_tp, s_t2, s_v2);
    if (nch == 1u) {
        if (threadIdx.x < 64u) finalize_A<MS, MR>(a, net_view(a), ri, r, tp, t2, v2, c0);
        return;
    }
    if (threadIdx.x == 0) {
        a.partA_chunk[r * nch + chunk] = PartA{tp, t2, v2, 0u};
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        s_last = atomicAdd(&a.decide_cnt[r], 1u) == nch - 1u ? 1u : 0u;
    }
    __syncthreads();
    if (!s_last || threadIdx.x >= 64u) return;
    // the last block of the realization: its first wavefront decides
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    const uint32_t lane = threadIdx.x;
    PartA p{kInf, kInf, 0xFFFFFFFFu, 0u};
    if (lane < nch) p = a.partA_chunk[r * nch + lane];
    tp = p.tp;
    t2 = p.t2;
    v2 = p.v2;
    reduce_partA(tp, t2, v2);
    if (lane == 0) a.decide_cnt[r] = 0u;
    finalize_A<MS, MR>(a, net_view(a), ri, r, tp, t2, v2, c0);
}

// ---------------------------------------------------------------------------
// phase B: leap + critical event + negative check, commit / leap-back
// ---------------------------------------------------------------------------
struct StepView {
    double tau;
    uint32_t step, attempt, real;
    int crit, ev_kind, ev_idx;
    uint32_t vstar, ev_dst;
};

template <int MS, int MR>
__device__ __forceinline__ bool leap_subvolume(const KernelArgs& a, const NetView& n, const RealInfo& ri,
                                               const uint32_t* X, uint32_t* Xn, uint32_t v,
                                               const StepView& sv, unsigned long long& fire_out,
                                               uint32_t& viol) {
    const uint32_t N = ri.N;
    const int nc = a.n_critical;
    uint32_t x[MS];
#pragma unroll
    for (int s = 0; s < MS; ++s) x[s] = s < n.S ? X[ri.x_off + (uint64_t)s * N + v] : 0u;
    const int c = n.C > 1 ? a.comp[ri.comp_off + v] : 0;
    double av[MR];
    uint32_t crit[(MR + 31) / 32];
    react_props<MR, MS>(n, x, c, nc, av, crit);
    long long nx[MS];
#pragma unroll
    for (int s = 0; s < MS; ++s) nx[s] = (long long)x[s];
    unsigned long long fire = 0;
    const double tau = sv.tau;
    const uint32_t k0 = a.seed_lo, k1 = a.seed_hi;

    // non-critical reactions: k_j ~ Poisson(tau a_j)
    {
        uint32_t blk = 0xFFFFFFFFu;
        U4 o{0, 0, 0, 0};
        const uint32_t c3base = ctr_word3(sv.real, sv.attempt, kReact, 0u);
        for (int j = 0; j < n.R; ++j) {
            if (is_crit<MR>(crit, j) || !(av[j] > 0.0)) continue;
            const uint32_t b = (uint32_t)j >> 4;
            if (b != blk) {
                o = philox10(v, kSelfId, sv.step, c3base | b, k0, k1);
                blk = b;
            }
            const long long k = leap_draw(tau * av[j], coarse_byte(o, (uint32_t)j), v, kSelfId, sv.step, c3base,
                                          k0, k1, (uint32_t)j);
            fire += (unsigned long long)k;
#pragma unroll
            for (int s = 0; s < MS; ++s)
                if (s < n.S) nx[s] += k * (long long)n.nu[s * n.R + j];
        }
    }

    const bool weighted = (ri.flags & 2u) != 0;
    const bool sym = (ri.flags & 1u) != 0;
    const uint32_t rb = a.rp[ri.rp_off + v], re = a.rp[ri.rp_off + v + 1];
    double dx[MS];
    bool act[MS];
#pragma unroll
    for (int s = 0; s < MS; ++s) {
        const double d = s < n.S ? n.d[c * n.S + s] : 0.0;
        act[s] = s < n.S && hop_active(d, x[s], nc);
        dx[s] = d * (double)x[s];
    }
    const uint32_t c3hop = ctr_word3(sv.real, sv.attempt, kHop, 0u);
    if (sym) {
        // One coarse block per (edge, species pair) serves both directions of
        // both species.  Pairs outer (unrolled: static indices keep the count
        // accumulators in registers), edges inner (rolled: few live values).
        const uint32_t* colp = a.col + ri.e_off;
        const float* wp = a.w + ri.e_off;
#pragma unroll
        for (int s0 = 0; s0 < MS; s0 += 2) {
            if (s0 >= n.S) break;
            const bool two = s0 + 1 < MS && s0 + 1 < n.S;
            const bool act0 = act[s0], act1 = two && act[s0 + 1 < MS ? s0 + 1 : s0];
            const double dx0 = dx[s0], dx1 = two ? dx[s0 + 1 < MS ? s0 + 1 : s0] : 0.0;
            const uint32_t* X0 = X + ri.x_off + (uint64_t)s0 * N;
            const uint32_t* X1 = two ? X0 + N : X0;
            const uint32_t cb = c3hop | ((uint32_t)s0 >> 3);
            const uint32_t w0 = 2u * (uint32_t)s0;
            long long d0 = 0, d1 = 0;
            for (uint32_t q = rb; q < re; ++q) {
                const uint32_t u = colp[q];
                const double wq = weighted ? (double)wp[q] : 1.0;
                const int cu = n.C > 1 ? a.comp[ri.comp_off + u] : 0;
                const uint32_t xu0 = X0[u];
                const uint32_t xu1 = two ? X1[u] : 0u;
                const double du0 = n.d[cu * n.S + s0];
                const double du1 = two ? n.d[cu * n.S + s0 + 1] : 0.0;
                const double lo0 = act0 ? tau * (dx0 * wq) : 0.0;
                const double lo1 = act1 ? tau * (dx1 * wq) : 0.0;
                const double li0 = hop_active(du0, xu0, nc) ? tau * ((du0 * (double)xu0) * wq) : 0.0;
                const double li1 = (two && hop_active(du1, xu1, nc)) ? tau * ((du1 * (double)xu1) * wq) : 0.0;
                if (!(lo0 > 0.0 || lo1 > 0.0 || li0 > 0.0 || li1 > 0.0)) continue;
                const uint32_t lo = v < u ? v : u, hi = v < u ? u : v;
                const uint32_t dout = v > u ? 1u : 0u, din = 1u - dout;
                const U4 o = philox10(lo, hi, sv.step, cb, k0, k1);
                auto hop = [&](double l, uint32_t w) {
                    return leap_draw(l, coarse_byte(o, w), lo, hi, sv.step, c3hop, k0, k1, w);
                };
                if (lo0 > 0.0) {
                    const long long k = hop(lo0, w0 + dout);
                    d0 -= k;
                    fire += (unsigned long long)k;
                }
                if (li0 > 0.0) d0 += hop(li0, w0 + din);
                if (lo1 > 0.0) {
                    const long long k = hop(lo1, w0 + 2u + dout);
                    d1 -= k;
                    fire += (unsigned long long)k;
                }
                if (li1 > 0.0) d1 += hop(li1, w0 + 2u + din);
            }
#pragma unroll
            for (int q = 0; q < MS; ++q) {
                if (q == s0) nx[q] += d0;
                if (q == s0 + 1) nx[q] += d1;
            }
        }
    } else {
        for (uint32_t q = rb; q < re; ++q) {
            const uint32_t u = a.col[ri.e_off + q];
            const double wq = weighted ? (double)a.w[ri.e_off + q] : 1.0;
            const uint32_t lo = v < u ? v : u, hi = v < u ? u : v;
            const uint32_t dout = v > u ? 1u : 0u;
#pragma unroll
            for (int s = 0; s < MS; ++s) {
                if (s >= n.S) break;
                if (!act[s]) continue;
                const uint32_t wd = 2u * (uint32_t)s + dout;
                const U4 o = philox10(lo, hi, sv.step, c3hop | (wd >> 4), k0, k1);
                const long long k =
                    leap_draw(tau * (dx[s] * wq), coarse_byte(o, wd), lo, hi, sv.step, c3hop, k0, k1, wd);
                nx[s] -= k;
                fire += (unsigned long long)k;
            }
        }
        const uint32_t ib = a.irp[ri.irp_off + v], ie = a.irp[ri.irp_off + v + 1];
        for (uint32_t q = ib; q < ie; ++q) {
            const uint32_t u = a.icol[ri.ie_off + q];
            const double wq = weighted ? (double)a.iw[ri.ie_off + q] : 1.0;
            const int cu = n.C > 1 ? a.comp[ri.comp_off + u] : 0;
            const uint32_t lo = v < u ? v : u, hi = v < u ? u : v;
            const uint32_t din = u > v ? 1u : 0u;
#pragma unroll
            for (int s = 0; s < MS; ++s) {
                if (s >= n.S) break;
                const uint32_t xu = X[ri.x_off + (uint64_t)s * N + u];
                const double du = n.d[cu * n.S + s];
                if (!hop_active(du, xu, nc)) continue;
                const uint32_t wd = 2u * (uint32_t)s + din;
                const U4 o = philox10(lo, hi, sv.step, c3hop | (wd >> 4), k0, k1);
                nx[s] += leap_draw(tau * ((du * (double)xu) * wq), coarse_byte(o, wd), lo, hi, sv.step, c3hop, k0,
                                   k1, wd);
            }
        }
    }

    // the critical event (one per global step, in the argmin-tau'' subvolume)
    if (sv.crit) {
        if (v == sv.vstar) {
            if (sv.ev_kind == 1) {
#pragma unroll
                for (int s = 0; s < MS; ++s)
                    if (s < n.S) nx[s] += n.nu[s * n.R + sv.ev_idx];
                fire += 1ull;
            } else if (sv.ev_kind == 2) {
#pragma unroll
                for (int s = 0; s < MS; ++s)
                    if (s == sv.ev_idx) nx[s] -= 1;
                fire += 1ull;
            }
        }
        if (sv.ev_kind == 2 && v == sv.ev_dst && sv.ev_dst != sv.vstar) {
#pragma unroll
            for (int s = 0; s < MS; ++s)
                if (s == sv.ev_idx) nx[s] += 1;
        }
    }
    bool bad = false;
#pragma unroll
    for (int s = 0; s < MS; ++s) {
        if (s >= n.S) break;
        bad = bad || nx[s] < 0 || nx[s] > 0xFFFFFFFFll;
        Xn[ri.x_off + (uint64_t)s * N + v] = (uint32_t)nx[s];
    }
    fire_out = fire;
    viol = a.judge_on ? leap_condition_broken<MS, MR>(n, c, nc, a.epsilon, a.theta, x, nx, av, crit) : 0u;
    return bad;
}

// ---------------------------------------------------------------------------
// Edge-parallel leap of one tile (symmetric realizations).
//
// Each undirected edge {v, u}, v < u, is drawn once, by the tile of its
// lower end: the tile's forward-CSR entries are contiguous, lane l takes
// entries l, l+64, ... (coalesced loads, no divergence from uneven degrees),
// each one 8-byte word that holds the neighbour, the owning subvolume's
// index in the tile and the weight, and draws both hop directions of every species from the
// edge's coarse Philox block; then lane i draws the reactions of subvolume
// v0+i (one coarse block per 16 reactions).
// Balances of the tile's own subvolumes collect in LDS, those of other
// tiles' subvolumes go straight to the global balances (gdelta); integer
// atomics keep every sum exact and order-free.  phase_b_finish_kernel adds
// the balances to the counts once every draw of the pass is in.
//
// Compacted draws.  Almost every draw of a step is 0 (hop means are ~1e-3
// on the headline mesh), so phase_b_kernel's walk only applies the zero
// shortcut to the coarse byte of each draw's word: a draw it cannot settle
// as 0 goes to the wavefront's LDS queue as its code and coarse byte, and
// after the walk the whole wavefront draws the queued ones, one per lane --
// the fine bits, the float squeeze, then the double search where the
// squeeze cannot decide -- so the fine block and the search run once per 64
// open draws instead of once per draw site.  A tile whose queue
// overflows, or that holds a draw whose mean may be 10 or more (PTRS; see
// kPtrsMaybeF), applies nothing
// from its queue and goes on the redo list: phase_b_redo_kernel walks it
// again and draws every draw the zero shortcut leaves open (the same test
// on the same inputs) with the complete sampler.  Draws are functions of
// their counters, so the two passes give the tile the values one pass
// would.  No call and no PTRS in phase_b_kernel keeps its register
// footprint, and so its occupancy, at that of the walk.
// ---------------------------------------------------------------------------
constexpr uint32_t kQueue = 256;               // queued draws per wavefront
constexpr uint32_t kQueueReact = 0x80000000u;  // reaction code: flag | lane << 8 | j
constexpr uint32_t kQueueDefer = kQueue + 1u;  // count bump that sends the tile to the redo pass
// A hop's float mean is within 8 * 2^-24 (relative) of its double mean, so a
// float mean at or below kPtrsMaybeF is a double mean below 10 (at most
// 9.99990 * (1 + 5e-7)): the walk defers a tile with a float mean above it,
// and every queued draw of a tile it does not defer has a double mean below
// 10 (a deferred tile whose means are below 10 after all gets the same
// values from the redo pass's complete sampler).
constexpr float kPtrsMaybeF = 9.9999f;
constexpr uint32_t kMaxTileEdges = 1u << 27;   // hop codes hold the entry offset in 27 bits

// Queue append by the whole wavefront (every lane calls it at the same
// point, qn is wave-uniform): the lanes whose draw is open take consecutive
// slots in lane order (a ballot and its prefix count, no LDS atomics).
__device__ __forceinline__ void q_append(uint2* s_q, uint32_t& qn, bool open, uint32_t code, uint32_t word) {
    const unsigned long long mq = __ballot(open);
    if (mq == 0ull) return;
    if (open) {
        const uint32_t slot =
            qn + __builtin_amdgcn_mbcnt_hi((uint32_t)(mq >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mq, 0u));
        if (slot < kQueue) s_q[slot] = make_uint2(code, word);
    }
    qn += (uint32_t)__popcll(mq);
}

__device__ __forceinline__ double hop_mean(double d, uint32_t x, double wq, double tau, int nc) {
    return hop_active(d, x, nc) ? tau * ((d * (double)x) * wq) : 0.0;
}

// The same mean in single precision for the fast path: tau, d and x rounded
// to float (w already is), three float products; within 8 2^-24 of it.
__device__ __forceinline__ float hop_mean_f(double d, uint32_t x, float w, float tauf, int nc) {
    return hop_active(d, x, nc) ? tauf * (((float)d * (float)x) * w) : 0.0f;
}

// Balance b of species s at subvolume u: LDS when u is in the tile.
__device__ __forceinline__ void add_balance(const KernelArgs& a, const RealInfo& ri,
                                            unsigned long long (*s_del)[kTile], uint32_t v0, int s, uint32_t u,
                                            long long b) {
    if (u - v0 < (uint32_t)kTile) atomicAdd(&s_del[s][u - v0], (unsigned long long)b);
    else atomicAdd(&a.gdelta[ri.x_off + (uint64_t)s * ri.N + u], (unsigned long long)b);
}

// The lane's draws: its forward entries (both directions of every species
// of an undirected edge), then the reactions of its subvolume.  Walk: the
// zero shortcut on the coarse bytes, draws it leaves open queued; returns
// the queue count (wave-uniform; above kQueue when the tile goes to the
// redo pass).  REDO: the open draws, with the complete sampler, applied.
// Control flow stays wave-uniform around every queue append: the loops run
// the same trip count on every lane, lanes without work carry flags.
template <int MS, int MR, bool REDO>
__device__ __forceinline__ uint32_t walk_fwd_draws(const KernelArgs& a, const NetView& n, const RealInfo& ri,
                                                   const uint32_t* X, uint32_t v0, const StepView& sv,
                                                   const uint32_t (*s_x)[kTile],
                                                   unsigned long long (*s_del)[kTile], uint32_t eb, uint32_t ee,
                                                   const int32_t* s_c, uint2* s_q, unsigned long long& fire) {
    static_assert(MS <= 8, "one coarse block per edge; hop codes hold w = 2s + direction in 4 bits");
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t N = ri.N;
    const int nc = a.n_critical;
    const uint64_t* cwp = a.fcw + ri.fe_off;  // neighbour, row in the tile and weight in one load
    const double tau = sv.tau;
    const float tauf = (float)tau;
    const uint32_t k0 = a.seed_lo, k1 = a.seed_hi;
    const uint32_t c3hop = ctr_word3(sv.real, sv.attempt, kHop, 0u);
    uint32_t qn = 0u;
    bool defer = false;
    // one compartment: the diffusion constants as float and the species
    // that hop at all (hop_mean_f's operands and test, loaded once)
    const uint32_t nc1 = nc > 1 ? (uint32_t)nc : 1u;
    float df[MS];
    bool dpos[MS];
#pragma unroll
    for (int s = 0; s < MS; ++s) {
        const double d = n.d[s < n.S ? s : 0];
        df[s] = (float)d;
        dpos[s] = s < n.S && d > 0.0;
    }
    // software pipeline: an entry's neighbour counts are requested one
    // iteration ahead, its neighbour id and weight two iterations ahead
    // (counts of a neighbour inside the tile are read from global memory
    // too: the staged copy holds the same values)
    auto counts = [&](uint32_t u, uint32_t (&xo)[MS]) {
#pragma unroll
        for (int s = 0; s < MS; ++s) xo[s] = X[ri.x_off + (uint64_t)(s < n.S ? s : n.S - 1) * N + u];
    };
    const uint32_t e0 = eb + lane;
    const uint64_t cw0 = e0 < ee ? cwp[e0] : (uint64_t)v0 | kOneW;
    const uint64_t cw1 = e0 + (uint32_t)kTile < ee ? cwp[e0 + kTile] : (uint64_t)v0 | kOneW;
    uint32_t u_cur = (uint32_t)cw0 & kFwdIdMask, u_next = (uint32_t)cw1 & kFwdIdMask;
    uint32_t i_cur = (uint32_t)cw0 >> 26, i_next = (uint32_t)cw1 >> 26;
    float w_cur = __uint_as_float((uint32_t)(cw0 >> 32));
    float w_next = __uint_as_float((uint32_t)(cw1 >> 32));
    uint32_t xn[MS];
#pragma unroll
    for (int s = 0; s < MS; ++s) xn[s] = 0u;
    if (e0 < ee) counts(u_cur, xn);
    const uint32_t nit = (ee - eb + (uint32_t)kTile - 1u) / (uint32_t)kTile;
    for (uint32_t it = 0; it < nit; ++it) {
        const uint32_t e = e0 + it * (uint32_t)kTile;
        const bool valid = e < ee;
        const uint32_t i = i_cur;  // the edge's row in the tile
        const uint32_t vo = v0 + i;
        const uint32_t u = u_cur;  // u > vo: vo is the edge's lower end
        const float wf = w_cur;
        uint32_t xc[MS];
#pragma unroll
        for (int s = 0; s < MS; ++s) xc[s] = xn[s];
        u_cur = u_next;
        w_cur = w_next;
        i_cur = i_next;
        if (e + (uint32_t)kTile < ee) {
            counts(u_cur, xn);
            if (e + 2u * (uint32_t)kTile < ee) {
                const uint64_t cw2 = cwp[e + 2u * kTile];
                u_next = (uint32_t)cw2 & kFwdIdMask;
                i_next = (uint32_t)cw2 >> 26;
                w_next = __uint_as_float((uint32_t)(cw2 >> 32));
            }
        }
        const bool local = u - v0 < (uint32_t)kTile;
        const int co = n.C > 1 ? s_c[i] : 0;
        const int cu = (n.C > 1 && valid) ? (local ? s_c[u - v0] : a.comp[ri.comp_off + u]) : 0;
        const uint32_t ecode = (e - eb) << 4;
        // the means of the edge's draws w = 2s + direction (0: vo -> u) and
        // their coarse bounds
        float lw[2 * MS], gw[2 * MS];
        bool any = false;
#pragma unroll
        for (int s = 0; s < MS; ++s) {
            const int sq = s < n.S ? s : 0;
            if (n.C > 1) {
                const bool on = valid && s < n.S;
                lw[2 * s] = on ? hop_mean_f(n.d[co * n.S + sq], s_x[sq][i], wf, tauf, nc) : 0.0f;
                lw[2 * s + 1] = on ? hop_mean_f(n.d[cu * n.S + sq], xc[s], wf, tauf, nc) : 0.0f;
            } else {
                const bool on = valid && dpos[s];
                const uint32_t xo = s_x[sq][i], xu = xc[s];
                lw[2 * s] = (on && xo >= nc1) ? tauf * ((df[s] * (float)xo) * wf) : 0.0f;
                lw[2 * s + 1] = (on && xu >= nc1) ? tauf * ((df[s] * (float)xu) * wf) : 0.0f;
            }
        }
#pragma unroll
        for (int w = 0; w < 2 * MS; ++w) {
            gw[w] = coarse_bound(lw[w]);
            any = any || lw[w] > 0.0f;
            if constexpr (!REDO) defer = defer || lw[w] > kPtrsMaybeF;
        }
        if (__ballot(any) == 0ull) continue;
        const U4 o = philox10(vo, u, sv.step, c3hop, k0, k1);  // the coarse block (MS <= 8)
#pragma unroll
        for (int w = 0; w < 2 * MS; ++w) {
            const uint32_t hb = coarse_byte(o, (uint32_t)w);
            const bool open = (float)hb > gw[w];
            if constexpr (REDO) {
                if (!open) continue;
                const int s = w >> 1;
                const bool fwd = (w & 1) == 0;
                const uint32_t fb = c3hop | (kFineBlock + ((uint32_t)w >> 2));
                const U4 f = philox10(vo, u, sv.step, fb, k0, k1);
                const uint32_t word = draw_word(hb, f, (uint32_t)w);
                if (poisson_zero_f(lw[w], word) >= 0) continue;
                const double lam = fwd ? hop_mean(n.d[co * n.S + s], s_x[s][i], (double)wf, tau, nc)
                                       : hop_mean(n.d[cu * n.S + s], xc[s], (double)wf, tau, nc);
                const long long k = poisson(lam, word, DrawId{vo, u, sv.step, fb, k0, k1, (uint32_t)w & 3u});
                if (k == 0) continue;
                fire += (unsigned long long)k;
                const long long bv = fwd ? -k : k;  // balance of vo
                atomicAdd(&s_del[s][i], (unsigned long long)bv);
                add_balance(a, ri, s_del, v0, s, u, -bv);
            } else {
                q_append(s_q, qn, open, ecode | (uint32_t)w, hb);
            }
        }
    }
    const uint32_t v = v0 + lane;
    const bool vok = v < N;
    // reactions of subvolume v, propensities and criticality as react_props
    // computes them (same operations, one reaction at a time)
    const int c = (n.C > 1 && vok) ? s_c[lane] : 0;
    uint32_t blk = 0xFFFFFFFFu;
    U4 o{0, 0, 0, 0};
    const uint32_t c3base = ctr_word3(sv.real, sv.attempt, kReact, 0u);
    for (int j = 0; j < n.R; ++j) {
        double aj = n.k[c * n.R + j];
        bool low = false;
        const int nr = n.nrs[j];
        for (int q = 0; q < nr; ++q) {
            const int sq = n.rs[j * 3 + q], rq = n.rr[j * 3 + q];
            const uint32_t xs = s_x[sq][lane];
            aj = aj * binom_d(xs, rq);
            low = low || below_nc(xs, rq, nc);
        }
        const bool live = vok && !(n.order[j] > 0 && low) && aj > 0.0;
        const uint32_t b = (uint32_t)j >> 4;
        if (b != blk) {
            o = philox10(v, kSelfId, sv.step, c3base | b, k0, k1);
            blk = b;
        }
        const uint32_t hb = coarse_byte(o, (uint32_t)j);
        const double lam = tau * aj;
        const bool open = live && !coarse_zero(lam, hb);
        if constexpr (REDO) {
            if (!open) continue;
            const long long k = leap_draw(lam, hb, v, kSelfId, sv.step, c3base, k0, k1, (uint32_t)j);
            if (k == 0) continue;
            fire += (unsigned long long)k;
            for (int s = 0; s < n.S; ++s) {
                const int nu = n.nu[s * n.R + j];
                if (nu != 0) atomicAdd(&s_del[s][lane], (unsigned long long)(k * (long long)nu));
            }
        } else {
            defer = defer || (open && !(lam < kPoissonInvMax));
            q_append(s_q, qn, open, kQueueReact | (lane << 8) | (uint32_t)j, hb);
        }
    }
    // a mean >= 10 (PTRS) anywhere: the tile goes to the redo pass
    if (__ballot(defer) != 0ull) qn += kQueueDefer;
    return qn;
}

// A queued hop, decoded: owner i, neighbour u, draw w = 2s + direction,
// species s and the mean in double.
struct QueuedHop {
    uint32_t i, u, w;
    int s;
    bool fwd;  // the hop v0 + i -> u
    double lam;
    float lf;  // the mean as the walk formed it in single precision
};

template <int MS>
__device__ __forceinline__ QueuedHop queued_hop(const KernelArgs& a, const NetView& n, const RealInfo& ri,
                                                const uint32_t* X, uint32_t v0, const StepView& sv,
                                                const uint32_t (*s_x)[kTile], uint32_t eb,
                                                const int32_t* s_c, uint32_t code) {
    QueuedHop h;
    const uint32_t e = eb + (code >> 4);
    h.w = code & 15u;
    const uint64_t cw = a.fcw[ri.fe_off + e];
    h.i = (uint32_t)cw >> 26;
    h.u = (uint32_t)cw & kFwdIdMask;
    h.s = (int)(h.w >> 1);
    h.fwd = (h.w & 1u) == 0u;
    const float wf = __uint_as_float((uint32_t)(cw >> 32));
    const double wq = (double)wf;
    double d;
    uint32_t xs;
    if (h.fwd) {
        d = n.d[(n.C > 1 ? s_c[h.i] : 0) * n.S + h.s];
        xs = s_x[h.s][h.i];
    } else {
        const bool local = h.u - v0 < (uint32_t)kTile;
        const int cu = n.C > 1 ? (local ? s_c[h.u - v0] : a.comp[ri.comp_off + h.u]) : 0;
        d = n.d[cu * n.S + h.s];
        xs = local ? s_x[h.s][h.u - v0] : X[ri.x_off + (uint64_t)h.s * ri.N + h.u];
    }
    h.lam = hop_mean(d, xs, wq, sv.tau, a.n_critical);
    h.lf = hop_mean_f(d, xs, wf, (float)sv.tau, a.n_critical);
    return h;
}

// One queued draw: its fine bits complete the word the walk queued the
// coarse byte of (one Philox call, whichever the kind), then the float
// squeeze on the mean rounded to float, the double search where it cannot
// decide (the drain runs only when every queued mean is below 10), applied
// to the balances.
template <int MS>
__device__ __forceinline__ void drain_fwd_draw(const KernelArgs& a, const NetView& n, const RealInfo& ri,
                                               const uint32_t* X, uint32_t v0, const StepView& sv,
                                               const uint32_t (*s_x)[kTile], unsigned long long (*s_del)[kTile],
                                               uint32_t eb, const int32_t* s_c, uint2 q,
                                               unsigned long long& fire) {
    const uint32_t code = q.x, hb = q.y;
    const bool react = (code & kQueueReact) != 0u;
    uint32_t i, c1, w, c3;
    float lf;
    double lam;
    QueuedHop h{};
    if (react) {
        i = (code >> 8) & 63u;
        w = code & 255u;
        double aj = n.k[(n.C > 1 ? s_c[i] : 0) * n.R + (int)w];
        const int nr = n.nrs[w];
        for (int r = 0; r < nr; ++r) aj = aj * binom_d(s_x[n.rs[w * 3 + r]][i], n.rr[w * 3 + r]);
        lam = sv.tau * aj;
        lf = (float)lam;
        c1 = kSelfId;
        c3 = ctr_word3(sv.real, sv.attempt, kReact, kFineBlock + (w >> 2));
    } else {
        h = queued_hop<MS>(a, n, ri, X, v0, sv, s_x, eb, s_c, code);
        i = h.i;
        w = h.w;
        lam = h.lam;
        lf = h.lf;
        c1 = h.u;
        c3 = ctr_word3(sv.real, sv.attempt, kHop, kFineBlock + (w >> 2));
    }
    const uint32_t word = draw_word(hb, philox10(v0 + i, c1, sv.step, c3, a.seed_lo, a.seed_hi), w);
    long long k = lf < kSqueezeMaxF ? poisson_squeeze(lf, __uint2float_rn(word) * 0x1p-32f) : -1;
    if (k < 0) k = inversion_search(lam, word);
    if (k == 0) return;
    fire += (unsigned long long)k;
    if (react) {
        for (int s = 0; s < n.S; ++s) {
            const int nu = n.nu[s * n.R + (int)w];
            if (nu != 0) atomicAdd(&s_del[s][i], (unsigned long long)(k * (long long)nu));
        }
    } else {
        const long long b = h.fwd ? -k : k;  // balance of v0 + i
        atomicAdd(&s_del[h.s][i], (unsigned long long)b);
        add_balance(a, ri, s_del, v0, h.s, h.u, -b);
    }
}

// The draws of one tile into the global balances.  Returns false when the
// tile's open draws are left to phase_b_redo_kernel (its settled draws are
// applied either way).
template <int MS, int MR, bool REDO>
__device__ __forceinline__ bool leap_tile_fwd(const KernelArgs& a, const NetView& n, const RealInfo& ri,
                                              const uint32_t* X, uint32_t v0, const StepView& sv,
                                              uint32_t (*s_x)[kTile], unsigned long long (*s_del)[kTile],
                                              int32_t* s_c, uint2* s_q,
                                              unsigned long long& fire) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t N = ri.N;
    const uint32_t v = v0 + lane;
    const bool active = v < N;
#pragma unroll
    for (int s = 0; s < MS; ++s) {
        s_x[s][lane] = (active && s < n.S) ? X[ri.x_off + (uint64_t)s * N + v] : 0u;
        s_del[s][lane] = 0ull;
    }
    s_c[lane] = (n.C > 1 && active) ? a.comp[ri.comp_off + v] : 0;
    const uint32_t eb = wave_uniform(a.frp[ri.rp_off + v0]);
    const uint32_t ee = wave_uniform(a.frp[ri.rp_off + min(v0 + (uint32_t)kTile, N)]);
    lds_wave_sync();
    const uint32_t qn = walk_fwd_draws<MS, MR, REDO>(a, n, ri, X, v0, sv, s_x, s_del, eb, ee, s_c, s_q, fire);
    lds_wave_sync();
    bool complete = true;
    if constexpr (!REDO) {
        const uint32_t cnt = wave_uniform(qn);
        // (a deferred tile's count is above kQueue: a draw whose mean may be
        // 10 or more needs PTRS, so nothing queued is drawn and the redo
        // pass draws all of it)
        complete = cnt <= kQueue;
        if (complete) {
            for (uint32_t q = lane; q < cnt; q += (uint32_t)kTile)
                drain_fwd_draw<MS>(a, n, ri, X, v0, sv, s_x, s_del, eb, s_c, s_q[q], fire);
            lds_wave_sync();
        }
    }
    if (active) {
#pragma unroll
        for (int s = 0; s < MS; ++s)
            if (s < n.S && s_del[s][lane] != 0ull) atomicAdd(&a.gdelta[ri.x_off + (uint64_t)s * N + v], s_del[s][lane]);
    }
    return complete;
}

// The end of the leap pass for one tile: new counts = counts + global
// balances (which are cleared again), the critical event, the negativity
// check, the post-leap leap condition (when the judge is on) and the tile's
// statistics.  nx returns the lane's new counts (any sign).
template <int MS, int MR>
__device__ __forceinline__ void finish_tile(const KernelArgs& a, const NetView& n, const RealInfo& ri,
                                            const uint32_t* X, uint32_t* Xn, uint32_t v0, const StepView& sv,
                                            unsigned long long& fire, uint32_t& bad, uint32_t& viol,
                                            uint32_t (&x)[MS], long long (&nx)[MS], bool judge) {
    const uint32_t v = v0 + (threadIdx.x & 63u);
#pragma unroll
    for (int s = 0; s < MS; ++s) {
        x[s] = 0u;
        nx[s] = 0;
    }
    if (v >= ri.N) return;
    const bool at_star = sv.crit && v == sv.vstar && (sv.ev_kind == 1 || sv.ev_kind == 2);
    const bool at_dst = sv.crit && sv.ev_kind == 2 && v == sv.ev_dst && sv.ev_dst != sv.vstar;
    if (at_star) fire += 1ull;
    bool neg = false;
#pragma unroll
    for (int s = 0; s < MS; ++s) {
        if (s >= n.S) break;
        const uint64_t at = ri.x_off + (uint64_t)s * ri.N + v;
        x[s] = X[at];
        const unsigned long long d = a.gdelta[at];
        if (d != 0ull) a.gdelta[at] = 0ull;
        long long y = (long long)x[s] + (long long)d;
        if (at_star) {
            if (sv.ev_kind == 1) y += n.nu[s * n.R + sv.ev_idx];
            else if (s == sv.ev_idx) y -= 1;
        }
        if (at_dst && s == sv.ev_idx) y += 1;
        neg = neg || y < 0 || y > 0xFFFFFFFFll;
        Xn[at] = (uint32_t)y;
        nx[s] = y;
    }
    bad = neg ? 1u : 0u;
    if (judge) {
        const int c = n.C > 1 ? a.comp[ri.comp_off + v] : 0;
        double av[MR];
        uint32_t crit[(MR + 31) / 32];
        react_props<MR, MS>(n, x, c, a.n_critical, av, crit);
        viol = leap_condition_broken<MS, MR>(n, c, a.n_critical, a.epsilon, a.theta, x, nx, av, crit);
    }
}

__device__ __forceinline__ void reduce_partB(unsigned long long& fire, uint32_t& bad, uint32_t& viol) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        fire += __shfl_xor(fire, off);
        bad |= (uint32_t)__shfl_xor((int)bad, off);
        viol += (uint32_t)__shfl_xor((int)viol, off);
    }
}

// Commits the step or leaps back, on c: returns true when the step was
// accepted and the realization runs on (its next step is then to be
// decided).  lead: the one thread that counts a finished realization in.
__device__ bool commit_step(const KernelArgs& a, const RealInfo& ri, unsigned long long fire, uint32_t bad,
                            uint32_t viol, RealCtl& c, bool lead) {
    // A learned proposal is judged by Cao's leap condition after the fact: it
    // is kept only without negative populations and with no more violating
    // subvolumes than the plain method produces -- mean mu = viol_rate_max N,
    // with a Poisson upper bound mu + 3 sqrt(mu) + 2 so that small
    // realizations are not rejected for single violations; otherwise the
    // step is redone with Cao's tau and fresh draws.
    const double mu = a.viol_rate_max * (double)ri.N;
    const bool learned_reject = c.learned && (bad || (double)viol > mu + 3.0 * sqrt(mu) + 2.0);
    if (bad || learned_reject) {
        c.attempt += 1u;
        if (learned_reject) {
            c.learned_rejects += 1;
            c.fallback = 1u;
            c.retry = 0u;
        } else {
            c.leapbacks += 1;
        }
        if (c.attempt > kMaxAttempt) {
            c.status = kErrLeapback;
            c.done = 1u;
            if (lead) atomicAdd(a.ndone, 1u);
            return false;
        }
        if (!learned_reject) {
            c.tau_p = c.tau_p * 0.5;
            decide(c);
            c.retry = 1u;
        }
        c.prop_ok = 0u;
        return false;
    }
    // the leap pass of a learned context proposed the next step's tau from
    // the state it produced (edge-parallel pass only)
    c.prop_ok = a.props_in_finish ? 1u : 0u;
    c.cur ^= 1u;
    c.t = c.last ? c.t_end : c.t + c.tau;
    c.step += 1u;
    c.steps += 1;
    c.firings += fire;
    c.subvolume_leaps += ri.N;
    c.violations += viol;
    if (c.crit) c.crit_events += 1;
    if (c.learned) c.learned_steps += 1;
    c.fallback = 0u;
    c.attempt = 0u;
    c.retry = 0u;
    if (c.t >= c.t_end || c.steps >= c.step_limit || c.t >= c.t_pause) {
        c.done = 1u;
        if (c.t >= c.t_end || c.steps >= c.step_limit) c.status = kDone;
        if (lead) atomicAdd(a.ndone, 1u);
        return false;
    }
    return true;
}

__device__ void finalize_B(const KernelArgs& a, const RealInfo& ri, uint32_t r, unsigned long long fire,
                           uint32_t bad, uint32_t viol, RealCtl c) {
    commit_step(a, ri, fire, bad, viol, c, true);
    a.ctl[r] = c;
}

__device__ __forceinline__ StepView step_view(const RealInfo& ri, cptr<RealCtl> cp) {
    StepView sv;
    sv.tau = cp->tau;
    sv.step = wave_uniform(cp->step);
    sv.attempt = wave_uniform(cp->attempt);
    sv.real = ri.id;
    sv.crit = (int)wave_uniform((uint32_t)cp->crit);
    sv.ev_kind = (int)wave_uniform((uint32_t)cp->ev_kind);
    sv.ev_idx = (int)wave_uniform((uint32_t)cp->ev_idx);
    sv.vstar = wave_uniform(cp->vstar);
    sv.ev_dst = wave_uniform(cp->ev_dst);
    return sv;
}

// Phase B of one tile.  EP: the draws of an edge-parallel tile into the
// global balances (REDO: the open draws, complete sampler); the tile's
// firings go to its partial and phase_b_finish_kernel completes the leap.
// Otherwise one lane walks one subvolume's out- and in-edges and finishes it.
template <int MS, int MR, bool EP, bool REDO>
__device__ __forceinline__ void phase_b_tile(const KernelArgs& a, const NetView& n, uint32_t t,
                                             uint32_t (*s_x)[kTile], unsigned long long (*s_del)[kTile],
                                             int32_t* s_c, uint2* s_q) {
    const uint32_t lane = threadIdx.x & 63u;
    const TileEnt te = ld_ro(a.tiles + t);
    const uint32_t r = wave_uniform(te.r);
    const RealInfo ri = ld_ro(a.info + r);
    const cptr<RealCtl> cp = ro(a.ctl + r);
    if (wave_uniform(cp->done)) return;
    const StepView sv = step_view(ri, cp);
    const uint32_t cur = wave_uniform(cp->cur);
    const uint32_t* X = cur ? a.x1 : a.x0;
    uint32_t* Xn = cur ? a.x0 : a.x1;
    unsigned long long fire = 0;
    uint32_t bad = 0, viol = 0;
    if constexpr (EP) {
        const bool complete = leap_tile_fwd<MS, MR, REDO>(a, n, ri, X, te.v0, sv, s_x, s_del, s_c, s_q, fire);
        reduce_partB(fire, bad, viol);
        if (lane == 0) {
            if (!complete) {
                const uint32_t slot = atomicAdd(a.redo_count, 1u);
                if (slot < a.ntiles) a.redo_tiles[slot] = t;  // one entry per tile per pass
            }
            // the redo pass adds the firings of the open draws
            a.partB_tile[t] = PartB{(REDO ? a.partB_tile[t].fire : 0ull) + fire, 0u, 0u};
        }
        return;
    } else {
        const uint32_t v = te.v0 + lane;
        if (v < ri.N) bad = leap_subvolume<MS, MR>(a, n, ri, X, Xn, v, sv, fire, viol) ? 1u : 0u;
    }
    reduce_partB(fire, bad, viol);
    if (lane == 0) a.partB_tile[t] = PartB{fire, bad, viol};
}

// The one-compartment edge-parallel leap of up to 5 species: asking for 6
// waves per SIMD holds it to 80 VGPRs (16 bytes of spill per lane), which
// measures faster than 85 VGPRs at 5 waves (the walk waits on gathers and
// LDS, more waves hide it).
template <int MS, int MR, bool EP, bool MC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(EP && !MC && MS <= 5 ? 6 : 1)))
void phase_b_kernel(KernelArgs a) {
    const uint32_t wv = wave_uniform(threadIdx.x >> 6);
    // LDS staging for the edge-parallel path (up to 4 waves per block)
    constexpr int LS = EP ? MS : 1;
    __shared__ uint32_t s_x[4][LS][kTile];
    __shared__ unsigned long long s_del[4][LS][kTile];
    __shared__ int32_t s_c[4][kTile];
    __shared__ uint2 s_q[4][EP ? kQueue : 1];
    for (uint32_t it = 0;; ++it) {
        const uint32_t t = tile_at(it, a.ntiles, a.waves_per_block);
        if (t == 0xFFFFFFFFu) break;
        KernelArgs al = a;  // network tables through per-tile opaque pointers, as in phase_a_kernel
        asm volatile("" : "+s"(al.net_i), "+s"(al.net_d));
        const NetView n = net_view<MC>(al);
        phase_b_tile<MS, MR, EP, false>(a, n, t, s_x[wv], s_del[wv], s_c[wv], s_q[wv]);
    }
}

// Tiles phase_b_kernel left on the redo list: their open draws, drawn with
// the complete sampler.  The list is emptied by decide_kernel before each leap pass.
template <int MS, int MR, bool MC>
__global__ __launch_bounds__(256) void phase_b_redo_kernel(KernelArgs a) {
    const NetView n = net_view<MC>(a);
    const uint32_t wv = wave_uniform(threadIdx.x >> 6);
    const uint32_t wpb = blockDim.x >> 6;
    __shared__ uint32_t s_x[4][MS][kTile];
    __shared__ unsigned long long s_del[4][MS][kTile];
    __shared__ int32_t s_c[4][kTile];
    const uint32_t cnt = min(wave_uniform(*a.redo_count), a.ntiles);
    if (blockIdx.x == 0 && threadIdx.x == 0 && cnt != 0u) *a.redo_total += cnt;
    for (uint32_t q = blockIdx.x * wpb + wv; q < cnt; q += gridDim.x * wpb)
        phase_b_tile<MS, MR, true, true>(a, n, wave_uniform(a.redo_tiles[q]), s_x[wv], s_del[wv], s_c[wv],
                                         nullptr);
}

// The end of the edge-parallel leap pass: every tile adds the global
// balances to its counts (see finish_tile) after all draws are in.  Up to 5
// species it fits 8 waves per SIMD (otherwise its scalar registers hold it
// at 7): more loads in flight for an HBM-bound kernel, 20.9 -> 20.5 us on
// the Brusselator mesh.
template <int MS, int MR, bool MC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MS <= 5 ? 8 : 1)))
void phase_b_finish_kernel(KernelArgs a) {
    const NetView n = net_view<MC>(a);
    const uint32_t lane = threadIdx.x & 63u;
    for (uint32_t it = 0;; ++it) {
        const uint32_t t = tile_at(it, a.ntiles, a.waves_per_block);
        if (t == 0xFFFFFFFFu) break;
        const TileEnt te = ld_ro(a.tiles + t);
        const uint32_t r = wave_uniform(te.r);
        const RealInfo ri = ld_ro(a.info + r);
        const cptr<RealCtl> cp = ro(a.ctl + r);
        if (wave_uniform(cp->done)) continue;
        const StepView sv = step_view(ri, cp);
        const uint32_t cur = wave_uniform(cp->cur);
        unsigned long long fire = 0;
        uint32_t bad = 0, viol = 0;
        uint32_t x[MS];
        long long nx[MS];
        finish_tile<MS, MR>(a, n, ri, cur ? a.x1 : a.x0, cur ? a.x0 : a.x1, te.v0, sv, fire, bad, viol, x, nx,
                            a.judge_on != 0);
        reduce_partB(fire, bad, viol);
        if (lane == 0) a.partB_tile[t] = PartB{a.partB_tile[t].fire + fire, bad, viol};
    }
}

// One block per realization: sum of firings / any negative over the tile
// partials of phase B, then commit the step or leap back.
__global__ __launch_bounds__(1024) void commit_kernel(KernelArgs a) {
    const uint32_t r = blockIdx.x;
    if (r >= a.nreal) return;
    const RealInfo ri = ld_ro(a.info + r);
    const RealCtl c0 = ld_ro(a.ctl + r);
    if (c0.done) return;
    __shared__ unsigned long long s_f[16];
    __shared__ uint32_t s_b[16], s_v[16];
    unsigned long long fire = 0;
    uint32_t bad = 0, viol = 0;
    for (uint32_t q0 = 0; q0 < ri.ntiles; q0 += 8u * blockDim.x) {
        PartB p[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const uint32_t q = q0 + (uint32_t)i * blockDim.x + threadIdx.x;
            p[i] = q < ri.ntiles ? a.partB_tile[ri.tile_begin + q] : PartB{0ull, 0u, 0u};
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            fire += p[i].fire;
            bad |= p[i].bad;
            viol += p[i].viol;
        }
    }
    reduce_partB(fire, bad, viol);
    const uint32_t wv = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    if (lane == 0) {
        s_f[wv] = fire;
        s_b[wv] = bad;
        s_v[wv] = viol;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t w = 1; w < (blockDim.x >> 6); ++w) {
            fire += s_f[w];
            bad |= s_b[w];
            viol += s_v[w];
        }
        finalize_B(a, ri, r, fire, bad, viol, c0);
    }
}

// The folded order of a plain-tau iteration: commits the step or leaps
// back from the leap pass's tile partials and, when the step is accepted
// and the realization runs on, decides the next step from the tile
// partials of the tau pass that ran on the new counts (phase_a_kernel with
// tau_spec) -- commit_kernel and decide_kernel in one launch.  Blocks and
// chunks as in decide_kernel (decide_chunks blocks per realization, the
// last one finishes; counter shared with decide_kernel, which never runs
// at the same time).  A leap-back drops the tau pass's partials (the
// retried leap reuses the halved tau).
template <int MS, int MR>
__global__ __launch_bounds__(decide_threads<MS>()) void commit_decide_kernel(KernelArgs a) {
    const uint32_t nch = a.decide_chunks;
    const uint32_t r = blockIdx.x / nch, chunk = blockIdx.x % nch;
    if (blockIdx.x == 0u && threadIdx.x == 0u) *a.redo_count = 0u;  // the coming leap pass's redo list
    if (r >= a.nreal) return;
    const RealInfo ri = ld_ro(a.info + r);
    const RealCtl c0 = ld_ro(a.ctl + r);
    __shared__ unsigned long long s_f[16];
    __shared__ uint32_t s_b[16], s_v[16];
    __shared__ double s_tp[16], s_t2[16];
    __shared__ uint32_t s_v2[16];
    __shared__ uint32_t s_last;
    unsigned long long fire = 0ull;
    uint32_t bad = 0u, viol = 0u;
    double tp = kInf, t2 = kInf;
    uint32_t v2 = 0xFFFFFFFFu;
    const uint32_t per = (ri.ntiles + nch - 1u) / nch;
    const uint32_t qb = min(chunk * per, ri.ntiles), qe = min(qb + per, ri.ntiles);
    const PartB* pb = a.partB_tile + ri.tile_begin;
    const PartA* pa = a.partA_tile + ri.tile_begin;
    // both partials' loads in flight together
    for (uint32_t q0 = qb; q0 < qe; q0 += 8u * blockDim.x) {
        PartB b[8];
        PartA p[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const uint32_t q = q0 + (uint32_t)i * blockDim.x + threadIdx.x;
            b[i] = q < qe ? pb[q] : PartB{0ull, 0u, 0u};
            p[i] = q < qe ? pa[q] : PartA{kInf, kInf, 0xFFFFFFFFu, 0u};
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            fire += b[i].fire;
            bad |= b[i].bad;
            viol += b[i].viol;
            merge_partA(tp, t2, v2, p[i].tp, p[i].t2, p[i].v2);
        }
    }
    if (c0.done) return;  // (the same in every block of the realization)
    reduce_partB(fire, bad, viol);
    const uint32_t wv = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    if (lane == 0) {
        s_f[wv] = fire;
        s_b[wv] = bad;
        s_v[wv] = viol;
    }
    reduce_block_partA<MS>(tp, t2, v2, s_tp, s_t2, s_v2);  // (its barrier publishes s_f / s_b / s_v)
    if (wv == 0)  // the first wavefront, every lane alike
        for (uint32_t w = 1; w < (blockDim.x >> 6); ++w) {
            fire += s_f[w];
            bad |= s_b[w];
            viol += s_v[w];
        }
    if (nch > 1u) {
        if (threadIdx.x == 0) {
            a.partA_chunk[r * nch + chunk] = PartA{tp, t2, v2, 0u};
            a.partB_chunk[r * nch + chunk] = PartB{fire, bad, viol};
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            s_last = atomicAdd(&a.decide_cnt[r], 1u) == nch - 1u ? 1u : 0u;
        }
        __syncthreads();
        if (!s_last || wv != 0) return;
        // the last block of the realization: its first wavefront finishes
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        PartA p{kInf, kInf, 0xFFFFFFFFu, 0u};
        PartB b{0ull, 0u, 0u};
        if (lane < nch) {
            p = a.partA_chunk[r * nch + lane];
            b = a.partB_chunk[r * nch + lane];
        }
        tp = p.tp;
        t2 = p.t2;
        v2 = p.v2;
        fire = b.fire;
        bad = b.bad;
        viol = b.viol;
        reduce_partA(tp, t2, v2);
        reduce_partB(fire, bad, viol);
        if (lane == 0) a.decide_cnt[r] = 0u;
    } else if (wv != 0) {
        return;
    }
    RealCtl c = c0;
    if (!commit_step(a, ri, fire, bad, viol, c, lane == 0u)) {
        if (lane == 0u) a.ctl[r] = c;
        return;
    }
    finalize_A<MS, MR>(a, net_view(a), ri, r, tp, t2, v2, c);
}

// Explicit instantiations, smallest fitting one is used: Schlogl-like
// (1 species, <= 4 reactions), <= 3 x 4, <= 5 x 8 (Brusselator), <= 8 x 16
// and the general case.
#define PLIS_INST(MS, MR)                                                   \
    template __global__ void phase_a_kernel<MS, MR, false>(KernelArgs);     \
    template __global__ void phase_a_kernel<MS, MR, true>(KernelArgs);      \
    template __global__ void decide_kernel<MS, MR>(KernelArgs);             \
    template __global__ void commit_decide_kernel<MS, MR>(KernelArgs);      \
    template __global__ void phase_b_kernel<MS, MR, false, false>(KernelArgs); \
    template __global__ void phase_b_kernel<MS, MR, false, true>(KernelArgs);
#define PLIS_INST_EP(MS, MR)                                                \
    template __global__ void phase_b_kernel<MS, MR, true, false>(KernelArgs); \
    template __global__ void phase_b_kernel<MS, MR, true, true>(KernelArgs);  \
    template __global__ void phase_b_redo_kernel<MS, MR, false>(KernelArgs);  \
    template __global__ void phase_b_redo_kernel<MS, MR, true>(KernelArgs);   \
    template __global__ void phase_b_finish_kernel<MS, MR, false>(KernelArgs); \
    template __global__ void phase_b_finish_kernel<MS, MR, true>(KernelArgs);
PLIS_INST(1, 4)
PLIS_INST(3, 4)
PLIS_INST(5, 8)
PLIS_INST(8, 16)
PLIS_INST(64, 128)
PLIS_INST_EP(1, 4)
PLIS_INST_EP(3, 4)
PLIS_INST_EP(5, 8)
PLIS_INST_EP(8, 16)

}  // namespace plis
