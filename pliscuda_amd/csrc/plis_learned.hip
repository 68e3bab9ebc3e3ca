// plis_learned.hip -- learned tau proposal: fused into the leap pass,
// standalone pass, and training records.
//
// The network replaces Cao's tau' selection: per subvolume it maps 12 local
// features (no neighbour gathers) to a log-tau proposal and a critical
// logit, on the wavefront's 64-subvolume tile with bf16 MFMA.  The
// proposal is tau_v = kappa exp(log tau), halved where the critical logit
// flags the subvolume (the network expects a leap at its proposal to break
// the leap condition there); the step takes tau = min_v tau_v.  The
// critical part stays exact (critical propensity, the tau'' clock and the
// critical event, as in the plain pass).
//
// learned_finish_kernel is the end of the edge-parallel leap pass with the
// proposal fused in: once a tile's new counts are formed it computes their
// features, runs the network and writes the tile partial the next step's
// decide kernel reads, so a learned step has no tau pass at all.  The
// partials are valid only if the step is accepted (commit sets prop_ok);
// after a rejection, a leap-back or a change of state made by the host,
// learned_tau_kernel recomputes them from the current state.  The leap pass
// then judges the step (finalize_B).
//
// records_kernel writes training records of one realization's state:
// features, log Cao tau'_v and the accept/reject label of a trial leap of v
// at tau'_v (the post-leap leap condition at v).
#pragma once
#include "plis_kernels.hip"
#include "plis_mlp.hip"

namespace plis {

// Philox purpose of the records' trial leaps (the step's purposes are 1..5).
constexpr uint32_t kTrial = 7u;

// Whether species s (a compile-time index) is a reactant of reaction j:
// scalar tests on the network tables.
__device__ __forceinline__ bool is_reactant(const NetView& n, int j, int s) {
    const int nr = n.nrs[j];
    return (nr > 0 && n.rs[j * 3] == s) || (nr > 1 && n.rs[j * 3 + 1] == s) || (nr > 2 && n.rs[j * 3 + 2] == s);
}

// Features of one subvolume from its counts x and their propensities av /
// critical flags crit (react_props; the host restatement computes the same
// quantities in double, single-precision sums and hardware logs are enough
// for a bf16 network):
//   f[s] = log(1 + x_s), f[5+s] = log(1 + sum_{j: r_sj>0} a_j + d_s x_s W),
//   f[10] = highest order among reactions with a_j > 0, f[11] = log prev tau.
// a0c_out: the exact critical propensity (FP64, as the plain tau pass);
// live: the subvolume has a non-critical reaction with a_j > 0 or an active
// hop -- without one its proposal is infinite, so a realization with no
// possible event stops as in the plain method.
template <int MS, int MR>
__device__ __forceinline__ void features_and_crit(const KernelArgs& a, const NetView& n, const RealInfo& ri,
                                                  const uint32_t (&x)[MS], const double (&av)[MR],
                                                  const uint32_t (&crit)[(MR + 31) / 32], uint32_t v,
                                                  float prev_logtau, float (&f)[12], double& a0c_out, bool& live) {
    const int c = n.C > 1 ? a.comp[ri.comp_off + v] : 0;
    const double W = a.wsum[ri.rp_off + v];
    float A[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    int maxo = 0;
    double a0c = 0.0;
    bool lv = false;
#pragma unroll
    for (int j = 0; j < MR; ++j) {
        if (j >= n.R) break;
        if (av[j] > 0.0 && n.order[j] > maxo) maxo = n.order[j];
        if (is_crit<MR>(crit, j)) a0c += av[j];
        else if (av[j] > 0.0) lv = true;
        const float aj = (float)av[j];
#pragma unroll
        for (int q = 0; q < 5; ++q)
            if (q < MS && is_reactant(n, j, q)) A[q] += aj;
    }
    const float Wf = (float)W;
#pragma unroll
    for (int s = 0; s < MS; ++s) {
        if (s >= n.S) break;
        const double d = n.d[c * n.S + s];
        if (d > 0.0 && !hop_active(d, x[s], a.n_critical)) a0c += (d * (double)x[s]) * W;
        if (hop_active(d, x[s], a.n_critical)) lv = true;
    }
#pragma unroll
    for (int s = 0; s < 5; ++s) {
        if (s < n.S && s < MS) {
            const int sq = s < MS ? s : 0;
            const float xf = (float)x[sq];
            const float As = A[s] + ((float)n.d[c * n.S + s] * xf) * Wf;
            f[s] = __logf(1.0f + xf);
            f[5 + s] = __logf(1.0f + As);
        } else {
            f[s] = 0.0f;
            f[5 + s] = 0.0f;
        }
    }
    f[10] = (float)maxo;
    f[11] = prev_logtau;
    a0c_out = a0c;
    live = lv;
}

// The proposals of one tile: features, the network on the tile, tau_v, the
// exact tau'' clock of the given step, then min / argmin into the tile
// partial.  Every lane of the wavefront takes part (the network runs on the
// whole tile); act = the lane holds a subvolume; av / crit: react_props of
// the lane's counts x.  W: the staged weights.
template <int MS, int MR, int ACT>
__device__ __forceinline__ void propose_tile(const KernelArgs& a, const NetView& n, const RealInfo& ri,
                                             const MlpLdsW& W, const uint32_t (&x)[MS], const double (&av)[MR],
                                             const uint32_t (&crit)[(MR + 31) / 32], bool act, uint32_t v,
                                             uint32_t step, double prev_tau, uint32_t t) {
    const MlpDev m{a.mlp_w32, (const __bf16*)a.mlp_q, a.mlp_norm};
    const float prev_logtau = prev_tau > 0.0 ? __logf((float)prev_tau) : 0.0f;
    float f[12];
    double a0c = 0.0;
    bool live = false;
    if (act) {
        features_and_crit<MS, MR>(a, n, ri, x, av, crit, v, prev_logtau, f, a0c, live);
    } else {
#pragma unroll
        for (int k = 0; k < 12; ++k) f[k] = 0.0f;
    }
    float o0, o1;
    mlp_infer_regs<ACT>(W, m, f, act, o0, o1);
    double tp = kInf, t2 = kInf;
    uint32_t v2 = v;
    if (act) {
        if (live) {
            const float ly = o0 * a.mlp_norm[25] + a.mlp_norm[24];
            tp = a.kappa * (double)__expf(ly);
            if (o1 > 0.0f) tp = tp * 0.5;
        }
        if (a0c > 0.0) {
            const U4 o = philox10(v, kSelfId, step, ctr_word3(ri.id, 0u, kTau2, 0u), a.seed_lo, a.seed_hi);
            t2 = -dlog(uniform01(o.x)) / a0c;
        }
    }
    reduce_partA(tp, t2, v2);
    if ((threadIdx.x & 63u) == 0u) a.partA_tile[t] = PartA{tp, t2, v2, 0u};
}

// The tau pass of a learned context, per tile: Cao's formula for a
// realization whose last proposal was rejected (fallback), the network on
// the current state where the partials do not hold its proposals (first
// step, after a leap-back, after the host changed the state or the
// network), nothing where the previous leap pass already proposed.  One
// launch covers all three; in steady state every tile is skipped, and a
// block with nothing to do leaves before staging the network.
template <int MS, int MR, bool MC, int ACT>
__global__ __launch_bounds__(256) void learned_tau_kernel(KernelArgs a) {
    __shared__ MlpLdsW W;
    __shared__ uint32_t s_need;
    const uint32_t lane = threadIdx.x & 63u;
    if (threadIdx.x == 0) s_need = 0u;
    __syncthreads();
    for (uint32_t it = 0;; ++it) {
        const uint32_t t = tile_at(it, a.ntiles, a.waves_per_block);
        if (t == 0xFFFFFFFFu) break;
        const cptr<RealCtl> cp = ro(a.ctl + wave_uniform(ld_ro(a.tiles + t).r));
        if (!wave_uniform(cp->done) && !wave_uniform(cp->retry) &&
            (wave_uniform(cp->fallback) || !wave_uniform(cp->prop_ok))) {
            if (lane == 0) atomicOr(&s_need, 1u);
            break;
        }
    }
    __syncthreads();
    if (!s_need) return;
    stage_weights(W, MlpDev{a.mlp_w32, (const __bf16*)a.mlp_q, a.mlp_norm});
    __syncthreads();
    const NetView n = net_view<MC>(a);
    for (uint32_t it = 0;; ++it) {
        const uint32_t t = tile_at(it, a.ntiles, a.waves_per_block);
        if (t == 0xFFFFFFFFu) break;
        const TileEnt te = ld_ro(a.tiles + t);
        const uint32_t r = wave_uniform(te.r);
        const RealInfo ri = ld_ro(a.info + r);
        const cptr<RealCtl> cp = ro(a.ctl + r);
        if (wave_uniform(cp->done) || wave_uniform(cp->retry)) continue;
        if (wave_uniform(cp->fallback)) {
            cao_tau_tile<MS, MR>(a, n, ri, cp, te, t, false);
            continue;
        }
        if (wave_uniform(cp->prop_ok)) continue;
        const uint32_t* X = wave_uniform(cp->cur) ? a.x1 : a.x0;
        const uint32_t v = te.v0 + lane;
        const bool act = v < ri.N;
        uint32_t x[MS];
#pragma unroll
        for (int s = 0; s < MS; ++s) x[s] = (act && s < n.S) ? X[ri.x_off + (uint64_t)s * ri.N + v] : 0u;
        double av[MR];
        uint32_t crit[(MR + 31) / 32];
        react_props<MR, MS>(n, x, n.C > 1 && act ? a.comp[ri.comp_off + v] : 0, a.n_critical, av, crit);
        propose_tile<MS, MR, ACT>(a, n, ri, W, x, av, crit, act, v, wave_uniform(cp->step), cp->tau, t);
    }
}

// End of the edge-parallel leap pass with the next step's proposals fused
// in (learned contexts): finish_tile, the post-leap leap condition and the
// network's proposals for step + 1 on the new counts, from one evaluation
// of their propensities.  Proposals of a realization whose step is rejected
// are never read (commit clears prop_ok).
template <int MS, int MR, bool MC, int ACT>
__global__ __launch_bounds__(256) void learned_finish_kernel(KernelArgs a) {
    __shared__ MlpLdsW W;
    stage_weights(W, MlpDev{a.mlp_w32, (const __bf16*)a.mlp_q, a.mlp_norm});
    __syncthreads();
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
        finish_tile<MS, MR>(a, n, ri, cur ? a.x1 : a.x0, cur ? a.x0 : a.x1, te.v0, sv, fire, bad, viol, x, nx, false);
        const uint32_t v = te.v0 + lane;
        const bool act = v < ri.N;
        const int c = n.C > 1 && act ? a.comp[ri.comp_off + v] : 0;
        uint32_t xn[MS];
#pragma unroll
        for (int s = 0; s < MS; ++s) xn[s] = (uint32_t)(nx[s] < 0 ? 0 : nx[s]);
        double av1[MR];
        uint32_t crit1[(MR + 31) / 32];
        react_props<MR, MS>(n, xn, c, a.n_critical, av1, crit1);
        if (act) {
            double av0[MR];
            uint32_t crit0[(MR + 31) / 32];
            react_props<MR, MS>(n, x, c, a.n_critical, av0, crit0);
            viol = leap_condition_broken<MS, MR>(n, c, a.n_critical, a.epsilon, a.theta, x, nx, av0, crit0, av1);
        }
        reduce_partB(fire, bad, viol);
        if (lane == 0) a.partB_tile[t] = PartB{a.partB_tile[t].fire + fire, bad, viol};
        propose_tile<MS, MR, ACT>(a, n, ri, W, xn, av1, crit1, act, v, sv.step + 1u, sv.tau, t);
    }
}

// Training records of realization rec_real's current state: features,
// log Cao tau'_v (capped at 1e6) and the reject label of a trial leap of v
// alone at tau'_v (non-critical reactions, out-hops and in-hops at the
// neighbours' propensities, drawn from the TRIAL streams with the complete
// sampler: counter (v, kSelfId or u, step, purpose kTrial), block 96 + j/4
// for reactions and 64 + s/2 for hops, word 2 (s & 1) + 0 out / 1 in);
// label 1 where the leap condition breaks at v.
template <int MS, int MR>
__global__ __launch_bounds__(256) void records_kernel(KernelArgs a) {
    const NetView n = net_view(a);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t k0 = a.seed_lo, k1 = a.seed_hi;
    for (uint32_t it = 0;; ++it) {
        const uint32_t t = tile_at(it, a.ntiles, a.waves_per_block);
        if (t == 0xFFFFFFFFu) break;
        const TileEnt te = a.tiles[t];
        if (te.r != a.rec_real) continue;
        const RealInfo ri = a.info[te.r];
        const RealCtl* cp = &a.ctl[te.r];
        const uint32_t* X = cp->cur ? a.x1 : a.x0;
        const double prev = cp->tau;
        const float prev_logtau = prev > 0.0 ? __logf((float)prev) : 0.0f;
        const uint32_t step = cp->step;
        const uint32_t v = te.v0 + lane;
        if (v >= ri.N) continue;
        const uint32_t N = ri.N;
        uint32_t x[MS];
#pragma unroll
        for (int s = 0; s < MS; ++s) x[s] = s < n.S ? X[ri.x_off + (uint64_t)s * N + v] : 0u;
        float f[12];
        double a0c, tp, a0c2;
        bool live;
        {
            double av[MR];
            uint32_t crit[(MR + 31) / 32];
            react_props<MR, MS>(n, x, n.C > 1 ? a.comp[ri.comp_off + v] : 0, a.n_critical, av, crit);
            features_and_crit<MS, MR>(a, n, ri, x, av, crit, v, prev_logtau, f, a0c, live);
        }
        tau_subvolume<MS, MR>(a, n, ri, X, v, tp, a0c2);
#pragma unroll
        for (int k = 0; k < 12; ++k) a.rec_feats[(size_t)v * 12 + k] = f[k];
        a.rec_ylog[v] = tp < 1e6 ? (float)dlog(tp) : (float)dlog(1e6);
        float label = 0.0f;
        if (tp < kInf) {
            const int c = n.C > 1 ? a.comp[ri.comp_off + v] : 0;
            double av[MR];
            uint32_t crit[(MR + 31) / 32];
            react_props<MR, MS>(n, x, c, a.n_critical, av, crit);
            long long nx[MS];
#pragma unroll
            for (int s = 0; s < MS; ++s) nx[s] = (long long)x[s];
            const uint32_t c3 = ctr_word3(ri.id, 0u, kTrial, 0u);
            for (int j = 0; j < n.R; ++j) {
                if (is_crit<MR>(crit, j) || !(av[j] > 0.0)) continue;
                const uint32_t cb = c3 | (96u + ((uint32_t)j >> 2)), wd = (uint32_t)j & 3u;
                const U4 o = philox10(v, kSelfId, step, cb, k0, k1);
                const long long k = poisson(tp * av[j], pick(o, wd), DrawId{v, kSelfId, step, cb, k0, k1, wd});
#pragma unroll
                for (int s = 0; s < MS; ++s)
                    if (s < n.S) nx[s] += k * (long long)n.nu[s * n.R + j];
            }
            const bool weighted = (ri.flags & 2u) != 0;
            const bool sym = (ri.flags & 1u) != 0;
            const uint32_t rb = a.rp[ri.rp_off + v], re = a.rp[ri.rp_off + v + 1];
#pragma unroll
            for (int s = 0; s < MS; ++s) {
                if (s >= n.S) break;
                const double d = n.d[c * n.S + s];
                if (!hop_active(d, x[s], a.n_critical)) continue;
                const double dx = d * (double)x[s];
                const uint32_t cb = c3 | (64u + ((uint32_t)s >> 1)), wd = 2u * ((uint32_t)s & 1u);
                for (uint32_t q = rb; q < re; ++q) {
                    const uint32_t u = a.col[ri.e_off + q];
                    const double wq = weighted ? (double)a.w[ri.e_off + q] : 1.0;
                    const U4 o = philox10(v, u, step, cb, k0, k1);
                    nx[s] -= poisson(tp * (dx * wq), pick(o, wd), DrawId{v, u, step, cb, k0, k1, wd});
                }
            }
            const uint32_t* icol = sym ? a.col + ri.e_off : a.icol + ri.ie_off;
            const float* iwp = sym ? a.w + ri.e_off : a.iw + ri.ie_off;
            const uint32_t* irp = sym ? a.rp + ri.rp_off : a.irp + ri.irp_off;
            for (uint32_t q = irp[v]; q < irp[v + 1]; ++q) {
                const uint32_t u = icol[q];
                const double wq = weighted ? (double)iwp[q] : 1.0;
                const int cu = n.C > 1 ? a.comp[ri.comp_off + u] : 0;
#pragma unroll
                for (int s = 0; s < MS; ++s) {
                    if (s >= n.S) break;
                    const uint32_t xu = X[ri.x_off + (uint64_t)s * N + u];
                    const double du = n.d[cu * n.S + s];
                    if (!hop_active(du, xu, a.n_critical)) continue;
                    const uint32_t cb = c3 | (64u + ((uint32_t)s >> 1)), wd = 2u * ((uint32_t)s & 1u) + 1u;
                    const U4 o = philox10(v, u, step, cb, k0, k1);
                    nx[s] += poisson(tp * ((du * (double)xu) * wq), pick(o, wd), DrawId{v, u, step, cb, k0, k1, wd});
                }
            }
            label = leap_condition_broken<MS, MR>(n, c, a.n_critical, a.epsilon, a.theta, x, nx, av, crit) ? 1.0f
                                                                                                            : 0.0f;
        }
        a.rec_crit[v] = label;
    }
}

#define PLIS_INST_L(MS, MR)                                                          \
    template __global__ void learned_tau_kernel<MS, MR, false, 0>(KernelArgs);      \
    template __global__ void learned_tau_kernel<MS, MR, false, 1>(KernelArgs);      \
    template __global__ void learned_tau_kernel<MS, MR, true, 0>(KernelArgs);       \
    template __global__ void learned_tau_kernel<MS, MR, true, 1>(KernelArgs);       \
    template __global__ void learned_finish_kernel<MS, MR, false, 0>(KernelArgs);   \
    template __global__ void learned_finish_kernel<MS, MR, false, 1>(KernelArgs);   \
    template __global__ void learned_finish_kernel<MS, MR, true, 0>(KernelArgs);    \
    template __global__ void learned_finish_kernel<MS, MR, true, 1>(KernelArgs);    \
    template __global__ void records_kernel<MS, MR>(KernelArgs);
PLIS_INST_L(1, 4)
PLIS_INST_L(3, 4)
PLIS_INST_L(5, 8)
PLIS_INST_L(8, 16)

}  // namespace plis
