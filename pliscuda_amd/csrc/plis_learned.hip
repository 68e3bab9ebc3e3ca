// plis_learned.hip -- learned tau pass and training records.
//
// learned_tau_kernel replaces the Cao part of the tau pass: per subvolume it
// builds 12 local features (no neighbour gathers), runs the MLP on the
// wavefront's 64-subvolume tile with bf16 MFMA, and proposes
// tau_v = kappa * exp(log tau prediction).  The critical part stays exact
// (critical propensity and the tau'' clock, as in the plain pass), so the
// critical event is drawn exactly as in the plain method.  The leap pass
// then judges the proposal (finalize_B).
//
// records_kernel writes (features, log Cao tau', critical flag) of every
// subvolume of one realization: the training records.
#pragma once
#include "plis_kernels.hip"
#include "plis_mlp.hip"

namespace plis {

// Features of the lane's subvolume (the host restatement computes the same
// quantities in double; single-precision hardware logs are enough here):
//   f[s] = log(1 + x_s), f[5+s] = log(1 + sum_{j: r_sj>0} a_j + d_s x_s W),
//   f[10] = highest order among reactions with a_j > 0, f[11] = log prev tau.
template <int MS, int MR>
__device__ __forceinline__ void features_and_crit(const KernelArgs& a, const NetView& n, const RealInfo& ri,
                                                  const uint32_t* X, uint32_t v, float prev_logtau, float (&f)[12],
                                                  double& a0c_out) {
    const uint32_t N = ri.N;
    uint32_t x[MS];
#pragma unroll
    for (int s = 0; s < MS; ++s) x[s] = s < n.S ? X[ri.x_off + (uint64_t)s * N + v] : 0u;
    const int c = n.C > 1 ? a.comp[ri.comp_off + v] : 0;
    double av[MR];
    uint32_t crit[(MR + 31) / 32];
    react_props<MR, MS>(n, x, c, a.n_critical, av, crit);
    const double W = a.wsum[ri.rp_off + v];
    double A[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    int maxo = 0;
    double a0c = 0.0;
#pragma unroll
    for (int j = 0; j < MR; ++j) {
        if (j >= n.R) break;
        if (av[j] > 0.0 && n.order[j] > maxo) maxo = n.order[j];
        if (is_crit<MR>(crit, j)) a0c += av[j];
        const int nr = n.nrs[j];
        for (int i = 0; i < nr; ++i) {
            const int s = n.rs[j * 3 + i];
#pragma unroll
            for (int q = 0; q < 5; ++q)
                if (q == s) A[q] += av[j];
        }
    }
#pragma unroll
    for (int s = 0; s < MS; ++s) {
        if (s >= n.S) break;
        const double d = n.d[c * n.S + s];
        if (d > 0.0 && !hop_active(d, x[s], a.n_critical)) a0c += (d * (double)x[s]) * W;
    }
#pragma unroll
    for (int s = 0; s < 5; ++s) {
        if (s < n.S && s < MS) {
            const double d = n.d[c * n.S + s];
            const double As = A[s] + (d * (double)x[s < MS ? s : 0]) * W;
            f[s] = __logf(1.0f + (float)x[s < MS ? s : 0]);
            f[5 + s] = __logf(1.0f + (float)As);
        } else {
            f[s] = 0.0f;
            f[5 + s] = 0.0f;
        }
    }
    f[10] = (float)maxo;
    f[11] = prev_logtau;
    a0c_out = a0c;
}

template <int MS, int MR>
__global__ __launch_bounds__(256) void learned_tau_kernel(KernelArgs a) {
    __shared__ MlpLdsInfer s_mlp[4];
    const NetView n = net_view(a);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wv = wave_uniform(threadIdx.x >> 6);
    MlpLdsInfer& S = s_mlp[wv];
    const MlpDev m{a.mlp_w32, (const __bf16*)a.mlp_q, a.mlp_norm};
    const float y_mu = a.mlp_norm[24], y_sig = a.mlp_norm[25];
    for (uint32_t it = 0;; ++it) {
        const uint32_t t = tile_at(it, a.ntiles, a.waves_per_block);
        if (t == 0xFFFFFFFFu) break;
        const TileEnt te = ld_ro(a.tiles + t);
        const uint32_t r = wave_uniform(te.r);
        const RealInfo ri = ld_ro(a.info + r);
        const cptr<RealCtl> cp = ro(a.ctl + r);
        if (wave_uniform(cp->done) || wave_uniform(cp->retry) || wave_uniform(cp->fallback)) continue;
        const uint32_t* X = wave_uniform(cp->cur) ? a.x1 : a.x0;
        const double prev = cp->tau;
        const float prev_logtau = prev > 0.0 ? __logf((float)prev) : 0.0f;
        const uint32_t v = te.v0 + lane;
        const bool act = v < ri.N;
        float f[12];
        double a0c = 0.0;
        if (act) {
            features_and_crit<MS, MR>(a, n, ri, X, v, prev_logtau, f, a0c);
        } else {
#pragma unroll
            for (int k = 0; k < 12; ++k) f[k] = 0.0f;
        }
        stage_inputs(S.h, nullptr, m, f, act, kLdH);
        float o0, o1;
        mlp_infer_tile(S, m, o0, o1);
        double tp = kInf, t2 = kInf;
        uint32_t v2 = v;
        if (act) {
            tp = a.kappa * dexp((double)(o0 * y_sig + y_mu));
            if (a0c > 0.0) {
                const uint32_t step = wave_uniform(cp->step);
                const U4 o = philox10(v, kSelfId, step, ctr_word3(ri.id, 0u, kTau2, 0u), a.seed_lo, a.seed_hi);
                t2 = -dlog(uniform01(o.x)) / a0c;
            }
        }
        reduce_partA(tp, t2, v2);
        if (lane == 0) a.partA_tile[t] = PartA{tp, t2, v2, 0u};
        wave_sync();
    }
}

template <int MS, int MR>
__global__ __launch_bounds__(256) void records_kernel(KernelArgs a) {
    const NetView n = net_view(a);
    const uint32_t lane = threadIdx.x & 63u;
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
        const uint32_t v = te.v0 + lane;
        if (v >= ri.N) continue;
        float f[12];
        double a0c, tp, a0c2;
        features_and_crit<MS, MR>(a, n, ri, X, v, prev_logtau, f, a0c);
        tau_subvolume<MS, MR>(a, n, ri, X, v, tp, a0c2);
#pragma unroll
        for (int k = 0; k < 12; ++k) a.rec_feats[(size_t)v * 12 + k] = f[k];
        a.rec_ylog[v] = tp < 1e6 ? (float)dlog(tp) : (float)dlog(1e6);
        a.rec_crit[v] = a0c > 0.0 ? 1.0f : 0.0f;
    }
}

#define PLIS_INST_L(MS, MR)                                                  \
    template __global__ void learned_tau_kernel<MS, MR>(KernelArgs);        \
    template __global__ void records_kernel<MS, MR>(KernelArgs);
PLIS_INST_L(1, 4)
PLIS_INST_L(3, 4)
PLIS_INST_L(5, 8)
PLIS_INST_L(8, 16)

}  // namespace plis
