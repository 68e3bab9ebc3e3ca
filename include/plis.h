/*
 * plis.h -- C ABI of libplis_hip.so, the gfx950 RDME tau-leap + SSA step.
 *
 * This is the drop-in boundary for PLISCuda's simulation path:
 *   PLISCuda::run_simulation(Topology t, State s, Reactions r, float* h_rrc,
 *       float* h_drc, float endTime, int compartments_count,
 *       int* subv_constants, float log_freq)
 *   (PLISCuda/include/cuda/driver.cuh:32-33, PLISCuda/src/driver.cu:12-408)
 * whose only caller is PLISCuda/src/main.cpp:53-63.  The C++ driver in
 * pliscuda_amd/csrc/host keeps that signature and calls the entry points
 * below; a caller in another language binds them directly (ctypes, cgo, JNI).
 *
 * Conventions: plain pointers and sizes only; every function returns 0 on
 * success or a negative PLIS_ERR_* code and never exits the process (the
 * reference calls exit() from gpuAssert, PLISCuda/include/cpp/cpp_utils.hpp:97-105);
 * plis_last_error() gives the message.  Host arrays are borrowed for the
 * duration of a call and copied; the library owns all device memory and frees
 * it in plis_destroy() (the reference never frees device memory).  One
 * context = one HIP stream; contexts are independent (no globals such as the
 * reference's __constant__ SBC/SPC/RC/NC/EPSILON, driver.cu:4-8), a context is
 * not thread-safe.
 */
#ifndef PLIS_H_
#define PLIS_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PLIS_OK 0
#define PLIS_ERR_ARG (-1)      /* invalid argument / model                 */
#define PLIS_ERR_HIP (-2)      /* HIP runtime error                        */
#define PLIS_ERR_LEAPBACK (-3) /* a step needed more than 63 leap-backs    */
#define PLIS_ERR_NODEV (-4)    /* no gfx950 device                         */
#define PLIS_ERR_STATE (-5)    /* call not valid in the context's state    */

/* Reaction network shared by all realizations of a context.
 * Replaces Reactions (PLISCuda/include/cpp/Reactions.hpp:12-223; layout
 * reactants/products [s*R + r], spec/reactions.md:58-80) and the per-
 * compartment constants h_rrc / h_drc (driver.cu:98-104, read by
 * cpp_utils.hpp:29-45).  Orders 0..3 are supported (the reference handles
 * 1..2, rates.cu:3-39).  Propensity a_j = k_j * prod_s C(x_s, r_sj)
 * (spec/rate_matrix.md:46-58); hop propensity along edge e of species s from
 * a subvolume of compartment c = d[c*S+s] * x_s * w_e (rates.cu:55 has w=1). */
typedef struct plis_network {
    int32_t n_species;
    int32_t n_reactions;
    int32_t n_compartments;
    const int32_t *reactants;  /* [s*R + j], >= 0                       */
    const int32_t *products;   /* [s*R + j], >= 0                       */
    const double *k;           /* [c*R + j] reaction constants          */
    const double *d;           /* [c*S + s] per-edge diffusion constants */
} plis_network;

/* One realization's topology.  Replaces Topology's 6-slot self-padded rows
 * (PLISCuda/include/cpp/Topology.hpp:98-191, driver.cu:150-161): CSR with any
 * degree.  Self loops are dropped, duplicate neighbours merge (weights add). */
typedef struct plis_topology {
    uint32_t n_subvolumes;
    uint32_t n_edges;
    const uint32_t *row_ptr;   /* N+1                                    */
    const uint32_t *col;       /* n_edges neighbour ids < N              */
    const float *weight;       /* n_edges > 0, or NULL for all 1         */
    const int32_t *compartment;/* N ids in [0, C), or NULL for all 0.
                                  Replaces subv_constants (driver.cu:138-140) */
} plis_topology;

typedef struct plis_params {
    double epsilon;            /* Cao epsilon (reference: 0.05 fixed, driver.cu:37) */
    int32_t n_critical;        /* critical threshold n_c (reference: 10, driver.cu:36) */
    int32_t waves_per_block;   /* 0 = default/autotuned                  */
    int32_t blocks_per_cu;     /* 0 = default/autotuned                  */
    int32_t use_graph;         /* 1 = replay the step as a hipGraph      */
    uint64_t seed;             /* Philox key (reference: clock64()*v, init.cu:8) */
} plis_params;

typedef struct plis_stats {
    double t;                  /* simulated time                         */
    double tau;                /* last step's tau                        */
    uint64_t steps;            /* accepted global steps                  */
    uint64_t leapbacks;        /* rejected attempts (negative population)*/
    uint64_t firings;          /* reaction events + molecule hops        */
    uint64_t subvolume_leaps;  /* subvolumes advanced by accepted steps  */
    uint64_t crit_events;      /* critical (SSA) events fired            */
    int32_t status;            /* 0 running/paused, 1 done, 2 no event possible
                                  (the reference's "min(tau) = +Inf" stop,
                                  driver.cu:269-272), <0 error           */
    int32_t pad;
    uint64_t learned_steps;    /* accepted steps whose tau came from the network */
    uint64_t learned_rejects;  /* network proposals rejected by the judge  */
    uint64_t violations;       /* subvolumes over the post-leap bound, summed
                                  over accepted steps                     */
} plis_stats;

typedef struct plis_ctx plis_ctx;

/* Create a context holding n_real realizations (allocation and uploads of
 * driver.cu:55-213).  topos[i] / x0[i] are realization i's topology and
 * initial counts (species-major [s*N + v], uint32; the reference's State is
 * int32 [s*N+v], State.hpp:112-217).  ids[i] is the realization id that keys
 * its random streams (NULL = 0..n_real-1); ids must be < 65536.  */
int plis_create(const plis_network *net, int32_t n_real, const plis_topology *topos,
                const uint32_t *const *x0, const uint32_t *ids, const plis_params *params,
                plis_ctx **out);

/* Advance every realization until t >= t_end (the last step is clamped to
 * t_end), or until it has made step_limit accepted steps in total.  This is
 * the step loop of driver.cu:259-376. */
int plis_run(plis_ctx *ctx, double t_end, uint64_t step_limit);

/* As plis_run, but also stop a realization after the first accepted step
 * that reaches t_pause (no clamping at t_pause); used for logFreq snapshots
 * (driver.cu:362-375) without perturbing the trajectory. */
int plis_run_until(plis_ctx *ctx, double t_end, double t_pause, uint64_t step_limit);

/* Set every realization's stopping rule (as plis_run_until) without
 * running; plis_iterate / plis_profile then advance armed realizations. */
int plis_arm(plis_ctx *ctx, double t_end, double t_pause, uint64_t step_limit);

/* Run exactly n step iterations (one tau pass + one leap pass per
 * iteration, whatever the realizations' clocks) -- the benchmark's unit.
 * With plain tau an iteration runs the leap pass, the tau pass on the counts
 * it wrote and one kernel that commits the step and decides the next; a
 * call (this one, plis_run_until, plis_profile*) first decides the step on
 * the current counts, which gives the values already in place, if any. */
int plis_iterate(plis_ctx *ctx, uint32_t n);

/* Time n iterations with HIP events recorded on the context's stream around
 * every kernel launch; out[0] = total ms, then the summed ms of the tau pass,
 * the decide kernel, the leap pass and the commit kernel (out[1..4]).  With
 * plain tau, out[4] times the kernel that commits and decides and out[2]
 * is 0 (the call's first decision is made before the timed iterations). */
int plis_profile(plis_ctx *ctx, uint32_t n, double *out);

/* As plis_profile, with the leap pass split by kernel: out[0..4] as there,
 * out[5] = phase_b_kernel, out[6] = phase_b_redo_kernel, out[7] =
 * phase_b_finish_kernel (summed ms).  A context with the per-subvolume leap
 * kernel (some topology not symmetric) reports it in out[5] and 0 in
 * out[6..7].  Diagnostics of this implementation; no counterpart in the
 * reference. */
int plis_profile_kernels(plis_ctx *ctx, uint32_t n, double *out);

int plis_set_state(plis_ctx *ctx, int32_t r, const uint32_t *x);
int plis_get_state(plis_ctx *ctx, int32_t r, uint32_t *x);
int plis_get_stats(plis_ctx *ctx, int32_t r, plis_stats *st);
int plis_sync(plis_ctx *ctx);

/* Per-subvolume tau', a0c (critical propensity) and tau'' of the current
 * state of every realization, concatenated in realization order (kernel-level
 * parity; the reference's DEBUG dumps of driver.cu:223-239). */
int plis_debug_tau(plis_ctx *ctx, double *tau_prime, double *a0c, double *tau2);

/* Launch shape in use: out[0] waves per block, out[1] blocks, out[2] tiles. */
int plis_launch_shape(plis_ctx *ctx, int32_t *out);

/* Checkpoint / resume.  No counterpart in the reference, whose snapshots
 * hold counts only (PLISCuda/src/driver.cu:362-375); the blob holds every
 * realization's counts, clock, step and attempt indices, tau state and
 * statistics, which with counter-based draws fix the rest of the trajectory
 * bit for bit.  plis_restore needs a context created from the same network,
 * topologies, realization ids, epsilon, n_c and seed; it fails with
 * PLIS_ERR_ARG otherwise.  The learned network is saved with plis_mlp_get. */
int plis_checkpoint_size(plis_ctx *ctx, uint64_t *bytes);
int plis_checkpoint(plis_ctx *ctx, void *buf, uint64_t bytes);
int plis_restore(plis_ctx *ctx, const void *buf, uint64_t bytes);

/* Diagnostic (no counterpart in the reference): number of tiles the
 * edge-parallel leap passes of this context have handed to their redo pass
 * since plis_create, i.e. tiles holding a Poisson draw of mean >= 10 or more
 * undecided fast-path draws than the per-wavefront queue holds.  Results do
 * not depend on it. */
int plis_redo_tiles(plis_ctx *ctx, uint64_t *total);

/* Diagnostic standing in for the reference's cuda-memcheck initcheck job
 * (PLISCuda/test/run.sh:100-120): fill every device buffer that a step
 * writes before it reads -- each realization's non-current count buffer,
 * the tile partials of the tau and leap passes, the redo list -- with
 * sentinels that are harmless if read but change the trajectory, and forget
 * any learned proposals.  A run after it must equal a run without it bit for
 * bit; a kernel that read an unwritten word would break that. */
int plis_debug_poison(plis_ctx *ctx);

/* Autotune waves-per-block x blocks-per-CU on this context's workload by
 * timing `iters` real iterations per candidate (results do not depend on
 * the launch shape, so the trajectory is unchanged). */
int plis_autotune(plis_ctx *ctx, uint32_t iters);

/* ---- learned tau proposal (no counterpart in the reference) ----
 * A 12-64-64-2 MLP (bf16 weights, fp32 accumulation on matrix cores; GELU
 * activations by default, ReLU selectable) maps local features of each
 * subvolume to a log-tau proposal and a critical-flag logit.  w32 = 5378
 * fp32 parameters W1[16][64] b1 W2[64][64] b2 W3[64][2] b3 (rows 12..15 of
 * W1 unused); norm = feature means[12], scales[12], target mean, target
 * scale (log tau standardisation). */
int plis_mlp_set(plis_ctx *ctx, const float *w32, const float *norm);
int plis_mlp_get(plis_ctx *ctx, float *w32);
/* activation of the network's hidden layers: 0 = ReLU, 1 = GELU (default;
 * evaluated as max(z,0) - |z| max(0, p(|z|)) with a fitted cubic p, within
 * 0.006 of the erf form everywhere) */
int plis_mlp_activation(plis_ctx *ctx, int32_t act);
/* outputs [n][2] (standardised log tau, critical logit) for feature rows [n][12] */
int plis_mlp_forward(plis_ctx *ctx, const float *feats, uint32_t n, float *out);
/* As plis_mlp_forward, but allocates nothing: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under
 * graph capture, serving loops, harnesses that forbid it).  The rows and
 * outputs are staged through a fixed per-GPU scratch array of the library
 * (held until the process ends), so n is at most 16384 (PLIS_ERR_ARG
 * beyond); calls that overlap on one GPU race on that scratch. */
int plis_mlp_forward_noalloc(plis_ctx *ctx, const float *feats, uint32_t n, float *out);
/* mean loss and its gradient over n (multiple of 64) records, no update;
 * loss per record = (o0 - ylog)^2 + softplus(o1) - ycrit o1 */
int plis_mlp_loss_grad(plis_ctx *ctx, const float *feats, const float *ylog, const float *ycrit,
                       uint32_t n, float *grad, double *loss);
/* `steps` Adam steps (lr, beta 0.9/0.999) on random minibatches of `batch`
 * records drawn from n on the GPU; losses[steps] optional */
int plis_mlp_train(plis_ctx *ctx, const float *feats, const float *ylog, const float *ycrit,
                   uint64_t n, uint32_t steps, uint32_t batch, float lr, uint64_t seed,
                   float *losses);
/* Training records of realization r's current state: features [N][12],
 * log of Cao's tau'_v [N], and the accept/reject label [N] of a trial leap
 * of subvolume v alone at tau'_v, drawn from its own streams (1 = the
 * post-leap leap condition below breaks at v: reject; 0 = accept).  The
 * condition's tolerance is the context's theta (plis_set_tau_mode). */
int plis_records(plis_ctx *ctx, int32_t r, float *feats, float *ylog, float *crit);
/* Post-leap leap condition (Cao, Gillespie & Petzold 2006 eq. 5 in
 * propensity form with its one-molecule floor, checked after the leap):
 * subvolume v breaks it when a non-critical channel j changes by
 * |a_j(x') - a_j(x)| > theta max(eps, max_i r_ij / x_i) a_j(x), or an active
 * hop channel of species s by |x'_s - x_s| > theta max(eps x_s, 1).
 * theta = 0 turns the condition off (plain runs); theta > 0 counts the
 * subvolumes that break it into plis_stats.violations.
 * learned = 1 (needs theta > 0 and a network): tau = min_v tau_v with
 * tau_v = kappa exp(network log tau), halved where the critical logit is
 * positive (the critical part stays exact); the proposals of the next step
 * are computed inside the leap pass from the state it produces.  A learned
 * step is kept only if no population goes negative and at most
 * mu + 3 sqrt(mu) + 2 subvolumes break the condition, mu = viol_rate_max N
 * (a Poisson upper bound around the plain method's own rate), otherwise it
 * is redone with Cao's tau and fresh draws (counted in learned_rejects). */
int plis_set_tau_mode(plis_ctx *ctx, int32_t learned, double kappa, double theta,
                      double viol_rate_max);

void plis_destroy(plis_ctx *ctx);

/* Message of the last failing call on ctx (ctx may be NULL for plis_create
 * failures; that message is thread-local). */
const char *plis_last_error(const plis_ctx *ctx);

/* Number of gfx950 devices visible (0 when none); name buffer optional. */
int plis_device_count(char *name, int32_t name_len);

#ifdef __cplusplus
}
#endif
#endif
