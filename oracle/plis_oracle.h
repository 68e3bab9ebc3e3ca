/*
 * plis_oracle.h -- CPU oracle for the RDME tau-leap + SSA step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker, never the product: only the
 * test suite, the smoke check and the CPU-baseline leg of the benchmark may
 * load it.  The product (libplis_hip.so) never links or calls anything here.
 *
 * It is an independent plain-C restatement of PLISCuda's step
 * (PLISCuda/src/{rates,tau,leap,ssa,init}.cu and the host loop
 * PLISCuda/src/driver.cu:259-376), with the reference's quirks corrected
 * (hops are reactions in Cao's mu/sigma^2, one critical event per global
 * step in the argmin-tau'' subvolume, leap-back halves tau', the last step is
 * clamped to t_end).  It draws from the same counter-based Philox4x32-10
 * streams as the GPU kernels, so a GPU run and an oracle run from the same
 * seed give bit-identical trajectories.
 *
 * It also holds an exact next-reaction (direct method) SSA for ensemble
 * statistics and the learned-tau MLP reference (forward, backward, Adam).
 */
#ifndef PLIS_ORACLE_H_
#define PLIS_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int32_t S, R, C;          /* species, reactions, compartments        */
    const int32_t *reac;      /* [s*R + j] reactant stoichiometry         */
    const int32_t *prod;      /* [s*R + j] product stoichiometry          */
    const double *k;          /* [c*R + j] reaction constants             */
    const double *d;          /* [c*S + s] per-edge diffusion constants   */
} orc_network;

typedef struct {
    uint32_t N;               /* subvolumes                               */
    uint32_t E;               /* directed edges in col / w                */
    const uint32_t *row_ptr;  /* N+1                                      */
    const uint32_t *col;      /* E neighbour ids (any order, dups allowed)*/
    const float *w;           /* E edge weights, NULL = all 1             */
    const int32_t *comp;      /* N compartment ids, NULL = all 0          */
} orc_topology;

typedef struct {
    double epsilon;
    int32_t n_critical;
    uint64_t seed;
} orc_params;

typedef struct {
    double t;
    uint64_t steps, leapbacks, firings, subvolume_leaps, crit_events;
    int32_t status;           /* 0 running/ok, 1 done, 2 absorbed, <0 err */
} orc_stats;

/* ---- numerics (exported so tests can pin them) ---- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double orc_exp(double x);
double orc_log(double x);
double orc_sqrt(double x);
double orc_logfact(int64_t k);
/* Poisson draw with mean lam; u0 is the draw's packed uniform word, the
 * PTRS stream is (ctr, key) as defined by the counter layout. */
int64_t orc_poisson(double lam, uint32_t u0, const uint32_t ctr[4], const uint32_t key[2],
                    uint32_t word);
/* Convenience: n draws with mean lam from distinct REACT streams (for
 * distribution tests).  out[i] uses subvolume id i. */
void orc_poisson_many(double lam, uint64_t seed, uint32_t n, int64_t *out);
/* Check of the kernels' float squeeze of the inversion search (see the
 * comment in plis_oracle.c): float e^-l, the squeeze decision (k or -1),
 * the kernels' whole fast path for one draw, and the worst |G_k - F_k| over
 * float means with bits b0, b0+stride, ... < b1. */
float orc_sq_exp_neg(float l);
int orc_sq_decide(float l, float u);
int orc_sq_decide_band(float l, float u, float band);
int orc_sq_poisson(double lam, uint32_t u0);
double orc_sq_maxerr(uint32_t b0, uint32_t b1, uint32_t stride);
void orc_sq_hop_check(const double *tau, const double *d, const uint32_t *x, const float *w,
                      const uint32_t *u0, int32_t n, int32_t *fast, int64_t *exact);

/* ---- per-subvolume kernels (for kernel-level parity) ---- */
/* tau' and a0c for every subvolume of state x (SoA [s*N+v]); tau2 uses the
 * step's TAU2 stream (realization r, step). NULL outputs are skipped. */
int orc_tau_all(const orc_network *net, const orc_topology *topo, const orc_params *p,
                const uint32_t *x, uint32_t realization, uint32_t step,
                double *tau_prime, double *a0c, double *tau2);

/* The critical event chosen at subvolume v for each of nsteps step indices
 * (kind 1 = reaction idx, 2 = hop of species idx to dst, 0 = none). */
int orc_critical_choice(const orc_network *net, const orc_topology *topo, const orc_params *p,
                        const uint32_t *x, uint32_t realization, uint32_t v, uint32_t nsteps,
                        const uint32_t *steps, int32_t *kind, int32_t *idx, uint32_t *dst);

/* ---- the plain step loop ---- */
/* Advance one realization until t >= t_end, steps >= step_limit, or
 * t >= t_pause (soft stop, no clamp).  x is SoA [s*N+v], in/out.
 * *t, *step, stats accumulate across calls. Returns 0 or <0 on error. */
int orc_run(const orc_network *net, const orc_topology *topo, const orc_params *p,
            uint32_t realization, uint32_t *x, double t_end, double t_pause,
            uint64_t step_limit, orc_stats *stats);

/* Many realizations in parallel (OpenMP over realizations).  x[i] points to
 * realization i's state, realization ids are ids[i]. */
int orc_run_batch(const orc_network *net, int32_t nreal, const orc_topology *topos,
                  const orc_params *p, const uint32_t *ids, uint32_t *const *x,
                  double t_end, uint64_t step_limit, orc_stats *stats);

/* ---- exact SSA (direct method over subvolumes, Gillespie 1977) ---- */
int orc_ssa_run(const orc_network *net, const orc_topology *topo, uint64_t seed,
                uint32_t realization, uint32_t *x, double t_end, uint64_t *events);
int orc_ssa_batch(const orc_network *net, const orc_topology *topo, uint64_t seed,
                  int32_t nreal, uint32_t *x_all /* nreal * S*N */, double t_end);

int orc_num_threads(void);

/* ---- learned tau (plis_mlp.c) ---- */
int orc_features(const orc_network *net, const orc_topology *topo, const orc_params *p,
                 const uint32_t *x, float prev_logtau, float *feats /* N*12 */);
/* post-leap leap condition (propensity form) per subvolume; returns the
 * number of subvolumes where it is broken */
int orc_violations(const orc_network *net, const orc_topology *topo, const orc_params *p,
                   const uint32_t *x, const uint32_t *xn, double theta, uint8_t *flags);
/* training records: features, log Cao tau', trial-leap reject label */
int orc_records(const orc_network *net, const orc_topology *topo, const orc_params *p,
                const uint32_t *x, uint32_t realization, uint32_t step, float prev_logtau,
                double theta, float *feats, float *ylog, float *label);
int orc_mlp_param_count(void);
/* act: 0 = ReLU, 1 = GELU (the clamped-cubic form of plis_mlp.c) */
void orc_mlp_forward(const float *w, const float *norm, const float *feats, int32_t n,
                     int32_t bfon, int32_t act, float *out);
double orc_mlp_loss_grad(const float *w, const float *norm, const float *feats, int32_t n,
                         int32_t bfon, int32_t act, const float *ylog, const float *ycrit,
                         double *grad);
float orc_gelu(float z);
void orc_adam_step(float *w, float *m, float *v, const float *g, int32_t count, int32_t t,
                   float lr, float b1, float b2, float eps);

#ifdef __cplusplus
}
#endif
#endif
