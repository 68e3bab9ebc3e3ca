/*
 * plis_oracle.c -- CPU oracle for the RDME tau-leap + SSA step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP kernels, never shipped or
 * called by the product.  See plis_oracle.h for what it restates.
 *
 * Numerics are written once here from published methods and must be applied
 * with exactly the same operation order as the GPU kernels (built with
 * -ffp-contract=off; every fused multiply-add is an explicit fma()):
 *   Philox4x32-10 ........ Salmon, Moraes, Dror, Shaw, "Parallel random numbers:
 *                          as easy as 1, 2, 3", SC'11.
 *   exp .................. Cody-Waite reduction by ln2 (hi/lo split), Taylor
 *                          polynomial of degree 13 in Horner form.
 *   log .................. mantissa in (sqrt(2)/2, sqrt(2)], log m = 2 atanh(s),
 *                          s = (m-1)/(m+1), odd series to s^21.
 *   sqrt ................. Newton-Heron iteration, 6 steps from an exponent-
 *                          halving first guess.
 *   Poisson .............. inversion by sequential search for mean < 10
 *                          (Devroye 1986, X.3); PTRS transformed rejection
 *                          with squeeze for mean >= 10 (Hoermann 1993).
 *   log k! ............... exact table for k < 10, Stirling series to 1/z^5.
 *   tau selection ........ Cao, Gillespie, Petzold, J. Chem. Phys. 124, 044109
 *                          (2006), eqs. 32a/32b/33 and the g_i table; critical
 *                          reactions L_j < n_c (Cao et al. 2005).
 */
#include "plis_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* constants (generated with 60-digit decimal arithmetic)               */
/* ------------------------------------------------------------------ */
#define LN2_HI 0x1.62e42fefa3800p-1 /* ln2 truncated to 42 significant bits */
#define LN2_LO 0x1.ef35793c76730p-45
#define INV_LN2 0x1.71547652b82fep+0
#define SQRT2 0x1.6a09e667f3bcdp+0
#define HALF_LOG_2PI 0x1.d67f1c864beb5p-1

static const double INVFACT[14] = {
    0x1.0000000000000p+0, 0x1.0000000000000p+0, 0x1.0000000000000p-1, 0x1.5555555555555p-3,
    0x1.5555555555555p-5, 0x1.1111111111111p-7, 0x1.6c16c16c16c17p-10, 0x1.a01a01a01a01ap-13,
    0x1.a01a01a01a01ap-16, 0x1.71de3a556c734p-19, 0x1.27e4fb7789f5cp-22, 0x1.ae64567f544e4p-26,
    0x1.1eed8eff8d898p-29, 0x1.6124613a86d09p-33};
static const double ATANH_C[11] = {
    0x1.0000000000000p+0, 0x1.5555555555555p-2, 0x1.999999999999ap-3, 0x1.2492492492492p-3,
    0x1.c71c71c71c71cp-4, 0x1.745d1745d1746p-4, 0x1.3b13b13b13b14p-4, 0x1.1111111111111p-4,
    0x1.e1e1e1e1e1e1ep-5, 0x1.af286bca1af28p-5, 0x1.8618618618618p-5};
static const double LOGFACT[10] = {
    0.0, 0.0, 0x1.62e42fefa39efp-1, 0x1.cab0bfa2a2002p+0, 0x1.96ca77c922cf9p+1,
    0x1.326643c4479c9p+2, 0x1.a51273acf01cap+2, 0x1.10ce1f32dcc30p+3, 0x1.5358e82fcb70dp+3,
    0x1.99a8921a7f7cfp+3};
#define C_1_12 0x1.5555555555555p-4
#define C_1_360 0x1.6c16c16c16c17p-9
#define C_1_1260 0x1.a01a01a01a01ap-11

/* Philox stream purposes (3 bits of counter word 3) */
enum { P_TAU2 = 1, P_CRIT = 2, P_REACT = 3, P_HOP = 4, P_EXTRA = 5, P_SSA = 6, P_TRIAL = 7 };
#define SELF_ID 0xFFFFFFFFu
#define MAX_ATTEMPT 63u
#define POISSON_INV_MAX 10.0

static inline uint64_t dbits(double x) { uint64_t b; memcpy(&b, &x, 8); return b; }
static inline double bitsd(uint64_t b) { double x; memcpy(&x, &b, 8); return x; }

/* ------------------------------------------------------------------ */
/* Philox4x32-10                                                       */
/* ------------------------------------------------------------------ */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int i = 0; i < 10; i++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline uint32_t ctr3(uint32_t real, uint32_t attempt, uint32_t purpose, uint32_t block)
{
    return (real << 16) | (attempt << 10) | (purpose << 7) | block;
}

static inline double uni(uint32_t w) { return ((double)w + 0.5) * 0x1p-32; }

/* ------------------------------------------------------------------ */
/* elementary functions                                                */
/* ------------------------------------------------------------------ */
double orc_exp(double x)
{
    if (x < -708.0) return 0.0;
    if (x > 709.0) return INFINITY;
    double n = rint(x * INV_LN2);
    double r = fma(-n, LN2_HI, x);
    r = fma(-n, LN2_LO, r);
    double p = INVFACT[13];
    for (int k = 12; k >= 0; k--) p = fma(p, r, INVFACT[k]);
    int64_t ni = (int64_t)n;
    double scale = bitsd((uint64_t)(ni + 1023) << 52);
    return p * scale;
}

double orc_log(double x)
{
    uint64_t b = dbits(x);
    int e = (int)((b >> 52) & 0x7FF) - 1023;
    double m = bitsd((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
    if (m > SQRT2) { m = m * 0.5; e += 1; }
    double s = (m - 1.0) / (m + 1.0);
    double z = s * s;
    double P = ATANH_C[10];
    for (int k = 9; k >= 0; k--) P = fma(P, z, ATANH_C[k]);
    double t = (2.0 * s) * P;
    double ed = (double)e;
    return fma(ed, LN2_HI, fma(ed, LN2_LO, t));
}

double orc_sqrt(double x)
{
    if (!(x > 0.0)) return 0.0;
    double y = bitsd((dbits(x) >> 1) + (0x3FF0000000000000ull >> 1));
    for (int i = 0; i < 6; i++) y = 0.5 * (y + x / y);
    return y;
}

double orc_logfact(int64_t k)
{
    if (k < 10) return LOGFACT[k];
    double z = (double)k + 1.0;
    double z2 = z * z;
    return (((double)k + 0.5) * orc_log(z) - z) + HALF_LOG_2PI
           + (C_1_12 - (C_1_360 - C_1_1260 / z2) / z2) / z;
}

/* ------------------------------------------------------------------ */
/* Poisson sampler                                                     */
/* ------------------------------------------------------------------ */
/* Extra uniform q of a draw whose main counter is ctr and word w. */
static inline double extra_uni(const uint32_t ctr[4], const uint32_t key[2], uint32_t word,
                               uint32_t q)
{
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], (ctr[3] & ~(7u << 7)) | ((uint32_t)P_EXTRA << 7)};
    uint32_t k[2] = {key[0] ^ (0x85EBCA6Bu * (4u * (q >> 2) + word + 1u)), key[1]};
    uint32_t o[4];
    orc_philox4x32_10(c, k, o);
    return uni(o[q & 3u]);
}

int64_t orc_poisson(double lam, uint32_t u0, const uint32_t ctr[4], const uint32_t key[2],
                    uint32_t word)
{
    if (!(lam > 0.0)) return 0;
    if (lam < POISSON_INV_MAX) {
        double u = uni(u0);
        double p = orc_exp(-lam);
        double F = p;
        int64_t k = 0;
        while (u > F && k < 1000) {
            k += 1;
            p = (p * lam) / (double)k;
            F = F + p;
        }
        return k;
    }
    /* PTRS (Hoermann 1993) */
    double slam = orc_sqrt(lam);
    double loglam = orc_log(lam);
    double b = 0.931 + 2.53 * slam;
    double a = -0.059 + 0.02483 * b;
    double invalpha = 1.1239 + 1.1328 / (b - 3.4);
    double vr = 0.9277 - 3.6224 / (b - 2.0);
    double loginvalpha = orc_log(invalpha);
    for (uint32_t it = 0; it < 100000u; it++) {
        double U = extra_uni(ctr, key, word, 2u * it) - 0.5;
        double V = extra_uni(ctr, key, word, 2u * it + 1u);
        double us = 0.5 - fabs(U);
        double kd = floor((2.0 * a / us + b) * U + lam + 0.43);
        if (us >= 0.07 && V <= vr) return (int64_t)kd;
        if (kd < 0.0 || (us < 0.013 && V > us)) continue;
        int64_t k = (int64_t)kd;
        double lhs = (orc_log(V) + loginvalpha) - orc_log(a / (us * us) + b);
        double rhs = (-lam + kd * loglam) - orc_logfact(k);
        if (lhs <= rhs) return k;
    }
    return (int64_t)floor(lam);
}

void orc_poisson_many(double lam, uint64_t seed, uint32_t n, int64_t *out)
{
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (uint32_t i = 0; i < n; i++) {
        uint32_t c[4] = {i, SELF_ID, 0u, ctr3(0, 0, P_REACT, 0)};
        uint32_t o[4];
        orc_philox4x32_10(c, key, o);
        out[i] = orc_poisson(lam, o[0], c, key, 0);
    }
}

/* ---- check of the kernels' single-precision squeeze -----------------
 * The kernels decide most inversion draws (mean < 10) in float: G_k, the
 * CDF accumulated in single precision at the float-rounded mean, is
 * compared with the float-rounded uniform with a guard band SQ_DELTA, and
 * only a uniform inside the band (or k > SQ_KMAX) goes to the double
 * search above.  The functions below restate that float evaluation with the
 * same IEEE operations (fmaf, rintf, round-to-nearest products and sums) so
 * that the tests can (1) bound max |G_k - F_k| over every float mean, which
 * is what makes the squeeze return the double search's result, and
 * (2) compare decisions draw by draw.  Checker only: orc_poisson above is
 * the reference's draw. */
#define SQ_DELTA 0x1p-15f
#define SQ_KMAX 15
static const float sq_recip[SQ_KMAX + 1] = {
    0.0f,           0x1p+0f,        0x1p-1f,        0x1.555556p-2f, 0x1p-2f,        0x1.99999ap-3f,
    0x1.555556p-3f, 0x1.24924ap-3f, 0x1p-3f,        0x1.c71c72p-4f, 0x1.99999ap-4f, 0x1.745d18p-4f,
    0x1.555556p-4f, 0x1.3b13b2p-4f, 0x1.24924ap-4f, 0x1.111112p-4f};

float orc_sq_exp_neg(float l)
{
    float n = rintf(l * 0x1.715476p+0f);
    float r = fmaf(n, 0x1.62e4p-1f, -l);
    r = fmaf(n, 0x1.7f7d1cp-20f, r);
    float p = 0x1.a01a02p-13f;
    p = fmaf(p, r, 0x1.6c16c2p-10f);
    p = fmaf(p, r, 0x1.111112p-7f);
    p = fmaf(p, r, 0x1.555556p-5f);
    p = fmaf(p, r, 0x1.555556p-3f);
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    union { int32_t i; float f; } sc;
    sc.i = (127 - (int32_t)n) << 23;
    return p * sc.f;
}

/* band = SQ_DELTA is the kernels' squeeze; other bands let the tests show
 * what a narrower guard band would get wrong. */
int orc_sq_decide_band(float l, float u, float band)
{
    float p = orc_sq_exp_neg(l);
    float G = p;
    for (int k = 0;; ++k) {
        if (u <= G - band) return k;
        if (!(u > G + band) || k == SQ_KMAX) return -1;
        p = (p * l) * sq_recip[k + 1];
        G = G + p;
    }
}

int orc_sq_decide(float l, float u) { return orc_sq_decide_band(l, u, SQ_DELTA); }

/* The kernels' full inversion fast path for one draw: 0 from the exact
 * zero shortcut, else the squeeze; -1 = left to the double search. */
int orc_sq_poisson(double lam, uint32_t u0)
{
    if (!(lam > 0.0) || !(lam < POISSON_INV_MAX)) return -1;
    double u = uni(u0);
    if (u <= (1.0 - lam) - 0x1p-50) return 0;
    return orc_sq_decide((float)lam, (float)u);
}

/* The kernels' fast path for a hop draw whose mean they form in single
 * precision (tau, d, x rounded to float, w already float, three float
 * products) and whose uniform is the word rounded to float: for each of n
 * draws, fast[i] = k or -1 (left to the double search), exact[i] = the
 * oracle's draw from the double mean tau ((d x) w) (-1 when that mean is
 * >= 10, which the fast path never settles). */
void orc_sq_hop_check(const double *tau, const double *d, const uint32_t *x, const float *w,
                      const uint32_t *u0, int32_t n, int32_t *fast, int64_t *exact)
{
    static const uint32_t zc[4] = {0, 0, 0, 0}, zk[2] = {0, 0};
    for (int32_t i = 0; i < n; i++) {
        double lam = tau[i] * ((d[i] * (double)x[i]) * (double)w[i]);
        float lf = (float)tau[i] * (((float)d[i] * (float)x[i]) * w[i]);
        exact[i] = lam < POISSON_INV_MAX ? orc_poisson(lam, u0[i], zc, zk, 0) : -1;
        if (!(lf > 0.0f)) { fast[i] = 0; continue; }
        float u = (float)u0[i] * 0x1p-32f;
        if (u <= (1.0f - lf) - 0x1p-19f) { fast[i] = 0; continue; }
        if (!(lf < 9.9999f)) { fast[i] = -1; continue; }
        fast[i] = orc_sq_decide(lf, u);
    }
}

/* max over float means l with bit patterns b0, b0+stride, ... < b1 and over
 * k <= SQ_KMAX of |G_k(l) - F_k(l)|, F_k the double search's CDF. */
double orc_sq_maxerr(uint32_t b0, uint32_t b1, uint32_t stride)
{
    double worst = 0.0;
    int64_t n = ((int64_t)b1 - (int64_t)b0 + stride - 1) / stride;
#pragma omp parallel for schedule(static, 65536) reduction(max : worst)
    for (int64_t i = 0; i < n; i++) {
        union { uint32_t b; float f; } cv;
        cv.b = b0 + (uint32_t)i * stride;
        float l = cv.f;
        double lam = (double)l;
        float gp = orc_sq_exp_neg(l);
        float G = gp;
        double p = orc_exp(-lam);
        double F = p;
        for (int k = 0;; ++k) {
            double e = fabs((double)G - F);
            if (e > worst) worst = e;
            if (k == SQ_KMAX) break;
            gp = (gp * l) * sq_recip[k + 1];
            G = G + gp;
            p = (p * lam) / (double)(k + 1);
            F = F + p;
        }
    }
    return worst;
}

/* ------------------------------------------------------------------ */
/* model preprocessing                                                 */
/* ------------------------------------------------------------------ */
typedef struct {
    int S, R, C;
    const double *k, *d;
    int32_t *nu;       /* [s*R + j] */
    int *order;        /* [j] */
    int *nrs;          /* [j] number of reactant species */
    int *rs, *rr;      /* [j*3 + i] reactant species / multiplicity */
    int *hor, *mult;   /* [s] */
    int *diff;         /* [s] species diffuses in some compartment */
} model_t;

typedef struct {
    uint32_t N;
    uint32_t *rp, *col;   /* canonical out-CSR */
    float *w;
    double *W;            /* out weight sum per v */
    uint32_t *irp, *isrc; /* in-CSR */
    float *iw;
    const int32_t *comp;
} csr_t;

static void model_free(model_t *m)
{
    free(m->nu); free(m->order); free(m->nrs); free(m->rs); free(m->rr);
    free(m->hor); free(m->mult); free(m->diff);
}

static int model_build(const orc_network *net, model_t *m)
{
    memset(m, 0, sizeof *m);
    int S = net->S, R = net->R, C = net->C;
    if (S < 1 || S > 64 || R < 0 || R > 128 || C < 1) return -1;
    m->S = S; m->R = R; m->C = C; m->k = net->k; m->d = net->d;
    m->nu = calloc((size_t)S * R + 1, sizeof(int32_t));
    m->order = calloc((size_t)R + 1, sizeof(int));
    m->nrs = calloc((size_t)R + 1, sizeof(int));
    m->rs = calloc((size_t)R * 3 + 1, sizeof(int));
    m->rr = calloc((size_t)R * 3 + 1, sizeof(int));
    m->hor = calloc((size_t)S, sizeof(int));
    m->mult = calloc((size_t)S, sizeof(int));
    m->diff = calloc((size_t)S, sizeof(int));
    for (int j = 0; j < R; j++) {
        int ord = 0, n = 0;
        for (int s = 0; s < S; s++) {
            int r = net->reac[s * R + j], pr = net->prod[s * R + j];
            if (r < 0 || pr < 0) return -1;
            m->nu[s * R + j] = pr - r;
            if (r > 0) {
                if (n >= 3) return -1;
                m->rs[j * 3 + n] = s; m->rr[j * 3 + n] = r; n++;
                ord += r;
            }
        }
        if (ord > 3) return -1;
        m->order[j] = ord; m->nrs[j] = n;
    }
    for (int s = 0; s < S; s++) {
        int hor = 0, mult = 0;
        for (int j = 0; j < R; j++) {
            int r = net->reac[s * R + j];
            if (r <= 0) continue;
            if (m->order[j] > hor) { hor = m->order[j]; mult = r; }
            else if (m->order[j] == hor && r > mult) mult = r;
        }
        int df = 0;
        for (int c = 0; c < C; c++) if (net->d[c * S + s] > 0.0) df = 1;
        m->diff[s] = df;
        if (df && hor == 0) { hor = 1; mult = 1; }
        m->hor[s] = hor; m->mult[s] = mult;
    }
    return 0;
}

static void csr_free(csr_t *g) { free(g->rp); free(g->col); free(g->w); free(g->W); free(g->irp); free(g->isrc); free(g->iw); }

typedef struct { uint32_t id; uint32_t pos; float w; } nbr_t;
static int nbr_cmp(const void *a, const void *b)
{
    const nbr_t *x = a, *y = b;
    if (x->id != y->id) return x->id < y->id ? -1 : 1;
    return x->pos < y->pos ? -1 : (x->pos > y->pos);
}

/* canonical form: self loops dropped, neighbours ascending, duplicates merged
 * (weights summed in double in input order, rounded to float). */
static int csr_build(const orc_topology *t, csr_t *g)
{
    memset(g, 0, sizeof *g);
    uint32_t N = t->N;
    g->N = N; g->comp = t->comp;
    g->rp = calloc((size_t)N + 1, sizeof(uint32_t));
    g->col = malloc(((size_t)t->E + 1) * sizeof(uint32_t));
    g->w = malloc(((size_t)t->E + 1) * sizeof(float));
    g->W = calloc((size_t)N, sizeof(double));
    nbr_t *tmp = malloc(((size_t)t->E + 1) * sizeof(nbr_t));
    uint32_t ne = 0;
    for (uint32_t v = 0; v < N; v++) {
        uint32_t b = t->row_ptr[v], e = t->row_ptr[v + 1], n = 0;
        for (uint32_t q = b; q < e; q++) {
            uint32_t u = t->col[q];
            if (u >= N) { free(tmp); return -1; }
            if (u == v) continue;
            float wq = t->w ? t->w[q] : 1.0f;
            if (!(wq > 0.0f)) { free(tmp); return -1; }
            tmp[n].id = u; tmp[n].pos = q; tmp[n].w = wq; n++;
        }
        qsort(tmp, n, sizeof(nbr_t), nbr_cmp);
        g->rp[v] = ne;
        for (uint32_t i = 0; i < n;) {
            uint32_t u = tmp[i].id;
            double acc = 0.0;
            uint32_t i2 = i;
            while (i2 < n && tmp[i2].id == u) { acc += (double)tmp[i2].w; i2++; }
            g->col[ne] = u;
            g->w[ne] = (i2 - i == 1) ? tmp[i].w : (float)acc;
            ne++;
            i = i2;
        }
    }
    g->rp[N] = ne;
    free(tmp);
    for (uint32_t v = 0; v < N; v++) {
        double W = 0.0;
        for (uint32_t q = g->rp[v]; q < g->rp[v + 1]; q++) W += (double)g->w[q];
        g->W[v] = W;
    }
    /* in-CSR: sources ascending */
    g->irp = calloc((size_t)N + 1, sizeof(uint32_t));
    g->isrc = malloc(((size_t)ne + 1) * sizeof(uint32_t));
    g->iw = malloc(((size_t)ne + 1) * sizeof(float));
    for (uint32_t q = 0; q < ne; q++) g->irp[g->col[q] + 1]++;
    for (uint32_t v = 0; v < N; v++) g->irp[v + 1] += g->irp[v];
    uint32_t *fill = calloc((size_t)N, sizeof(uint32_t));
    for (uint32_t u = 0; u < N; u++)
        for (uint32_t q = g->rp[u]; q < g->rp[u + 1]; q++) {
            uint32_t v = g->col[q];
            uint32_t pos = g->irp[v] + fill[v]++;
            g->isrc[pos] = u; g->iw[pos] = g->w[q];
        }
    free(fill);
    return 0;
}

/* ------------------------------------------------------------------ */
/* per-subvolume quantities                                            */
/* ------------------------------------------------------------------ */
static inline double binom_d(uint32_t x, int r)
{
    if ((int64_t)x < r) return 0.0;
    double xd = (double)x;
    if (r == 1) return xd;
    if (r == 2) return (xd * (xd - 1.0)) * 0.5;
    return ((xd * (xd - 1.0)) * (xd - 2.0)) / 6.0;
}

static inline int comp_of(const csr_t *g, uint32_t v) { return g->comp ? g->comp[v] : 0; }
static inline uint32_t X(const uint32_t *x, uint32_t N, int s, uint32_t v) { return x[(size_t)s * N + v]; }

/* reaction propensities and critical flags at v */
static void react_props(const model_t *m, const csr_t *g, const orc_params *p, const uint32_t *x,
                        uint32_t v, double *a, int *crit)
{
    int c = comp_of(g, v);
    for (int j = 0; j < m->R; j++) {
        double aj = m->k[c * m->R + j];
        int64_t L = -1; /* -1 = unbounded */
        for (int i = 0; i < m->nrs[j]; i++) {
            int s = m->rs[j * 3 + i], r = m->rr[j * 3 + i];
            uint32_t xs = X(x, g->N, s, v);
            aj = aj * binom_d(xs, r);
            int64_t l = (int64_t)(xs / (uint32_t)r);
            if (L < 0 || l < L) L = l;
        }
        a[j] = aj;
        crit[j] = (m->order[j] > 0 && L < (int64_t)p->n_critical);
    }
}

static inline int hop_active(const model_t *m, const orc_params *p, double d, uint32_t xs)
{
    (void)m;
    return d > 0.0 && (int64_t)xs >= (int64_t)p->n_critical && xs > 0;
}

static inline double g_cao(const model_t *m, int s, uint32_t xs)
{
    double xd = (double)xs;
    int hor = m->hor[s], mu = m->mult[s];
    if (hor == 1) return 1.0;
    if (hor == 2) return mu == 2 ? 2.0 + 1.0 / (xd - 1.0) : 2.0;
    if (mu == 3) return (3.0 + 1.0 / (xd - 1.0)) + 2.0 / (xd - 2.0);
    if (mu == 2) return 1.5 * (2.0 + 1.0 / (xd - 1.0));
    return 3.0;
}

/* tau' and a0c at v.  Cao 2006 eq. 33 with hops as first-order channels:
 * out-hops of s from v have nu=-1 and total propensity (d*x)*W_v, in-hops from
 * u have nu=+1 and propensity (d_u*x_u)*w_uv. */
static void tau_v(const model_t *m, const csr_t *g, const orc_params *p, const uint32_t *x,
                  uint32_t v, double *tau_prime, double *a0c_out)
{
    double a[128];
    int crit[128];
    react_props(m, g, p, x, v, a, crit);
    int c = comp_of(g, v);
    double W = g->W[v];
    double a0c = 0.0;
    for (int j = 0; j < m->R; j++) if (crit[j]) a0c += a[j];
    for (int s = 0; s < m->S; s++) {
        uint32_t xs = X(x, g->N, s, v);
        double d = m->d[c * m->S + s];
        if (d > 0.0 && !hop_active(m, p, d, xs)) a0c += (d * (double)xs) * W;
    }
    double tp = INFINITY;
    for (int s = 0; s < m->S; s++) {
        if (m->hor[s] == 0) continue;
        uint32_t xs = X(x, g->N, s, v);
        double mu = 0.0, sig = 0.0;
        for (int j = 0; j < m->R; j++) {
            int nu = m->nu[s * m->R + j];
            if (crit[j] || nu == 0) continue;
            mu += (double)nu * a[j];
            sig += (double)(nu * nu) * a[j];
        }
        double d = m->d[c * m->S + s];
        if (hop_active(m, p, d, xs)) {
            double aout = (d * (double)xs) * W;
            mu -= aout;
            sig += aout;
        }
        for (uint32_t q = g->irp[v]; q < g->irp[v + 1]; q++) {
            uint32_t u = g->isrc[q];
            uint32_t xu = X(x, g->N, s, u);
            double du = m->d[comp_of(g, u) * m->S + s];
            if (!hop_active(m, p, du, xu)) continue;
            double ain = (du * (double)xu) * (double)g->iw[q];
            mu += ain;
            sig += ain;
        }
        double gg = g_cao(m, s, xs);
        double mm = fmax((p->epsilon * (double)xs) / gg, 1.0);
        double t1 = mm / fabs(mu);
        double t2 = (mm * mm) / sig;
        tp = fmin(tp, fmin(t1, t2));
    }
    *tau_prime = tp;
    *a0c_out = a0c;
}

static inline double tau2_of(double a0c, const orc_params *p, uint32_t real, uint32_t step, uint32_t v)
{
    if (!(a0c > 0.0)) return INFINITY;
    uint32_t key[2] = {(uint32_t)p->seed, (uint32_t)(p->seed >> 32)};
    uint32_t c[4] = {v, SELF_ID, step, ctr3(real, 0, P_TAU2, 0)};
    uint32_t o[4];
    orc_philox4x32_10(c, key, o);
    return -orc_log(uni(o[0])) / a0c;
}

int orc_tau_all(const orc_network *net, const orc_topology *topo, const orc_params *p,
                const uint32_t *x, uint32_t realization, uint32_t step,
                double *tau_prime, double *a0c, double *tau2)
{
    model_t m; csr_t g;
    if (model_build(net, &m)) { model_free(&m); return -1; }
    if (csr_build(topo, &g)) { model_free(&m); csr_free(&g); return -1; }
#pragma omp parallel for schedule(static)
    for (int64_t vv = 0; vv < (int64_t)g.N; vv++) {
        uint32_t v = (uint32_t)vv;
        double tp, ac;
        tau_v(&m, &g, p, x, v, &tp, &ac);
        if (tau_prime) tau_prime[v] = tp;
        if (a0c) a0c[v] = ac;
        if (tau2) tau2[v] = tau2_of(ac, p, realization, step, v);
    }
    model_free(&m); csr_free(&g);
    return 0;
}

/* ------------------------------------------------------------------ */
/* the step                                                            */
/* ------------------------------------------------------------------ */
typedef struct {
    double t, t_end, tau_p, tau2, tau;
    uint32_t step, attempt, vstar;
    int crit, last;
    int ev_kind;           /* 0 none, 1 reaction, 2 hop */
    int ev_idx;            /* reaction j or species s */
    uint32_t ev_dst;       /* hop destination */
} ctl_t;

/* choose the critical event at v* (always chosen when tau2 is finite so that
 * a leap-back that makes it fire can reuse it) */
static void choose_event(const model_t *m, const csr_t *g, const orc_params *p, const uint32_t *x,
                         uint32_t real, ctl_t *st)
{
    uint32_t v = st->vstar;
    double a[128];
    int crit[128];
    react_props(m, g, p, x, v, a, crit);
    int c = comp_of(g, v);
    double W = g->W[v];
    double a0c = 0.0;
    for (int j = 0; j < m->R; j++) if (crit[j]) a0c += a[j];
    for (int s = 0; s < m->S; s++) {
        uint32_t xs = X(x, g->N, s, v);
        double d = m->d[c * m->S + s];
        if (d > 0.0 && !hop_active(m, p, d, xs)) a0c += (d * (double)xs) * W;
    }
    uint32_t key[2] = {(uint32_t)p->seed, (uint32_t)(p->seed >> 32)};
    uint32_t cc[4] = {v, SELF_ID, st->step, ctr3(real, 0, P_CRIT, 0)};
    uint32_t o[4];
    orc_philox4x32_10(cc, key, o);
    double target = uni(o[0]) * a0c;
    double acc = 0.0;
    int kind = 0, idx = -1;
    for (int j = 0; j < m->R && kind == 0; j++) {
        if (!crit[j] || !(a[j] > 0.0)) continue;
        acc += a[j];
        idx = j;
        if (target < acc) kind = 1;
    }
    int lastkind = idx >= 0 ? 1 : 0;
    for (int s = 0; s < m->S && kind == 0; s++) {
        uint32_t xs = X(x, g->N, s, v);
        double d = m->d[c * m->S + s];
        if (!(d > 0.0) || hop_active(m, p, d, xs)) continue;
        double ao = (d * (double)xs) * W;
        if (!(ao > 0.0)) continue;
        acc += ao;
        idx = s; lastkind = 2;
        if (target < acc) kind = 2;
    }
    if (kind == 0) kind = lastkind;
    st->ev_kind = kind;
    st->ev_idx = idx;
    st->ev_dst = v;
    if (kind == 2) {
        double tw = uni(o[1]) * W;
        double accw = 0.0;
        uint32_t dst = v;
        for (uint32_t q = g->rp[v]; q < g->rp[v + 1]; q++) {
            accw += (double)g->w[q];
            dst = g->col[q];
            if (tw < accw) break;
        }
        st->ev_dst = dst;
    }
}

/* The critical event the step loop would choose at subvolume v of state x
 * for each of nsteps step indices (test access to choose_event): kind 1 =
 * reaction idx, kind 2 = hop of species idx to dst, 0 = none. */
int orc_critical_choice(const orc_network *net, const orc_topology *topo, const orc_params *p,
                        const uint32_t *x, uint32_t realization, uint32_t v, uint32_t nsteps,
                        const uint32_t *steps, int32_t *kind, int32_t *idx, uint32_t *dst)
{
    model_t m; csr_t g;
    if (model_build(net, &m)) { model_free(&m); return -1; }
    if (csr_build(topo, &g) || v >= g.N) { model_free(&m); csr_free(&g); return -1; }
    for (uint32_t i = 0; i < nsteps; i++) {
        ctl_t st;
        memset(&st, 0, sizeof st);
        st.vstar = v;
        st.step = steps[i];
        choose_event(&m, &g, p, x, realization, &st);
        kind[i] = st.ev_kind;
        idx[i] = st.ev_idx;
        dst[i] = st.ev_dst;
    }
    model_free(&m); csr_free(&g);
    return 0;
}

static void decide(ctl_t *st)
{
    double tau = fmin(st->tau_p, st->tau2);
    st->crit = (st->tau2 <= st->tau_p);
    st->last = 0;
    if (st->t + tau >= st->t_end) {
        tau = st->t_end - st->t;
        st->crit = 0;
        st->last = 1;
    }
    st->tau = tau;
}

/* Uniform word of leap draw w (hop: w = 2s + direction of edge {c0 < c1},
 * direction 0 for the hop c0 -> c1; reaction: w = j, c1 = SELF_ID): its top
 * byte is byte w & 15 (bytes 0..3 of word 0, then words 1..3) of the coarse
 * block c3base | (w >> 4), its low 24 bits word w & 3 of the fine block
 * c3base | (FINE_BLOCK + (w >> 2)), whose counter fctr receives (the PTRS
 * stream's).  The kernels settle a draw from its top byte alone when every
 * word with that byte gives 0 (plis_numerics.h, coarse_zero); the
 * draws themselves are the inversion / PTRS of this word. */
#define FINE_BLOCK 64u
static uint32_t leap_word(uint32_t c0, uint32_t c1, uint32_t step, uint32_t c3base, const uint32_t key[2],
                          uint32_t w, uint32_t fctr[4])
{
    uint32_t cc[4] = {c0, c1, step, c3base | (w >> 4)};
    uint32_t o[4];
    orc_philox4x32_10(cc, key, o);
    uint32_t hb = (o[(w >> 2) & 3u] >> (8u * (w & 3u))) & 0xFFu;
    fctr[0] = c0;
    fctr[1] = c1;
    fctr[2] = step;
    fctr[3] = c3base | (FINE_BLOCK + (w >> 2));
    orc_philox4x32_10(fctr, key, o);
    return (hb << 24) | (o[w & 3u] & 0xFFFFFFu);
}

/* one leap at v; returns 1 on a negative (or overflowing) population */
static int leap_v(const model_t *m, const csr_t *g, const orc_params *p, const uint32_t *x,
                  uint32_t *xn, uint32_t v, uint32_t real, const ctl_t *st, uint64_t *fire)
{
    double a[128];
    int crit[128];
    int64_t nx[64];
    react_props(m, g, p, x, v, a, crit);
    uint32_t N = g->N;
    int c = comp_of(g, v);
    uint32_t key[2] = {(uint32_t)p->seed, (uint32_t)(p->seed >> 32)};
    for (int s = 0; s < m->S; s++) nx[s] = X(x, N, s, v);
    uint64_t f = 0;
    double tau = st->tau;
    /* non-critical reactions */
    for (int j = 0; j < m->R; j++) {
        if (crit[j] || !(a[j] > 0.0)) continue;
        uint32_t fc[4];
        uint32_t wd = leap_word(v, SELF_ID, st->step, ctr3(real, st->attempt, P_REACT, 0), key, (uint32_t)j, fc);
        int64_t k = orc_poisson(tau * a[j], wd, fc, key, (uint32_t)j & 3u);
        f += (uint64_t)k;
        for (int s = 0; s < m->S; s++) nx[s] += k * (int64_t)m->nu[s * m->R + j];
    }
    /* out-hops */
    for (int s = 0; s < m->S; s++) {
        uint32_t xs = X(x, N, s, v);
        double d = m->d[c * m->S + s];
        if (!hop_active(m, p, d, xs)) continue;
        double dx = d * (double)xs;
        for (uint32_t q = g->rp[v]; q < g->rp[v + 1]; q++) {
            uint32_t u = g->col[q];
            uint32_t w = 2u * (uint32_t)s + (v > u ? 1u : 0u);
            uint32_t fc[4];
            uint32_t wd = leap_word(v < u ? v : u, v < u ? u : v, st->step, ctr3(real, st->attempt, P_HOP, 0),
                                    key, w, fc);
            double lam = tau * (dx * (double)g->w[q]);
            int64_t k = orc_poisson(lam, wd, fc, key, w & 3u);
            nx[s] -= k;
            f += (uint64_t)k;
        }
    }
    /* in-hops, regenerated from the sender's stream */
    for (int s = 0; s < m->S; s++) {
        for (uint32_t q = g->irp[v]; q < g->irp[v + 1]; q++) {
            uint32_t u = g->isrc[q];
            uint32_t xu = X(x, N, s, u);
            double du = m->d[comp_of(g, u) * m->S + s];
            if (!hop_active(m, p, du, xu)) continue;
            uint32_t w = 2u * (uint32_t)s + (u > v ? 1u : 0u);
            uint32_t fc[4];
            uint32_t wd = leap_word(v < u ? v : u, v < u ? u : v, st->step, ctr3(real, st->attempt, P_HOP, 0),
                                    key, w, fc);
            double lam = tau * ((du * (double)xu) * (double)g->iw[q]);
            int64_t k = orc_poisson(lam, wd, fc, key, w & 3u);
            nx[s] += k;
        }
    }
    /* the critical event */
    if (st->crit) {
        if (v == st->vstar) {
            if (st->ev_kind == 1) {
                for (int s = 0; s < m->S; s++) nx[s] += m->nu[s * m->R + st->ev_idx];
                f += 1;
            } else if (st->ev_kind == 2) {
                nx[st->ev_idx] -= 1;
                f += 1;
            }
        }
        if (st->ev_kind == 2 && v == st->ev_dst && st->ev_dst != st->vstar) nx[st->ev_idx] += 1;
    }
    int bad = 0;
    for (int s = 0; s < m->S; s++) {
        if (nx[s] < 0 || nx[s] > 0xFFFFFFFFll) bad = 1;
        xn[(size_t)s * N + v] = (uint32_t)nx[s];
    }
    *fire = f;
    return bad;
}

static int run_one(const model_t *m, const csr_t *g, const orc_params *p, uint32_t real,
                   uint32_t *x, double t_end, double t_pause, uint64_t step_limit,
                   orc_stats *stats, int par)
{
    uint32_t N = g->N;
    size_t SN = (size_t)m->S * N;
    uint32_t *xn = malloc(SN * sizeof(uint32_t) + 4);
    double *tp = malloc((size_t)N * sizeof(double) + 8);
    double *t2 = malloc((size_t)N * sizeof(double) + 8);
    int *bad = malloc((size_t)N * sizeof(int) + 4);
    uint64_t *fire = malloc((size_t)N * sizeof(uint64_t) + 8);
    ctl_t st;
    memset(&st, 0, sizeof st);
    st.t = stats->t;
    st.t_end = t_end;
    st.step = (uint32_t)stats->steps;
    int rc = 0;
    if (stats->status == 2) goto out;
    while (st.t < t_end && stats->steps < step_limit && st.t < t_pause) {
        /* phase A */
#pragma omp parallel for schedule(static) if (par)
        for (int64_t vv = 0; vv < (int64_t)N; vv++) {
            uint32_t v = (uint32_t)vv;
            double a0c;
            tau_v(m, g, p, x, v, &tp[v], &a0c);
            t2[v] = tau2_of(a0c, p, real, st.step, v);
        }
        double tpg = INFINITY, t2g = INFINITY;
        uint32_t vs = 0xFFFFFFFFu;
        for (uint32_t v = 0; v < N; v++) {
            tpg = fmin(tpg, tp[v]);
            if (t2[v] < t2g) { t2g = t2[v]; vs = v; }
        }
        if (tpg == INFINITY && t2g == INFINITY) { stats->status = 2; break; }
        st.tau_p = tpg; st.tau2 = t2g; st.vstar = vs; st.attempt = 0;
        st.ev_kind = 0;
        decide(&st);
        /* needed only if tau'' <= tau': a leap-back halves tau' (and may lift
         * the clamp to t_end) but never makes a larger tau'' fire */
        if (t2g < INFINITY && t2g <= tpg) choose_event(m, g, p, x, real, &st);
        for (;;) {
            /* phase B */
#pragma omp parallel for schedule(static) if (par)
            for (int64_t vv = 0; vv < (int64_t)N; vv++) {
                uint32_t v = (uint32_t)vv;
                bad[v] = leap_v(m, g, p, x, xn, v, real, &st, &fire[v]);
            }
            int anybad = 0;
            uint64_t f = 0;
            for (uint32_t v = 0; v < N; v++) { anybad |= bad[v]; f += fire[v]; }
            if (!anybad) {
                memcpy(x, xn, SN * sizeof(uint32_t));
                st.t = st.last ? st.t_end : st.t + st.tau;
                st.step++;
                stats->steps++;
                stats->firings += f;
                stats->subvolume_leaps += N;
                if (st.crit) stats->crit_events++;
                break;
            }
            stats->leapbacks++;
            st.attempt++;
            if (st.attempt > MAX_ATTEMPT) { rc = -3; stats->status = -3; goto out; }
            st.tau_p = st.tau_p * 0.5;
            decide(&st);
        }
    }
    if (stats->status == 0 && (st.t >= t_end || stats->steps >= step_limit)) stats->status = 1;
out:
    stats->t = st.t;
    free(xn); free(tp); free(t2); free(bad); free(fire);
    return rc;
}

int orc_run(const orc_network *net, const orc_topology *topo, const orc_params *p,
            uint32_t realization, uint32_t *x, double t_end, double t_pause,
            uint64_t step_limit, orc_stats *stats)
{
    model_t m; csr_t g;
    if (realization > 0xFFFFu) return -1;
    if (model_build(net, &m)) { model_free(&m); return -1; }
    if (csr_build(topo, &g)) { model_free(&m); csr_free(&g); return -1; }
    int rc = run_one(&m, &g, p, realization, x, t_end, t_pause, step_limit, stats, 1);
    model_free(&m); csr_free(&g);
    return rc;
}

int orc_run_batch(const orc_network *net, int32_t nreal, const orc_topology *topos,
                  const orc_params *p, const uint32_t *ids, uint32_t *const *x,
                  double t_end, uint64_t step_limit, orc_stats *stats)
{
    model_t m;
    if (model_build(net, &m)) { model_free(&m); return -1; }
    int rc = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(min : rc)
    for (int32_t i = 0; i < nreal; i++) {
        csr_t g;
        int r = csr_build(&topos[i], &g);
        if (!r) r = run_one(&m, &g, p, ids ? ids[i] : (uint32_t)i, x[i], t_end, INFINITY,
                            step_limit, &stats[i], 0);
        csr_free(&g);
        if (r < rc) rc = r;
    }
    model_free(&m);
    return rc;
}

/* ------------------------------------------------------------------ */
/* exact SSA: direct method with a sum tree over subvolumes            */
/* ------------------------------------------------------------------ */
typedef struct { uint32_t n; double *t; } sumtree_t;

static void st_set(sumtree_t *T, uint32_t i, double v)
{
    uint32_t k = i + T->n;
    T->t[k] = v;
    for (k >>= 1; k >= 1; k >>= 1) T->t[k] = T->t[2 * k] + T->t[2 * k + 1];
}

static uint32_t st_find(const sumtree_t *T, double target)
{
    uint32_t k = 1;
    while (k < T->n) {
        if (target < T->t[2 * k] || T->t[2 * k + 1] <= 0.0) k = 2 * k;
        else { target -= T->t[2 * k]; k = 2 * k + 1; }
    }
    return k - T->n;
}

static double total_rate_v(const model_t *m, const csr_t *g, const uint32_t *x, uint32_t v)
{
    int c = comp_of(g, v);
    double a = 0.0;
    for (int j = 0; j < m->R; j++) {
        double aj = m->k[c * m->R + j];
        for (int i = 0; i < m->nrs[j]; i++) aj *= binom_d(X(x, g->N, m->rs[j * 3 + i], v), m->rr[j * 3 + i]);
        a += aj;
    }
    for (int s = 0; s < m->S; s++) a += (m->d[c * m->S + s] * (double)X(x, g->N, s, v)) * g->W[v];
    return a;
}

static int ssa_one(const model_t *m, const csr_t *g, uint64_t seed, uint32_t real, uint32_t *x,
                   double t_end, uint64_t *events)
{
    uint32_t N = g->N, n = 1;
    while (n < N) n <<= 1;
    sumtree_t T = {n, calloc(2 * (size_t)n, sizeof(double))};
    for (uint32_t v = 0; v < N; v++) st_set(&T, v, total_rate_v(m, g, x, v));
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    double t = 0.0;
    uint64_t ev = 0;
    for (;;) {
        double a0 = T.t[1];
        if (!(a0 > 0.0)) break;
        uint32_t cc[4] = {(uint32_t)ev, (uint32_t)(ev >> 32), 0u, ctr3(real, 0, P_SSA, 0)};
        uint32_t o[4];
        orc_philox4x32_10(cc, key, o);
        t += -orc_log(uni(o[0])) / a0;
        if (t >= t_end) break;
        double target = uni(o[1]) * a0;
        uint32_t v = st_find(&T, target);
        /* channel within v */
        int c = comp_of(g, v);
        double av = total_rate_v(m, g, x, v);
        double tg = uni(o[2]) * av, acc = 0.0;
        int done = 0;
        for (int j = 0; j < m->R && !done; j++) {
            double aj = m->k[c * m->R + j];
            for (int i = 0; i < m->nrs[j]; i++) aj *= binom_d(X(x, N, m->rs[j * 3 + i], v), m->rr[j * 3 + i]);
            if (!(aj > 0.0)) continue;
            acc += aj;
            if (tg < acc) {
                for (int s = 0; s < m->S; s++) x[(size_t)s * N + v] = (uint32_t)((int64_t)x[(size_t)s * N + v] + m->nu[s * m->R + j]);
                done = 1;
            }
        }
        uint32_t dst = v;
        for (int s = 0; s < m->S && !done; s++) {
            double ao = (m->d[c * m->S + s] * (double)X(x, N, s, v)) * g->W[v];
            if (!(ao > 0.0)) continue;
            acc += ao;
            if (tg < acc) {
                double tw = uni(o[3]) * g->W[v], aw = 0.0;
                for (uint32_t q = g->rp[v]; q < g->rp[v + 1]; q++) {
                    aw += (double)g->w[q];
                    dst = g->col[q];
                    if (tw < aw) break;
                }
                x[(size_t)s * N + v] -= 1;
                x[(size_t)s * N + dst] += 1;
                done = 1;
            }
        }
        if (!done) { ev++; continue; } /* rounding: no channel chosen, redraw */
        st_set(&T, v, total_rate_v(m, g, x, v));
        if (dst != v) st_set(&T, dst, total_rate_v(m, g, x, dst));
        ev++;
    }
    free(T.t);
    if (events) *events = ev;
    return 0;
}

int orc_ssa_run(const orc_network *net, const orc_topology *topo, uint64_t seed,
                uint32_t realization, uint32_t *x, double t_end, uint64_t *events)
{
    model_t m; csr_t g;
    if (model_build(net, &m)) { model_free(&m); return -1; }
    if (csr_build(topo, &g)) { model_free(&m); csr_free(&g); return -1; }
    int rc = ssa_one(&m, &g, seed, realization, x, t_end, events);
    model_free(&m); csr_free(&g);
    return rc;
}

int orc_ssa_batch(const orc_network *net, const orc_topology *topo, uint64_t seed,
                  int32_t nreal, uint32_t *x_all, double t_end)
{
    model_t m; csr_t g;
    if (model_build(net, &m)) { model_free(&m); return -1; }
    if (csr_build(topo, &g)) { model_free(&m); csr_free(&g); return -1; }
    size_t SN = (size_t)m.S * g.N;
#pragma omp parallel for schedule(dynamic, 1)
    for (int32_t i = 0; i < nreal; i++) ssa_one(&m, &g, seed, (uint32_t)i, x_all + SN * (size_t)i, t_end, NULL);
    model_free(&m); csr_free(&g);
    return 0;
}

/* ------------------------------------------------------------------ */
/* learned-tau inputs: per-subvolume features and post-leap check      */
/* ------------------------------------------------------------------ */
int orc_features(const orc_network *net, const orc_topology *topo, const orc_params *p,
                 const uint32_t *x, float prev_logtau, float *feats)
{
    model_t m; csr_t g;
    if (model_build(net, &m)) { model_free(&m); return -1; }
    if (csr_build(topo, &g)) { model_free(&m); csr_free(&g); return -1; }
#pragma omp parallel for schedule(static)
    for (int64_t vv = 0; vv < (int64_t)g.N; vv++) {
        uint32_t v = (uint32_t)vv;
        double a[128];
        int crit[128];
        react_props(&m, &g, p, x, v, a, crit);
        int c = comp_of(&g, v);
        float *f = feats + (size_t)v * 12;
        int maxo = 0;
        for (int j = 0; j < m.R; j++)
            if (a[j] > 0.0 && m.order[j] > maxo) maxo = m.order[j];
        for (int s = 0; s < 5; s++) {
            if (s >= m.S) { f[s] = 0.0f; f[5 + s] = 0.0f; continue; }
            uint32_t xs = X(x, g.N, s, v);
            double A = 0.0;
            for (int j = 0; j < m.R; j++)
                if (net->reac[s * m.R + j] > 0) A += a[j];
            A += (m.d[c * m.S + s] * (double)xs) * g.W[v];
            f[s] = (float)orc_log(1.0 + (double)xs);
            f[5 + s] = (float)orc_log(1.0 + A);
        }
        f[10] = (float)maxo;
        f[11] = prev_logtau;
    }
    model_free(&m); csr_free(&g);
    return 0;
}

/* Post-leap leap condition in propensity form (Cao, Gillespie & Petzold
 * 2006, eq. 5, with their one-molecule floor; checked after the leap as in
 * Anderson 2008): every non-critical channel j of v with a_j(x) > 0 must
 * satisfy
 *     |a_j(x') - a_j(x)| <= theta * max(eps, phi_j(x)) * a_j(x),
 * phi_j(x) = max over reactants i of r_ij / x_i, the relative change one
 * molecule of i makes in a_j; a zero-order channel never changes.  An active
 * hop channel of species s (propensity d_s x_s w_e) changes by the factor
 * x'_s / x_s, so its condition is |x'_s - x_s| <= theta * max(eps x_s, 1).
 * nx holds v's S new counts (any sign); a negative count breaks it too. */
#define LC_SLACK (1.0 + 0x1p-30)

static int leap_condition_broken(const model_t *m, const csr_t *g, const orc_params *p,
                                 const uint32_t *x, uint32_t v, const int64_t *nx, double theta)
{
    double a[128];
    int crit[128];
    uint32_t xn[64];
    for (int s = 0; s < m->S; s++) {
        if (nx[s] < 0 || nx[s] > 0xFFFFFFFFll) return 1;
        xn[s] = (uint32_t)nx[s];
    }
    react_props(m, g, p, x, v, a, crit);
    int c = comp_of(g, v);
    const double eps = p->epsilon;
    for (int j = 0; j < m->R; j++) {
        if (crit[j] || !(a[j] > 0.0) || m->nrs[j] == 0) continue;
        double aj = m->k[c * m->R + j];
        for (int i = 0; i < m->nrs[j]; i++) aj = aj * binom_d(xn[m->rs[j * 3 + i]], m->rr[j * 3 + i]);
        /* D > theta max(eps, max_i r_i / x_i) a_j, without divides; the
         * relative slack 2^-30 keeps a change that meets a bound exactly (a
         * one-molecule change at the floor) on the holding side whatever the
         * rounding of the products */
        double D = fabs(aj - a[j]), t = (theta * a[j]) * LC_SLACK;
        int over = D > t * eps;
        for (int i = 0; i < m->nrs[j]; i++)
            over = over && D * (double)X(x, g->N, m->rs[j * 3 + i], v) > t * (double)m->rr[j * 3 + i];
        if (over) return 1;
    }
    for (int s = 0; s < m->S; s++) {
        uint32_t xs = X(x, g->N, s, v);
        if (!hop_active(m, p, m->d[c * m->S + s], xs)) continue;
        if (fabs((double)xn[s] - (double)xs) > (theta * fmax(eps * (double)xs, 1.0)) * LC_SLACK) return 1;
    }
    return 0;
}

int orc_violations(const orc_network *net, const orc_topology *topo, const orc_params *p,
                   const uint32_t *x, const uint32_t *xn, double theta, uint8_t *flags)
{
    model_t m; csr_t g;
    if (model_build(net, &m)) { model_free(&m); return -1; }
    if (csr_build(topo, &g)) { model_free(&m); csr_free(&g); return -1; }
    int64_t count = 0;
#pragma omp parallel for schedule(static) reduction(+ : count)
    for (int64_t vv = 0; vv < (int64_t)g.N; vv++) {
        uint32_t v = (uint32_t)vv;
        int64_t nx[64];
        for (int s = 0; s < m.S; s++) nx[s] = X(xn, g.N, s, v);
        int bad = leap_condition_broken(&m, &g, p, x, v, nx, theta);
        flags[v] = (uint8_t)bad;
        count += bad;
    }
    model_free(&m); csr_free(&g);
    return (int)count;
}

/* Training records of state x: per subvolume the features, log of Cao's
 * tau'_v (capped at 1e6) and the accept/reject label of a trial leap of v
 * alone with tau = tau'_v: its non-critical reactions, out-hops and in-hops
 * (at the neighbours' propensities) are drawn from the TRIAL streams
 * (counter (v, SELF_ID or u, step, purpose TRIAL), block 96 + j/4 for
 * reactions, 64 + s/2 for hops, word 2(s&1) + 0 out / 1 in), and label = 1 where the
 * leap condition above is broken at v (reject), 0 where it holds.  A
 * subvolume without non-critical channels (tau' infinite) is labelled 0. */
int orc_records(const orc_network *net, const orc_topology *topo, const orc_params *p,
                const uint32_t *x, uint32_t realization, uint32_t step, float prev_logtau,
                double theta, float *feats, float *ylog, float *label)
{
    if (orc_features(net, topo, p, x, prev_logtau, feats)) return -1;
    model_t m; csr_t g;
    if (model_build(net, &m)) { model_free(&m); return -1; }
    if (csr_build(topo, &g)) { model_free(&m); csr_free(&g); return -1; }
    const uint32_t key[2] = {(uint32_t)p->seed, (uint32_t)(p->seed >> 32)};
    const uint32_t N = g.N;
#pragma omp parallel for schedule(static)
    for (int64_t vv = 0; vv < (int64_t)N; vv++) {
        uint32_t v = (uint32_t)vv;
        double tp, a0c;
        tau_v(&m, &g, p, x, v, &tp, &a0c);
        ylog[v] = tp < 1e6 ? (float)orc_log(tp) : (float)orc_log(1e6);
        if (!(tp < INFINITY)) { label[v] = 0.0f; continue; }
        double a[128];
        int crit[128];
        int64_t nx[64];
        react_props(&m, &g, p, x, v, a, crit);
        int c = comp_of(&g, v);
        for (int s = 0; s < m.S; s++) nx[s] = X(x, N, s, v);
        for (int j = 0; j < m.R; j++) {
            if (crit[j] || !(a[j] > 0.0)) continue;
            uint32_t cc[4] = {v, SELF_ID, step, ctr3(realization, 0, P_TRIAL, 96u + ((uint32_t)j >> 2))};
            uint32_t o[4];
            orc_philox4x32_10(cc, key, o);
            int64_t k = orc_poisson(tp * a[j], o[j & 3], cc, key, (uint32_t)j & 3u);
            for (int s = 0; s < m.S; s++) nx[s] += k * (int64_t)m.nu[s * m.R + j];
        }
        for (int s = 0; s < m.S; s++) {
            uint32_t xs = X(x, N, s, v);
            double d = m.d[c * m.S + s];
            if (!hop_active(&m, p, d, xs)) continue;
            double dx = d * (double)xs;
            for (uint32_t q = g.rp[v]; q < g.rp[v + 1]; q++) {
                uint32_t u = g.col[q];
                uint32_t wd = 2u * ((uint32_t)s & 1u);
                uint32_t cc[4] = {v, u, step, ctr3(realization, 0, P_TRIAL, 64u + ((uint32_t)s >> 1))};
                uint32_t o[4];
                orc_philox4x32_10(cc, key, o);
                nx[s] -= orc_poisson(tp * (dx * (double)g.w[q]), o[wd], cc, key, wd);
            }
        }
        for (uint32_t q = g.irp[v]; q < g.irp[v + 1]; q++) {
            uint32_t u = g.isrc[q];
            int cu = comp_of(&g, u);
            for (int s = 0; s < m.S; s++) {
                uint32_t xu = X(x, N, s, u);
                double du = m.d[cu * m.S + s];
                if (!hop_active(&m, p, du, xu)) continue;
                uint32_t wd = 2u * ((uint32_t)s & 1u) + 1u;
                uint32_t cc[4] = {v, u, step, ctr3(realization, 0, P_TRIAL, 64u + ((uint32_t)s >> 1))};
                uint32_t o[4];
                orc_philox4x32_10(cc, key, o);
                nx[s] += orc_poisson(tp * ((du * (double)xu) * (double)g.iw[q]), o[wd], cc, key, wd);
            }
        }
        label[v] = leap_condition_broken(&m, &g, p, x, v, nx, theta) ? 1.0f : 0.0f;
    }
    model_free(&m); csr_free(&g);
    return 0;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
