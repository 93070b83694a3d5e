/*
 * busseq/model.h -- the BUSseq model as used by the sampler: prior constants,
 * random-stream phases, the per-element log-likelihood-ratio terms and the
 * per-parameter Metropolis-Hastings prior/proposal terms.
 *
 * Shared by the gfx950 kernels and by host code, so that each term is one
 * expression with one evaluation order everywhere (compile with
 * -ffp-contract=off).  The per-element terms exist in two precisions:
 * single (_f32, the mixed-precision hot path) and double (_f64).
 *
 * Notation: cell i in batch b with subtype k = W_i, gene g, count x = X_ig,
 * mean mu = exp(alpha_g + beta_kg + nu_bg + delta_i), size phi = phi_bg.
 * A random-walk proposal shifts log(mu) by dlt; c = expm1(dlt).  With
 * r = mu / (phi + mu),
 *   log NB(x; mu e^dlt, phi) - log NB(x; mu, phi) = x*dlt - (phi+x)*log1p(r*c),
 * which is the quantity the reference fills per element in fill_prop_alpha,
 * fill_prop_beta, fill_prop_nu, fill_prop_delta and fill_prop_W
 * (BUSseq_gpu.cu:608-623, 676-694, 715-728, 746-761, 831-844), rewritten in a
 * form that keeps full relative precision in single precision.
 */
#ifndef BUSSEQ_MODEL_H
#define BUSSEQ_MODEL_H

#include "numerics.h"

/* hyperparameters: BUSseq_gpu.cu:12-32 */
#define BQ_P_INIT_BOUND 0.5
#define BQ_TAU0_INIT 0.005
#define BQ_GAMMA0_INIT 0.0
#define BQ_GAMMA1_INIT (-0.3)
#define BQ_PHI_INIT 5.0
#define BQ_XI_PI 2.0
#define BQ_SIGMA2_ALPHA 5.0
#define BQ_SIGMA2_NU 5.0
#define BQ_SIGMA2_DELTA 5.0
#define BQ_SIGMA2_GAMMA0 3.0
#define BQ_KAPPA_GAMMA1 0.001
#define BQ_TAU_GAMMA1 0.01
#define BQ_KAPPA_PHI 1.0
#define BQ_TAU_PHI 0.1
#define BQ_A_P 1.0
#define BQ_B_P 3.0
#define BQ_A_TAU0 2.0
#define BQ_B_TAU0 0.01
#define BQ_TAU1 100.0
#define BQ_RW_SD 0.1          /* random-walk proposal sd (cu:518, 605, 673, 712, 742) */
#define BQ_LOG10 2.302585092994046

/* Bins of the per-batch histogram of X used by the gamma updates; counts at
 * or above this value contribute their gamma terms directly. */
#define BQ_HIST 1024

/* random-stream phases: counter word 3 = phase << 20 | block */
enum {
    BQ_PH_INIT_W = 1,   /* (i, 0)   W0 categorical draw                   */
    BQ_PH_INIT_L = 2,   /* (g, k)   L0 Bernoulli draw                     */
    BQ_PH_INIT_P = 3,   /* (0, 0)   p0 ~ U(0, 0.5)                        */
    BQ_PH_W = 8,        /* (i, 0)   subtype categorical draw              */
    BQ_PH_PI = 9,       /* (b, 0)   Dirichlet draw of pi_b                */
    BQ_PH_XZ = 10,      /* (i, g)   Z draw, NB proposal, MH uniform       */
    BQ_PH_ALPHA_P = 11, BQ_PH_ALPHA_A = 12,   /* (g, 0)                   */
    BQ_PH_GAMMA0_P = 13, BQ_PH_GAMMA0_A = 14, /* (b, 0)                   */
    BQ_PH_GAMMA1_P = 15, BQ_PH_GAMMA1_A = 16, /* (b, 0)                   */
    BQ_PH_L = 17,       /* (g, k)                                         */
    BQ_PH_PTAU = 18,    /* (0, 0)   p then tau0                           */
    BQ_PH_BETA_P = 19, BQ_PH_BETA_A = 20,     /* (g, k)                   */
    BQ_PH_NU_P = 21, BQ_PH_NU_A = 22,         /* (g, b)                   */
    BQ_PH_DELTA_P = 23, BQ_PH_DELTA_A = 24,   /* (i, 0)                   */
    BQ_PH_PHI_P = 25, BQ_PH_PHI_A = 26,       /* (g, b)                   */
    BQ_PH_SIM_PARAM = 40, BQ_PH_SIM_CELL = 41, BQ_PH_SIM_COUNT = 42
};
#define BQ_SWEEP_INIT 0xFFFFFFFFu

/* ------------------------------------------------------------------------ */
/* per-element terms                                                          */
/* ------------------------------------------------------------------------ */
#define BQ_DEFINE_ELEMENT_TERMS(T, SFX, LOG1P, LGDIFF, FMA)                         \
/* mean: ((e^alpha * e^beta) * e^nu) * e^delta */                                  \
BQ_FN T bq_mu_##SFX(T ea, T eb, T en, T ed) { return ((ea * eb) * en) * ed; }     \
/* subtype-k logit relative to subtype 0: x*beta_k - (phi+x)*log1p(r0*expm1(beta_k)) */ \
BQ_FN T bq_tw_gen_##SFX(T x, T phi, T r0, T beta, T em)                            \
{ return FMA(-(phi + x), LOG1P(r0 * em), x * beta); }                             \
/* random-walk MH term for a log-mean shift dlt, c = expm1(dlt); l = log1p(r*c) */ \
BQ_FN T bq_tmh_l_##SFX(T x, T phi, T dlt, T l) { return FMA(-(phi + x), l, x * dlt); } \
BQ_FN T bq_tmh_##SFX(T x, T phi, T r, T dlt, T c)                                  \
{ return bq_tmh_l_##SFX(x, phi, dlt, LOG1P(r * c)); }                              \
/* shift dlt applied on top of an accepted shift whose log1p term is l1:           \
 * c12 = expm1(dlt1 + dlt) */                                                       \
BQ_FN T bq_tmh2_##SFX(T x, T phi, T r, T dlt, T c12, T l1)                         \
{ return FMA(-(phi + x), LOG1P(r * c12) - l1, x * dlt); }                         \
/* phi MH term per cell (BUSseq_gpu.cu:799-802) rewritten as                        \
 *   [lgamma(phi'+x)-lgamma(phi+x)] - [lgamma(phi')-lgamma(phi)] - x*log(phi'/phi)  \
 *   + (phi+x)*log1p(d*mu/(phi*(phi'+mu))) - d*log1p(mu/phi'),  d = phi'-phi.       \
 * dop = d/phi, iphip = 1/phi', lrat = log(phi'/phi), dlg0 = lgamma(phi')-lgamma(phi) */ \
BQ_FN T bq_tphi_gen_##SFX(int xi, T x, T phi, T phip, T mu, T d, T dop, T iphip, T lrat, T dlg0) \
{                                                                                   \
    T l1 = LOG1P((dop * mu) / (phip + mu));                                        \
    T l2 = LOG1P(mu * iphip);                                                      \
    T t = FMA(-d, l2, FMA(phi + x, l1, -(x * lrat)));                              \
    if (xi > 0) t = t + (LGDIFF(phi + x, d) - dlg0);                               \
    return t;                                                                      \
}

/* r = mu / (phi + mu).  In single precision as mu times the reciprocal of
 * phi + mu (two roundings instead of the quotient's one, a relative error
 * below 2^-22): bq_rcpf's four instructions and a minimum in place of the
 * division's eleven, in the per-element hot path of every streaming pass.
 * mu (1/z) <= 1 holds by itself while z = phi + mu >= mu is not held
 * (the product is below 1 + 2^-24); the minimum keeps r <= 1, which the
 * subtype term needs (r expm1(beta) > -1), also for mu above bq_rcpf's hold
 * at 2^125. */
BQ_FN float bq_ratio_f32(float mu, float phi) { return __builtin_fminf(mu * bq_rcpf(phi + mu), 1.0f); }
BQ_FN double bq_ratio_f64(double mu, double phi) { return mu / (phi + mu); }
/* The single-precision random-walk terms' logarithm log1p(r c): r = mu/(phi + mu)
 * in [0, 1] and c = expm1 of one proposal step (alpha, beta, nu, delta) or of
 * two (beta and nu together).  A step is BQ_RW_SD times a normal deviate of
 * bq_normal, whose magnitude is at most ndtri(2^-33) = 6.34 (it inverts a
 * uniform in [2^-33, 1 - 2^-33]), so |step| <= 0.634 and r c lies in
 * (-0.72, 2.56): inside bq_log1pf_mid's range, which needs no branch. */
BQ_FN float bq_log1p_mh_f32(float y) { return bq_log1pf_mid(y); }

BQ_DEFINE_ELEMENT_TERMS(float, f32, bq_log1p_mh_f32, bq_lgamma_diff32, __builtin_fmaf)
BQ_DEFINE_ELEMENT_TERMS(double, f64, bq_log1p, bq_lgamma_diff64, __builtin_fma)

/* The double-precision subtype logit and phi terms take the branch-free
 * bq_log1p_mid: their arguments (r0 expm1(beta) > -1; d mu / (phi (phi' + mu))
 * > -1; mu / phi' > 0) often leave bq_log1p's fast range, where a branch would
 * run both paths for the wave; the random-walk terms' arguments r c (|c| about
 * 0.1) stay inside it and keep bq_log1p. */
BQ_FN double bq_tw_f64(double x, double phi, double r0, double beta, double em)
{ return __builtin_fma(-(phi + x), bq_log1p_mid(r0 * em), x * beta); }

/* The single-precision subtype logit term takes expm1(beta) held to
 * [-1 + 2^-24, 2^100] (it moves only beta beyond about +-69), so that r0
 * expm1(beta) with 0 <= r0 <= 1 stays where the branch-free log1pf holds; the
 * value is clamped once, where the parameter copy is written (bq_tw_em_f32),
 * not per cell x gene. */
BQ_FN float bq_tw_clamp_f32(float em) { return __builtin_fminf(__builtin_fmaxf(em, -0.99999994f), 0x1p100f); }
BQ_FN float bq_tw_em_f32(double beta) { return bq_tw_clamp_f32((float)bq_expm1(beta)); }
BQ_FN float bq_tw_f32(float x, float phi, float r0, float beta, float emc)
{
    return __builtin_fmaf(-(phi + x), bq_log1pf_mid(r0 * emc), x * beta);
}

/*
 * The phi term of one cell (BUSseq_gpu.cu:799-802) in two parts, each summed on
 * its own:
 *   bq_tphi_nb: (phi+x)*log1p(d*mu/(phi*(phi'+mu))) - d*log1p(mu/phi') - x*log(phi'/phi),
 *     for every cell, summed per tile like the other MH terms;
 *   bq_tphi_lg: [lgamma(phi'+x) - lgamma(phi+x)] - [lgamma(phi') - lgamma(phi)], for a
 *     count x >= 1.  It depends on x, phi and phi' only, and each count's value enters
 *     the exact fixed-point sum of its (batch, gene) separately, so the counts equal to
 *     v may be added as their number times the value for v.
 * The single-precision lgamma part takes the Stirling difference directly where both
 * arguments are at least 1 (both reciprocals from one, bq_rcpf) and bq_lgamma_diff32
 * otherwise; the single-precision NB part takes 1/(phi' + mu) from bq_rcpf.
 */
BQ_FN double bq_tphi_nb_f64(double x, double phi, double phip, double mu, double d, double dop, double iphip,
                            double lrat)
{
    double l1 = bq_log1p_mid((dop * mu) / (phip + mu));
    double l2 = bq_log1p_mid(mu * iphip);
    return __builtin_fma(-d, l2, __builtin_fma(phi + x, l1, -(x * lrat)));
}

BQ_FN double bq_tphi_lg_f64(double x, double phi, double d, double dlg0)
{
    return bq_lgamma_diff64(phi + x, d) - dlg0;
}

BQ_FN float bq_tphi_nb_f32(float x, float phi, float phip, float mu, float d, float dop, float iphip, float lrat)
{
    const float l1 = bq_log1pf((dop * mu) * bq_rcpf(phip + mu));
    const float l2 = bq_log1pf(mu * iphip);
    return __builtin_fmaf(-d, l2, __builtin_fmaf(phi + x, l1, -(x * lrat)));
}

BQ_FN float bq_tphi_lg_f32(float x, float phi, float d, float dlg0)
{
    const float z = phi + x, zp = z + d;
    float dl;
    if (z >= 1.0f && zp >= 1.0f) {
        const float w = bq_rcpf(z * zp);   /* 1/(z zp) */
        const float iz = zp * w, izp = z * w;
        const float core = __builtin_fmaf(d, bq_logf_norm(zp), __builtin_fmaf(z - 0.5f, bq_log1pf(d * iz), -d));
        dl = core + (bq_stirling_s32(izp) - bq_stirling_s32(iz));
    } else {
        dl = bq_lgamma_diff32(z, d);
    }
    return dl - dlg0;
}

/* ------------------------------------------------------------------------ */
/* per-parameter MH terms (double precision)                                  */
/* ------------------------------------------------------------------------ */
/* Gaussian prior N(m, s2): log p(new) - log p(old)  (cu:628, 733, 773, 581)  */
BQ_FN double bq_lr_normal(double old_v, double new_v, double m, double s2)
{
    double a = old_v - m, b = new_v - m;
    return (a * a - b * b) / (2.0 * s2);
}

/* beta_kg prior: spike (tau0) or slab (tau1) by L (cu:700-703) */
BQ_FN double bq_lr_beta(double old_v, double new_v, int L, double tau0)
{
    return (old_v * old_v - new_v * new_v) / (2.0 * (L ? BQ_TAU1 : tau0));
}

/* -gamma1 ~ Gamma(kappa, rate tau) prior with Gamma(10 g, scale 0.1) proposal (cu:591-596);
 * g = -gamma1, gn = -gamma1' */
BQ_FN double bq_lr_gamma1(double g, double gn)
{
    double lg = bq_log(g), lgn = bq_log(gn);
    return (BQ_KAPPA_GAMMA1 - 1.0) * (lgn - lg)
         + BQ_TAU_GAMMA1 * (g - gn)
         - bq_lgamma(10.0 * gn) + (10.0 * gn - 1.0) * lg + 10.0 * gn * (BQ_LOG10 + 1.0)
         + bq_lgamma(10.0 * g) - (10.0 * g - 1.0) * lgn - 10.0 * g * (BQ_LOG10 + 1.0);
}

/* phi ~ Gamma(kappa, rate tau) prior with Gamma(phi, 1) proposal (cu:809-814) */
BQ_FN double bq_lr_phi(double phi, double phin)
{
    return (BQ_KAPPA_PHI - phi) * bq_log(phin) - (BQ_KAPPA_PHI - phin) * bq_log(phi)
         + (1.0 - BQ_TAU_PHI) * (phin - phi) + (bq_lgamma(phi) - bq_lgamma(phin));
}

/* spike-and-slab log odds of L_kg = 1 (cu:640-642) */
BQ_FN double bq_logodds_L(double p, double tau0, double beta)
{
    return bq_log(p) - bq_log(1.0 - p) - (bq_log(BQ_TAU1) - bq_log(tau0)) / 2.0
         + (beta * beta) * (1.0 / (2.0 * tau0) - 1.0 / (2.0 * BQ_TAU1));
}

/* Bernoulli decision with log odds lo: draw u, L = u > 1/(1+e^lo) (cu:643) */
BQ_FN int bq_bern_logodds(double u, double lo)
{
    return u > 1.0 / (1.0 + bq_exp(lo));
}

/* 1 + e^{-(g0 + g1 x)}: a factor of the X MH ratio (cu:504).  The exponent is
 * clamped at 700 so that neither the factor nor u times it overflows; the
 * clamp changes the ratio only where g0 + g1 x < -700. */
BQ_FN double bq_xz_odds(double g0, double g1, double x)
{
    double e = -(g0 + g1 * x);
    if (e > 700.0) e = 700.0;
    return 1.0 + bq_exp(e);
}

/*
 * One X/Z update of an entry with Y = 0 and encoded state s (cu:484-512): Z is
 * redrawn when X = 0 (P(Z = 1) = 1 - qz, qz = 1/(1 + e^{g0})); with Z = 1 a
 * proposal x' ~ NB(mu, phi) is accepted with probability
 * min(1, (1 + e^{-g0-g1 X}) / (1 + e^{-g0-g1 x'})), tested as
 * u (1 + e^{-g0-g1 x'}) <= 1 + e^{-g0-g1 X}.  Returns the new encoded state
 * (0, or ~X with Z = 1).
 */
BQ_FN int32_t bq_xz_update(int32_t s, double mu, double phi, double g0, double g1, double qz, bq_rng* r)
{
    int xo = s < 0 ? ~s : 0;
    int z = xo == 0 ? (bq_uniform(r) > qz) : 1;
    if (!z) return 0;
    int xp = bq_rnbinom_xz(r, mu, phi);
    int xi = bq_uniform(r) * bq_xz_odds(g0, g1, (double)xp) <= bq_xz_odds(g0, g1, (double)xo) ? xp : xo;
    return ~xi;
}

/* per-histogram-bin gamma terms: sp(g0 + g1 x) */
BQ_FN double bq_sp_lin(double g0, double g1, double x) { return bq_softplus(g0 + g1 * x); }

/*
 * BIC building block: log-likelihood of one count y of cell i, gene g under
 * subtype k with the dropout marginalised (BUSseq_gpu.cu:967-1007):
 *   y > 0: log P(Z=0 | X=y) + log NB(y; mu, phi)
 *   y = 0: log sum_{x < floor(3 mu)} P(Z=1 | x) NB(x; mu, phi), with the x = 0
 *          term NB(0) and the reference's truncation of the sum at floor(3 mu);
 * log p and log(1-p) are clipped at -250 as in the reference.  The lgamma
 * ratio of the NB pmf is carried by the recurrence
 *   c_x = c_{x-1} + log((phi + x - 1) / x),  c_0 = 0.
 */
BQ_FN double bq_bic_elem(double a, double bk, double nb, double dl, double g0, double g1,
                         double phi, int y)
{
    double mu = bq_exp(((a + bk) + nb) + dl);
    double pb = mu / (mu + phi);
    double logp, log1mp;
    double lim = bq_exp(-250.0);
    if (pb < lim) { logp = -250.0; log1mp = bq_log(1.0 - pb); }
    else if (1.0 - pb < lim) { logp = bq_log(pb); log1mp = -250.0; }
    else { logp = bq_log(pb); log1mp = bq_log(1.0 - pb); }
    if (y > 0) {
        return -bq_softplus(g0 + g1 * y) + bq_lgamma(phi + y) - bq_lgamma(1.0 + y) - bq_lgamma(phi)
               + y * logp + phi * log1mp;
    }
    double xm = 3.0 * mu;
    /* a NaN mean (a chain that has left the finite range) sums nothing here
     * instead of running to the 2e9 cap: the element is NaN either way */
    int xmax = xm < 2.0e9 ? (int)xm : (xm >= 2.0e9 ? 2000000000 : 0);
    double s = phi * log1mp, c = 0.0;
    for (int x = 1; x < xmax; ++x) {
        c += bq_log((phi + (x - 1)) / (double)x);
        double lin = g0 + g1 * x;
        double tmp = (lin - bq_softplus(lin)) + c + x * logp + phi * log1mp;
        if (tmp > s) s = tmp + bq_log1p(bq_exp(s - tmp));
        else s = s + bq_log1p(bq_exp(tmp - s));
    }
    return s;
}

#endif /* BUSSEQ_MODEL_H */
