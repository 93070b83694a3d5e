/*
 * busseq/numerics.h -- the numerics contract shared by the gfx950 kernels and
 * the host code: Philox4x32-10 counter RNG, software exp/log/log1p/lgamma built
 * only from IEEE-exact operations (+ - * / fma, integer bit operations), the
 * samplers (uniform, normal, gamma, Poisson, negative binomial, beta) and the
 * exact fixed-point accumulator used for every long MH sum.
 *
 * Everything here is written once and compiled twice: by hipcc for gfx950
 * device code and by gcc/g++ for host code.  Both compilers must be run with
 * floating-point contraction off and without fast-math, so that every
 * function returns the same bits on the CPU and on the GPU.  That is what
 * makes the integer draws (X, Z, W) reproducible bit for bit across devices,
 * kernel schedules and chain batching.
 *
 * Published methods used (no vendor math library is reproduced):
 *   Philox4x32-10        Salmon, Moraes, Dror, Shaw, SC'11 (Random123).
 *   normal deviates      inverse CDF, Wichura 1988, Algorithm AS 241 (PPND16).
 *   gamma deviates       Marsaglia & Tsang 2000, with the U^(1/a) boost for a<1.
 *   Poisson deviates     sequential inversion (lambda < 10), Hoermann 1993 PTRS.
 *   negative binomial    sequential inversion of the CDF (small means), else
 *                        the gamma-Poisson mixture.
 *   log / exp            Cody-Waite argument reduction; polynomials fitted by
 *                        busseq_gpu_amd/csrc/gen_coeffs.py (Chebyshev, mpmath).
 *   lgamma               shift recurrence + Stirling series (Bernoulli numbers).
 *
 * The reference samplers this replaces are cuRAND's XORWOW stream and the
 * Ahrens-Dieter gamma in BUSseq_gpu.cu:34-218; their streams are not
 * reproduced (only their distributions).
 */
#ifndef BUSSEQ_NUMERICS_H
#define BUSSEQ_NUMERICS_H

#include <stdint.h>

#ifdef __HIPCC__
#define BQ_FN static inline __host__ __device__ __attribute__((always_inline))
#else
#define BQ_FN static inline __attribute__((always_inline))
#endif

/* A constant table read by host and device code: in the device's constant
 * memory, so that a read at a compile-time index is a scalar load.  Polynomial
 * coefficients taken from such a table enter each fused multiply-add as a
 * scalar-register operand, instead of occupying vector registers.  Only the
 * device pass of a HIP source declares it __constant__: the host pass's copy
 * of a __constant__ variable does not hold the initialiser, so host code there
 * reads an ordinary initialised array. */
#ifdef __HIP_DEVICE_COMPILE__
#define BQ_TABLE static __constant__ const
#else
#define BQ_TABLE static const
#endif

/* An int that holds the same value in every active lane of a wave (the step
 * counter of a search loop that all lanes enter together).  Device code reads
 * it from the first active lane so that table reads indexed by it become
 * scalar loads; host code uses it as is.  The value is the same either way. */
#ifdef __HIP_DEVICE_COMPILE__
#define BQ_WAVE_UNIFORM(k) __builtin_amdgcn_readfirstlane(k)
#else
#define BQ_WAVE_UNIFORM(k) (k)
#endif

#define BQ_INF (__builtin_inf())
#define BQ_NAN (__builtin_nan(""))
#define BQ_INFF (__builtin_inff())
#define BQ_NANF (__builtin_nanf(""))

/* ------------------------------------------------------------------------ */
/* bit casts                                                                 */
/* ------------------------------------------------------------------------ */
typedef union { double d; uint64_t u; int64_t i; } bq_d64;
typedef union { float f; uint32_t u; int32_t i; } bq_f32;

BQ_FN uint64_t bq_dbits(double x) { bq_d64 v; v.d = x; return v.u; }
BQ_FN double bq_bitsd(uint64_t u) { bq_d64 v; v.u = u; return v.d; }
BQ_FN uint32_t bq_fbits(float x) { bq_f32 v; v.f = x; return v.u; }
BQ_FN float bq_bitsf(uint32_t u) { bq_f32 v; v.u = u; return v.f; }

/* ------------------------------------------------------------------------ */
/* Philox4x32-10                                                             */
/* ------------------------------------------------------------------------ */
typedef struct { uint32_t v0, v1, v2, v3; } bq_u32x4;

BQ_FN bq_u32x4 bq_philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                uint32_t k0, uint32_t k1)
{
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    bq_u32x4 o; o.v0 = c0; o.v1 = c1; o.v2 = c2; o.v3 = c3;
    return o;
}

/*
 * One random stream = (key, counter prefix).  key = (seed, chain);
 * counter = (a, b, sweep, phase<<20 | block).  Each refill consumes one block
 * of four 32-bit words.  The buffer is a shift register (no dynamic indexing),
 * so it stays in VGPRs on the GPU.
 */
typedef struct {
    uint32_t k0, k1, c0, c1, c2, c3;
    uint32_t b0, b1, b2, b3;
    int n;
} bq_rng;

BQ_FN void bq_rng_init(bq_rng* r, uint32_t seed, uint32_t chain, uint32_t a, uint32_t b,
                       uint32_t sweep, uint32_t phase)
{
    r->k0 = seed; r->k1 = chain; r->c0 = a; r->c1 = b; r->c2 = sweep;
    r->c3 = phase << 20;
    r->b0 = r->b1 = r->b2 = r->b3 = 0u;
    r->n = 0;
}

BQ_FN uint32_t bq_next32(bq_rng* r)
{
    if (r->n == 0) {
        bq_u32x4 o = bq_philox4x32_10(r->c0, r->c1, r->c2, r->c3, r->k0, r->k1);
        r->c3 += 1u;
        r->b0 = o.v0; r->b1 = o.v1; r->b2 = o.v2; r->b3 = o.v3;
        r->n = 4;
    }
    uint32_t w = r->b0;
    r->b0 = r->b1; r->b1 = r->b2; r->b2 = r->b3;
    r->n -= 1;
    return w;
}

/* uniform on the open interval (0,1) from one 32-bit word, centred in its
 * 2^-32 cell: ((w + 1/2) 2^-32) is exact in double.  A resolution of 2^-32
 * bounds the bias of every Bernoulli, inversion and accept/reject decision
 * that uses it by 2^-32. */
BQ_FN double bq_uniform_of(uint32_t w)
{
    return ((double)w + 0.5) * 2.3283064365386962890625e-10; /* 2^-32 */
}

BQ_FN double bq_uniform(bq_rng* r) { return bq_uniform_of(bq_next32(r)); }

/* ------------------------------------------------------------------------ */
/* double precision elementary functions                                     */
/* ------------------------------------------------------------------------ */
#define BQ_LN2_HI 0.6931467056274414        /* 0x3fe62e4200000000, 21-bit mantissa */
#define BQ_LN2_LO 4.7493250390316726e-07
/* ln 2 (hi, lo) times 2^-52: the exponent k enters as the double (k << 52), which
 * is exact, and k * ln2 is the same real product, so the same rounding, without
 * the shift back */
#define BQ_LN2_HI_S52 (BQ_LN2_HI * 0x1p-52)
#define BQ_LN2_LO_S52 (BQ_LN2_LO * 0x1p-52)
#define BQ_INV_LN2 1.4426950408889634
#define BQ_HALF_LOG_2PI 0.9189385332046728
#define BQ_LOG1P_LO (-0.29289321881345248)   /* sqrt(1/2) - 1 */
#define BQ_LOG1P_HI 0.41421356237309515      /* sqrt(2) - 1   */

/* (log1p(f) - f) / f^2 on [sqrt(1/2)-1, sqrt(2)-1], degree 19, rel err 4e-17 */
BQ_TABLE double bq_log1p_q64_c[20] = {
    2.63421908043126387633e-2, -5.96236671431324208315e-2, 6.45516604239284724764e-2,
    -5.66719174788978263813e-2, 5.68275041380634990396e-2, -6.20216185875966524721e-2,
    6.68287182705409738004e-2, -7.14852763864529533455e-2, 7.69162297923115106935e-2,
    -8.33298977274350011202e-2, 9.09092238018320702997e-2, -1.00000116843998718141e-1,
    1.1111111090821723657e-1, -1.24999997803443355093e-1, 1.42857142827786720574e-1,
    -1.66666666687299985794e-1, 2.00000000000323002736e-1, -2.4999999999992503219e-1,
    3.33333333333332593185e-1, -5.0e-1,
};
BQ_FN double bq_log1p_q64(double f)
{
    double p = bq_log1p_q64_c[0];
    for (int j = 1; j < 20; ++j) p = __builtin_fma(p, f, bq_log1p_q64_c[j]);
    return p;
}

/* split a positive finite normal x into 2^e * m, m in [sqrt(1/2), sqrt(2)) */
BQ_FN double bq_split64(double x, int* e)
{
    uint64_t i = bq_dbits(x);
    int64_t t = (int64_t)(i - 0x3FE6A09E667F3BCDull);
    int64_t k = t >> 52;                       /* arithmetic shift */
    *e = (int)k;
    return bq_bitsd(i - ((uint64_t)k << 52));
}

BQ_FN double bq_log(double x)
{
    if (!(x > 0.0)) return x == 0.0 ? -BQ_INF : BQ_NAN;
    if (x == BQ_INF) return BQ_INF;
    int eadj = 0;
    if (x < 2.2250738585072014e-308) { x *= 18014398509481984.0; eadj = -54; } /* 2^54 */
    int e;
    double m = bq_split64(x, &e);
    double f = m - 1.0;
    double r = __builtin_fma(f * f, bq_log1p_q64(f), f);
    double ed = (double)(e + eadj);
    return __builtin_fma(ed, BQ_LN2_HI, __builtin_fma(ed, BQ_LN2_LO, r));
}

BQ_FN double bq_log1p(double y)
{
    int e = 0;
    double f = y;
    if (!(y >= BQ_LOG1P_LO && y <= BQ_LOG1P_HI)) {
        if (!(y > -1.0)) return y == -1.0 ? -BQ_INF : BQ_NAN;
        if (y == BQ_INF) return BQ_INF;
        double m = bq_split64(1.0 + y, &e);
        f = m - 1.0;
    }
    double r = __builtin_fma(f * f, bq_log1p_q64(f), f);
    double ed = (double)e;
    return __builtin_fma(ed, BQ_LN2_HI, __builtin_fma(ed, BQ_LN2_LO, r));
}

/* log1p(y) for finite y > -1 without the special-value and range branches of
 * bq_log1p, for per-element terms whose arguments are known to lie there:
 * 1 + y = 2^k m with m in [sqrt(1/2), sqrt(2)), k from the rounded sum t = 1 + y,
 * and the reduced argument f = (1 + y) 2^-k - 1 = fma(y, 2^-k, 2^-k - 1) formed
 * from the unrounded y with one rounding (f = y exactly for k = 0) */
BQ_FN double bq_log1p_mid(double y)
{
    double t = 1.0 + y;
    uint64_t kb = (bq_dbits(t) - 0x3FE6A09E667F3BCDull) & 0xFFF0000000000000ull;   /* k << 52 */
    double s = bq_bitsd(0x3FF0000000000000ull - kb);                                  /* 2^-k */
    double f = __builtin_fma(y, s, s - 1.0);
    double eds = (double)(int64_t)kb;   /* k 2^52 */
    double r = __builtin_fma(f * f, bq_log1p_q64(f), f);
    return __builtin_fma(eds, BQ_LN2_HI_S52, __builtin_fma(eds, BQ_LN2_LO_S52, r));
}

/* (exp(r) - 1 - r) / r^2 on |r| <= ln2/2, degree 11, rel err 6e-19 */
BQ_TABLE double bq_exp_e64_c[12] = {
    1.60867819486130656973e-10, 2.0918368776542248092e-9, 2.50520706768841624642e-8,
    2.75572626528368004013e-7, 2.75573192476098404626e-6, 2.48015873370290303277e-5,
    1.98412698412631517252e-4, 1.38888888888788540436e-3, 8.33333333333333321769e-3,
    4.16666666666666782315e-2, 1.66666666666666657415e-1, 5.0e-1,
};
BQ_FN double bq_exp_e64(double r)
{
    double p = bq_exp_e64_c[0];
    for (int j = 1; j < 12; ++j) p = __builtin_fma(p, r, bq_exp_e64_c[j]);
    return p;
}

BQ_FN double bq_pow2i(int k) /* 2^k for k in [-1022, 1023] */
{
    return bq_bitsd((uint64_t)(int64_t)(k + 1023) << 52);
}

BQ_FN double bq_exp(double x)
{
    if (x != x) return x;
    if (x > 709.782712893384) return BQ_INF;
    if (x < -745.2) return 0.0;
    double kd = __builtin_rint(x * BQ_INV_LN2);
    int k = (int)kd;
    double r = (x - kd * BQ_LN2_HI) - kd * BQ_LN2_LO;
    double p = 1.0 + (r + (r * r) * bq_exp_e64(r));
    if (k > 1000 || k < -1000) {
        int k1 = k / 2;
        return (p * bq_pow2i(k1)) * bq_pow2i(k - k1);
    }
    return p * bq_pow2i(k);
}

BQ_FN double bq_expm1(double x)
{
    if (x >= -0.3465735902799727 && x <= 0.3465735902799727)
        return x + (x * x) * bq_exp_e64(x);
    if (x < -40.0) return -1.0;
    return bq_exp(x) - 1.0;
}

/* log(1 + e^t) */
BQ_FN double bq_softplus(double t)
{
    if (t > 0.0) return t + bq_log1p(bq_exp(-t));
    return bq_log1p(bq_exp(t));
}

/* Stirling remainder sum_{n>=1} B_2n / (2n(2n-1) w^(2n-1)) for w >= 10 (8 terms, error < 3e-17) */
BQ_FN double bq_stirling_s64(double iw)
{
    double z = iw * iw;
    double p = -0.029550653594771242;
    p = __builtin_fma(p, z, 0.00641025641025641);
    p = __builtin_fma(p, z, -0.0019175269175269176);
    p = __builtin_fma(p, z, 0.0008417508417508417);
    p = __builtin_fma(p, z, -0.0005952380952380953);
    p = __builtin_fma(p, z, 0.0007936507936507937);
    p = __builtin_fma(p, z, -0.002777777777777778);
    p = __builtin_fma(p, z, 0.08333333333333333);
    return p * iw;
}

/* w S(w) for 1 <= w <= 10 from t = 1/w: the degree-20 Chebyshev interpolant on
 * t in [0.1, 1] (busseq_gpu_amd/csrc/gen_coeffs.py, STIRLING_T_F64; abs err of
 * S = t Q(t) below 1.1e-17 with the rounded coefficients) */
BQ_FN double bq_stirling_q64_t(double t)
{
    double p = -3.43114355651204472287e-5;
    p = __builtin_fma(p, t, 4.2004461293514551521e-4);
    p = __builtin_fma(p, t, -2.41949115872181234541e-3);
    p = __builtin_fma(p, t, 8.71150149019403637063e-3);
    p = __builtin_fma(p, t, -2.19576369828119316585e-2);
    p = __builtin_fma(p, t, 4.10800329988699128303e-2);
    p = __builtin_fma(p, t, -5.89271741857547368126e-2);
    p = __builtin_fma(p, t, 6.58297525301950320342e-2);
    p = __builtin_fma(p, t, -5.72932315726428567637e-2);
    p = __builtin_fma(p, t, 3.80745376197201002011e-2);
    p = __builtin_fma(p, t, -1.82074013536179177275e-2);
    p = __builtin_fma(p, t, 5.30895500113091195482e-3);
    p = __builtin_fma(p, t, -5.12545438014087558279e-4);
    p = __builtin_fma(p, t, 2.72928504362822963239e-4);
    p = __builtin_fma(p, t, -6.38592612214043784009e-4);
    p = __builtin_fma(p, t, 5.36487805755688477614e-6);
    p = __builtin_fma(p, t, 7.93144926255619697394e-4);
    p = __builtin_fma(p, t, 3.50466592371063494979e-8);
    p = __builtin_fma(p, t, -2.77777945586209912276e-3);
    p = __builtin_fma(p, t, 4.94941088224170359986e-11);
    p = __builtin_fma(p, t, 8.33333333326573694189e-2);
    return p;
}

/* Stirling remainder S(w) for w >= 1 from iw = 1/w: the fit below 10, the series above */
BQ_FN double bq_stirling_s64_w1(double w, double iw)
{
    return w >= 10.0 ? bq_stirling_s64(iw) : iw * bq_stirling_q64_t(iw);
}

/* log Gamma(z) for z > 0 (priors, the Poisson sampler, the BIC: scalar or rare
 * per element, so the shift to 10 and the series serve) */
BQ_FN double bq_lgamma(double z)
{
    if (!(z > 0.0)) return z == 0.0 ? BQ_INF : BQ_NAN;
    if (z == BQ_INF) return BQ_INF;
    double prod = 1.0, w = z;
    int shifted = 0;
    while (w < 10.0) { prod *= w; w += 1.0; shifted = 1; }
    double lw = bq_log(w);
    double s = bq_stirling_s64(1.0 / w);
    double st = ((w - 0.5) * lw - w) + (BQ_HALF_LOG_2PI + s);
    return shifted ? st - bq_log(prod) : st;
}

/* ------------------------------------------------------------------------ */
/* single precision elementary functions (per-element hot path)              */
/* ------------------------------------------------------------------------ */
#define BQ_LN2_HI_F 0.693115234375f
#define BQ_LN2_LO_F 3.194618329871446e-05f
/* ln 2 (hi, lo) times 2^-23: the exponent k enters as the float (k << 23), which
 * is exact, and k * ln2 is the same real product, so the same rounding, without
 * the shift back */
#define BQ_LN2_HI_F23 (BQ_LN2_HI_F * 0x1p-23f)
#define BQ_LN2_LO_F23 (BQ_LN2_LO_F * 0x1p-23f)
#define BQ_LN2_F23 (0.693147182464599609375f * 0x1p-23f)   /* ln 2 rounded to single, times 2^-23 */
#define BQ_LOG1P_LO_F (-0.29289321f)
#define BQ_LOG1P_HI_F 0.41421356f

/* 1/z for a positive z held to [2^-126, 2^125] (one v_med3_f32), rounded to
 * nearest as the IEEE division 1.0f / z returns it.  On gfx950 the hardware
 * reciprocal y = v_rcp_f32(z) followed by one Newton step
 * fma(fma(-z, y, 1), y, y) (Markstein's final correction) gives that value for
 * every normal z in [2^-126, 2^126): all 2113929216 positive ones were compared
 * with the division on the device, and the library repeats that comparison on
 * each device before its first single-precision sampler is created
 * (busseq_rcp_check), so that host and device keep returning the same bits.
 * The host takes the division of the same held value.  Four instructions
 * instead of the division's eleven.  z is a positive finite value in every
 * caller (phi + mu, (phi' + mu)(phi + x)(phi' + x)). */
#define BQ_RCP_LO 0x1p-126f
#define BQ_RCP_HI 0x1p125f
BQ_FN float bq_rcpf(float z)
{
#ifdef __HIP_DEVICE_COMPILE__
    const float zc = __builtin_amdgcn_fmed3f(z, BQ_RCP_LO, BQ_RCP_HI);   /* = the host's min(max()) for non-NaN z */
    const float y = __builtin_amdgcn_rcpf(zc);
    return __builtin_fmaf(__builtin_fmaf(-zc, y, 1.0f), y, y);
#else
    const float zc = __builtin_fminf(__builtin_fmaxf(z, BQ_RCP_LO), BQ_RCP_HI);
    return 1.0f / zc;
#endif
}

/* (log1p(f) - f) / f^2 on [sqrt(1/2)-1, sqrt(2)-1], degree 8, rel err 1e-8 */
BQ_FN float bq_log1p_q32(float f)
{
    float p = -7.68349170684814453125e-2f;
    p = __builtin_fmaf(p, f, 1.27387985587120056152e-1f);
    p = __builtin_fmaf(p, f, -1.31315439939498901367e-1f);
    p = __builtin_fmaf(p, f, 1.42030596733093261719e-1f);
    p = __builtin_fmaf(p, f, -1.66260927915573120117e-1f);
    p = __builtin_fmaf(p, f, 2.00012877583503723145e-1f);
    p = __builtin_fmaf(p, f, -2.50007987022399902344e-1f);
    p = __builtin_fmaf(p, f, 3.33333283662796020508e-1f);
    p = __builtin_fmaf(p, f, -4.99999970197677612305e-1f);
    return p;
}

/* the same on the same interval, degree 7, rel err 5.6e-8: for bq_log1pf_mid,
 * one fused multiply-add fewer in each of the subtype draw's K - 1 logarithms
 * per cell x gene */
BQ_FN float bq_log1p_q32_mid(float f)
{
    float p = 8.79710465669631958008e-2f;
    p = __builtin_fmaf(p, f, -1.41752481460571289063e-1f);
    p = __builtin_fmaf(p, f, 1.48381710052490234375e-1f);
    p = __builtin_fmaf(p, f, -1.65869489312171936035e-1f);
    p = __builtin_fmaf(p, f, 1.99716746807098388672e-1f);
    p = __builtin_fmaf(p, f, -2.50011831521987915039e-1f);
    p = __builtin_fmaf(p, f, 3.33336770534515380859e-1f);
    p = __builtin_fmaf(p, f, -4.99999970197677612305e-1f);
    return p;
}

/* split a positive finite normal x into 2^k * m, m in [sqrt(1/2), sqrt(2)); kb = k << 23 */
BQ_FN float bq_split32s(float x, int32_t* kb)
{
    uint32_t i = bq_fbits(x);
    uint32_t k23 = (i - 0x3F3504F3u) & 0xFF800000u;
    *kb = (int32_t)k23;
    return bq_bitsf(i - k23);
}

/* branch-free: every lane runs the same instructions; special inputs are
 * patched by selects at the end (x <= 0, +inf, NaN; subnormals rescaled) */
BQ_FN float bq_logf(float x)
{
    int sub = x < 1.17549435e-38f;
    float xs = sub ? x * 33554432.0f : x; /* 2^25 */
    int32_t kb;
    float m = bq_split32s(xs, &kb);
    float f = m - 1.0f;
    float r = __builtin_fmaf(f * f, bq_log1p_q32(f), f);
    float efs = (float)(kb + (sub ? -(25 << 23) : 0));   /* k 2^23 */
    float res = __builtin_fmaf(efs, BQ_LN2_HI_F23, __builtin_fmaf(efs, BQ_LN2_LO_F23, r));
    if (!(x > 0.0f)) res = x == 0.0f ? -BQ_INFF : BQ_NANF;
    if (x == BQ_INFF) res = BQ_INFF;
    return res;
}

/* log1p for the arguments the fast path does not take: y <= -1, NaN, +inf and
 * 1 + y >= sqrt(2) 2^125 (k >= 126), where 1 + y equals y up to a relative 2^-125 */
/* logf for a positive, finite, normal argument (no special-value handling), for
 * the phi term's lgamma difference: the degree-7 polynomial and k ln 2 in one
 * fused multiply-add, as bq_log1pf (bq_logf keeps the degree-8 polynomial and
 * the two-part ln 2, within 1 ulp) */
BQ_FN float bq_logf_norm(float x)
{
    int32_t kb;
    float m = bq_split32s(x, &kb);
    float f = m - 1.0f;
    float r = __builtin_fmaf(f * f, bq_log1p_q32_mid(f), f);
    float efs = (float)kb;   /* k 2^23 */
    return __builtin_fmaf(efs, BQ_LN2_F23, r);
}

BQ_FN float bq_log1pf_rare(float y)
{
    if (!(y > -1.0f)) return y == -1.0f ? -BQ_INFF : BQ_NANF;
    return bq_logf(1.0f + y);
}

/*
 * log1p(y) in single precision.  1 + y = 2^k m with m in [sqrt(1/2), sqrt(2)),
 * k taken from the rounded sum t = 1 + y; the reduced argument
 *   f = (1 + y) 2^-k - 1 = fma(y, 2^-k, 2^-k - 1)
 * is formed from the unrounded y with a single rounding (2^-k - 1 is exact
 * for the k that occur here or carries an error far below f's own), so no
 * correction term for the rounding of 1 + y is needed, and for |y| small
 * (k = 0) f = y exactly.  The degree-7 polynomial of bq_log1p_q32_mid and
 * k ln 2 in one fused multiply-add with ln 2 rounded to single precision
 * (1.9e-9 |k|, at most 0.13 ulp, exact for k = 0): within 2 ulp.
 */
BQ_FN float bq_log1pf(float y)
{
    float t = 1.0f + y;
    uint32_t kb = (bq_fbits(t) - 0x3F3504F3u) & 0xFF800000u;   /* k << 23 */
    float s = bq_bitsf(0x3F800000u - kb);                        /* 2^-k */
    float f = __builtin_fmaf(y, s, s - 1.0f);
    float efs = (float)(int32_t)kb;   /* k 2^23 */
    float r = __builtin_fmaf(f * f, bq_log1p_q32_mid(f), f);
    float res = __builtin_fmaf(efs, BQ_LN2_F23, r);
    /* 0 < t < sqrt(2) 2^125 (y > -1, k <= 125, not NaN) as one unsigned compare of the bits */
    if (bq_fbits(t) - 1u >= 0x7E3504F3u - 1u) res = bq_log1pf_rare(y);
    return res;
}

/* log1p(y) for y in (-1, sqrt(2) 2^125), and NaN for NaN: bq_log1pf without
 * its branch to the special values, for callers that keep y in that range
 * (within 2 ulp; the degree-7 polynomial's mean error is 0.3 ulp, the
 * degree-8 one's 0.25) */
BQ_FN float bq_log1pf_mid(float y)
{
    float t = 1.0f + y;
    uint32_t kb = (bq_fbits(t) - 0x3F3504F3u) & 0xFF800000u;
    float s = bq_bitsf(0x3F800000u - kb);
    float f = __builtin_fmaf(y, s, s - 1.0f);
    float efs = (float)(int32_t)kb;   /* k 2^23 */
    float r = __builtin_fmaf(f * f, bq_log1p_q32_mid(f), f);
    /* k ln 2 + r in one rounding with ln 2 rounded to single precision
     * (|error| 1.9e-9 |k|: at most 0.13 ulp of the result; exact for k = 0,
     * most of the subtype term's arguments) */
    return __builtin_fmaf(efs, BQ_LN2_F23, r);
}

/* Stirling remainder S(w) = lgamma(w) - (w - 1/2) log w + w - log(2 pi)/2 for
 * w >= 1 in single precision, from iw = 1/w: iw Q(iw^2), Q the degree-6
 * Chebyshev interpolant of w S(w) in u = 1/w^2 on [0, 1]
 * (busseq_gpu_amd/csrc/gen_coeffs.py; abs err of S < 6.1e-9 with the rounded
 * coefficients, < 1e-8 evaluated in single precision).  Valid down to w = 1,
 * so the arguments of the phi term, phi + x with x >= 1, are never shifted. */
BQ_FN float bq_stirling_s32(float iw)
{
    float z = iw * iw;
    float p = 3.96799550799187272787e-5f;
    p = __builtin_fmaf(p, z, -1.71036081155762076378e-4f);
    p = __builtin_fmaf(p, z, 3.40970960678532719612e-4f);
    p = __builtin_fmaf(p, z, -4.83421026729047298431e-4f);
    p = __builtin_fmaf(p, z, 7.78905232436954975128e-4f);
    p = __builtin_fmaf(p, z, -2.77695455588400363922e-3f);
    p = __builtin_fmaf(p, z, 8.33333283662796020508e-2f);
    return p * iw;
}

/*
 * lgamma(z + d) - lgamma(z) for z > 0 and z + d > 0, computed without the
 * cancellation of two large lgamma values:
 *   d*log(z+d) + (z-1/2)*log1p(d/z) - d + S(z+d) - S(z)   (Stirling),
 * with both arguments first shifted up to >= 1 by the recurrence where
 * needed (never for the phi term: there z, z + d > 1).
 */
BQ_FN float bq_lgamma_diff32(float z, float d)
{
    float zp = z + d;
    float pz = 1.0f, pzp = 1.0f;
    int shifted = 0;
    while (z < 1.0f || zp < 1.0f) { pz *= z; pzp *= zp; z += 1.0f; zp += 1.0f; shifted = 1; }
    /* both reciprocals from one division (z, zp >= 1; ~2 ulp each) */
    float w = 1.0f / (z * zp);
    float iz = zp * w, izp = z * w;
    float core = __builtin_fmaf(d, bq_logf_norm(zp), __builtin_fmaf(z - 0.5f, bq_log1pf(d * iz), -d));
    float dl = core + (bq_stirling_s32(izp) - bq_stirling_s32(iz));
    if (shifted) dl = dl - bq_logf(pzp / pz);
    return dl;
}

/* the double-precision form: the same Stirling difference with S from
 * bq_stirling_s64_w1, valid from 1, so the phi term's arguments phi + x
 * (x >= 1) are never shifted (a shift to 10 ran up to nine steps of a loop and
 * two more logarithms on most lanes) */
BQ_FN double bq_lgamma_diff64(double z, double d)
{
    double zp = z + d;
    double pz = 1.0, pzp = 1.0;
    int shifted = 0;
    while (z < 1.0 || zp < 1.0) { pz *= z; pzp *= zp; z += 1.0; zp += 1.0; shifted = 1; }
    double iz = 1.0 / z, izp = 1.0 / zp;
    double core = __builtin_fma(d, bq_log(zp), __builtin_fma(z - 0.5, bq_log1p_mid(d * iz), -d));
    double dl = core + (bq_stirling_s64_w1(zp, izp) - bq_stirling_s64_w1(z, iz));
    if (shifted) dl = dl - (bq_log(pzp) - bq_log(pz));
    return dl;
}

/* ------------------------------------------------------------------------ */
/* normal deviates: inverse CDF, Wichura (1988) AS 241 PPND16                */
/* ------------------------------------------------------------------------ */
BQ_FN double bq_ndtri(double p)
{
    double q = p - 0.5;
    if (q >= -0.425 && q <= 0.425) {
        double r = 0.180625 - q * q;
        double num = (((((((2.5090809287301226727e+3 * r + 3.3430575583588128105e+4) * r
                     + 6.7265770927008700853e+4) * r + 4.5921953931549871457e+4) * r
                     + 1.3731693765509461125e+4) * r + 1.9715909503065514427e+3) * r
                     + 1.3314166789178437745e+2) * r + 3.3871328727963666080e+0);
        double den = (((((((5.2264952788528545610e+3 * r + 2.8729085735721942674e+4) * r
                     + 3.9307895800092710610e+4) * r + 2.1213794301586595867e+4) * r
                     + 5.3941960214247511077e+3) * r + 6.8718700749205790830e+2) * r
                     + 4.2313330701600911252e+1) * r + 1.0);
        return q * num / den;
    }
    double pm = q < 0.0 ? p : 1.0 - p;
    double r = __builtin_sqrt(-bq_log(pm));
    double val;
    if (r <= 5.0) {
        r -= 1.6;
        double num = (((((((7.74545014278341407640e-4 * r + 2.27238449892691845833e-2) * r
                     + 2.41780725177450611770e-1) * r + 1.27045825245236838258e+0) * r
                     + 3.64784832476320460504e+0) * r + 5.76949722146069140550e+0) * r
                     + 4.63033784615654529590e+0) * r + 1.42343711074968357734e+0);
        double den = (((((((1.05075007164441684324e-9 * r + 5.47593808499534494600e-4) * r
                     + 1.51986665636164571966e-2) * r + 1.48103976427480074590e-1) * r
                     + 6.89767334985100004550e-1) * r + 1.67638483018380384940e+0) * r
                     + 2.05319162663775882187e+0) * r + 1.0);
        val = num / den;
    } else {
        r -= 5.0;
        double num = (((((((2.01033439929228813265e-7 * r + 2.71155556874348757815e-5) * r
                     + 1.24266094738807843860e-3) * r + 2.65321895265761230930e-2) * r
                     + 2.96560571828504891230e-1) * r + 1.78482653991729133580e+0) * r
                     + 5.46378491116411436990e+0) * r + 6.65790464350110377720e+0);
        double den = (((((((2.04426310338993978564e-15 * r + 1.42151175831644588870e-7) * r
                     + 1.84631831751005468180e-5) * r + 7.86869131145613259100e-4) * r
                     + 1.48753612908506148525e-2) * r + 1.36929880922735805310e-1) * r
                     + 5.99832206555887937690e-1) * r + 1.0);
        val = num / den;
    }
    return q < 0.0 ? -val : val;
}

BQ_FN double bq_normal(bq_rng* r) { return bq_ndtri(bq_uniform(r)); }

/* ------------------------------------------------------------------------ */
/* samplers                                                                  */
/* ------------------------------------------------------------------------ */
#define BQ_MAX_REJECT 1000

/* Gamma(shape a, scale s): Marsaglia & Tsang (2000); a < 1 via Gamma(a+1)*U^(1/a) */
BQ_FN double bq_rgamma(bq_rng* r, double a, double scale)
{
    if (!(a > 0.0) || !(scale > 0.0)) return 0.0;
    double a1 = a < 1.0 ? a + 1.0 : a;
    double d = a1 - 1.0 / 3.0;
    double c = 1.0 / __builtin_sqrt(9.0 * d);
    double v = 1.0;
    for (int it = 0; it < BQ_MAX_REJECT; ++it) {
        double x = bq_normal(r);
        double t = 1.0 + c * x;
        if (t <= 0.0) continue;
        v = t * t * t;
        double u = bq_uniform(r);
        double x2 = x * x;
        if (u < 1.0 - 0.0331 * (x2 * x2)) break;
        if (bq_log(u) < 0.5 * x2 + d * ((1.0 - v) + bq_log(v))) break;
    }
    double g = d * v;
    if (a < 1.0) g *= bq_exp(bq_log(bq_uniform(r)) / a);
    return g * scale;
}

/* 1/k for k = 0..256 (entry 0 unused), correctly rounded doubles */
#define BQ_INV_N 257
BQ_TABLE double bq_inv_tab[BQ_INV_N] = {
#define BQ_I4(a) 1.0 / (a), 1.0 / (a + 1), 1.0 / (a + 2), 1.0 / (a + 3)
#define BQ_I16(a) BQ_I4(a), BQ_I4(a + 4), BQ_I4(a + 8), BQ_I4(a + 12)
    0.0, 1.0, 1.0 / 2, 1.0 / 3, BQ_I4(4), BQ_I4(8), BQ_I4(12),
    BQ_I16(16), BQ_I16(32), BQ_I16(48), BQ_I16(64), BQ_I16(80), BQ_I16(96), BQ_I16(112),
    BQ_I16(128), BQ_I16(144), BQ_I16(160), BQ_I16(176), BQ_I16(192), BQ_I16(208), BQ_I16(224), BQ_I16(240),
    1.0 / 256
#undef BQ_I16
#undef BQ_I4
};

/* Poisson(lam): inversion for lam < 10, Hoermann (1993) PTRS otherwise */
BQ_FN int bq_rpois(bq_rng* r, double lam)
{
    if (!(lam > 0.0)) return 0;
    if (lam < 10.0) {
        double p = bq_exp(-lam);
        double s = p;
        double u = bq_uniform(r);
        int k = 0;
        while (u > s && k < 200) { k += 1; p *= lam * bq_inv_tab[BQ_WAVE_UNIFORM(k)]; s += p; }
        return k;
    }
    if (lam > 1.0e9) lam = 1.0e9;
    double slam = __builtin_sqrt(lam);
    double loglam = bq_log(lam);
    double b = 0.931 + 2.53 * slam;
    double a = -0.059 + 0.02483 * b;
    double invalpha = 1.1239 + 1.1328 / (b - 3.4);
    double vr = 0.9277 - 3.6224 / (b - 2.0);
    double loginva = bq_log(invalpha);
    double k = __builtin_floor(lam);
    for (int it = 0; it < BQ_MAX_REJECT; ++it) {
        double U = bq_uniform(r) - 0.5;
        double V = bq_uniform(r);
        double us = 0.5 - (U < 0.0 ? -U : U);
        k = __builtin_floor((2.0 * a / us + b) * U + lam + 0.43);
        if (us >= 0.07 && V <= vr) break;
        if (k < 0.0 || (us < 0.013 && V > us)) continue;
        if (bq_log(V) + loginva - bq_log(a / (us * us) + b)
            <= -lam + k * loglam - bq_lgamma(k + 1.0)) break;
    }
    if (k < 0.0) k = 0.0;
    return (int)k;
}

/* negative binomial with mean mu and size phi, as Gamma-Poisson (cu:199-210) */
BQ_FN int bq_rnbinom(bq_rng* r, double mu, double phi)
{
    double lam = bq_rgamma(r, phi, mu / phi);
    return bq_rpois(r, lam);
}

/*
 * Negative binomial NB(mean mu, size phi) by sequential inversion of its CDF:
 *   p_0 = (phi / (phi + mu))^phi,  p_k = p_{k-1} q (phi + k - 1) / k,  q = mu / (phi + mu),
 * returning the first k with u <= p_0 + ... + p_k.  One uniform per draw.
 * bq_nb_inv_ok admits mu <= 30 with q <= 0.85; there P(X > 255) < 2e-13, far
 * below 2^-33, the smallest distance of a 32-bit uniform from 1, and the
 * search ends by k = 210 even at that extreme u (with a relative rounding
 * error of the partial sums below 1e-11), so the k < 255 guard never binds.
 */
BQ_FN int bq_nb_inv_ok(double mu, double phi) { return mu <= 30.0 && 3.0 * mu <= 17.0 * phi; }

/* p_k = p_{k-1} c_k with c_k = q ((phi + k - 1) (1/k)); c_{k+1} is formed in
 * step k, so that its table read overlaps the step (inv needs 1/256 for the
 * last step). */
BQ_FN int bq_rnbinom_inv_t(double u, double mu, double phi, const double* inv)
{
    double q = mu / (phi + mu);
    /* -q in (-0.85, 0): often below bq_log1p's fast range, so the branch-free form */
    double p = bq_exp(phi * bq_log1p_mid(-q));
    double s = p;
    int k = 0;
    double c = q * (phi * inv[1]);
    while (u > s && k < 255) {
        k += 1;
        p *= c;
        s += p;
        c = q * ((phi + (double)k) * inv[BQ_WAVE_UNIFORM(k + 1)]);
    }
    return k;
}

/* inv: the reciprocals 1/k of bq_inv_tab, or a copy of them (device code may
 * pass one staged in LDS) */
BQ_FN int bq_rnbinom_inv(double u, double mu, double phi) { return bq_rnbinom_inv_t(u, mu, phi, bq_inv_tab); }

/* The proposal draw of the X/Z step: the first uniform of the stream drives
 * the inversion where bq_nb_inv_ok holds; elsewhere it is skipped and the
 * gamma-Poisson mixture runs on the following draws. */
BQ_FN int bq_rnbinom_xz(bq_rng* r, double mu, double phi)
{
    double u = bq_uniform(r);
    if (bq_nb_inv_ok(mu, phi)) return bq_rnbinom_inv(u, mu, phi);
    return bq_rnbinom(r, mu, phi);
}

/* Beta(a, b) as a ratio of gamma deviates (cu:212-218) */
BQ_FN double bq_rbeta(bq_rng* r, double a, double b)
{
    double x = bq_rgamma(r, a, 1.0);
    double y = bq_rgamma(r, b, 1.0);
    return x / (x + y);
}

/* ------------------------------------------------------------------------ */
/* exact fixed-point accumulation: value = sum * 2^-32, wrapping uint64 adds  */
/* ------------------------------------------------------------------------ */
#define BQ_FIX_ONE 4294967296.0
#define BQ_FIX_INV 2.3283064365386963e-10
#define BQ_FIX_CLAMP 4.6e18

BQ_FN uint64_t bq_fix(double v)
{
    double s = v * BQ_FIX_ONE;
    if (!(s < BQ_FIX_CLAMP)) s = (s != s) ? -BQ_FIX_CLAMP : BQ_FIX_CLAMP;
    if (s < -BQ_FIX_CLAMP) s = -BQ_FIX_CLAMP;
    return (uint64_t)(int64_t)__builtin_rint(s);
}

BQ_FN double bq_unfix(uint64_t acc) { return (double)(int64_t)acc * BQ_FIX_INV; }

#endif /* BUSSEQ_NUMERICS_H */
