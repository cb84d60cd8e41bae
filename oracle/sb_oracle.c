/*
 * sb_oracle.c -- CPU restatement of the SB solver loop.  TEST INFRASTRUCTURE:
 * the checker for the HIP path, never part of it (see sb_oracle.h).
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -mfma -fopenmp).  Every
 * floating-point expression below is written op by op; contraction is off so
 * that only the explicit fmaf/fma calls fuse.
 */
#include "sb_oracle.h"

#include <ctype.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- RNG and init ------------------------------------------------------ */

/* SB/CUDA_GPU/m2_opt.cu:38-46 */
uint32_t orc_xorshift32(uint32_t *state)
{
    uint32_t v = *state;
    v ^= v << 13;
    v ^= v >> 17;
    v ^= v << 5;
    *state = v;
    return v;
}

/* SB/CUDA_GPU/m2_opt.cu:54-55: (xs / (float)MAX) / 10.0, stored as float.
 * (float)4294967295.0 rounds to 2^32, so the FP32 division is the exact
 * scaling (float)v * 2^-32 (also under the reference's fast-math division,
 * whose reciprocal of a power of two is exact). */
float orc_ref_draw_value(uint32_t v)
{
    float q = (float)v / 4294967296.0f;
    double d = (double)q / 10.0;
    return (float)d;
}

void orc_init_reference(int N, int R, int ld, uint32_t seed, float *x, float *y)
{
    for (int r = 0; r < R; ++r)
        for (int i = N; i < ld; ++i) {
            x[(size_t)r * ld + i] = 0.0f;
            y[(size_t)r * ld + i] = 0.0f;
        }
    for (int i = 0; i < N; ++i) {
        /* m2_opt.cu:136: initRand[i] = i; state carries over trials (:152). */
        uint32_t st = (uint32_t)i + seed * (uint32_t)N;
        for (int r = 0; r < R; ++r) {
            x[(size_t)r * ld + i] = orc_ref_draw_value(orc_xorshift32(&st));
            y[(size_t)r * ld + i] = orc_ref_draw_value(orc_xorshift32(&st));
        }
    }
}

uint32_t orc_keyed_u24(uint64_t seed, uint32_t replica, uint32_t spin, uint32_t comp)
{
    uint64_t key = ((uint64_t)replica << 32) | (uint64_t)spin;
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * (2ull * key + (uint64_t)comp + 1ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (uint32_t)(z >> 40);
}

static float keyed_value(uint32_t u24)
{
    float v = (float)u24 * 0x1p-23f; /* exact, in [0, 2) */
    v = v - 1.0f;                    /* exact, in [-1, 1) */
    return v * 0.1f;                 /* one rounding */
}

void orc_init_uniform(int N, int R, int ld, uint64_t seed, float *x, float *y)
{
    for (int r = 0; r < R; ++r) {
        for (int i = 0; i < ld; ++i) {
            size_t o = (size_t)r * ld + i;
            if (i < N) {
                x[o] = keyed_value(orc_keyed_u24(seed, (uint32_t)r, (uint32_t)i, 0));
                y[o] = keyed_value(orc_keyed_u24(seed, (uint32_t)r, (uint32_t)i, 1));
            } else {
                x[o] = 0.0f;
                y[o] = 0.0f;
            }
        }
    }
}

/* ---- parameters -------------------------------------------------------- */

double orc_sigma(int N, const float *J, int ldj)
{
    if (N < 2) return 0.0;
    double s = 0.0;
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            if (i != j) {
                double v = (double)J[(size_t)i * ldj + j];
                s += v * v;
            }
    return sqrt(s / ((double)N * (double)(N - 1)));
}

double orc_c0_auto(int variant, int N, double a0, const float *J, int ldj)
{
    double sg = orc_sigma(N, J, ldj);
    if (sg <= 0.0) return 0.0;
    if (variant == ORC_ASB || variant == ORC_ASB_M1) return 0.7 * a0 / (sg * sqrt((double)N));
    return 0.5 / (sqrt((double)N) * sg);
}

void orc_coeffs(int variant, int k, int steps, double dt, double a0, double c0,
                float out[4])
{
    if (variant == ORC_ASB_M1) {
        out[0] = out[1] = out[2] = 0.0f;
        out[3] = (float)k / (float)steps; /* Bifurcation_m1.cu.c:165 amplitude */
        return;
    }
    if (variant == ORC_ASB) {
        float p = (float)k / (float)steps; /* m2_opt.cu:158 */
        out[0] = 0.0f;
        out[1] = (float)(dt * c0); /* m2_opt.cu:148: coeff1 = DeltaT * xi */
        out[2] = 0.0f;
        out[3] = p;
        return;
    }
    double p = (double)k / (double)steps;
    double a = a0 * p;
    out[0] = (float)(dt * (a0 - a));
    out[1] = (float)(dt * c0);
    out[2] = (float)(dt * a0);
    out[3] = (float)p;
}

/* ---- field ------------------------------------------------------------- */

static inline float sgnf(float v) { return v > 0.0f ? 1.0f : -1.0f; }

/* Rows are processed eight at a time so the sequential fmaf chains of
 * different rows overlap in the CPU pipeline; each row's own order is the
 * plain j = 0..N-1 chain of Bifurcation_m2.cu.c:93-97.  row_threads > 1
 * splits the row blocks over OpenMP threads (same sums). */
static void field_rows_f32(int N, const float *J, int ldj, const float *xs, float *F, int row_threads)
{
    const int nb = N / 8;
#ifdef _OPENMP
#pragma omp parallel for num_threads(row_threads) schedule(static) if (row_threads > 1)
#endif
    for (int b = 0; b < nb; ++b) {
        const int i = 8 * b;
        float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f, a4 = 0.f, a5 = 0.f, a6 = 0.f, a7 = 0.f;
        const float *r0 = J + (size_t)i * ldj;
        for (int j = 0; j < N; ++j) {
            float v = xs[j];
            a0 = fmaf(r0[j], v, a0);
            a1 = fmaf(r0[(size_t)ldj + j], v, a1);
            a2 = fmaf(r0[2 * (size_t)ldj + j], v, a2);
            a3 = fmaf(r0[3 * (size_t)ldj + j], v, a3);
            a4 = fmaf(r0[4 * (size_t)ldj + j], v, a4);
            a5 = fmaf(r0[5 * (size_t)ldj + j], v, a5);
            a6 = fmaf(r0[6 * (size_t)ldj + j], v, a6);
            a7 = fmaf(r0[7 * (size_t)ldj + j], v, a7);
        }
        F[i] = a0; F[i + 1] = a1; F[i + 2] = a2; F[i + 3] = a3;
        F[i + 4] = a4; F[i + 5] = a5; F[i + 6] = a6; F[i + 7] = a7;
    }
    for (int i = 8 * nb; i < N; ++i) {
        float a = 0.f;
        const float *row = J + (size_t)i * ldj;
        for (int j = 0; j < N; ++j) a = fmaf(row[j], xs[j], a);
        F[i] = a;
    }
}

/* The same per-row fmaf accumulation regrouped into 64 lane sums: lane l
 * takes the groups of 16 consecutive k starting at 16 (l + 64 it), it = 0,
 * 1, ..., each in ascending k, and the 64 sums are combined pairwise at lane
 * distances 32, 16, 8, 4, 2, 1.  Positions k in [N, round_up(N, 16)) enter as
 * fmaf(0, pad, acc) with pad = xs_pad[k] (the zero padding of the state, or
 * its sign for dSB).  This is the summation order of the thin kernels; xs_pad
 * holds round_up(N, 16) values. */
static void field_rows_f32_lanes(int N, const float *J, int ldj, const float *xs_pad, float *F)
{
    for (int i = 0; i < N; ++i) {
        const float *row = J + (size_t)i * ldj;
        float p[64];
        for (int l = 0; l < 64; ++l) {
            float a = 0.f;
            for (int k0 = 16 * l; k0 < N; k0 += 1024)
                for (int e = 0; e < 16; ++e) {
                    const int k = k0 + e;
                    a = fmaf(k < N ? row[k] : 0.0f, xs_pad[k], a);
                }
            p[l] = a;
        }
        for (int m = 32; m >= 1; m >>= 1) {
            float q[64];
            for (int l = 0; l < 64; ++l) q[l] = p[l] + p[l ^ m];
            memcpy(p, q, sizeof(p));
        }
        F[i] = p[0];
    }
}

/* Position in [0, round_up(N, 32)) of the t-th product of a row's MFMA chain. */
static inline int mfma_k(int t) { return (t & ~7) + ((t & 1) ? 4 : 0) + ((t & 7) >> 1); }

/* The MFMA order (orc_field_f32_mfma), eight rows at a time (independent
 * chains, each in its own order); xs_pad holds round_up(N, 32) values. */
static void field_rows_f32_mfma(int N, const float *J, int ldj, const float *xs_pad, float *F, int row_threads)
{
    const int Kp = (N + 31) / 32 * 32;
    const int nb = (N + 7) / 8;
#ifdef _OPENMP
#pragma omp parallel for num_threads(row_threads) schedule(static) if (row_threads > 1)
#endif
    for (int b = 0; b < nb; ++b) {
        const int i0 = 8 * b, nr = N - i0 < 8 ? N - i0 : 8;
        float a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int t = 0; t < Kp; ++t) {
            const int k = mfma_k(t);
            const float v = xs_pad[k];
            for (int q = 0; q < nr; ++q) {
                const float jv = k < N ? J[(size_t)(i0 + q) * ldj + k] : 0.0f;
                a[q] = fmaf(v, jv, a[q]);
            }
        }
        for (int q = 0; q < nr; ++q) F[i0 + q] = a[q];
    }
}

/* ---- The split form's BF16 MFMA (v_mfma_f32_32x32x16_bf16) ---------------- */

static uint32_t f32_bits(float v)
{
    uint32_t u;
    memcpy(&u, &v, 4);
    return u;
}

/* Nearest-even rounding of an FP32 value to BF16 (returned as a float). */
static float bf16_rne(float v)
{
    uint32_t u = f32_bits(v);
    if ((u & 0x7f800000u) == 0x7f800000u) return v;  /* inf / nan: not used */
    u = (u + 0x7fffu + ((u >> 16) & 1u)) & 0xffff0000u;
    float r;
    memcpy(&r, &u, 4);
    return r;
}

/* |v| = m * 2^(e - frac) for a nonzero float with `frac` fraction bits kept
 * (23 for FP32, 7 for a BF16 value); subnormals keep e = -126, no hidden bit. */
static void float_parts(float v, int frac, int *e, uint32_t *m)
{
    const uint32_t u = f32_bits(v);
    const int be = (int)((u >> 23) & 0xffu);
    const uint32_t man = (u & 0x7fffffu) >> (23 - frac);
    *e = be ? be - 127 : -126;
    *m = be ? man | (1u << frac) : man;
}

static float round_scaled(__int128 s, int L)
{
    if (s == 0) return 0.0f;
    const int neg = s < 0;
    unsigned __int128 m = neg ? (unsigned __int128)(-s) : (unsigned __int128)s;
    const uint64_t mh = (uint64_t)(m >> 64), ml = (uint64_t)m;
    const int nb = mh ? 128 - __builtin_clzll(mh) : 64 - __builtin_clzll(ml);
    if (nb > 24) {
        int sh = nb - 24;
        const unsigned __int128 one = 1;
        unsigned __int128 q = m >> sh, rem = m & ((one << sh) - 1), half = one << (sh - 1);
        if (rem > half || (rem == half && (q & 1))) q += 1;
        if (q >> 24) {
            q >>= 1;
            sh += 1;
        }
        m = q;
        L += sh;
    }
    /* m < 2^24: normalise and assemble the FP32 bits (ldexpf outside the
     * normal range). */
    uint32_t mm = (uint32_t)m;
    const int nbm = 32 - __builtin_clz(mm);
    const int e = L + nbm - 1;
    if (e >= -126 && e <= 127) {
        mm <<= 24 - nbm;
        const uint32_t bits = ((uint32_t)neg << 31) | ((uint32_t)(e + 127) << 23) | (mm & 0x7fffffu);
        float r;
        memcpy(&r, &bits, 4);
        return r;
    }
    const float v = ldexpf((float)mm, L);
    return neg ? -v : v;
}

/* One step of the MFMA: acc plus 8 products (see orc_mfma_bf16). */
static float mfma_bf16_step(float c, const float *a, const float *b)
{
    int ek[8], sk[8], ep = INT32_MIN;
    uint32_t mk[8];
    for (int k = 0; k < 8; ++k) {
        sk[k] = -1;
        if (a[k] == 0.0f || b[k] == 0.0f) continue;
        int ea, eb;
        uint32_t ma, mb;
        float_parts(a[k], 7, &ea, &ma);
        float_parts(b[k], 7, &eb, &mb);
        ek[k] = ea + eb;
        mk[k] = ma * mb;  /* |a b| = mk * 2^(ek - 14) */
        sk[k] = (int)((f32_bits(a[k]) ^ f32_bits(b[k])) >> 31);
        if (ek[k] > ep) ep = ek[k];
    }
    if (ep == INT32_MIN) return c;
    int ec = INT32_MIN;
    uint32_t mc = 0;
    if (c != 0.0f) float_parts(c, 23, &ec, &mc);
    const int E = ec > ep ? ec : ep;
    if (E - ep > 90) return c;  /* every product lies far below the 32 bits kept of acc */
    const int L1 = ep - 24;
    __int128 s = 0;
    if (c != 0.0f) {
        const int sh = ec - 23 - L1; /* acc = mc * 2^(ec - 23), floored to 2^L1 */
        const int cneg = (int)(f32_bits(c) >> 31);
        __int128 v;
        if (sh >= 0) {
            v = (__int128)mc << sh;
            if (cneg) v = -v;
        } else if (-sh >= 64) {
            v = cneg ? -1 : 0;
        } else {
            const uint64_t d = (uint64_t)1 << -sh;
            v = cneg ? -(__int128)(((uint64_t)mc + d - 1) >> -sh) : (__int128)((uint64_t)mc >> -sh);
        }
        s += v;
    }
    for (int k = 0; k < 8; ++k) {
        if (sk[k] < 0) continue;
        const int t = ek[k] - ep + 10; /* product / 2^L1 = mk * 2^t, truncated toward zero */
        __int128 v = t >= 0 ? (__int128)mk[k] << t : (-t >= 32 ? 0 : (__int128)(mk[k] >> -t));
        s += sk[k] ? -v : v;
    }
    /* The sum keeps 32 significant bits: floored to 2^(e(s) - 31). */
    int L = L1;
    if (s != 0) {
        const unsigned __int128 m = s < 0 ? (unsigned __int128)(-s) : (unsigned __int128)s;
        const uint64_t mh = (uint64_t)(m >> 64), ml = (uint64_t)m;
        const int nb = mh ? 128 - __builtin_clzll(mh) : 64 - __builtin_clzll(ml);
        const int L2 = L1 + nb - 1 - 31;
        if (L2 > L1) {
            s >>= (L2 - L1); /* arithmetic shift: floor */
            L = L2;
        }
    }
    return round_scaled(s, L);
}

float orc_mfma_bf16(float c, const float *a, const float *b)
{
    return mfma_bf16_step(mfma_bf16_step(c, a, b), a + 8, b + 8);
}

/* Rows of the split field; Jb holds BF16-rounded J, hi/lo hold
 * round_up(N, 16) split values of x (zero past N).  lo == NULL: dSB (one pass). */
static void field_rows_split(int N, const float *Jb, int ldj, const float *hi, const float *lo, float *F,
                             int row_threads)
{
    const int Kp = (N + 15) / 16 * 16;
#ifdef _OPENMP
#pragma omp parallel for num_threads(row_threads) schedule(static) if (row_threads > 1)
#endif
    for (int i = 0; i < N; ++i) {
        const float *row = Jb + (size_t)i * ldj;
        float jr[16], acc = 0.0f;
        for (int g = 0; g < Kp; g += 16) {
            for (int q = 0; q < 16; ++q) jr[q] = g + q < N ? row[g + q] : 0.0f;
            acc = orc_mfma_bf16(acc, hi + g, jr);
            if (lo) acc = orc_mfma_bf16(acc, lo + g, jr);
        }
        F[i] = acc;
    }
}

/* BF16-rounded copy of J (the split form's J). */
static float *bf16_copy(int N, const float *J, int ldj)
{
    float *Jb = (float *)malloc(sizeof(float) * (size_t)N * (size_t)ldj);
    if (!Jb) return NULL;
    for (size_t q = 0; q < (size_t)N * (size_t)ldj; ++q) Jb[q] = bf16_rne(J[q]);
    return Jb;
}

static void field_order(int order, int variant, int N, const float *J, int ldj, const float *x,
                        float *s, float *F, int row_threads);

void orc_field_split(int N, int R, const float *J, int ldj, const float *x, int ld, int use_sign, float *F)
{
    float *Jb = bf16_copy(N, J, ldj);
    float *s = (float *)malloc(sizeof(float) * (size_t)(2 * N + 64));
    if (Jb && s)
        for (int r = 0; r < R; ++r)
            field_order(ORC_ORDER_SPLIT, use_sign ? ORC_DSB : ORC_BSB, N, Jb, ldj, x + (size_t)r * ld, s,
                        F + (size_t)r * ld, 1);
    free(s);
    free(Jb);
}

void orc_field_f32_mfma(int N, int R, const float *J, int ldj, const float *x, int ld,
                        int use_sign, float *F)
{
    float *s = (float *)malloc(sizeof(float) * (size_t)(N + 32));
    for (int r = 0; r < R; ++r)
        field_order(ORC_ORDER_MFMA, use_sign ? ORC_DSB : ORC_BSB, N, J, ldj, x + (size_t)r * ld, s,
                    F + (size_t)r * ld, 1);
    free(s);
}

void orc_field_f32_lanes(int N, int R, const float *J, int ldj, const float *x, int ld,
                         int use_sign, float *F)
{
    float *s = (float *)malloc(sizeof(float) * (size_t)(N + 16));
    for (int r = 0; r < R; ++r)
        field_order(ORC_ORDER_LANES, use_sign ? ORC_DSB : ORC_BSB, N, J, ldj, x + (size_t)r * ld, s,
                    F + (size_t)r * ld, 1);
    free(s);
}

void orc_field_f32(int N, int R, const float *J, int ldj, const float *x, int ld,
                   int use_sign, float *F)
{
    float *s = use_sign ? (float *)malloc(sizeof(float) * (size_t)(N > 0 ? N : 1)) : NULL;
    for (int r = 0; r < R; ++r) {
        const float *xr = x + (size_t)r * ld;
        const float *src = xr;
        if (use_sign) {
            for (int j = 0; j < N; ++j) s[j] = sgnf(xr[j]);
            src = s;
        }
        field_rows_f32(N, J, ldj, src, F + (size_t)r * ld, 1);
    }
    free(s);
}

/* FP64 field, each row the sequential j = 0..N-1 fma chain; eight rows at a
 * time (independent chains), row blocks over row_threads OpenMP threads. */
static void field_rows_f64(int N, const float *J, int ldj, const double *xs, double *F, int row_threads)
{
    const int nb = N / 8;
#ifdef _OPENMP
#pragma omp parallel for num_threads(row_threads) schedule(static) if (row_threads > 1)
#endif
    for (int b = 0; b < nb; ++b) {
        const int i = 8 * b;
        double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        const float *r0 = J + (size_t)i * ldj;
        for (int j = 0; j < N; ++j) {
            const double v = xs[j];
            for (int q = 0; q < 8; ++q) a[q] = fma((double)r0[(size_t)q * ldj + j], v, a[q]);
        }
        for (int q = 0; q < 8; ++q) F[i + q] = a[q];
    }
    for (int i = 8 * nb; i < N; ++i) {
        const float *row = J + (size_t)i * ldj;
        double a = 0.0;
        for (int j = 0; j < N; ++j) a = fma((double)row[j], xs[j], a);
        F[i] = a;
    }
}

static void field_f64_one(int N, const float *J, int ldj, const double *xr, int use_sign, double *s,
                          double *F, int row_threads)
{
    const double *src = xr;
    if (use_sign) {
        for (int j = 0; j < N; ++j) s[j] = xr[j] > 0.0 ? 1.0 : -1.0;
        src = s;
    }
    field_rows_f64(N, J, ldj, src, F, row_threads);
}

void orc_field_f64(int N, int R, const float *J, int ldj, const double *x, int ld,
                   int use_sign, double *F)
{
    double *s = (double *)malloc(sizeof(double) * (size_t)(N > 0 ? N : 1));
    for (int r = 0; r < R; ++r)
        field_f64_one(N, J, ldj, x + (size_t)r * ld, use_sign, s, F + (size_t)r * ld, 1);
    free(s);
}

/* ---- updates ----------------------------------------------------------- */

/* bSB / dSB, Goto et al. 2021: y += (-(a0 - a) x + c0 F) dt; x += a0 y dt;
 * inelastic wall at |x| = 1.  With alpha = dt (a0 - a), beta = dt c0,
 * gamma = dt a0 this is the op sequence below. */
void orc_bsb_update_f32(float F, float alpha, float beta, float gamma, float *x,
                        float *y)
{
    float yn = fmaf(beta, F, fmaf(-alpha, *x, *y));
    float xn = fmaf(gamma, yn, *x);
    if (fabsf(xn) > 1.0f) {
        xn = copysignf(1.0f, xn);
        yn = 0.0f;
    }
    *x = xn;
    *y = yn;
}

static void bsb_update_f64(double F, double alpha, double beta, double gamma,
                           double *x, double *y)
{
    double yn = fma(beta, F, fma(-alpha, *x, *y));
    double xn = fma(gamma, yn, *x);
    if (fabs(xn) > 1.0) {
        xn = copysign(1.0, xn);
        yn = 0.0;
    }
    *x = xn;
    *y = yn;
}

/* SB/CUDA_GPU/m2_opt.cu:58-75.  detune, deltaT and K are double literals in
 * the reference, so each statement is evaluated in FP64 from the FP32
 * operands and rounded to FP32 on assignment:
 *   lx = lx + detune * ly * deltaT;
 *   ly = ly - deltaT * (K * lx * lx * lx + (detune - amplitude) * lx); */
void orc_asb_local_f32(float *x, float *y, float p, double a0, double K,
                       double dts, int M)
{
    float lx = *x, ly = *y;
    for (int m = 0; m < M; ++m) {
        double t = a0 * (double)ly;
        t = t * dts;
        lx = (float)((double)lx + t);
        double xd = (double)lx;
        double cube = K * xd;
        cube = cube * xd;
        cube = cube * xd;
        double lin = a0 - (double)p;
        lin = lin * xd;
        double u = cube + lin;
        u = dts * u;
        ly = (float)((double)ly - u);
    }
    *x = lx;
    *y = ly;
}

static void asb_local_f64(double *x, double *y, double p, double a0, double K,
                          double dts, int M)
{
    double lx = *x, ly = *y;
    for (int m = 0; m < M; ++m) {
        lx = lx + (a0 * ly) * dts;
        ly = ly - dts * (((K * lx) * lx) * lx + (a0 - p) * lx);
    }
    *x = lx;
    *y = ly;
}

/* SB/CUDA_GPU/Bifurcation_m1.cu.c:55-69 (UpdateX).  detune is an int and
 * DeltaT a double literal: detune * y is an FP32 product, the rest is
 * evaluated in FP64 and rounded to FP32 on assignment. */
float orc_m1_x_f32(float x, float y, double a0, double dt)
{
    float t = (float)a0 * y;
    return (float)((double)t * dt + (double)x);
}

/* SB/CUDA_GPU/Bifurcation_m1.cu.c:71-92 (UpdateY), with the coupling sum F
 * given:
 *   newY = -K * (x * x * x) * DeltaT + amplitude * x * DeltaT - detune * x * DeltaT;
 *   s = s * xi * DeltaT;  newY += s;  newY += oldY;
 * x * x * x, (-K) * that, amplitude * x and detune * x are FP32 products; each
 * is then multiplied by DeltaT in FP64 and the three are summed in FP64 as
 * (A + B) - C and rounded to FP32; s * xi * DeltaT is FP64, rounded to FP32;
 * the two += are FP32 additions. */
float orc_m1_y_f32(float F, float x, float y, float p, double a0, double K, double dt, double xi)
{
    float x3 = x * x;
    x3 = x3 * x;
    float kx3 = -(float)K * x3;
    float px = p * x;
    float dx = (float)a0 * x;
    double A = (double)kx3 * dt;
    double B = (double)px * dt;
    double C = (double)dx * dt;
    double u = A + B;
    u = u - C;
    float ny = (float)u;
    double sd = (double)F * xi;
    sd = sd * dt;
    float s = (float)sd;
    ny = ny + s;
    ny = ny + y;
    return ny;
}

/* ---- solver loops ------------------------------------------------------ */

/* Field of one replica in the chosen order; s has round_up(N, 16) entries.
 * row_threads > 1 splits the rows of the SEQ order over OpenMP threads. */
static void field_order(int order, int variant, int N, const float *J, int ldj, const float *x,
                        float *s, float *F, int row_threads)
{
    const int Kp = (N + 15) / 16 * 16;
    const int sign = variant == ORC_DSB;
    if (order == ORC_ORDER_SEQ) {
        const float *src = x;
        if (sign) {
            for (int j = 0; j < N; ++j) s[j] = sgnf(x[j]);
            src = s;
        }
        field_rows_f32(N, J, ldj, src, F, row_threads);
        return;
    }
    if (order == ORC_ORDER_SPLIT) {
        /* J is the BF16-rounded copy; s holds hi then lo, Kp values each. */
        float *hi = s, *lo = s + Kp;
        for (int j = 0; j < Kp; ++j) {
            const float v = j < N ? x[j] : 0.0f;
            if (sign) {
                hi[j] = j < N ? sgnf(v) : 0.0f;
            } else {
                hi[j] = bf16_rne(v);
                lo[j] = bf16_rne(v - hi[j]);
            }
        }
        field_rows_split(N, J, ldj, hi, sign ? NULL : lo, F, row_threads);
        return;
    }
    if (order == ORC_ORDER_MFMA) {
        const int Kp32 = (N + 31) / 32 * 32;
        for (int j = 0; j < Kp32; ++j) {
            const float v = j < N ? x[j] : 0.0f;
            s[j] = sign ? sgnf(v) : v;
        }
        field_rows_f32_mfma(N, J, ldj, s, F, row_threads);
        return;
    }
    for (int j = 0; j < Kp; ++j) {
        const float v = j < N ? x[j] : 0.0f;
        s[j] = sign ? sgnf(v) : v;
    }
    field_rows_f32_lanes(N, J, ldj, s, F);
}

/* External field: F_i = (J.x)_i + h_i, one FP32 addition after the sum. */
static void add_field(int N, const float *h, float *F)
{
    if (h)
        for (int i = 0; i < N; ++i) F[i] = F[i] + h[i];
}

static int run_f32_replica(int order, int variant, int N, const float *J, int ldj, const float *h,
                           float *x, float *y, float *F, float *s, int k0, int k1, int steps,
                           double dt, double a0, double c0, double K, int M, int row_threads)
{
    double dts = dt / (double)M;
    for (int k = k0; k < k1; ++k) {
        float cf[4];
        orc_coeffs(variant, k, steps, dt, a0, c0, cf);
        if (variant == ORC_ASB_M1) {
            /* Bifurcation_m1.cu.c:166-169: UpdateX over all spins, then
             * UpdateY with the coupling of the new x. */
            const float p = (float)k / (float)steps;
            for (int i = 0; i < N; ++i) x[i] = orc_m1_x_f32(x[i], y[i], a0, dt);
            field_order(order, ORC_BSB, N, J, ldj, x, s, F, row_threads);
            add_field(N, h, F);
            for (int i = 0; i < N; ++i) y[i] = orc_m1_y_f32(F[i], x[i], y[i], p, a0, K, dt, c0);
        } else if (variant == ORC_ASB) {
            /* m2_opt.cu:168-169: local sub-steps, then the kick with new x */
            for (int i = 0; i < N; ++i) orc_asb_local_f32(&x[i], &y[i], cf[3], a0, K, dts, M);
            field_order(order, variant, N, J, ldj, x, s, F, row_threads);
            add_field(N, h, F);
            for (int i = 0; i < N; ++i) y[i] = fmaf(cf[1], F[i], y[i]);
        } else {
            field_order(order, variant, N, J, ldj, x, s, F, row_threads);
            add_field(N, h, F);
            for (int i = 0; i < N; ++i) orc_bsb_update_f32(F[i], cf[0], cf[1], cf[2], &x[i], &y[i]);
        }
    }
    return ORC_OK;
}

static int check_run_args(int variant, int N, int R, int ldj, int ld, int k0, int k1,
                          int steps, int M)
{
    if (variant < ORC_ASB || variant > ORC_ASB_M1) return ORC_E_INVALID;
    if (N < 1 || R < 0 || ldj < N || ld < N || steps < 1) return ORC_E_INVALID;
    if (k0 < 0 || k1 < k0 || k1 > steps) return ORC_E_INVALID;
    if (variant == ORC_ASB && M < 1) return ORC_E_INVALID;
    return ORC_OK;
}

int orc_run_f32(int variant, int N, int R, const float *J, int ldj, float *x,
                float *y, int ld, int k0, int k1, int steps, double dt, double a0,
                double c0, double K, int M, int nthreads)
{
    return orc_run_f32_order(ORC_ORDER_SEQ, variant, N, R, J, ldj, NULL, x, y, ld, k0, k1, steps,
                             dt, a0, c0, K, M, nthreads);
}

int orc_run_f32_order(int order, int variant, int N, int R, const float *J, int ldj, const float *h,
                      float *x, float *y, int ld, int k0, int k1, int steps, double dt, double a0,
                      double c0, double K, int M, int nthreads)
{
    int rc = check_run_args(variant, N, R, ldj, ld, k0, k1, steps, M);
    if (rc) return rc;
    if (order != ORC_ORDER_SEQ && order != ORC_ORDER_LANES && order != ORC_ORDER_MFMA && order != ORC_ORDER_SPLIT)
        return ORC_E_INVALID;
    if (!orc_has_fma()) return ORC_E_NOFMA;
    if (nthreads < 1) nthreads = 1;
    /* The split form multiplies BF16-rounded J. */
    float *Jb = NULL;
    if (order == ORC_ORDER_SPLIT) {
        Jb = bf16_copy(N, J, ldj);
        if (!Jb) return ORC_E_INVALID;
        J = Jb;
    }
    if (nthreads > 1 && R < nthreads && order != ORC_ORDER_LANES) {
        /* Fewer replicas than threads (e.g. R = 1): replicas one after the
         * other, the rows of each field over the threads (same sums). */
        float *F = (float *)malloc(sizeof(float) * (size_t)N);
        float *s = (float *)malloc(sizeof(float) * (size_t)(2 * N + 64));
        for (int r = 0; r < R; ++r)
            run_f32_replica(order, variant, N, J, ldj, h, x + (size_t)r * ld, y + (size_t)r * ld,
                            F, s, k0, k1, steps, dt, a0, c0, K, M, nthreads);
        free(F);
        free(s);
        free(Jb);
        return ORC_OK;
    }
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads)
#endif
    {
        float *F = (float *)malloc(sizeof(float) * (size_t)N);
        float *s = (float *)malloc(sizeof(float) * (size_t)(2 * N + 64));
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int r = 0; r < R; ++r)
            run_f32_replica(order, variant, N, J, ldj, h, x + (size_t)r * ld, y + (size_t)r * ld,
                            F, s, k0, k1, steps, dt, a0, c0, K, M, 1);
        free(F);
        free(s);
    }
    free(Jb);
    return ORC_OK;
}

int orc_run_f64(int variant, int N, int R, const float *J, int ldj, double *x,
                double *y, int ld, int k0, int k1, int steps, double dt, double a0,
                double c0, double K, int M, int nthreads)
{
    return orc_run_f64_h(variant, N, R, J, ldj, NULL, x, y, ld, k0, k1, steps, dt, a0, c0, K, M,
                         nthreads);
}

int orc_run_f64_h(int variant, int N, int R, const float *J, int ldj, const float *h, double *x,
                  double *y, int ld, int k0, int k1, int steps, double dt, double a0,
                  double c0, double K, int M, int nthreads)
{
    int rc = check_run_args(variant, N, R, ldj, ld, k0, k1, steps, M);
    if (rc) return rc;
    double dts = dt / (double)M;
    if (nthreads < 1) nthreads = 1;
    /* Fewer replicas than threads: replica-parallel over R threads, the rows
     * of each field over the remaining share (same sums either way). */
    const int rep_threads = R < nthreads ? (R > 0 ? R : 1) : nthreads;
    const int row_threads = nthreads / rep_threads > 1 ? nthreads / rep_threads : 1;
#ifdef _OPENMP
    omp_set_max_active_levels(2);
#pragma omp parallel num_threads(rep_threads)
#endif
    {
        double *F = (double *)malloc(sizeof(double) * (size_t)N);
        double *sg = (double *)malloc(sizeof(double) * (size_t)N);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int r = 0; r < R; ++r) {
            double *xr = x + (size_t)r * ld, *yr = y + (size_t)r * ld;
            for (int k = k0; k < k1; ++k) {
                double p = (double)k / (double)steps;
                if (variant == ORC_ASB_M1) {
                    /* the same step in FP64 throughout */
                    for (int i = 0; i < N; ++i) xr[i] = xr[i] + a0 * yr[i] * dt;
                    field_f64_one(N, J, ldj, xr, 0, sg, F, row_threads);
                    if (h)
                        for (int i = 0; i < N; ++i) F[i] += (double)h[i];
                    for (int i = 0; i < N; ++i) {
                        const double xi_ = xr[i];
                        yr[i] = (-K * xi_ * xi_ * xi_ * dt + p * xi_ * dt - a0 * xi_ * dt) + F[i] * c0 * dt + yr[i];
                    }
                } else if (variant == ORC_ASB) {
                    for (int i = 0; i < N; ++i) asb_local_f64(&xr[i], &yr[i], p, a0, K, dts, M);
                    field_f64_one(N, J, ldj, xr, 0, sg, F, row_threads);
                    if (h)
                        for (int i = 0; i < N; ++i) F[i] += (double)h[i];
                    for (int i = 0; i < N; ++i) yr[i] = fma(dt * c0, F[i], yr[i]);
                } else {
                    field_f64_one(N, J, ldj, xr, variant == ORC_DSB, sg, F, row_threads);
                    if (h)
                        for (int i = 0; i < N; ++i) F[i] += (double)h[i];
                    double alpha = dt * (a0 - a0 * p), beta = dt * c0, gamma = dt * a0;
                    for (int i = 0; i < N; ++i) bsb_update_f64(F[i], alpha, beta, gamma, &xr[i], &yr[i]);
                }
            }
        }
        free(F);
        free(sg);
    }
    return ORC_OK;
}

/* ---- readout and energy ------------------------------------------------ */

void orc_spins_f32(int N, int R, const float *x, int ld, int8_t *s)
{
    for (int r = 0; r < R; ++r)
        for (int i = 0; i < N; ++i)
            s[(size_t)r * N + i] = x[(size_t)r * ld + i] > 0.0f ? 1 : -1;
}

/* SB/CUDA_GPU/m2_opt.cu:192-202 with an int64 accumulator. */
int orc_energy_i64(int N, int R, const float *J, int ldj, const int8_t *s, int64_t *H)
{
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            float v = J[(size_t)i * ldj + j];
            if (v != floorf(v) || fabsf(v) > 16777216.0f) return ORC_E_INVALID;
        }
    for (int r = 0; r < R; ++r) {
        const int8_t *sr = s + (size_t)r * N;
        int64_t E = 0;
        for (int i = 0; i < N; ++i)
            for (int j = i + 1; j < N; ++j)
                E += (int64_t)sr[i] * (int64_t)sr[j] * (int64_t)J[(size_t)i * ldj + j];
        H[r] = -E;
    }
    return ORC_OK;
}

void orc_energy_f64(int N, int R, const float *J, int ldj, const int8_t *s, double *H)
{
    for (int r = 0; r < R; ++r) {
        const int8_t *sr = s + (size_t)r * N;
        double E = 0.0;
        for (int i = 0; i < N; ++i)
            for (int j = i + 1; j < N; ++j)
                E += (double)(sr[i] * sr[j]) * (double)J[(size_t)i * ldj + j];
        H[r] = -E;
    }
}

/* The external-field term -sum_i h_i s_i added to H (int64 when h is
 * integer-valued; returns ORC_E_INVALID otherwise). */
int orc_energy_field_i64(int N, int R, const float *h, const int8_t *s, int64_t *H)
{
    for (int i = 0; i < N; ++i)
        if (h[i] != floorf(h[i])) return ORC_E_INVALID;
    for (int r = 0; r < R; ++r) {
        int64_t E = 0;
        for (int i = 0; i < N; ++i) E += (int64_t)s[(size_t)r * N + i] * (int64_t)h[i];
        H[r] -= E;
    }
    return ORC_OK;
}

void orc_energy_field_f64(int N, int R, const float *h, const int8_t *s, double *H)
{
    for (int r = 0; r < R; ++r) {
        double E = 0.0;
        for (int i = 0; i < N; ++i) E += (double)s[(size_t)r * N + i] * (double)h[i];
        H[r] -= E;
    }
}

int orc_ground_state(int N, const float *J, int ldj, double *H, int8_t *s)
{
    if (N < 1 || N > 24) return ORC_E_INVALID;
    double best = INFINITY;
    uint32_t best_m = 0;
    for (uint32_t m = 0; m < (1u << N); ++m) {
        double E = 0.0;
        for (int i = 0; i < N; ++i) {
            int si = (m >> i) & 1 ? 1 : -1;
            for (int j = i + 1; j < N; ++j) {
                int sj = (m >> j) & 1 ? 1 : -1;
                E += (double)(si * sj) * (double)J[(size_t)i * ldj + j];
            }
        }
        if (-E < best) {
            best = -E;
            best_m = m;
        }
    }
    *H = best;
    for (int i = 0; i < N; ++i) s[i] = (best_m >> i) & 1 ? 1 : -1;
    return ORC_OK;
}

/* ---- edge list --------------------------------------------------------- */

typedef struct {
    const char *p, *end;
} cursor;

static void skip_ws(cursor *c)
{
    while (c->p < c->end && isspace((unsigned char)*c->p)) ++c->p;
}

static int read_long(cursor *c, long *out)
{
    skip_ws(c);
    if (c->p >= c->end) return 0;
    char buf[32];
    size_t n = 0;
    while (c->p < c->end && n < sizeof(buf) - 1 && !isspace((unsigned char)*c->p))
        buf[n++] = *c->p++;
    buf[n] = 0;
    char *e = NULL;
    long v = strtol(buf, &e, 10);
    if (e == buf || *e) return -1;
    *out = v;
    return 1;
}

static int read_double(cursor *c, double *out)
{
    skip_ws(c);
    if (c->p >= c->end) return 0;
    char buf[64];
    size_t n = 0;
    while (c->p < c->end && n < sizeof(buf) - 1 && !isspace((unsigned char)*c->p))
        buf[n++] = *c->p++;
    buf[n] = 0;
    char *e = NULL;
    double v = strtod(buf, &e);
    if (e == buf || *e || !isfinite(v)) return -1;
    *out = v;
    return 1;
}

int orc_edgelist_header(const char *text, size_t len, int *N)
{
    cursor c = {text, text + len};
    long n;
    if (read_long(&c, &n) != 1 || n < 1 || n > (1L << 30)) return ORC_E_PARSE;
    *N = (int)n;
    return ORC_OK;
}

int orc_edgelist_fill(const char *text, size_t len, int one_based, int N, float *J,
                      int ldj)
{
    return orc_edgelist_fill_h(text, len, one_based, N, J, ldj, NULL);
}

int orc_edgelist_fill_h(const char *text, size_t len, int one_based, int N, float *J,
                        int ldj, float *h)
{
    if (h) memset(h, 0, sizeof(float) * (size_t)N);
    cursor c = {text, text + len};
    long n;
    if (read_long(&c, &n) != 1 || n != N) return ORC_E_PARSE;
    for (int i = 0; i < N; ++i) memset(J + (size_t)i * ldj, 0, sizeof(float) * (size_t)N);
    for (;;) {
        long a, b;
        double w;
        int ra = read_long(&c, &a);
        if (ra == 0) break;
        if (ra < 0 || read_long(&c, &b) != 1 || read_double(&c, &w) != 1) return ORC_E_PARSE;
        if (one_based) {
            a -= 1;
            b -= 1;
        }
        if (a < 0 || b < 0 || a >= N || b >= N) return ORC_E_RANGE;
        if (a == b) {
            if (!h) return ORC_E_INVALID;
            h[a] = (float)w; /* "i i w": external field */
            continue;
        }
        J[(size_t)a * ldj + b] = (float)w;
        J[(size_t)b * ldj + a] = (float)w;
    }
    return ORC_OK;
}

int orc_has_fma(void)
{
#if defined(__x86_64__) || defined(__i386__)
    __builtin_cpu_init();
    return __builtin_cpu_supports("fma") ? 1 : 0;
#else
    return 1;
#endif
}
