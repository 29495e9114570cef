/*
 * CPU oracle for the NeuronGPU per-timestep hot path (see oracle.h).
 * TEST INFRASTRUCTURE ONLY: never linked or loaded by the product library.
 *
 * Built twice: -DORC_REAL_DOUBLE=0 (FP32, bit-comparable with the HIP path)
 * and -DORC_REAL_DOUBLE=1 (FP64 accuracy reference).  Compile with
 * -ffp-contract=off: every fused multiply-add below is an explicit fma call,
 * matching the HIP kernels operation for operation.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#if ORC_REAL_DOUBLE
typedef double REAL;
#define FMA(a, b, c) fma((a), (b), (c))
#define EXPR(x) exp(x)
#define POW_DEC(e) pow((e), -0.25)
#define POW_INC(e) pow((e), -0.2)
#define ROUNDR(x) round(x)
#define SQRTR(x) sqrt(x)
#else
typedef float REAL;
#define FMA(a, b, c) fmaf((a), (b), (c))
#define EXPR(x) orc_expf(x)
/* err^(-1/4) and err^(-1/5) by one polynomial sequence (orc_pow_ctrl) */
#define POW_DEC(e) orc_pow_ctrl((e), -0.25f)
#define POW_INC(e) orc_pow_ctrl((e), -0.2f)
#define ROUNDR(x) roundf(x)
#define SQRTR(x) sqrtf(x)
#endif

/* input accumulator type: REAL, or double with ORC_ACC_DOUBLE (the
 * reference accumulates the step's synaptic input in double,
 * src/get_spike.cu:70-72 and :134-138) */
#ifndef ORC_ACC_DOUBLE
#define ORC_ACC_DOUBLE 0
#endif
#if ORC_ACC_DOUBLE
typedef double ACC;
#else
typedef REAL ACC;
#endif

#define MAX_PORT 20
#define MAX_ATTEMPTS_PER_STEP 100000
#define POISSON_TABLE_MAX 256
#define MAX_SYN_DT 16384
#define LAST_SPIKE_TIME_GUARD 0x70000000LL

int orc_real_size(void) { return (int)sizeof(REAL); }

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).                   */
/* ------------------------------------------------------------------ */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* ------------------------------------------------------------------ */
/* Threefry2x32-20 (Salmon, Moraes, Dror, Shaw, SC'11): Threefish rounds */
/* (add, rotate left, xor) with rotations 13 15 26 6 17 29 16 24 and a   */
/* key injection after every fourth round from the schedule             */
/* (k0, k1, 0x1BD11BDA ^ k0 ^ k1), plus the injection count on word 1.  */
/* ------------------------------------------------------------------ */
void orc_threefry2x32_20(const uint32_t ctr[2], const uint32_t key[2],
                         uint32_t out[2]) {
  static const int rot[8] = {13, 15, 26, 6, 17, 29, 16, 24};
  const uint32_t ks[3] = {key[0], key[1], 0x1BD11BDAu ^ key[0] ^ key[1]};
  uint32_t x0 = ctr[0] + ks[0], x1 = ctr[1] + ks[1];
  for (int r = 0; r < 20; r++) {
    x0 += x1;
    x1 = (x1 << rot[r % 8]) | (x1 >> (32 - rot[r % 8]));
    x1 ^= x0;
    if (r % 4 == 3) {
      uint32_t s = (uint32_t)(r + 1) / 4;
      x0 += ks[s % 3];
      x1 += ks[(s + 1) % 3] + s;
    }
  }
  out[0] = x0;
  out[1] = x1;
}

/* Poisson draw word of the direct connections and CreatePoissonGenerator
 * nodes: Threefry2x32-20 at counter (index, T mod 2^32), key
 * (seed_lo ^ tag, seed_hi + T / 2^32), word 0. */
static uint32_t orc_pois_word(uint64_t seed, uint32_t index, long long T,
                              uint32_t tag) {
  uint32_t ctr[2] = {index, (uint32_t)T};
  uint32_t key[2] = {(uint32_t)seed ^ tag,
                     (uint32_t)(seed >> 32) + (uint32_t)((uint64_t)T >> 32)};
  uint32_t out[2];
  orc_threefry2x32_20(ctr, key, out);
  return out[0];
}

/* Poisson CDF table: tab[k] = ceil(2^32 F(k)) - 1 clamped to 2^32-1, where
 * F is the Poisson(lambda) CDF computed in double by the pmf recurrence.
 * A draw is X = min{k : u <= tab[k]} for u uniform on 32 bits
 * (inversion by sequential search, Devroye 1986 ch. X.3).
 * Returns the table length, or -1 if max_len entries do not reach F = 1. */
int orc_poisson_table(double lambda, uint32_t *tab, int max_len) {
  double p = exp(-lambda);
  double F = p;
  for (int k = 0; k < max_len; k++) {
    double v = ceil(F * 4294967296.0) - 1.0;
    uint32_t t;
    if (v >= 4294967295.0) t = 0xFFFFFFFFu;
    else if (v <= 0.0) t = 0u;
    else t = (uint32_t)v;
    tab[k] = t;
    if (t == 0xFFFFFFFFu) return k + 1;
    p *= lambda / (double)(k + 1);
    F += p;
  }
  return -1;
}

int orc_poisson_sample(const uint32_t *tab, uint32_t u) {
  int k = 0;
  while (u > tab[k]) k++;
  return k;
}

/* ------------------------------------------------------------------ */
/* FP32 exp and log (own implementation, mirrored in the HIP kernels).  */
/* exp: Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2, exp(r) by its  */
/* degree-7 Taylor polynomial (Horner, fma), scaling by 2^k with ldexp    */
/* (exact, or rounded once where the result is subnormal).              */
/* ------------------------------------------------------------------ */
static inline float orc_u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline uint32_t orc_f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

float orc_expf(float x) {
  if (x != x) return x;
  if (x > 88.8f) return orc_u2f(0x7F800000u);
  if (x < -104.0f) return 0.0f;
  float kf = rintf(x * 1.44269502f);
  float r = fmaf(kf, -0.693145751953125f, x);
  r = fmaf(kf, -1.42860677e-06f, r);
  float p = 1.98412701e-04f;          /* 1/5040 */
  p = fmaf(p, r, 1.38888892e-03f);    /* 1/720 */
  p = fmaf(p, r, 8.33333377e-03f);    /* 1/120 */
  p = fmaf(p, r, 4.16666679e-02f);    /* 1/24 */
  p = fmaf(p, r, 1.66666672e-01f);    /* 1/6 */
  p = fmaf(p, r, 0.5f);
  p = fmaf(p, r, 1.0f);
  p = fmaf(p, r, 1.0f);
  return ldexpf(p, (int)kf);
}

/* log: x = 2^e m, m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s),
 * s = (m-1)/(m+1), atanh by its odd series to s^9; e ln2 added in two parts.
 * Defined for normal positive x; +inf -> +inf; NaN -> NaN. */
float orc_logf(float x) {
  if (x != x) return x;
  if (x <= 0.0f) return orc_u2f(0xFF800000u);
  if (x == orc_u2f(0x7F800000u)) return x;
  uint32_t b = orc_f2u(x);
  int e = (int)((b >> 23) & 0xFFu) - 127;
  float m = orc_u2f((b & 0x007FFFFFu) | 0x3F800000u);
  if (m > 1.41421354f) { m = m * 0.5f; e = e + 1; }
  float s = (m - 1.0f) / (m + 1.0f);
  float s2 = s * s;
  float q = 1.11111112e-01f;          /* 1/9 */
  q = fmaf(q, s2, 1.42857149e-01f);   /* 1/7 */
  q = fmaf(q, s2, 2.00000003e-01f);   /* 1/5 */
  q = fmaf(q, s2, 3.33333343e-01f);   /* 1/3 */
  float two_s = 2.0f * s;
  float lm = fmaf(two_s * s2, q, two_s);
  float ef = (float)e;
  return fmaf(ef, 0.693145751953125f, fmaf(ef, 1.42860677e-06f, lm));
}

/* err^c of the step-size controller (src/rk5.h:122-146: pow(err, -0.20)
 * after an accepted attempt, pow(err, -0.25) after a rejected one), with no
 * division: err = 2^k m, m in [sqrt(1/2), sqrt(2)); log2 m = t Q(t),
 * t = m - 1; Q of degree 7, a least-squares fit of log2(1+t)/t with float
 * coefficients; y = c k + c log2 m; 2^y = 2^j R(y - j), j = rint(y), R of
 * degree 5 fitted to 2^f on [-1/2, 1/2].  +inf -> 0, NaN -> NaN. */
float orc_pow_ctrl(float e, float c) {
  if (e != e) return e;
  if (e == orc_u2f(0x7F800000u)) return 0.0f;
  uint32_t b = orc_f2u(e);
  int k = (int)((b >> 23) & 0xFFu) - 127;
  float m = orc_u2f((b & 0x007FFFFFu) | 0x3F800000u);
  if (m > 1.41421354f) { m = m * 0.5f; k = k + 1; }
  float t = m - 1.0f;
  float q = -0.14237648f;
  q = fmaf(q, t, 0.23247829f);
  q = fmaf(q, t, -0.24932154f);
  q = fmaf(q, t, 0.28731275f);
  q = fmaf(q, t, -0.36022335f);
  q = fmaf(q, t, 0.48091587f);
  q = fmaf(q, t, -0.72135293f);
  q = fmaf(q, t, 1.442695f);
  float y = fmaf(c, (float)k, c * (t * q));
  float j = rintf(y);
  float f = y - j;
  float r = 0.0013266971f;
  r = fmaf(r, f, 0.00967546f);
  r = fmaf(r, f, 0.055507425f);
  r = fmaf(r, f, 0.24022122f);
  r = fmaf(r, f, 0.69314694f);
  r = fmaf(r, f, 1.0000001f);
  return ldexpf(r, (int)j);
}

/* STDP rule in double, src/stdp.cu:24-51.  param7 = tau_plus, tau_minus,
 * lambda, alpha, mu_plus, mu_minus, Wmax. */
float orc_stdp_update(float wf, float Dt, const float *param7) {
  double tau_plus = param7[0], tau_minus = param7[1], lambda = param7[2];
  double alpha = param7[3], mu_plus = param7[4], mu_minus = param7[5];
  double Wmax = param7[6];
  double w = wf, w1;
  if (Dt >= 0) {
    double fact = lambda * exp(-(double)Dt / tau_plus);
    w1 = w + fact * Wmax * pow(1.0 - w / Wmax, mu_plus);
  } else {
    double fact = -alpha * lambda * exp((double)Dt / tau_minus);
    w1 = w + fact * Wmax * pow(w / Wmax, mu_minus);
  }
  w1 = w1 > 0.0 ? w1 : 0.0;
  w1 = w1 < Wmax ? w1 : Wmax;
  return (float)w1;
}

/* g0 normalisation of the beta conductance, src/aeif_cond_beta.cu:58-87,
 * evaluated in double. */
double orc_beta_g0(double tau_rise, double tau_decay) {
  double denom1 = tau_decay - tau_rise;
  double denom2 = 0.0;
  if (denom1 != 0.0) {
    double t_p = tau_decay * tau_rise * log(tau_decay / tau_rise) / denom1;
    denom2 = exp(-t_p / tau_decay) - exp(-t_p / tau_rise);
  }
  if (denom2 == 0.0) return M_E / tau_decay;
  return (1.0 / tau_rise - 1.0 / tau_decay) / denom2;
}

/* Exact-integration propagators, src/iaf_psc_exp.cu:53-91, in double.
 * out6 = P20, P11ex, P11in, P21ex, P21in, P22 */
static double orc_p32(double tau_syn, double tau, double C, double h) {
  double P32_linear = 1.0 / (2.0 * C * tau * tau) * h * h * (tau_syn - tau) *
                      exp(-h / tau);
  double P32_singular = h / C * exp(-h / tau);
  double P32 = -tau / (C * (1.0 - tau / tau_syn)) * exp(-h / tau_syn) *
               expm1(h * (1.0 / tau_syn - 1.0 / tau));
  double dev_P32 = fabs(P32 - P32_singular);
  if (tau == tau_syn ||
      (fabs(tau - tau_syn) < 0.1 && dev_P32 > 2.0 * fabs(P32_linear)))
    return P32_singular;
  return P32;
}

void orc_iaf_propagators(double h, double tau_m, double C_m, double tau_ex,
                         double tau_in, float out6[6]) {
  double P22 = exp(-h / tau_m);
  out6[0] = (float)(tau_m / C_m * (1.0 - P22));
  out6[1] = (float)exp(-h / tau_ex);
  out6[2] = (float)exp(-h / tau_in);
  out6[3] = (float)orc_p32(tau_ex, tau_m, C_m, h);
  out6[4] = (float)orc_p32(tau_in, tau_m, C_m, h);
  out6[5] = (float)P22;
}

/* ------------------------------------------------------------------ */
/* Cash-Karp tableau, src/rk5.cu:23-60 (constants formed in double and   */
/* rounded once to REAL).                                               */
/* ------------------------------------------------------------------ */
static const REAL a21 = (REAL)0.2;
static const REAL a31 = (REAL)(3.0 / 40.0), a32 = (REAL)(9.0 / 40.0);
static const REAL a41 = (REAL)0.3, a42 = (REAL)-0.9, a43 = (REAL)1.2;
static const REAL a51 = (REAL)(-11.0 / 54.0), a52 = (REAL)2.5,
                  a53 = (REAL)(-70.0 / 27.0), a54 = (REAL)(35.0 / 27.0);
static const REAL a61 = (REAL)(1631.0 / 55296.0), a62 = (REAL)(175.0 / 512.0),
                  a63 = (REAL)(575.0 / 13824.0),
                  a64 = (REAL)(44275.0 / 110592.0),
                  a65 = (REAL)(253.0 / 4096.0);
static const REAL a71 = (REAL)(37.0 / 378.0), a73 = (REAL)(250.0 / 621.0),
                  a74 = (REAL)(125.0 / 594.0), a76 = (REAL)(512.0 / 1771.0);
static const REAL e1 = (REAL)(37.0 / 378.0 - 2825.0 / 27648.0);
static const REAL e3 = (REAL)(250.0 / 621.0 - 18575.0 / 48384.0);
static const REAL e4 = (REAL)(125.0 / 594.0 - 13525.0 / 55296.0);
static const REAL e5 = (REAL)(-277.00 / 14336.0);
static const REAL e6 = (REAL)(512.0 / 1771.0 - 0.25);
static const REAL coeff = (REAL)0.9;
static const REAL err_min = (REAL)1.889568e-4;
static const REAL scal_min = (REAL)0.1;

typedef void (*orc_rhs_fn)(int nvar, const REAL *y, REAL *dy, const void *ctx);

typedef struct {
  long n_acc, n_rej, n_att;
} orc_rk5_count;

/* One adaptive step, src/rk5.h:57-147.  Returns 0, or 1 if the attempt
 * cap was hit (never in normal operation). */
static int orc_rk5_step(int nvar, double *x, REAL *y, REAL *h, REAL h_min,
                        REAL h_max, REAL eps, orc_rhs_fn rhs, const void *ctx,
                        orc_rk5_count *cnt) {
  REAL k1[64], k2[64], k3[64], k4[64], k5[64], k6[64], yn[64], scal[64];
  REAL err = 0;
  double h_min_acc = (double)h_min * (1.0 + 1.0e-5);
  rhs(nvar, y, k1, ctx);
  for (int i = 0; i < nvar; i++) {
    REAL s = (REAL)fabs((double)y[i]) + (REAL)fabs((double)(k1[i] * (*h)));
    scal[i] = s + scal_min;
  }
  for (;;) {
    REAL hh = *h;
    if (hh > h_max) hh = h_max;
    if (hh < h_min) hh = h_min;
    *h = hh;
    for (int i = 0; i < nvar; i++) {
      REAL t = a21 * k1[i];
      yn[i] = FMA(hh, t, y[i]);
    }
    rhs(nvar, yn, k2, ctx);
    for (int i = 0; i < nvar; i++) {
      REAL t = a31 * k1[i];
      t = FMA(a32, k2[i], t);
      yn[i] = FMA(hh, t, y[i]);
    }
    rhs(nvar, yn, k3, ctx);
    for (int i = 0; i < nvar; i++) {
      REAL t = a41 * k1[i];
      t = FMA(a42, k2[i], t);
      t = FMA(a43, k3[i], t);
      yn[i] = FMA(hh, t, y[i]);
    }
    rhs(nvar, yn, k4, ctx);
    for (int i = 0; i < nvar; i++) {
      REAL t = a51 * k1[i];
      t = FMA(a52, k2[i], t);
      t = FMA(a53, k3[i], t);
      t = FMA(a54, k4[i], t);
      yn[i] = FMA(hh, t, y[i]);
    }
    rhs(nvar, yn, k5, ctx);
    for (int i = 0; i < nvar; i++) {
      REAL t = a61 * k1[i];
      t = FMA(a62, k2[i], t);
      t = FMA(a63, k3[i], t);
      t = FMA(a64, k4[i], t);
      t = FMA(a65, k5[i], t);
      yn[i] = FMA(hh, t, y[i]);
    }
    rhs(nvar, yn, k6, ctx);
    for (int i = 0; i < nvar; i++) {
      REAL t = a71 * k1[i];
      t = FMA(a73, k3[i], t);
      t = FMA(a74, k4[i], t);
      t = FMA(a76, k6[i], t);
      yn[i] = FMA(hh, t, y[i]);
    }
    /* max_i |hh t_i| / scal_i: the largest quotient found by
     * cross-multiplication (a NaN takes the place), only it divided */
    REAL num = 0, den = 1;
    for (int i = 0; i < nvar; i++) {
      REAL t = e1 * k1[i];
      t = FMA(e3, k3[i], t);
      t = FMA(e4, k4[i], t);
      t = FMA(e5, k5[i], t);
      t = FMA(e6, k6[i], t);
      REAL v = (REAL)fabs((double)(hh * t));
      if (!(v * den <= num * scal[i])) {
        num = v;
        den = scal[i];
      }
    }
    err = num / (den * eps);
    cnt->n_att++;
    if (err <= (REAL)1 || (double)hh <= h_min_acc) break;
    cnt->n_rej++;
    if (cnt->n_att > MAX_ATTEMPTS_PER_STEP) return 1;
    REAL h_new = (hh * coeff) * POW_DEC(err);
    REAL h_lo = (REAL)0.1 * hh;
    *h = (h_new > h_lo) ? h_new : h_lo;
  }
  cnt->n_acc++;
  *x += (double)(*h);
  if (err > err_min) {
    *h = ((*h) * coeff) * POW_INC(err);
  } else {
    *h = (REAL)5 * (*h);
  }
  for (int i = 0; i < nvar; i++) y[i] = yn[i];
  return 0;
}

/* ---- linear test ODE for order / tolerance checks ---- */
typedef struct { REAL lambda; } lin_ctx;
static void lin_rhs(int nvar, const REAL *y, REAL *dy, const void *ctx) {
  const lin_ctx *c = (const lin_ctx *)ctx;
  for (int i = 0; i < nvar; i++) dy[i] = c->lambda * y[i];
}

double orc_rk5_linear_fixed(double lambda, double y0, double T, int n_steps) {
  lin_ctx c = {(REAL)lambda};
  REAL y = (REAL)y0;
  double x = 0.0;
  REAL h = (REAL)(T / n_steps);
  orc_rk5_count cnt = {0, 0, 0};
  for (int s = 0; s < n_steps; s++) {
    REAL hh = h;
    /* eps huge -> every attempt accepted; h_min = h_max = h fixes the step */
    orc_rk5_step(1, &x, &y, &hh, h, h, (REAL)1e30, lin_rhs, &c, &cnt);
  }
  return (double)y;
}

double orc_rk5_linear_adaptive(double lambda, double y0, double T, double h0,
                               double h_min, double eps, long *n_acc,
                               long *n_rej) {
  lin_ctx c = {(REAL)lambda};
  REAL y = (REAL)y0;
  double x = 0.0;
  REAL h = (REAL)h0;
  orc_rk5_count cnt = {0, 0, 0};
  while (x < T - h_min) {
    REAL hmax = (REAL)(T - x);
    cnt.n_att = 0;
    orc_rk5_step(1, &x, &y, &h, (REAL)h_min, hmax, (REAL)eps, lin_rhs, &c,
                 &cnt);
  }
  *n_acc = cnt.n_acc;
  *n_rej = cnt.n_rej;
  return (double)y;
}

/* ------------------------------------------------------------------ */
/* AdEx conductance RHS, src/aeif_cond_beta_kernel.h:153-176 (alpha:     */
/* src/aeif_cond_alpha_kernel.h:149-170 with tau_rise = tau_decay).     */
/* ------------------------------------------------------------------ */
typedef struct {
  REAL V_th, Delta_T, g_L, E_L, C_m, a, b, tau_w, I_e, V_peak, V_reset,
      t_ref, den_delay;
  REAL inv_Delta_T, inv_C_m, inv_tau_w;
  REAL E_rev[MAX_PORT], inv_tau_rise[MAX_PORT], inv_tau_decay[MAX_PORT],
      g0[MAX_PORT];
  REAL refr;
  int n_port;
} aeif_par;

static void aeif_rhs(int nvar, const REAL *y, REAL *dy, const void *ctx) {
  const aeif_par *p = (const aeif_par *)ctx;
  (void)nvar;
  int refr = p->refr > (REAL)0;
  REAL V = refr ? p->V_reset : ((y[0] < p->V_peak) ? y[0] : p->V_peak);
  REAL I_syn = 0;
  for (int i = 0; i < p->n_port; i++)
    I_syn = FMA(y[2 + 2 * i], p->E_rev[i] - V, I_syn);
  REAL V_spike = p->Delta_T * EXPR((V - p->V_th) * p->inv_Delta_T);
  REAL t = (V - p->E_L) - V_spike;
  t = (-p->g_L) * t;
  t = t + I_syn;
  t = t - y[1];
  t = t + p->I_e;
  /* refractory: dV = t times a zero 1/C_m (the kernels' form) */
  dy[0] = t * (refr ? (REAL)0 : p->inv_C_m);
  REAL u = p->a * (V - p->E_L);
  u = u - y[1];
  dy[1] = u * p->inv_tau_w;
  for (int i = 0; i < p->n_port; i++) {
    REAL g = y[2 + 2 * i], g1 = y[3 + 2 * i];
    dy[3 + 2 * i] = (-g1) * p->inv_tau_rise[i];
    dy[2 + 2 * i] = FMA(-g, p->inv_tau_decay[i], g1);
  }
}

/* AdEx with current-based synapses, src/aeif_psc_alpha_kernel.h:132-155,
 * src/aeif_psc_exp_kernel.h:126-146, src/aeif_psc_delta_kernel.h:118-131;
 * izhikevich_cond_beta, src/izhikevich_cond_beta_kernel.h:145-165 (its
 * membrane equation in double: the reference's literals are double). */
typedef struct {
  int model, n_port;
  REAL V_th, Delta_T, inv_Delta_T, g_L, E_L, inv_C_m, a, b, inv_tau_w, I_e,
      V_peak, V_reset, t_ref, c, d;
  REAL inv_tau_syn[MAX_PORT], w_in[MAX_PORT];
  REAL E_rev[MAX_PORT], inv_tau_rise[MAX_PORT], inv_tau_decay[MAX_PORT],
      g0[MAX_PORT];
  REAL refr;
} rk5m_par;

static void psc_rhs(int nvar, const REAL *y, REAL *dy, const void *ctx) {
  const rk5m_par *p = (const rk5m_par *)ctx;
  (void)nvar;
  int vp = p->model == ORC_AEIF_PSC_ALPHA ? 2
           : p->model == ORC_AEIF_PSC_EXP ? 1 : 0;
  int refr = p->refr > (REAL)0;
  REAL V = refr ? p->V_reset : ((y[0] < p->V_peak) ? y[0] : p->V_peak);
  REAL I_syn = 0;
  if (vp > 0)
    for (int q = 0; q < p->n_port; q++) I_syn = I_syn + y[2 + vp * q];
  REAL V_spike = p->Delta_T == (REAL)0
                     ? (REAL)0
                     : p->Delta_T * EXPR((V - p->V_th) * p->inv_Delta_T);
  REAL t = (V - p->E_L) - V_spike;
  t = (-p->g_L) * t;
  t = t + I_syn;
  t = t - y[1];
  t = t + p->I_e;
  dy[0] = refr ? (REAL)0 : t * p->inv_C_m;
  REAL u = p->a * (V - p->E_L);
  u = u - y[1];
  dy[1] = u * p->inv_tau_w;
  for (int q = 0; q < p->n_port; q++) {
    if (vp == 2) {
      REAL I = y[2 + 2 * q], I1 = y[3 + 2 * q];
      dy[3 + 2 * q] = (-I1) * p->inv_tau_syn[q];
      dy[2 + 2 * q] = FMA(-I, p->inv_tau_syn[q], I1);
    } else if (vp == 1) {
      dy[2 + q] = (-y[2 + q]) * p->inv_tau_syn[q];
    }
  }
}

static void izh_rhs(int nvar, const REAL *y, REAL *dy, const void *ctx) {
  const rk5m_par *p = (const rk5m_par *)ctx;
  (void)nvar;
  int refr = p->refr > (REAL)0;
  REAL V = refr ? p->c : y[0];
  REAL I_syn = 0;
  for (int q = 0; q < p->n_port; q++)
    I_syn = FMA(y[2 + 2 * q], p->E_rev[q] - V, I_syn);
  double Vd = (double)V;
  double t = 0.04 * Vd;
  t = t * Vd;
  t = fma(5.0, Vd, t);
  t = t + 140.0;
  t = t - (double)y[1];
  t = t + (double)I_syn;
  t = t + (double)p->I_e;
  dy[0] = refr ? (REAL)0 : (REAL)t;
  dy[1] = p->a * FMA(p->b, V, -y[1]);
  for (int q = 0; q < p->n_port; q++) {
    REAL g = y[2 + 2 * q], g1 = y[3 + 2 * q];
    dy[3 + 2 * q] = (-g1) * p->inv_tau_rise[q];
    dy[2 + 2 * q] = FMA(-g, p->inv_tau_decay[q], g1);
  }
}

/* g0 of izhikevich_cond_beta in float, as the reference computes it
 * (src/izhikevich_cond_beta.cu:57-84) */
static float orc_izh_g0(float tau_rise, float tau_decay) {
  float denom1 = tau_decay - tau_rise;
  float denom2 = 0.0f;
  if (denom1 != 0.0f) {
    float t_p = tau_decay * tau_rise * logf(tau_decay / tau_rise) / denom1;
    denom2 = expf(-t_p / tau_decay) - expf(-t_p / tau_rise);
  }
  if (denom2 == 0.0f) return (float)(M_E / tau_decay);
  return (float)((1.0 / tau_rise - 1.0 / tau_decay) / denom2);
}

static int is_rk5m(int model) {
  return model == ORC_AEIF_PSC_ALPHA || model == ORC_AEIF_PSC_EXP ||
         model == ORC_AEIF_PSC_DELTA || model == ORC_IZH_COND_BETA;
}

/* ------------------------------------------------------------------ */
/* Network                                                              */
/* ------------------------------------------------------------------ */
typedef struct {
  int model, n, n_port, i0, nvar, nparam;
  REAL *var;   /* [nvar][n] */
  REAL *param; /* [nparam][n] */
  double *x;
  REAL *h;
  double h_min_rel, h0_rel, eps;
  REAL h_min, h0, epsr;
  /* spike generator */
  int *sg_n, *sg_next;
  float **sg_t, **sg_h;
  int **sg_idx;
  /* poisson generator */
  float *pg_rate, *pg_origin, *pg_start, *pg_stop;
  float lp_lam;      /* CreatePoissonGenerator: mean count per step */
  uint32_t *lp_tab;  /* and its CDF table */
} orc_group;

typedef struct {
  int src, tgt;
  uint8_t port, syn_group;
  float weight, delay;
  int d;      /* delay index d_int */
  long ins;   /* insertion index */
} orc_conn;

typedef struct {
  long step;
  int node, count;
  REAL height;
} orc_spike;

struct orc_net {
  double dt_d;
  float dt;
  double t0;
  uint64_t seed;
  int n_group;
  orc_group groups[128];
  int n_node;
  int *node_group;
  /* connections */
  orc_conn *conn;
  long n_conn, cap_conn;
  /* canonical CSR */
  long *src_ptr;       /* per node: first conn */
  int max_d;
  /* STDP */
  int n_syn_group;
  float syn_param[64][7];
  uint16_t *spike_time;
  long *rev_ptr;
  long *rev_idx;
  int any_plastic;
  /* direct (poisson) connections, canonical order */
  long n_dir;
  long *dir_conn;   /* index into conn */
  /* per node */
  long long *last_spike, *last_rev;
  ACC *acc;    /* [max_port][n_node] */
  int max_port;
  int *den_idx;
  int max_den;
  long *last_push; /* per node: index in spk of its spike in the current step */
  uint32_t *pg_tab; /* per poisson node: CDF table (POISSON_TABLE_MAX) */
  int *pg_ok;
  /* spikes */
  orc_spike *spk;
  long n_spk, cap_spk;
  long *step_first; /* first spike index of each global step */
  long cap_steps;
  long gs;          /* global step counter */
  /* run state */
  double neural_time;
  long long time_idx;
  /* records */
  int n_rec;
  int rec_node[256], rec_var[256], rec_is_param[256];
  double *rec;
  long n_rec_rows, cap_rec_rows;
  int calibrated, started;
  long cnt_acc, cnt_rej, cnt_events, cnt_pois;
};

orc_net *orc_net_create(double dt, double t0, uint64_t seed) {
  orc_net *net = (orc_net *)calloc(1, sizeof(orc_net));
  net->dt = (float)dt;
  net->dt_d = (double)net->dt;
  net->t0 = t0;
  net->seed = seed;
  return net;
}

static void grp_free(orc_group *g) {
  free(g->var); free(g->param); free(g->x); free(g->h);
  if (g->sg_n) {
    for (int i = 0; i < g->n; i++) {
      free(g->sg_t[i]); free(g->sg_h[i]); free(g->sg_idx[i]);
    }
    free(g->sg_n); free(g->sg_next); free(g->sg_t); free(g->sg_h);
    free(g->sg_idx);
  }
  free(g->pg_rate); free(g->pg_origin); free(g->pg_start); free(g->pg_stop);
  free(g->lp_tab);
}

void orc_net_free(orc_net *net) {
  if (!net) return;
  for (int g = 0; g < net->n_group; g++) grp_free(&net->groups[g]);
  free(net->node_group); free(net->conn); free(net->src_ptr);
  free(net->spike_time); free(net->rev_ptr); free(net->rev_idx);
  free(net->dir_conn); free(net->last_spike); free(net->last_rev);
  free(net->acc); free(net->den_idx); free(net->last_push);
  free(net->pg_tab); free(net->pg_ok); free(net->spk); free(net->step_first);
  free(net->rec);
  free(net);
}

/* variable/parameter layout per model (reference enum order) */
static void model_sizes(int model, int n_port, int *nvar, int *nparam) {
  switch (model) {
  case ORC_AEIF_COND_BETA: *nvar = 2 + 2 * n_port; *nparam = 14 + 4 * n_port; break;
  case ORC_AEIF_COND_ALPHA: *nvar = 2 + 2 * n_port; *nparam = 14 + 3 * n_port; break;
  case ORC_IAF_PSC_EXP: *nvar = 4; *nparam = 16; break;
  case ORC_POISSON_GENERATOR: *nvar = 0; *nparam = 4; break;
  case ORC_PARROT_NEURON: *nvar = 3; *nparam = 2; break;
  case ORC_AEIF_PSC_ALPHA: *nvar = 2 + 2 * n_port; *nparam = 14 + 2 * n_port; break;
  case ORC_AEIF_PSC_EXP: *nvar = 2 + n_port; *nparam = 14 + n_port; break;
  case ORC_AEIF_PSC_DELTA: *nvar = 2; *nparam = 14; break;
  case ORC_IZH_COND_BETA: *nvar = 2 + 2 * n_port; *nparam = 9 + 4 * n_port; break;
  /* group parameters are stored as per-neuron columns after the scalar
   * ones; the update reads them from neuron 0 */
  case ORC_IAF_PSC_EXP_G: *nvar = 3; *nparam = 1 + 7; break;
  case ORC_IAF_PSC_EXP_HC: *nvar = 3; *nparam = 1; break;
  case ORC_IZHIKEVICH: *nvar = 4; *nparam = 2 + 6; break;
  case ORC_IZH_PSC_EXP:
  case ORC_IZH_PSC_EXP_2S:
  case ORC_IZH_PSC_EXP_5S: *nvar = 4; *nparam = 2 + 7; break;
  /* port_input, port_value per port; den_delay, port_weight per port */
  case ORC_EXT_NEURON: *nvar = 2 * n_port; *nparam = 1 + n_port; break;
  default: *nvar = 0; *nparam = 0; break;
  }
}

int orc_net_add_group(orc_net *net, int model, int n, int n_port) {
  orc_group *g = &net->groups[net->n_group];
  memset(g, 0, sizeof(*g));
  g->model = model;
  g->n = n;
  if (model == ORC_IAF_PSC_EXP) n_port = 2;
  if (model == ORC_SPIKE_GENERATOR || model == ORC_POISSON_GENERATOR ||
      model == ORC_POISSON_LEGACY)
    n_port = 0;
  if (model == ORC_PARROT_NEURON) n_port = 2;
  if (model >= ORC_IAF_PSC_EXP_G && model <= ORC_IZH_PSC_EXP_5S) n_port = 1;
  g->n_port = n_port;
  g->i0 = net->n_node;
  model_sizes(model, n_port, &g->nvar, &g->nparam);
  g->var = (REAL *)calloc((size_t)g->nvar * n + 1, sizeof(REAL));
  g->param = (REAL *)calloc((size_t)g->nparam * n + 1, sizeof(REAL));
  g->x = (double *)calloc(n, sizeof(double));
  g->h = (REAL *)calloc(n, sizeof(REAL));
  g->h_min_rel = 1.0e-3;
  g->h0_rel = 1.0e-2;
  g->eps = 1.0e-6;
  /* defaults, src/aeif_cond_beta.cu:25-56, src/iaf_psc_exp.cu:134-158 */
  if (model == ORC_AEIF_COND_BETA || model == ORC_AEIF_COND_ALPHA) {
    REAL def[14] = {-50.4f, 2.0f, 30.0f, -70.6f, 281.0f, 4.0f, 80.5f, 144.0f,
                    0.0f, 0.0f, -60.0f, 0.0f, 0.0f, 0.0f};
    for (int p = 0; p < 14; p++)
      for (int i = 0; i < n; i++) g->param[(size_t)p * n + i] = def[p];
    for (int i = 0; i < n; i++) g->var[i] = (REAL)-70.6f;
    int pp = (model == ORC_AEIF_COND_BETA) ? 4 : 3;
    for (int q = 0; q < n_port; q++)
      for (int i = 0; i < n; i++) {
        g->param[(size_t)(14 + pp * q + 0) * n + i] = 0;
        if (model == ORC_AEIF_COND_BETA) {
          g->param[(size_t)(14 + pp * q + 1) * n + i] = 2.0f;
          g->param[(size_t)(14 + pp * q + 2) * n + i] = 20.0f;
        } else {
          g->param[(size_t)(14 + pp * q + 1) * n + i] = 2.0f;
        }
      }
  } else if (model == ORC_AEIF_PSC_ALPHA || model == ORC_AEIF_PSC_EXP ||
             model == ORC_AEIF_PSC_DELTA) {
    /* src/aeif_psc_alpha.cu:24-48 (and the exp / delta variants) */
    REAL def[14] = {-50.4f, 2.0f, 30.0f, -70.6f, 281.0f, 4.0f, 80.5f, 144.0f,
                    0.0f, 0.0f, -60.0f, 0.0f, 0.0f, 0.0f};
    for (int p = 0; p < 14; p++)
      for (int i = 0; i < n; i++) g->param[(size_t)p * n + i] = def[p];
    for (int i = 0; i < n; i++) g->var[i] = (REAL)-70.6f;
    int pp = model == ORC_AEIF_PSC_ALPHA ? 2 : (model == ORC_AEIF_PSC_EXP ? 1 : 0);
    for (int q = 0; q < n_port && pp > 0; q++)
      for (int i = 0; i < n; i++) g->param[(size_t)(14 + pp * q) * n + i] = 0.2f;
  } else if (model == ORC_IZH_COND_BETA) {
    /* src/izhikevich_cond_beta.cu:24-50 */
    REAL def[9] = {30.0f, 0.02f, 0.2f, -65.0f, 8.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    for (int p = 0; p < 9; p++)
      for (int i = 0; i < n; i++) g->param[(size_t)p * n + i] = def[p];
    for (int i = 0; i < n; i++) {
      g->var[i] = (REAL)-70.0f;
      g->var[(size_t)n + i] = (REAL)(0.2f * -70.0f);
    }
    for (int q = 0; q < n_port; q++)
      for (int i = 0; i < n; i++) {
        g->param[(size_t)(9 + 4 * q + 1) * n + i] = 2.0f;
        g->param[(size_t)(9 + 4 * q + 2) * n + i] = 20.0f;
      }
  } else if (model == ORC_IAF_PSC_EXP) {
    REAL def[10] = {10.0f, 250.0f, -70.0f, 0.0f, 15.0f, 0.0f, 2.0f, 2.0f,
                    2.0f, 0.0f};
    for (int p = 0; p < 10; p++)
      for (int i = 0; i < n; i++) g->param[(size_t)p * n + i] = def[p];
  } else if (model == ORC_EXT_NEURON) {
    /* port_weight 1 (src/ext_neuron.cu:71-77) */
    for (int q = 0; q < n_port; q++)
      for (int i = 0; i < n; i++) g->param[(size_t)(1 + q) * n + i] = (REAL)1;
  } else if (model == ORC_IAF_PSC_EXP_G) {
    /* src/iaf_psc_exp_g.cu:96-118 */
    REAL def[7] = {10.0f, 250.0f, -65.0f, 15.0f, 0.0f, 0.5f, 2.0f};
    for (int p = 0; p < 7; p++)
      for (int i = 0; i < n; i++) g->param[(size_t)(1 + p) * n + i] = def[p];
  } else if (model >= ORC_IZHIKEVICH && model <= ORC_IZH_PSC_EXP_5S) {
    /* src/izhikevich.cu:96-110, src/izhikevich_psc_exp.cu:100-116 */
    REAL def[7] = {30.0f, 0.02f, 0.2f, -65.0f, 8.0f, 2.0f, 0.0f};
    int psc = model != ORC_IZHIKEVICH;
    for (int p = 0, q = 0; p < 7; p++) {
      if (!psc && p == 5) continue;   /* tau_syn */
      for (int i = 0; i < n; i++) g->param[(size_t)(2 + q) * n + i] = def[p];
      q++;
    }
    for (int i = 0; i < n; i++) {
      g->var[(size_t)1 * n + i] = (REAL)-70.0f;
      g->var[(size_t)2 * n + i] = (REAL)(-70.0f * 0.2f);
    }
  } else if (model == ORC_SPIKE_GENERATOR) {
    g->sg_n = (int *)calloc(n, sizeof(int));
    g->sg_next = (int *)calloc(n, sizeof(int));
    g->sg_t = (float **)calloc(n, sizeof(float *));
    g->sg_h = (float **)calloc(n, sizeof(float *));
    g->sg_idx = (int **)calloc(n, sizeof(int *));
  } else if (model == ORC_POISSON_GENERATOR) {
    g->pg_rate = (float *)calloc(n, sizeof(float));
    g->pg_origin = (float *)calloc(n, sizeof(float));
    g->pg_start = (float *)calloc(n, sizeof(float));
    g->pg_stop = (float *)malloc(n * sizeof(float));
    for (int i = 0; i < n; i++) g->pg_stop[i] = 1.0e30f;
  }
  net->node_group = (int *)realloc(net->node_group,
                                   (size_t)(net->n_node + n) * sizeof(int));
  for (int i = 0; i < n; i++) net->node_group[net->n_node + i] = net->n_group;
  net->n_node += n;
  if (n_port > net->max_port) net->max_port = n_port;
  return net->groups[net->n_group++].i0;
}

int orc_net_legacy_rate(orc_net *net, int group, float rate) {
  orc_group *g = &net->groups[group];
  /* float product, then a double division stored as float
   * (src/neurongpu.cu:261) */
  float dt = (float)net->dt_d;
  g->lp_lam = (float)((double)(rate * dt) / 1000.0);
  free(g->lp_tab);
  g->lp_tab = (uint32_t *)malloc(POISSON_TABLE_MAX * sizeof(uint32_t));
  for (int k = 0; k < POISSON_TABLE_MAX; k++) g->lp_tab[k] = 0xFFFFFFFFu;
  if (!(g->lp_lam >= 0.0f)) return -1;
  return orc_poisson_table((double)g->lp_lam, g->lp_tab, POISSON_TABLE_MAX) < 0
             ? -1 : 0;
}

int orc_net_n_var(orc_net *net, int group) { return net->groups[group].nvar; }
int orc_net_n_param(orc_net *net, int group) { return net->groups[group].nparam; }

void orc_net_set_var(orc_net *net, int group, int idx, const void *vals) {
  orc_group *g = &net->groups[group];
  memcpy(g->var + (size_t)idx * g->n, vals, (size_t)g->n * sizeof(REAL));
}
void orc_net_get_var(orc_net *net, int group, int idx, void *vals) {
  orc_group *g = &net->groups[group];
  memcpy(vals, g->var + (size_t)idx * g->n, (size_t)g->n * sizeof(REAL));
}
void orc_net_set_param(orc_net *net, int group, int idx, const void *vals) {
  orc_group *g = &net->groups[group];
  memcpy(g->param + (size_t)idx * g->n, vals, (size_t)g->n * sizeof(REAL));
}
void orc_net_get_param(orc_net *net, int group, int idx, void *vals) {
  orc_group *g = &net->groups[group];
  memcpy(vals, g->param + (size_t)idx * g->n, (size_t)g->n * sizeof(REAL));
}
void orc_net_set_rk5(orc_net *net, int group, double h_min_rel, double h0_rel,
                     double eps) {
  orc_group *g = &net->groups[group];
  g->h_min_rel = h_min_rel;
  g->h0_rel = h0_rel;
  g->eps = eps;
}

void orc_net_spike_gen(orc_net *net, int node, int n, const float *times_ms,
                       const float *heights) {
  orc_group *g = &net->groups[net->node_group[node]];
  int i = node - g->i0;
  free(g->sg_t[i]); free(g->sg_h[i]); free(g->sg_idx[i]);
  g->sg_n[i] = n;
  g->sg_t[i] = (float *)malloc(n * sizeof(float));
  g->sg_h[i] = (float *)malloc(n * sizeof(float));
  g->sg_idx[i] = (int *)malloc(n * sizeof(int));
  for (int k = 0; k < n; k++) {
    g->sg_t[i][k] = times_ms[k];
    g->sg_h[i][k] = heights ? heights[k] : 1.0f;
  }
}

void orc_net_poisson(orc_net *net, int node, float rate, float origin,
                     float start, float stop) {
  orc_group *g = &net->groups[net->node_group[node]];
  int i = node - g->i0;
  g->pg_rate[i] = rate;
  g->pg_origin[i] = origin;
  g->pg_start[i] = start;
  g->pg_stop[i] = stop;
}

int orc_net_syn_group_stdp(orc_net *net, const float *param7) {
  memcpy(net->syn_param[net->n_syn_group], param7, 7 * sizeof(float));
  return ++net->n_syn_group;
}

void orc_net_connect(orc_net *net, int n, const int *src, const int *tgt,
                     const uint8_t *port, const uint8_t *syn_group,
                     const float *weight, const float *delay) {
  if (net->n_conn + n > net->cap_conn) {
    long cap = (net->cap_conn + n) * 2;
    net->conn = (orc_conn *)realloc(net->conn, cap * sizeof(orc_conn));
    net->cap_conn = cap;
  }
  for (int k = 0; k < n; k++) {
    orc_conn *c = &net->conn[net->n_conn];
    c->src = src[k];
    c->tgt = tgt[k];
    c->port = port ? port[k] : 0;
    c->syn_group = syn_group ? syn_group[k] : 0;
    c->weight = weight[k];
    c->delay = delay[k];
    /* src/connect.cu:37: d_int = round(delay/dt) - 1 (float arithmetic) */
    c->d = (int)roundf(delay[k] / net->dt) - 1;
    c->ins = net->n_conn;
    net->n_conn++;
  }
}

int orc_net_record(orc_net *net, int node, int var_idx, int is_param) {
  net->rec_node[net->n_rec] = node;
  net->rec_var[net->n_rec] = var_idx;
  net->rec_is_param[net->n_rec] = is_param;
  return net->n_rec++;
}

static int conn_cmp(const void *pa, const void *pb) {
  const orc_conn *a = (const orc_conn *)pa, *b = (const orc_conn *)pb;
  if (a->src != b->src) return a->src < b->src ? -1 : 1;
  if (a->d != b->d) return a->d < b->d ? -1 : 1;
  return a->ins < b->ins ? -1 : (a->ins > b->ins);
}

static REAL *P(orc_group *g, int idx, int i) {
  return &g->param[(size_t)idx * g->n + i];
}
static REAL *V(orc_group *g, int idx, int i) {
  return &g->var[(size_t)idx * g->n + i];
}

static void record_row(orc_net *net, double t) {
  if (net->n_rec == 0) return;
  if (net->n_rec_rows + 1 > net->cap_rec_rows) {
    long cap = net->cap_rec_rows * 2 + 1024;
    net->rec = (double *)realloc(net->rec, cap * (1 + net->n_rec) * sizeof(double));
    net->cap_rec_rows = cap;
  }
  double *row = net->rec + net->n_rec_rows * (1 + net->n_rec);
  row[0] = (double)(float)t;
  for (int r = 0; r < net->n_rec; r++) {
    orc_group *g = &net->groups[net->node_group[net->rec_node[r]]];
    int i = net->rec_node[r] - g->i0;
    REAL v = net->rec_is_param[r] ? *P(g, net->rec_var[r], i)
                                  : *V(g, net->rec_var[r], i);
    row[1 + r] = (double)v;
  }
  net->n_rec_rows++;
}

int orc_net_calibrate(orc_net *net) {
  long nc = net->n_conn;
  for (long k = 0; k < nc; k++)
    if (net->conn[k].delay < net->dt) return -1; /* src/connect.cu:29 */
  /* canonical order (source, delay index, insertion): stable counting
   * sort by source, then stable sort by delay inside each source */
  net->src_ptr = (long *)calloc((size_t)net->n_node + 1, sizeof(long));
  net->max_d = 0;
  for (long k = 0; k < nc; k++) {
    net->src_ptr[net->conn[k].src + 1]++;
    if (net->conn[k].d + 1 > net->max_d) net->max_d = net->conn[k].d + 1;
    if (net->conn[k].syn_group > 0) net->any_plastic = 1;
  }
  for (int i = 0; i < net->n_node; i++) net->src_ptr[i + 1] += net->src_ptr[i];
  {
    orc_conn *tmp = (orc_conn *)malloc((nc + 1) * sizeof(orc_conn));
    long *fill = (long *)malloc(((size_t)net->n_node + 1) * sizeof(long));
    memcpy(fill, net->src_ptr, (size_t)net->n_node * sizeof(long));
    for (long k = 0; k < nc; k++) tmp[fill[net->conn[k].src]++] = net->conn[k];
    free(fill);
    free(net->conn);
    net->conn = tmp;
    net->cap_conn = nc + 1;
    for (int s = 0; s < net->n_node; s++) {
      long b = net->src_ptr[s], e = net->src_ptr[s + 1];
      int sorted = 1;
      for (long k = b + 1; k < e; k++)
        if (net->conn[k].d < net->conn[k - 1].d) { sorted = 0; break; }
      if (!sorted) qsort(net->conn + b, e - b, sizeof(orc_conn), conn_cmp);
    }
  }
  /* reverse plastic index, src/spike_buffer.cu:423-470 */
  net->rev_ptr = (long *)calloc(net->n_node + 1, sizeof(long));
  long n_rev = 0;
  for (long k = 0; k < nc; k++)
    if (net->conn[k].syn_group > 0) { net->rev_ptr[net->conn[k].tgt + 1]++; n_rev++; }
  for (int i = 0; i < net->n_node; i++) net->rev_ptr[i + 1] += net->rev_ptr[i];
  net->rev_idx = (long *)malloc((n_rev + 1) * sizeof(long));
  long *fill = (long *)calloc((size_t)net->n_node + 1, sizeof(long));
  for (long k = 0; k < nc; k++)
    if (net->conn[k].syn_group > 0) {
      int t = net->conn[k].tgt;
      net->rev_idx[net->rev_ptr[t] + fill[t]++] = k;
    }
  free(fill);
  net->spike_time = (uint16_t *)malloc((nc + 1) * sizeof(uint16_t));
  for (long k = 0; k < nc; k++) net->spike_time[k] = 0x8000; /* rev_spike.cu:153 */
  /* direct connections of poisson generators, canonical order */
  net->dir_conn = (long *)malloc((nc + 1) * sizeof(long));
  net->n_dir = 0;
  for (long k = 0; k < nc; k++) {
    orc_group *g = &net->groups[net->node_group[net->conn[k].src]];
    if (g->model == ORC_POISSON_GENERATOR) net->dir_conn[net->n_dir++] = k;
  }
  net->last_spike = (long long *)malloc(net->n_node * sizeof(long long));
  net->last_rev = (long long *)malloc(net->n_node * sizeof(long long));
  for (int i = 0; i < net->n_node; i++) {
    net->last_spike[i] = LAST_SPIKE_TIME_GUARD;
    net->last_rev[i] = LAST_SPIKE_TIME_GUARD;
  }
  int mp = net->max_port > 0 ? net->max_port : 1;
  net->acc = (ACC *)calloc((size_t)mp * net->n_node, sizeof(ACC));
  net->den_idx = (int *)calloc(net->n_node, sizeof(int));
  net->last_push = (long *)malloc(net->n_node * sizeof(long));
  for (int i = 0; i < net->n_node; i++) net->last_push[i] = -1;
  net->pg_tab = (uint32_t *)calloc((size_t)net->n_node * POISSON_TABLE_MAX + 1,
                                   sizeof(uint32_t));
  net->pg_ok = (int *)calloc(net->n_node + 1, sizeof(int));
  /* group calibration */
  for (int gi = 0; gi < net->n_group; gi++) {
    orc_group *g = &net->groups[gi];
    /* h_min = h_min_rel * dt in float (src/aeif_cond_beta.cu:153-160) */
    g->h_min = (REAL)((float)g->h_min_rel * net->dt);
    g->h0 = (REAL)((float)g->h0_rel * net->dt);
    g->epsr = (REAL)g->eps;
    if (g->model == ORC_AEIF_COND_BETA || g->model == ORC_AEIF_COND_ALPHA) {
      for (int i = 0; i < g->n; i++) {
        g->x[i] = net->t0;
        g->h[i] = g->h0;
        *P(g, 12, i) = 0; /* refractory_step */
        for (int q = 0; q < g->n_port; q++) {
          if (g->model == ORC_AEIF_COND_BETA) {
            double tr = *P(g, 14 + 4 * q + 1, i), td = *P(g, 14 + 4 * q + 2, i);
            *P(g, 14 + 4 * q + 3, i) = (REAL)orc_beta_g0(tr, td);
          } else {
            double ts = *P(g, 14 + 3 * q + 1, i);
            *P(g, 14 + 3 * q + 2, i) = (REAL)(M_E / ts);
          }
        }
      }
    } else if (is_rk5m(g->model)) {
      int izh = g->model == ORC_IZH_COND_BETA;
      for (int i = 0; i < g->n; i++) {
        g->x[i] = net->t0;
        g->h[i] = g->h0;
        *P(g, izh ? 7 : 12, i) = 0; /* refractory_step */
        if (!izh && *P(g, 1, i) <= (REAL)0) *P(g, 9, i) = *P(g, 0, i);
        for (int q = 0; q < g->n_port; q++) {
          if (g->model == ORC_AEIF_PSC_ALPHA) {
            double ts = *P(g, 14 + 2 * q, i);
            *P(g, 14 + 2 * q + 1, i) = (REAL)(M_E / ts);
          } else if (izh) {
            *P(g, 9 + 4 * q + 3, i) = (REAL)orc_izh_g0(
                (float)*P(g, 9 + 4 * q + 1, i), (float)*P(g, 9 + 4 * q + 2, i));
          }
        }
      }
    } else if (g->model == ORC_IAF_PSC_EXP) {
      for (int i = 0; i < g->n; i++) {
        float pr[6];
        orc_iaf_propagators(net->dt_d, *P(g, 0, i), *P(g, 1, i), *P(g, 6, i),
                            *P(g, 7, i), pr);
        for (int k = 0; k < 6; k++) *P(g, 10 + k, i) = (REAL)pr[k];
      }
    } else if (g->model == ORC_SPIKE_GENERATOR) {
      for (int i = 0; i < g->n; i++) {
        for (int k = 0; k < g->sg_n[i]; k++)
          g->sg_idx[i][k] =
              (int)roundf((g->sg_t[i][k] - (float)net->t0) / net->dt);
        g->sg_next[i] = 0;
      }
    }
    /* dendritic delay index, src/spike_buffer.cu:150-160 */
    int den_p = -1;
    if (g->model == ORC_AEIF_COND_BETA || g->model == ORC_AEIF_COND_ALPHA) den_p = 13;
    if (g->model == ORC_IAF_PSC_EXP) den_p = 9;
    if (g->model == ORC_PARROT_NEURON) den_p = 1;
    if (g->model == ORC_AEIF_PSC_ALPHA || g->model == ORC_AEIF_PSC_EXP ||
        g->model == ORC_AEIF_PSC_DELTA) den_p = 13;
    if (g->model == ORC_IZH_COND_BETA) den_p = 8;
    if (den_p >= 0)
      for (int i = 0; i < g->n; i++)
        net->den_idx[g->i0 + i] = (int)ROUNDR(*P(g, den_p, i) / (REAL)net->dt);
  }
  net->max_den = 0;
  for (int i = 0; i < net->n_node; i++)
    if (net->den_idx[i] > net->max_den) net->max_den = net->den_idx[i];
  net->neural_time = net->t0;
  net->gs = 0;
  net->cap_steps = 1024;
  net->step_first = (long *)calloc(net->cap_steps + 1, sizeof(long));
  net->calibrated = 1;
  return 0;
}

static void push_spike(orc_net *net, int node, REAL height, long long T,
                       double t) {
  (void)t;
  net->last_spike[node] = T;
  if (net->den_idx[node] == 0) net->last_rev[node] = T;
  /* merge with an earlier spike of the same node in this step */
  long lp = net->last_push[node];
  if (lp >= net->step_first[net->gs] && lp < net->n_spk &&
      net->spk[lp].node == node && net->spk[lp].step == net->gs) {
    net->spk[lp].height += height;
    net->spk[lp].count += 1;
    return;
  }
  net->last_push[node] = net->n_spk;
  if (net->n_spk + 1 > net->cap_spk) {
    net->cap_spk = net->cap_spk * 2 + 4096;
    net->spk = (orc_spike *)realloc(net->spk, net->cap_spk * sizeof(orc_spike));
  }
  orc_spike *s = &net->spk[net->n_spk++];
  s->step = net->gs;
  s->node = node;
  s->count = 1;
  s->height = height;
}

typedef struct { int n_spk; REAL h; } nspk_t;

/* RK5Update, src/rk5.h:149-162, with ExternalUpdate
 * src/aeif_cond_beta_kernel.h:178-214 */
static int aeif_update_one(orc_net *net, orc_group *g, int i, double x1,
                           orc_rk5_count *cnt) {
  aeif_par p;
  int beta = g->model == ORC_AEIF_COND_BETA;
  int pp = beta ? 4 : 3;
  p.n_port = g->n_port;
  p.V_th = *P(g, 0, i); p.Delta_T = *P(g, 1, i); p.g_L = *P(g, 2, i);
  p.E_L = *P(g, 3, i); p.C_m = *P(g, 4, i); p.a = *P(g, 5, i);
  p.b = *P(g, 6, i); p.tau_w = *P(g, 7, i); p.I_e = *P(g, 8, i);
  p.V_peak = *P(g, 9, i); p.V_reset = *P(g, 10, i); p.t_ref = *P(g, 11, i);
  p.refr = *P(g, 12, i);
  p.inv_Delta_T = (REAL)1 / p.Delta_T;
  p.inv_C_m = (REAL)1 / p.C_m;
  p.inv_tau_w = (REAL)1 / p.tau_w;
  for (int q = 0; q < g->n_port; q++) {
    p.E_rev[q] = *P(g, 14 + pp * q, i);
    if (beta) {
      p.inv_tau_rise[q] = (REAL)1 / *P(g, 14 + pp * q + 1, i);
      p.inv_tau_decay[q] = (REAL)1 / *P(g, 14 + pp * q + 2, i);
    } else {
      REAL its = (REAL)1 / *P(g, 14 + pp * q + 1, i);
      p.inv_tau_rise[q] = its;
      p.inv_tau_decay[q] = its;
    }
  }
  int nvar = g->nvar;
  REAL y[64];
  for (int k = 0; k < nvar; k++) y[k] = *V(g, k, i);
  double x = g->x[i];
  REAL h = g->h[i];
  int nsp = 0;
  int end = 0;
  cnt->n_att = 0;
  while (!end) {
    REAL hmax = (REAL)(x1 - x);
    if (orc_rk5_step(nvar, &x, y, &h, g->h_min, hmax, g->epsr, aeif_rhs, &p,
                     cnt)) {
      x = x1; /* attempt cap reached: give up on this step */
    }
    end = (x >= x1 - (double)g->h_min);
    /* ExternalUpdate */
    if (y[0] < (REAL)-1.0e3) { y[0] = p.V_reset; y[1] = 0; continue; }
    if (y[1] < (REAL)-1.0e6 || y[1] > (REAL)1.0e6) {
      y[0] = p.V_reset; y[1] = 0; continue;
    }
    if (p.refr > (REAL)0) {
      y[0] = p.V_reset;
      if (end) p.refr -= (REAL)1;
    } else if (y[0] >= p.V_peak) {
      nsp++;
      y[0] = p.V_reset;
      y[1] = y[1] + p.b;
      REAL r = ROUNDR(p.t_ref / (REAL)net->dt);
      int ri = (int)r;
      p.refr = (REAL)(ri < 0 ? 0 : ri);
    }
  }
  for (int k = 0; k < nvar; k++) *V(g, k, i) = y[k];
  *P(g, 12, i) = p.refr;
  g->x[i] = x;
  g->h[i] = h;
  return nsp;
}

/* RK5Update (src/rk5.h:149-162) of the current-based AdEx models and
 * izhikevich_cond_beta with their ExternalUpdate
 * (src/aeif_psc_alpha_kernel.h:157-201, src/izhikevich_cond_beta_kernel.h:
 * 167-206) */
static int rk5m_update_one(orc_net *net, orc_group *g, int i, double x1,
                           orc_rk5_count *cnt) {
  rk5m_par p;
  memset(&p, 0, sizeof(p));
  p.model = g->model;
  p.n_port = g->n_port;
  int izh = g->model == ORC_IZH_COND_BETA;
  REAL V_reset, thresh, jump, t_ref;
  int refr_idx;
  if (izh) {
    p.V_th = *P(g, 0, i); p.a = *P(g, 1, i); p.b = *P(g, 2, i);
    p.c = *P(g, 3, i); p.d = *P(g, 4, i); p.I_e = *P(g, 5, i);
    p.t_ref = *P(g, 6, i);
    refr_idx = 7;
    for (int q = 0; q < g->n_port; q++) {
      p.E_rev[q] = *P(g, 9 + 4 * q, i);
      p.inv_tau_rise[q] = (REAL)1 / *P(g, 9 + 4 * q + 1, i);
      p.inv_tau_decay[q] = (REAL)1 / *P(g, 9 + 4 * q + 2, i);
    }
    V_reset = p.c; thresh = p.V_th; jump = p.d; t_ref = p.t_ref;
  } else {
    p.V_th = *P(g, 0, i); p.Delta_T = *P(g, 1, i); p.g_L = *P(g, 2, i);
    p.E_L = *P(g, 3, i);
    p.inv_Delta_T = (REAL)1 / p.Delta_T;
    p.inv_C_m = (REAL)1 / *P(g, 4, i);
    p.a = *P(g, 5, i); p.b = *P(g, 6, i);
    p.inv_tau_w = (REAL)1 / *P(g, 7, i);
    p.I_e = *P(g, 8, i); p.V_peak = *P(g, 9, i); p.V_reset = *P(g, 10, i);
    p.t_ref = *P(g, 11, i);
    refr_idx = 12;
    for (int q = 0; q < g->n_port; q++) {
      if (g->model == ORC_AEIF_PSC_ALPHA)
        p.inv_tau_syn[q] = (REAL)1 / *P(g, 14 + 2 * q, i);
      else if (g->model == ORC_AEIF_PSC_EXP)
        p.inv_tau_syn[q] = (REAL)1 / *P(g, 14 + q, i);
    }
    V_reset = p.V_reset; thresh = p.V_peak; jump = p.b; t_ref = p.t_ref;
  }
  p.refr = *P(g, refr_idx, i);
  int nvar = g->nvar;
  REAL y[64];
  for (int k = 0; k < nvar; k++) y[k] = *V(g, k, i);
  double x = g->x[i];
  REAL h = g->h[i];
  int nsp = 0;
  int end = 0;
  cnt->n_att = 0;
  while (!end) {
    REAL hmax = (REAL)(x1 - x);
    if (orc_rk5_step(nvar, &x, y, &h, g->h_min, hmax, g->epsr,
                     izh ? izh_rhs : psc_rhs, &p, cnt)) {
      x = x1;
    }
    end = (x >= x1 - (double)g->h_min);
    if (y[0] < (REAL)-1.0e3) { y[0] = V_reset; y[1] = 0; continue; }
    if (y[1] < (REAL)-1.0e6 || y[1] > (REAL)1.0e6) {
      y[0] = V_reset; y[1] = 0; continue;
    }
    if (p.refr > (REAL)0) {
      y[0] = V_reset;
      if (end) p.refr -= (REAL)1;
    } else if (y[0] >= thresh) {
      nsp++;
      y[0] = V_reset;
      y[1] = y[1] + jump;
      REAL r = ROUNDR(t_ref / (REAL)net->dt);
      int ri = (int)r;
      p.refr = (REAL)(ri < 0 ? 0 : ri);
    }
  }
  for (int k = 0; k < nvar; k++) *V(g, k, i) = y[k];
  *P(g, refr_idx, i) = p.refr;
  g->x[i] = x;
  g->h[i] = h;
  return nsp;
}

/* src/iaf_psc_exp.cu:94-119 */
static int iaf_update_one(orc_net *net, orc_group *g, int i) {
  REAL *I_ex = V(g, 0, i), *I_in = V(g, 1, i), *Vm = V(g, 2, i),
       *refr = V(g, 3, i);
  REAL P20 = *P(g, 10, i), P11ex = *P(g, 11, i), P11in = *P(g, 12, i);
  REAL P21ex = *P(g, 13, i), P21in = *P(g, 14, i), P22 = *P(g, 15, i);
  if (*refr > (REAL)0) {
    *refr -= (REAL)1;
  } else {
    REAL v = *Vm * P22;
    v = FMA(*I_ex, P21ex, v);
    v = FMA(*I_in, P21in, v);
    v = FMA(*P(g, 3, i), P20, v);
    *Vm = v;
  }
  *I_ex = *I_ex * P11ex;
  *I_in = *I_in * P11in;
  if (*Vm >= *P(g, 4, i)) {
    *Vm = *P(g, 5, i);
    *refr = (REAL)(int)ROUNDR(*P(g, 8, i) / (REAL)net->dt);
    return 1;
  }
  return 0;
}

/* Constants of one step of a fixed-step point model, as the reference's
 * Update() forms them on the host in float (src/iaf_psc_exp_g.cu:136-150,
 * src/izhikevich.cu:135-143, src/izhikevich_psc_exp_2s.cu:143-152). */
typedef struct {
  REAL P11, P22, P21, P20, Theta, V_reset;
  REAL V_th, a, b, c, d, h, C_syn;
  int n_refr, steps;
} point_k;

static void point_consts(orc_net *net, orc_group *g, point_k *k) {
  memset(k, 0, sizeof(*k));
  REAL dt = (REAL)net->dt;
  if (g->model == ORC_IAF_PSC_EXP_G) {
    REAL tau_m = *P(g, 1, 0), C_m = *P(g, 2, 0), tau_syn = *P(g, 6, 0);
#if ORC_REAL_DOUBLE
    k->P11 = exp(-dt / tau_syn);
    k->P22 = exp(-dt / tau_m);
#else
    k->P11 = expf(-dt / tau_syn);
    k->P22 = expf(-dt / tau_m);
#endif
    k->P21 = (REAL)orc_p32(tau_syn, tau_m, C_m, dt);
    k->P20 = (REAL)((double)(tau_m / C_m) * (1.0 - (double)k->P22));
    k->Theta = *P(g, 4, 0);
    k->V_reset = *P(g, 5, 0);
    k->n_refr = (int)ROUNDR(*P(g, 7, 0) / dt);
  } else if (g->model >= ORC_IZHIKEVICH) {
    k->V_th = *P(g, 2, 0);
    k->a = *P(g, 3, 0);
    k->b = *P(g, 4, 0);
    k->c = *P(g, 5, 0);
    k->d = *P(g, 6, 0);
    k->steps = g->model == ORC_IZH_PSC_EXP_2S ? 2
             : g->model == ORC_IZH_PSC_EXP_5S ? 5 : 1;
    k->h = dt / (REAL)k->steps;
    if (g->model == ORC_IZHIKEVICH) {
      k->n_refr = (int)ROUNDR(*P(g, 7, 0) / k->h);
    } else {
#if ORC_REAL_DOUBLE
      k->C_syn = exp(-k->h / *P(g, 7, 0));
#else
      k->C_syn = expf(-k->h / *P(g, 7, 0));
#endif
      k->n_refr = (int)ROUNDR(*P(g, 8, 0) / k->h);
    }
  }
}

/* 0.04 v v + 5 v + 140 - u in double, in the reference's order */
static double izh_drive(REAL v, REAL u) {
  double vd = (double)v;
  double s = 0.04 * vd;
  s = s * vd;
  s = s + 5.0 * vd;
  s = s + 140.0;
  s = s - (double)u;
  return s;
}

/* One step of a point model; returns 1 on a spike. */
static int point_update_one(orc_group *g, const point_k *k, int i) {
  REAL *I = V(g, 0, i), *Vm = V(g, 1, i);
  REAL I_e = *P(g, 0, i);
  if (g->model == ORC_IAF_PSC_EXP_G || g->model == ORC_IAF_PSC_EXP_HC) {
    /* src/iaf_psc_exp_g.cu:34-61, src/iaf_psc_exp_hc.cu:29-53 */
    REAL *refr = V(g, 2, i);
    int hc = g->model == ORC_IAF_PSC_EXP_HC;
    if (*refr > (REAL)0) {
      *refr -= (REAL)1;
    } else if (hc) {
      /* src/iaf_psc_exp_hc_params.h: double constants */
      double v = (double)*Vm * 9.9004983E-01;
      v = v + (double)*I * 3.6067175E-04;
      v = v + (double)I_e * 3.9800665E-04;
      *Vm = (REAL)v;
    } else {
      REAL v = *Vm * k->P22;
      v = FMA(*I, k->P21, v);
      v = FMA(I_e, k->P20, v);
      *Vm = v;
    }
    *I = hc ? (REAL)((double)*I * 8.1873075E-01) : *I * k->P11;
    int fire = hc ? (double)*Vm >= 1.5E+01 : *Vm >= k->Theta;
    if (fire) {
      *Vm = hc ? (REAL)0 : k->V_reset;
      *refr = (REAL)(hc ? 20 : k->n_refr);
      return 1;
    }
    return 0;
  }
  /* izhikevich models: src/izhikevich.cu:35-72,
   * src/izhikevich_psc_exp_2s.cu:41-83 */
  REAL *u = V(g, 2, i), *refr = V(g, 3, i);
  REAL ha = k->h * k->a;
  if (g->model == ORC_IZHIKEVICH) {
    if (*refr > (REAL)0) {
      *refr -= (REAL)1;
    } else {
      REAL v0 = *Vm, u0 = *u;
      double t = (double)k->h * (izh_drive(v0, u0) + (double)I_e);
      t = t + (double)*I;
      *Vm = (REAL)((double)*Vm + t);
      *u = *u + ha * (k->b * v0 - u0);
    }
    *I = (REAL)0;
  } else if (*refr > (REAL)0) {
    *refr -= (REAL)1;
    for (int s = 0; s < k->steps; s++) *I = *I * k->C_syn;
  } else {
    for (int s = 0; s < k->steps; s++) {
      REAL v0 = *Vm, u0 = *u;
      double r = izh_drive(v0, u0) + (double)*I;
      r = r + (double)I_e;
      *Vm = (REAL)((double)*Vm + (double)k->h * r);
      *u = *u + ha * (k->b * v0 - u0);
      *I = *I * k->C_syn;
    }
  }
  if (*Vm >= k->V_th) {
    *Vm = k->c;
    *u = *u + k->d;
    *refr = (REAL)(k->n_refr < 0 ? 0 : k->n_refr);
    return 1;
  }
  return 0;
}

static void apply_stdp(orc_net *net, long c, float Dt) {
  int sg = net->conn[c].syn_group;
  net->conn[c].weight = orc_stdp_update(net->conn[c].weight, Dt,
                                        net->syn_param[sg - 1]);
}

static void orc_step(orc_net *net, long it, double t0_run, long long T_base,
                     int n_threads) {
  long long T = T_base + it + 1;
  double t = t0_run + net->dt_d * (double)(it + 1);
  int nn = net->n_node;
  if (net->gs + 2 > net->cap_steps) {
    long cap = net->cap_steps * 2;
    net->step_first = (long *)realloc(net->step_first, (cap + 1) * sizeof(long));
    net->cap_steps = cap;
  }
  net->step_first[net->gs] = net->n_spk;
  /* phase 1: back-propagated spikes with dendritic delay > 0,
   * src/spike_buffer.cu:276-278,299-301 */
  for (int den = 1; den <= net->max_den; den++) {
    long s = net->gs - den;
    if (s < 0) break;
    for (long k = net->step_first[s]; k < net->step_first[s + 1]; k++)
      if (net->den_idx[net->spk[k].node] == den)
        net->last_rev[net->spk[k].node] = T;
  }
  net->time_idx = T;
  net->neural_time = t;
  /* 16-bit spike-time wrap, src/neurongpu.cu:472-479 */
  if (net->any_plastic) {
    if ((T & 0xffff) == 0x8000) {
      for (long c = 0; c < net->n_conn; c++)
        if (net->spike_time[c] >= 0x8000) net->spike_time[c] = 0;
    } else if ((T & 0xffff) == 0) {
      for (long c = 0; c < net->n_conn; c++)
        if (net->spike_time[c] < 0x8000) net->spike_time[c] = 0x8000;
    }
  }
  /* phase 3: node updates */
  for (int gi = 0; gi < net->n_group; gi++) {
    orc_group *g = &net->groups[gi];
    if (g->model == ORC_AEIF_COND_BETA || g->model == ORC_AEIF_COND_ALPHA) {
      int *ns = (int *)calloc(g->n, sizeof(int));
      long tot_acc = 0, tot_rej = 0;
#ifdef _OPENMP
#pragma omp parallel for num_threads(n_threads) schedule(static) \
    reduction(+ : tot_acc, tot_rej)
#endif
      for (int i = 0; i < g->n; i++) {
        orc_rk5_count cnt = {0, 0, 0};
        ns[i] = aeif_update_one(net, g, i, t, &cnt);
        tot_acc += cnt.n_acc;
        tot_rej += cnt.n_rej;
      }
      net->cnt_acc += tot_acc;
      net->cnt_rej += tot_rej;
      for (int i = 0; i < g->n; i++)
        for (int s = 0; s < ns[i]; s++)
          push_spike(net, g->i0 + i, (REAL)1, T, t);
      free(ns);
    } else if (is_rk5m(g->model)) {
      int *ns = (int *)calloc(g->n, sizeof(int));
      long tot_acc = 0, tot_rej = 0;
#ifdef _OPENMP
#pragma omp parallel for num_threads(n_threads) schedule(static) \
    reduction(+ : tot_acc, tot_rej)
#endif
      for (int i = 0; i < g->n; i++) {
        orc_rk5_count cnt = {0, 0, 0};
        ns[i] = rk5m_update_one(net, g, i, t, &cnt);
        tot_acc += cnt.n_acc;
        tot_rej += cnt.n_rej;
      }
      net->cnt_acc += tot_acc;
      net->cnt_rej += tot_rej;
      for (int i = 0; i < g->n; i++)
        for (int s = 0; s < ns[i]; s++)
          push_spike(net, g->i0 + i, (REAL)1, T, t);
      free(ns);
    } else if (g->model == ORC_IAF_PSC_EXP) {
      for (int i = 0; i < g->n; i++)
        if (iaf_update_one(net, g, i)) push_spike(net, g->i0 + i, (REAL)1, T, t);
    } else if (g->model == ORC_EXT_NEURON) {
      /* src/ext_neuron.cu:23-39: port_value = port_input; port_input = 0 */
      for (int q = 0; q < g->n_port; q++)
        for (int i = 0; i < g->n; i++) {
          *V(g, 2 * q + 1, i) = *V(g, 2 * q, i);
          *V(g, 2 * q, i) = (REAL)0;
        }
    } else if (g->model >= ORC_IAF_PSC_EXP_G && g->model <= ORC_IZH_PSC_EXP_5S) {
      point_k k;
      point_consts(net, g, &k);
      for (int i = 0; i < g->n; i++)
        if (point_update_one(g, &k, i)) push_spike(net, g->i0 + i, (REAL)1, T, t);
    } else if (g->model == ORC_SPIKE_GENERATOR) {
      for (int i = 0; i < g->n; i++) {
        int k = g->sg_next[i];
        if (k < g->sg_n[i] && g->sg_idx[i][k] == it) {
          push_spike(net, g->i0 + i, (REAL)g->sg_h[i][k], T, t);
          g->sg_next[i]++;
        }
      }
    } else if (g->model == ORC_POISSON_LEGACY) {
      /* src/poisson.cu:27-54: one count per node and step, dropped when
       * above 5 lambda, pushed as the spike height.  The push precedes the
       * time update of the step (src/neurongpu.cu:461-470), so the spike is
       * stamped with the previous time index. */
      for (int i = 0; i < g->n; i++) {
        int node = g->i0 + i;
        uint32_t u = orc_pois_word((uint64_t)net->seed, (uint32_t)node, T,
                                   0x4C504F49u);
        int k = orc_poisson_sample(g->lp_tab, u);
        if ((float)k > g->lp_lam * 5.0f) k = 0;
        if (k > 0) push_spike(net, node, (REAL)k, T - 1, t);
      }
    } else if (g->model == ORC_PARROT_NEURON) {
      /* src/parrot_neuron.cu:48-70 */
      for (int i = 0; i < g->n; i++) {
        REAL *in = V(g, 0, i);
        REAL sh = *in;
        *V(g, 2, i) = sh;
        if (sh != (REAL)0) {
          if (*P(g, 0, i) == (REAL)0) sh = (REAL)1;
          *in = (REAL)0;
          push_spike(net, g->i0 + i, sh, T, t);
        }
      }
    }
  }
  record_row(net, t);
  /* phase 4: delivery, src/get_spike.cu:48-85 */
  ACC *acc = net->acc;
  for (int d = 0; d < net->max_d; d++) {
    long s = net->gs - 1 - d;
    if (s < 0) break;
    long k0 = net->step_first[s], k1 = net->step_first[s + 1];
    for (long k = k0; k < k1; k++) {
      orc_spike *sp = &net->spk[k];
      long lo = net->src_ptr[sp->node], hi = net->src_ptr[sp->node + 1];
      while (lo < hi) { /* first conn with delay index >= d */
        long mid = (lo + hi) / 2;
        if (net->conn[mid].d < d) lo = mid + 1; else hi = mid;
      }
      for (long c = lo; c < net->src_ptr[sp->node + 1]; c++) {
        orc_conn *cn = &net->conn[c];
        if (cn->d != d) break;
        net->cnt_events++;
        REAL w = (REAL)cn->weight;
        acc[(size_t)cn->port * nn + cn->tgt] += (ACC)(sp->height * w);
        if (cn->syn_group > 0) {
          for (int rep = 0; rep < sp->count; rep++) {
            net->spike_time[c] = (uint16_t)(T & 0xffff);
            long long Dt_int = T - net->last_rev[cn->tgt];
            if (Dt_int > 0 && Dt_int < MAX_SYN_DT)
              apply_stdp(net, c, -net->dt * (float)Dt_int);
          }
        }
      }
    }
  }
  /* phase 5: Poisson direct connections, src/poiss_gen.cu:43-74 */
  if (net->n_dir > 0) {
    float time_step = (float)(net->dt_d / 1000.0);
    for (long j = 0; j < net->n_dir; j++) {
      orc_conn *cn = &net->conn[net->dir_conn[j]];
      orc_group *g = &net->groups[net->node_group[cn->src]];
      int i = cn->src - g->i0;
      float ddelay = net->dt * (float)(cn->d + 1);
      double t_rel = t - (double)g->pg_origin[i] - (double)ddelay;
      if (t_rel >= (double)g->pg_start[i] && t_rel <= (double)g->pg_stop[i]) {
        (void)time_step;
        if (!net->pg_ok[cn->src]) continue;
        const uint32_t *tab = net->pg_tab + (size_t)cn->src * POISSON_TABLE_MAX;
        uint32_t u = orc_pois_word((uint64_t)net->seed, (uint32_t)j, T,
                                   0x504F4953u);
        int n = orc_poisson_sample(tab, u);
        net->cnt_pois++;
        if (n > 0)
          acc[(size_t)cn->port * nn + cn->tgt] += (ACC)(cn->weight * (float)n);
      }
    }
  }
  /* phase 6: fold into port inputs, src/get_spike.cu:116-140 */
  for (int gi = 0; gi < net->n_group; gi++) {
    orc_group *g = &net->groups[gi];
    if (g->model == ORC_AEIF_COND_BETA || g->model == ORC_AEIF_COND_ALPHA) {
      int pp = g->model == ORC_AEIF_COND_BETA ? 4 : 3;
      int g0i = g->model == ORC_AEIF_COND_BETA ? 3 : 2;
      for (int q = 0; q < g->n_port; q++)
        for (int i = 0; i < g->n; i++) {
          REAL a = (REAL)acc[(size_t)q * nn + g->i0 + i];
          REAL *g1 = V(g, 3 + 2 * q, i);
          *g1 = FMA(a, *P(g, 14 + pp * q + g0i, i), *g1);
        }
    } else if (g->model == ORC_AEIF_PSC_ALPHA) {
      /* I1_syn += input * I0 */
      for (int q = 0; q < g->n_port; q++)
        for (int i = 0; i < g->n; i++) {
          REAL a = (REAL)acc[(size_t)q * nn + g->i0 + i];
          REAL *I1 = V(g, 3 + 2 * q, i);
          *I1 = FMA(a, *P(g, 14 + 2 * q + 1, i), *I1);
        }
    } else if (g->model == ORC_AEIF_PSC_EXP) {
      /* I_syn += input (weight 1) */
      for (int q = 0; q < g->n_port; q++)
        for (int i = 0; i < g->n; i++) {
          REAL *I = V(g, 2 + q, i);
          *I = *I + (REAL)acc[(size_t)q * nn + g->i0 + i];
        }
    } else if (g->model == ORC_AEIF_PSC_DELTA) {
      /* every port into V_m (weight 1, src/aeif_psc_delta.cu:112-120) */
      for (int q = 0; q < g->n_port; q++)
        for (int i = 0; i < g->n; i++) {
          REAL *Vm = V(g, 0, i);
          *Vm = *Vm + (REAL)acc[(size_t)q * nn + g->i0 + i];
        }
    } else if (g->model == ORC_IZH_COND_BETA) {
      for (int q = 0; q < g->n_port; q++)
        for (int i = 0; i < g->n; i++) {
          REAL a = (REAL)acc[(size_t)q * nn + g->i0 + i];
          REAL *g1 = V(g, 3 + 2 * q, i);
          *g1 = FMA(a, *P(g, 9 + 4 * q + 3, i), *g1);
        }
    } else if (g->model == ORC_IAF_PSC_EXP) {
      for (int q = 0; q < 2; q++)
        for (int i = 0; i < g->n; i++) {
          REAL *I = V(g, q, i);
          *I = *I + (REAL)acc[(size_t)q * nn + g->i0 + i];
        }
    } else if (g->model == ORC_EXT_NEURON) {
      /* port_input += input * port_weight */
      for (int q = 0; q < g->n_port; q++)
        for (int i = 0; i < g->n; i++) {
          REAL *in = V(g, 2 * q, i);
          *in = FMA((REAL)acc[(size_t)q * nn + g->i0 + i], *P(g, 1 + q, i), *in);
        }
    } else if (g->model >= ORC_IAF_PSC_EXP_G && g->model <= ORC_IZH_PSC_EXP_5S) {
      /* one port, weight 1, into I_syn */
      for (int i = 0; i < g->n; i++) {
        REAL *I = V(g, 0, i);
        *I = *I + (REAL)acc[g->i0 + i];
      }
    } else if (g->model == ORC_PARROT_NEURON) {
      /* port 0 weight 1 into input_spike_height; port 1 weight 0 */
      for (int i = 0; i < g->n; i++) {
        REAL *I = V(g, 0, i);
        *I = *I + (REAL)acc[g->i0 + i];
      }
    }
  }
  memset(acc, 0, (size_t)(net->max_port > 0 ? net->max_port : 1) * nn * sizeof(ACC));
  /* phase 7: post-synaptic STDP, src/rev_spike.cu:41-75 */
  if (net->any_plastic) {
    for (int node = 0; node < nn; node++) {
      if (net->last_rev[node] != T) continue;
      for (long r = net->rev_ptr[node]; r < net->rev_ptr[node + 1]; r++) {
        long c = net->rev_idx[r];
        uint16_t st = net->spike_time[c];
        uint16_t ti = (uint16_t)(T & 0xffff);
        uint16_t Dt_int = (uint16_t)(ti - st);
        if (Dt_int < MAX_SYN_DT) apply_stdp(net, c, net->dt * (float)Dt_int);
      }
    }
  }
  net->gs++;
  net->step_first[net->gs] = net->n_spk;
}

int orc_net_run(orc_net *net, long n_steps, int n_threads) {
  if (!net->calibrated) {
    int r = orc_net_calibrate(net);
    if (r) return r;
  }
  if (!net->started) {
    record_row(net, net->neural_time); /* src/neurongpu.cu:361-364 */
    net->started = 1;
  }
  if (n_threads < 1) n_threads = 1;
  double t0_run = net->neural_time;
  /* Poisson tables for this run (rates may change between runs) */
  {
    float time_step = (float)(net->dt_d / 1000.0);
    for (int gi = 0; gi < net->n_group; gi++) {
      orc_group *g = &net->groups[gi];
      if (g->model != ORC_POISSON_GENERATOR) continue;
      for (int i = 0; i < g->n; i++) {
        float lam = time_step * g->pg_rate[i];
        int len = orc_poisson_table((double)lam,
                                    net->pg_tab + (size_t)(g->i0 + i) * POISSON_TABLE_MAX,
                                    POISSON_TABLE_MAX);
        net->pg_ok[g->i0 + i] = len > 0;
      }
    }
  }
  long long T_base = (long long)llround(t0_run / net->dt_d);
  for (long it = 0; it < n_steps; it++) orc_step(net, it, t0_run, T_base, n_threads);
  return 0;
}

long orc_net_n_spikes(orc_net *net) { return net->n_spk; }
void orc_net_get_spikes(orc_net *net, long *step, int *node, int *count) {
  for (long k = 0; k < net->n_spk; k++) {
    step[k] = net->spk[k].step;
    node[k] = net->spk[k].node;
    count[k] = net->spk[k].count;
  }
}
void orc_net_get_spike_heights(orc_net *net, double *height) {
  for (long k = 0; k < net->n_spk; k++) height[k] = (double)net->spk[k].height;
}
int orc_net_n_records(orc_net *net) { return net->n_rec; }
long orc_net_n_record_rows(orc_net *net) { return net->n_rec_rows; }
void orc_net_get_records(orc_net *net, double *out) {
  memcpy(out, net->rec, net->n_rec_rows * (1 + net->n_rec) * sizeof(double));
}
long orc_net_n_conn(orc_net *net) { return net->n_conn; }
void orc_net_get_conn(orc_net *net, int *src, int *tgt, int *port, int *delay,
                      float *weight) {
  for (long k = 0; k < net->n_conn; k++) {
    src[k] = net->conn[k].src;
    tgt[k] = net->conn[k].tgt;
    port[k] = net->conn[k].port;
    delay[k] = net->conn[k].d;
    weight[k] = net->conn[k].weight;
  }
}
void orc_net_counters(orc_net *net, long *out4) {
  out4[0] = net->cnt_acc;
  out4[1] = net->cnt_rej;
  out4[2] = net->cnt_events;
  out4[3] = net->cnt_pois;
}
