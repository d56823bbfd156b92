/*
 * oracle.c -- CPU restatement of one VIRAL viscous hydro step (Milne KT and
 * SHASTA paths of NEWEST10, Cartesian KT path of TXYZ5).
 *
 * TEST INFRASTRUCTURE: the parity oracle (see oracle.h).  Built with
 * -O2 -ffp-contract=off so every expression rounds in the order written.
 *
 * Reference files followed (paths relative to the reference root):
 *   RK stage flow ........ NEWEST10/time.h:2-180,300-478
 *   centred flux ......... NEWEST10/hydroKT.h:16-98
 *   eigenvalues .......... NEWEST10/hydroKT.h:176-258
 *   WENO reconstruction .. NEWEST10/hydroKT.h:103-171, NEWEST10/utils.h:57-160
 *   max eigenvalue ....... NEWEST10/hydroKT.h:280-329
 *   Rusanov face flux .... NEWEST10/hydroKT.h:444-506
 *   Var copy ............. NEWEST10/hydroKT.h:364-403
 *   sources .............. NEWEST10/Source.h:3-101 (velocity gradients),
 *                          :208-221 (geometric), :347-517 (shear)
 *   root solve ........... NEWEST10/root.h:212-450
 *   EOS .................. NEWEST10/maindef.h:68-72,320-324 (conformal),
 *                          NEWEST10/s95p.h:13-123 (s95p)
 *   ghost fill ........... NEWEST10/boundary.h:1-275
 *   Cartesian ............ TXYZ5/hydroKT.h:16-62 (flux), TXYZ5/Source.h:
 *                          136-157,270-330 (sources), TXYZ5/time.h:100-165,
 *                          TXYZ5/multiroot.h:187-479 (4-D root),
 *                          TXYZ5/boundary.h; derived shear components from
 *                          NEWEST5/maindef.h:497-501 at tau = 1
 *
 * Deliberate differences from the reference text (all below 1e-15 relative):
 *  - x^(1/4) of the conformal temperature is sqrt(sqrt(x)) instead of
 *    pow(x,0.25) (<= 1 ulp apart; pinned in tests/test_oracle_math.py);
 *  - s95p's exp/pow use oracle/vmath.h instead of libm (pinned likewise);
 *  - the Israel-Stewart shear source is evaluated in covariant tensor form
 *    (see shear_source below) instead of the reference's expanded
 *    computer-algebra output; agreement with the reference's own CalcSource
 *    is pinned by tests/test_oracle_vs_reference.py through oracle/_ref;
 *  - the Cartesian root restates GSL's gnewton and LU routines (GSL is not
 *    in the reference tree; version unpinned), and flags a cell whose
 *    solution is not finite, which the reference would carry on silently.
 */
#include "oracle.h"
#include "vmath.h"
#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define B VIRAL_BORDER
#define SVAR 15
#define GEVFM 0.1973
#define PIE 3.141592653589793

/* Planted mistakes, for the tests that show each pin of the RK
 * orchestration, the k0 broadcast and the GUBSER mirror fails on a wrong
 * oracle (tests/test_oracle_pins.py builds one library per value into a
 * temporary directory).  0, the only value any product or check library is
 * built with, plants nothing. */
#ifndef OR_MUTANT
#define OR_MUTANT 0
#endif
static const int MUT = OR_MUTANT;

/* Work-sharing of the per-cell loops for the all-core CPU baseline (built
 * with -fopenmp; the test build has no -fopenmp and runs them in order).
 * Every loop marked OMP_FOR writes only its own cells from inputs no other
 * iteration writes, so its results do not depend on the thread count; the
 * first failing cell is a min-reduction over cell indices, which grow in the
 * reference's loop order. */
#define OMP_FOR _Pragma("omp parallel for schedule(static)")

struct or_ctx {
  viral_desc d;
  int gx, gy, gz;          /* ghosted dims */
  int il, ir, jl, jr, kl, kr, k0;
  size_t n;                /* cells */
  double* f[VIRAL_NFIELDS];
  double* var;             /* Var  [15][n]  */
  double* var0;            /* Var0 [15][n]  */
  double* var1;            /* Var1 [15][n]  */
  double* var2;            /* Var2 [15][n]  */
  double* res;             /* Result [15][n] */
  double* src;             /* Source [15][n] */
  /* KT work arrays */
  double* F;               /* centred flux of one direction [15][n] */
  double* A;               /* eigenvalue of one direction [n] */
  double* FL; double* FR;  /* [15][n] */
  double* VL; double* VR;  /* [15][n] */
  double* AL; double* AR;  /* [n] */
  double* AmL; double* AmR;/* [n] */
  double* flux;            /* [15][n] */
  uint8_t* flags;
  int32_t* iters[3];
  int bad_cell, bad_stage; unsigned bad_flags;
  double bmax;             /* DebugMSG's BMax of the last step (VIRAL_Q_FIX_VE) */
  int bmax_valid;          /* 0 after the state was set from outside */
};

static inline size_t IX(const or_ctx* c, int i, int j, int k) { return ((size_t)i * c->gy + j) * c->gz + k; }

/* Rows of the 15-row Var/flux/source layout a path evolves: Milne all 15
 * (T^{tau mu}, 10 pi, PI); Cartesian (TXYZ5 SVAR = 10, TXYZ5/maindef.h:196)
 * T^{t mu}, the 5 evolved pi in rows 4..8 and PI in row 14. */
static inline int row_used(const or_ctx* c, int l) { return c->d.coords == VIRAL_MILNE || l < 9 || l == 14; }

/* Full Cartesian pi^{mu nu} (slots {tt,tx,ty,tz,xx,xy,xz,yy,yz,zz}) from the
 * five evolved components p = {xx, yy, xy, xz, yz} (TXYZ5/maindef.h:240):
 *   u_mu pi^{mu x} = 0 -> pi^{tx} = (u^x pi^{xx} + u^y pi^{xy} + u^z pi^{xz})/u^t
 *   u_mu pi^{mu y} = 0 -> pi^{ty} = (u^x pi^{xy} + u^y pi^{yy} + u^z pi^{yz})/u^t
 *   u_mu pi^{mu t} = 0 and pi^mu_mu = 0 -> pi^{zz} below, pi^{tt} = xx+yy+zz
 *   u_mu pi^{mu z} = 0 -> pi^{tz} = (u^x pi^{xz} + u^y pi^{yz} + u^z pi^{zz})/u^t
 * the A1..A5 of NEWEST5/maindef.h:497-501 at tau = 1. */
void or_cart_pi10(const double u[4], const double* p, double pi10[10])
{
  const double u0 = u[0], u1 = u[1], u2 = u[2], u3 = u[3];
  const double p1 = p[0], p2 = p[1], p3 = p[2], p4 = p[3], p5 = p[4];
  const double zz = (p1 * u1 * u1 + 2.0 * p3 * u1 * u2 + p2 * u2 * u2 + 2.0 * u3 * (p4 * u1 + p5 * u2) -
                     u0 * u0 * (p1 + p2)) /
                    (u0 * u0 - u3 * u3);
  pi10[0] = p1 + p2 + zz;
  pi10[1] = (p1 * u1 + p3 * u2 + p4 * u3) / u0;
  pi10[2] = (p3 * u1 + p2 * u2 + p5 * u3) / u0;
  pi10[3] = (p4 * u1 + p5 * u2 + zz * u3) / u0;
  pi10[4] = p1;
  pi10[5] = p3;
  pi10[6] = p4;
  pi10[7] = p2;
  pi10[8] = p5;
  pi10[9] = zz;
}

/* slot of the full tensor holding Cartesian evolved component l */
static const int CART5[5] = {4, 7, 5, 6, 8};

static void cell_pi10(const or_ctx* c, size_t x, const double u[4], double pi10[10])
{
  if (c->d.coords == VIRAL_MILNE) {
    for (int l = 0; l < 10; l++) pi10[l] = c->f[VIRAL_F_PI0 + l][x];
  } else {
    double p[5];
    for (int l = 0; l < 5; l++) p[l] = c->f[VIRAL_F_PI0 + l][x];
    or_cart_pi10(u, p, pi10);
  }
}

/* ------------------------------------------------------------------ EOS */

/* FACTOR of NEWEST10/maindef.h:10, evaluated left to right as written */
static double conformal_factor(const viral_desc* d) { return 3.0 * d->dof * PIE * PIE / 90.0; }

/* s95p closed form, NEWEST10/s95p.h:1-123.  The break points and fit
 * coefficients must be those of the reference; the e == E_n gaps fall through
 * to the last branch exactly as the reference's strict comparisons do (q4). */
#define S95_E1 0.5028563305441270
#define S95_E2 1.62
#define S95_E3 1.86
#define S95_E4 9.9878355786273545
#define S95_ES1 0.1270769021427449
#define S95_ES2 0.446707952467404
#define S95_ES3 1.9402832534193788
#define S95_ES4 3.7292474570977285

static double s95p_p(double e)
{
  if (e < 1e-5) return 0.143375 * e;
  double x = e * GEVFM, r;
  if (x < S95_E1)
    r = 0.3299 * (vm_exp(0.4346 * x) - 1);
  else if (x > S95_E1 && x < S95_E2)
    r = 1.024e-7 * vm_exp(6.041 * x) + 0.007273 + 0.14578 * x;
  else if (x > S95_E2 && x < S95_E3)
    r = 0.30195 * vm_exp(0.31308 * x) - 0.256232;
  else if (x > S95_E3 && x < S95_E4)
    r = 0.332 * x - 0.3223 * vm_pow(x, 0.4585) + 0.1167 * vm_pow(x, -1.233) - 0.003906 * x * vm_exp(-0.05697 * x) +
        0.1436 * x * vm_exp(-0.9131 * x);
  else
    r = 0.3327 * x - 0.3223 * vm_pow(x, 0.4585) - 0.003906 * x * vm_exp(-0.05697 * x);
  return r / GEVFM;
}

static double s95p_a(double e)
{
  if (e < 1e-5) return 0.143375;
  double x = e * GEVFM, r;
  if (x < S95_E1)
    r = 1 * (0.3299 * (0.4346 * vm_exp(0.4346 * x)));
  else if (x > S95_E1 && x < S95_E2)
    r = 1 * (1.024e-7 * 6.041 * vm_exp(6.041 * x) + 0.14578);
  else if (x > S95_E2 && x < S95_E3)
    r = 1 * (0.30195 * 0.31308 * vm_exp(0.31308 * x));
  else if (x > S95_E3 && x < S95_E4)
    r = 1 * (0.332 - 0.3223 * 0.4585 * vm_pow(x, (0.4585 - 1)) + 0.1167 * (-1.233) * vm_pow(x, (-1.233 - 1)) -
             0.003906 * (vm_exp(-0.05697 * x) - 0.05697 * x * vm_exp(-0.05697 * x)) +
             0.1436 * (vm_exp(-0.9131 * x) - 0.9131 * x * vm_exp(-0.9131 * x)));
  else
    r = 1 * (0.3327 - 0.3223 * 0.4585 * vm_pow(x, (0.4585 - 1)) -
             0.003906 * (vm_exp(-0.05697 * x) - 0.05697 * x * vm_exp(-0.05697 * x)));
  if ((x / GEVFM) < 0.12948 * 1e-14) r = 0.286743;
  return r;
}

static double s95p_s(double e)
{
  double x = e * GEVFM, r;
  if (x < S95_ES1)
    r = 12.2304 * vm_pow(x, 1.16849);
  else if (x > S95_ES1 && x < S95_ES2)
    r = 11.9279 * vm_pow(x, 1.15635);
  else if (x > S95_ES2 && x < S95_ES3)
    r = 0.0580578 + 11.833 * vm_pow(x, 1.16187);
  else if (x > S95_ES3 && x < S95_ES4)
    r = 18.202 * x - 62.021814 - 4.85479 * vm_exp(-2.72407e-11 * vm_pow(x, 4.54886)) +
        65.1272 * vm_pow(x, -0.128012) * vm_exp(-0.00369624 * vm_pow(x, 1.18735)) - 4.75253 * vm_pow(x, -1.18423);
  else
    r = 18.202 * x - 63.0218 - 4.85479 * vm_exp(-2.72407e-11 * vm_pow(x, 4.54886)) +
        65.1272 * vm_pow(x, -0.128012) * vm_exp(-0.00369624 * vm_pow(x, 1.18735));
  return vm_pow(r, 0.75);
}

static double s95p_T(double e)
{
  double x = e * GEVFM, r;
  if (x < 0.5143939846236409)
    r = 0.203054 * vm_pow(x, 0.30679);
  else
    r = (x + GEVFM * s95p_p(e)) / s95p_s(e);
  return r / GEVFM;
}

double or_eos_p(const viral_desc* d, double e) { return d->eos == VIRAL_EOS_S95P ? s95p_p(e) : e / 3.0; }
double or_eos_dpde(const viral_desc* d, double e) { return d->eos == VIRAL_EOS_S95P ? s95p_a(e) : 1.0 / 3.0; }
double or_eos_T(const viral_desc* d, double e)
{
  return d->eos == VIRAL_EOS_S95P ? s95p_T(e) : sqrt(sqrt(e / conformal_factor(d)));
}
/* FS: conformal (e+P)*pow(T,-1) -- the compiler folds pow(T,-1) to 1.0/T --
 * or s95p_s(e) */
double or_eos_s(const viral_desc* d, double e, double p, double T)
{
  return d->eos == VIRAL_EOS_S95P ? s95p_s(e) : (e + p) * (1.0 / T);
}

static double eta_of(const viral_desc* d, double s)
{
  return d->eta_law == VIRAL_ETA_S_OVER_4PI ? d->eta_coef * (s / (4.0 * PIE)) : d->eta_coef * s;
}
static double tpi_of(const viral_desc* d, double eta, double p, double e)
{
  switch (d->tpi_law) {
    case VIRAL_TPI_OVER_P: return d->tpi_coef * (eta / p);
    case VIRAL_TPI_CONST: return d->tpi_coef;
    default: return d->tpi_coef * eta / (e + p);
  }
}
/* FZeta / FtauPI laws of the -DBULK build, NEWEST10/maindef.h:82-83,140-141,
 * 191-192,289-290 */
static double zeta_of(const viral_desc* d, double s)
{
  return d->zeta_law == VIRAL_ZETA_S_OVER_4PI ? d->zeta_coef * (s / (4.0 * PIE)) : d->zeta_coef;
}
static double tpib_of(const viral_desc* d, double zeta, double p)
{
  return d->tpib_law == VIRAL_TPIB_OVER_P ? d->tpib_coef * (zeta / p) : d->tpib_coef;
}

/* ------------------------------------------------------------------ WENO */
/* NEWEST10/utils.h:102-160: WENO5-JS, d = (0.3, 0.6, 0.1), eps = 1e-6, p = 2.
 * L is the value at the right face of the centre cell, R at its left face. */
#define WENO_EPS 1e-6
static inline void weno_beta(double qm2, double qm1, double qc, double qp1, double qp2, double b[3])
{
  double a0 = qc - 2 * qp1 + qp2, c0 = 3 * qc - 4 * qp1 + qp2;
  double a1 = qm1 - 2 * qc + qp1, c1 = qm1 - qp1;
  double a2 = qm2 - 2 * qm1 + qc, c2 = qm2 - 4 * qm1 + 3 * qc;
  b[0] = (13. / 12.) * (a0 * a0) + 0.25 * (c0 * c0);
  b[1] = (13. / 12.) * (a1 * a1) + 0.25 * (c1 * c1);
  b[2] = (13. / 12.) * (a2 * a2) + 0.25 * (c2 * c2);
}

double or_weno_L(double qm2, double qm1, double qc, double qp1, double qp2)
{
  double b[3], q[3], al[3], sum, left = 0;
  const double dd[3] = {0.3, 0.6, 0.1};
  weno_beta(qm2, qm1, qc, qp1, qp2, b);
  q[0] = (2 * qc + 5 * qp1 - 1 * qp2) / 6.0;
  q[1] = (-1 * qm1 + 5 * qc + 2 * qp1) / 6.0;
  q[2] = (2 * qm2 - 7 * qm1 + 11 * qc) / 6.0;
  for (int c = 0; c < 3; c++) { double t = WENO_EPS + b[c]; al[c] = dd[c] / (t * t); }
  sum = 0;
  for (int c = 0; c < 3; c++) sum += al[c];
  for (int c = 0; c < 3; c++) left += (al[c] / sum) * q[c];
  return left;
}

double or_weno_R(double qm2, double qm1, double qc, double qp1, double qp2)
{
  double b[3], q[3], al[3], sum, right = 0;
  const double dd[3] = {0.1, 0.6, 0.3};
  weno_beta(qm2, qm1, qc, qp1, qp2, b);
  q[0] = (11 * qc - 7 * qp1 + 2 * qp2) / 6.0;
  q[1] = (2 * qm1 + 5 * qc - 1 * qp1) / 6.0;
  q[2] = (-1 * qm2 + 5 * qm1 + 2 * qc) / 6.0;
  for (int c = 0; c < 3; c++) { double t = WENO_EPS + b[c]; al[c] = dd[c] / (t * t); }
  sum = 0;
  for (int c = 0; c < 3; c++) sum += al[c];
  for (int c = 0; c < 3; c++) right += (al[c] / sum) * q[c];
  return right;
}

/* NEWEST10/utils.h:57-98 */
double or_weno_der(double qm2, double qm1, double qc, double qp1, double qp2, double dx)
{
  double left = or_weno_L(qm2, qm1, qc, qp1, qp2);
  double right = or_weno_R(qm2, qm1, qc, qp1, qp2);
  return (left - right) / dx;
}

double or_vm_exp(double x) { return vm_exp(x); }
double or_vm_log(double x) { return vm_log(x); }
double or_vm_pow(double x, double y) { return vm_pow(x, y); }

/* ------------------------------------------------------------------ setup */

or_ctx* or_create(const viral_desc* d)
{
  or_ctx* c = (or_ctx*)calloc(1, sizeof(or_ctx));
  c->d = *d;
  c->gx = d->nx + 2 * B;
  c->gy = d->ny + 2 * B;
  c->gz = d->lbi ? 1 : d->nz + 2 * B;
  c->il = B; c->ir = B + d->nx;
  c->jl = B; c->jr = B + d->ny;
  if (d->lbi) { c->kl = 0; c->kr = 1; c->k0 = 0; }
  else { c->kl = B; c->kr = B + d->nz; c->k0 = d->k0 >= 0 ? d->k0 : (c->gz - 1) / 2; }
  c->n = (size_t)c->gx * c->gy * c->gz;
  size_t n = c->n;
  for (int f = 0; f < VIRAL_NFIELDS; f++) c->f[f] = (double*)calloc(n, sizeof(double));
  c->var = (double*)calloc(SVAR * n, sizeof(double));
  c->var0 = (double*)calloc(SVAR * n, sizeof(double));
  c->var1 = (double*)calloc(SVAR * n, sizeof(double));
  c->var2 = (double*)calloc(SVAR * n, sizeof(double));
  c->res = (double*)calloc(SVAR * n, sizeof(double));
  c->src = (double*)calloc(SVAR * n, sizeof(double));
  c->F = (double*)calloc(SVAR * n, sizeof(double));
  c->A = (double*)calloc(n, sizeof(double));
  c->FL = (double*)calloc(SVAR * n, sizeof(double));
  c->FR = (double*)calloc(SVAR * n, sizeof(double));
  c->VL = (double*)calloc(SVAR * n, sizeof(double));
  c->VR = (double*)calloc(SVAR * n, sizeof(double));
  c->AL = (double*)calloc(n, sizeof(double));
  c->AR = (double*)calloc(n, sizeof(double));
  c->AmL = (double*)calloc(n, sizeof(double));
  c->AmR = (double*)calloc(n, sizeof(double));
  c->flux = (double*)calloc(SVAR * n, sizeof(double));
  c->flags = (uint8_t*)calloc(n, 1);
  for (int s = 0; s < 3; s++) c->iters[s] = (int32_t*)calloc(n, sizeof(int32_t));
  c->bad_cell = -1;
  return c;
}

void or_destroy(or_ctx* c)
{
  if (!c) return;
  for (int f = 0; f < VIRAL_NFIELDS; f++) free(c->f[f]);
  free(c->var); free(c->var0); free(c->var1); free(c->var2); free(c->res); free(c->src);
  free(c->F); free(c->A); free(c->FL); free(c->FR); free(c->VL); free(c->VR);
  free(c->AL); free(c->AR); free(c->AmL); free(c->AmR); free(c->flux); free(c->flags);
  for (int s = 0; s < 3; s++) free(c->iters[s]);
  free(c);
}

double* or_field(or_ctx* c, int f) { return (f >= 0 && f < VIRAL_NFIELDS) ? c->f[f] : NULL; }
size_t or_cells(const or_ctx* c) { return c->n; }
const uint8_t* or_flags(const or_ctx* c) { return c->flags; }
const int32_t* or_iters(const or_ctx* c, int stage) { return (stage >= 0 && stage < 3) ? c->iters[stage] : NULL; }

int or_first_bad(const or_ctx* c, int* i, int* j, int* k, int* stage, unsigned* flags)
{
  if (c->bad_cell < 0) return -1;
  size_t q = (size_t)c->bad_cell;
  *k = (int)(q % c->gz); q /= c->gz;
  *j = (int)(q % c->gy); q /= c->gy;
  *i = (int)q;
  *stage = c->bad_stage;
  *flags = c->bad_flags;
  return 0;
}

/* ------------------------------------------------------------------ sources */

/* Israel-Stewart bulk relaxation source of PI (conservative form
 * d_tau PI + d_i(v^i PI) = S), the physics of NEWEST10/Source.h:519-527
 * (Milne) and TXYZ5/Source.h:327-333 (Cartesian):
 *   u.d PI = -(PI + zeta theta)/tau_PI - 4/3 PI theta,
 *   theta  = d_mu u^mu, plus u^tau/tau in Milne coordinates (the Christoffel
 *            G^eta_{eta tau} = 1/tau),
 *   S      = u.d PI / u^0 + PI d_i(u^i/u^0).
 * du[a][b] = d_b u^a. */
static double bulk_source(int milne, double tau, const double u[4], double PI, double du[4][4], double zeta,
                          double tPI)
{
  double iu0 = 1.0 / u[0];
  double divv = 0;
  for (int i = 1; i < 4; i++) divv += du[i][i] * iu0 - u[i] * du[0][i] * (iu0 * iu0);
  double theta = du[0][0] + du[1][1] + du[2][2] + du[3][3];
  if (milne) theta = theta + u[0] / tau;
  double rhs = -(PI + zeta * theta) / tPI - (MUT == 11 ? 0.0 : 4.0 / 3.0) * PI * theta;
  return rhs * iu0 + PI * divv;
}

/* Israel-Stewart shear relaxation source for the ten pi^{mu nu} (conservative
 * form: d_tau pi + d_i(v^i pi) = S), the physics of NEWEST10/Source.h:448-516:
 *
 *   u.d pi^{mn} = -(pi^{mn} - 2 eta sigma^{mn})/tau_pi - 4/3 pi^{mn} theta
 *                 - (u^m pi^{na} + u^n pi^{ma}) u.d u_a
 *   sigma^{mn}  = Delta^{mn}_{ab} d^a u^b  (double projector; the four
 *                 components of d_b u^a are treated as independent, as the
 *                 reference's du[a][b] are)
 *   S^{mn}      = (u.d pi^{mn} - Christoffel part of u.d pi^{mn}) / u^0
 *                 + pi^{mn} d_i(u^i/u^0)
 * with d the covariant derivative of the metric diag(1,-1,-1,-tau^2)
 * (Milne; Christoffels G^t_ee = tau, G^e_te = G^e_et = 1/tau) or
 * diag(1,-1,-1,-1) (Cartesian, no Christoffels).
 * du[a][b] = d_b u^a (b = 0 time, 1 x, 2 y, 3 eta). */
static const int PIDX[4][4] = {{0, 1, 2, 3}, {1, 4, 5, 6}, {2, 5, 7, 8}, {3, 6, 8, 9}};

static void shear_source(int milne, double tau, const double u[4], const double pi10[10], double du[4][4], double SH,
                         double tpi, double out[10])
{
  double g[4] = {1.0, -1.0, -1.0, milne ? -(tau * tau) : -1.0}; /* g_{mm} */
  double gi[4] = {1.0, -1.0, -1.0, milne ? -(1.0 / (tau * tau)) : -1.0};
  double pi[4][4];
  for (int m = 0; m < 4; m++)
    for (int n = 0; n < 4; n++) pi[m][n] = pi10[PIDX[m][n]];
  double ul[4];
  for (int m = 0; m < 4; m++) ul[m] = g[m] * u[m];
  /* d_i(u^i/u^0) from the partial derivatives */
  double iu0 = 1.0 / u[0];
  double divv = 0;
  for (int i = 1; i < 4; i++) divv += du[i][i] * iu0 - u[i] * du[0][i] * (iu0 * iu0);
  /* covariant derivative Dc[a][b] = d_b u^a (Christoffel terms added) */
  double Dc[4][4];
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < 4; b++) Dc[a][b] = du[a][b];
  double itau = milne ? 1.0 / tau : 0.0;
  if (milne) {
    Dc[0][3] = Dc[0][3] + tau * u[3];
    Dc[3][0] = Dc[3][0] + itau * u[3];
    Dc[3][3] = Dc[3][3] + itau * u[0];
  }
  double theta = Dc[0][0] + Dc[1][1] + Dc[2][2] + Dc[3][3];
  /* D u_a = g_aa u^b d_b u^a */
  double Dul[4];
  for (int a = 0; a < 4; a++) {
    double s = 0;
    for (int b = 0; b < 4; b++) s += u[b] * Dc[a][b];
    Dul[a] = g[a] * s;
  }
  /* G^{ab} = g^{aa} d_a u^b */
  double G[4][4];
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < 4; b++) G[a][b] = gi[a] * Dc[b][a];
  /* Delta^{mn}_{ab} G^{ab} with Delta^m_a = delta - u^m u_a expands to
   *   G^{mn} - u^n (G u_)^m - u^m (u_ G)^n + u^m u^n (u_ G u_),
   * and Delta_{ab} G^{ab} = sum_a g_aa G^{aa} - u_ G u_ */
  double Gu[4], uG[4], H[4];
  for (int m = 0; m < 4; m++) {
    double s1 = 0, s2 = 0;
    for (int b = 0; b < 4; b++) {
      s1 += G[m][b] * ul[b];
      s2 += ul[b] * G[b][m];
    }
    Gu[m] = s1;
    uG[m] = s2;
  }
  double uGu = 0, trg = 0;
  for (int a = 0; a < 4; a++) {
    uGu += ul[a] * Gu[a];
    trg += g[a] * G[a][a];
  }
  double tr = trg - uGu;
  for (int m = 0; m < 4; m++) H[m] = Gu[m] + uG[m];
  double c3 = uGu + (1.0 / 3.0) * tr;
  /* pi^{ma} D u_a */
  double PD[4];
  for (int m = 0; m < 4; m++) {
    double s = 0;
    for (int a = 0; a < 4; a++) s += pi[m][a] * Dul[a];
    PD[m] = s;
  }
  for (int m = 0; m < 4; m++)
    for (int n = m; n < 4; n++) {
      double sigma = 0.5 * (G[m][n] + G[n][m]) - 0.5 * (u[n] * H[m] + u[m] * H[n]) + u[m] * u[n] * c3;
      if (m == n) sigma = sigma - (1.0 / 3.0) * gi[m] * tr;
      double pu = u[m] * PD[n] + u[n] * PD[m];
      double rhs = -(pi[m][n] - 2.0 * SH * sigma) / tpi - (4.0 / 3.0) * pi[m][n] * theta - pu;
      double chr = 0;
      if (milne) {
        /* u^l (G^m_{la} pi^{an} + G^n_{la} pi^{ma}) */
        double cm = 0, cn = 0;
        if (m == 0) cm = u[3] * tau * pi[3][n];
        if (m == 3) cm = u[0] * itau * pi[3][n] + u[3] * itau * pi[0][n];
        if (n == 0) cn = u[3] * tau * pi[m][3];
        if (n == 3) cn = u[0] * itau * pi[m][3] + u[3] * itau * pi[m][0];
        chr = cm + cn;
      }
      out[PIDX[m][n]] = (rhs - chr) * iu0 + pi[m][n] * divv;
    }
}

void or_shear_source(int milne, double tau, const double u[4], const double pi10[10], const double du16[16],
                     double SH, double tpi, double out[10])
{
  double du[4][4];
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < 4; b++) du[a][b] = du16[4 * a + b];
  shear_source(milne, tau, u, pi10, du, SH, tpi, out);
}

/* 2 eta sigma^{mn}, the Navier-Stokes shear stress of CalcNS
 * (NEWEST10/Source.h:643-751, a computer-algebra expansion): sigma as in
 * shear_source above, sigma^{mn} = Delta^{mn}_{ab} D^a u^b with the Milne
 * Christoffel terms. */
static void ns_shear(int milne, double tau, const double u[4], double du[4][4], double SH, double out[10])
{
  double g[4] = {1.0, -1.0, -1.0, milne ? -(tau * tau) : -1.0};
  double gi[4] = {1.0, -1.0, -1.0, milne ? -(1.0 / (tau * tau)) : -1.0};
  double ul[4];
  for (int m = 0; m < 4; m++) ul[m] = g[m] * u[m];
  double Dc[4][4];
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < 4; b++) Dc[a][b] = du[a][b];
  double itau = milne ? 1.0 / tau : 0.0;
  if (milne) {
    Dc[0][3] = Dc[0][3] + tau * u[3];
    Dc[3][0] = Dc[3][0] + itau * u[3];
    Dc[3][3] = Dc[3][3] + itau * u[0];
  }
  double G[4][4];
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < 4; b++) G[a][b] = gi[a] * Dc[b][a];
  double Gu[4], uG[4], H[4];
  for (int m = 0; m < 4; m++) {
    double s1 = 0, s2 = 0;
    for (int b = 0; b < 4; b++) {
      s1 += G[m][b] * ul[b];
      s2 += ul[b] * G[b][m];
    }
    Gu[m] = s1;
    uG[m] = s2;
  }
  double uGu = 0, trg = 0;
  for (int a = 0; a < 4; a++) {
    uGu += ul[a] * Gu[a];
    trg += g[a] * G[a][a];
  }
  double tr = trg - uGu;
  for (int m = 0; m < 4; m++) H[m] = Gu[m] + uG[m];
  double c3 = uGu + (1.0 / 3.0) * tr;
  for (int m = 0; m < 4; m++)
    for (int n = m; n < 4; n++) {
      double sigma = 0.5 * (G[m][n] + G[n][m]) - 0.5 * (u[n] * H[m] + u[m] * H[n]) + u[m] * u[n] * c3;
      if (m == n) sigma = sigma - (1.0 / 3.0) * gi[m] * tr;
      out[PIDX[m][n]] = 2.0 * SH * sigma;
    }
}

/* The velocity gradients CalcDer4Vel (NEWEST10/Source.h:3-101) leaves on a
 * freshly allocated (zeroed) grid, at cell (i,j,k): the time derivative on the
 * interior; the x derivative on interior y-z rows for 2 <= i < gx-2, the y
 * derivative on interior x-z rows for 2 <= j < gy-2; the eta derivative on
 * interior x-y columns for every k, its stencil clamped at both ends
 * (qm2 = qm1 = qc at k = 0, qm2 = qm1 = q(k-1) at k = 1, and mirrored at the
 * top); zero everywhere else. */
static void ns_du(const or_ctx* c, double ts, int i, int j, int k, double du[4][4])
{
  const viral_desc* d = &c->d;
  const int iin = i >= c->il && i < c->ir, jin = j >= c->jl && j < c->jr, kin = k >= c->kl && k < c->kr;
  const size_t x = IX(c, i, j, k), sx = (size_t)c->gy * c->gz, sy = (size_t)c->gz;
  for (int l = 0; l < 4; l++) {
    const double* ul = c->f[VIRAL_F_U0 + l];
    du[l][0] = (iin && jin && kin) ? (ul[x] - c->f[VIRAL_F_PU0 + l][x]) / ts : 0.0;
    du[l][1] = (jin && kin && i >= 2 && i < c->gx - 2)
                   ? or_weno_der(ul[x - 2 * sx], ul[x - sx], ul[x], ul[x + sx], ul[x + 2 * sx], d->dx)
                   : 0.0;
    du[l][2] = (iin && kin && j >= 2 && j < c->gy - 2)
                   ? or_weno_der(ul[x - 2 * sy], ul[x - sy], ul[x], ul[x + sy], ul[x + 2 * sy], d->dy)
                   : 0.0;
    du[l][3] = 0.0;
    if (!d->lbi && iin && jin) {
      const int Z = c->gz;
      double qm2, qm1, qc = ul[x], qp1, qp2;
      if (k > 1 && k < Z - 2) {
        qm1 = ul[x - 1]; qm2 = ul[x - 2]; qp1 = ul[x + 1]; qp2 = ul[x + 2];
      } else if (k == 0) {
        qm2 = qm1 = qc; qp1 = ul[x + 1]; qp2 = ul[x + 2];
      } else if (k == 1) {
        qm2 = qm1 = ul[x - 1]; qp1 = ul[x + 1]; qp2 = ul[x + 2];
      } else if (k == Z - 2) {
        qm1 = ul[x - 1]; qm2 = ul[x - 2]; qp2 = qp1 = ul[x + 1];
      } else {
        qm1 = ul[x - 1]; qm2 = ul[x - 2]; qp2 = qp1 = qc;
      }
      du[l][3] = or_weno_der(qm2, qm1, qc, qp1, qp2, d->deta);
    }
  }
}

/* CalcNS / NSInit / ZeroInit (NEWEST10/Source.h:542-562, 643-776) over every
 * ghosted cell: mode 0 writes the Navier-Stokes pi to nspi (10 x cells, and
 * with bulk PI = -zeta theta as an 11th row, Source.h:740-745); mode 1 sets
 * pi and PI to them (PI = 0 without bulk), mode 2 both to zero; both then set
 * T^{tau mu} = (e + P - PI) u^0 u^mu - (P - PI) g^{tau mu} + pi^{tau mu}.
 * Milne only (TXYZ5 declares NSInit but does not define it). */
int or_ns_init(or_ctx* c, double tau, double ts, int mode, double* nspi)
{
  const viral_desc* d = &c->d;
  if (d->coords != VIRAL_MILNE || mode < 0 || mode > 2 || (mode == 0 && !nspi)) return -1;
  if (mode != 0) c->bmax_valid = 0;
  double** f = c->f;
  const size_t n = c->n;
  for (int i = 0; i < c->gx; i++)
    for (int j = 0; j < c->gy; j++)
      for (int k = 0; k < c->gz; k++) {
        const size_t x = IX(c, i, j, k);
        double pn[10] = {0}, nsPI = 0.0;
        if (mode != 2) {
          double du[4][4], u[4];
          ns_du(c, ts, i, j, k, du);
          for (int l = 0; l < 4; l++) u[l] = f[VIRAL_F_U0 + l][x];
          const double e = f[VIRAL_F_EN][x], P = f[VIRAL_F_P][x];
          const double T = or_eos_T(d, e);
          const double s = or_eos_s(d, e, P, T);
          ns_shear(1, tau, u, du, eta_of(d, s), pn);
          if (d->bulk) {
            /* theta = d_mu u^mu + u^tau/tau */
            const double theta = du[0][0] + du[1][1] + du[2][2] + du[3][3] + u[0] / tau;
            nsPI = -(zeta_of(d, s) * theta);
          }
        }
        if (mode == 0) {
          for (int l = 0; l < 10; l++) nspi[l * n + x] = pn[l];
          if (d->bulk) nspi[10 * n + x] = nsPI;
          continue;
        }
        for (int l = 0; l < 10; l++) f[VIRAL_F_PI0 + l][x] = pn[l];
        f[VIRAL_F_BULK][x] = nsPI;
        const double e = f[VIRAL_F_EN][x], P = f[VIRAL_F_P][x], PI = nsPI;
        const double u0 = f[VIRAL_F_U0][x];
        f[VIRAL_F_T00][x] = -P + PI + (e + P - PI) * (u0 * u0) + pn[0];
        for (int l = 1; l < 4; l++) f[VIRAL_F_T00 + l][x] = (e + P - PI) * u0 * f[VIRAL_F_U0 + l][x] + pn[l];
      }
  return 0;
}

/* ---------------------------------------------------------------- Landau matching */

/* the reference's velocity test on an eigenvector x (NEWEST10/ginit.h:
 * 567-585): x^0 != 0, v = x^i/x^0 with |v_x|, |v_y|, |v_eta| < 1 and
 * sqrt(v_x^2 + v_y^2 + v_eta^2) < 1 */
static int landau_passes(const double x[4], double v[3])
{
  if (x[0] == 0.0) return 0;
  v[0] = x[1] / x[0];
  v[1] = x[2] / x[0];
  v[2] = x[3] / x[0];
  return fabs(v[0]) < 1 && fabs(v[1]) < 1 && fabs(v[2]) < 1 && sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) < 1;
}

/* 3x3 inverse by cofactors; 0 when singular or not finite */
static int inv3(const double a[3][3], double r[3][3])
{
  const double c00 = a[1][1] * a[2][2] - a[1][2] * a[2][1], c01 = a[1][2] * a[2][0] - a[1][0] * a[2][2],
               c02 = a[1][0] * a[2][1] - a[1][1] * a[2][0];
  const double det = a[0][0] * c00 + a[0][1] * c01 + a[0][2] * c02;
  if (det == 0.0 || !isfinite(det)) return 0;
  const double id = 1.0 / det;
  r[0][0] = c00 * id;
  r[1][0] = c01 * id;
  r[2][0] = c02 * id;
  r[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) * id;
  r[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) * id;
  r[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) * id;
  r[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) * id;
  r[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) * id;
  r[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) * id;
  return 1;
}

/* Landau matching of one cell, the eigenvalue problem of FindVar
 * (NEWEST10/ginit.h:521-629): M x = lambda x with M^a_b = T^{ab} g_bb,
 * g = diag(1, -1, -1, -tau^2).  The reference takes all four eigenpairs from
 * GSL's nonsymmetric solver and keeps those whose eigenvector is real and
 * passes landau_passes; exactly one must, and then E = |lambda|.
 *
 * Here (own method): the timelike eigenvector by Newton's method on v with
 * x = (1, v): r^i = (Mx)^i - (Mx)^0 v^i = 0, J^i_k = M^i_k - M^0_k v^i -
 * (Mx)^0 delta^i_k, from v^i = T^{0i} / (4/3 T^{00}), steps halved while the
 * new v is not timelike.  The other three eigenvectors lie in the
 * g-orthogonal complement of u, which M maps to itself (M is g-self-adjoint):
 * with the basis b_k = e_k - u_k u (k = x, y, eta), M b = b C, and C is
 * symmetric after the change of basis by the Cholesky factor of -b^T g b, so
 * cyclic Jacobi rotations give its eigenvectors.  They are spacelike, so
 * they can pass the test only when tau > 1 (the test ignores the tau^2 of
 * g_etaeta); for tau <= 1 they are not formed.  Returns the number of
 * eigenvectors passing (1 on success), with E and v of the passing one. */
static int landau_cell(double tau, const double T[10], double* E, double vout[3])
{
  const double g[4] = {1.0, -1.0, -1.0, -(tau * tau)};
  double M[4][4];
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < 4; b++) M[a][b] = T[PIDX[a][b]] * g[b];
  double v[3];
  for (int i = 0; i < 3; i++) v[i] = M[i + 1][0] / ((4.0 / 3.0) * M[0][0]);
  if (!(v[0] * v[0] + v[1] * v[1] + (tau * tau) * (v[2] * v[2]) < 1.0)) v[0] = v[1] = v[2] = 0.0;
  double e = 0.0;
  int ok = 0;
  for (int it = 0; it < 64; it++) {
    double Mx[4];
    for (int a = 0; a < 4; a++) Mx[a] = M[a][0] + M[a][1] * v[0] + M[a][2] * v[1] + M[a][3] * v[2];
    e = Mx[0];
    double r[3], J[3][3], Ji[3][3];
    for (int i = 0; i < 3; i++) {
      r[i] = Mx[i + 1] - e * v[i];
      for (int k = 0; k < 3; k++) J[i][k] = M[i + 1][k + 1] - M[0][k + 1] * v[i] - (i == k ? e : 0.0);
    }
    if (!inv3(J, Ji)) break;
    double dv[3];
    for (int i = 0; i < 3; i++) dv[i] = -(Ji[i][0] * r[0] + Ji[i][1] * r[1] + Ji[i][2] * r[2]);
    double t = 1.0, vn[3];
    for (int h = 0; h < 60; h++) {
      for (int i = 0; i < 3; i++) vn[i] = v[i] + t * dv[i];
      if (vn[0] * vn[0] + vn[1] * vn[1] + (tau * tau) * (vn[2] * vn[2]) < 1.0) break;
      t = 0.5 * t;
    }
    if (!(vn[0] * vn[0] + vn[1] * vn[1] + (tau * tau) * (vn[2] * vn[2]) < 1.0)) break;
    const double step = fmax(fabs(vn[0] - v[0]), fmax(fabs(vn[1] - v[1]), fabs(vn[2] - v[2])));
    v[0] = vn[0];
    v[1] = vn[1];
    v[2] = vn[2];
    if (!isfinite(step)) break;
    if (step <= 1e-15) {
      ok = 1;
      break;
    }
  }
  if (!ok) return 0;
  {
    /* an eigenvector to working accuracy */
    double Mx[4], sc = 0.0, res = 0.0;
    for (int a = 0; a < 4; a++) {
      Mx[a] = M[a][0] + M[a][1] * v[0] + M[a][2] * v[1] + M[a][3] * v[2];
      for (int b = 0; b < 4; b++) sc = fmax(sc, fabs(M[a][b]));
    }
    e = Mx[0];
    for (int i = 0; i < 3; i++) res = fmax(res, fabs(Mx[i + 1] - e * v[i]));
    if (!(res <= 1e-11 * sc)) return 0;
  }
  int count = 0;
  {
    const double x[4] = {1.0, v[0], v[1], v[2]};
    double vv[3];
    if (landau_passes(x, vv)) {
      count++;
      *E = fabs(e);
      vout[0] = vv[0];
      vout[1] = vv[1];
      vout[2] = vv[2];
    }
  }
  if (tau > 1.0) {
    const double u0 = 1.0 / sqrt(1.0 - v[0] * v[0] - v[1] * v[1] - (tau * tau) * (v[2] * v[2]));
    const double u[4] = {u0, u0 * v[0], u0 * v[1], u0 * v[2]};
    double Bs[4][3], MB[4][3], S[3][3], Si[3][3], Cm[3][3], H[3][3];
    for (int a = 0; a < 4; a++)
      for (int k = 0; k < 3; k++) Bs[a][k] = (a == k + 1 ? 1.0 : 0.0) - g[k + 1] * u[k + 1] * u[a];
    for (int a = 0; a < 4; a++)
      for (int k = 0; k < 3; k++) MB[a][k] = M[a][0] * Bs[0][k] + M[a][1] * Bs[1][k] + M[a][2] * Bs[2][k] + M[a][3] * Bs[3][k];
    for (int i = 0; i < 3; i++)
      for (int k = 0; k < 3; k++) S[i][k] = Bs[i + 1][k];
    if (!inv3(S, Si)) return 0;
    for (int i = 0; i < 3; i++)
      for (int k = 0; k < 3; k++) Cm[i][k] = Si[i][0] * MB[1][k] + Si[i][1] * MB[2][k] + Si[i][2] * MB[3][k];
    for (int k = 0; k < 3; k++)
      for (int l = 0; l < 3; l++)
        H[k][l] = -(g[0] * Bs[0][k] * Bs[0][l] + g[1] * Bs[1][k] * Bs[1][l] + g[2] * Bs[2][k] * Bs[2][l] +
                    g[3] * Bs[3][k] * Bs[3][l]);
    /* H = L L^T */
    double L[3][3] = {{0}};
    for (int i = 0; i < 3; i++)
      for (int j = 0; j <= i; j++) {
        double s = H[i][j];
        for (int q = 0; q < j; q++) s -= L[i][q] * L[j][q];
        if (i == j) {
          if (!(s > 0.0)) return 0;
          L[i][i] = sqrt(s);
        } else {
          L[i][j] = s / L[j][j];
        }
      }
    double Li[3][3], LiT[3][3], A[3][3], tmp[3][3], V[3][3];
    if (!inv3(L, Li)) return 0;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) LiT[i][j] = Li[j][i];
    /* A = L^T C L^{-T}, symmetrised against rounding */
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) tmp[i][j] = Cm[i][0] * LiT[0][j] + Cm[i][1] * LiT[1][j] + Cm[i][2] * LiT[2][j];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) A[i][j] = L[0][i] * tmp[0][j] + L[1][i] * tmp[1][j] + L[2][i] * tmp[2][j];
    for (int i = 0; i < 3; i++)
      for (int j = i + 1; j < 3; j++) A[i][j] = A[j][i] = 0.5 * (A[i][j] + A[j][i]);
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) V[i][j] = i == j ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 12; sweep++) {
      const double off = fabs(A[0][1]) + fabs(A[0][2]) + fabs(A[1][2]);
      const double dia = fabs(A[0][0]) + fabs(A[1][1]) + fabs(A[2][2]);
      if (!(off > 1e-17 * dia)) break;
      for (int pq = 0; pq < 3; pq++) {
        const int P = pq == 2 ? 1 : 0, Q = pq == 0 ? 1 : 2;
        if (A[P][Q] == 0.0) continue;
        const double th = (A[Q][Q] - A[P][P]) / (2.0 * A[P][Q]);
        const double tt = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
        const double cc = 1.0 / sqrt(tt * tt + 1.0), ss = tt * cc;
        for (int q = 0; q < 3; q++) {
          const double aP = A[q][P], aQ = A[q][Q];
          A[q][P] = cc * aP - ss * aQ;
          A[q][Q] = ss * aP + cc * aQ;
        }
        for (int q = 0; q < 3; q++) {
          const double aP = A[P][q], aQ = A[Q][q];
          A[P][q] = cc * aP - ss * aQ;
          A[Q][q] = ss * aP + cc * aQ;
        }
        for (int q = 0; q < 3; q++) {
          const double vP = V[q][P], vQ = V[q][Q];
          V[q][P] = cc * vP - ss * vQ;
          V[q][Q] = ss * vP + cc * vQ;
        }
      }
    }
    for (int c = 0; c < 3; c++) {
      double y[3], x[4], vv[3];
      for (int k = 0; k < 3; k++) y[k] = LiT[k][0] * V[0][c] + LiT[k][1] * V[1][c] + LiT[k][2] * V[2][c];
      for (int a = 0; a < 4; a++) x[a] = Bs[a][0] * y[0] + Bs[a][1] * y[1] + Bs[a][2] * y[2];
      if (landau_passes(x, vv)) {
        count++;
        *E = fabs(A[c][c]);
        vout[0] = vv[0];
        vout[1] = vv[1];
        vout[2] = vv[2];
      }
    }
  }
  return count;
}

/* The per-cell part of ginit (NEWEST10/ginit.h:876-990) on a T^{mu nu}
 * field (tmunu: 10 x cells, pi index map order, contravariant at tau):
 * Landau matching, then e, Temp = T(e), P = P(e), v and u; at_start 0 (the
 * call at tau - ts) stores prevu = u; at_start 1 sets PI = P - e/3 (minus a
 * third of the ideal part's trace), pi = T - T_ideal - PI (g^{-1} - u u),
 * then PI = 0 (kept with bulk) and T^{tau mu} from e, P, u and pi.  Every ghosted
 * cell; a cell without exactly one passing eigenvector ("EIGEN VALUE SEARCH
 * FAILED", where the reference exits) is left alone and counted.  Returns the
 * number of such cells and the first in loop order in *first (-1 if none). */
long or_landau(or_ctx* c, double tau, const double* tmunu, int at_start, long long* first)
{
  const viral_desc* d = &c->d;
  double** f = c->f;
  const size_t n = c->n;
  long nbad = 0;
  *first = -1;
  c->bmax_valid = 0;
  const double tau2 = tau * tau;
  const double guu[4] = {1.0, -1.0, -1.0, -1.0 / tau2};
  for (size_t x = 0; x < n; x++) {
    double T[10], E = 0.0, v[3] = {0, 0, 0};
    for (int l = 0; l < 10; l++) T[l] = tmunu[l * n + x];
    if (landau_cell(tau, T, &E, v) != 1) {
      if (nbad++ == 0) *first = (long long)x;
      continue;
    }
    const double e = E, P = or_eos_p(d, E);
    f[VIRAL_F_EN][x] = e;
    f[VIRAL_F_TEMP][x] = or_eos_T(d, e);
    f[VIRAL_F_P][x] = P;
    f[VIRAL_F_VX][x] = v[0];
    f[VIRAL_F_VY][x] = v[1];
    f[VIRAL_F_VE][x] = v[2];
    const double u0 = 1.0 / sqrt(1.0 - v[0] * v[0] - v[1] * v[1] - tau2 * v[2] * v[2]);
    const double u[4] = {u0, u0 * v[0], u0 * v[1], u0 * v[2]};
    for (int l = 0; l < 4; l++) f[VIRAL_F_U0 + l][x] = u[l];
    if (!at_start) {
      for (int l = 0; l < 4; l++) f[VIRAL_F_PU0 + l][x] = u[l];
      continue;
    }
    double tid[4][4];
    for (int a = 0; a < 4; a++)
      for (int b = 0; b < 4; b++) tid[a][b] = (e + P) * u[a] * u[b] - P * (a == b ? guu[a] : 0.0);
    const double temp2 = tid[0][0] - tid[1][1] - tid[2][2] - tau * tau * tid[3][3];
    const double PIt = -temp2 / 3.0;
    double pn[10];
    for (int a = 0; a < 4; a++)
      for (int b = a; b < 4; b++) {
        const double PImat = PIt * ((a == b ? guu[a] : 0.0) - u[a] * u[b]);
        pn[PIDX[a][b]] = T[PIDX[a][b]] - tid[a][b] - PImat;
      }
    for (int l = 0; l < 10; l++) f[VIRAL_F_PI0 + l][x] = pn[l];
    const double PI = d->bulk ? PIt : 0.0; /* ginit.h:996-999 zeroes it without BULK */
    f[VIRAL_F_BULK][x] = PI;
    f[VIRAL_F_T00][x] = -P + PI + (e + P - PI) * (u0 * u0) + pn[0];
    for (int l = 1; l < 4; l++) f[VIRAL_F_T00 + l][x] = (e + P - PI) * u0 * u[l] + pn[l];
  }
  return nbad;
}

/* geometric + shear sources on the interior, NEWEST10/Source.h:195-536
 * (Cartesian: TXYZ5/Source.h:136-157 zero geometric terms, 270-330 shear for
 * the five evolved components; under q9 only ghosted k < nz) */
void or_calc_source(or_ctx* c, double tau, double ts, double* S)
{
  const viral_desc* d = &c->d;
  const int milne = d->coords == VIRAL_MILNE;
  double tau2 = tau * tau;
  size_t n = c->n;
  double** f = c->f;
  OMP_FOR
  for (int i = c->il; i < c->ir; i++)
    for (int j = c->jl; j < c->jr; j++)
      for (int k = c->kl; k < c->kr; k++) {
        size_t x = IX(c, i, j, k);
        double e = f[VIRAL_F_EN][x], P = f[VIRAL_F_P][x], PI = f[VIRAL_F_BULK][x];
        double u0 = f[VIRAL_F_U0][x], u3 = f[VIRAL_F_U3][x];
        double p4 = f[VIRAL_F_PI0 + 3][x], p10 = f[VIRAL_F_PI0 + 9][x];
        if (milne) {
          S[0 * n + x] = -(P - PI) - tau2 * (p10 + (e + P - PI) * u3 * u3);
          S[3 * n + x] = (-2.0 * (p4 + (e + P - PI) * u0 * u3));
        } else {
          S[0 * n + x] = 0;
          S[3 * n + x] = 0;
        }
        S[1 * n + x] = 0;
        S[2 * n + x] = 0;
        S[14 * n + x] = 0;
      }
  /* CON (SHASTA) pressure/shear gradient sources, NEWEST10/Source.h:223-336:
   * S^mu += d/dx^i of q^mu_i, WENO derivative over the 5-cell stencil */
  if (d->scheme == VIRAL_SHASTA) {
    int ndir = d->lbi ? 2 : 3;
    const double hh[3] = {d->dx, d->dy, d->deta};
    for (int dir = 0; dir < ndir; dir++) {
      size_t st = dir == 0 ? IX(c, 1, 0, 0) : dir == 1 ? IX(c, 0, 1, 0) : 1;
      OMP_FOR
      for (int i = c->il; i < c->ir; i++)
        for (int j = c->jl; j < c->jr; j++)
          for (int k = c->kl; k < c->kr; k++) {
            size_t x = IX(c, i, j, k);
            double q[4][5];
            for (int o = 0; o < 5; o++) {
              size_t y = x + (o - 2) * st;
              double P = f[VIRAL_F_P][y], PI = f[VIRAL_F_BULK][y];
              double u0 = f[VIRAL_F_U0][y], ud = f[VIRAL_F_U1 + dir][y];
              double p2 = f[VIRAL_F_PI0 + 1][y], p3 = f[VIRAL_F_PI0 + 2][y], p4 = f[VIRAL_F_PI0 + 3][y];
              double tt = milne ? tau : 1.0;
              q[0][o] = -tt * (0);
              if (dir == 0) {
                double p5 = f[VIRAL_F_PI0 + 4][y], p6 = f[VIRAL_F_PI0 + 5][y], p7 = f[VIRAL_F_PI0 + 6][y];
                q[1][o] = -tt * (P - PI - (ud / u0) * p2 + p5);
                q[2][o] = -tt * (-(ud / u0) * p3 + p6);
                q[3][o] = -tt * (-(ud / u0) * p4 + p7);
              } else if (dir == 1) {
                double p6 = f[VIRAL_F_PI0 + 5][y], p8 = f[VIRAL_F_PI0 + 7][y], p9 = f[VIRAL_F_PI0 + 8][y];
                q[1][o] = -tt * (-(ud / u0) * p2 + p6);
                q[2][o] = -tt * (P - PI - (ud / u0) * p3 + p8);
                q[3][o] = -tt * (-(ud / u0) * p4 + p9);
              } else {
                double p7 = f[VIRAL_F_PI0 + 6][y], p9 = f[VIRAL_F_PI0 + 8][y], p10 = f[VIRAL_F_PI0 + 9][y];
                q[1][o] = -tt * (-(ud / u0) * p2 + p7);
                q[2][o] = -tt * (-(ud / u0) * p3 + p9);
                q[3][o] = milne ? -tt * ((P - PI) / (tau * tau) - (ud / u0) * p4 + p10)
                                : -tt * ((P - PI) - (ud / u0) * p4 + p10);
              }
            }
            for (int m = 0; m < 4; m++)
              S[m * n + x] += or_weno_der(q[m][0], q[m][1], q[m][2], q[m][3], q[m][4], hh[dir]);
          }
    }
  }
  /* velocity gradients, NEWEST10/Source.h:3-101 */
  const double hx[3] = {d->dx, d->dy, d->deta};
  OMP_FOR
  for (int i = c->il; i < c->ir; i++)
    for (int j = c->jl; j < c->jr; j++)
      for (int k = c->kl; k < c->kr; k++) {
        size_t x = IX(c, i, j, k);
        double du[4][4];
        double u[4], pi[10];
        for (int l = 0; l < 4; l++) {
          u[l] = f[VIRAL_F_U0 + l][x];
          du[l][0] = (u[l] - f[VIRAL_F_PU0 + l][x]) / ts;
          const double* ul = f[VIRAL_F_U0 + l];
          size_t sx = IX(c, 1, 0, 0), sy = IX(c, 0, 1, 0), sz = 1;
          du[l][1] = or_weno_der(ul[x - 2 * sx], ul[x - sx], ul[x], ul[x + sx], ul[x + 2 * sx], hx[0]);
          du[l][2] = or_weno_der(ul[x - 2 * sy], ul[x - sy], ul[x], ul[x + sy], ul[x + 2 * sy], hx[1]);
          if (!d->lbi)
            du[l][3] = or_weno_der(ul[x - 2 * sz], ul[x - sz], ul[x], ul[x + sz], ul[x + 2 * sz], hx[2]);
          else
            du[l][3] = 0;
        }
        cell_pi10(c, x, u, pi);
        double e = f[VIRAL_F_EN][x], P = f[VIRAL_F_P][x];
        double T = or_eos_T(d, e);
        double s = or_eos_s(d, e, P, T);
        double SH = eta_of(d, s);
        double tpi = tpi_of(d, SH, P, e);
        double out[10];
        shear_source(milne, tau, u, pi, du, SH, tpi, out);
        double sb = 0.0;
        if (d->bulk) {
          const double BU = zeta_of(d, s);
          sb = bulk_source(milne, tau, u, f[VIRAL_F_BULK][x], du, BU, tpib_of(d, BU, P));
        }
        if (milne) {
          for (int l = 0; l < 10; l++) S[(4 + l) * n + x] = out[l];
          S[14 * n + x] = sb;
        } else {
          /* TXYZ5 evaluates the bulk row in the same slice loop (Source.h:272,327) */
          const int live = !(d->quirks & VIRAL_Q_SHEAR_SLICES) || k < d->nz;
          for (int l = 0; l < 5; l++) S[(4 + l) * n + x] = live ? out[CART5[l]] : 0.0;
          for (int l = 5; l < 10; l++) S[(4 + l) * n + x] = 0.0;
          S[14 * n + x] = live ? sb : 0.0;
        }
      }
}

/* ------------------------------------------------------------------ KT */

/* NEWEST10/hydroKT.h:16-98, direction dir (0 x, 1 y, 2 eta), all cells */
static void centre_flux(or_ctx* c, int dir, double tau, double* F)
{
  const int milne = c->d.coords == VIRAL_MILNE;
  double** f = c->f;
  size_t n = c->n;
  OMP_FOR
  for (size_t x = 0; x < n; x++) {
    double e = f[VIRAL_F_EN][x], P = f[VIRAL_F_P][x], PI = f[VIRAL_F_BULK][x];
    double u[4] = {f[VIRAL_F_U0][x], f[VIRAL_F_U1][x], f[VIRAL_F_U2][x], f[VIRAL_F_U3][x]};
    double p[10];
    cell_pi10(c, x, u, p);
    double ud = u[dir + 1];
    double tf = milne ? tau : 1.0;
    int row = dir + 1;
    /* T rows: tau * T^{dir,mu} */
    for (int mu = 0; mu < 4; mu++) {
      double piv = p[PIDX[row][mu]];
      double v;
      if (mu == row) {
        double pp = (P - PI);
        if (dir == 2 && milne) pp = ((P - PI) / (tau * tau));
        v = piv + pp + ud * u[mu] * (e + P - PI);
      } else {
        v = piv + ud * u[mu] * (e + P - PI);
      }
      F[mu * n + x] = milne ? tf * (v) : v;
    }
    if (milne) {
      for (int l = 0; l < 10; l++) F[(4 + l) * n + x] = (p[l] * ud) / u[0];
    } else {  /* TXYZ5/hydroKT.h:31-35 */
      for (int l = 0; l < 5; l++) F[(4 + l) * n + x] = (p[CART5[l]] * ud) / u[0];
      for (int l = 5; l < 10; l++) F[(4 + l) * n + x] = 0.0;
    }
    F[14 * n + x] = (PI * ud) / u[0];
  }
}

void or_centre_flux(or_ctx* c, int dir, double tau, double* F) { centre_flux(c, dir, tau, F); }

/* WENO of one quantity along a direction over interior rows; L stored at the
 * cell, R stored at the cell before it (NEWEST10/hydroKT.h:103-171). */
static void weno_rows(or_ctx* c, int dir, const double* q, double* L, double* R)
{
  int G = dir == 0 ? c->gx : dir == 1 ? c->gy : c->gz;
  size_t st = dir == 0 ? IX(c, 1, 0, 0) : dir == 1 ? IX(c, 0, 1, 0) : 1;
  int a0, a1, b0, b1; /* the two other directions' interior ranges */
  if (dir == 0) { a0 = c->jl; a1 = c->jr; b0 = c->kl; b1 = c->kr; }
  else if (dir == 1) { a0 = c->il; a1 = c->ir; b0 = c->kl; b1 = c->kr; }
  else { a0 = c->il; a1 = c->ir; b0 = c->jl; b1 = c->jr; }
  OMP_FOR
  for (int a = a0; a < a1; a++)
    for (int b = b0; b < b1; b++) {
      size_t base = dir == 0 ? IX(c, 0, a, b) : dir == 1 ? IX(c, a, 0, b) : IX(c, a, b, 0);
      for (int p = 1; p < G; p++) {
        double qm2, qm1, qc, qp1, qp2;
        size_t x = base + (size_t)p * st;
        qc = q[x];
        if (p > 1 && p < G - 2) { qm1 = q[x - st]; qm2 = q[x - 2 * st]; qp1 = q[x + st]; qp2 = q[x + 2 * st]; }
        else if (p == 1) { qm2 = qm1 = q[x - st]; qp1 = q[x + st]; qp2 = q[x + 2 * st]; }
        else if (p == G - 2) { qm2 = q[x - 2 * st]; qm1 = q[x - st]; qp1 = qp2 = q[x + st]; }
        else { qm2 = q[x - 2 * st]; qm1 = q[x - st]; qp1 = qp2 = qc; }
        L[x] = or_weno_L(qm2, qm1, qc, qp1, qp2);
        R[x - st] = or_weno_R(qm2, qm1, qc, qp1, qp2);
      }
    }
}

/* SHASTA flux-corrected transport, NEWEST10/hydroshasta.h:8-32 (transport
 * velocities), 81-98 (stage), 101-284 (one direction): low-order transported
 * UTD with Q+- = (1/2 -+ v lambda)/(1 +- (v_{+-1} - v) lambda), flux-limited
 * antidiffusion A = s max(0, min(s dUTD_{+1}, |dUTD/8|, s dUTD_{-1})), and
 * the rate (-U + UTD - A_i + A_{i-1}) / taustep. */
static void shasta_explicit(or_ctx* c, double tau, double taustep, double* Result)
{
  const viral_desc* d = &c->d;
  const int milne = d->coords == VIRAL_MILNE;
  size_t n = c->n;
  double** f = c->f;
  memset(Result, 0, SVAR * n * sizeof(double));
  OMP_FOR
  for (size_t x = 0; x < n; x++) {
    double tf = milne ? tau : 1.0;
    for (int mu = 0; mu < 4; mu++) c->var[mu * n + x] = milne ? tf * f[VIRAL_F_T00 + mu][x] : f[VIRAL_F_T00 + mu][x];
    for (int l = 0; l < 10; l++) c->var[(4 + l) * n + x] = f[VIRAL_F_PI0 + l][x];
    c->var[14 * n + x] = f[VIRAL_F_BULK][x];
  }
  int ndir = d->lbi ? 2 : 3;
  const double hx[3] = {d->dx, d->dy, d->deta};
  double* NV = c->F;    /* transport velocity of one direction, [15][n] */
  double* UTD = c->FL;  /* [n] */
  double* AD = c->FR;   /* [n] */
  for (int dir = 0; dir < ndir; dir++) {
    /* CalcNumVel: row 0 moves with T^{dir,tau}/T^{tau tau}, the rest with u^dir/u^0 */
    OMP_FOR
    for (size_t x = 0; x < n; x++) {
      double e = f[VIRAL_F_EN][x], P = f[VIRAL_F_P][x], PI = f[VIRAL_F_BULK][x];
      double u0 = f[VIRAL_F_U0][x], ud = f[VIRAL_F_U1 + dir][x];
      double pd = f[VIRAL_F_PI0 + 1 + dir][x];
      NV[0 * n + x] = (pd + ud * u0 * (e + P - PI)) / f[VIRAL_F_T00][x];
      for (int l = 1; l < SVAR; l++) NV[l * n + x] = ud / u0;
    }
    int G = dir == 0 ? c->gx : dir == 1 ? c->gy : c->gz;
    size_t st = dir == 0 ? IX(c, 1, 0, 0) : dir == 1 ? IX(c, 0, 1, 0) : 1;
    double lam = taustep / hx[dir];
    int a0, a1, b0, b1;
    if (dir == 0) { a0 = c->jl; a1 = c->jr; b0 = c->kl; b1 = c->kr; }
    else if (dir == 1) { a0 = c->il; a1 = c->ir; b0 = c->kl; b1 = c->kr; }
    else { a0 = c->il; a1 = c->ir; b0 = c->jl; b1 = c->jr; }
    int lo = dir == 0 ? c->il : dir == 1 ? c->jl : c->kl;
    int hi = dir == 0 ? c->ir : dir == 1 ? c->jr : c->kr;
    for (int l = 0; l < SVAR; l++) {
      const double* U = c->var + l * n;
      const double* V = NV + l * n;
      OMP_FOR
      for (int a = a0; a < a1; a++)
        for (int b = b0; b < b1; b++) {
          size_t base = dir == 0 ? IX(c, 0, a, b) : dir == 1 ? IX(c, a, 0, b) : IX(c, a, b, 0);
          for (int p = 1; p < G - 1; p++) {
            size_t x = base + (size_t)p * st;
            double Vc = V[x], Vp = V[x + st], Vm = V[x - st];
            double Qp = (0.5 - Vc * lam) / (1.0 + (Vp - Vc) * lam);
            double Qm = (0.5 + Vc * lam) / (1.0 - (Vm - Vc) * lam);
            UTD[x] = 0.5 * ((Qp * Qp * (U[x + st] - U[x])) + (Qm * Qm * (U[x - st] - U[x]))) + (Qp + Qm) * U[x];
          }
          for (int p = 2; p < G - 3; p++) {
            size_t x = base + (size_t)p * st;
            double signa = 1;
            double temp = 1 * 0.125 * (UTD[x + st] - UTD[x]);
            if (temp < 0) signa = -1;
            double term1 = signa * (UTD[x + 2 * st] - UTD[x + st]);
            double term3 = signa * (UTD[x] - UTD[x - st]);
            double mn = fmin(fmin(term1, fabs(temp)), term3);
            AD[x] = signa * fmax(0, mn);
          }
          for (int p = lo; p < hi; p++) {
            size_t x = base + (size_t)p * st;
            double pr = (-U[x] + UTD[x] - AD[x] + AD[x - st]) / taustep;
            Result[l * n + x] += pr;
          }
        }
    }
  }
}

/* NEWEST10/hydroKT.h:406-436 with fvX/fvY/fvZ (444-506) */
void or_hydro_explicit(or_ctx* c, double tau, double taustep, double* Result)
{
  if (c->d.scheme == VIRAL_SHASTA) {
    shasta_explicit(c, tau, taustep, Result);
    return;
  }
  const viral_desc* d = &c->d;
  const int milne = d->coords == VIRAL_MILNE;
  size_t n = c->n;
  double** f = c->f;
  /* ClearResultVariable + CopyPrimaryVariablesToVar (all cells) */
  memset(Result, 0, SVAR * n * sizeof(double));
  OMP_FOR
  for (size_t x = 0; x < n; x++) {
    double tf = milne ? tau : 1.0;
    for (int mu = 0; mu < 4; mu++) c->var[mu * n + x] = milne ? tf * f[VIRAL_F_T00 + mu][x] : f[VIRAL_F_T00 + mu][x];
    for (int l = 0; l < 10; l++) c->var[(4 + l) * n + x] = f[VIRAL_F_PI0 + l][x];
    c->var[14 * n + x] = f[VIRAL_F_BULK][x];
  }
  int ndir = d->lbi ? 2 : 3;
  const double hx[3] = {d->dx, d->dy, d->deta};
  for (int dir = 0; dir < ndir; dir++) {
    /* eigenvalue u^dir/u^0: interior only under q1 (ghost entries stay 0) */
    memset(c->A, 0, n * sizeof(double));
    OMP_FOR
    for (int i = 0; i < c->gx; i++)
      for (int j = 0; j < c->gy; j++)
        for (int k = 0; k < c->gz; k++) {
          int inside = i >= c->il && i < c->ir && j >= c->jl && j < c->jr && k >= c->kl && k < c->kr;
          if (!inside && (d->quirks & VIRAL_Q_GHOST_EIGEN)) continue;
          size_t x = IX(c, i, j, k);
          c->A[x] = f[VIRAL_F_U1 + dir][x] / f[VIRAL_F_U0][x];
        }
    centre_flux(c, dir, tau, c->F);
    memset(c->FL, 0, SVAR * n * sizeof(double));
    memset(c->FR, 0, SVAR * n * sizeof(double));
    memset(c->VL, 0, SVAR * n * sizeof(double));
    memset(c->VR, 0, SVAR * n * sizeof(double));
    memset(c->AL, 0, n * sizeof(double));
    memset(c->AR, 0, n * sizeof(double));
    for (int l = 0; l < SVAR; l++) {
      if (!row_used(c, l)) continue;
      weno_rows(c, dir, c->F + l * n, c->FL + l * n, c->FR + l * n);
      weno_rows(c, dir, c->var + l * n, c->VL + l * n, c->VR + l * n);
    }
    weno_rows(c, dir, c->A, c->AL, c->AR);
    /* max eigenvalue at the interior cells (and at the cell before the row
     * when q1 is off) */
    memset(c->AmL, 0, n * sizeof(double));
    memset(c->AmR, 0, n * sizeof(double));
    int lo_ext = (d->quirks & VIRAL_Q_GHOST_EIGEN) ? 0 : 1;
    OMP_FOR
    for (int i = c->il - (dir == 0 ? lo_ext : 0); i < c->ir; i++)
      for (int j = c->jl - (dir == 1 ? lo_ext : 0); j < c->jr; j++)
        for (int k = c->kl - (dir == 2 ? lo_ext : 0); k < c->kr; k++) {
          size_t x = IX(c, i, j, k);
          double ml = 0, mr = 0;
          if (fabs(c->AL[x]) > ml) ml = fabs(c->AL[x]);
          if (fabs(c->AR[x]) > mr) mr = fabs(c->AR[x]);
          c->AmL[x] = ml;
          c->AmR[x] = mr;
        }
    size_t st = dir == 0 ? IX(c, 1, 0, 0) : dir == 1 ? IX(c, 0, 1, 0) : 1;
    /* face fluxes at faces p+1/2 for p = lo-1 .. hi-1 */
    for (int l = 0; l < SVAR; l++)
      OMP_FOR
      for (int i = c->il - (dir == 0); i < c->ir; i++)
        for (int j = c->jl - (dir == 1); j < c->jr; j++)
          for (int k = c->kl - (dir == 2); k < c->kr; k++) {
            size_t x = IX(c, i, j, k);
            c->flux[l * n + x] = 0.5 * (c->FL[l * n + x] + c->FR[l * n + x]) -
                                 0.5 * (fmax(c->AmL[x], c->AmR[x])) * (c->VR[l * n + x] - c->VL[l * n + x]);
          }
    for (int l = 0; l < SVAR; l++)
      OMP_FOR
      for (int i = c->il; i < c->ir; i++)
        for (int j = c->jl; j < c->jr; j++)
          for (int k = c->kl; k < c->kr; k++) {
            if (!row_used(c, l)) continue;
            size_t x = IX(c, i, j, k);
            double pr = -(1.0 / hx[dir]) * (c->flux[l * n + x] - c->flux[l * n + x - st]);
            Result[l * n + x] += pr;
          }
  }
}

/* ------------------------------------------------------------------ root */

/* f and f' of NEWEST10/root.h:212-285 */
static inline void feps(const viral_desc* d, double eps, double Ttt, double pi1, double PI, double M2, double* fv,
                        double* dfv)
{
  double Pe = or_eos_p(d, eps);
  *fv = eps - (Ttt - pi1) + (M2 / ((Ttt - pi1) + Pe - PI));
  double den = Ttt - pi1 + Pe - PI;
  *dfv = 1 - ((M2 / (den * den)) * or_eos_dpde(d, eps));
}

static int newton_shortcut = 1;
static long newton_cycles[5];

/* GSL newton semantics (roots/newton.c): x_{n+1} = x_n - f/f', stop when
 * x_{n+1} == x_n (gsl_root_test_delta with epsabs 0, epsrel 1e-30, as called
 * at NEWEST10/root.h:354), error when f' == 0 or f, f' are not finite. */
double or_newton_cell(const viral_desc* d, double tau, double T00, double T10, double T20, double T30, double pi0,
                      double pi1, double pi2, double pi3, double PI, double guess, int32_t* iters, unsigned* flags)
{
  const int milne = d->coords == VIRAL_MILNE;
  double M2 = milne ? (T10 - pi1) * (T10 - pi1) + (T20 - pi2) * (T20 - pi2) + tau * tau * (T30 - pi3) * (T30 - pi3)
                    : (T10 - pi1) * (T10 - pi1) + (T20 - pi2) * (T20 - pi2) + (T30 - pi3) * (T30 - pi3);
  double x = guess, fv, dfv;
  feps(d, x, T00, pi0, PI, M2, &fv, &dfv);
  int it = 0;
  unsigned fl = 0;
  double h1 = x, h2 = x, h3 = x, h4 = x; /* x_{it-1} .. x_{it-4} */
  for (;;) {
    it++;
    if (dfv == 0.0) { fl |= VIRAL_FLAG_SOLVER; break; }
    double xn = x - (fv / dfv);
    feps(d, xn, T00, pi0, PI, M2, &fv, &dfv);
    double xp = x;
    x = xn;
    if (!isfinite(fv) || !isfinite(dfv)) { fl |= VIRAL_FLAG_SOLVER; break; }
    if (x == xp) break;
    if (it >= d->max_iter) { fl |= VIRAL_FLAG_MAXITER; break; }
    h4 = h3; h3 = h2; h2 = h1; h1 = xp;
    if (newton_shortcut) {
      /* The Newton map depends on x alone, so x_it == x_{it-p} means the
       * iterates repeat with period p until max_iter: jump to x_{max_iter}
       * = x_{it-p+r}, r = (max_iter - it) mod p.  Same result and count as
       * running the loop out (pinned by tests/test_oracle_root.py). */
      int per = (it >= 2 && x == h2) ? 2 : (it >= 3 && x == h3) ? 3 : (it >= 4 && x == h4) ? 4 : 0;
      if (per) {
        int r = (d->max_iter - it) % per;
        const double hist[5] = {x, h1, h2, h3, h4};
        x = hist[per - r];
        it = d->max_iter;
        fl |= VIRAL_FLAG_MAXITER;
        newton_cycles[per]++;
        break;
      }
    }
  }
  *iters = it;
  *flags = fl;
  return x;
}

/* Safeguard of the root solve, VIRAL_Q_ROOT_SAFEGUARD (not a reference
 * behaviour: the reference aborts): the root of f(e) = e - A + M2/(A + P(e)
 * - PI) on [0, A] by the safeguarded Newton method -- a Newton step while it
 * stays inside the bracket and shrinks |f| fast enough, otherwise a
 * bisection; the bracket follows the sign of f.  A bracket exists exactly
 * when f(0) < 0 <= f(A), i.e. M2 < A (A - PI) with P(0) = 0.  Stops when the
 * step or the bracket falls below 4 ulp of e, or after 200 iterations.  The
 * device version (viral_amd/csrc/physics.h, root_bracketed) evaluates the
 * same expressions in the same order. */
static void brk_f(const viral_desc* d, double M2, double A, double PI, double x, double* fx, double* dfx)
{
  const double Q = A + or_eos_p(d, x) - PI;
  *fx = x - A + M2 / Q;
  *dfx = 1.0 - M2 * or_eos_dpde(d, x) / (Q * Q);
}

static int root_bracketed(const viral_desc* d, double M2, double A, double PI, double* root, int32_t* iters)
{
  double lo = 0.0, hi = A, fx, dfx;
  if (!(A > 0.0)) return 0;
  brk_f(d, M2, A, PI, lo, &fx, &dfx);
  if (!(fx < 0.0)) return 0;
  brk_f(d, M2, A, PI, hi, &fx, &dfx);
  if (!(fx >= 0.0)) return 0;
  if (fx == 0.0) {
    *root = hi;
    *iters = 0;
    return 1;
  }
  double x = 0.5 * (lo + hi), dx = hi - lo, dxold = dx;
  brk_f(d, M2, A, PI, x, &fx, &dfx);
  for (int it = 1; it <= 200; it++) {
    const int leaves = ((x - hi) * dfx - fx) * ((x - lo) * dfx - fx) > 0.0;
    if (leaves || !(fabs(2.0 * fx) <= fabs(dxold * dfx))) {
      dxold = dx;
      dx = 0.5 * (hi - lo);
      x = lo + dx;
    } else {
      dxold = dx;
      dx = fx / dfx;
      x = x - dx;
    }
    brk_f(d, M2, A, PI, x, &fx, &dfx);
    if (fx < 0.0) lo = x;
    else hi = x;
    if (fx == 0.0 || fabs(dx) <= 4e-16 * fabs(x) || hi - lo <= 4e-16 * hi) {
      *root = x;
      *iters = it;
      return 1;
    }
  }
  *root = x;
  *iters = 200;
  return 1;
}

double or_root_safeguard(const viral_desc* d, double M2, double A, double PI, int32_t* iters)
{
  double r;
  int32_t it = 0;
  if (!root_bracketed(d, M2, A, PI, &r, &it)) r = NAN;
  if (iters) *iters = it;
  return r;
}

void or_set_newton_shortcut(int on) { newton_shortcut = on; }
long or_newton_cycle_count(int period) { return (period >= 0 && period < 5) ? newton_cycles[period] : -1; }

/* ---- Cartesian root (TXYZ5/multiroot.h) */

/* conservation_f of TXYZ5/multiroot.h:203-240: residuals of T^{t mu} =
 * (e+P-PI) u^t u^mu - (P-PI) g^{t mu} + pi^{t mu} in the unknowns
 * (e, u^x, u^y, u^z), u^t = sqrt(1 + u.u), multiplied out as the reference
 * writes them.  prm = {W, X, Y, Z, p1..p5, PI}, W..Z = T^{tt}..T^{tz}. */
static void cart_f(const viral_desc* d, const double* prm, const double* x, double* y)
{
  const double W = prm[0], X = prm[1], Y = prm[2], Z = prm[3];
  const double p1 = prm[4], p2 = prm[5], p3 = prm[6], p4 = prm[7], p5 = prm[8], PI = prm[9];
  const double e = x[0], u1 = x[1], u2 = x[2], u3 = x[3];
  const double P = or_eos_p(d, e);
  const double q = 1 + u1 * u1 + u2 * u2 + u3 * u3;
  const double sq = sqrt(q);
  y[0] = -(u1 * X) - u2 * Y - u3 * Z + (-e + W) * sq;
  y[1] = -(p3 * u2) - p4 * u3 - P * u1 * q - u1 * (p1 + e * q - PI * q) + X * sq;
  y[2] = -(p3 * u1) - (e + p2 - PI) * u2 - p5 * u3 - (e - PI) * u2 * (u1 * u1 + u2 * u2 + u3 * u3) - P * u2 * q + Y * sq;
  y[3] = -(p4 * u1) - p5 * u2 - P * u3 + u3 * (p1 + p2 + PI - W) + Z * sq;
}

/* conservation_df of TXYZ5/multiroot.h:243-319: J[r][c] = d y_r / d x_c */
static void cart_df(const viral_desc* d, const double* prm, const double* x, double J[4][4])
{
  const double W = prm[0], X = prm[1], Y = prm[2], Z = prm[3];
  const double p1 = prm[4], p2 = prm[5], p3 = prm[6], p4 = prm[7], p5 = prm[8], PI = prm[9];
  const double e = x[0], u1 = x[1], u2 = x[2], u3 = x[3];
  const double P = or_eos_p(d, e);
  const double a = or_eos_dpde(d, e);
  const double q = 1 + u1 * u1 + u2 * u2 + u3 * u3;
  const double sq = sqrt(q), isq = 1.0 / sqrt(q);
  const double w = e + P - PI;
  J[0][0] = -sq;
  J[0][1] = -X + u1 * (-e + W) * isq;
  J[0][2] = -Y + u2 * (-e + W) * isq;
  J[0][3] = -Z + u3 * (-e + W) * isq;
  J[1][0] = -((1 + a) * u1 * q);
  J[1][1] = -p1 - w * (1 + 3 * u1 * u1 + u2 * u2 + u3 * u3) + u1 * X * isq;
  J[1][2] = -p3 - 2 * w * u1 * u2 + u2 * X * isq;
  J[1][3] = -p4 - 2 * w * u1 * u3 + u3 * X * isq;
  J[2][0] = -((1 + a) * u2 * q);
  J[2][1] = -p3 - 2 * w * u1 * u2 + u1 * Y * isq;
  J[2][2] = -p2 - w * (1 + u1 * u1 + 3 * u2 * u2 + u3 * u3) + u2 * Y * isq;
  J[2][3] = -p5 - 2 * w * u2 * u3 + u3 * Y * isq;
  J[3][0] = (-a * u3);
  J[3][1] = -p4 + u1 * Z * isq;
  J[3][2] = -p5 + u2 * Z * isq;
  J[3][3] = -P + p1 + p2 + PI - W + u3 * Z * isq;
}

void or_cart_f(const viral_desc* d, const double* prm, const double* x, double* y) { cart_f(d, prm, x, y); }
void or_cart_df(const viral_desc* d, const double* prm, const double* x, double* J16)
{
  double J[4][4];
  cart_df(d, prm, x, J);
  for (int r = 0; r < 4; r++)
    for (int c = 0; c < 4; c++) J16[4 * r + c] = J[r][c];
}

static double enorm4(const double* f)
{
  double e2 = 0;
  for (int i = 0; i < 4; i++) e2 += f[i] * f[i];
  return sqrt(e2);
}

#define GN_MAX_ITER 100 /* TXYZ5/multiroot.h:388 */

/* gsl_multiroot_fdfsolver_gnewton driven as TXYZ5/multiroot.h:385-418:
 * per iteration an LU solve J d = f with partial pivoting (GSL
 * gsl_linalg_LU_decomp / LU_solve), trial x - t d from t = 1, and while the
 * residual norm rises and t > DBL_EPSILON the step shrinks by
 * t *= (sqrt(1 + 6 theta) - 1)/(3 theta), theta = |f_trial|/|f|; stop when
 * gsl_multiroot_test_delta(dx, x, 0, 1e-30) holds for every component or after
 * 100 iterations.  A zero pivot (GSL's "matrix is singular" error, which
 * aborts the reference) or a non-finite result sets VIRAL_FLAG_SOLVER. */
int or_cart_gnewton(const viral_desc* d, const double* prm, double* x, unsigned* flags)
{
  double f[4], J[4][4];
  unsigned fl = 0;
  cart_f(d, prm, x, f);
  cart_df(d, prm, x, J);
  double phi = enorm4(f);
  int iter = 0;
  for (;;) {
    iter++;
    double LU[4][4];
    int perm[4] = {0, 1, 2, 3};
    memcpy(LU, J, sizeof(LU));
    for (int j = 0; j < 3; j++) {
      double mx = fabs(LU[j][j]);
      int ip = j;
      for (int i = j + 1; i < 4; i++) {
        double aij = fabs(LU[i][j]);
        if (aij > mx) { mx = aij; ip = i; }
      }
      if (ip != j) {
        for (int k = 0; k < 4; k++) { double t = LU[j][k]; LU[j][k] = LU[ip][k]; LU[ip][k] = t; }
        int t = perm[j]; perm[j] = perm[ip]; perm[ip] = t;
      }
      double ajj = LU[j][j];
      if (ajj != 0.0)
        for (int i = j + 1; i < 4; i++) {
          double aij = LU[i][j] / ajj;
          LU[i][j] = aij;
          for (int k = j + 1; k < 4; k++) LU[i][k] = LU[i][k] - aij * LU[j][k];
        }
    }
    if (LU[0][0] == 0.0 || LU[1][1] == 0.0 || LU[2][2] == 0.0 || LU[3][3] == 0.0) { fl |= VIRAL_FLAG_SOLVER; break; }
    double dv[4];
    for (int i = 0; i < 4; i++) dv[i] = f[perm[i]];
    for (int i = 1; i < 4; i++) {
      double t = dv[i];
      for (int j = 0; j < i; j++) t -= LU[i][j] * dv[j];
      dv[i] = t;
    }
    dv[3] = dv[3] / LU[3][3];
    for (int i = 2; i >= 0; i--) {
      double t = dv[i];
      for (int j = i + 1; j < 4; j++) t -= LU[i][j] * dv[j];
      dv[i] = t / LU[i][i];
    }
    double t = 1, phi0 = phi, phi1, xt[4];
    for (;;) {
      for (int i = 0; i < 4; i++) xt[i] = x[i] - t * dv[i];
      cart_f(d, prm, xt, f);
      phi1 = enorm4(f);
      if (phi1 > phi0 && t > DBL_EPSILON) {
        double theta = phi1 / phi0;
        double u = (sqrt(1.0 + 6.0 * theta) - 1.0) / (3.0 * theta);
        t *= u;
        continue;
      }
      break;
    }
    double dx[4];
    for (int i = 0; i < 4; i++) {
      x[i] = xt[i];
      dx[i] = -t * dv[i];
    }
    cart_df(d, prm, x, J);
    phi = phi1;
    int ok = 1;
    for (int i = 0; i < 4; i++) {
      double tol = 1e-30 * fabs(x[i]);
      if (!(fabs(dx[i]) < tol || dx[i] == 0)) { ok = 0; break; }
    }
    if (ok) break;
    if (iter >= GN_MAX_ITER) { fl |= VIRAL_FLAG_MAXITER; break; }
  }
  for (int i = 0; i < 4; i++)
    if (!isfinite(x[i])) fl |= VIRAL_FLAG_SOLVER;
  *flags = fl;
  return iter;
}

/* TXYZ5/multiroot.h:322-479 on the interior: gnewton from the guess
 * (e, u^x, u^y, u^z) of the cell, then Temp, u^t = sqrt(1 + u.u), v = u/u^t,
 * P; a cell with |v| >= 1 or e <= 0 is where the reference exits. */
static int cart_root(or_ctx* c, int32_t* iters, uint8_t* flags)
{
  const viral_desc* d = &c->d;
  double** f = c->f;
  int bad = -1;
  OMP_FOR
  for (int i = c->il; i < c->ir; i++)
    for (int j = c->jl; j < c->jr; j++)
      for (int k = c->kl; k < c->kr; k++) {
        size_t x = IX(c, i, j, k);
        double prm[10] = {f[VIRAL_F_T00][x], f[VIRAL_F_T10][x], f[VIRAL_F_T20][x], f[VIRAL_F_T30][x],
                          f[VIRAL_F_PI0][x], f[VIRAL_F_PI0 + 1][x], f[VIRAL_F_PI0 + 2][x], f[VIRAL_F_PI0 + 3][x],
                          f[VIRAL_F_PI0 + 4][x], f[VIRAL_F_BULK][x]};
        double r[4] = {f[VIRAL_F_EN][x], f[VIRAL_F_U1][x], f[VIRAL_F_U2][x], f[VIRAL_F_U3][x]};
        unsigned fl;
        int it = or_cart_gnewton(d, prm, r, &fl);
        f[VIRAL_F_EN][x] = r[0];
        f[VIRAL_F_TEMP][x] = or_eos_T(d, r[0]);
        f[VIRAL_F_U1][x] = r[1];
        f[VIRAL_F_U2][x] = r[2];
        f[VIRAL_F_U3][x] = r[3];
        f[VIRAL_F_U0][x] = sqrt(1 + r[1] * r[1] + r[2] * r[2] + r[3] * r[3]);
        if (iters) iters[x] = it;
        if (flags) flags[x] = (uint8_t)fl;
      }
  int badx = INT_MAX;
  _Pragma("omp parallel for schedule(static) reduction(min : badx)")
  for (int i = c->il; i < c->ir; i++)
    for (int j = c->jl; j < c->jr; j++)
      for (int k = c->kl; k < c->kr; k++) {
        size_t x = IX(c, i, j, k);
        double En = f[VIRAL_F_EN][x], u0 = f[VIRAL_F_U0][x];
        double vx = f[VIRAL_F_U1][x] / u0, vy = f[VIRAL_F_U2][x] / u0, ve = f[VIRAL_F_U3][x] / u0;
        unsigned fl = flags ? flags[x] : 0;
        if (sqrt(vx * vx + vy * vy + ve * ve) >= 1) fl |= VIRAL_FLAG_ACAUSAL;
        if (En <= 0) fl |= VIRAL_FLAG_ENEG;
        f[VIRAL_F_VX][x] = vx;
        f[VIRAL_F_VY][x] = vy;
        f[VIRAL_F_VE][x] = ve;
        f[VIRAL_F_P][x] = or_eos_p(d, En);
        if (d->t_freeze > 0 && f[VIRAL_F_TEMP][x] < d->t_freeze) fl |= VIRAL_FLAG_FREEZEOUT;
        if (flags) flags[x] = (uint8_t)fl;
        if ((fl & (VIRAL_FLAG_ACAUSAL | VIRAL_FLAG_ENEG | VIRAL_FLAG_SOLVER)) && (int)x < badx) badx = (int)x;
      }
  bad = badx == INT_MAX ? -1 : badx;
  return bad;
}

/* NEWEST10/root.h:301-450 on the interior */
int or_root(or_ctx* c, double tau, int32_t* iters, uint8_t* flags)
{
  const viral_desc* d = &c->d;
  const int milne = d->coords == VIRAL_MILNE;
  double** f = c->f;
  if (!milne) return cart_root(c, iters, flags);
  c->bmax_valid = 0;
  double t2 = milne ? tau * tau : 1.0;
  int bad = -1;
  OMP_FOR
  for (int i = c->il; i < c->ir; i++)
    for (int j = c->jl; j < c->jr; j++)
      for (int k = c->kl; k < c->kr; k++) {
        size_t x = IX(c, i, j, k);
        int32_t it;
        unsigned fl;
        double r = or_newton_cell(d, tau, f[VIRAL_F_T00][x], f[VIRAL_F_T10][x], f[VIRAL_F_T20][x], f[VIRAL_F_T30][x],
                                  f[VIRAL_F_PI0][x], f[VIRAL_F_PI0 + 1][x], f[VIRAL_F_PI0 + 2][x],
                                  f[VIRAL_F_PI0 + 3][x], f[VIRAL_F_BULK][x], f[VIRAL_F_EN][x], &it, &fl);
        if ((d->quirks & VIRAL_Q_ROOT_SAFEGUARD) && ((fl & VIRAL_FLAG_SOLVER) || !(r > 0))) {
          const double a1 = f[VIRAL_F_T10][x] - f[VIRAL_F_PI0 + 1][x], a2 = f[VIRAL_F_T20][x] - f[VIRAL_F_PI0 + 2][x],
                       a3 = f[VIRAL_F_T30][x] - f[VIRAL_F_PI0 + 3][x];
          const double M2 = a1 * a1 + a2 * a2 + tau * tau * a3 * a3;
          double er;
          int32_t ir;
          if (root_bracketed(d, M2, f[VIRAL_F_T00][x] - f[VIRAL_F_PI0][x], f[VIRAL_F_BULK][x], &er, &ir)) {
            r = er;
            it = it + ir;
            fl = (fl & ~(VIRAL_FLAG_SOLVER | VIRAL_FLAG_MAXITER)) | VIRAL_FLAG_RESCUED;
          }
        }
        f[VIRAL_F_EN][x] = r;
        f[VIRAL_F_TEMP][x] = or_eos_T(d, r);
        f[VIRAL_F_P][x] = or_eos_p(d, r);
        if (iters) iters[x] = it;
        if (flags) flags[x] = (uint8_t)fl;
      }
  int badx = INT_MAX;
  _Pragma("omp parallel for schedule(static) reduction(min : badx)")
  for (int i = c->il; i < c->ir; i++)
    for (int j = c->jl; j < c->jr; j++)
      for (int k = c->kl; k < c->kr; k++) {
        size_t x = IX(c, i, j, k);
        double Ttt = f[VIRAL_F_T00][x], Ttx = f[VIRAL_F_T10][x], Tty = f[VIRAL_F_T20][x], Tte = f[VIRAL_F_T30][x];
        double e = f[VIRAL_F_EN][x], P = f[VIRAL_F_P][x], PI = f[VIRAL_F_BULK][x];
        double pi1 = f[VIRAL_F_PI0][x], pi2 = f[VIRAL_F_PI0 + 1][x], pi3 = f[VIRAL_F_PI0 + 2][x],
               pi4 = f[VIRAL_F_PI0 + 3][x];
        double vx = (Ttx - pi2) / (Ttt - pi1 + P - PI);
        double vy = (Tty - pi3) / (Ttt - pi1 + P - PI);
        double ve = (Tte - pi4) / (Ttt - pi1 + P - PI);
        unsigned fl = flags ? flags[x] : 0;
        if (vx * vx + vy * vy + t2 * ve * ve >= 1) fl |= VIRAL_FLAG_ACAUSAL;
        if (e <= 0) fl |= VIRAL_FLAG_ENEG;
        double g = 1.0 / sqrt(1 - vx * vx - vy * vy - t2 * ve * ve);
        f[VIRAL_F_U0][x] = g;
        f[VIRAL_F_U1][x] = g * vx;
        f[VIRAL_F_U2][x] = g * vy;
        f[VIRAL_F_U3][x] = g * ve;
        f[VIRAL_F_VX][x] = vx;
        f[VIRAL_F_VY][x] = vy;
        f[VIRAL_F_VE][x] = ve;
        if (d->t_freeze > 0 && f[VIRAL_F_TEMP][x] < d->t_freeze) fl |= VIRAL_FLAG_FREEZEOUT;
        if (flags) flags[x] = (uint8_t)fl;
        if ((fl & (VIRAL_FLAG_ACAUSAL | VIRAL_FLAG_ENEG | VIRAL_FLAG_SOLVER)) && (int)x < badx) badx = (int)x;
      }
  bad = badx == INT_MAX ? -1 : badx;
  if ((d->quirks & VIRAL_Q_K0_BROADCAST) && !d->lbi && milne) {
    OMP_FOR
    for (int i = c->il; i < c->ir; i++)
      for (int j = c->jl; j < c->jr; j++)
        for (int k = c->kl; k < c->kr; k++) {
          if (k == c->k0) continue;
          size_t x = IX(c, i, j, k), x0 = IX(c, i, j, MUT == 7 ? c->k0 + 1 : c->k0);
          double vx = f[VIRAL_F_VX][x0], vy = f[VIRAL_F_VY][x0], ve = f[VIRAL_F_VE][x0];
          f[VIRAL_F_VE][x] = ve;
          f[VIRAL_F_VX][x] = vx;
          f[VIRAL_F_VY][x] = vy;
          double u0 = 1.0 / sqrt(1.0 - vx * vx - vy * vy - (MUT == 8 ? 1.0 : t2) * ve * ve);
          f[VIRAL_F_U0][x] = u0;
          f[VIRAL_F_U1][x] = u0 * vx;
          f[VIRAL_F_U2][x] = u0 * vy;
          f[VIRAL_F_U3][x] = u0 * ve;
        }
  }
  return bad;
}

/* ------------------------------------------------------------------ ghosts */

static const int BFIELDS[] = {VIRAL_F_T00, VIRAL_F_T10, VIRAL_F_T20, VIRAL_F_T30, VIRAL_F_EN,  VIRAL_F_VX,
                              VIRAL_F_VY,  VIRAL_F_VE,  VIRAL_F_P,   VIRAL_F_U0,  VIRAL_F_U1,  VIRAL_F_U2,
                              VIRAL_F_U3,  VIRAL_F_PU0, VIRAL_F_PU1, VIRAL_F_PU2, VIRAL_F_PU3, VIRAL_F_PI0,
                              VIRAL_F_PI0 + 1, VIRAL_F_PI0 + 2, VIRAL_F_PI0 + 3, VIRAL_F_PI0 + 4,
                              VIRAL_F_PI0 + 5, VIRAL_F_PI0 + 6, VIRAL_F_PI0 + 7, VIRAL_F_PI0 + 8,
                              VIRAL_F_PI0 + 9, VIRAL_F_BULK};
#define NBF ((int)(sizeof(BFIELDS) / sizeof(BFIELDS[0])))

/* NEWEST10/boundary.h:1-275: zero-gradient copy into the x, y (and eta)
 * ghost layers of interior rows; corners are never written (q8). */
void or_boundary(or_ctx* c)
{
  double** f = c->f;
  const int mirror = (c->d.quirks & VIRAL_Q_GUBSER_MIRROR) != 0 && c->d.coords == VIRAL_MILNE;
  OMP_FOR
  for (int j = c->jl; j < c->jr; j++)
    for (int k = c->kl; k < c->kr; k++)
      for (int g = 0; g < B; g++) {
        size_t xl = IX(c, g, j, k), sl = IX(c, c->il, j, k);
        size_t xr = IX(c, c->ir + g, j, k), sr = IX(c, c->ir - 1, j, k);
        for (int q = 0; q < NBF; q++) {
          f[BFIELDS[q]][xl] = f[BFIELDS[q]][sl];
          f[BFIELDS[q]][xr] = f[BFIELDS[q]][sr];
        }
        if (mirror) {
          size_t ml = IX(c, MUT == 9 ? c->il + g : c->il + B - g - 1, j, k), mr = IX(c, c->ir - g - 1, j, k);
          f[VIRAL_F_U1][xl] = f[VIRAL_F_U1][ml];
          if (MUT != 10) f[VIRAL_F_U2][xl] = f[VIRAL_F_U2][ml];
          f[VIRAL_F_U1][xr] = f[VIRAL_F_U1][mr];
          f[VIRAL_F_U2][xr] = f[VIRAL_F_U2][mr];
        }
      }
  OMP_FOR
  for (int i = c->il; i < c->ir; i++)
    for (int k = c->kl; k < c->kr; k++)
      for (int g = 0; g < B; g++) {
        size_t xl = IX(c, i, g, k), sl = IX(c, i, c->jl, k);
        size_t xr = IX(c, i, c->jr + g, k), sr = IX(c, i, c->jr - 1, k);
        for (int q = 0; q < NBF; q++) {
          f[BFIELDS[q]][xl] = f[BFIELDS[q]][sl];
          f[BFIELDS[q]][xr] = f[BFIELDS[q]][sr];
        }
        if (mirror) {
          size_t ml = IX(c, i, c->jl + B - g - 1, k), mr = IX(c, i, c->jr - g - 1, k);
          f[VIRAL_F_U1][xl] = f[VIRAL_F_U1][ml];
          f[VIRAL_F_U2][xl] = f[VIRAL_F_U2][ml];
          f[VIRAL_F_U1][xr] = f[VIRAL_F_U1][mr];
          f[VIRAL_F_U2][xr] = f[VIRAL_F_U2][mr];
        }
      }
  if (!c->d.lbi)
    OMP_FOR
    for (int i = c->il; i < c->ir; i++)
      for (int j = c->jl; j < c->jr; j++)
        for (int g = 0; g < B; g++) {
          size_t xl = IX(c, i, j, g), sl = IX(c, i, j, c->kl);
          size_t xr = IX(c, i, j, c->kr + g), sr = IX(c, i, j, c->kr - 1);
          for (int q = 0; q < NBF; q++) {
            f[BFIELDS[q]][xl] = f[BFIELDS[q]][sl];
            f[BFIELDS[q]][xr] = f[BFIELDS[q]][sr];
          }
        }
}

double or_max_temp_gev(or_ctx* c)
{
  double tmax = 0;
  for (int i = c->il; i < c->ir; i++)
    for (int j = c->jl; j < c->jr; j++)
      for (int k = c->kl; k < c->kr; k++) {
        double t = GEVFM * or_eos_T(&c->d, c->f[VIRAL_F_EN][IX(c, i, j, k)]);
        if (t > tmax) tmax = t;
      }
  return tmax;
}

/* ------------------------------------------------------------------ RK */

/* T = Var/tau, pi = Var, PI = Var on the interior (NEWEST10/time.h:100-162;
 * Cartesian T = Var and five pi, TXYZ5/time.h:100-165) */
static void update_primary(or_ctx* c, const double* V, double tau)
{
  const int milne = c->d.coords == VIRAL_MILNE;
  size_t n = c->n;
  OMP_FOR
  for (int i = c->il; i < c->ir; i++)
    for (int j = c->jl; j < c->jr; j++)
      for (int k = c->kl; k < c->kr; k++) {
        size_t x = IX(c, i, j, k);
        for (int mu = 0; mu < 4; mu++) c->f[VIRAL_F_T00 + mu][x] = milne ? (V[mu * n + x]) / (tau) : V[mu * n + x];
        for (int l = 0; l < (milne ? 10 : 5); l++) c->f[VIRAL_F_PI0 + l][x] = (V[(4 + l) * n + x]);
        c->f[VIRAL_F_BULK][x] = (V[14 * n + x]);
      }
}

static void update_prevu(or_ctx* c)
{
  OMP_FOR
  for (int i = c->il; i < c->ir; i++)
    for (int j = c->jl; j < c->jr; j++)
      for (int k = c->kl; k < c->kr; k++) {
        size_t x = IX(c, i, j, k);
        for (int l = 0; l < 4; l++) c->f[VIRAL_F_PU0 + l][x] = c->f[VIRAL_F_U0 + l][x];
      }
}

/* one stage: sources + fluxes at (tau_s, ts_src), L = Result + Source */
static void stage_rhs(or_ctx* c, double tau_s, double ts_src, double taustep)
{
  or_calc_source(c, tau_s, ts_src, c->src);
  or_hydro_explicit(c, tau_s, taustep, c->res);
  size_t n = c->n;
  for (int l = 0; l < SVAR; l++)
    OMP_FOR
    for (int i = c->il; i < c->ir; i++)
      for (int j = c->jl; j < c->jr; j++)
        for (int k = c->kl; k < c->kr; k++) {
          size_t x = IX(c, i, j, k);
          c->res[l * n + x] = (c->res[l * n + x] + c->src[l * n + x]);
        }
}

#define FORI for (int i = c->il; i < c->ir; i++) for (int j = c->jl; j < c->jr; j++) for (int k = c->kl; k < c->kr; k++)
/* FORI with the iterations shared among threads: for loops whose cells are independent */
#define FORI_PAR OMP_FOR FORI

static int stage_root(or_ctx* c, double tau_r, int stage)
{
  uint8_t* fl = (uint8_t*)malloc(c->n);
  memset(fl, 0, c->n);
  int bad = or_root(c, tau_r, c->iters[stage], fl);
  FORI_PAR {
    size_t x = IX(c, i, j, k);
    c->flags[x] = (uint8_t)((c->flags[x] & ~VIRAL_FLAG_FREEZEOUT) | fl[x]);
  }
  if (bad >= 0 && c->bad_cell < 0) {
    c->bad_cell = bad;
    c->bad_stage = stage;
    c->bad_flags = fl[bad];
  }
  free(fl);
  or_boundary(c);
  return bad;
}

/* BMax of DebugMSG (NEWEST10/debug.h:30-74): the largest
 * sqrt(vx^2 + vy^2 + tau^2 ve^2) over the interior, from 0 (NaN never counts) */
static double debug_bmax(const or_ctx* c, double tau)
{
  double m = 0.0;
  for (int i = c->il; i < c->ir; i++)
    for (int j = c->jl; j < c->jr; j++)
      for (int k = c->kl; k < c->kr; k++) {
        const size_t x = IX(c, i, j, k);
        const double vx = c->f[VIRAL_F_VX][x], vy = c->f[VIRAL_F_VY][x], ve = c->f[VIRAL_F_VE][x];
        const double b = sqrt(vx * vx + vy * vy + tau * tau * ve * ve);
        if (b > m) m = b;
      }
  return m;
}

void or_state_changed(or_ctx* c) { c->bmax_valid = 0; }

static int step_core(or_ctx* c, double tau, double ts);

/* One step; with VIRAL_Q_FIX_VE the FIX build's end of step: FixTwo with the
 * previous BMax, then DebugMSG's BMax at the step's start time
 * (NEWEST10/time.h:385-392) */
int or_step(or_ctx* c, double tau, double ts)
{
  const int fix = (c->d.quirks & VIRAL_Q_FIX_VE) && c->d.coords == VIRAL_MILNE;
  if (fix && !c->bmax_valid) {
    c->bmax = debug_bmax(c, tau);
    c->bmax_valid = 1;
  }
  const int r = step_core(c, tau, ts);
  if (fix) {
    if (c->bmax > 0.98)
      for (int i = c->il; i < c->ir; i++)
        for (int j = c->jl; j < c->jr; j++)
          for (int k = c->kl; k < c->kr; k++) c->f[VIRAL_F_VE][IX(c, i, j, k)] = 0;
    c->bmax = debug_bmax(c, tau);
  }
  return r;
}

static int step_core(or_ctx* c, double tau, double ts)
{
  const viral_desc* d = &c->d;
  size_t n = c->n;
  memset(c->flags, 0, n);
  for (int s = 0; s < 3; s++) memset(c->iters[s], 0, n * sizeof(int32_t));
  c->bad_cell = -1;
  double* L = c->res;
  double sub = d->rk == 3 ? ts / 3.0 : d->rk == 2 ? ts / 2 : ts;
  if (d->quirks & VIRAL_Q_SHASTA_FULLSTEP) sub = ts; /* the stable variant (include/viral.h) */
  /* stage 1 */
  stage_rhs(c, tau, ts, sub);
  memcpy(c->var0, c->var, SVAR * n * sizeof(double)); /* CalcVar0 (interior entries used) */
  for (int l = 0; l < SVAR; l++) FORI_PAR {
      size_t x = l * n + IX(c, i, j, k);
      c->var1[x] = c->var0[x] + ts * (L[x]);
    }
  update_primary(c, c->var1, tau + ts);
  update_prevu(c);
  stage_root(c, tau + ts, 0);
  if (d->rk == 1) return c->bad_cell >= 0 ? VIRAL_ERR_PHYSICS : VIRAL_OK;
  /* stage 2 */
  stage_rhs(c, tau + ts, ts, sub);
  if (d->rk == 2) {
    const double w2 = MUT == 5 ? 1.0 : 0.5;
    for (int l = 0; l < SVAR; l++) FORI_PAR {
        size_t x = l * n + IX(c, i, j, k);
        c->var2[x] = 0.5 * c->var0[x] + 0.5 * c->var1[x] + w2 * ts * L[x];
      }
    update_primary(c, c->var2, tau + ts);
    if (MUT == 6) update_prevu(c);
    stage_root(c, tau + ts, 1);
    return c->bad_cell >= 0 ? VIRAL_ERR_PHYSICS : VIRAL_OK;
  }
  const double a0 = MUT == 1 ? 0.5 : 0.75, a1 = MUT == 1 ? 0.5 : 0.25;
  for (int l = 0; l < SVAR; l++) FORI_PAR {
      size_t x = l * n + IX(c, i, j, k);
      c->var2[x] = a0 * c->var0[x] + a1 * c->var1[x] + a1 * ts * L[x];
    }
  const double tmid = MUT == 2 ? tau + ts : tau + ts / 2.0;
  update_primary(c, c->var2, tmid);
  if (MUT == 6) update_prevu(c);
  stage_root(c, tmid, 1);
  /* stage 3 */
  stage_rhs(c, tmid, MUT == 3 ? ts : ts / 2.0, sub);
  const double b0 = MUT == 4 ? 0.25 : 1.0 / 3.0, b1 = MUT == 4 ? 0.75 : 2.0 / 3.0;
  for (int l = 0; l < SVAR; l++) FORI_PAR {
      size_t x = l * n + IX(c, i, j, k);
      c->var1[x] = b0 * c->var0[x] + b1 * c->var2[x] + b1 * ts * L[x];
    }
  update_primary(c, c->var1, tau + ts);
  stage_root(c, tau + ts, 2);
  return c->bad_cell >= 0 ? VIRAL_ERR_PHYSICS : VIRAL_OK;
}

/* ------------------------------------------------------------------ per-step diagnostics */

/* Running maximum of the reference's `if (x > m) m = x` from m = 0: the
 * first cell in loop order that reaches the maximum keeps its index. */
static inline void keep_first_max(double x, size_t at, double* m, long long* where)
{
  if (x > *m) { *m = x; *where = (long long)at; }
}

/* Root check after a step.
 * Milne, NEWEST10/root.h:1-199 -- six statistics:
 *   0  |e - (T^tt - pi^tt) + M^2 / ((T^tt - pi^tt) + P(e) - PI)|, M^2 as in
 *      the root solve (root.h:20-44)
 *   1-4  |pi^{mu nu} u_nu| for mu = t, x, y, eta (transversality, root.h:
 *        72-75), 5  |pi^mu_mu| (root.h:90)
 * Cartesian, TXYZ5/multiroot.h:7-150 -- nine statistics: 0-3 the residuals
 * of the four conservation equations at (e, u) (multiroot.h:45-48, with
 * pow(x, 0.5) as sqrt and pow(u, 2) as u*u), 4-7 transversality, 8 the
 * trace, whose maximum follows the reference's `fabs(trace) > trace` test
 * (multiroot.h:109): the |trace| of the last cell in loop order with a
 * negative trace, 0 if none.
 * mx: maxima (at: first cell in loop order reaching each, -1 when it is 0;
 * for Cartesian statistic 8 the cell the value came from); avg: sum of the
 * absolute values over the interior / number of interior cells, summed in
 * loop order.  Returns the number of statistics. */
int or_check_root(or_ctx* c, double tau, double* mx, double* avg, long long* at)
{
  const viral_desc* d = &c->d;
  double** f = c->f;
  const int ns = d->coords == VIRAL_MILNE ? 6 : 9;
  double sum[9];
  long count = 0;
  for (int s = 0; s < ns; s++) { mx[s] = 0; sum[s] = 0; at[s] = -1; }
  FORI {
    size_t x = IX(c, i, j, k);
    count++;
    double e = f[VIRAL_F_EN][x], PI = f[VIRAL_F_BULK][x], P = or_eos_p(d, e);
    double W = f[VIRAL_F_T00][x], X = f[VIRAL_F_T00 + 1][x], Y = f[VIRAL_F_T00 + 2][x], Z = f[VIRAL_F_T00 + 3][x];
    if (d->coords == VIRAL_MILNE) {
      double Ttt = W, Ttx = X, Tty = Y, Tte = Z;
      double pi1 = f[VIRAL_F_PI0][x], pi2 = f[VIRAL_F_PI0 + 1][x], pi3 = f[VIRAL_F_PI0 + 2][x],
             pi4 = f[VIRAL_F_PI0 + 3][x];
      double M2 = (Ttx - pi2) * (Ttx - pi2) + (Tty - pi3) * (Tty - pi3) + tau * tau * (Tte - pi4) * (Tte - pi4);
      double res = e - (Ttt - pi1) + (M2 / ((Ttt - pi1) + P - PI));
      keep_first_max(fabs(res), x, &mx[0], &at[0]);
      sum[0] += fabs(res);
    } else {
      double u1 = f[VIRAL_F_U0 + 1][x], u2 = f[VIRAL_F_U0 + 2][x], u3 = f[VIRAL_F_U0 + 3][x];
      double p1 = f[VIRAL_F_PI0][x], p2 = f[VIRAL_F_PI0 + 1][x], p3 = f[VIRAL_F_PI0 + 2][x],
             p4 = f[VIRAL_F_PI0 + 3][x], p5 = f[VIRAL_F_PI0 + 4][x];
      double Q = 1 + u1 * u1 + u2 * u2 + u3 * u3, S3 = u1 * u1 + u2 * u2 + u3 * u3, rq = sqrt(Q);
      double y[4];
      y[0] = -(u1 * X) - u2 * Y - u3 * Z + (-e + W) * rq;
      y[1] = -(p3 * u2) - p4 * u3 - P * u1 * Q - u1 * (p1 + e * Q - PI * Q) + X * rq;
      y[2] = -(p3 * u1) - (e + p2 - PI) * u2 - p5 * u3 - (e - PI) * u2 * S3 - P * u2 * Q + Y * rq;
      y[3] = -(p4 * u1) - p5 * u2 - P * u3 + u3 * (p1 + p2 + PI - W) + Z * rq;
      for (int l = 0; l < 4; l++) {
        keep_first_max(fabs(y[l]), x, &mx[l], &at[l]);
        sum[l] += fabs(y[l]);
      }
    }
  }
  const int r0 = d->coords == VIRAL_MILNE ? 1 : 4;
  FORI {
    size_t x = IX(c, i, j, k);
    double u[4], p[10];
    for (int m = 0; m < 4; m++) u[m] = f[VIRAL_F_U0 + m][x];
    if (d->coords == VIRAL_MILNE)
      for (int l = 0; l < 10; l++) p[l] = f[VIRAL_F_PI0 + l][x];
    else
      cell_pi10(c, x, u, p);
    double u0 = u[0], u1 = u[1], u2 = u[2], u3 = u[3];
    double rows[4], trace;
    if (d->coords == VIRAL_MILNE) {
      rows[0] = p[0] * u0 - p[1] * u1 - p[2] * u2 - p[3] * u3 * tau * tau;
      rows[1] = p[1] * u0 - p[4] * u1 - p[5] * u2 - p[6] * u3 * tau * tau;
      rows[2] = p[2] * u0 - p[5] * u1 - p[7] * u2 - p[8] * u3 * tau * tau;
      rows[3] = p[3] * u0 - p[6] * u1 - p[8] * u2 - p[9] * u3 * tau * tau;
      trace = p[0] - p[4] - p[7] - p[9] * tau * tau;
    } else {
      rows[0] = p[0] * u0 - p[1] * u1 - p[2] * u2 - p[3] * u3;
      rows[1] = p[1] * u0 - p[4] * u1 - p[5] * u2 - p[6] * u3;
      rows[2] = p[2] * u0 - p[5] * u1 - p[7] * u2 - p[8] * u3;
      rows[3] = p[3] * u0 - p[6] * u1 - p[8] * u2 - p[9] * u3;
      trace = p[0] - p[4] - p[7] - p[9];
    }
    for (int l = 0; l < 4; l++) {
      keep_first_max(fabs(rows[l]), x, &mx[r0 + l], &at[r0 + l]);
      sum[r0 + l] += fabs(rows[l]);
    }
    if (d->coords == VIRAL_MILNE) {
      keep_first_max(fabs(trace), x, &mx[5], &at[5]);
    } else if (fabs(trace) > trace) {
      mx[8] = fabs(trace);
      at[8] = (long long)x;
    }
    sum[r0 + 4] += fabs(trace);
  }
  for (int s = 0; s < ns; s++) avg[s] = sum[s] / count;
  return ns;
}

/* |v| of a cell: sqrt(vx^2 + vy^2 + tau^2 ve^2) (NEWEST10/debug.h:35,
 * utils.h:242); without tau in Cartesian coordinates (TXYZ5/debug.h:54) */
static inline double speed(const or_ctx* c, size_t x, double tau)
{
  double vx = c->f[VIRAL_F_VX][x], vy = c->f[VIRAL_F_VX + 1][x], ve = c->f[VIRAL_F_VX + 2][x];
  if (c->d.coords == VIRAL_MILNE) return sqrt(vx * vx + vy * vy + tau * tau * ve * ve);
  return sqrt(vx * vx + vy * vy + ve * ve);
}

/* End-of-step extrema, NEWEST10/debug.h:3-100 (TXYZ5/debug.h): the largest
 * |v| (`b >= vmax` from 0: the last cell in loop order reaching it), the
 * largest e (`>` from 0: the first cell) and the smallest e below 0 (`<`
 * from 0: the first cell; 0 and no cell when none is negative).
 * out = {vmax, vx, vy, ve of that cell, emax, emin}; at = {vmax cell, emax
 * cell, emin cell}, ghosted linear indices, -1 where none.  vmax is the
 * reference's BMax. */
void or_debug_msg(or_ctx* c, double tau, double out[6], long long at[3])
{
  double vmax = 0, emax = 0, emin = 0;
  at[0] = at[1] = at[2] = -1;
  FORI {
    size_t x = IX(c, i, j, k);
    double b = speed(c, x, tau), e = c->f[VIRAL_F_EN][x];
    if (e > emax) { emax = e; at[1] = (long long)x; }
    if (e < emin) { emin = e; at[2] = (long long)x; }
    if (b >= vmax) { vmax = b; at[0] = (long long)x; }
  }
  out[0] = vmax;
  for (int m = 0; m < 3; m++) out[1 + m] = at[0] >= 0 ? c->f[VIRAL_F_VX + m][at[0]] : 0.0;
  out[4] = emax;
  out[5] = emin;
}

/* Start-of-step physics check, NEWEST10/utils.h:231-258 (TXYZ5/utils.h:
 * 223-248): interior cells with |v| >= 1 (n[0]) and with e <= 0 (n[1]), and
 * the first of each in loop order (-1: none). */
void or_check_physics(or_ctx* c, double tau, long long n[2], long long first[2])
{
  n[0] = n[1] = 0;
  first[0] = first[1] = -1;
  FORI {
    size_t x = IX(c, i, j, k);
    if (speed(c, x, tau) >= 1) { if (!n[0]++) first[0] = (long long)x; }
    if (c->f[VIRAL_F_EN][x] <= 0) { if (!n[1]++) first[1] = (long long)x; }
  }
}

/* Output records of NEWEST10/write.h: PRINTVAR = 17 doubles per cell,
 * {e, T, P, vx, vy, ve, pi[0..9], PI} (write.h:1,87-103).  Full grid
 * (WriteResults, write.h:286-365): every freq-th cell in x and y and every
 * freqz-th in eta, in (i, j, k) order, which is the order of the file's
 * records (OFFSET, write.h:267-284, when freq divides nx and ny and freqz
 * divides nz).  xy != 0 (WriteResultsXY, write.h:52-119): the k0 slice in
 * (i, j) order.  Writes the records to out (may be NULL) and returns their
 * doubles count; 0 when the strides do not divide the grid. */
size_t or_output(or_ctx* c, int freq, int freqz, int xy, double* out)
{
  const viral_desc* d = &c->d;
  if (freq < 1 || freqz < 1 || d->nx % freq || d->ny % freq || (!xy && d->nz % freqz)) return 0;
  size_t q = 0;
  for (int i = c->il; i < c->ir; i += freq)
    for (int j = c->jl; j < c->jr; j += freq)
      for (int k = xy ? c->k0 : c->kl; xy ? k == c->k0 : k < c->kr; k += xy ? 1 : freqz) {
        size_t x = IX(c, i, j, k);
        if (out) {
          double* r = out + q;
          r[0] = c->f[VIRAL_F_EN][x];
          r[1] = c->f[VIRAL_F_TEMP][x];
          r[2] = c->f[VIRAL_F_P][x];
          for (int m = 0; m < 3; m++) r[3 + m] = c->f[VIRAL_F_VX + m][x];
          for (int l = 0; l < 10; l++) r[6 + l] = c->f[VIRAL_F_PI0 + l][x];
          r[16] = c->f[VIRAL_F_BULK][x];
        }
        q += 17;
      }
  return q;
}
