// physics.h -- per-cell device physics of the VIRAL viscous step (gfx950, FP64).
//
// Every function here evaluates its expression in the reference's operation
// order, compiled with -ffp-contract=off: IEEE add/mul/div/sqrt round the same
// on gfx950 as on the host, so in exact mode the GPU step reproduces the CPU
// restatement bit for bit.  Reference lines are cited per function.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

namespace viral {

constexpr double GEVFM = 0.1973;               // NEWEST10/main.cpp:44
constexpr double PIE = 3.141592653589793;      // NEWEST10/main.cpp:45
constexpr double WENO_EPS = 1e-6;              // NEWEST10/maindef.h:355

// ---------------------------------------------------------------- exp/log/pow
// The s95p EOS (NEWEST10/s95p.h) calls exp/pow.  These are built from IEEE
// operations that round identically on host and device, so host and device
// agree to the bit: exp by Cody-Waite reduction (ln2 split with an exact
// hi*k product) and the degree-15 Taylor polynomial of e^r, |r| <= ln2/2;
// log by frexp reduction to [sqrt(1/2), sqrt(2)) and the atanh series of
// log m in s = (m-1)/(m+1) to s^25; pow(x,y) = exp(y log x).
#define VM_LN2_HI 6.93147180369123816490e-01
#define VM_LN2_LO 1.90821492927058770002e-10
#define VM_INV_LN2 1.44269504088896338700e+00

__host__ __device__ __forceinline__ double dexp(double x)
{
  if (x != x) return x;
  if (x > 709.0) return HUGE_VAL;
  if (x < -745.0) return 0.0;
  double k = rint(x * VM_INV_LN2);
  double r = (x - k * VM_LN2_HI) - k * VM_LN2_LO;
  double p = 1.0 / 1307674368000.0;
  p = p * r + 1.0 / 87178291200.0;
  p = p * r + 1.0 / 6227020800.0;
  p = p * r + 1.0 / 479001600.0;
  p = p * r + 1.0 / 39916800.0;
  p = p * r + 1.0 / 3628800.0;
  p = p * r + 1.0 / 362880.0;
  p = p * r + 1.0 / 40320.0;
  p = p * r + 1.0 / 5040.0;
  p = p * r + 1.0 / 720.0;
  p = p * r + 1.0 / 120.0;
  p = p * r + 1.0 / 24.0;
  p = p * r + 1.0 / 6.0;
  p = p * r + 0.5;
  p = p * r + 1.0;
  p = p * r + 1.0;
  return ldexp(p, (int)k);
}

__host__ __device__ __forceinline__ double dlog(double x)
{
  if (x != x || x < 0.0) return NAN;
  if (x == 0.0) return -HUGE_VAL;
  if (x == HUGE_VAL) return x;
  int e;
  double m = frexp(x, &e);
  if (m < 0.70710678118654752440) {
    m = m * 2.0;
    e = e - 1;
  }
  double s = (m - 1.0) / (m + 1.0);
  double z = s * s;
  double p = 1.0 / 25.0;
  p = p * z + 1.0 / 23.0;
  p = p * z + 1.0 / 21.0;
  p = p * z + 1.0 / 19.0;
  p = p * z + 1.0 / 17.0;
  p = p * z + 1.0 / 15.0;
  p = p * z + 1.0 / 13.0;
  p = p * z + 1.0 / 11.0;
  p = p * z + 1.0 / 9.0;
  p = p * z + 1.0 / 7.0;
  p = p * z + 1.0 / 5.0;
  p = p * z + 1.0 / 3.0;
  double lm = 2.0 * s + 2.0 * s * (z * p);
  double de = (double)e;
  return de * VM_LN2_HI + (de * VM_LN2_LO + lm);
}

__host__ __device__ __forceinline__ double dpow(double x, double y) { return dexp(y * dlog(x)); }

// ---------------------------------------------------------------- parameters
// Tabulated s95p EOS: nodes e = 2^(emin+o) (1 + q/K), q = -1..K+1, in each
// of noct octaves, holding (p, dp/de, T, s); 4-point Lagrange interpolation
// inside an octave, closed form where a node stencil straddles a branch point
// of the piecewise EOS (flag array, one byte per interval).
// Element i of a device array reached through a parameter struct: a global
// load in device code (a pointer read from memory is generic to the compiler,
// which would emit a flat load), a plain load on the host.
template <class T>
__host__ __device__ __forceinline__ T gload(const T* p, long i)
{
#if defined(__HIP_DEVICE_COMPILE__)
  return ((const __attribute__((address_space(1))) T*)p)[i];
#else
  return p[i];
#endif
}

struct EosTab {
  const double* v;             // [noct][K+3][4]
  const unsigned char* closed; // [noct][K]
  int emin, noct, K;
};

struct Params {
  int milne, lbi, eos, eta_law, tpi_law, max_iter, rk, scheme;
  unsigned quirks;
  double dx, dy, deta;
  double eta_coef, tpi_coef;
  int bulk, zeta_law, tpib_law;  // bulk viscosity (viral_desc.bulk and its laws)
  double zeta_coef, tpib_coef;
  double factor;    // conformal FACTOR = 3*DOF*pi^2/90 (NEWEST10/maindef.h:10)
  double t_freeze;
  int use_tab;      // fast mode with the tabulated s95p EOS
  EosTab tab;
};

// ---------------------------------------------------------------- EOS
// s95p closed form, NEWEST10/s95p.h:13-123 (break points and fit coefficients
// are the reference's; strict comparisons keep its e == E_n fall-through).
#define S95_E1 0.5028563305441270
#define S95_E2 1.62
#define S95_E3 1.86
#define S95_E4 9.9878355786273545
#define S95_ES1 0.1270769021427449
#define S95_ES2 0.446707952467404
#define S95_ES3 1.9402832534193788
#define S95_ES4 3.7292474570977285

__host__ __device__ __forceinline__ double s95p_p(double e)
{
  if (e < 1e-5) return 0.143375 * e;
  double x = e * GEVFM, r;
  if (x < S95_E1)
    r = 0.3299 * (dexp(0.4346 * x) - 1);
  else if (x > S95_E1 && x < S95_E2)
    r = 1.024e-7 * dexp(6.041 * x) + 0.007273 + 0.14578 * x;
  else if (x > S95_E2 && x < S95_E3)
    r = 0.30195 * dexp(0.31308 * x) - 0.256232;
  else if (x > S95_E3 && x < S95_E4)
    r = 0.332 * x - 0.3223 * dpow(x, 0.4585) + 0.1167 * dpow(x, -1.233) - 0.003906 * x * dexp(-0.05697 * x) +
        0.1436 * x * dexp(-0.9131 * x);
  else
    r = 0.3327 * x - 0.3223 * dpow(x, 0.4585) - 0.003906 * x * dexp(-0.05697 * x);
  return r / GEVFM;
}

__host__ __device__ __forceinline__ double s95p_a(double e)
{
  if (e < 1e-5) return 0.143375;
  double x = e * GEVFM, r;
  if (x < S95_E1)
    r = 1 * (0.3299 * (0.4346 * dexp(0.4346 * x)));
  else if (x > S95_E1 && x < S95_E2)
    r = 1 * (1.024e-7 * 6.041 * dexp(6.041 * x) + 0.14578);
  else if (x > S95_E2 && x < S95_E3)
    r = 1 * (0.30195 * 0.31308 * dexp(0.31308 * x));
  else if (x > S95_E3 && x < S95_E4)
    r = 1 * (0.332 - 0.3223 * 0.4585 * dpow(x, (0.4585 - 1)) + 0.1167 * (-1.233) * dpow(x, (-1.233 - 1)) -
             0.003906 * (dexp(-0.05697 * x) - 0.05697 * x * dexp(-0.05697 * x)) +
             0.1436 * (dexp(-0.9131 * x) - 0.9131 * x * dexp(-0.9131 * x)));
  else
    r = 1 * (0.3327 - 0.3223 * 0.4585 * dpow(x, (0.4585 - 1)) -
             0.003906 * (dexp(-0.05697 * x) - 0.05697 * x * dexp(-0.05697 * x)));
  if ((x / GEVFM) < 0.12948 * 1e-14) r = 0.286743;
  return r;
}

__host__ __device__ __forceinline__ double s95p_s(double e)
{
  double x = e * GEVFM, r;
  if (x < S95_ES1)
    r = 12.2304 * dpow(x, 1.16849);
  else if (x > S95_ES1 && x < S95_ES2)
    r = 11.9279 * dpow(x, 1.15635);
  else if (x > S95_ES2 && x < S95_ES3)
    r = 0.0580578 + 11.833 * dpow(x, 1.16187);
  else if (x > S95_ES3 && x < S95_ES4)
    r = 18.202 * x - 62.021814 - 4.85479 * dexp(-2.72407e-11 * dpow(x, 4.54886)) +
        65.1272 * dpow(x, -0.128012) * dexp(-0.00369624 * dpow(x, 1.18735)) - 4.75253 * dpow(x, -1.18423);
  else
    r = 18.202 * x - 63.0218 - 4.85479 * dexp(-2.72407e-11 * dpow(x, 4.54886)) +
        65.1272 * dpow(x, -0.128012) * dexp(-0.00369624 * dpow(x, 1.18735));
  return dpow(r, 0.75);
}

__host__ __device__ __forceinline__ double s95p_T(double e)
{
  double x = e * GEVFM, r;
  if (x < 0.5143939846236409)
    r = 0.203054 * dpow(x, 0.30679);
  else
    r = (x + GEVFM * s95p_p(e)) / s95p_s(e);
  return r / GEVFM;
}

// e-derivatives of the s95p entropy-density and temperature fits, taken
// branch by branch (the derivative of the branch s95p_s / s95p_T evaluates
// at e; s95p_a is already that of s95p_p).  Used by the adjoint.
__host__ __device__ __forceinline__ double s95p_ds(double e)
{
  const double x = e * GEVFM;
  double r, dr;
  if (x < S95_ES1) {
    r = 12.2304 * dpow(x, 1.16849);
    dr = 12.2304 * 1.16849 * dpow(x, 0.16849);
  } else if (x > S95_ES1 && x < S95_ES2) {
    r = 11.9279 * dpow(x, 1.15635);
    dr = 11.9279 * 1.15635 * dpow(x, 0.15635);
  } else if (x > S95_ES2 && x < S95_ES3) {
    r = 0.0580578 + 11.833 * dpow(x, 1.16187);
    dr = 11.833 * 1.16187 * dpow(x, 0.16187);
  } else {
    const double a4 = dpow(x, 4.54886), g = dexp(-2.72407e-11 * a4);
    const double b1 = dpow(x, 1.18735), h = dexp(-0.00369624 * b1);
    const double c0 = dpow(x, -0.128012);
    const bool mid = x > S95_ES3 && x < S95_ES4;
    r = 18.202 * x - (mid ? 62.021814 : 63.0218) - 4.85479 * g + 65.1272 * c0 * h -
        (mid ? 4.75253 * dpow(x, -1.18423) : 0.0);
    dr = 18.202 + 4.85479 * 2.72407e-11 * 4.54886 * g * a4 / x +
         65.1272 * h * c0 * (-0.128012 / x - 0.00369624 * 1.18735 * b1 / x) +
         (mid ? 4.75253 * 1.18423 * dpow(x, -2.18423) : 0.0);
  }
  // s = r^(3/4)
  return 0.75 * dpow(r, -0.25) * dr * GEVFM;
}

__host__ __device__ __forceinline__ double s95p_dT(double e)
{
  const double x = e * GEVFM;
  if (x < 0.5143939846236409) return 0.203054 * 0.30679 * dpow(x, 0.30679 - 1.0);
  // T = (e + p(e)) / s(e)
  const double P = s95p_p(e), a = s95p_a(e), s = s95p_s(e), ds = s95p_ds(e);
  return ((1.0 + a) * s - (e + P) * ds) / (s * s);
}

// EOS(e), DPDE(e), FT(e), FS(e,P,T): NEWEST10/maindef.h:68-72 (conformal,
// with x^(1/4) as sqrt(sqrt(x))) and :274-278 (s95p)
__host__ __device__ __forceinline__ double eos_p(const Params& p, double e) { return p.eos ? s95p_p(e) : e / 3.0; }
__host__ __device__ __forceinline__ double eos_dpde(const Params& p, double e) { return p.eos ? s95p_a(e) : 1.0 / 3.0; }
__host__ __device__ __forceinline__ double eos_T(const Params& p, double e)
{
  return p.eos ? s95p_T(e) : sqrt(sqrt(e / p.factor));
}
__host__ __device__ __forceinline__ double eos_s(const Params& p, double e, double P, double T)
{
  return p.eos ? s95p_s(e) : (e + P) * (1.0 / T);
}
// Interpolated (p, dp/de, T, s) at e from the table; false where the table
// does not apply (outside it, or an interval next to a branch point).
__host__ __device__ __forceinline__ bool eos_tab_lookup(const EosTab& t, double e, double out[4])
{
  if (!(e >= 1e-5)) return false;
  int E;
  const double m = frexp(e, &E);  // e = m 2^E, m in [0.5, 1)
  const int oct = E - 1 - t.emin;
  if (oct < 0 || oct >= t.noct) return false;
  const double u = (2.0 * m - 1.0) * t.K;
  int q = (int)u;
  if (q > t.K - 1) q = t.K - 1;
  // The interval's flag and its four points are loaded together and the flag
  // is tested after the interpolation (a closed interval's points are in the
  // table too), so a lookup costs one memory round trip rather than two --
  // the Newton iteration chains one lookup per iterate.
  const bool closed = gload(t.closed, (long)oct * t.K + q) != 0;
  const double x = u - q;
  const double xm = x - 1.0, xp = x + 1.0, x2 = x - 2.0;
  const double w0 = -(x * xm * x2) * (1.0 / 6.0);
  const double w1 = (xp * xm * x2) * 0.5;
  const double w2 = -(xp * x * x2) * 0.5;
  const double w3 = (xp * x * xm) * (1.0 / 6.0);
  const double* b = t.v + ((long)oct * (t.K + 3) + q) * 4;
#pragma unroll
  for (int c = 0; c < 4; c++)
    out[c] = fma(w0, gload(b, c), fma(w1, gload(b, 4 + c), fma(w2, gload(b, 8 + c), w3 * gload(b, 12 + c))));
  return !closed;
}

// e-derivatives of the interpolated (p, dp/de, T, s), for the adjoint: the
// derivatives of the four cubic Lagrange weights times du/de = 2K 2^-E.
__host__ __device__ __forceinline__ bool eos_tab_lookup_d(const EosTab& t, double e, double out[4], double dout[4])
{
  if (!eos_tab_lookup(t, e, out)) return false;
  int E;
  const double m = frexp(e, &E);
  const int oct = E - 1 - t.emin;
  const double u = (2.0 * m - 1.0) * t.K;
  int q = (int)u;
  if (q > t.K - 1) q = t.K - 1;
  const double x = u - q, x2 = x * x;
  const double d0 = -(3.0 * x2 - 6.0 * x + 2.0) * (1.0 / 6.0);
  const double d1 = (3.0 * x2 - 4.0 * x - 1.0) * 0.5;
  const double d2 = -(3.0 * x2 - 2.0 * x - 2.0) * 0.5;
  const double d3 = (3.0 * x2 - 1.0) * (1.0 / 6.0);
  const double dude = ldexp(2.0 * t.K, -E);
  const double* b = t.v + ((long)oct * (t.K + 3) + q) * 4;
#pragma unroll
  for (int c = 0; c < 4; c++)
    dout[c] = (d0 * gload(b, c) + d1 * gload(b, 4 + c) + d2 * gload(b, 8 + c) + d3 * gload(b, 12 + c)) * dude;
  return true;
}

// closed-form (p, dp/de, T, s)
__host__ __device__ __forceinline__ void s95p_all(double e, double out[4])
{
  out[0] = s95p_p(e);
  out[1] = s95p_a(e);
  out[2] = s95p_T(e);
  out[3] = s95p_s(e);
}

// Feta / Ftaupi laws, NEWEST10/maindef.h:76-77,187-188,282-283,328-329
__host__ __device__ __forceinline__ double eta_of(const Params& p, double s)
{
  return p.eta_law == 1 ? p.eta_coef * (s / (4.0 * PIE)) : p.eta_coef * s;
}
__host__ __device__ __forceinline__ double tpi_of(const Params& p, double eta, double P, double e)
{
  if (p.tpi_law == 1) return p.tpi_coef * (eta / P);
  if (p.tpi_law == 2) return p.tpi_coef;
  return p.tpi_coef * eta / (e + P);
}
// FZeta / FtauPI laws of the -DBULK build, NEWEST10/maindef.h:82-83,140-141,191-192,289-290
__host__ __device__ __forceinline__ double zeta_of(const Params& p, double s)
{
  return p.zeta_law == 1 ? p.zeta_coef * (s / (4.0 * PIE)) : p.zeta_coef;
}
__host__ __device__ __forceinline__ double tpib_of(const Params& p, double zeta, double P)
{
  return p.tpib_law == 1 ? p.tpib_coef * (zeta / P) : p.tpib_coef;
}

// ---------------------------------------------------------------- WENO5-JS
// NEWEST10/utils.h:102-160: L is the value at the right face of the centre
// cell, R at its left face; d = (0.3, 0.6, 0.1), eps = 1e-6, p = 2.
__host__ __device__ __forceinline__ void weno_beta(double qm2, double qm1, double qc, double qp1, double qp2, double& b0,
                                          double& b1, double& b2)
{
  double a0 = qc - 2 * qp1 + qp2, c0 = 3 * qc - 4 * qp1 + qp2;
  double a1 = qm1 - 2 * qc + qp1, c1 = qm1 - qp1;
  double a2 = qm2 - 2 * qm1 + qc, c2 = qm2 - 4 * qm1 + 3 * qc;
  b0 = (13. / 12.) * (a0 * a0) + 0.25 * (c0 * c0);
  b1 = (13. / 12.) * (a1 * a1) + 0.25 * (c1 * c1);
  b2 = (13. / 12.) * (a2 * a2) + 0.25 * (c2 * c2);
}

__host__ __device__ __forceinline__ double weno_L(double qm2, double qm1, double qc, double qp1, double qp2)
{
  double b0, b1, b2;
  weno_beta(qm2, qm1, qc, qp1, qp2, b0, b1, b2);
  double q0 = (2 * qc + 5 * qp1 - 1 * qp2) / 6.0;
  double q1 = (-1 * qm1 + 5 * qc + 2 * qp1) / 6.0;
  double q2 = (2 * qm2 - 7 * qm1 + 11 * qc) / 6.0;
  double t0 = WENO_EPS + b0, t1 = WENO_EPS + b1, t2 = WENO_EPS + b2;
  double al0 = 0.3 / (t0 * t0), al1 = 0.6 / (t1 * t1), al2 = 0.1 / (t2 * t2);
  double sum = 0;
  sum += al0;
  sum += al1;
  sum += al2;
  double left = 0;
  left += (al0 / sum) * q0;
  left += (al1 / sum) * q1;
  left += (al2 / sum) * q2;
  return left;
}

__host__ __device__ __forceinline__ double weno_R(double qm2, double qm1, double qc, double qp1, double qp2)
{
  double b0, b1, b2;
  weno_beta(qm2, qm1, qc, qp1, qp2, b0, b1, b2);
  double q0 = (11 * qc - 7 * qp1 + 2 * qp2) / 6.0;
  double q1 = (2 * qm1 + 5 * qc - 1 * qp1) / 6.0;
  double q2 = (-1 * qm2 + 5 * qm1 + 2 * qc) / 6.0;
  double t0 = WENO_EPS + b0, t1 = WENO_EPS + b1, t2 = WENO_EPS + b2;
  double al0 = 0.1 / (t0 * t0), al1 = 0.6 / (t1 * t1), al2 = 0.3 / (t2 * t2);
  double sum = 0;
  sum += al0;
  sum += al1;
  sum += al2;
  double right = 0;
  right += (al0 / sum) * q0;
  right += (al1 / sum) * q1;
  right += (al2 / sum) * q2;
  return right;
}

// NEWEST10/utils.h:57-98
__host__ __device__ __forceinline__ double weno_der(double qm2, double qm1, double qc, double qp1, double qp2, double dx)
{
  double left = weno_L(qm2, qm1, qc, qp1, qp2);
  double right = weno_R(qm2, qm1, qc, qp1, qp2);
  return (left - right) / dx;
}

// ---------------------------------------------------------------- shear source
// Israel-Stewart relaxation of pi^{mu nu} in conservative form (the physics of
// NEWEST10/Source.h:448-516):
//   u.d pi = -(pi - 2 eta sigma)/tau_pi - 4/3 pi theta - (u^m pi^{na} + u^n pi^{ma}) u.d u_a,
//   sigma^{mn} = Delta^{mn}_{ab} d^a u^b (double projector, du[a][b] = d_b u^a
//   taken as independent), S = (u.d pi - Christoffel part)/u^0 + pi d_i(u^i/u^0).
// Operation order matches the CPU restatement exactly.
__host__ __device__ __forceinline__ int pidx(int m, int n)
{
  // {{0,1,2,3},{1,4,5,6},{2,5,7,8},{3,6,8,9}}
  const int a = m < n ? m : n, b = m < n ? n : m;
  return a == 0 ? b : a == 1 ? 3 + b : a == 2 ? 5 + b : 9;
}

template <bool FAST = false>
__host__ __device__ __forceinline__ void shear_source(int milne, double tau, const double u[4], const double pi10[10],
                                             const double du[4][4], double SH, double tpi, double out[10]);

// ---------------------------------------------------------------- root solve
// f(e), f'(e) of NEWEST10/root.h:212-285 and the GSL Newton iteration
// (x_{n+1} = x_n - f/f', stop on two equal iterates, error on f' == 0 or a
// non-finite f or f'), NEWEST10/root.h:301-364.
__host__ __device__ __forceinline__ double newton_cell(const Params& p, double M2, double Ttt, double pi1, double PI,
                                              double guess, int& iters, unsigned& flags)
{
  double x = guess, fv, dfv;
  {
    double Pe = eos_p(p, x);
    fv = x - (Ttt - pi1) + (M2 / ((Ttt - pi1) + Pe - PI));
    double den = Ttt - pi1 + Pe - PI;
    dfv = 1 - ((M2 / (den * den)) * eos_dpde(p, x));
  }
  int it = 0;
  unsigned fl = 0;
  double h1 = x, h2 = x, h3 = x, h4 = x;  // x_{it-1} .. x_{it-4}
  for (;;) {
    it++;
    if (dfv == 0.0) {
      fl |= 4u;
      break;
    }
    double xn = x - (fv / dfv);
    double Pe = eos_p(p, xn);
    fv = xn - (Ttt - pi1) + (M2 / ((Ttt - pi1) + Pe - PI));
    double den = Ttt - pi1 + Pe - PI;
    dfv = 1 - ((M2 / (den * den)) * eos_dpde(p, xn));
    double xp = x;
    x = xn;
    if (!isfinite(fv) || !isfinite(dfv)) {
      fl |= 4u;
      break;
    }
    if (x == xp) break;
    if (it >= p.max_iter) {
      fl |= 8u;
      break;
    }
    h4 = h3;
    h3 = h2;
    h2 = h1;
    h1 = xp;
    // The Newton map depends on x alone: x_it == x_{it-p} means the iterates
    // repeat with period p until max_iter, so jump to x_{max_iter} =
    // x_{it-p+r}, r = (max_iter - it) mod p.  GSL's stopping rule (two equal
    // iterates) leaves many cells in a period-2 orbit between neighbouring
    // doubles; running those loops out would stall every wavefront.
    int per = (it >= 2 && x == h2) ? 2 : (it >= 3 && x == h3) ? 3 : (it >= 4 && x == h4) ? 4 : 0;
    if (per) {
      int r = (p.max_iter - it) % per;
      int q = per - r;
      x = q == 1 ? h1 : q == 2 ? h2 : q == 3 ? h3 : h4;
      it = p.max_iter;
      fl |= 8u;
      break;
    }
  }
  iters = it;
  flags = fl;
  return x;
}


// ---------------------------------------------------------------- fast forms
// 1/x from the hardware estimate (v_rcp_f64) refined by two Newton-Raphson
// steps r <- r + r(1 - x r): within an ulp or two of the IEEE quotient for
// the normal, finite denominators the fast path feeds it.
__host__ __device__ __forceinline__ double frcp(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
#else
  return 1.0 / x;  // host builds (the adjoint checks) use the IEEE quotient
#endif
}

// 1/x with one Newton-Raphson step: v_rcp_f64 is good to about 2^-24 and the
// step squares that, to within about ten ulp of the quotient (2.2e-15 relative
// at most over 4M random normal x of both signs and exponents -300..300).  For
// the WENO weights' normalisation, where a few ulp do not matter.
__host__ __device__ __forceinline__ double frcp1(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
  const double r = __builtin_amdgcn_rcp(x);
  return fma(r, fma(-x, r, 1.0), r);
#else
  return 1.0 / x;  // host builds (the adjoint checks) use the IEEE quotient
#endif
}

// Same WENO5-JS weights written with one division per reconstruction:
//   w_c = d_c / t_c / sum  ==  d_c P_c / sum_c d_c P_c,  t_c = (eps+b_c)^2,
//   P_0 = t_1 t_2, P_1 = t_0 t_2, P_2 = t_0 t_1,
// with the weights' common factors taken out (d = (3, 6, 1)/10 for L and
// (1, 6, 3)/10 for R; the t_c scaled by 4, which leaves the normalised
// weights unchanged), the 1/6 of the candidate polynomials folded into the
// denominator, L and R sharing the smoothness indicators and one reciprocal,
// and the candidates built from the indicators' differences where that saves
// an operation (r1 = l1 + 3 c1, r0 = 2 c0 + 5 q + q+1, l2 = 2 c2 + 5 q + q-1).
// 59 FP64 operations.  Agrees with weno_L/weno_R to about ten ulp (tests hold
// fast mode to 1e-10 relative per step).
__device__ __forceinline__ void weno_LR_fast(double qm2, double qm1, double qc, double qp1, double qp2, double& L,
                                             double& R)
{
  const double a0 = fma(-2.0, qp1, qc) + qp2, c0 = fma(3.0, qc, fma(-4.0, qp1, qp2));
  const double a1 = fma(-2.0, qc, qm1) + qp1, c1 = qm1 - qp1;
  const double a2 = fma(-2.0, qm1, qm2) + qc, c2 = fma(3.0, qc, fma(-4.0, qm1, qm2));
  // 4 (eps + b_c) = 4 eps + (13/3) a_c^2 + c_c^2
  const double k4 = 13.0 / 3.0, e4 = 4.0 * WENO_EPS;
  double t0 = fma(k4 * a0, a0, fma(c0, c0, e4));
  double t1 = fma(k4 * a1, a1, fma(c1, c1, e4));
  double t2 = fma(k4 * a2, a2, fma(c2, c2, e4));
  t0 = t0 * t0;
  t1 = t1 * t1;
  t2 = t2 * t2;
  const double P0 = t1 * t2, P1 = t0 * t2, P2 = t0 * t1;
  // candidate values times 6
  const double l0 = fma(2.0, qc, fma(5.0, qp1, -qp2));
  const double l1 = fma(5.0, qc, fma(2.0, qp1, -qm1));
  const double l2 = fma(2.0, c2, fma(5.0, qc, qm1));
  const double r0 = fma(2.0, c0, fma(5.0, qc, qp1));
  const double r1 = fma(3.0, c1, l1);
  const double r2 = fma(5.0, qm1, fma(2.0, qc, -qm2));
  // L = (3 P0 l0 + 6 P1 l1 + P2 l2) / (6 Dl),  R = (P0 r0 + 6 P1 r1 + 3 P2 r2) / (6 Dr)
  const double A0 = 3.0 * P0, B1 = 6.0 * P1, C2 = 3.0 * P2;
  const double Dl = A0 + B1 + P2, Dr = P0 + B1 + C2;
  // one reciprocal for both: 1/(6 Dl) = Dr/(6 Dl Dr), 1/(6 Dr) = Dl/(6 Dl Dr)
  const double iDD = frcp1(6.0 * Dl * Dr);
  L = fma(A0, l0, fma(B1, l1, P2 * l2)) * (Dr * iDD);
  R = fma(P0, r0, fma(B1, r1, C2 * r2)) * (Dl * iDD);
}

__device__ __forceinline__ double weno_L_fast(double qm2, double qm1, double qc, double qp1, double qp2)
{
  const double a0 = fma(-2.0, qp1, qc) + qp2, c0 = fma(3.0, qc, fma(-4.0, qp1, qp2));
  const double a1 = fma(-2.0, qc, qm1) + qp1, c1 = qm1 - qp1;
  const double a2 = fma(-2.0, qm1, qm2) + qc, c2 = fma(3.0, qc, fma(-4.0, qm1, qm2));
  const double k = 13.0 / 12.0;
  double t0 = WENO_EPS + fma(k * a0, a0, 0.25 * (c0 * c0));
  double t1 = WENO_EPS + fma(k * a1, a1, 0.25 * (c1 * c1));
  double t2 = WENO_EPS + fma(k * a2, a2, 0.25 * (c2 * c2));
  t0 = t0 * t0;
  t1 = t1 * t1;
  t2 = t2 * t2;
  const double wl0 = 0.3 * (t1 * t2), wl1 = 0.6 * (t0 * t2), wl2 = 0.1 * (t0 * t1);
  const double l0 = fma(2.0, qc, fma(5.0, qp1, -qp2));
  const double l1 = fma(5.0, qc, fma(2.0, qp1, -qm1));
  const double l2 = fma(2.0, qm2, fma(-7.0, qm1, 11.0 * qc));
  return fma(wl0, l0, fma(wl1, l1, wl2 * l2)) * frcp(6.0 * (wl0 + wl1 + wl2));
}

__device__ __forceinline__ double weno_R_fast(double qm2, double qm1, double qc, double qp1, double qp2)
{
  const double a0 = fma(-2.0, qp1, qc) + qp2, c0 = fma(3.0, qc, fma(-4.0, qp1, qp2));
  const double a1 = fma(-2.0, qc, qm1) + qp1, c1 = qm1 - qp1;
  const double a2 = fma(-2.0, qm1, qm2) + qc, c2 = fma(3.0, qc, fma(-4.0, qm1, qm2));
  const double k = 13.0 / 12.0;
  double t0 = WENO_EPS + fma(k * a0, a0, 0.25 * (c0 * c0));
  double t1 = WENO_EPS + fma(k * a1, a1, 0.25 * (c1 * c1));
  double t2 = WENO_EPS + fma(k * a2, a2, 0.25 * (c2 * c2));
  t0 = t0 * t0;
  t1 = t1 * t1;
  t2 = t2 * t2;
  const double wr0 = 0.1 * (t1 * t2), wr1 = 0.6 * (t0 * t2), wr2 = 0.3 * (t0 * t1);
  const double r0 = fma(11.0, qc, fma(-7.0, qp1, 2.0 * qp2));
  const double r1 = fma(2.0, qm1, fma(5.0, qc, -qp1));
  const double r2 = fma(5.0, qm1, fma(2.0, qc, -qm2));
  return fma(wr0, r0, fma(wr1, r1, wr2 * r2)) * frcp(6.0 * (wr0 + wr1 + wr2));
}

template <bool FAST>
__device__ __forceinline__ double wL(double a, double b, double c, double d, double e)
{
  return FAST ? weno_L_fast(a, b, c, d, e) : weno_L(a, b, c, d, e);
}
template <bool FAST>
__device__ __forceinline__ double wR(double a, double b, double c, double d, double e)
{
  return FAST ? weno_R_fast(a, b, c, d, e) : weno_R(a, b, c, d, e);
}
template <bool FAST>
__device__ __forceinline__ double wder(double a, double b, double c, double d, double e, double h, double ih)
{
  if (FAST) {
    double L, R;
    weno_LR_fast(a, b, c, d, e, L, R);
    return (L - R) * ih;
  }
  return weno_der(a, b, c, d, e, h);
}


// Fast-mode Newton: same iteration, stopping rule and period shortcut as
// newton_cell, with reciprocal forms and, for s95p with the table, the
// interpolated p(e) and dp/de.
__device__ __forceinline__ void eos_pa_fast(const Params& p, double e, double& P, double& a)
{
  if (!p.eos) {
    P = e * (1.0 / 3.0);
    a = 1.0 / 3.0;
    return;
  }
  double v[4];
  if (p.use_tab && eos_tab_lookup(p.tab, e, v)) {
    P = v[0];
    a = v[1];
    return;
  }
  P = s95p_p(e);
  a = s95p_a(e);
}

// Safeguard of the root solve (no counterpart in the reference, which aborts
// instead): the root of f(e) = e - A + M2/(A + P(e) - PI) on [0, A] by the
// safeguarded Newton method (a Newton step while it stays inside the bracket
// and shrinks |f| fast enough, a bisection of the bracket otherwise; the
// bracket follows the sign of f).  f(0) < 0 <= f(A) holds exactly when
// M2 < A (A - PI) for P(0) = 0, i.e. for a causal cell; false when no bracket
// exists.  Stops when the step or the bracket falls below 4 ulp of e, or
// after 200 iterations.
__host__ __device__ inline bool root_bracketed(const Params& p, double M2, double A, double PI, double& root,
                                               int& iters)
{
  double lo = 0.0, hi = A, fx, dfx;
  auto F = [&](double x) {
    const double Q = A + eos_p(p, x) - PI;
    fx = x - A + M2 / Q;
    dfx = 1.0 - M2 * eos_dpde(p, x) / (Q * Q);
  };
  if (!(A > 0.0)) return false;
  F(lo);
  if (!(fx < 0.0)) return false;
  F(hi);
  if (!(fx >= 0.0)) return false;
  if (fx == 0.0) {
    root = hi;
    iters = 0;
    return true;
  }
  double x = 0.5 * (lo + hi), dx = hi - lo, dxold = dx;
  F(x);
  for (int it = 1; it <= 200; it++) {
    const bool leaves = ((x - hi) * dfx - fx) * ((x - lo) * dfx - fx) > 0.0;
    if (leaves || !(fabs(2.0 * fx) <= fabs(dxold * dfx))) {
      dxold = dx;
      dx = 0.5 * (hi - lo);
      x = lo + dx;
    } else {
      dxold = dx;
      dx = fx / dfx;
      x = x - dx;
    }
    F(x);
    if (fx < 0.0) lo = x;
    else hi = x;
    if (fx == 0.0 || fabs(dx) <= 4e-16 * fabs(x) || hi - lo <= 4e-16 * hi) {
      root = x;
      iters = it;
      return true;
    }
  }
  root = x;
  iters = 200;
  return true;
}

// eos_pa_fast that keeps the table's (p, dp/de, T, s) at e in tv (and sets
// have) where the table applies.
__device__ __forceinline__ void eos_pa_keep(const Params& p, double e, double& P, double& a, double tv[4], bool& have)
{
  if (p.eos && p.use_tab && eos_tab_lookup(p.tab, e, tv)) {
    P = tv[0];
    a = tv[1];
    have = true;
    return;
  }
  have = false;
  eos_pa_fast(p, e, P, a);
}

// The fast Newton iteration (see eos_pa_fast).  tv and have_tv: the table's
// values at the returned root when the last evaluation was at that root (every
// exit but the period shortcut), so the caller's P and T there need no second
// lookup.
__device__ __forceinline__ double newton_cell_fast(const Params& p, double M2, double Ttt, double pi1, double PI,
                                                   double guess, int& iters, unsigned& flags, double tv[4],
                                                   bool& have_tv)
{
  const double A = Ttt - pi1;
  double x = guess, fv, dfv;
  {
    double Pe, a;
    eos_pa_keep(p, x, Pe, a, tv, have_tv);
    const double r = frcp(A + Pe - PI);
    fv = x - A + M2 * r;
    dfv = 1 - (M2 * r * r) * a;
  }
  int it = 0;
  unsigned fl = 0;
  double h1 = x, h2 = x, h3 = x, h4 = x;
  for (;;) {
    it++;
    if (dfv == 0.0) {
      fl |= 4u;
      break;
    }
    double xn = x - fv * frcp(dfv);
    double Pe, a;
    eos_pa_keep(p, xn, Pe, a, tv, have_tv);
    const double r = frcp(A + Pe - PI);
    fv = xn - A + M2 * r;
    dfv = 1 - (M2 * r * r) * a;
    double xp = x;
    x = xn;
    if (!isfinite(fv) || !isfinite(dfv)) {
      fl |= 4u;
      break;
    }
    if (x == xp) break;
    if (it >= p.max_iter) {
      fl |= 8u;
      break;
    }
    h4 = h3;
    h3 = h2;
    h2 = h1;
    h1 = xp;
    int per = (it >= 2 && x == h2) ? 2 : (it >= 3 && x == h3) ? 3 : (it >= 4 && x == h4) ? 4 : 0;
    if (per) {
      int rr = (p.max_iter - it) % per;
      int q = per - rr;
      x = q == 1 ? h1 : q == 2 ? h2 : q == 3 ? h3 : h4;
      it = p.max_iter;
      fl |= 8u;
      have_tv = false;
      break;
    }
  }
  iters = it;
  flags = fl;
  return x;
}

// Fast-mode T(e) and s(e) (and P when asked) for the sources and the output.
__device__ __forceinline__ void eos_Ts_fast(const Params& p, double e, double P, double& T, double& s)
{
  if (!p.eos) {
    T = sqrt(sqrt(e * (1.0 / p.factor)));
    s = (e + P) * frcp(T);
    return;
  }
  double v[4];
  if (p.use_tab && eos_tab_lookup(p.tab, e, v)) {
    T = v[2];
    s = v[3];
    return;
  }
  T = s95p_T(e);
  s = s95p_s(e);
}


// Israel-Stewart shear source, register-light form of the double projector:
//   Delta^{mn}_{ab} G^{ab} = G^{mn} - u^n (G u_)^m - u^m (u_ G)^n + u^m u^n (u_ G u_),
//   Delta_{ab} G^{ab} = sum_a g_aa G^{aa} - u_ G u_,   G^{ab} = g^{aa} d_a u^b.
// Exact mode follows the CPU restatement's operation order; FAST replaces
// the divisions by reciprocals.
template <bool FAST>
__host__ __device__ __forceinline__ void shear_source(int milne, double tau, const double u[4], const double pi10[10],
                                             const double du[4][4], double SH, double tpi, double out[10])
{
  const double g[4] = {1.0, -1.0, -1.0, milne ? -(tau * tau) : -1.0};
  const double gi[4] = {1.0, -1.0, -1.0, milne ? -(1.0 / (tau * tau)) : -1.0};
  double ul[4];
#pragma unroll
  for (int m = 0; m < 4; m++) ul[m] = g[m] * u[m];
  const double iu0 = FAST ? frcp(u[0]) : 1.0 / u[0];
  double divv = 0;
#pragma unroll
  for (int i = 1; i < 4; i++) divv += du[i][i] * iu0 - u[i] * du[0][i] * (iu0 * iu0);
  // Dc = du with the Christoffel part and G^{ab} = g^{aa} Dc[b][a], evaluated
  // where used (the same values each time; the loops are unrolled, so the
  // index tests fold away and no 4x4 copies stay live)
  const double itau = milne ? 1.0 / tau : 0.0;
  auto Dc = [&](int a, int b) -> double {
    double v = du[a][b];
    if (milne) {
      if (a == 0 && b == 3) v = v + tau * u[3];
      if (a == 3 && b == 0) v = v + itau * u[3];
      if (a == 3 && b == 3) v = v + itau * u[0];
    }
    return v;
  };
  auto G = [&](int a, int b) -> double { return gi[a] * Dc(b, a); };
  const double theta = Dc(0, 0) + Dc(1, 1) + Dc(2, 2) + Dc(3, 3);
  double Dul[4];
#pragma unroll
  for (int a = 0; a < 4; a++) {
    double s = 0;
#pragma unroll
    for (int b = 0; b < 4; b++) s += u[b] * Dc(a, b);
    Dul[a] = g[a] * s;
  }
  double Gu[4], uG[4], H[4];
#pragma unroll
  for (int m = 0; m < 4; m++) {
    double s1 = 0, s2 = 0;
#pragma unroll
    for (int b = 0; b < 4; b++) {
      s1 += G(m, b) * ul[b];
      s2 += ul[b] * G(b, m);
    }
    Gu[m] = s1;
    uG[m] = s2;
  }
  double uGu = 0, trg = 0;
#pragma unroll
  for (int a = 0; a < 4; a++) {
    uGu += ul[a] * Gu[a];
    trg += g[a] * G(a, a);
  }
  const double tr = trg - uGu;
#pragma unroll
  for (int m = 0; m < 4; m++) H[m] = Gu[m] + uG[m];
  const double c3 = uGu + (1.0 / 3.0) * tr;
  double PD[4];
#pragma unroll
  for (int m = 0; m < 4; m++) {
    double s = 0;
#pragma unroll
    for (int a = 0; a < 4; a++) s += pi10[pidx(m, a)] * Dul[a];
    PD[m] = s;
  }
  const double itpi = FAST ? frcp(tpi) : 0.0;
#pragma unroll
  for (int m = 0; m < 4; m++)
#pragma unroll
    for (int n = m; n < 4; n++) {
      const double pmn = pi10[pidx(m, n)];
      double sigma = 0.5 * (G(m, n) + G(n, m)) - 0.5 * (u[n] * H[m] + u[m] * H[n]) + u[m] * u[n] * c3;
      if (m == n) sigma = sigma - (1.0 / 3.0) * gi[m] * tr;
      const double pu = u[m] * PD[n] + u[n] * PD[m];
      const double rel = FAST ? -(pmn - 2.0 * SH * sigma) * itpi : -(pmn - 2.0 * SH * sigma) / tpi;
      const double rhs = rel - (4.0 / 3.0) * pmn * theta - pu;
      double chr = 0;
      if (milne) {
        double cm = 0, cn = 0;
        if (m == 0) cm = u[3] * tau * pi10[pidx(3, n)];
        if (m == 3) cm = u[0] * itau * pi10[pidx(3, n)] + u[3] * itau * pi10[pidx(0, n)];
        if (n == 0) cn = u[3] * tau * pi10[pidx(m, 3)];
        if (n == 3) cn = u[0] * itau * pi10[pidx(m, 3)] + u[3] * itau * pi10[pidx(m, 0)];
        chr = cm + cn;
      }
      out[pidx(m, n)] = (rhs - chr) * iu0 + pmn * divv;
    }
}

// ---------------------------------------------------------------- Cartesian (TXYZ5)
// Full pi^{mu nu} (slots tt,tx,ty,tz,xx,xy,xz,yy,yz,zz) from the five evolved
// components {xx, yy, xy, xz, yz} (TXYZ5/maindef.h:240) by u_mu pi^{mu nu} = 0
// and pi^mu_mu = 0 -- the A1..A5 of NEWEST5/maindef.h:497-501 at tau = 1.
// Same operation order as the CPU restatement.
__host__ __device__ __forceinline__ void cart_pi10(const double u[4], double p1, double p2, double p3, double p4, double p5,
                                          double pi10[10])
{
  const double u0 = u[0], u1 = u[1], u2 = u[2], u3 = u[3];
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

// slot of the full tensor holding evolved Cartesian component l
__host__ __device__ __forceinline__ int cart5(int l) { return l == 0 ? 4 : l == 1 ? 7 : l == 2 ? 5 : l == 3 ? 6 : 8; }

// conservation_f / conservation_df of TXYZ5/multiroot.h:203-319 (residuals of
// T^{t mu} in the unknowns e, u^x, u^y, u^z), prm = {T^tt, T^tx, T^ty, T^tz,
// p1..p5, PI}.
__host__ __device__ __forceinline__ void cart_f(const Params& p, const double prm[10], const double x[4], double y[4])
{
  const double W = prm[0], X = prm[1], Y = prm[2], Z = prm[3];
  const double p1 = prm[4], p2 = prm[5], p3 = prm[6], p4 = prm[7], p5 = prm[8], PI = prm[9];
  const double e = x[0], u1 = x[1], u2 = x[2], u3 = x[3];
  const double P = eos_p(p, e);
  const double q = 1 + u1 * u1 + u2 * u2 + u3 * u3;
  const double sq = sqrt(q);
  y[0] = -(u1 * X) - u2 * Y - u3 * Z + (-e + W) * sq;
  y[1] = -(p3 * u2) - p4 * u3 - P * u1 * q - u1 * (p1 + e * q - PI * q) + X * sq;
  y[2] = -(p3 * u1) - (e + p2 - PI) * u2 - p5 * u3 - (e - PI) * u2 * (u1 * u1 + u2 * u2 + u3 * u3) - P * u2 * q + Y * sq;
  y[3] = -(p4 * u1) - p5 * u2 - P * u3 + u3 * (p1 + p2 + PI - W) + Z * sq;
}

__host__ __device__ __forceinline__ void cart_df(const Params& p, const double prm[10], const double x[4], double J[4][4])
{
  const double W = prm[0], X = prm[1], Y = prm[2], Z = prm[3];
  const double p1 = prm[4], p2 = prm[5], p3 = prm[6], p4 = prm[7], p5 = prm[8], PI = prm[9];
  const double e = x[0], u1 = x[1], u2 = x[2], u3 = x[3];
  const double P = eos_p(p, e);
  const double a = eos_dpde(p, e);
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

__host__ __device__ __forceinline__ double enorm4(const double f[4])
{
  double e2 = 0;
#pragma unroll
  for (int i = 0; i < 4; i++) e2 += f[i] * f[i];
  return sqrt(e2);
}

// GSL gnewton as driven by TXYZ5/multiroot.h:385-418 (max 100 iterations,
// test_delta(dx, x, 0, 1e-30)): LU with partial pivoting (first largest
// pivot; the row swaps are applied to the right-hand side as they happen,
// which equals GSL's permutation of it), forward/back substitution, and the
// step halving rule t *= (sqrt(1+6 theta)-1)/(3 theta) while the residual
// norm rises and t > DBL_EPSILON.  All indices are compile-time (no scratch).
// A zero pivot or a non-finite root sets the solver flag (4).
__host__ __device__ __forceinline__ int cart_gnewton(const Params& p, const double prm[10], double x[4], unsigned& flags)
{
  double f[4], J[4][4];
  unsigned fl = 0;
  cart_f(p, prm, x, f);
  cart_df(p, prm, x, J);
  double phi = enorm4(f);
  int iter = 0;
  for (;;) {
    iter++;
    double LU[4][4], b[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      b[i] = f[i];
#pragma unroll
      for (int k = 0; k < 4; k++) LU[i][k] = J[i][k];
    }
#pragma unroll
    for (int j = 0; j < 3; j++) {
      double mx = fabs(LU[j][j]);
      int ip = j;
#pragma unroll
      for (int i = j + 1; i < 4; i++) {
        const double aij = fabs(LU[i][j]);
        if (aij > mx) {
          mx = aij;
          ip = i;
        }
      }
#pragma unroll
      for (int i = j + 1; i < 4; i++) {
        if (ip == i) {
#pragma unroll
          for (int k = 0; k < 4; k++) {
            const double t = LU[j][k];
            LU[j][k] = LU[i][k];
            LU[i][k] = t;
          }
          const double t = b[j];
          b[j] = b[i];
          b[i] = t;
        }
      }
      const double ajj = LU[j][j];
      if (ajj != 0.0) {
#pragma unroll
        for (int i = j + 1; i < 4; i++) {
          const double aij = LU[i][j] / ajj;
          LU[i][j] = aij;
#pragma unroll
          for (int k = j + 1; k < 4; k++) LU[i][k] = LU[i][k] - aij * LU[j][k];
        }
      }
    }
    if (LU[0][0] == 0.0 || LU[1][1] == 0.0 || LU[2][2] == 0.0 || LU[3][3] == 0.0) {
      fl |= 4u;
      break;
    }
#pragma unroll
    for (int i = 1; i < 4; i++) {
      double t = b[i];
#pragma unroll
      for (int j = 0; j < i; j++) t -= LU[i][j] * b[j];
      b[i] = t;
    }
    b[3] = b[3] / LU[3][3];
#pragma unroll
    for (int i = 2; i >= 0; i--) {
      double t = b[i];
#pragma unroll
      for (int j = i + 1; j < 4; j++) t -= LU[i][j] * b[j];
      b[i] = t / LU[i][i];
    }
    double t = 1, phi0 = phi, phi1, xt[4];
    for (;;) {
#pragma unroll
      for (int i = 0; i < 4; i++) xt[i] = x[i] - t * b[i];
      cart_f(p, prm, xt, f);
      phi1 = enorm4(f);
      if (phi1 > phi0 && t > 2.220446049250313080847e-16) {
        const double theta = phi1 / phi0;
        const double u = (sqrt(1.0 + 6.0 * theta) - 1.0) / (3.0 * theta);
        t *= u;
        continue;
      }
      break;
    }
    bool ok = true;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      x[i] = xt[i];
      const double dxi = -t * b[i];
      const double tol = 1e-30 * fabs(x[i]);
      ok = ok && (fabs(dxi) < tol || dxi == 0);
    }
    cart_df(p, prm, x, J);
    phi = phi1;
    if (ok) break;
    if (iter >= 100) {
      fl |= 8u;
      break;
    }
  }
#pragma unroll
  for (int i = 0; i < 4; i++)
    if (!isfinite(x[i])) fl |= 4u;
  flags = fl;
  return iter;
}

// Israel-Stewart bulk relaxation source of PI in conservative form, the
// physics of NEWEST10/Source.h:519-527 (Milne) and TXYZ5/Source.h:327-333:
//   u.d PI = -(PI + zeta theta)/tau_PI - 4/3 PI theta,
//   theta = d_mu u^mu (+ u^tau/tau in Milne),  S = u.d PI / u^0 + PI d_i(u^i/u^0).
// du[a][b] = d_b u^a.  Exact mode follows the CPU restatement's operation order.
template <bool FAST>
__host__ __device__ __forceinline__ double bulk_source(int milne, double tau, const double u[4], double PI,
                                                       const double du[4][4], double zeta, double tPI)
{
  const double iu0 = FAST ? frcp(u[0]) : 1.0 / u[0];
  double divv = 0;
#pragma unroll
  for (int i = 1; i < 4; i++) divv += du[i][i] * iu0 - u[i] * du[0][i] * (iu0 * iu0);
  double theta = du[0][0] + du[1][1] + du[2][2] + du[3][3];
  if (milne) theta = theta + u[0] / tau;
  const double rhs = FAST ? -(PI + zeta * theta) * frcp(tPI) - (4.0 / 3.0) * PI * theta
                          : -(PI + zeta * theta) / tPI - (4.0 / 3.0) * PI * theta;
  return rhs * iu0 + PI * divv;
}

// ---------------------------------------------------------------- initial state
// -zeta theta, the Navier-Stokes bulk pressure of CalcNS under BULK
// (NEWEST10/Source.h:740-745), theta = d_mu u^mu + u^tau/tau (Milne).
__host__ __device__ __forceinline__ double ns_bulk(double tau, const double u[4], const double du[4][4], double zeta)
{
  const double theta = du[0][0] + du[1][1] + du[2][2] + du[3][3] + u[0] / tau;
  return -(zeta * theta);
}

// 2 eta sigma^{mn}, the Navier-Stokes shear stress of CalcNS
// (NEWEST10/Source.h:643-751, written there as a computer-algebra expansion):
// sigma^{mn} = Delta^{mn}_{ab} D^a u^b with the Milne Christoffel terms, in
// the operation order of the CPU restatement (shear_source's sigma).
__host__ __device__ __forceinline__ void ns_shear(double tau, const double u[4], const double du[4][4], double SH,
                                                  double out[10])
{
  const double g[4] = {1.0, -1.0, -1.0, -(tau * tau)};
  const double gi[4] = {1.0, -1.0, -1.0, -(1.0 / (tau * tau))};
  double ul[4];
  for (int m = 0; m < 4; m++) ul[m] = g[m] * u[m];
  double Dc[4][4];
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < 4; b++) Dc[a][b] = du[a][b];
  const double itau = 1.0 / tau;
  Dc[0][3] = Dc[0][3] + tau * u[3];
  Dc[3][0] = Dc[3][0] + itau * u[3];
  Dc[3][3] = Dc[3][3] + itau * u[0];
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
  const double tr = trg - uGu;
  for (int m = 0; m < 4; m++) H[m] = Gu[m] + uG[m];
  const double c3 = uGu + (1.0 / 3.0) * tr;
  for (int m = 0; m < 4; m++)
    for (int n = m; n < 4; n++) {
      double sigma = 0.5 * (G[m][n] + G[n][m]) - 0.5 * (u[n] * H[m] + u[m] * H[n]) + u[m] * u[n] * c3;
      if (m == n) sigma = sigma - (1.0 / 3.0) * gi[m] * tr;
      out[pidx(m, n)] = 2.0 * SH * sigma;
    }
}

// ---------------------------------------------------------------- Landau matching
// The same algorithm, operation for operation, as the CPU restatement's
// landau_cell, so that the two agree to the bit.
/* the reference's velocity test on an eigenvector x (NEWEST10/ginit.h:
 * 567-585): x^0 != 0, v = x^i/x^0 with |v_x|, |v_y|, |v_eta| < 1 and
 * sqrt(v_x^2 + v_y^2 + v_eta^2) < 1 */
__host__ __device__ inline int landau_passes(const double x[4], double v[3])
{
  if (x[0] == 0.0) return 0;
  v[0] = x[1] / x[0];
  v[1] = x[2] / x[0];
  v[2] = x[3] / x[0];
  return fabs(v[0]) < 1 && fabs(v[1]) < 1 && fabs(v[2]) < 1 && sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) < 1;
}

/* 3x3 inverse by cofactors; 0 when singular or not finite */
__host__ __device__ inline int inv3(const double a[3][3], double r[3][3])
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
__host__ __device__ inline int landau_cell(double tau, const double T[10], double* E, double vout[3])
{
  const double g[4] = {1.0, -1.0, -1.0, -(tau * tau)};
  double M[4][4];
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < 4; b++) M[a][b] = T[pidx(a, b)] * g[b];
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
    double L[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
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

}  // namespace viral
