// adjoint.h -- hand-written reverse-mode (vector-Jacobian) forms of the
// pieces of one viscous KT stage in Milne and Cartesian coordinates: WENO5
// reconstruction, the centred flux / Var / eigenvalue of a cell, the
// Israel-Stewart shear source, the geometric source, the EOS and transport
// coefficients, the velocities and the roots by the implicit-function theorem
// at the converged root (Milne's 1-D Newton: de = -(df/dtheta) dtheta /
// (df/de); Cartesian 4-D gnewton: dx = -J^{-1} (dF/dprm) dprm), and the
// Cartesian shear tensor built from its five evolved components.
//
// Each *_vjp function takes the adjoint of its outputs and ADDS the adjoint of
// its inputs into the caller's accumulators.  All functions are host- and
// device-callable so they can be checked against finite differences on the CPU.
#pragma once

#include "physics.h"

namespace viral {

// The reverse pass matches finite differences, not a bitwise operation order,
// so its expressions may contract into FMAs (the forward kernels are compiled
// without contraction to keep the reference's rounding).
#pragma clang fp contract(fast)

#define VHD __host__ __device__ __forceinline__

// d(L or R)/d(x0..x4) of the WENO5-JS reconstruction of NEWEST10/utils.h:
// 102-160 on the stencil x = (q_{c-2}, .., q_{c+2}).  With w_c = a_c / sum a,
// a_c = d_c / (eps + b_c)^2 and candidate values q_c (linear in x):
//   dOut/dx_i = sum_c w_c dq_c/dx_i + (1/sum a) sum_c (q_c - Out) da_c/dx_i,
//   da_c/dx_i = -2 a_c / (eps + b_c) db_c/dx_i,
//   b_c = 13/12 A_c^2 + 1/4 C_c^2 (A_c, C_c linear in x).
// Returns the reconstruction itself (to rounding).  One reciprocal per
// candidate and one for the weight sum: the adjoint needs accuracy, not the
// forward's operation order.
VHD double weno_grad(const double x[5], bool right, double g[5])
{
  // linear forms of A_c and C_c (smoothness indicators), candidates q_c
  const double Ac[3][5] = {{0, 0, 1, -2, 1}, {0, 1, -2, 1, 0}, {1, -2, 1, 0, 0}};
  const double Cc[3][5] = {{0, 0, 3, -4, 1}, {0, 1, 0, -1, 0}, {1, -4, 3, 0, 0}};
  const double QL[3][5] = {{0, 0, 2, 5, -1}, {0, -1, 5, 2, 0}, {2, -7, 11, 0, 0}};
  const double QR[3][5] = {{0, 0, 11, -7, 2}, {0, 2, 5, -1, 0}, {-1, 5, 2, 0, 0}};
  const double dL[3] = {0.3, 0.6, 0.1}, dR[3] = {0.1, 0.6, 0.3};
  double A[3], C[3], al[3], k2[3], q[3], sum = 0;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    A[c] = C[c] = q[c] = 0;
#pragma unroll
    for (int i = 0; i < 5; i++) {
      A[c] += Ac[c][i] * x[i];
      C[c] += Cc[c][i] * x[i];
      q[c] += (right ? QR[c][i] : QL[c][i]) * x[i];
    }
    q[c] = q[c] * (1.0 / 6.0);
    const double it = frcp(WENO_EPS + (13. / 12.) * (A[c] * A[c]) + 0.25 * (C[c] * C[c]));
    al[c] = (right ? dR[c] : dL[c]) * (it * it);
    k2[c] = -2.0 * al[c] * it;  // da_c/db_c
    sum += al[c];
  }
  const double rs = frcp(sum);
  double w[3], out = 0;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    w[c] = al[c] * rs;
    out += w[c] * q[c];
  }
  double r[3];
#pragma unroll
  for (int c = 0; c < 3; c++) r[c] = (q[c] - out) * k2[c] * rs;
#pragma unroll
  for (int i = 0; i < 5; i++) {
    double gi = 0;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const double db = (13. / 6.) * A[c] * Ac[c][i] + 0.5 * C[c] * Cc[c][i];
      gi += w[c] * ((right ? QR[c][i] : QL[c][i]) * (1.0 / 6.0)) + r[c] * db;
    }
    g[i] = gi;
  }
  return out;
}

// Both WENO5-JS reconstructions of one cell from its stencil x (L, the value
// at the cell's right face, and R, at its left face) sharing the smoothness
// indicators, and the vector-Jacobian product of the pair:
//   g_i = sL dL/dx_i + sR dR/dx_i
//       = sum_c (sL wL_c dqL_c/dx_i + sR wR_c dqR_c/dx_i) + sum_c k_c db_c/dx_i,
//   k_c = -2 / (eps + b_c) (sL (qL_c - L) aL_c / SL + sR (qR_c - R) aR_c / SR),
// with the notation of weno_grad.  One set of b_c, 1/(eps + b_c) and db/dx
// serves both sides; the three 1/(eps + b_c) come from one reciprocal of
// their product and the two weight sums from one more, and the sparse
// stencil coefficients are written out.
VHD void weno_lr_vjp(const double x[5], double sL, double sR, double& L, double& R, double g[5])
{
  const double qm2 = x[0], qm1 = x[1], qc = x[2], qp1 = x[3], qp2 = x[4];
  // smoothness indicators b_c = 13/12 a_c^2 + 1/4 c_c^2 and s_c = 4 (eps + b_c)
  const double a0 = qc - 2.0 * qp1 + qp2, c0 = 3.0 * qc - 4.0 * qp1 + qp2;
  const double a1 = qm1 - 2.0 * qc + qp1, c1 = qm1 - qp1;
  const double a2 = qm2 - 2.0 * qm1 + qc, c2 = qm2 - 4.0 * qm1 + 3.0 * qc;
  const double k4 = 13.0 / 3.0, e4 = 4.0 * WENO_EPS;
  const double s0 = k4 * a0 * a0 + c0 * c0 + e4;
  const double s1 = k4 * a1 * a1 + c1 * c1 + e4;
  const double s2 = k4 * a2 * a2 + c2 * c2 + e4;
  // 1/s_c from one reciprocal of the product
  const double s12 = s1 * s2, inv3 = frcp1(s0 * s12);
  const double i0 = s12 * inv3, i1 = s0 * s2 * inv3, i2 = s0 * s1 * inv3;
  const double j0 = i0 * i0, j1 = i1 * i1, j2 = i2 * i2;
  // unnormalised weights d_c / s_c^2 (d = 0.3, 0.6, 0.1 for L; 0.1, 0.6, 0.3 for R)
  const double uL0 = 0.3 * j0, uL1 = 0.6 * j1, uL2 = 0.1 * j2;
  const double uR0 = 0.1 * j0, uR1 = 0.6 * j1, uR2 = 0.3 * j2;
  const double DL = uL0 + uL1 + uL2, DR = uR0 + uR1 + uR2;
  const double inv2 = frcp1(DL * DR);
  const double rL = DR * inv2, rR = DL * inv2;
  const double wL0 = uL0 * rL, wL1 = uL1 * rL, wL2 = uL2 * rL;
  const double wR0 = uR0 * rR, wR1 = uR1 * rR, wR2 = uR2 * rR;
  // candidate values
  const double o6 = 1.0 / 6.0;
  const double qL0 = (2.0 * qc + 5.0 * qp1 - qp2) * o6, qL1 = (-qm1 + 5.0 * qc + 2.0 * qp1) * o6,
               qL2 = (2.0 * qm2 - 7.0 * qm1 + 11.0 * qc) * o6;
  const double qR0 = (11.0 * qc - 7.0 * qp1 + 2.0 * qp2) * o6, qR1 = (2.0 * qm1 + 5.0 * qc - qp1) * o6,
               qR2 = (-qm2 + 5.0 * qm1 + 2.0 * qc) * o6;
  L = wL0 * qL0 + wL1 * qL1 + wL2 * qL2;
  R = wR0 * qR0 + wR1 * qR1 + wR2 * qR2;
  // k_c = -2/(eps + b_c) (sL (qL_c - L) wL_c + sR (qR_c - R) wR_c), 1/(eps + b_c) = 4/s_c;
  // db_c = 13/6 a_c da_c + 1/2 c_c dc_c
  const double m0 = -8.0 * i0 * (sL * (qL0 - L) * wL0 + sR * (qR0 - R) * wR0);
  const double m1 = -8.0 * i1 * (sL * (qL1 - L) * wL1 + sR * (qR1 - R) * wR1);
  const double m2 = -8.0 * i2 * (sL * (qL2 - L) * wL2 + sR * (qR2 - R) * wR2);
  const double ka0 = (13.0 / 6.0) * a0 * m0, ka1 = (13.0 / 6.0) * a1 * m1, ka2 = (13.0 / 6.0) * a2 * m2;
  const double kc0 = 0.5 * c0 * m0, kc1 = 0.5 * c1 * m1, kc2 = 0.5 * c2 * m2;
  // the candidates' linear parts, weighted
  const double hL0 = sL * wL0 * o6, hL1 = sL * wL1 * o6, hL2 = sL * wL2 * o6;
  const double hR0 = sR * wR0 * o6, hR1 = sR * wR1 * o6, hR2 = sR * wR2 * o6;
  g[0] = 2.0 * hL2 - hR2 + ka2 + kc2;
  g[1] = -hL1 - 7.0 * hL2 + 2.0 * hR1 + 5.0 * hR2 + ka1 - 2.0 * ka2 + kc1 - 4.0 * kc2;
  g[2] = 2.0 * hL0 + 5.0 * hL1 + 11.0 * hL2 + 11.0 * hR0 + 5.0 * hR1 + 2.0 * hR2 + ka0 - 2.0 * ka1 + ka2 +
         3.0 * kc0 + 3.0 * kc2;
  g[3] = 5.0 * hL0 + 2.0 * hL1 - 7.0 * hR0 - hR1 - 2.0 * ka0 + ka1 - 4.0 * kc0 - kc1;
  g[4] = -hL0 + 2.0 * hR0 + ka0 + kc0;
}

// P(e), T(e), s(e) and their e-derivatives as the forward step evaluates
// them: the table's interpolant where fast mode reads the table (use_tab is
// set only in fast mode), the closed forms elsewhere.  The conformal s =
// (e + P)/T depends on the state's P and is left to the caller (s = ds = 0).
VHD void eos_d(const Params& p, double e, double& P, double& dP, double& T, double& dT, double& s, double& ds)
{
  if (!p.eos) {
    P = e / 3.0;
    dP = 1.0 / 3.0;
    T = eos_T(p, e);
    dT = T / (4.0 * e);  // T = (e/FACTOR)^(1/4)
    s = ds = 0.0;
    return;
  }
  double v[4], dv[4];
  if (p.use_tab && eos_tab_lookup_d(p.tab, e, v, dv)) {
    P = v[0];
    dP = dv[0];
    T = v[2];
    dT = dv[2];
    s = v[3];
    ds = dv[3];
    return;
  }
  P = s95p_p(e);
  dP = s95p_a(e);
  T = s95p_T(e);
  dT = s95p_dT(e);
  s = s95p_s(e);
  ds = s95p_ds(e);
}

// u = gamma (1, v), gamma = (1 - vx^2 - vy^2 - t2 ve^2)^(-1/2):
// adds d<ub, u>/dv to vb
VHD void gamma_vjp(const double v[3], double t2, const double ub[4], double vb[3])
{
  const double gm = 1.0 / sqrt(1 - v[0] * v[0] - v[1] * v[1] - t2 * v[2] * v[2]);
  const double s = ub[0] + ub[1] * v[0] + ub[2] * v[1] + ub[3] * v[2];
  const double g3 = gm * gm * gm;
  vb[0] += gm * ub[1] + g3 * v[0] * s;
  vb[1] += gm * ub[2] + g3 * v[1] * s;
  vb[2] += gm * ub[3] + g3 * t2 * v[2] * s;
}

// Centred KT flux F_l (15 rows), Var V_l and eigenvalue A = u^d/u^0 of one
// cell for direction d (NEWEST10/hydroKT.h:16-98, 176-258, 364-381):
//   T rows  F_l = tau (pi^{d+1,l} + [l = d+1] pp + u^{d+1} u^l w),  V_l = tau T^{tau l}
//   pi rows F_l = pi u^{d+1}/u^0, V = pi;  PI row likewise;
//   w = e + P - PI, pp = P - PI (eta direction: (P - PI)/tau^2).
// Adds the adjoint of (e, P, PI, u, pi, T) given Fb, Vb, Ab.
VHD void flux_cell_vjp(int d, double tau, double e, double P, double PI, const double u[4], const double pi10[10],
                       const double Fb[15], const double Vb[15], double Ab, double& eb, double& Pb, double& PIb,
                       double ub[4], double pib[10], double Tb[4])
{
  const int row = d + 1;
  const double ud = u[row], u0 = u[0];
  const double w = e + P - PI;
  double wb = 0, ppb = 0, udb = 0;
#pragma unroll
  for (int l = 0; l < 4; l++) {
    const double vb = tau * Fb[l];
    pib[pidx(row, l)] += vb;
    if (l == row) ppb += vb;
    udb += vb * u[l] * w;
    ub[l] += vb * ud * w;
    wb += vb * ud * u[l];
  }
  const double iu0 = frcp(u0);
#pragma unroll
  for (int l = 0; l < 11; l++) {
    const double f = l < 10 ? pi10[l] : PI;
    const double fb = Fb[4 + l];
    if (l < 10) pib[l] += fb * ud * iu0;
    else PIb += fb * ud * iu0;
    udb += fb * f * iu0;
    ub[0] -= fb * f * ud * iu0 * iu0;
  }
  const double k = d == 2 ? 1.0 / (tau * tau) : 1.0;
  Pb += ppb * k;
  PIb -= ppb * k;
  eb += wb;
  Pb += wb;
  PIb -= wb;
#pragma unroll
  for (int l = 0; l < 4; l++) Tb[l] += tau * Vb[l];
#pragma unroll
  for (int l = 0; l < 10; l++) pib[l] += Vb[4 + l];
  PIb += Vb[14];
  udb += Ab * iu0;
  ub[0] -= Ab * ud * iu0 * iu0;
  ub[row] += udb;
}

// Reverse of shear_source (physics.h): given ob = d(J)/d(out), adds the
// adjoint of u, pi, du, SH and tpi.  The forward intermediates are recomputed
// in the same order.
VHD void shear_source_vjp(int milne, double tau, const double u[4], const double pi10[10], const double du[4][4],
                          double SH, double tpi, const double ob[10], double ub[4], double pib[10], double dub[4][4],
                          double& SHb, double& tpib)
{
  const double g[4] = {1.0, -1.0, -1.0, milne ? -(tau * tau) : -1.0};
  const double gi[4] = {1.0, -1.0, -1.0, milne ? -(1.0 / (tau * tau)) : -1.0};
  double ul[4];
  for (int m = 0; m < 4; m++) ul[m] = g[m] * u[m];
  const double iu0 = frcp(u[0]);
  const double itpi = frcp(tpi);
  double divv = 0;
  for (int i = 1; i < 4; i++) divv += du[i][i] * iu0 - u[i] * du[0][i] * (iu0 * iu0);
  const double itau = milne ? 1.0 / tau : 0.0;
  // Dc = du with the Christoffel part, G^{ab} = g^{aa} Dc[b][a]; evaluated on
  // the fly (the loops are unrolled, so the index tests fold away)
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
  for (int m = 0; m < 4; m++) H[m] = Gu[m] + uG[m];
  const double c3 = uGu + (1.0 / 3.0) * tr;
  double PD[4];
  for (int m = 0; m < 4; m++) {
    double s = 0;
    for (int a = 0; a < 4; a++) s += pi10[pidx(m, a)] * Dul[a];
    PD[m] = s;
  }
  // ---- reverse sweep
  double ulb[4] = {0, 0, 0, 0}, Hb[4] = {0, 0, 0, 0}, PDb[4] = {0, 0, 0, 0}, Dulb[4] = {0, 0, 0, 0};
  // the adjoint of Dc; that of G^{ab} = g^{aa} Dc[b][a] lands on Dc[b][a]
  double Dcb[4][4];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) Dcb[a][b] = 0;
  double iu0b = 0, divvb = 0, thetab = 0, c3b = 0, trb = 0;
#pragma unroll
  for (int m = 0; m < 4; m++)
#pragma unroll
    for (int n = m; n < 4; n++) {
      const int q = pidx(m, n);
      const double o = ob[q];
      const double pmn = pi10[q];
      double sigma = 0.5 * (G(m, n) + G(n, m)) - 0.5 * (u[n] * H[m] + u[m] * H[n]) + u[m] * u[n] * c3;
      if (m == n) sigma = sigma - (1.0 / 3.0) * gi[m] * tr;
      const double pu = u[m] * PD[n] + u[n] * PD[m];
      const double rel = -(pmn - 2.0 * SH * sigma) * itpi;
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
      // out = (rhs - chr) iu0 + pmn divv
      const double rhsb = o * iu0, chrb = -o * iu0;
      iu0b += o * (rhs - chr);
      pib[q] += o * divv;
      divvb += o * pmn;
      if (milne) {
        if (m == 0) { ub[3] += chrb * tau * pi10[pidx(3, n)]; pib[pidx(3, n)] += chrb * u[3] * tau; }
        if (m == 3) {
          ub[0] += chrb * itau * pi10[pidx(3, n)]; pib[pidx(3, n)] += chrb * u[0] * itau;
          ub[3] += chrb * itau * pi10[pidx(0, n)]; pib[pidx(0, n)] += chrb * u[3] * itau;
        }
        if (n == 0) { ub[3] += chrb * tau * pi10[pidx(m, 3)]; pib[pidx(m, 3)] += chrb * u[3] * tau; }
        if (n == 3) {
          ub[0] += chrb * itau * pi10[pidx(m, 3)]; pib[pidx(m, 3)] += chrb * u[0] * itau;
          ub[3] += chrb * itau * pi10[pidx(m, 0)]; pib[pidx(m, 0)] += chrb * u[3] * itau;
        }
      }
      // rhs = rel - 4/3 pmn theta - pu
      const double relb = rhsb;
      pib[q] += -(4.0 / 3.0) * theta * rhsb;
      thetab += -(4.0 / 3.0) * pmn * rhsb;
      const double pub = -rhsb;
      // rel = -(pmn - 2 SH sigma) / tpi
      pib[q] += -relb * itpi;
      SHb += 2.0 * sigma * relb * itpi;
      const double sigb = 2.0 * SH * relb * itpi;
      tpib += (pmn - 2.0 * SH * sigma) * (itpi * itpi) * relb;
      // pu = u^m PD^n + u^n PD^m
      ub[m] += pub * PD[n];
      PDb[n] += pub * u[m];
      ub[n] += pub * PD[m];
      PDb[m] += pub * u[n];
      // sigma
      Dcb[n][m] += gi[m] * (0.5 * sigb);
      Dcb[m][n] += gi[n] * (0.5 * sigb);
      ub[n] -= 0.5 * sigb * H[m];
      Hb[m] -= 0.5 * sigb * u[n];
      ub[m] -= 0.5 * sigb * H[n];
      Hb[n] -= 0.5 * sigb * u[m];
      ub[m] += sigb * u[n] * c3;
      ub[n] += sigb * u[m] * c3;
      c3b += sigb * u[m] * u[n];
      if (m == n) trb -= (1.0 / 3.0) * gi[m] * sigb;
    }
  // PD^m = pi^{ma} Dul_a
  for (int m = 0; m < 4; m++)
    for (int a = 0; a < 4; a++) {
      pib[pidx(m, a)] += PDb[m] * Dul[a];
      Dulb[a] += PDb[m] * pi10[pidx(m, a)];
    }
  double uGub = c3b;
  trb += c3b * (1.0 / 3.0);
  double Gub[4], uGb[4];
  for (int m = 0; m < 4; m++) Gub[m] = uGb[m] = Hb[m];
  const double trgb = trb;
  uGub -= trb;
  for (int a = 0; a < 4; a++) Dcb[a][a] += gi[a] * (g[a] * trgb);
  for (int a = 0; a < 4; a++) {
    ulb[a] += uGub * Gu[a];
    Gub[a] += uGub * ul[a];
  }
#pragma unroll
  for (int m = 0; m < 4; m++)
#pragma unroll
    for (int b = 0; b < 4; b++) {
      Dcb[b][m] += gi[m] * (Gub[m] * ul[b]);
      ulb[b] += Gub[m] * G(m, b);
      ulb[b] += uGb[m] * G(b, m);
      Dcb[m][b] += gi[b] * (uGb[m] * ul[b]);
    }
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) {
      ub[b] += g[a] * Dulb[a] * Dc(a, b);
      Dcb[a][b] += g[a] * Dulb[a] * u[b];
    }
  for (int a = 0; a < 4; a++) Dcb[a][a] += thetab;
  if (milne) {
    ub[3] += Dcb[0][3] * tau + Dcb[3][0] * itau;
    ub[0] += Dcb[3][3] * itau;
  }
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < 4; b++) dub[a][b] += Dcb[a][b];
  for (int i = 1; i < 4; i++) {
    dub[i][i] += divvb * iu0;
    dub[0][i] -= divvb * u[i] * iu0 * iu0;
    ub[i] -= divvb * du[0][i] * iu0 * iu0;
    iu0b += divvb * (du[i][i] - 2.0 * u[i] * du[0][i] * iu0);
  }
  for (int m = 0; m < 4; m++) ub[m] += g[m] * ulb[m];
  ub[0] -= iu0b * iu0 * iu0;
}

VHD void shear_transport_vjp(const Params& p, int milne, double tau, double e, double P, const double u[4],
                             const double pi10[10], const double du[4][4], const double Sb10[10], double& eb,
                             double& Pb, double ub[4], double pib[10], double dub[4][4], double& cetab, double& ctpib);
VHD void bulk_vjp(const Params& p, int milne, double tau, double e, double P, const double u[4], double PI,
                  const double du[4][4], double Sb14, double& eb, double& Pb, double& PIb, double ub[4],
                  double dub[4][4]);

// Sources of one cell (NEWEST10/Source.h:208-221 geometric, :347-517 shear)
// given the velocity gradients du and the transport coefficients: adds the
// adjoint of (e, P, PI, u, pi, du) and of the viscosity / relaxation
// coefficients (eta_coef, tpi_coef) for the source adjoint Sb (rows 0..13;
// the bulk row 14 is the caller's, bulk_vjp).
// The entropy density is (e + P)/T(e) for the conformal EOS and s(e) for
// s95p, as the forward evaluates it (eos_d).
VHD void source_vjp(const Params& p, double tau, double e, double P, double PI, const double u[4], const double pi10[10],
                    const double du[4][4], const double Sb[15], double& eb, double& Pb, double& PIb, double ub[4],
                    double pib[10], double dub[4][4], double& cetab, double& ctpib)
{
  const double w = e + P - PI;
  double wb = 0;
  // S0 = -(P - PI) - tau^2 (pi^{ee} + w u3 u3)
  Pb -= Sb[0];
  PIb += Sb[0];
  pib[9] -= tau * tau * Sb[0];
  wb -= tau * tau * u[3] * u[3] * Sb[0];
  ub[3] -= 2.0 * tau * tau * w * u[3] * Sb[0];
  // S3 = -2 (pi^{te} + w u0 u3)
  pib[3] -= 2.0 * Sb[3];
  wb -= 2.0 * u[0] * u[3] * Sb[3];
  ub[0] -= 2.0 * w * u[3] * Sb[3];
  ub[3] -= 2.0 * w * u[0] * Sb[3];
  eb += wb;
  Pb += wb;
  PIb -= wb;
  shear_transport_vjp(p, 1, tau, e, P, u, pi10, du, Sb + 4, eb, Pb, ub, pib, dub, cetab, ctpib);
  // the bulk row (Sb[14]) is the caller's: bulk_vjp
}

// The shear source with its transport coefficients eta(s) and tau_pi in
// reverse (milne 0: Cartesian): given the adjoint Sb10 of its ten outputs,
// adds the adjoint of (e, P, u, pi10, du) and of eta_coef, tpi_coef.
VHD void shear_transport_vjp(const Params& p, int milne, double tau, double e, double P, const double u[4],
                             const double pi10[10], const double du[4][4], const double Sb10[10], double& eb,
                             double& Pb, double ub[4], double pib[10], double dub[4][4], double& cetab, double& ctpib)
{
  double Pe, dPe, T, dT, s, ds;
  eos_d(p, e, Pe, dPe, T, dT, s, ds);
  if (!p.eos) s = (e + P) / T;
  const double SH = eta_of(p, s);
  const double tpi = tpi_of(p, SH, P, e);
  double SHb = 0, tpib = 0;
  shear_source_vjp(milne, tau, u, pi10, du, SH, tpi, Sb10, ub, pib, dub, SHb, tpib);
  if (p.tpi_law == 1) {  // c eta / P
    SHb += tpib * p.tpi_coef / P;
    Pb -= tpib * p.tpi_coef * SH / (P * P);
    ctpib += tpib * SH / P;
  } else if (p.tpi_law == 2) {
    ctpib += tpib;
  } else {  // c eta / (e + P)
    const double ep = e + P;
    SHb += tpib * p.tpi_coef / ep;
    eb -= tpib * p.tpi_coef * SH / (ep * ep);
    Pb -= tpib * p.tpi_coef * SH / (ep * ep);
    ctpib += tpib * SH / ep;
  }
  double sb;
  if (p.eta_law == 1) {
    sb = SHb * p.eta_coef / (4.0 * PIE);
    cetab += SHb * s / (4.0 * PIE);
  } else {
    sb = SHb * p.eta_coef;
    cetab += SHb * s;
  }
  if (!p.eos) {  // s = (e + P) / T(e)
    eb += sb / T - sb * (e + P) / (T * T) * dT;
    Pb += sb / T;
  } else {
    eb += sb * ds;
  }
}

// The Israel-Stewart bulk source (bulk_source, physics.h; NEWEST10/Source.h:
// 519-527, TXYZ5/Source.h:327-333) with zeta(s) and tau_Pi(zeta, P) in
// reverse (milne 0: Cartesian): given the adjoint Sb14 of the PI row, adds
// the adjoint of (e, P, PI, u, du).
//   out = rhs / u^0 + PI divv,  rhs = -(PI + zeta theta)/tau_Pi - 4/3 PI theta,
//   theta = sum_a du[a][a] (+ u^0/tau in Milne),
//   divv = sum_i (du[i][i]/u^0 - u^i du[0][i]/(u^0)^2).
VHD void bulk_vjp(const Params& p, int milne, double tau, double e, double P, const double u[4], double PI,
                  const double du[4][4], double Sb14, double& eb, double& Pb, double& PIb, double ub[4],
                  double dub[4][4])
{
  if (Sb14 == 0.0) return;
  double Pe, dPe, T, dT, s, ds;
  eos_d(p, e, Pe, dPe, T, dT, s, ds);
  if (!p.eos) s = (e + P) / T;
  const double zeta = zeta_of(p, s);
  const double tPI = tpib_of(p, zeta, P);
  const double iu0 = 1.0 / u[0];
  double divv = 0, sdd = 0, sud = 0;
  for (int i = 1; i < 4; i++) {
    divv += du[i][i] * iu0 - u[i] * du[0][i] * (iu0 * iu0);
    sdd += du[i][i];
    sud += u[i] * du[0][i];
  }
  double theta = du[0][0] + du[1][1] + du[2][2] + du[3][3];
  if (milne) theta = theta + u[0] / tau;
  const double itp = 1.0 / tPI;
  const double rhs = -(PI + zeta * theta) * itp - (4.0 / 3.0) * PI * theta;
  // out = rhs iu0 + PI divv
  const double rhsb = Sb14 * iu0;
  double iu0b = Sb14 * rhs;
  PIb += Sb14 * divv;
  const double divvb = Sb14 * PI;
  iu0b += divvb * (sdd - 2.0 * iu0 * sud);
  for (int i = 1; i < 4; i++) {
    dub[i][i] += divvb * iu0;
    dub[0][i] -= divvb * u[i] * iu0 * iu0;
    ub[i] -= divvb * du[0][i] * iu0 * iu0;
  }
  // rhs
  PIb += rhsb * (-itp - (4.0 / 3.0) * theta);
  const double thetab = rhsb * (-zeta * itp - (4.0 / 3.0) * PI);
  double zetab = rhsb * (-theta * itp);
  const double tPIb = rhsb * (PI + zeta * theta) * itp * itp;
  for (int a = 0; a < 4; a++) dub[a][a] += thetab;
  if (milne) ub[0] += thetab / tau;
  ub[0] -= iu0b * iu0 * iu0;
  // tau_Pi = c zeta / P (law 1) or c
  if (p.tpib_law == 1) {
    zetab += tPIb * p.tpib_coef / P;
    Pb -= tPIb * p.tpib_coef * zeta / (P * P);
  }
  // zeta = c s / 4 pi (law 1) or c
  if (p.zeta_law == 1) {
    const double sb = zetab * p.zeta_coef / (4.0 * PIE);
    if (!p.eos) {  // s = (e + P) / T(e)
      eb += sb / T - sb * (e + P) / (T * T) * dT;
      Pb += sb / T;
    } else {
      eb += sb * ds;
    }
  }
}

// Velocities and root of one cell (NEWEST10/root.h:212-424) in reverse.
// Forward: T0 = V[0..3]/tn, pi = V[4..13], PI = V[14]; e solves
//   f(e) = e - A + M2/(A + P(e) - PI) = 0,  A = T0_0 - pi_0,
//   M2 = sum_i g_i (T0_i - pi_i)^2 (g = 1, 1, tn^2);
// v_i = (T0_i - pi_i)/D, D = A + P(e) - PI; u = gamma (1, v); P, Temp of e.
// Given the adjoint of (e, P, Temp, v, u) and of the stored T0, pi, PI,
// writes Vb (15).  The root enters through the implicit-function theorem at
// the converged e: de = -(f_A dA + f_M2 dM2 + f_PI dPI)/f_e.
VHD void update_vjp(const Params& p, double tn, const double T0[4], const double pi10[10], double PI, double e,
                    double ebo, double Pbo, double Tempbo, const double vbo[3], const double ubo[4], const double T0bo[4],
                    const double pibo[10], double PIbo, double Vb[15])
{
  const double t2 = tn * tn;
  // P(e) and Temp(e) as the forward stores them; a = dP/de of that P (the
  // interpolant's derivative with the table), which is also f_e's
  double P, a, Te, dT, se, dse;
  eos_d(p, e, P, a, Te, dT, se, dse);
  const double N[3] = {T0[1] - pi10[1], T0[2] - pi10[2], T0[3] - pi10[3]};
  const double D = T0[0] - pi10[0] + P - PI, iD = 1.0 / D;
  const double v[3] = {N[0] * iD, N[1] * iD, N[2] * iD};
  double vb[3] = {vbo[0], vbo[1], vbo[2]};
  gamma_vjp(v, t2, ubo, vb);
  double T0b[4] = {T0bo[0], T0bo[1], T0bo[2], T0bo[3]};
  double pib[10];
  for (int l = 0; l < 10; l++) pib[l] = pibo[l];
  double PIb = PIbo;
  double eb = ebo + Pbo * a + Tempbo * dT;
  double Db = 0;
  for (int i = 0; i < 3; i++) {
    const double Nb = vb[i] * iD;
    Db -= vb[i] * v[i] * iD;
    T0b[i + 1] += Nb;
    pib[i + 1] -= Nb;
  }
  T0b[0] += Db;
  pib[0] -= Db;
  eb += Db * a;
  PIb -= Db;
  // implicit-function theorem at the root
  const double A = T0[0] - pi10[0];
  const double M2 = N[0] * N[0] + N[1] * N[1] + t2 * N[2] * N[2];
  const double Q = A + P - PI, iQ = 1.0 / Q, iQ2 = iQ * iQ;
  const double fe = 1.0 - M2 * a * iQ2;
  const double lam = -eb / fe;
  const double Ab = lam * (-1.0 - M2 * iQ2);
  const double M2b = lam * iQ;
  PIb += lam * M2 * iQ2;
  T0b[0] += Ab;
  pib[0] -= Ab;
  const double gg[3] = {1.0, 1.0, t2};
  for (int i = 0; i < 3; i++) {
    const double nb = M2b * 2.0 * gg[i] * N[i];
    T0b[i + 1] += nb;
    pib[i + 1] -= nb;
  }
  const double itn = 1.0 / tn;
  for (int l = 0; l < 4; l++) Vb[l] = T0b[l] * itn;
  for (int l = 0; l < 10; l++) Vb[4 + l] = pib[l];
  Vb[14] = PIb;
}

// ---------------------------------------------------------------- Cartesian (TXYZ5)

// The full pi^{mu nu} of the Cartesian step (cart_pi10, from u and the five
// evolved components pe = {xx, yy, xy, xz, yz}) in reverse: given the adjoint
// ob of its ten slots, adds the adjoint of u and pe.
VHD void cart_pi10_vjp(const double u[4], const double pe[5], const double ob[10], double ub[4], double peb[5])
{
  const double u0 = u[0], u1 = u[1], u2 = u[2], u3 = u[3];
  const double p1 = pe[0], p2 = pe[1], p3 = pe[2], p4 = pe[3], p5 = pe[4];
  const double num =
      p1 * u1 * u1 + 2.0 * p3 * u1 * u2 + p2 * u2 * u2 + 2.0 * u3 * (p4 * u1 + p5 * u2) - u0 * u0 * (p1 + p2);
  const double iden = 1.0 / (u0 * u0 - u3 * u3);
  const double zz = num * iden;
  const double iu0 = 1.0 / u0;
  // slots tx, ty, tz = n_i / u0
  const double n1 = p1 * u1 + p3 * u2 + p4 * u3, n2 = p3 * u1 + p2 * u2 + p5 * u3, n3 = p4 * u1 + p5 * u2 + zz * u3;
  const double b1 = ob[1] * iu0, b2 = ob[2] * iu0, b3 = ob[3] * iu0;
  const double zzb = ob[0] + ob[9] + b3 * u3;
  double p1b = ob[0] + ob[4] + b1 * u1;
  double p2b = ob[0] + ob[7] + b2 * u2;
  double p3b = ob[5] + b1 * u2 + b2 * u1;
  double p4b = ob[6] + b1 * u3 + b3 * u1;
  double p5b = ob[8] + b2 * u3 + b3 * u2;
  double u0b = -(b1 * n1 + b2 * n2 + b3 * n3) * iu0;
  double u1b = b1 * p1 + b2 * p3 + b3 * p4;
  double u2b = b1 * p3 + b2 * p2 + b3 * p5;
  double u3b = b1 * p4 + b2 * p5 + b3 * zz;
  // zz = num / den, den = u0^2 - u3^2
  const double numb = zzb * iden, denb = -zzb * zz * iden;
  p1b += numb * (u1 * u1 - u0 * u0);
  p2b += numb * (u2 * u2 - u0 * u0);
  p3b += numb * 2.0 * u1 * u2;
  p4b += numb * 2.0 * u3 * u1;
  p5b += numb * 2.0 * u3 * u2;
  u1b += numb * 2.0 * (p1 * u1 + p3 * u2 + p4 * u3);
  u2b += numb * 2.0 * (p3 * u1 + p2 * u2 + p5 * u3);
  u3b += numb * 2.0 * (p4 * u1 + p5 * u2) - denb * 2.0 * u3;
  u0b += -numb * 2.0 * u0 * (p1 + p2) + denb * 2.0 * u0;
  ub[0] += u0b;
  ub[1] += u1b;
  ub[2] += u2b;
  ub[3] += u3b;
  peb[0] += p1b;
  peb[1] += p2b;
  peb[2] += p3b;
  peb[3] += p4b;
  peb[4] += p5b;
}

// Sources of one Cartesian cell (TXYZ5/Source.h:152-155, 272-333): no
// geometric rows; the shear source's xx, yy, xy, xz, yz slots feed Var rows
// 4..8 (zero where the reference's slice loop stops short, `live` false).
// Given Sb (15 rows), adds the adjoint of (e, P, u, pe, du) and of the shear
// coefficients; the bulk row is the caller's (bulk_vjp).
VHD void source_vjp_cart(const Params& p, double e, double P, const double u[4], const double pe[5], bool live,
                         const double du[4][4], const double Sb[15], double& eb, double& Pb, double ub[4],
                         double peb[5], double dub[4][4], double& cetab, double& ctpib)
{
  if (!live) return;
  double pi10[10], Sb10[10], pib[10];
  cart_pi10(u, pe[0], pe[1], pe[2], pe[3], pe[4], pi10);
  for (int l = 0; l < 10; l++) Sb10[l] = pib[l] = 0.0;
  for (int l = 0; l < 5; l++) Sb10[cart5(l)] = Sb[4 + l];
  shear_transport_vjp(p, 0, 1.0, e, P, u, pi10, du, Sb10, eb, Pb, ub, pib, dub, cetab, ctpib);
  cart_pi10_vjp(u, pe, pib, ub, peb);
  // the bulk row (Sb[14], also zero where `live` is false) is the caller's: bulk_vjp
}

// Solves A^T y = b for a 4x4 A by Gaussian elimination with partial pivoting.
// Every index is a compile-time constant once the loops unroll (rows are
// swapped by comparing the pivot row with each candidate), so the matrix
// stays in registers.
VHD void solve4t(const double A[4][4], const double b[4], double y[4])
{
  double M[4][5];
#pragma unroll
  for (int i = 0; i < 4; i++) {
#pragma unroll
    for (int k = 0; k < 4; k++) M[i][k] = A[k][i];
    M[i][4] = b[i];
  }
#pragma unroll
  for (int j = 0; j < 4; j++) {
    int ip = j;
    double mx = fabs(M[j][j]);
#pragma unroll
    for (int i = j + 1; i < 4; i++)
      if (fabs(M[i][j]) > mx) {
        mx = fabs(M[i][j]);
        ip = i;
      }
#pragma unroll
    for (int i = j + 1; i < 4; i++)
      if (ip == i) {
#pragma unroll
        for (int k = 0; k < 5; k++) {
          const double t = M[j][k];
          M[j][k] = M[i][k];
          M[i][k] = t;
        }
      }
    const double inv = 1.0 / M[j][j];
#pragma unroll
    for (int i = j + 1; i < 4; i++) {
      const double f = M[i][j] * inv;
#pragma unroll
      for (int k = j; k < 5; k++) M[i][k] -= f * M[j][k];
    }
  }
#pragma unroll
  for (int i = 3; i >= 0; i--) {
    double t = M[i][4];
#pragma unroll
    for (int k = i + 1; k < 4; k++) t -= M[i][k] * y[k];
    y[i] = t / M[i][i];
  }
}

// The Cartesian update and its 4-D root (TXYZ5/time.h:100-165,
// TXYZ5/multiroot.h:203-479) in reverse.  Forward: prm = (T^{t mu}, p1..p5,
// PI) are the stage's Var rows {0..8, 14}; x = (e, u^x, u^y, u^z) solves
// F(x; prm) = 0 (cart_f); the stored outputs are T^{t mu}, the five pi and PI
// (= prm), e, P(e), Temp(e) (closed forms: the Cartesian root and its outputs
// do not read the table), v = u / u^t, u^t = sqrt(1 + u.u) and u.  Given the
// adjoint of those outputs, writes Vb (15 rows; rows 9..13 zero).  The root
// enters by the implicit-function theorem at the converged x:
// J^T lam = xbar with J = dF/dx (cart_df), prm_bar = -(dF/dprm)^T lam.
VHD void cart_update_vjp(const Params& p, const double prm[10], const double x[4], const double T0bo[4],
                         const double pebo[5], double PIbo, double ebo, double Pbo, double Tempbo, const double vbo[3],
                         const double ubo[4], double Vb[15])
{
  Params pc = p;
  pc.use_tab = 0;
  double Pe, a, Te, dT, se, dse;
  eos_d(pc, x[0], Pe, a, Te, dT, se, dse);
  const double u1 = x[1], u2 = x[2], u3 = x[3];
  const double q = 1.0 + u1 * u1 + u2 * u2 + u3 * u3, sq = sqrt(q), isq = 1.0 / sq;
  // adjoint of x from e, P, Temp, u^t, u, v
  const double vs = (vbo[0] * u1 + vbo[1] * u2 + vbo[2] * u3) * isq * isq * isq;
  double xb[4];
  xb[0] = ebo + Pbo * a + Tempbo * dT;
  xb[1] = ubo[1] + ubo[0] * u1 * isq + vbo[0] * isq - u1 * vs;
  xb[2] = ubo[2] + ubo[0] * u2 * isq + vbo[1] * isq - u2 * vs;
  xb[3] = ubo[3] + ubo[0] * u3 * isq + vbo[2] * isq - u3 * vs;
  double J[4][4], lam[4];
  cart_df(pc, prm, x, J);
  solve4t(J, xb, lam);
  // dF/dprm of cart_f, prm = (W, X, Y, Z, p1, p2, p3, p4, p5, PI)
  const double l0 = lam[0], l1 = lam[1], l2 = lam[2], l3 = lam[3];
  Vb[0] = T0bo[0] - (l0 * sq - l3 * u3);
  Vb[1] = T0bo[1] - (-l0 * u1 + l1 * sq);
  Vb[2] = T0bo[2] - (-l0 * u2 + l2 * sq);
  Vb[3] = T0bo[3] - (-l0 * u3 + l3 * sq);
  Vb[4] = pebo[0] - (-l1 * u1 + l3 * u3);
  Vb[5] = pebo[1] - (-l2 * u2 + l3 * u3);
  Vb[6] = pebo[2] - (-l1 * u2 - l2 * u1);
  Vb[7] = pebo[3] - (-l1 * u3 - l3 * u1);
  Vb[8] = pebo[4] - (-l2 * u3 - l3 * u2);
  for (int l = 9; l < 14; l++) Vb[l] = 0.0;
  Vb[14] = PIbo - (l1 * u1 * q + l2 * u2 * q + l3 * u3);
}

#undef VHD

#pragma clang fp contract(off)

}  // namespace viral
