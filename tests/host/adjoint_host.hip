// Host build of the adjoint primitives (viral_amd/csrc/adjoint.h) and of the
// forward pieces they reverse, for finite-difference checks on the CPU.
// Test infrastructure; compiled by __graft_entry__.build() into
// tests/host/_build/.
#include <hip/hip_runtime.h>

#include "../../viral_amd/csrc/adjoint.h"
#include "../../viral_amd/csrc/eos_table.h"

using namespace viral;

// the library's s95p table, built on the host
static EosTab host_table()
{
  static std::vector<double> v;
  static std::vector<unsigned char> c;
  if (v.empty()) build_eos_table(v, c);
  return EosTab{v.data(), c.data(), kTabEmin, kTabNoct, kTabK};
}

static Params make_params(int eta_law, double eta_coef, int tpi_law, double tpi_coef, int eos = 0)
{
  Params p{};
  p.milne = 1;
  p.lbi = 0;
  p.eos = eos;
  p.eta_law = eta_law;
  p.tpi_law = tpi_law;
  p.max_iter = 5000;
  p.rk = 2;
  p.scheme = 0;
  p.quirks = 0;
  p.dx = p.dy = p.deta = 0.1;
  p.eta_coef = eta_coef;
  p.tpi_coef = tpi_coef;
  p.factor = 3.0 * 42.25 * PIE * PIE / 90.0;
  p.t_freeze = 0;
  p.use_tab = 0;
  return p;
}

extern "C" {

// closed-form P, T, s of the EOS and their e-derivatives (eos_d)
void ah_eos(int eos, double e, double* out /* P, T, s */)
{
  Params p = make_params(0, 0.2, 0, 5.0, eos);
  out[0] = eos_p(p, e);
  out[1] = eos_T(p, e);
  out[2] = eos ? s95p_s(e) : 0.0;
}
void ah_eos_d(int eos, double e, double* out /* P, dP, T, dT, s, ds */)
{
  Params p = make_params(0, 0.2, 0, 5.0, eos);
  eos_d(p, e, out[0], out[1], out[2], out[3], out[4], out[5]);
}
// the table's interpolant and its e-derivative; returns 0 where the table
// does not apply
int ah_tab(double e, double* out)
{
  return eos_tab_lookup(host_table(), e, out) ? 1 : 0;
}
int ah_tab_d(double e, double* out, double* dout)
{
  return eos_tab_lookup_d(host_table(), e, out, dout) ? 1 : 0;
}

double ah_weno(const double* x, int right)
{
  return right ? weno_R(x[0], x[1], x[2], x[3], x[4]) : weno_L(x[0], x[1], x[2], x[3], x[4]);
}
void ah_weno_grad(const double* x, int right, double* g) { weno_grad(x, right != 0, g); }
void ah_weno_lr_vjp(const double* x, double sL, double sR, double* LR, double* g)
{
  weno_lr_vjp(x, sL, sR, LR[0], LR[1], g);
}

void ah_shear(int milne, double tau, const double* u, const double* pi10, const double* du16, double SH, double tpi,
              double* out)
{
  double du[4][4];
  for (int a = 0; a < 16; a++) du[a / 4][a % 4] = du16[a];
  shear_source<false>(milne, tau, u, pi10, du, SH, tpi, out);
}
void ah_shear_vjp(int milne, double tau, const double* u, const double* pi10, const double* du16, double SH, double tpi,
                  const double* ob, double* ub, double* pib, double* dub16, double* SHb, double* tpib)
{
  double du[4][4], dub[4][4];
  for (int a = 0; a < 16; a++) {
    du[a / 4][a % 4] = du16[a];
    dub[a / 4][a % 4] = 0;
  }
  shear_source_vjp(milne, tau, u, pi10, du, SH, tpi, ob, ub, pib, dub, *SHb, *tpib);
  for (int a = 0; a < 16; a++) dub16[a] += dub[a / 4][a % 4];
}

// forward of flux_cell_vjp: out = F(15), V(15), A
void ah_flux_cell(int d, double tau, double e, double P, double PI, const double* u, const double* pi10, const double* T,
                  double* out)
{
  const int row = d + 1;
  const double w = e + P - PI;
  const double pp = d == 2 ? (P - PI) / (tau * tau) : (P - PI);
  for (int l = 0; l < 4; l++) {
    double v = pi10[pidx(row, l)] + u[row] * u[l] * w;
    if (l == row) v = pi10[pidx(row, l)] + pp + u[row] * u[l] * w;
    out[l] = tau * v;
    out[15 + l] = tau * T[l];
  }
  for (int l = 0; l < 10; l++) {
    out[4 + l] = (pi10[l] * u[row]) / u[0];
    out[15 + 4 + l] = pi10[l];
  }
  out[14] = (PI * u[row]) / u[0];
  out[29] = PI;
  out[30] = u[row] / u[0];
}
void ah_flux_cell_vjp(int d, double tau, double e, double P, double PI, const double* u, const double* pi10,
                      const double* ob /* 31 */, double* inb /* e,P,PI,u4,pi10,T4 = 21 */)
{
  double eb = 0, Pb = 0, PIb = 0, ub[4] = {0, 0, 0, 0}, pib[10] = {0}, Tb[4] = {0, 0, 0, 0};
  flux_cell_vjp(d, tau, e, P, PI, u, pi10, ob, ob + 15, ob[30], eb, Pb, PIb, ub, pib, Tb);
  inb[0] = eb;
  inb[1] = Pb;
  inb[2] = PIb;
  for (int l = 0; l < 4; l++) inb[3 + l] = ub[l];
  for (int l = 0; l < 10; l++) inb[7 + l] = pib[l];
  for (int l = 0; l < 4; l++) inb[17 + l] = Tb[l];
}

// forward of source_vjp: the 15 source rows of one Milne KT cell
void ah_source(int eos, int eta_law, double eta_coef, int tpi_law, double tpi_coef, double tau, double e, double P,
               double PI, const double* u, const double* pi10, const double* du16, double* S)
{
  Params p = make_params(eta_law, eta_coef, tpi_law, tpi_coef, eos);
  double du[4][4];
  for (int a = 0; a < 16; a++) du[a / 4][a % 4] = du16[a];
  const double tau2 = tau * tau;
  S[0] = -(P - PI) - tau2 * (pi10[9] + (e + P - PI) * u[3] * u[3]);
  S[3] = (-2.0 * (pi10[3] + (e + P - PI) * u[0] * u[3]));
  S[1] = S[2] = S[14] = 0;
  const double T = eos_T(p, e);
  const double s = eos_s(p, e, P, T);
  const double SH = eta_of(p, s);
  const double tpi = tpi_of(p, SH, P, e);
  shear_source<false>(1, tau, u, pi10, du, SH, tpi, S + 4);
}
void ah_source_vjp(int eos, int eta_law, double eta_coef, int tpi_law, double tpi_coef, double tau, double e,
                   double P, double PI, const double* u, const double* pi10, const double* du16, const double* Sb,
                   double* inb /* e,P,PI,u4,pi10,du16,ceta,ctpi = 35 */)
{
  Params p = make_params(eta_law, eta_coef, tpi_law, tpi_coef, eos);
  double du[4][4], dub[4][4];
  for (int a = 0; a < 16; a++) {
    du[a / 4][a % 4] = du16[a];
    dub[a / 4][a % 4] = 0;
  }
  double eb = 0, Pb = 0, PIb = 0, ub[4] = {0, 0, 0, 0}, pib[10] = {0}, ceb = 0, ctb = 0;
  source_vjp(p, tau, e, P, PI, u, pi10, du, Sb, eb, Pb, PIb, ub, pib, dub, ceb, ctb);
  inb[0] = eb;
  inb[1] = Pb;
  inb[2] = PIb;
  for (int l = 0; l < 4; l++) inb[3 + l] = ub[l];
  for (int l = 0; l < 10; l++) inb[7 + l] = pib[l];
  for (int a = 0; a < 16; a++) inb[17 + a] = dub[a / 4][a % 4];
  inb[33] = ceb;
  inb[34] = ctb;
}

// forward of update_vjp: from V (15) at tn, the root (from guess) and
// outputs e, P, Temp, v(3), u(4), T0(4), pi(10), PI = 25 values
int ah_update(int eos, double tn, const double* V, double guess, double* out)
{
  Params p = make_params(0, 0.2, 0, 5.0, eos);
  double T0[4];
  for (int l = 0; l < 4; l++) T0[l] = V[l] / tn;
  double pi[10];
  for (int l = 0; l < 10; l++) pi[l] = V[4 + l];
  const double PIn = V[14];
  const double M2 = (T0[1] - pi[1]) * (T0[1] - pi[1]) + (T0[2] - pi[2]) * (T0[2] - pi[2]) +
                    tn * tn * (T0[3] - pi[3]) * (T0[3] - pi[3]);
  int it;
  unsigned fl;
  const double en = newton_cell(p, M2, T0[0], pi[0], PIn, guess, it, fl);
  const double Pn = eos_p(p, en);
  const double vx = (T0[1] - pi[1]) / (T0[0] - pi[0] + Pn - PIn);
  const double vy = (T0[2] - pi[2]) / (T0[0] - pi[0] + Pn - PIn);
  const double ve = (T0[3] - pi[3]) / (T0[0] - pi[0] + Pn - PIn);
  const double gm = 1.0 / sqrt(1 - vx * vx - vy * vy - tn * tn * ve * ve);
  out[0] = en;
  out[1] = Pn;
  out[2] = eos_T(p, en);
  out[3] = vx;
  out[4] = vy;
  out[5] = ve;
  out[6] = gm;
  out[7] = gm * vx;
  out[8] = gm * vy;
  out[9] = gm * ve;
  for (int l = 0; l < 4; l++) out[10 + l] = T0[l];
  for (int l = 0; l < 10; l++) out[14 + l] = pi[l];
  out[24] = PIn;
  return (int)fl;
}
void ah_update_vjp(int eos, double tn, const double* V, double e, const double* ob /* 25 */, double* Vb)
{
  Params p = make_params(0, 0.2, 0, 5.0, eos);
  double T0[4], pi[10];
  for (int l = 0; l < 4; l++) T0[l] = V[l] / tn;
  for (int l = 0; l < 10; l++) pi[l] = V[4 + l];
  update_vjp(p, tn, T0, pi, V[14], e, ob[0], ob[1], ob[2], ob + 3, ob + 6, ob + 10, ob + 14, ob[24], Vb);
}

// ---- Cartesian (TXYZ5)
void ah_cart_pi10(const double* u, const double* pe, double* out)
{
  cart_pi10(u, pe[0], pe[1], pe[2], pe[3], pe[4], out);
}
void ah_cart_pi10_vjp(const double* u, const double* pe, const double* ob, double* inb /* u4, pe5 */)
{
  double ub[4] = {0, 0, 0, 0}, peb[5] = {0, 0, 0, 0, 0};
  cart_pi10_vjp(u, pe, ob, ub, peb);
  for (int l = 0; l < 4; l++) inb[l] = ub[l];
  for (int l = 0; l < 5; l++) inb[4 + l] = peb[l];
}
// forward of cart_update_vjp: from prm (10) the gnewton root started at
// guess (4); outputs e, P, Temp, v(3), u(4), T(4), pe(5), PI = 20 values;
// returns the solver flags
int ah_cart_update(int eos, const double* prm, const double* guess, double* out)
{
  Params p = make_params(0, 0.2, 0, 5.0, eos);
  p.milne = 0;
  double x[4] = {guess[0], guess[1], guess[2], guess[3]};
  unsigned fl;
  cart_gnewton(p, prm, x, fl);
  const double c0 = sqrt(1 + x[1] * x[1] + x[2] * x[2] + x[3] * x[3]);
  out[0] = x[0];
  out[1] = eos_p(p, x[0]);
  out[2] = eos_T(p, x[0]);
  for (int l = 0; l < 3; l++) out[3 + l] = x[1 + l] / c0;
  out[6] = c0;
  for (int l = 0; l < 3; l++) out[7 + l] = x[1 + l];
  for (int l = 0; l < 4; l++) out[10 + l] = prm[l];
  for (int l = 0; l < 5; l++) out[14 + l] = prm[4 + l];
  out[19] = prm[9];
  return (int)fl;
}
void ah_cart_update_vjp(int eos, const double* prm, const double* x, const double* ob /* 20 */, double* Vb)
{
  Params p = make_params(0, 0.2, 0, 5.0, eos);
  p.milne = 0;
  cart_update_vjp(p, prm, x, ob + 10, ob + 14, ob[19], ob[0], ob[1], ob[2], ob + 3, ob + 6, Vb);
}
// forward of source_vjp_cart: the Cartesian source rows 4..8 of one cell
void ah_source_cart(int eos, int eta_law, double eta_coef, int tpi_law, double tpi_coef, double e, double P,
                    const double* u, const double* pe, const double* du16, double* S5)
{
  Params p = make_params(eta_law, eta_coef, tpi_law, tpi_coef, eos);
  p.milne = 0;
  double du[4][4], pi10[10], out[10];
  for (int a = 0; a < 16; a++) du[a / 4][a % 4] = du16[a];
  cart_pi10(u, pe[0], pe[1], pe[2], pe[3], pe[4], pi10);
  const double T = eos_T(p, e);
  const double s = eos_s(p, e, P, T);
  const double SH = eta_of(p, s);
  const double tpi = tpi_of(p, SH, P, e);
  shear_source<false>(0, 1.0, u, pi10, du, SH, tpi, out);
  for (int l = 0; l < 5; l++) S5[l] = out[cart5(l)];
}
void ah_source_cart_vjp(int eos, int eta_law, double eta_coef, int tpi_law, double tpi_coef, double e, double P,
                        const double* u, const double* pe, const double* du16, const double* Sb5,
                        double* inb /* e, P, u4, pe5, du16, ceta, ctpi = 29 */)
{
  Params p = make_params(eta_law, eta_coef, tpi_law, tpi_coef, eos);
  p.milne = 0;
  double du[4][4], dub[4][4], Sb[15] = {0};
  for (int a = 0; a < 16; a++) {
    du[a / 4][a % 4] = du16[a];
    dub[a / 4][a % 4] = 0;
  }
  for (int l = 0; l < 5; l++) Sb[4 + l] = Sb5[l];
  double eb = 0, Pb = 0, ub[4] = {0, 0, 0, 0}, peb[5] = {0, 0, 0, 0, 0}, ceb = 0, ctb = 0;
  source_vjp_cart(p, e, P, u, pe, true, du, Sb, eb, Pb, ub, peb, dub, ceb, ctb);
  inb[0] = eb;
  inb[1] = Pb;
  for (int l = 0; l < 4; l++) inb[2 + l] = ub[l];
  for (int l = 0; l < 5; l++) inb[6 + l] = peb[l];
  for (int a = 0; a < 16; a++) inb[11 + a] = dub[a / 4][a % 4];
  inb[27] = ceb;
  inb[28] = ctb;
}

// ---- bulk source: forward (as the step evaluates it) and reverse
static Params bulk_params(int eos, int zeta_law, double zeta_coef, int tpib_law, double tpib_coef, int milne)
{
  Params p = make_params(0, 0.2, 0, 5.0, eos);
  p.milne = milne;
  p.bulk = 1;
  p.zeta_law = zeta_law;
  p.zeta_coef = zeta_coef;
  p.tpib_law = tpib_law;
  p.tpib_coef = tpib_coef;
  return p;
}
double ah_bulk(int eos, int zeta_law, double zeta_coef, int tpib_law, double tpib_coef, int milne, double tau,
               double e, double P, const double* u, double PI, const double* du16)
{
  Params p = bulk_params(eos, zeta_law, zeta_coef, tpib_law, tpib_coef, milne);
  double du[4][4];
  for (int a = 0; a < 16; a++) du[a / 4][a % 4] = du16[a];
  const double T = eos_T(p, e);
  const double s = eos_s(p, e, P, T);
  const double BU = zeta_of(p, s);
  return bulk_source<false>(milne, tau, u, PI, du, BU, tpib_of(p, BU, P));
}
void ah_bulk_vjp(int eos, int zeta_law, double zeta_coef, int tpib_law, double tpib_coef, int milne, double tau,
                 double e, double P, const double* u, double PI, const double* du16, double Sb,
                 double* inb /* e, P, PI, u4, du16 = 23 */)
{
  Params p = bulk_params(eos, zeta_law, zeta_coef, tpib_law, tpib_coef, milne);
  double du[4][4], dub[4][4];
  for (int a = 0; a < 16; a++) {
    du[a / 4][a % 4] = du16[a];
    dub[a / 4][a % 4] = 0;
  }
  double eb = 0, Pb = 0, PIb = 0, ub[4] = {0, 0, 0, 0};
  bulk_vjp(p, milne, tau, e, P, u, PI, du, Sb, eb, Pb, PIb, ub, dub);
  inb[0] = eb;
  inb[1] = Pb;
  inb[2] = PIb;
  for (int l = 0; l < 4; l++) inb[3 + l] = ub[l];
  for (int a = 0; a < 16; a++) inb[7 + a] = dub[a / 4][a % 4];
}

}  // extern "C"
