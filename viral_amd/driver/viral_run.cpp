/*
 * viral_run -- host driver of a hydro evolution over the C ABI of libviral
 * (include/viral.h).  It keeps the flow of the reference's main program
 * (NEWEST10/main.cpp:87-207) and init (NEWEST10/init.h:548-698):
 *
 *   initial state -> DebugMSG / MaxTempGev / CheckRoot report -> first output
 *   loop: CheckPhysics, step, DebugMSG (inside the reference's TVDRK*,
 *         NEWEST10/time.h:303,334 ...), tau += dtau, CheckRoot, MaxTempGev
 *         line, output every print interval (and after the first step), stop
 *         rules
 *
 * with the device doing the step, the reductions and the packing of the
 * output records; the host only writes the files.  The compile-time profile
 * macros of NEWEST10/maindef.h become --profile presets that options
 * override.  Initial states (the reference's init branches):
 *   initvar  e = 16 exp(-r^2) with the eta plateau cut-off, at rest
 *            (NEWEST10/init.h:150-214), the default profile
 *   gubser   viscous Gubser flow (NEWEST10/init.h:421-541): T and the shear
 *            components from the file the reference reads (--gubser-file,
 *            8 doubles per ghosted cell, rows of y) or, by default, from the
 *            semi-analytic solution of the Gubser-frame equations; velocities
 *            and prevu from the Gubser flow as init.h:470-530 sets them
 *   glauber  optical Glauber overlap with the plateau eta profile (the shape
 *            of NEWEST10/ginit.h's input; ginit itself needs GSL)
 *   bjorken  uniform e = 30 GeV/fm^3 at rest (NEWEST10/init.h:293-353; the
 *            BULKTEST profile's initBulktest, init.h:356-413, is the same state)
 *
 * Output files: <out>/tau%2.3ffm.bin, the reference's records (write.h).  The
 * bjorken and bulktest profiles also write temp.txt (main.cpp:124-141): one
 * line "tau  value" before every step, value the centre cell's
 * (HydroGrid[XCM/2][YCM/2][ZCMA/2]) temperature in GeV (BJORKEN) or bulk
 * pressure PI (BULKTEST), in the stream's default 6-digit format.
 */
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "../../include/viral.h"

namespace {

constexpr double GEVFM = 0.1973;  // NEWEST10/main.cpp:44
constexpr double PIE = 3.141592653589793;
constexpr int B = VIRAL_BORDER;
constexpr int NF = VIRAL_NFIELDS;

struct Options {
  std::string profile = "default";
  viral_desc d{};
  double tau0 = 0.6, ts = 0.002;  // TAUSTART, TS
  double pfreq = 0.01;            // PFREQ
  int freq = 1, freqz = 1;        // FREQ, FREQZ
  std::string init = "initvar";
  std::string gubser_file;
  double gubser_rho0 = -4.5, gubser_that0 = 0.1;
  double b = 7.0, e_peak_gev = 30.0, plateau = 1.5, falloff = 1.0;
  std::string out = "res";
  int max_steps = -1;        // < 0: run to the profile's stop rule
  double tau_stop = 25.3;    // main.cpp:203
  double tau_stop_profile = 0;  // GUBSER: 9 (main.cpp:189)
  double tmin_mev = 0;       // GINIT: 120 MeV (main.cpp:195)
  int checks = 1;            // CheckRoot / DebugMSG / CheckPhysics lines
  int write = 1;
  std::string dump_state;    // write the initial 29-field state here
  int dry_run = 0;           // initial state and first output only, no device
  int temp_kind = 0;         // temp.txt: 0 none, 1 centre T in GeV (BJORKEN), 2 centre PI (BULKTEST)
  std::string temp_file = "temp.txt";
};

void usage()
{
  std::printf(
      "usage: viral_run [--profile default|gubser|ginit|bjorken|bulktest] [options]\n"
      "  grid:      --nx N --ny N --nz N --lbi 0|1 --dx F --dy F --deta F --coords milne|cartesian\n"
      "  step:      --rk 1|2|3 --scheme kt|shasta --fast 0|1 --eos conformal|s95p --eos-table 0|1\n"
      "             --eta-law times_s|s_over_4pi --eta-coef F --tpi-law over_ep|over_p|const --tpi-coef F\n"
      "             --quirks BITS --fix 0|1 --max-iter N\n"
      "             --bulk 0|1 --zeta-law const|s_over_4pi --zeta-coef F --tpib-law const|over_p --tpib-coef F\n"
      "  time:      --tau0 F --dtau F --max-steps N --tau-stop F --tmin-mev F\n"
      "  init:      --init initvar|gubser|glauber|bjorken --gubser-file PATH --gubser-rho0 F --gubser-that0 F\n"
      "             --b F --e-peak F --plateau F --falloff F\n"
      "  output:    --out DIR --pfreq F --freq N --freqz N --write 0|1 --checks 0|1\n"
      "             --dump-state PATH --dry-run --temp-file PATH --temp none|t|pi\n");
}

// maindef.h profiles as presets
void set_profile(Options& o, const std::string& name)
{
  o.profile = name;
  viral_desc& d = o.d;
  if (name == "default") {  // NEWEST10/maindef.h:36-87 (one rank: XL=YL=4, ZL=5, XS=ZS=0.1)
    viral_desc_default(&d, 40, 40, 101);
    d.dx = d.dy = d.deta = 0.1;
    d.eta_law = VIRAL_ETA_S_OVER_4PI; d.eta_coef = 1.0;
    d.tpi_law = VIRAL_TPI_OVER_P; d.tpi_coef = 1.5;
    o.tau0 = 0.6; o.ts = 0.002; o.pfreq = 0.01; o.freq = 1; o.freqz = 1;
    o.init = "initvar";
  } else if (name == "gubser") {  // :298-339, all NPX x NPY = 6 x 6 ranks' cells on one device
    viral_desc_default(&d, 600, 600, 1);
    d.lbi = 1;
    d.dx = d.dy = 0.05; d.deta = 0.1;
    d.eta_law = VIRAL_ETA_TIMES_S; d.eta_coef = 0.2;
    d.tpi_law = VIRAL_TPI_OVER_EP; d.tpi_coef = 5.0;
    d.quirks |= VIRAL_Q_GUBSER_MIRROR;
    o.tau0 = 1.0; o.ts = 0.005; o.pfreq = 0.1; o.freq = 2; o.freqz = 1;
    o.init = "gubser";
    o.tau_stop_profile = 9.0;
  } else if (name == "ginit") {  // :236-294, NPX x NPY = 10 x 10 ranks of 30 x 30
    viral_desc_default(&d, 300, 300, 101);
    d.eos = VIRAL_EOS_S95P;
    d.dx = d.dy = d.deta = 0.1;
    d.eta_law = VIRAL_ETA_S_OVER_4PI; d.eta_coef = 1.0;
    d.tpi_law = VIRAL_TPI_OVER_P; d.tpi_coef = 1.5;
    o.tau0 = 0.1; o.ts = 0.002; o.pfreq = 0.05; o.freq = 5; o.freqz = 1;
    o.init = "glauber";
    o.tmin_mev = 120.0;
  } else if (name == "bjorken") {  // :197-232 (XL=YL=1, ZL=2, XS=ZS=0.1)
    viral_desc_default(&d, 10, 10, 41);
    d.dx = d.dy = d.deta = 0.1;
    d.eta_law = VIRAL_ETA_TIMES_S; d.eta_coef = 0.2;
    d.tpi_law = VIRAL_TPI_CONST; d.tpi_coef = 0.005;
    o.tau0 = 0.6; o.ts = 0.01; o.pfreq = 0.2; o.freq = 1; o.freqz = 1;
    o.init = "bjorken";
    o.tau_stop_profile = 10.0;
    o.temp_kind = 1;
  } else if (name == "bulktest") {  // :157-196, the bjorken grid with -DBULK (zeta = 1, tau_PI = 0.1)
    set_profile(o, "bjorken");
    o.profile = name;
    d.bulk = 1;
    d.zeta_law = VIRAL_ZETA_CONST; d.zeta_coef = 1.0;
    d.tpib_law = VIRAL_TPIB_CONST; d.tpib_coef = 0.1;
    o.pfreq = 0.1; o.freq = 2; o.freqz = 1;
    o.temp_kind = 2;
  } else {
    std::fprintf(stderr, "unknown profile %s\n", name.c_str());
    std::exit(2);
  }
}

bool parse(int argc, char** argv, Options& o)
{
  set_profile(o, "default");
  // the profile first, so that other options override it
  for (int a = 1; a + 1 < argc; a++)
    if (!std::strcmp(argv[a], "--profile")) set_profile(o, argv[a + 1]);
  viral_desc& d = o.d;
  for (int a = 1; a < argc; a++) {
    std::string k = argv[a];
    if (k == "--dry-run") { o.dry_run = 1; continue; }
    if (k == "--help" || k == "-h") { usage(); std::exit(0); }
    if (a + 1 >= argc) { std::fprintf(stderr, "option %s needs a value\n", k.c_str()); return false; }
    std::string v = argv[++a];
    const double f = std::atof(v.c_str());
    const int n = std::atoi(v.c_str());
    if (k == "--profile") {}
    else if (k == "--nx") d.nx = n;
    else if (k == "--ny") d.ny = n;
    else if (k == "--nz") d.nz = n;
    else if (k == "--lbi") d.lbi = n;
    else if (k == "--dx") d.dx = f;
    else if (k == "--dy") d.dy = f;
    else if (k == "--deta") d.deta = f;
    else if (k == "--coords") d.coords = v == "cartesian" ? VIRAL_CARTESIAN : VIRAL_MILNE;
    else if (k == "--rk") d.rk = n;
    else if (k == "--scheme") d.scheme = v == "shasta" ? VIRAL_SHASTA : VIRAL_KT;
    else if (k == "--fast") d.fast_math = n;
    else if (k == "--eos") d.eos = v == "s95p" ? VIRAL_EOS_S95P : VIRAL_EOS_CONFORMAL;
    else if (k == "--eos-table") d.eos_table = n;
    else if (k == "--eta-law") d.eta_law = v == "times_s" ? VIRAL_ETA_TIMES_S : VIRAL_ETA_S_OVER_4PI;
    else if (k == "--eta-coef") d.eta_coef = f;
    else if (k == "--tpi-law") d.tpi_law = v == "over_ep" ? VIRAL_TPI_OVER_EP : v == "const" ? VIRAL_TPI_CONST : VIRAL_TPI_OVER_P;
    else if (k == "--tpi-coef") d.tpi_coef = f;
    else if (k == "--quirks") d.quirks = (unsigned)std::strtoul(v.c_str(), nullptr, 0);
    else if (k == "--fix") d.quirks = n ? (d.quirks | VIRAL_Q_FIX_VE) : (d.quirks & ~VIRAL_Q_FIX_VE);  // -DFIX
    else if (k == "--max-iter") d.max_iter = n;
    else if (k == "--bulk") d.bulk = n;
    else if (k == "--zeta-law") d.zeta_law = v == "const" ? VIRAL_ZETA_CONST : VIRAL_ZETA_S_OVER_4PI;
    else if (k == "--zeta-coef") d.zeta_coef = f;
    else if (k == "--tpib-law") d.tpib_law = v == "const" ? VIRAL_TPIB_CONST : VIRAL_TPIB_OVER_P;
    else if (k == "--tpib-coef") d.tpib_coef = f;
    else if (k == "--temp-file") o.temp_file = v;
    else if (k == "--temp") o.temp_kind = v == "t" ? 1 : v == "pi" ? 2 : 0;
    else if (k == "--tau0") o.tau0 = f;
    else if (k == "--dtau") o.ts = f;
    else if (k == "--max-steps") o.max_steps = n;
    else if (k == "--tau-stop") o.tau_stop = f;
    else if (k == "--tmin-mev") o.tmin_mev = f;
    else if (k == "--init") o.init = v;
    else if (k == "--gubser-file") o.gubser_file = v;
    else if (k == "--gubser-rho0") o.gubser_rho0 = f;
    else if (k == "--gubser-that0") o.gubser_that0 = f;
    else if (k == "--b") o.b = f;
    else if (k == "--e-peak") o.e_peak_gev = f;
    else if (k == "--plateau") o.plateau = f;
    else if (k == "--falloff") o.falloff = f;
    else if (k == "--out") o.out = v;
    else if (k == "--pfreq") o.pfreq = f;
    else if (k == "--freq") o.freq = n;
    else if (k == "--freqz") o.freqz = n;
    else if (k == "--write") o.write = n;
    else if (k == "--checks") o.checks = n;
    else if (k == "--dump-state") o.dump_state = v;
    else { std::fprintf(stderr, "unknown option %s\n", k.c_str()); usage(); return false; }
  }
  if (d.lbi) d.nz = 1;
  return true;
}

// ------------------------------------------------------------------ grid

struct Grid {
  int gx, gy, gz, il, ir, jl, jr, kl, kr, k0;
  size_t n;
  std::vector<double> X, Y, E;  // cell positions per index (NEWEST10/init.h:44-81)
  size_t at(int i, int j, int k) const { return ((size_t)i * gy + j) * gz + k; }
};

Grid make_grid(const viral_desc& d)
{
  Grid g;
  g.gx = d.nx + 2 * B;
  g.gy = d.ny + 2 * B;
  g.gz = d.lbi ? 1 : d.nz + 2 * B;
  g.il = B; g.ir = B + d.nx;
  g.jl = B; g.jr = B + d.ny;
  if (d.lbi) { g.kl = 0; g.kr = 1; g.k0 = 0; }
  else { g.kl = B; g.kr = B + d.nz; g.k0 = d.k0 >= 0 ? d.k0 : (g.gz - 1) / 2; }
  g.n = (size_t)g.gx * g.gy * g.gz;
  // XSTART = -GRIDXMAX = -(points/2)*XS, positions continue into the ghosts
  const double x0 = -(double)(d.nx / 2) * d.dx, y0 = -(double)(d.ny / 2) * d.dy;
  const double z0 = d.lbi ? 0.0 : -(double)(d.nz / 2) * d.deta;
  for (int i = 0; i < g.gx; i++) g.X.push_back(x0 + (i - B) * d.dx);
  for (int j = 0; j < g.gy; j++) g.Y.push_back(y0 + (j - B) * d.dy);
  for (int k = 0; k < g.gz; k++) g.E.push_back(d.lbi ? 0.0 : z0 + (k - B) * d.deta);
  return g;
}

typedef std::vector<double> State;  // NF * n, field-major
inline double& F(State& s, const Grid& g, int f, size_t x) { return s[(size_t)f * g.n + x]; }

// P and Temp from the library's EOS (the reference's EOS/FT profile inlines)
void eos_fill(const viral_desc& d, const Grid& g, State& s)
{
  std::vector<double> P(g.n), a(g.n), T(g.n), S(g.n);
  viral_eos_eval(&d, g.n, &s[(size_t)VIRAL_F_EN * g.n], P.data(), a.data(), T.data(), S.data());
  for (size_t x = 0; x < g.n; x++) {
    F(s, g, VIRAL_F_P, x) = P[x];
    F(s, g, VIRAL_F_TEMP, x) = T[x];
  }
}

// T^{tau mu} from (e, P, PI, u, pi), NEWEST10/init.h:205-209 (and 536-540)
void assemble_T(const Grid& g, State& s)
{
  for (size_t x = 0; x < g.n; x++) {
    const double e = F(s, g, VIRAL_F_EN, x), P = F(s, g, VIRAL_F_P, x), PI = F(s, g, VIRAL_F_BULK, x);
    const double u0 = F(s, g, VIRAL_F_U0, x), u1 = F(s, g, VIRAL_F_U1, x), u2 = F(s, g, VIRAL_F_U2, x),
                 u3 = F(s, g, VIRAL_F_U3, x);
    F(s, g, VIRAL_F_T00, x) = -P + PI + (e + P - PI) * std::pow(u0, 2) + F(s, g, VIRAL_F_PI0, x);
    F(s, g, VIRAL_F_T10, x) = (e + P - PI) * u0 * u1 + F(s, g, VIRAL_F_PI0 + 1, x);
    F(s, g, VIRAL_F_T20, x) = (e + P - PI) * u0 * u2 + F(s, g, VIRAL_F_PI0 + 2, x);
    F(s, g, VIRAL_F_T30, x) = (e + P - PI) * u0 * u3 + F(s, g, VIRAL_F_PI0 + 3, x);
  }
}

// fluid at rest: v = 0, u = prevu = (1/sqrt(1 - v^2), 0, 0, 0) as init.h:177-193
void at_rest(const Grid& g, State& s, double tau)
{
  for (size_t x = 0; x < g.n; x++) {
    const double vx = 0, vy = 0, ve = 0;
    F(s, g, VIRAL_F_VX, x) = vx; F(s, g, VIRAL_F_VY, x) = vy; F(s, g, VIRAL_F_VE, x) = ve;
    const double u0 = 1.0 / (std::sqrt(1 - vx * vx - vy * vy - tau * tau * ve * ve));
    F(s, g, VIRAL_F_U0, x) = u0;
    F(s, g, VIRAL_F_U1, x) = u0 * vx;
    F(s, g, VIRAL_F_U2, x) = u0 * vy;
    F(s, g, VIRAL_F_U3, x) = u0 * ve;
    for (int m = 0; m < 4; m++) F(s, g, VIRAL_F_PU0 + m, x) = F(s, g, VIRAL_F_U0 + m, x);
  }
}

// NEWEST10/init.h:150-214
void init_initvar(const viral_desc& d, const Grid& g, State& s, double tau)
{
  for (int i = 0; i < g.gx; i++)
    for (int j = 0; j < g.gy; j++)
      for (int k = 0; k < g.gz; k++) {
        const size_t x = g.at(i, j, k);
        const double X = g.X[i], Y = g.Y[j], eta = g.E[k];
        const double r = std::sqrt(X * X + Y * Y);
        double en = 16 * std::exp(-r * r);
        const double AbsE = std::fabs(eta), EtaF = 4, sig = 1;
        const double arg = (AbsE - (EtaF / 2));
        const double cutoff = std::exp(-std::pow(arg / (std::sqrt(2) * sig), 2) * (arg >= 0 ? 1 : 0));
        en *= cutoff;
        F(s, g, VIRAL_F_EN, x) = en;
      }
  eos_fill(d, g, s);
  at_rest(g, s, tau);
  assemble_T(g, s);
}

// NEWEST10/init.h:293-353: e = 30 GeV/fm^3 everywhere, at rest
void init_bjorken(const viral_desc& d, const Grid& g, State& s, double tau)
{
  for (size_t x = 0; x < g.n; x++) F(s, g, VIRAL_F_EN, x) = 30 / GEVFM;
  eos_fill(d, g, s);
  at_rest(g, s, tau);
  assemble_T(g, s);
}

// Woods-Saxon thickness, rho0 = 1 (shape only): Riemann sum over z in [-20, 20]
double thickness(double x, double y, double R, double a)
{
  const int nz = 801;
  const double zmax = 20.0, dz = 2 * zmax / (nz - 1);
  double t = 0;
  for (int q = 0; q < nz; q++) {
    const double z = -zmax + q * dz;
    const double r = std::sqrt(x * x + y * y + z * z);
    t += 1.0 / (1.0 + std::exp((r - R) / a));
  }
  return t * dz;
}

// plateau of half-width `plateau` with Gaussian tails of width `falloff`
double eta_profile(double eta, double plateau, double falloff)
{
  const double over = std::fmax(std::fabs(eta) - plateau, 0.0);
  return std::exp(-over * over / (2.0 * falloff * falloff));
}

// optical Glauber T_A(x+b/2, y) T_B(x-b/2, y) x H(eta), peak e_peak GeV/fm^3, at rest
void init_glauber(const Options& o, const viral_desc& d, const Grid& g, State& s, double tau)
{
  const double R = 6.62, a = 0.546;
  std::vector<double> tr((size_t)g.gx * g.gy);
  double mx = 0;
  for (int i = 0; i < g.gx; i++)
    for (int j = 0; j < g.gy; j++) {
      const double v = thickness(g.X[i] + o.b / 2.0, g.Y[j], R, a) * thickness(g.X[i] - o.b / 2.0, g.Y[j], R, a);
      tr[(size_t)i * g.gy + j] = v;
      mx = std::fmax(mx, v);
    }
  double emax = 0;
  for (int i = 0; i < g.gx; i++)
    for (int j = 0; j < g.gy; j++)
      for (int k = 0; k < g.gz; k++) {
        const double e = (o.e_peak_gev / GEVFM) * (tr[(size_t)i * g.gy + j] / mx) * eta_profile(g.E[k], o.plateau, o.falloff);
        F(s, g, VIRAL_F_EN, g.at(i, j, k)) = e;
        emax = std::fmax(emax, e);
      }
  for (size_t x = 0; x < g.n; x++) F(s, g, VIRAL_F_EN, x) = std::fmax(F(s, g, VIRAL_F_EN, x), 1e-7 * emax);
  eos_fill(d, g, s);
  at_rest(g, s, tau);
  assemble_T(g, s);
}

// ------------------------------------------------------------------ Gubser flow

// Viscous Gubser flow of conformal Israel-Stewart hydro with eta = c s,
// tau_pi = 5 eta / (e + p) (NEWEST10/maindef.h:328-329): in de Sitter time rho
//   dT/drho = T tanh(rho) (p/3 - 2/3)
//   dp/drho = -p T / (5 c) + tanh(rho) (4/15 - 4/3 p^2)
// for T = tau * temperature and p = the eta-eta shear over (e + P)
// (Marrochio et al., Phys. Rev. C 91, 014903), integrated forward from rho0
// by classical RK4 and evaluated by cubic Hermite interpolation.
struct GubserODE {
  double c, rho0, h;
  std::vector<double> T, p;
  void rhs(double rho, double t, double q, double& dt, double& dq) const
  {
    const double th = std::tanh(rho);
    dt = t * th * (q / 3.0 - 2.0 / 3.0);
    dq = -q * t / (5.0 * c) + th * (4.0 / 15.0 - 4.0 / 3.0 * q * q);
  }
  GubserODE(double c_, double rho0_, double that0, double rho_max) : c(c_), rho0(rho0_), h(1e-4)
  {
    const int n = (int)std::ceil((rho_max - rho0) / h) + 2;
    T.resize(n); p.resize(n);
    T[0] = that0; p[0] = 0.0;
    for (int s = 0; s + 1 < n; s++) {
      const double r = rho0 + s * h, t = T[s], q = p[s];
      double k1t, k1q, k2t, k2q, k3t, k3q, k4t, k4q;
      rhs(r, t, q, k1t, k1q);
      rhs(r + h / 2, t + h / 2 * k1t, q + h / 2 * k1q, k2t, k2q);
      rhs(r + h / 2, t + h / 2 * k2t, q + h / 2 * k2q, k3t, k3q);
      rhs(r + h, t + h * k3t, q + h * k3q, k4t, k4q);
      T[s + 1] = t + h / 6 * (k1t + 2 * k2t + 2 * k3t + k4t);
      p[s + 1] = q + h / 6 * (k1q + 2 * k2q + 2 * k3q + k4q);
    }
  }
  bool eval(double rho, double& t, double& q) const
  {
    const double u = (rho - rho0) / h;
    const int s = (int)std::floor(u);
    if (s < 0 || s + 1 >= (int)T.size()) return false;
    const double w = u - s;
    double d0t, d0q, d1t, d1q;
    rhs(rho0 + s * h, T[s], p[s], d0t, d0q);
    rhs(rho0 + (s + 1) * h, T[s + 1], p[s + 1], d1t, d1q);
    const double h00 = (1 + 2 * w) * (1 - w) * (1 - w), h10 = w * (1 - w) * (1 - w), h01 = w * w * (3 - 2 * w),
                 h11 = w * w * (w - 1);
    t = h00 * T[s] + h10 * h * d0t + h01 * T[s + 1] + h11 * h * d1t;
    q = h00 * p[s] + h10 * h * d0q + h01 * p[s + 1] + h11 * h * d1q;
    return true;
  }
};

// NEWEST10/init.h:421-541, q = 1/fm
bool init_gubser(const Options& o, const viral_desc& d, const Grid& g, State& s, double tau, double ts)
{
  const double q = 1;
  const double factor = 3.0 * d.dof * PIE * PIE / 90.0;
  if (!o.gubser_file.empty()) {
    // 8 doubles per ghosted cell: T, pi^tt, pi^tx, pi^ty, pi^xx, pi^xy, pi^yy, pi^ee;
    // rows of y, each gx cells long (init.h:428-458 for one rank)
    FILE* fh = std::fopen(o.gubser_file.c_str(), "rb");
    if (!fh) { std::fprintf(stderr, "Could not read gubser temperatures\n"); return false; }
    std::vector<double> buf((size_t)8 * g.gx);
    for (int j = 0; j < g.gy; j++) {
      if (std::fread(buf.data(), sizeof(double), buf.size(), fh) != buf.size()) {
        std::fprintf(stderr, "Could not read gubser temperatures\n");
        std::fclose(fh);
        return false;
      }
      for (int i = 0; i < g.gx; i++) {
        const size_t x = g.at(i, j, g.k0);
        const double temp = buf[8 * i + 0];
        F(s, g, VIRAL_F_EN, x) = factor * std::pow(temp, 4);
        const int slot[7] = {0, 1, 2, 4, 5, 7, 9};
        for (int m = 0; m < 7; m++) F(s, g, VIRAL_F_PI0 + slot[m], x) = buf[8 * i + 1 + m];
      }
    }
    std::fclose(fh);
  } else {
    double rho_min = 0, rho_max = 0;
    for (int i = 0; i < g.gx; i++)
      for (int j = 0; j < g.gy; j++) {
        const double r = std::sqrt(g.X[i] * g.X[i] + g.Y[j] * g.Y[j]);
        const double rho = -std::asinh((1 - q * q * tau * tau + q * q * r * r) / (2 * q * tau));
        rho_min = std::fmin(rho_min, rho);
        rho_max = std::fmax(rho_max, rho);
      }
    // the solution through (rho0, That0) with pbar = 0 there; a grid that
    // reaches further back starts earlier, at That0 scaled as the ideal flow
    // scales (That ~ cosh(rho)^(-2/3))
    double rho0 = o.gubser_rho0, that0 = o.gubser_that0;
    if (rho_min < rho0) {
      const double r0 = rho_min - 0.05;
      that0 = that0 * std::pow(std::cosh(rho0) / std::cosh(r0), 2.0 / 3.0);
      rho0 = r0;
      std::printf("gubser: the grid reaches rho = %g; the solution starts at rho = %g with T_hat = %g\n", rho_min,
                  rho0, that0);
    }
    GubserODE ode(d.eta_coef, rho0, that0, rho_max + 0.01);
    for (int i = 0; i < g.gx; i++)
      for (int j = 0; j < g.gy; j++)
        for (int k = 0; k < g.gz; k++) {
          const size_t x = g.at(i, j, k);
          const double X = g.X[i], Y = g.Y[j], r = std::sqrt(X * X + Y * Y);
          const double rho = -std::asinh((1 - q * q * tau * tau + q * q * r * r) / (2 * q * tau));
          double That, pbar;
          ode.eval(rho, That, pbar);
          const double T = That / tau, e = factor * std::pow(T, 4), P = e / 3.0;
          const double kap = std::atanh((2 * q * q * tau * r) / (1 + q * q * tau * tau + q * q * r * r));
          const double u0 = std::cosh(kap), ur = std::sinh(kap);
          const double ux = r > 1e-7 ? X * ur / (r + 1e-15) : 0.0, uy = r > 1e-7 ? Y * ur / (r + 1e-15) : 0.0;
          // pi = 3/2 P_eta e_eta e_eta + 1/2 P_eta (g - u u), P_eta = pbar (e + P)
          const double hh = 0.5 * pbar * (e + P);
          F(s, g, VIRAL_F_EN, x) = e;
          double* pi = &F(s, g, VIRAL_F_PI0, x);
          const size_t n = g.n;
          pi[0] = hh * (1 - u0 * u0);
          pi[1 * n] = -hh * u0 * ux;
          pi[2 * n] = -hh * u0 * uy;
          pi[3 * n] = 0;
          pi[4 * n] = hh * (-1 - ux * ux);
          pi[5 * n] = -hh * ux * uy;
          pi[6 * n] = 0;
          pi[7 * n] = hh * (-1 - uy * uy);
          pi[8 * n] = 0;
          pi[9 * n] = 2 * hh / (tau * tau);
        }
  }
  // velocities and prevu of the Gubser flow (init.h:470-530)
  for (int i = 0; i < g.gx; i++)
    for (int j = 0; j < g.gy; j++)
      for (int k = 0; k < g.gz; k++) {
        const size_t x = g.at(i, j, k);
        const double X = g.X[i], Y = g.Y[j], r = std::sqrt(X * X + Y * Y);
        double kappa = std::atanh((2 * q * q * tau * r) / (1 + q * q * tau * tau + q * q * r * r));
        if (std::fabs(r) > 1e-7) {
          F(s, g, VIRAL_F_VX, x) = ((X * std::tanh(kappa)) / (r + 1e-15));
          F(s, g, VIRAL_F_VY, x) = ((Y * std::tanh(kappa)) / (r + 1e-15));
          F(s, g, VIRAL_F_U1, x) = ((X * std::sinh(kappa)) / (r + 1e-15));
          F(s, g, VIRAL_F_U2, x) = ((Y * std::sinh(kappa)) / (r + 1e-15));
        } else {
          F(s, g, VIRAL_F_VX, x) = F(s, g, VIRAL_F_VY, x) = 0;
          F(s, g, VIRAL_F_U1, x) = F(s, g, VIRAL_F_U2, x) = 0;
        }
        F(s, g, VIRAL_F_U0, x) = std::cosh(kappa);
        kappa = std::atanh((2 * q * q * (tau - ts) * r) / (1 + q * q * (tau - ts) * (tau - ts) + q * q * r * r));
        if (std::fabs(r) > 1e-7) {
          F(s, g, VIRAL_F_PU1, x) = ((X * std::sinh(kappa)) / (r + 1e-15));
          F(s, g, VIRAL_F_PU2, x) = ((Y * std::sinh(kappa)) / (r + 1e-15));
        } else {
          F(s, g, VIRAL_F_PU1, x) = F(s, g, VIRAL_F_PU2, x) = 0;
        }
        F(s, g, VIRAL_F_PU0, x) = std::cosh(kappa);
      }
  eos_fill(d, g, s);
  assemble_T(g, s);
  return true;
}

// ------------------------------------------------------------------ output

std::string out_name(const Options& o, double tau)
{
  char buf[64];
  std::snprintf(buf, sizeof buf, "tau%2.3ffm.bin", tau);  // NEWEST10/write.h:296-298
  return o.out + "/" + buf;
}

bool write_file(const std::string& path, const double* rec, size_t n)
{
  FILE* fh = std::fopen(path.c_str(), "wb");
  if (!fh) { std::fprintf(stderr, "cannot write %s\n", path.c_str()); return false; }
  const bool ok = std::fwrite(rec, sizeof(double), n, fh) == n;
  std::fclose(fh);
  return ok;
}

// host packing of the records of a host state (the initial state in a dry run)
void host_records(const Options& o, const viral_desc& d, const Grid& g, const State& s, std::vector<double>& rec)
{
  rec.clear();
  const bool xy = d.lbi;
  for (int i = g.il; i < g.ir; i += o.freq)
    for (int j = g.jl; j < g.jr; j += o.freq)
      for (int k = xy ? g.k0 : g.kl; xy ? k == g.k0 : k < g.kr; k += xy ? 1 : o.freqz) {
        const size_t x = g.at(i, j, k);
        const int fields[17] = {VIRAL_F_EN, VIRAL_F_TEMP, VIRAL_F_P, VIRAL_F_VX, VIRAL_F_VY, VIRAL_F_VE,
                                VIRAL_F_PI0, VIRAL_F_PI0 + 1, VIRAL_F_PI0 + 2, VIRAL_F_PI0 + 3, VIRAL_F_PI0 + 4,
                                VIRAL_F_PI0 + 5, VIRAL_F_PI0 + 6, VIRAL_F_PI0 + 7, VIRAL_F_PI0 + 8,
                                VIRAL_F_PI0 + 9, VIRAL_F_BULK};
        for (int q = 0; q < 17; q++) rec.push_back(s[(size_t)fields[q] * g.n + x]);
      }
}

// ------------------------------------------------------------------ reports

void report_checkroot(viral_ctx* ctx, const viral_desc& d, double tau)
{
  double mx[9], avg[9];
  int ns = 0;
  if (viral_check_root(ctx, 0, tau, mx, avg, nullptr, &ns) != VIRAL_OK) return;
  std::printf("\n\n");
  if (d.coords == VIRAL_MILNE) {  // NEWEST10/root.h:162-196
    const char* names[6] = {"ROOT CHECK", "ROW 0 TRANSVERSALITY", "ROW 1 TRANSVERSALITY", "ROW 2 TRANSVERSALITY",
                            "ROW 3 TRANSVERSALITY", "TRACE CHECK"};
    for (int s = 0; s < 6; s++) std::printf("%s -- MAX %e &  Avg %e\n", names[s], mx[s], avg[s]);
  } else {  // TXYZ5/multiroot.h:122-150
    for (int s = 0; s < 4; s++) std::printf("ROOT CHECK EQN %d -- MAX %e &     Avg %e\n", s + 1, mx[s], avg[s]);
    std::printf("TRACE -- MAX %e &    Avg %e\n", mx[8], avg[8]);
  }
}

void report_debug(viral_ctx* ctx, const viral_desc& d, const Grid& g, double tau)
{
  double out[6];
  int ijk[9];
  if (viral_debug_extrema(ctx, 0, tau, out, ijk) != VIRAL_OK) return;
  const double tz = d.coords == VIRAL_MILNE ? tau : 1.0;
  auto pos = [&](int q, double& X, double& Y, double& E) {
    X = ijk[3 * q] >= 0 ? g.X[ijk[3 * q]] : 0.0;
    Y = ijk[3 * q] >= 0 ? g.Y[ijk[3 * q + 1]] : 0.0;
    E = ijk[3 * q] >= 0 ? g.E[ijk[3 * q + 2]] : 0.0;
  };
  double X, Y, E;
  pos(0, X, Y, E);  // NEWEST10/debug.h:75-98
  std::printf("Max vel ( %f , %f , %f ) & mag %f @ ( %f , %f , %f )\n", out[1], out[2], tz * out[3], out[0], X, Y, E);
  pos(1, X, Y, E);
  std::printf("Max Energy ( %f ) @ ( %f , %f , %f )\n", out[4], X, Y, E);
  pos(2, X, Y, E);
  std::printf("Min Energy ( %f ) @ ( %f , %f , %f )\n", out[5], X, Y, E);
}

void report_physics(viral_ctx* ctx, const Grid& g, double tau)
{
  long long count[2];
  int ijk[6];
  if (viral_check_physics(ctx, 0, tau, count, ijk) != VIRAL_OK) return;
  if (count[0])
    std::printf("stage 0 Causality violated in %lld cells, first @ %f, %f, %f\n", count[0], g.X[ijk[0]], g.Y[ijk[1]],
                g.E[ijk[2]]);
  if (count[1])
    std::printf("stage 0 Energy Negative in %lld cells, first @ %f, %f, %f\n", count[1], g.X[ijk[3]], g.Y[ijk[4]],
                g.E[ijk[5]]);
}

}  // namespace

int main(int argc, char** argv)
{
  Options o;
  if (!parse(argc, argv, o)) return 2;
  viral_desc& d = o.d;
  const Grid g = make_grid(d);
  double tau = o.tau0;
  const double ts = o.ts;

  // banner (NEWEST10/init.h:617-628 for one rank)
  std::printf("profile %s: points (%d , %d , %d), spacing (%g , %g , %g), k0 %d, tau0 %g, dtau %g, RK%d %s %s\n",
              o.profile.c_str(), d.nx, d.ny, d.lbi ? 1 : d.nz, d.dx, d.dy, d.deta, g.k0, tau, ts, d.rk,
              d.scheme == VIRAL_SHASTA ? "SHASTA" : "KT", d.coords == VIRAL_MILNE ? "Milne" : "Cartesian");

  State st((size_t)NF * g.n, 0.0);
  if (o.init == "initvar") init_initvar(d, g, st, tau);
  else if (o.init == "bjorken") init_bjorken(d, g, st, tau);
  else if (o.init == "glauber") init_glauber(o, d, g, st, tau);
  else if (o.init == "gubser") { if (!init_gubser(o, d, g, st, tau, ts)) return 1; }
  else { std::fprintf(stderr, "unknown init %s\n", o.init.c_str()); return 2; }

  if (!o.dump_state.empty() && !write_file(o.dump_state, st.data(), st.size())) return 1;

  size_t nrec = 0;
  if (viral_output_size(&d, o.freq, o.freqz, d.lbi, &nrec) != VIRAL_OK) {
    std::fprintf(stderr, "output strides %d/%d do not divide the grid\n", o.freq, o.freqz);
    return 2;
  }
  if (o.dry_run) {
    std::vector<double> rec;
    host_records(o, d, g, st, rec);
    if (o.write && !write_file(out_name(o, tau), rec.data(), rec.size())) return 1;
    return 0;
  }

  viral_ctx* ctx = nullptr;
  if (viral_create(&d, 1, &ctx) != VIRAL_OK) {
    std::fprintf(stderr, "viral_create: %s\n", viral_last_error(ctx));
    return 1;
  }
  std::vector<const double*> fields(NF);
  for (int f = 0; f < NF; f++) fields[f] = &st[(size_t)f * g.n];
  if (viral_upload_state(ctx, 0, fields.data()) != VIRAL_OK) {
    std::fprintf(stderr, "viral_upload_state: %s\n", viral_last_error(ctx));
    return 1;
  }
  double* rec = nullptr;
  if (viral_host_alloc(nrec * sizeof(double), (void**)&rec) != VIRAL_OK) {
    std::fprintf(stderr, "viral_host_alloc failed\n");
    return 1;
  }
  auto output = [&](double t) -> bool {
    if (!o.write) return true;
    if (viral_output(ctx, 0, o.freq, o.freqz, d.lbi, rec, 1) != VIRAL_OK || viral_sync(ctx) != VIRAL_OK) {
      std::fprintf(stderr, "viral_output: %s\n", viral_last_error(ctx));
      return false;
    }
    return write_file(out_name(o, t), rec, nrec);
  };

  // end of init (init.h:688-697)
  double tmax = 0;
  if (o.checks) report_debug(ctx, d, g, tau);
  viral_max_temp_gev(ctx, 0, &tmax);
  std::printf("Time Step is %g @  TempMax is %.4f MeV at TAU -->%.4f This is tau step no. 0\n", ts, 1000 * tmax, tau);
  if (o.checks) report_checkroot(ctx, d, tau);
  if (!output(tau)) return 1;

  // temp.txt of the BJORKEN / BULKTEST builds (main.cpp:124-141)
  FILE* temp = nullptr;
  std::vector<double> fbuf;
  const size_t centre = g.at(g.gx / 2, g.gy / 2, d.lbi ? 0 : d.nz / 2);
  if (o.temp_kind) {
    temp = std::fopen(o.temp_file.c_str(), "w");
    if (!temp) {
      std::fprintf(stderr, "cannot open %s\n", o.temp_file.c_str());
      return 1;
    }
    fbuf.resize(g.n);
  }
  auto temp_line = [&](double t) -> bool {
    if (!temp) return true;
    const int f = o.temp_kind == 1 ? VIRAL_F_EN : VIRAL_F_BULK;
    if (viral_download_field(ctx, 0, f, fbuf.data()) != VIRAL_OK) return false;
    double v = fbuf[centre];
    if (o.temp_kind == 1) {  // fmtoMev(FT(En))/1000: T in GeV (maindef.h:481-485)
      double p, a, T, s_;
      if (viral_eos_eval(&d, 1, &v, &p, &a, &T, &s_) != VIRAL_OK) return false;
      v = (1000 * (T * GEVFM)) / 1000;
    }
    std::fprintf(temp, "%g  %g\n", t, v);
    return true;
  };

  const clock_t start = std::clock();
  double tauPrint = o.tau0;
  int rc = 0;
  for (long l = 1;; l++) {
    if (!temp_line(tau)) { rc = 1; break; }
    if (o.checks) report_physics(ctx, g, tau);
    const int r = viral_step(ctx, tau, ts);
    if (r != VIRAL_OK) {
      int i, j, k, stage;
      unsigned fl;
      if (r == VIRAL_ERR_PHYSICS && viral_first_bad_cell(ctx, 0, &i, &j, &k, &stage, &fl) == VIRAL_OK)
        std::printf("step %ld: stage %d cell (%d, %d, %d) at (%f, %f, %f) flags %u: %s\n", l, stage, i, j, k, g.X[i],
                    g.Y[j], g.E[k], fl, viral_last_error(ctx));
      else
        std::printf("step %ld failed: %s\n", l, viral_last_error(ctx));
      rc = 3;
      break;
    }
    if (o.checks) report_debug(ctx, d, g, tau);
    tau += ts;
    if (o.checks) report_checkroot(ctx, d, tau);
    viral_max_temp_gev(ctx, 0, &tmax);
    const double tmaxMev = 1000 * tmax;
    std::printf("Time Step is %g @  TempMax is %.4f MeV at TAU -->%.4f This is tau step no. %ld at time %.5f seconds\n",
                ts, tmaxMev, tau, l, (double)(std::clock() - start) / CLOCKS_PER_SEC);
    std::fflush(stdout);
    // main.cpp:176-186
    if (std::fabs((tau - tauPrint) - o.pfreq) < 1e-6) {
      tauPrint += o.pfreq;
      if (!output(tau)) { rc = 1; break; }
    }
    if (std::fabs((tau - (o.tau0 + ts))) < 1e-6) {
      if (d.lbi) tauPrint += o.pfreq;  // the LBI branch also advances the print clock (main.cpp:180)
      if (!output(tau)) { rc = 1; break; }
    }
    if (o.tau_stop_profile > 0 && tau > o.tau_stop_profile) break;  // main.cpp:188-200
    if (o.tmin_mev > 0 && tmaxMev < o.tmin_mev) break;
    if (tau > o.tau_stop) break;
    if (o.max_steps >= 0 && l >= o.max_steps) break;
  }
  if (temp) std::fclose(temp);
  viral_host_free(rec);
  viral_destroy(ctx);
  return rc;
}
