/*
 * ref_harness.cpp -- drives the reference's own step functions, compiled in
 * place from /root/reference/NEWEST10 (see oracle/build_ref.sh).
 *
 * TEST INFRASTRUCTURE.  The reference as a whole cannot be built here:
 * main.cpp, root.h and ginit.h need GSL, which the image lacks.  The other
 * step headers (maindef.h, s95p.h, alloc.h, init.h, utils.h, pack.h,
 * boundary.h, debug.h, Source.h, write.h, hydroKT.h, hydroshasta.h) need only MPI, which the
 * image has (MPICH under /opt/conda).  This harness includes them exactly as
 * NEWEST10/main.cpp:1-81 does (same defines, same order, minus the GSL
 * headers and root.h), fills the grid with a seeded state and calls
 *   CalcSource       (NEWEST10/Source.h:195)
 *   CalcCentreFlux + hydroExplicit (NEWEST10/hydroKT.h:16,406)
 *   boundary         (NEWEST10/boundary.h:266)
 *   s95p_p/a/s/T     (NEWEST10/s95p.h)
 * and writes inputs and outputs to a binary file that the tests compare with
 * the oracle.  init() is never called (it would need root.h's CheckRoot); the
 * globals it would set (index ranges, k0, MPI coordinates of a single rank)
 * are set here the way NEWEST10/init.h:548-645 sets them for one rank.
 *
 * Usage: ref_<profile> <out.bin> <seed> <tau> <ts> <mode>
 *   mode 0: source + flux + boundary on a random state
 *   mode 1: s95p samples
 *   mode 2: the driver-loop diagnostics and output on a random state with a
 *           few planted cells (|v| >= 1, e <= 0, and a repeated largest e and
 *           |v|): DebugMSG (debug.h:3) and CheckPhysics (utils.h:231) print
 *           to stdout, BMax and MaxTempGev (utils.h:263) go to the file, and
 *           WriteResults / WriteResultsXY (write.h:52,286) write
 *           res/tau<tau>fm.bin under the working directory.
 *   mode 3: the initial state of the profile's own init (initvar,
 *           init.h:150-214, for the default profile), after the cell
 *           positions init() sets (init.h:558-566).
 *   mode 4: CalcNS (Source.h:643-751) on a random state: the Navier-Stokes
 *           pi (nspi, nsPI), then the state after NSInit (Source.h:755-776).
 *   mode 5: the state after ZeroInit (Source.h:542-562) of a random state.
 *   mode 6: under mpiexec with NPX*NPY ranks (the GUBSER profile's 6x6):
 *           mpi_init (init.h:118-141), then boundary() on the same random
 *           state in every rank; the two corner ranks write the state before
 *           and after it to <out>.<xcord>_<ycord>.
 */
#include <mpi.h>
#include <iostream>
#include <fstream>
#include <iomanip>
#include <stdlib.h>
#include <math.h>
#include <time.h>
#include <sstream>
#include <string>
#include <cstring>
#include <float.h>
#include <vector>
#include <stdio.h>
#include <stdint.h>

#define GEVFM 0.1973
#define PIE 3.141592653589793

using namespace std;

#include "s95p.h"
#include "maindef.h"
#include "alloc.h"
#include "init.h"
#include "utils.h"
#include "pack.h"
#include "boundary.h"
#include "debug.h"
#include "Source.h"
#include "write.h"
#ifdef KT
#include "hydroKT.h"
#endif
#ifdef SHAS
#include "hydroshasta.h"
#endif

/* splitmix64, used only to fill the grid */
static uint64_t rng_state;
static double urand()
{
  uint64_t z = (rng_state += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  z = z ^ (z >> 31);
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

static void put(FILE* f, const double* p, size_t n) { fwrite(p, sizeof(double), n, f); }

int main(int argc, char** argv)
{
  if (argc < 6) { fprintf(stderr, "usage: %s out seed tau ts mode\n", argv[0]); return 2; }
  const char* out = argv[1];
  rng_state = strtoull(argv[2], 0, 10);
  double tau_ = atof(argv[3]);
  double ts_ = atof(argv[4]);
  int mode = atoi(argv[5]);
  /* mode 6 runs on several ranks: each corner rank opens its own file below */
  FILE* f = mode == 6 ? NULL : fopen(out, "wb");
  if (!f && mode != 6) return 3;

  if (mode == 1) {
    int n = 4000;
    int hdr[2] = {1, n};
    fwrite(hdr, sizeof(int), 2, f);
    for (int q = 0; q < n; q++) {
      /* energies from 1e-7 to ~1e3 1/fm^4, log-spaced with jitter, plus the break points */
      double e = pow(10.0, -7.0 + 10.0 * (q + urand()) / n);
      double v[5] = {e, s95p_p(e), s95p_a(e), s95p_s(e), s95p_T(e)};
      put(f, v, 5);
    }
    fclose(f);
    return 0;
  }

  MPI_Init(&argc, &argv);
  rank = 0;
  size = 1;
  /* init() makes mpi_grid a Cartesian communicator over the ranks
   * (NEWEST10/init.h:121-144); with one process that is the world */
  mpi_grid = MPI_COMM_WORLD;
  /* boundary() asks MYXLEFT/MYXRIGHT (NEWEST10/init.h:98-103), which are
   * BOUNDARY on both sides only for a one-rank profile (NPX == NPY == 1);
   * other profiles skip the boundary call below. */
  xcord = 0;
  ycord = 0;
  zcord = 0;
  if (mode == 6) {
    /* run under mpiexec with NPX*NPY ranks: the reference's own mpi_init
     * (NEWEST10/init.h:118-141) builds the Cartesian communicator and sets
     * this rank's coordinates, so boundary() fills exactly the faces that lie
     * on the domain edge (MYXLEFT/MYXRIGHT/MYYLEFT/MYYRIGHT == BOUNDARY) */
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    mpi_init();
  }
  const int one_rank = (NPX == 1 && NPY == 1);
  StoreInHeap();
  AllocatePack();
  il = jl = BORDER;
  ir = il + XCMA;
  jr = jl + YCMA;
#if !defined LBI
  kl = BORDER;
  kr = kl + ZCMA;
  k0 = int((ZCM - 1) / 2);
#else
  kl = 0;
  kr = 1;
  k0 = 0;
#endif
  ts = ts_;
  tau = tau_;

  /* a smooth field with seeded random modes and a little cell noise */
  double ph[12];
  for (int q = 0; q < 12; q++) ph[q] = 6.283185307179586 * urand();
  for (int i = 0; i < XCM; i++)
    for (int j = 0; j < YCM; j++)
      for (int k = 0; k < ZCM; k++) {
        cell& c = HydroGrid[i][j][k];
        double x = 0.13 * i, y = 0.11 * j, z = 0.07 * k;
        double e = 4.0 + 2.0 * sin(x + ph[0]) * cos(y + ph[1]) + 0.8 * sin(z + ph[2]) + 0.2 * urand();
        c.En = e;
        c.P = EOS(e);
        c.PI = 0;
#ifdef BULK
        /* the -DBULK build (BULKTEST profile) evolves PI: give it a value */
        c.PI = 0.05 * e * (urand() - 0.5);
#endif
        double u1 = 0.3 * sin(0.7 * x + ph[3]) + 0.05 * urand();
        double u2 = 0.3 * cos(0.6 * y + ph[4]) + 0.05 * urand();
        double u3 = 0.1 * sin(0.9 * z + ph[5]) / tau + 0.01 * urand();
        c.u[1] = u1; c.u[2] = u2; c.u[3] = u3;
        c.u[0] = sqrt(1 + u1 * u1 + u2 * u2 + tau * tau * u3 * u3);
        for (int l = 0; l < 4; l++) c.prevu[l] = c.u[l] * (1.0 + 0.02 * (urand() - 0.5));
        for (int l = 0; l < 10; l++) c.pi[l] = 0.2 * e * (urand() - 0.5) * (l == 3 || l == 6 || l == 8 ? 1.0 / tau : 1.0) *
                                                (l == 9 ? 1.0 / tau : 1.0);
        c.T00 = (e + c.P - c.PI) * c.u[0] * c.u[0] - (c.P - c.PI) + c.pi[0];
        c.T10 = (e + c.P - c.PI) * c.u[0] * u1 + c.pi[1];
        c.T20 = (e + c.P - c.PI) * c.u[0] * u2 + c.pi[2];
        c.T30 = (e + c.P - c.PI) * c.u[0] * u3 + c.pi[3];
        c.Vx = u1 / c.u[0]; c.Vy = u2 / c.u[0]; c.Ve = u3 / c.u[0];
        c.Temp = FT(e);
        /* cell positions as init() sets them (NEWEST10/init.h:44-81) */
        c.X = XCORDWB(i);
        c.Y = YCORDWB(j);
        c.eta = ZCORDWB(k);
      }
  if (mode == 3) {
#if !defined(GINIT) && !defined(BJORKEN) && !defined(GUBSER) && !defined(FLUCT)
    for (int i = 0; i < XCM; i++)
      for (int j = 0; j < YCM; j++)
        for (int k = 0; k < ZCM; k++) {
          cell& c = HydroGrid[i][j][k];
          c.X = XCORDWB(i);
          c.Y = YCORDWB(j);
          c.eta = ZCORDWB(k);
          c.r = sqrt(c.X * c.X + c.Y * c.Y);
        }
    initvar(HydroGrid, tau, ts);
#else
    fclose(f);
    MPI_Finalize();
    return 4;
#endif
  }
  if (mode == 2) {
    /* planted cells, in interior loop order: two with the same largest
     * |v| >= 1, two with the same largest e, two with e <= 0 */
    int ci[6][3] = {{BORDER + 1, BORDER + 2, kl}, {ir - 2, jr - 1, kr - 1}, {BORDER + 3, BORDER, kl},
                    {ir - 1, BORDER + 1, kl}, {BORDER + 2, BORDER + 4, kr - 1}, {ir - 3, jr - 2, kl}};
    for (int q = 0; q < 6; q++) {
      cell& c = HydroGrid[ci[q][0]][ci[q][1]][ci[q][2]];
      if (q < 2) { c.Vx = 0.8; c.Vy = -0.7; c.Ve = 0.0; }
      else if (q < 4) c.En = 90.0;
      else c.En = q == 4 ? -0.5 : 0.0;
    }
  }
  int hdr[8] = {0, XCM, YCM, ZCM, BORDER, ZCMA, one_rank, 0};
#ifdef LBI
  hdr[7] = 1;
#endif
  if (mode == 6) {
    /* every rank holds the same seeded state; the ranks at the two corners of
     * the rank grid write it before and after boundary() to
     * <out>.<xcord>_<ycord> (the (0,0) rank fills the left x and y ghosts,
     * the (NPX-1,NPY-1) rank the right ones) */
    const int corner = (xcord == 0 && ycord == 0) || (xcord == NPX - 1 && ycord == NPY - 1);
    if (corner) {
      char name[4096];
      snprintf(name, sizeof(name), "%s.%d_%d", out, xcord, ycord);
      f = fopen(name, "wb");
      if (!f) MPI_Abort(MPI_COMM_WORLD, 3);
      hdr[6] = size;
      fwrite(hdr, sizeof(int), 8, f);
      double th6[2] = {tau, ts};
      fwrite(th6, sizeof(double), 2, f);
    }
  } else {
    fwrite(hdr, sizeof(int), 8, f);
  }
  if (mode == 6) {
    size_t n6 = (size_t)XCM * YCM * ZCM;
    std::vector<double> b6(n6);
#define DUMP6(EXPR)                                                    \
  do {                                                                  \
    size_t q = 0;                                                       \
    for (int i = 0; i < XCM; i++)                                       \
      for (int j = 0; j < YCM; j++)                                     \
        for (int k = 0; k < ZCM; k++) { cell& c = HydroGrid[i][j][k]; b6[q++] = (EXPR); } \
    if (f) put(f, &b6[0], n6);                                          \
  } while (0)
#define DUMPSTATE6()                                                    \
  do {                                                                  \
    DUMP6(c.T00); DUMP6(c.T10); DUMP6(c.T20); DUMP6(c.T30);             \
    for (int l = 0; l < 10; l++) DUMP6(c.pi[l]);                        \
    DUMP6(c.PI); DUMP6(c.En); DUMP6(c.P); DUMP6(c.Temp); DUMP6(c.Vx); DUMP6(c.Vy); DUMP6(c.Ve); \
    for (int l = 0; l < 4; l++) DUMP6(c.u[l]);                          \
    for (int l = 0; l < 4; l++) DUMP6(c.prevu[l]);                      \
  } while (0)
    DUMPSTATE6();
    boundary(HydroGrid);
    DUMPSTATE6();
    if (f) fclose(f);
    MPI_Finalize();
    return 0;
  }
  double th[2] = {tau, ts};
  put(f, th, 2);
  size_t n = (size_t)XCM * YCM * ZCM;
  std::vector<double> buf(n);
#define DUMP(EXPR)                                                      \
  do {                                                                  \
    size_t q = 0;                                                       \
    for (int i = 0; i < XCM; i++)                                       \
      for (int j = 0; j < YCM; j++)                                     \
        for (int k = 0; k < ZCM; k++) { cell& c = HydroGrid[i][j][k]; buf[q++] = (EXPR); } \
    put(f, &buf[0], n);                                                 \
  } while (0)
  /* inputs in the order of include/viral.h's field enum */
  DUMP(c.T00); DUMP(c.T10); DUMP(c.T20); DUMP(c.T30);
  for (int l = 0; l < 10; l++) DUMP(c.pi[l]);
  DUMP(c.PI); DUMP(c.En); DUMP(c.P); DUMP(c.Temp); DUMP(c.Vx); DUMP(c.Vy); DUMP(c.Ve);
  for (int l = 0; l < 4; l++) DUMP(c.u[l]);
  for (int l = 0; l < 4; l++) DUMP(c.prevu[l]);

  if (mode == 3) {
    fclose(f);
    MPI_Finalize();
    return 0;
  }
  if (mode == 4 || mode == 5) {
    /* on the freshly allocated grid, as at init: du is zero wherever
     * CalcDer4Vel does not write it */
    if (mode == 4) {
      CalcNS(HydroGrid, tau, ts);
      for (int l = 0; l < 10; l++) DUMP(c.nspi[l]);
      DUMP(c.nsPI);
      NSInit(HydroGrid, tau, ts);
    } else {
      ZeroInit(HydroGrid, tau, ts);
    }
    DUMP(c.T00); DUMP(c.T10); DUMP(c.T20); DUMP(c.T30);
    for (int l = 0; l < 10; l++) DUMP(c.pi[l]);
    DUMP(c.PI); DUMP(c.En); DUMP(c.P); DUMP(c.Temp); DUMP(c.Vx); DUMP(c.Vy); DUMP(c.Ve);
    for (int l = 0; l < 4; l++) DUMP(c.u[l]);
    for (int l = 0; l < 4; l++) DUMP(c.prevu[l]);
    fclose(f);
    MPI_Finalize();
    return 0;
  }
  if (mode == 2) {
    DebugMSG(HydroGrid, tau);
    CheckPhysics(HydroGrid, tau, 0);
    double extra[6] = {BMax, MaxTempGev(HydroGrid), (double)NPX, (double)NPY, (double)FREQ, (double)FREQZ};
    put(f, extra, 6);
    fclose(f);
#ifdef LBI
    WriteResultsXY(tau, HydroGrid);
#else
    WriteResults(tau, HydroGrid);
#endif
    MPI_Finalize();
    return 0;
  }
  CalcSource(HydroGrid, tau, ts);
  for (int l = 0; l < SVAR; l++) DUMP(c.Source[l]);
#ifdef KT
  CalcCentreFlux(HydroGrid, tau);
  hydroExplicit(HydroGrid, tau, ts / 2);
  for (int l = 0; l < SVAR; l++) DUMP(c.Result[l]);
#elif defined SHAS
  CalcNumVel(HydroGrid, tau);
  hydroExplicit(HydroGrid, tau, ts / 2);
  for (int l = 0; l < SVAR; l++) DUMP(c.Result[l]);
#else
  for (int l = 0; l < SVAR; l++) DUMP(0.0);
#endif
  if (one_rank) boundary(HydroGrid);
  DUMP(c.T00); DUMP(c.T10); DUMP(c.T20); DUMP(c.T30);
  for (int l = 0; l < 10; l++) DUMP(c.pi[l]);
  DUMP(c.PI); DUMP(c.En); DUMP(c.P); DUMP(c.Temp); DUMP(c.Vx); DUMP(c.Vy); DUMP(c.Ve);
  for (int l = 0; l < 4; l++) DUMP(c.u[l]);
  for (int l = 0; l < 4; l++) DUMP(c.prevu[l]);
  fclose(f);
  MPI_Finalize();
  return 0;
}
