/*
 * oracle.h -- CPU restatement of the VIRAL viscous hydro step.
 *
 * TEST INFRASTRUCTURE.  This is the parity oracle: only tests/, the smoke
 * check and bench.py's CPU-baseline leg may load it, and only as the checker
 * or the timed CPU baseline -- never as the product path.
 *
 * It restates NEWEST10 (Milne) / TXYZ5 (Cartesian) one pass at a time, in the
 * reference's loop structure and operation order, on the structure-of-arrays
 * ghosted layout of include/viral.h.  Each function cites the reference lines
 * it follows in oracle.c.
 */
#ifndef VIRAL_ORACLE_H
#define VIRAL_ORACLE_H
#include <stdint.h>
#include "../include/viral.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_ctx or_ctx;

or_ctx* or_create(const viral_desc* d);
void or_destroy(or_ctx* c);
/* state field f (ghosted, reference order) */
double* or_field(or_ctx* c, int f);
size_t or_cells(const or_ctx* c);

/* one full RK step (NEWEST10/time.h:300-478); returns VIRAL_OK or VIRAL_ERR_PHYSICS */
int or_step(or_ctx* c, double tau, double dtau);
/* the state was written from outside (VIRAL_Q_FIX_VE takes BMax anew) */
void or_state_changed(or_ctx* c);
const uint8_t* or_flags(const or_ctx* c);
const int32_t* or_iters(const or_ctx* c, int stage);
int or_first_bad(const or_ctx* c, int* i, int* j, int* k, int* stage, unsigned* flags);

/* unit-level entry points (used to pin the oracle against the reference) */
void or_calc_source(or_ctx* c, double tau, double ts, double* source /* 15*cells, field-major */);
void or_hydro_explicit(or_ctx* c, double tau, double taustep, double* result /* 15*cells */);
/* root solve + velocities (+ k0 broadcast) on the interior at tau; per-cell
 * iteration counts into iters (may be NULL); returns VIRAL_OK/VIRAL_ERR_PHYSICS */
int or_root(or_ctx* c, double tau, int32_t* iters, uint8_t* flags);
void or_boundary(or_ctx* c);
double or_max_temp_gev(or_ctx* c);

/* EOS and transport, scalar (for tests) */
double or_eos_p(const viral_desc* d, double e);
double or_eos_dpde(const viral_desc* d, double e);
double or_eos_T(const viral_desc* d, double e);
double or_eos_s(const viral_desc* d, double e, double p, double T);

/* Single-cell Newton solve (NEWEST10/root.h:301-364): returns the root and
 * writes the iteration count and solver flag bits. */
double or_newton_cell(const viral_desc* d, double tau, double T00, double T10, double T20, double T30,
                      double pi0, double pi1, double pi2, double pi3, double PI, double guess,
                      int32_t* iters, unsigned* flags);

/* The VIRAL_Q_ROOT_SAFEGUARD solve of one cell: the root of f on [0, A] by
 * bracketed Newton-bisection; NaN when no bracket exists (M2 >= A (A - PI)). */
double or_root_safeguard(const viral_desc* d, double M2, double A, double PI, int32_t* iters);

/* 1 (default): detect period-2..4 orbits of the Newton iteration and jump to
 * the iterate the loop would end on; 0: run the loop to max_iter. */
void or_set_newton_shortcut(int on);
long or_newton_cycle_count(int period);

/* Cartesian pieces (TXYZ5): the full pi^{mu nu} (slots tt,tx,ty,tz,xx,xy,xz,
 * yy,yz,zz) from the five evolved components {xx,yy,xy,xz,yz}; the 4-D root
 * solve of one cell, prm = {T^tt, T^tx, T^ty, T^tz, p1..p5, PI}, x = guess
 * (e, u^x, u^y, u^z) in / root out; returns the iteration count. */
void or_cart_pi10(const double u[4], const double* p5, double pi10[10]);
int or_cart_gnewton(const viral_desc* d, const double* prm, double* x, unsigned* flags);
/* its residuals and Jacobian (J16[4r+c] = d y_r / d x_c) */
void or_cart_f(const viral_desc* d, const double* prm, const double* x, double* y);
void or_cart_df(const viral_desc* d, const double* prm, const double* x, double* J16);
/* centred fluxes of direction dir on every cell, [15][cells] */
void or_centre_flux(or_ctx* c, int dir, double tau, double* F);

/* The Israel-Stewart shear source of one cell (du16[4a+b] = d_b u^a). */
void or_shear_source(int milne, double tau, const double u[4], const double pi10[10], const double du16[16],
                     double SH, double tpi, double out[10]);

/* WENO kernels of NEWEST10/utils.h:57-160 */
double or_weno_L(double qm2, double qm1, double qc, double qp1, double qp2);
double or_weno_R(double qm2, double qm1, double qc, double qp1, double qp2);
double or_weno_der(double qm2, double qm1, double qc, double qp1, double qp2, double dx);

/* Per-step diagnostics of the driver loop.  Cell positions are ghosted
 * linear indices ((i*gy + j)*gz + k), -1 where there is none.
 * or_check_root: CheckRoot (NEWEST10/root.h:1-199: 6 statistics; TXYZ5/
 * multiroot.h:7-150: 9); returns the count.  or_debug_msg: DebugMSG extrema
 * (NEWEST10/debug.h:3-100), out = {vmax, vx, vy, ve, emax, emin}, at =
 * {vmax, emax, emin cells}.  or_check_physics: CheckPhysics (utils.h:
 * 231-258), counts and first cells of |v| >= 1 and e <= 0. */
int or_check_root(or_ctx* c, double tau, double* mx, double* avg, long long* at);
void or_debug_msg(or_ctx* c, double tau, double out[6], long long at[3]);
void or_check_physics(or_ctx* c, double tau, long long n[2], long long first[2]);
/* write.h output records (17 doubles per output cell); returns the count of
 * doubles, 0 when the strides do not divide the grid; out may be NULL */
size_t or_output(or_ctx* c, int freq, int freqz, int xy, double* out);

/* CalcNS (mode 0: Navier-Stokes pi into nspi, 10 x cells), NSInit (1) and
 * ZeroInit (2) over every ghosted cell (NEWEST10/Source.h:542-776); -1 on a
 * Cartesian context or bad mode. */
int or_ns_init(or_ctx* c, double tau, double ts, int mode, double* nspi);

/* Landau matching of a T^{mu nu} field (10 x cells) and the per-cell part of
 * ginit (NEWEST10/ginit.h:521-629, 876-990); returns the number of cells
 * without exactly one passing eigenvector, the first of them in *first. */
long or_landau(or_ctx* c, double tau, const double* tmunu, int at_start, long long* first);

/* deterministic math (oracle/vmath.h) exported for tests */
double or_vm_exp(double x);
double or_vm_log(double x);
double or_vm_pow(double x, double y);

#ifdef __cplusplus
}
#endif
#endif
