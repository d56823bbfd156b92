/*
 * viral.h -- C ABI of the MI355X-native viscous hydrodynamics step.
 *
 * The reference (rjfries/VIRAL) has no plugin API; its step is the call
 *   TVDRK2 / TVDRK3 / FirstOrder(GRID HydroGrid, double tau, double ts)
 * (NEWEST10/time.h:300,338,397, called from NEWEST10/main.cpp:144-152), which
 * reads and writes a global array-of-structs grid plus the globals il..kr, k0
 * (NEWEST10/init.h:7,16) and compile-time profile macros
 * (NEWEST10/maindef.h:36-339).  This header replaces that seam: the grid
 * becomes a context owning structure-of-arrays device buffers, the profile
 * macros become fields of viral_desc, and every entry point returns a status
 * instead of calling exit().
 *
 * Cell layout (all host/device buffers): ghosted grid of
 *   Gx = nx + 2*B, Gy = ny + 2*B, Gz = (lbi ? 1 : nz + 2*B), B = VIRAL_BORDER,
 * linear index ((i*Gy + j)*Gz + k) -- the reference's HydroGrid[i][j][k]
 * order (NEWEST10/maindef.h:463), eta fastest.  Interior cells are
 * i in [B, B+nx), j in [B, B+ny), k in [B, B+nz) (k = 0 when lbi), matching
 * il..ir, jl..jr, kl..kr of NEWEST10/init.h:570-581.
 *
 * All quantities are in the reference's units: fm, 1/fm, 1/fm^4.
 */
#ifndef VIRAL_H
#define VIRAL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VIRAL_BORDER 3

/* coordinates (NEWEST10 = Milne, TXYZ5 = Cartesian) */
enum { VIRAL_MILNE = 0, VIRAL_CARTESIAN = 1 };
/* flux scheme: -DKT (hydroKT.h) or -DSHAS (hydroshasta.h, implies CON) */
enum { VIRAL_KT = 0, VIRAL_SHASTA = 1 };
/* equation of state: conformal profile inlines or s95p.h */
enum { VIRAL_EOS_CONFORMAL = 0, VIRAL_EOS_S95P = 1 };
/* shear viscosity law, NEWEST10/maindef.h:76,134,282 (law 1) and :187,229,328 (law 0) */
enum { VIRAL_ETA_TIMES_S = 0,      /* eta = c * s              */
       VIRAL_ETA_S_OVER_4PI = 1 }; /* eta = c * (s / (4 pi))   */
/* relaxation time law, NEWEST10/maindef.h:329 (0), :77,135,283 (1), :188,230 (2) */
enum { VIRAL_TPI_OVER_EP = 0,      /* tau_pi = (c*eta)/(e+p)   */
       VIRAL_TPI_OVER_P = 1,       /* tau_pi = c*(eta/p)       */
       VIRAL_TPI_CONST = 2 };      /* tau_pi = c               */
/* bulk viscosity law FZeta, NEWEST10/maindef.h:191 (const) and :82,140,289 (c s/(4 pi)) */
enum { VIRAL_ZETA_CONST = 0,       /* zeta = c                 */
       VIRAL_ZETA_S_OVER_4PI = 1 };/* zeta = c * (s / (4 pi))  */
/* bulk relaxation time law FtauPI, NEWEST10/maindef.h:192 (const) and :83,141,290 (c zeta/p) */
enum { VIRAL_TPIB_CONST = 0,       /* tau_PI = c               */
       VIRAL_TPIB_OVER_P = 1 };    /* tau_PI = c * zeta / p    */

/* Reference behaviours kept as switchable modes (all on = "as reference"). */
#define VIRAL_Q_GHOST_EIGEN   (1u << 0) /* q1: eigenvalues only in the interior; first face of each row gets zero
                                           dissipation (NEWEST10/hydroKT.h:182-184,284,453-455) */
#define VIRAL_Q_K0_BROADCAST  (1u << 1) /* q2: copy v of slice k0 to all eta slices after each root solve
                                           (NEWEST10/root.h:426-449) */
#define VIRAL_Q_DPDE_OF_P     (1u << 2) /* q3: a = DPDE(P) in the CON eigenvalues (NEWEST10/maindef.h:490) */
#define VIRAL_Q_GUBSER_MIRROR (1u << 3) /* q7: mirror u^1,u^2 into x/y ghosts (NEWEST10/boundary.h:32-37,77-81,
                                           118-122,160-165); off by default: only the GUBSER profile has it;
                                           needs nx, ny >= 3 (viral_create refuses narrower Milne grids) */
#define VIRAL_Q_SHEAR_SLICES  (1u << 4) /* q9, Cartesian only: shear sources only in ghosted eta slices k < nz,
                                           as the reference's loop k = 0..ZCMA-1 (TXYZ5/Source.h:272); the top 3
                                           interior slices keep the zero sources of the freshly allocated grid */
#define VIRAL_Q_ROOT_SAFEGUARD (1u << 5) /* not a reference behaviour (the reference aborts): a Milne cell whose
                                           Newton solve fails (VIRAL_FLAG_SOLVER, or a root e <= 0) is solved again
                                           by bracketed Newton-bisection on [0, A], A = T^tt - pi^tt, where a root
                                           exists (M^2 < A (A - PI)); such cells carry VIRAL_FLAG_RESCUED.  Off by
                                           default; cells whose Newton solve succeeds are unchanged either way */
#define VIRAL_Q_FIX_VE        (1u << 6) /* the FIX build (NEWEST10/time.h:280-290, 328, 385, 471; FixTwo): at
                                           the end of every step, after the ghost fill, zero the stored v_eta of
                                           every interior cell when BMax > 0.98, BMax being DebugMSG's maximum of
                                           sqrt(vx^2 + vy^2 + tau^2 ve^2) over the interior (debug.h:30-74) at the
                                           end of the previous step (or of the uploaded state, at the first
                                           step's tau), then take BMax of the new state.  Milne only; off in every
                                           reference profile */
#define VIRAL_Q_SHASTA_FULLSTEP (1u << 7) /* not a reference behaviour: the SHASTA rate uses the full step dtau as
                                           its sub-step instead of the reference's dtau/2 (RK2) or dtau/3 (RK3)
                                           (q6, NEWEST10/time.h:352,411).  The reference's choice divides each
                                           direction's 1/8 diffusion by the sub-step, which makes the 3-D
                                           checkerboard mode of a fluid at rest grow by 2.5 per RK2 step and 8.6
                                           per RK3 step (its antidiffusion limiter is off at extrema); with dtau
                                           it decays (0.625, 0.0625).  Off by default */
#define VIRAL_Q_DEFAULT (VIRAL_Q_GHOST_EIGEN | VIRAL_Q_K0_BROADCAST | VIRAL_Q_DPDE_OF_P | VIRAL_Q_SHEAR_SLICES)

/* State fields, each one ghosted array of doubles.  Names follow the
 * reference's cell members (NEWEST10/maindef.h:399-405).
 *
 * Cartesian (TXYZ5) events use the same arrays: T^{t mu} in VIRAL_F_T00..T30,
 * the five evolved shear components of TXYZ5/maindef.h:240 (pi[0..4] =
 * pi^{xx}, pi^{yy}, pi^{xy}, pi^{xz}, pi^{yz}) in VIRAL_F_PI0..PI0+4, and
 * VIRAL_F_VE = v^z.  Slots VIRAL_F_PI0+5..PI0+9 are not evolved (the step only
 * copies them into ghosts).  The other five components follow from
 * u_mu pi^{mu nu} = 0 and pi^mu_mu = 0. */
enum {
  VIRAL_F_T00 = 0, VIRAL_F_T10, VIRAL_F_T20, VIRAL_F_T30,       /* T^{tau mu}                     */
  VIRAL_F_PI0,                                                  /* pi[0..9] = pi^{tt,tx,ty,te,xx,xy,xe,yy,ye,ee}
                                                                   (NEWEST10/time.h:194)          */
  VIRAL_F_BULK = VIRAL_F_PI0 + 10,                              /* PI (bulk)                      */
  VIRAL_F_EN, VIRAL_F_P, VIRAL_F_TEMP,                          /* En, P, Temp                    */
  VIRAL_F_VX, VIRAL_F_VY, VIRAL_F_VE,                           /* Vx, Vy, Ve                     */
  VIRAL_F_U0, VIRAL_F_U1, VIRAL_F_U2, VIRAL_F_U3,               /* u[4]                           */
  VIRAL_F_PU0, VIRAL_F_PU1, VIRAL_F_PU2, VIRAL_F_PU3,           /* prevu[4]                       */
  VIRAL_NFIELDS                                                 /* = 29                           */
};

/* per-cell flag bits written by every root solve */
#define VIRAL_FLAG_ACAUSAL    (1u << 0) /* v^2 >= 1            (NEWEST10/root.h:396)           */
#define VIRAL_FLAG_ENEG       (1u << 1) /* e <= 0              (NEWEST10/root.h:396)           */
#define VIRAL_FLAG_SOLVER     (1u << 2) /* f' == 0 or non-finite f, f' in the Newton step
                                           (the GSL errors that abort the reference)           */
#define VIRAL_FLAG_MAXITER    (1u << 3) /* Newton hit max_iter without two equal iterates      */
#define VIRAL_FLAG_FREEZEOUT  (1u << 4) /* Temp < t_freeze                                     */
#define VIRAL_FLAG_RESCUED    (1u << 5) /* root found by the VIRAL_Q_ROOT_SAFEGUARD bracketed solve */

/* status codes */
#define VIRAL_OK            0
#define VIRAL_ERR_ARG      -1
#define VIRAL_ERR_HIP      -2
#define VIRAL_ERR_PHYSICS  -3 /* a cell raised ACAUSAL, ENEG or SOLVER; see viral_first_bad_cell */
#define VIRAL_ERR_UNSUPPORTED -4

typedef struct viral_desc {
  int nx, ny, nz;           /* interior cells (XCMA, YCMA, ZCMA)                          */
  int lbi;                  /* 1: boost-invariant build (-DLBI): nz == 1, no eta ghosts   */
  int coords;               /* VIRAL_MILNE / VIRAL_CARTESIAN                              */
  int scheme;               /* VIRAL_KT / VIRAL_SHASTA                                    */
  int rk;                   /* 1, 2, 3 (FirstOrder, TVDRK2, TVDRK3)                        */
  int eos;                  /* VIRAL_EOS_*                                                */
  double dx, dy, deta;      /* XS, YS, ZS                                                 */
  double dof;               /* DOF of the conformal EOS (42.25, NEWEST10/maindef.h:9)     */
  int eta_law;  double eta_coef;
  int tpi_law;  double tpi_coef;
  unsigned quirks;          /* VIRAL_Q_* bits                                             */
  int k0;                   /* ghosted eta index broadcast by q2; < 0: (Gz-1)/2 as the
                               reference (NEWEST10/init.h:641-645)                        */
  int max_iter;             /* Newton iteration cap (5000, NEWEST10/root.h:346)           */
  double t_freeze;          /* freeze-out temperature in 1/fm for VIRAL_FLAG_FREEZEOUT    */
  int fast_math;            /* 0: arithmetic in the reference's operation order, bitwise
                               equal to the CPU restatement; 1: reciprocal-multiply/FMA
                               forms, within 1e-10 relative per step                     */
  int eos_table;            /* fast mode only: 1 evaluates s95p from a table (4-point
                               interpolation, 1024 nodes per octave of e, closed form next
                               to its branch points and below 1.2e-4 fm^-4) instead of
                               the closed form                                            */
  int bulk;                 /* 1: the -DBULK build (BULKTEST profile, NEWEST10/maindef.h:
                               157-158): Israel-Stewart bulk source of PI
                               (NEWEST10/Source.h:519-527; TXYZ5/Source.h:327-333), the
                               Navier-Stokes PI = -zeta theta of NSInit (Source.h:740-745)
                               and the kept PI of the Landau matching (ginit.h:996-999).
                               0: PI's source is zero and the initial states set PI = 0  */
  int zeta_law;  double zeta_coef;  /* VIRAL_ZETA_*: bulk viscosity (FZeta)            */
  int tpib_law;  double tpib_coef;  /* VIRAL_TPIB_*: bulk relaxation time (FtauPI)     */
} viral_desc;

/* Fill *d with the reference's defaults for an nx*ny*nz Milne KT RK2 grid. */
void viral_desc_default(viral_desc* d, int nx, int ny, int nz);

/* Ghosted sizes of a descriptor. */
size_t viral_cells(const viral_desc* d);
void viral_ghost_dims(const viral_desc* d, int* gx, int* gy, int* gz);

typedef struct viral_ctx viral_ctx;

/* Create a context on the current HIP device holding nev events described by
 * descs[0..nev-1] (an ensemble: events may differ in grid size, EOS and
 * transport coefficients; scheme, rk, fast_math and coords must agree).
 * Replaces init()'s allocation (NEWEST10/init.h:548-552, NEWEST10/alloc.h:33).  Returns VIRAL_OK or an error. */
int viral_create(const viral_desc* descs, int nev, viral_ctx** out);
void viral_destroy(viral_ctx* ctx);

/* Copy the VIRAL_NFIELDS state arrays of event ev between host and device.
 * fields[f] points to viral_cells(desc_ev) doubles (ghosted, reference order).
 * Replaces direct HydroGrid access by init (NEWEST10/init.h:653-686) and
 * output (NEWEST10/write.h:87-103). */
int viral_upload_state(viral_ctx* ctx, int ev, const double* const* fields);
int viral_download_state(viral_ctx* ctx, int ev, double* const* fields);
/* One field only (f in [0, VIRAL_NFIELDS)). */
int viral_download_field(viral_ctx* ctx, int ev, int f, double* out);

/* Advance every event by one full RK step from tau to tau+dtau, including the
 * ghost fill after each stage.  Replaces TVDRK2/TVDRK3/FirstOrder
 * (NEWEST10/time.h:300-478) and the halo + boundary calls inside them
 * (NEWEST10/pack.h:689, NEWEST10/boundary.h:266).  Returns VIRAL_OK, or
 * VIRAL_ERR_PHYSICS where the reference would have exited
 * (NEWEST10/root.h:396-421). */
int viral_step(viral_ctx* ctx, double tau, double dtau);

/* nsteps steps from tau, tau advancing by dtau between them as in the
 * reference's time loop (NEWEST10/main.cpp:144-156), with one host
 * synchronisation at the end.  *done (may be NULL) receives the number of
 * steps that completed before the first failing one (nsteps if none failed);
 * on VIRAL_ERR_PHYSICS the state is that of the last step run and
 * viral_first_bad_cell describes the first failing step.  The reference
 * checks every step and exits on the first failure; drivers that act on a
 * failure mid-run call viral_step once per step instead. */
int viral_steps(viral_ctx* ctx, double tau, double dtau, int nsteps, int* done);

/* Per-cell flags (uint8, ghosted layout, OR over the stages of the last step)
 * and Newton iteration counts of stage s (0-based) of the last step (int32). */
int viral_flags(viral_ctx* ctx, int ev, uint8_t* out);
int viral_iterations(viral_ctx* ctx, int ev, int stage, int32_t* out);
/* First interior cell, in the reference's i,j,k loop order, that raised
 * ACAUSAL/ENEG/SOLVER in the last step's first failing stage; -1 if none.
 * Writes the ghosted (i,j,k), the 0-based stage and that cell's flags in
 * that stage. */
int viral_first_bad_cell(viral_ctx* ctx, int ev, int* i, int* j, int* k, int* stage, unsigned* flags);

/* RootSearchForEnVelUsingDerivatives (NEWEST10/root.h:212-450) alone, on the
 * event's current state in place: every interior cell's e (Newton from its
 * current e, with the VIRAL_Q_ROOT_SAFEGUARD re-solve when that quirk is set),
 * P, Temp, v and u from its T^{tau mu}, pi and PI at tau, then the k0
 * broadcast.  Iteration counts and flags land where stage 0 of a step puts
 * them (viral_iterations(ctx, ev, 0), viral_flags); a cell raising ACAUSAL,
 * ENEG or SOLVER gives VIRAL_ERR_PHYSICS and viral_first_bad_cell.  Milne
 * events only (VIRAL_ERR_UNSUPPORTED otherwise); ghosts are not filled. */
int viral_root(viral_ctx* ctx, int ev, double tau);

/* Navier-Stokes initialisation of the event's current state at tau, the time
 * derivative of u taken from prevu over ts (replaces CalcNS, NSInit and
 * ZeroInit, NEWEST10/Source.h:542-776, called from init() at
 * NEWEST10/init.h:658-686).  Every ghosted cell, with the velocity
 * gradients CalcDer4Vel (Source.h:3-101) leaves on a freshly zeroed grid:
 *   VIRAL_NS_OUTPUT  the Navier-Stokes pi = 2 eta sigma^{mu nu} (CalcNS's nspi)
 *                    into nspi[0..9], host arrays of viral_cells() doubles in
 *                    the pi index map, and with desc.bulk the Navier-Stokes
 *                    PI = -zeta theta (nsPI, Source.h:740-745) into nspi[10];
 *                    the state is unchanged
 *   VIRAL_NS_INIT    NSInit: pi = that, PI = nsPI (0 without bulk),
 *                    T^{tau mu} recomputed from e, P, u, pi and PI
 *   VIRAL_ZERO_INIT  ZeroInit: pi = 0, PI = 0, T^{tau mu} recomputed
 * Exact arithmetic and the closed-form EOS in every mode.  Milne events only
 * (VIRAL_ERR_UNSUPPORTED otherwise: TXYZ5 declares but never defines them).
 * VIRAL_NS_OUTPUT allocates 10 (11) x cells doubles of staging on first use
 * (kept until viral_destroy, shared with viral_output). */
enum { VIRAL_NS_OUTPUT = 0, VIRAL_NS_INIT = 1, VIRAL_ZERO_INIT = 2 };
int viral_ns_init(viral_ctx* ctx, int ev, double tau, double ts, int mode, double* const* nspi);

/* Landau matching of a T^{mu nu} field onto the event's current state
 * (replaces FindVar and the per-cell part of ginit, NEWEST10/ginit.h:521-629
 * and 876-990, which init() calls at tau - ts and at tau, init.h:658-659).
 * tmunu[0..9]: host arrays of viral_cells() doubles, the contravariant
 * components {tt, tx, ty, teta, xx, xy, xeta, yy, yeta, etaeta} at tau, for
 * every ghosted cell.  Each cell takes the eigenvector of T^mu_nu that
 * FindVar would select (real, with |v_i| < 1 and |v| < 1 for v = x^i/x^0;
 * exactly one such) and sets e = |lambda|, Temp, P, v and u; at_start = 0
 * (the call at tau - ts) also sets prevu = u and nothing else; at_start = 1
 * sets pi = T - T_ideal - PI (g^{-1} - u u) with PI = P - e/3, then PI = 0
 * (kept when desc.bulk, as ginit.h:996-999) and T^{tau mu} from e, P, u and pi.  Where no single eigenvector passes
 * (the reference prints EIGEN VALUE SEARCH FAILED and exits) the cell keeps
 * its state and gets VIRAL_FLAG_SOLVER in viral_flags, and the call returns
 * VIRAL_ERR_PHYSICS with viral_first_bad_cell naming the first such cell in
 * loop order.  Overwrites the flags of every cell.  Exact arithmetic and the
 * closed-form EOS.  Milne events only.  Allocates 10 x cells doubles of
 * staging on first use (shared with viral_output). */
int viral_landau(viral_ctx* ctx, int ev, double tau, const double* const* tmunu, int at_start);

/* Device reductions the reference's driver runs every step.  Each reads the
 * event's current state on the device (no state download).  Cells are
 * reported as ghosted (i, j, k) triples, -1 -1 -1 where there is none.
 * Sums are combined in a fixed order, so repeated calls give equal results.
 *
 * MaxTempGev (NEWEST10/utils.h:263-283, called at main.cpp:163): max over
 * the interior of GEVFM*T(e). */
int viral_max_temp_gev(viral_ctx* ctx, int ev, double* out);
/* CheckRoot (NEWEST10/root.h:1-199, called at main.cpp:158 with tau after
 * the step; Cartesian TXYZ5/multiroot.h:7-150).  Milne: 6 statistics,
 *   0 |e - (T^tt - pi^tt) + M^2 / ((T^tt - pi^tt) + P(e) - PI)|,
 *   1-4 |pi^{mu nu} u_nu| for mu = t, x, y, eta, 5 |pi^mu_mu|;
 * Cartesian: 9, 0-3 the residuals of the four conservation equations at
 * (e, u), 4-7 transversality, 8 the trace as the reference keeps it (its
 * `fabs(trace) > trace` test: |trace| of the last cell in loop order with
 * a negative trace, 0 if none).  max[s]: maximum over the interior, with
 * ijk[3s..3s+2] the first cell in loop order reaching it; avg[s]: sum over
 * the interior / number of interior cells (the reference's Avg for one
 * rank).  Arrays hold 9 entries (27 for ijk, which may be NULL); *nstat
 * receives the count.  P(e) is the closed-form EOS. */
int viral_check_root(viral_ctx* ctx, int ev, double tau, double* max, double* avg, int* ijk, int* nstat);
/* viral_check_root without any allocation: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results, bit for bit.  Its block partials and combined
 * statistics live in a fixed device array of the library (DIAG_BLOCKS = 512
 * blocks x 9 statistics + 9, 110.8 KB per GPU, held until the process ends;
 * every grid fits, since the reduction strides over the interior with at most
 * 512 blocks, and a layout that would not fit is rejected with
 * VIRAL_ERR_ARG).  All calls on one GPU share that array and each call
 * overwrites what it reads: calls that overlap on one GPU (several threads,
 * or contexts on several streams) race. */
int viral_check_root_noalloc(viral_ctx* ctx, int ev, double tau, double* max, double* avg, int* ijk, int* nstat);
/* DebugMSG (NEWEST10/debug.h:3-100, called at the end of every step,
 * time.h:334,391,477; Cartesian TXYZ5/debug.h with |v| free of tau):
 * out[0] the largest |v| = sqrt(vx^2 + vy^2 + tau^2 ve^2) (the reference's
 * BMax; the last cell in loop order reaching it), out[1..3] that cell's
 * (vx, vy, ve), out[4] the largest e (first cell), out[5] the smallest e
 * below 0 (first cell; 0 when none).  ijk: 9 ints (vmax, emax, emin cells),
 * may be NULL. */
int viral_debug_extrema(viral_ctx* ctx, int ev, double tau, double* out, int* ijk);
/* CheckPhysics (NEWEST10/utils.h:231-258, called at the start of every step,
 * time.h:303,341,399): count[0] interior cells with |v| >= 1, count[1] with
 * e <= 0; ijk: the first of each in loop order (6 ints, may be NULL). */
int viral_check_physics(viral_ctx* ctx, int ev, double tau, long long* count, int* ijk);

/* Output records of NEWEST10/write.h (WriteResults :286-365, WriteResultsXY
 * :52-119): 17 doubles per output cell, {e, T, P, vx, vy, ve, pi[0..9], PI}
 * (write.h:87-103), for every freq-th cell in x and y and every freqz-th in
 * eta, in (i, j, k) order -- exactly the bytes of the reference's
 * res/tau<tau>fm.bin for one rank.  xy (or an LBI grid): the k0 slice only.
 * The strides must divide the grid (VIRAL_ERR_ARG otherwise).
 * viral_output_size gives the count of doubles.  viral_output packs the
 * records on the device and copies them to out; with async != 0 it returns
 * once the copy is queued on the context's stream (out must then be pinned
 * memory from viral_host_alloc, and stay untouched until viral_sync), so the
 * caller can write the previous file while the next step runs. */
int viral_output_size(const viral_desc* d, int freq, int freqz, int xy, size_t* ndoubles);
int viral_output(viral_ctx* ctx, int ev, int freq, int freqz, int xy, double* out, int async);
/* Waits for the context's queued work (steps, async output copies). */
int viral_sync(viral_ctx* ctx);
/* Pinned host memory for asynchronous output copies. */
int viral_host_alloc(size_t bytes, void** p);
int viral_host_free(void* p);

/* Host evaluation of the equation of state a descriptor selects: pressure,
 * dp/de, temperature and entropy density at n energy densities (1/fm^4).
 * The reference's EOS/DPDE/FT/FS profile inlines (NEWEST10/maindef.h:68-72,
 * 274-278) and s95p_p/a/T/s (NEWEST10/s95p.h); with eos_table set, the
 * same table the fast-mode kernels read. */
int viral_eos_eval(const viral_desc* d, size_t n, const double* e, double* p, double* dpde, double* T, double* s);

/* ---- Launch layer (no counterpart in the reference, a CPU code with no
 * tiling, batching or launches).
 *
 * Every RK stage is a fixed sequence of kernel phases (flux x, flux y, flux
 * eta, sources, update, k0 broadcast, ghost fill).  Each phase splits every
 * event into tiles (one block's worth of work); the tile table lists, per
 * phase, the first tile of each event.
 *
 *   batched        1: each phase is ONE launch over all events of the context,
 *                  a block finding its (event, tile) in the table; 0: one
 *                  launch per event and phase.  Default: 1 when nev > 1.
 *   graph          1: the launches of one RK step are captured in a hipGraph
 *                  once (per starting state copy) and replayed every step; step
 *                  times reach the kernels through a device-resident clock.
 *                  Per-kernel timers force direct launches while on.
 *   blocks_per_cu  0: a block per tile; k > 0: persistent launch of k blocks per
 *                  compute unit striding over the tiles.
 *   fusion         stage-fusion depth: 0 sources in their own kernel, 1 sources
 *                  evaluated inside the update kernel.
 *   xcd            1: permute tile ids so each XCD (blocks b with equal b % 8)
 *                  receives a contiguous range, sharing halos in its L2.
 *   flux_nt[3]     threads per block of the fast KT flux kernel per direction:
 *                  64, 128, 256 or 512 (1, 2, 4, 8 wavefronts).
 * The launch configuration never changes results: every cell and face sees
 * the same arithmetic whatever the tiling, batching or fusion. */
typedef struct viral_launch {
  int batched, graph, blocks_per_cu, fusion, xcd;
  int flux_nt[3];
} viral_launch;
int viral_get_launch(viral_ctx* ctx, viral_launch* out);
int viral_set_launch(viral_ctx* ctx, const viral_launch* in);

/* kernel phases of the tile table */
enum { VIRAL_PHASE_FLUX_X = 0, VIRAL_PHASE_FLUX_Y, VIRAL_PHASE_FLUX_ETA, VIRAL_PHASE_CELLS, VIRAL_PHASE_K0,
       VIRAL_PHASE_GHOST, VIRAL_NPHASE };
/* Tiles of one event in a phase for flux tile shape (tp, tk) (used by the fast
 * KT flux phases only; host arithmetic, no device needed); -1 on bad input. */
long long viral_tile_count(const viral_desc* d, int phase, int tp, int tk);
/* The context's tile table of a phase: start[0..nev], event e owning tiles
 * [start[e], start[e+1]). */
int viral_tile_table(viral_ctx* ctx, int phase, long long* start);
/* Tile that block slot w of a `total`-tile launch processes when the XCD order
 * is on (a permutation of [0, total)); -1 on bad input. */
long long viral_xcd_tile(long long w, long long total);

/* LDS tile shape of the fast KT flux kernel of direction dir (0 x, 1 y,
 * 2 eta): tp cells along the direction (tp-5 faces, tp-6 written cells per
 * tile) times tk lines, tp >= 8, tp*tk at most 4*flux_nt[dir] for flux_nt 64
 * and 128, 2*flux_nt[dir] for 256 and 512, and at most 1024.
 * viral_autotune times candidates on the current state at the clock of the
 * last step (reps launches each) and keeps the fastest: with ev >= 0 the tile
 * shapes of event ev; with ev < 0 the whole launch configuration (per
 * direction the block size and one shape for all events, then fusion depth,
 * persistent blocks and XCD order).  It overwrites the last step's flags and
 * iteration counts, never the state. */
int viral_set_tile(viral_ctx* ctx, int ev, int dir, int tp, int tk);
int viral_get_tile(viral_ctx* ctx, int ev, int dir, int* tp, int* tk);
int viral_autotune(viral_ctx* ctx, int ev, int reps);

/* ---- Adjoint (reverse mode) of a recorded run.  No counterpart in the
 * reference (it has no derivatives).
 *
 * viral_adjoint_begin(ctx, every) starts recording: it keeps a device copy of
 * the current state and, from then on, the (tau, dtau) of every step taken by
 * viral_step / viral_steps and a copy of the state every `every` steps.
 * viral_adjoint(ctx, seed, grad, pgrad) then runs the recorded steps in
 * reverse: between checkpoints it recomputes the stage states forward and
 * applies hand-written adjoint kernels (KT flux and WENO reconstruction,
 * sources, EOS -- the closed forms or the derivative of the table's
 * interpolant --, velocities; the root by the implicit-function theorem at
 * the converged root: the 1-D Newton of the Milne step, the 4-D gnewton of the
 * Cartesian one).  seed[f] holds dJ/d(final state field f) (VIRAL_NFIELDS
 * ghosted arrays), grad[f] receives dJ/d(recorded initial state field f), each
 * field taken as an independent input, and pgrad[0], pgrad[1] receive dJ/d
 * eta_coef and dJ/d tpi_coef.  The state is left as it was.  Single-event KT
 * contexts in Milne or Cartesian coordinates, conformal or s95p EOS, with or
 * without the bulk source (not SHASTA or the FIX build: VIRAL_ERR_UNSUPPORTED).
 * viral_adjoint_end frees the recording.  With checkpoints the recording
 * also keeps the stage states of the current segment (the steps since the
 * last checkpoint) in a ring of every * RK stages state copies, so the last
 * segment is reversed without a recompute; viral_adjoint recomputes each
 * earlier segment j on a second stream of the context while it reverses
 * segment j + 1, into two device buffers of (every * RK stages + 1) state
 * copies that it allocates and frees itself.  The recorded steps do not
 * replay a hipGraph.
 *
 * every == 0 keeps every stage's output state instead of checkpoints (the
 * stages write straight into the recording; no recompute in viral_adjoint):
 * NFIELDS * cells * 8 bytes per RK stage of every recorded step, allocated by
 * each viral_steps call for its steps (63 GB for 200 RK2 steps of a
 * 128x128x32 grid).  The recorded steps do not replay a hipGraph. */
int viral_adjoint_begin(viral_ctx* ctx, int every);
int viral_adjoint(viral_ctx* ctx, const double* const* seed, double* const* grad, double* pgrad);
int viral_adjoint_end(viral_ctx* ctx);

/* Per-kernel timing with HIP events recorded on the context's stream around
 * every launch (instrumentation; the reference only times whole steps with
 * clock(), NEWEST10/main.cpp:107,170-171).  viral_timers(ctx, 1) clears and
 * starts recording; viral_timer_read returns the summed duration and launch
 * count of kernel id k in [0, viral_kernel_count()).  The last id,
 * "adjoint_device", times the device part of each viral_adjoint call (its
 * recompute and reverse kernels, without the seed's upload and the
 * gradient's download). */
int viral_timers(viral_ctx* ctx, int enable);
int viral_timer_read(viral_ctx* ctx, int kernel, double* total_ms, long long* launches);
int viral_kernel_count(void);
const char* viral_kernel_name(int kernel);

/* Error text of the last failing call on ctx (never NULL). */
const char* viral_last_error(viral_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif /* VIRAL_H */
