/*
 * tpx.h -- C ABI of the MI355X (gfx950) implementation of the RK4 time step
 * of the triply periodic, staggered-grid, low-Mach, particle-laden solver.
 *
 * The reference exposes this path as a C++ class API called by its driver
 * (simulation.cpp:57-93): grid's substep methods (grid.h:112-134),
 * poisson::Solve (poisson.h:79), particle's methods (particle.h:121-138) and
 * the tensor0/tensor1 reductions.  Each entry point below replaces one of
 * those calls; the C++ classes in triplyperiodic-modified_amd/host forward
 * to them with the reference's names.
 *
 * Conventions
 *  - Every call returns 0 on success and a negative TPX_E* code on failure;
 *    tpx_last_error() gives the message.  No call exits the process or lets a
 *    C++ exception escape.
 *  - Host arrays are caller-owned.  Grid arrays use the reference's global
 *    C order [nz][ny][nx] (x fastest).  With bs > 0 they are the reference's
 *    ghosted local arrays (nx+2bs)(ny+2bs)(nz+2bs): ghosts are ignored on
 *    upload and filled with periodic images on download.
 *  - All device state stays resident between calls.  Calls on one context
 *    are ordered on its HIP stream; contexts are independent; one context
 *    must not be used from two host threads at once.
 *  - Grid sizes must be powers of two in [8, 512] (the FFT Poisson solve).
 */
#ifndef TPX_H
#define TPX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TPX_OK 0
#define TPX_E_ARG (-1)      /* invalid argument */
#define TPX_E_HIP (-2)      /* HIP runtime error */
#define TPX_E_STATE (-3)    /* call not valid in the current state */
#define TPX_E_UNSUPP (-4)   /* feature outside this library (HYPRE, ...) */
#define TPX_E_OOM (-5)      /* device allocation failed */

/* The input-file parameters the step reads (params.cpp:23-59, params.h:8-158).
 * Field meaning and units are the reference's. */
typedef struct tpx_params {
  int nx, ny, nz;                 /* Nx_tot Ny_tot Nz_tot */
  double lx, ly, lz;              /* Lx Ly Lz */
  int bs;                         /* bordersize of host arrays */
  double dt, threshold;
  int iteration1, iteration2;
  double epsilon;
  int twc, particle_gravity, cooling;
  double a1, a2, rho0;
  double u0, v0, w0;              /* mean velocity kept by make_mean_U0 */
  double t0, mu0, k;
  double gx1, gy1, gz1, gx2, gy2, gz2;
  double cp, cv, np0, dp, rhop, cvp, epsilonp, nu, i01, i02;
  int solve_for_scalar;           /* solve_for_scalar: advance the passive scalar */
  int s1_type;                    /* S1_type != 0 (the source form of params.h:78) */
  double d_m;                     /* D_M: its molecular diffusivity */
  int solve_for_vector;           /* solve_for_vector: advance the passive vector */
  int s2_type;                    /* S2_type != 0 (the source form of params.h:80) */
  double eta;                     /* eta0: the vector's diffusivity (params.h:79) */
  double rho_forv;                /* Rho_forV: its constant density (params.h:100) */
} tpx_params;

typedef struct tpx_ctx tpx_ctx;

/* field identifiers (grid.h:56-98) */
typedef enum tpx_field {
  TPX_RU = 0, TPX_RU_INT, TPX_RU_NEW, TPX_RU_NP1,  /* tensor1, comp 0..2 */
  TPX_RHS_RU,                                       /* tensor1: deposited drag */
  TPX_RV, TPX_RV_INT, TPX_RV_NEW, TPX_RV_NP1,      /* tensor1: the passive vector (grid.h:65-68) */
  TPX_P = 16, TPX_RHS_POIS, TPX_C, TPX_RHO, TPX_RHO_INT, TPX_RHO_NEW,
  TPX_RHO_NP1, TPX_DIVERGENCE, TPX_RHS_PART_TEMP,
  TPX_PASSIVE_SCALAR, TPX_PASSIVE_SCALAR_INT, TPX_PASSIVE_SCALAR_NEW, TPX_PASSIVE_SCALAR_NP1,
  TPX_Q, TPX_RHS_POIS_Q                             /* the passive vector's pressure and its RHS */
} tpx_field;

/* particle stage identifiers (particle.h:63-113) */
typedef enum tpx_stage { TPX_BASE = 0, TPX_INT = 1, TPX_NEW = 2, TPX_NP1 = 3 } tpx_stage;

/* scalar identifiers for tpx_get_scalars */
typedef enum tpx_scalar {
  TPX_S_P0 = 0, TPX_S_P0_INT, TPX_S_P0_NEW, TPX_S_P0_NP1, TPX_S_T_CUR,
  TPX_S_NUM_TIMESTEP, TPX_S_NUM_ITERATION, TPX_S_MEAN_ENERGY, TPX_S_A,
  TPX_S_I0, TPX_S_ITERATION, TPX_S_GX, TPX_S_GY, TPX_S_GZ, TPX_S_DP0_DT,
  TPX_S_TOTAL_ITERATIONS,   /* Poisson iterations summed over all solves */
  TPX_S_COUNT
} tpx_scalar;

/* ---- lifetime: replaces the grid/poisson/particle constructors
 * (grid.cpp:12-59, poisson.cpp:11-92 incl. fft_3d_create_plan at :87,
 * particle.cpp:14-42).  np_capacity bounds the particle count. */
int tpx_create(const tpx_params *p, int64_t np_capacity, tpx_ctx **out);
int tpx_destroy(tpx_ctx *ctx);
const char *tpx_last_error(const tpx_ctx *ctx);
/* the last error of a failed tpx_create (no context exists) */
const char *tpx_create_error(void);

/* ---- data movement: replaces communicator::read/write (communicator.cpp:65-106)
 * and the public tensor pointers.  comp selects a tensor1 component. */
int tpx_upload_field(tpx_ctx *ctx, int field, int comp, const double *host, int bs);
int tpx_download_field(tpx_ctx *ctx, int field, int comp, double *host, int bs);

/* replaces particle::Load_All / load_random (particle.cpp:899-963): sets the
 * base state; tau / eps (per-particle response time and absorption
 * efficiency) may be NULL to use the scalar Tp / epsilonp of the params. */
int tpx_upload_particles(tpx_ctx *ctx, int64_t n, const double *x, const double *y,
                         const double *z, const double *u, const double *v, const double *w,
                         const double *T, const int32_t *flag, const double *tau,
                         const double *eps);
int tpx_num_particles(tpx_ctx *ctx, int64_t *n);
/* one stage of x,y,z,u,v,w,T (any pointer NULL to skip) and the flags, in the
 * device's current particle order (compare by flag). */
int tpx_download_particles(tpx_ctx *ctx, int stage, double *x, double *y, double *z,
                           double *u, double *v, double *w, double *T, int32_t *flag);
/* per-particle RHS_T (particle.h:113) in the same order */
int tpx_download_particle_rhs_t(tpx_ctx *ctx, double *rhs_t);

/* restart numbers (grid.cpp:232-236): P0, T_cur, num_timestep */
int tpx_set_restart_numbers(tpx_ctx *ctx, double P0, double T_cur, int num_timestep);
int tpx_get_scalars(tpx_ctx *ctx, double *out /* TPX_S_COUNT */);

/* ---- grid methods (grid.h:112-134).  rk is grid::RK4_count. */
int tpx_initialize(tpx_ctx *ctx);                    /* grid.cpp:460-486 */
int tpx_params_update(tpx_ctx *ctx);                 /* params.cpp:75-86 at T_cur */
int tpx_update_rho(tpx_ctx *ctx, int rk);            /* grid.cpp:608-614 */
int tpx_update_p0(tpx_ctx *ctx, int rk);             /* grid.cpp:816-830 */
int tpx_update_particle(tpx_ctx *ctx, int rk);       /* grid.cpp:865-888 */
/* grid.cpp:711-745 (C_Source :686-708); tpx_step runs it after
 * Update_Particle when solve_for_scalar is set (simulation.cpp:65) */
int tpx_update_passive_scalar(tpx_ctx *ctx, int rk);
/* the passive vector (simulation.cpp:66-71; tpx_step runs the four after
 * Update_Passive_Scalar when solve_for_vector is set) */
int tpx_update_rv_woq(tpx_ctx *ctx, int rk);         /* grid.cpp:646-692, V_Source :616-641 */
int tpx_compute_rhs_pois_q(tpx_ctx *ctx, int rk);    /* grid.cpp:694-701 */
int tpx_solve_poisson_q(tpx_ctx *ctx, int rk, int *iterations); /* grid.cpp:703-707 */
int tpx_update_rv_wq(tpx_ctx *ctx, int rk);          /* grid.cpp:709-715 */
int tpx_update_ru_wop(tpx_ctx *ctx, int rk);         /* grid.cpp:776-814 */
int tpx_compute_div_u_new(tpx_ctx *ctx, int rk);     /* grid.cpp:832-839 */
int tpx_compute_rhs_pois(tpx_ctx *ctx, int rk);      /* grid.cpp:841-849 */
/* grid.cpp:851-855 -> poisson::Solve/FFT (poisson.cpp:115-117,391-462);
 * *iterations receives poisson::num_iteration() (may be NULL) */
int tpx_solve_poisson(tpx_ctx *ctx, int rk, int *iterations);
int tpx_update_ru_wp(tpx_ctx *ctx, int rk);          /* grid.cpp:857-863 */
/* poisson::Solve(tensor1& C, tensor0& P, tensor0& RHS) with the caller's face
 * coefficient C (poisson.h:79, poisson.cpp:115-117 -> FFT :391-462): iterates
 *   r = (div_F2C(C grad_C2F P) - RHS) cc,  cc = -1 / I_F2C(C),
 *   P += -Re IFFT(FFT(r) / lambda) / N^3 (zero mode passed through)
 * until Iteration iterations have run or max|dp| <= epsilon (the context's
 * params and params::update state).  cx, cy, cz, P and rhs are host arrays in
 * the ghosted layout with border bs (ghosts ignored on input, filled
 * periodically in P on output); P is the initial guess and receives the
 * result; RHS is used as given (the reference's callers remove its mean).
 * *iterations receives poisson::num_iteration() (may be NULL).  The context's
 * grid state is not touched. */
int tpx_poisson_solve(tpx_ctx *ctx, const double *cx, const double *cy, const double *cz, double *P,
                      const double *rhs, int bs, int *iterations);
/* The same solve on dense device arrays (nz*ny*nx doubles each, x fastest, no
 * ghosts: the caller's inputs cx, cy, cz, rhs and its P, read as the initial
 * guess and overwritten), allocating nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The residual it iterates on lives in a file-scope static device array of
 * 2^21 doubles (16 MiB) per GPU, held until the process ends, so grids of
 * more than 2^21 cells (128^3) are rejected with TPX_E_ARG.  Every call on one
 * GPU shares that GPU's array: calls that overlap on one GPU (two contexts on
 * two host threads) race.  The call returns when the solve has ended. */
int tpx_poisson_solve_device(tpx_ctx *ctx, const double *cx, const double *cy, const double *cz, double *P,
                             const double *rhs, int *iterations);
int tpx_time_advance_rk4(tpx_ctx *ctx);              /* grid.cpp:890-900 */
/* tensor1::make_mean_U0 on RU_np1 (tensor1.cpp:406-411), a = rho0*(U0,V0,W0) */
int tpx_make_mean_u0(tpx_ctx *ctx, const double a[3]);
int tpx_time_advance(tpx_ctx *ctx);                  /* grid.cpp:595-606 */

/* ---- the whole step: simulation.cpp:58-91 with Update_Particle enabled,
 * nsteps times, with the substeps fused where no result changes. */
int tpx_step(tpx_ctx *ctx, int nsteps);

/* ---- a batch of independent boxes (an ensemble; no counterpart in the
 * reference, which runs one box per job): every context advances nsteps
 * whole steps exactly as tpx_step would, with up to `concurrency` contexts
 * in flight at once, each driven by its own host thread on its own stream
 * (<= 0: 16).  Contexts must be distinct.  Returns the first error any
 * context reported; tpx_last_error of that context has the message. */
int tpx_ensemble_step(tpx_ctx *const *ctxs, int n, int nsteps, int concurrency);

/* ---- a batch of boxes stepped by batched kernels (no counterpart in the
 * reference): every kernel of a step is launched once for all boxes, the box
 * in the grid's third dimension and each box's arguments in a table.  The
 * contexts keep their own state, which the other entries read and write as
 * usual between batch steps; each box advances exactly as tpx_step would
 * advance it alone.  They must be distinct, on one device, with equal grid
 * sizes and the same switches that select kernels (TWC, heat exchange on or
 * off, particles present, per-particle tau / eps, Poisson pass count);
 * particle counts, box lengths and all coefficients may differ.  Kernel timing
 * must be off.  tpx_batch_create reports why it failed through
 * tpx_batch_error(); tpx_batch_step through tpx_batch_last_error(). */
/* ---- adjoint of the step (no counterpart in the reference).
 * Objective J = Var(C) over the cells, C the cloud-in-cell concentration of
 * the particles' positions after nsteps steps (particle.cpp:522-555 weights).
 * tpx_adjoint_record advances the context nsteps whole steps exactly as
 * tpx_step would, taping the gas stage fields each particle advance reads and
 * the gas at each step's start (RU, P), checkpointing the particles every
 * `every` steps, and returns J.  For one- or two-way coupling, with or
 * without the heat exchange (under two-way coupling the drag deposits are
 * taped too, with the heat exchange every substep's RHS_Part_Temp, its sum
 * and P0, and the particles' stages are reversed between the gas substeps).
 * tpx_adjoint_gradient then writes dJ/dtau_p for every particle and, if
 * dJ_dstate is not null, dJ/d of the particles' state when recording began
 * (6 arrays of np: x, y, z, u, v, w), both in the context's present particle
 * order (the order tpx_download_particles uses; match by flag): each value
 * belongs to the particle now at that position.
 * tpx_adjoint_gradient_full adds the gas: dJ/dA (the forcing coefficient in
 * effect, one value) and dJ/d of the gas state when recording began, RU
 * (3 x n), Rho (n) and P (n), in tpx_download_field's order with bs = 0; any
 * output may be null, and the gas sweep runs only if a gas output is asked
 * for.  It takes each step of the gas from the tape (kept, or recomputed
 * from the step's start) and reverses it
 * (the momentum, density, projection and Poisson-iteration transposes), and
 * leaves the context's gas where the recorded run ended.  The Poisson solves
 * are differentiated with their recorded iteration counts.
 * Under two-way coupling or with the heat exchange tpx_adjoint_gradient runs
 * the gas sweep as well (the gas depends on the particles).
 * tpx_adjoint_gradient_heat (runs with the heat exchange) writes, from the
 * last sweep (running one if none has run since the recording), dJ/d of each
 * particle's temperature when recording began and dJ/d of its absorption
 * efficiency eps (np values each, in the present particle order), and dJ/d
 * of the initial thermodynamic pressure P0 (one value); the recording's
 * first RHS_T is taken as the exchange of its initial state, as
 * tpx_initialize and every step's end form it.  Any output may be null.
 * tpx_adjoint_release frees the tape (nsteps x 20 grid fields, + 12 under
 * two-way coupling, + 4 with the heat exchange) and the sweep's work arrays. */
int tpx_adjoint_record(tpx_ctx *ctx, int nsteps, int every, double *J);
int tpx_adjoint_gradient(tpx_ctx *ctx, double *dJ_dtau, double *dJ_dstate);
int tpx_adjoint_gradient_full(tpx_ctx *ctx, double *dJ_dtau, double *dJ_dstate, double *dJ_dA, double *dJ_dRU,
                              double *dJ_dRho, double *dJ_dP);
int tpx_adjoint_gradient_heat(tpx_ctx *ctx, double *dJ_dT, double *dJ_deps, double *dJ_dP0);
/* How tpx_adjoint_record keeps the gas for the gas sweep: 0 (default) keeps
 * every substep's momentum, density and Poisson iterates when they fit in
 * half the free device memory, else recomputes each step from its start;
 * 1 always keeps them; 2 always recomputes; 3 checkpoints the gas state
 * (RU, RU_int, Rho, P) every `every` steps, as the particles are, keeps no
 * per-step gas field but the coupling deposits of two-way coupling and the
 * heat exchange, and has the sweep replay each checkpoint segment's gas from
 * its checkpoint into a tape of that segment's steps before reversing them
 * (the sweep then also runs for tpx_adjoint_gradient alone).  Applies to the
 * next recording. */
int tpx_adjoint_set_gas_tape(tpx_ctx *ctx, int mode);
int tpx_adjoint_release(tpx_ctx *ctx);
/* Transpose checks of the Poisson solve's adjoint, as the sweep runs it (no
 * counterpart in the reference; for tests).  x, y, Ax, ATy, ATy_rho: n values
 * each in tpx_download_field's order with bs = 0.  op 0: Ax = M x and
 * ATy = M y, M the iteration's spectral step (transform, division by the
 * modified wavenumbers, inverse, -1/ncell), which the sweep takes as its own
 * transpose.  op 1: Ax = P_1 = x + M(cc (div(C grad x) - 0)), one Poisson
 * iteration from P = x with a zero right-hand side at the context's Rho
 * (C = 1/I_C2F(Rho), cc = -1/I_F2C(C)); ATy and ATy_rho = the adjoints of x
 * and of Rho for the output adjoint y, from the sweep's kernels.  So
 * <Ax, y> = <x, ATy> and d<Ax, y>/dRho = ATy_rho. */
int tpx_adjoint_check(tpx_ctx *ctx, int op, const double *x, const double *y, double *Ax, double *ATy,
                      double *ATy_rho);

typedef struct tpx_batch tpx_batch;
int tpx_batch_create(tpx_ctx *const *ctxs, int n, tpx_batch **out);
int tpx_batch_step(tpx_batch *b, int nsteps);
int tpx_batch_destroy(tpx_batch *b);
const char *tpx_batch_error(void);
const char *tpx_batch_last_error(tpx_batch *b);

/* ---- reductions of tensor0 (tensor0.cpp:364-396,466-478) on a field */
int tpx_field_sum(tpx_ctx *ctx, int field, int comp, double *out);
int tpx_field_max_abs(tpx_ctx *ctx, int field, int comp, double *out);
/* particle::part2gas_concentration (particle.cpp:522-555) into TPX_C, at x_int */
int tpx_deposit_concentration(tpx_ctx *ctx);

/* ---- timing support: the HIP stream of the context (a hipStream_t) and
 * per-kernel-class device times accumulated with HIP events since the last
 * reset, in milliseconds, indexed by tpx_kclass. */
typedef enum tpx_kclass {
  TPX_K_GRID_RHS = 0,   /* density + momentum right-hand side (G1+G4) */
  TPX_K_DIV_RHS,        /* low-Mach divergence + Poisson RHS (G5+G6) */
  TPX_K_FFT,            /* Poisson iteration, first pass: residual, x transform, first y step (G8/G9) */
  TPX_K_PROJECT,        /* pressure-gradient projection (G10) */
  TPX_K_PARTICLE,       /* kept for numbering: the temperature gather (P8) runs inside TPX_K_DEPOSIT */
  TPX_K_DEPOSIT,        /* deposits (P5, P6, P9) with the temperature gather (P8) */
  TPX_K_SORT,           /* cell sort of the particles */
  TPX_K_OTHER,          /* reductions, scalars, rotations */
  TPX_K_ADVANCE,        /* particle RK4 advance with its velocity gather (P1-P4, P7) */
  TPX_K_FFT_YZ,         /* Poisson iteration, middle pass: y and z transforms, division by lambda */
  TPX_K_FFT_INV,        /* Poisson iteration, last pass: inverse y and x transforms, P += dp, loop test */
  TPX_K_COUNT
} tpx_kclass;
int tpx_get_stream(tpx_ctx *ctx, void **stream);
/* ---- grid::Statistics (grid.cpp:902-982) on the device.
 * tpx_statistics_enable(ctx, 1) makes every step keep what the statistics
 * read from inside its last substep (the concentration C at that substep's
 * particle positions, the mean square of its pre-projection velocity
 * RU_new / I_C2F(Rho_new) — the reference's stale U — and the maximum of the
 * gathered gas x velocity); tpx_statistics then reduces the rest over the
 * state at the step boundary and writes TPX_ST_COUNT values (reductions on
 * the device, scalars to the host).  The passive vector is not built: its
 * entries are 0, as they are in the reference with it off. */
enum {
  TPX_ST_TG_MEAN = 0,   /* mean of P0 / (R Rho) */
  TPX_ST_TP_MEAN,       /* mean particle temperature */
  TPX_ST_HT_MEAN,       /* mean of RHS_Part_Temp (mean_energy_transferred) */
  TPX_ST_TKE2,          /* mean_squares of the last substep's pre-projection U */
  TPX_ST_TKE_U, TPX_ST_TKE_V, TPX_ST_TKE_W,  /* mean of U_c RU_c, U = RU / I_C2F(Rho) */
  TPX_ST_PASSIVE_SCALAR_MEAN,  /* mean_squares of the passive scalar (0 if it has no buffers) */
  TPX_ST_TKEV_V1, TPX_ST_TKEV_V2, TPX_ST_TKEV_V3, TPX_ST_VV,  /* 0: no passive vector */
  TPX_ST_P0,
  TPX_ST_C_MAX, TPX_ST_C_MIN, TPX_ST_C_MEAN,
  TPX_ST_RHO_MAX, TPX_ST_RHO_MIN, TPX_ST_RHO_MEAN,
  TPX_ST_PARTICLE_MAX_CFL,  /* 0, as in the reference */
  TPX_ST_GAS_MAX_CFL,       /* max over axes of max(U_c) dt / d_c */
  TPX_ST_GAS_MAX_DIFF_CFL,  /* dt / (dx^2 Rho_min / Mu0) * 6 */
  TPX_ST_NUM_ITERATION,
  TPX_ST_BALANCE_INDEX,     /* Np / N0 on one process */
  TPX_ST_U_MAX,             /* max over axes of max(U_c) */
  TPX_ST_VP_MAX,            /* max particle x velocity */
  TPX_ST_UG_MAX,            /* max gathered gas x velocity, last substep */
  TPX_ST_TP_MAX,            /* max particle temperature */
  TPX_ST_COUNT
};
int tpx_statistics_enable(tpx_ctx *ctx, int on);
int tpx_statistics(tpx_ctx *ctx, double *out /* TPX_ST_COUNT */);
/* particle::trajectory (particle.cpp:717-745): the positions x, y, z of the
 * particles whose flag is m*stride, m < count, into xyz[3m..3m+2] (NaN where no
 * particle has that flag); the reference tracks flags 0, K, 2K, ... below N0
 * with K = ceil(N0 / Np_track) */
int tpx_track_particles(tpx_ctx *ctx, int stride, int count, double *xyz);

/* ---- deposit tile shapes (no counterpart in the reference).
 * The heat and velocity deposit kernels march `zc` planes with `rows` whole
 * x-rows per block.  tpx_tune_tiles starts an online tuner: the next
 * tpx_step calls launch those kernels with each candidate shape in turn
 * (rows 1, 2, 4, 8 x zc 2..32 that tile the grid and fit in LDS), timed with
 * events, and keep the fastest once each has run twice (results do not
 * depend on the shape).  tpx_get_tiles writes {heat rows, heat zc, velocity rows,
 * velocity zc, tuning still running}; 0 means the built-in shape.
 * tpx_set_tiles sets the four values (0, 0 for the built-in shape), e.g. to
 * give every box of an ensemble the shape tuned on one. */
int tpx_tune_tiles(tpx_ctx *ctx);
int tpx_get_tiles(tpx_ctx *ctx, int *out /* 5 */);
int tpx_set_tiles(tpx_ctx *ctx, const int *in /* 4 */);

int tpx_timing_enable(tpx_ctx *ctx, int on);
int tpx_timing_read(tpx_ctx *ctx, double *ms /* TPX_K_COUNT */, int64_t *launches /* TPX_K_COUNT */);
int tpx_synchronize(tpx_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
