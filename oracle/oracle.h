/*
 * oracle.h -- CPU restatement of the reference solver's RK4 time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * implementation and the CPU timing baseline of bench.py.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product path (libtpx) never links or calls it.
 *
 * It restates, operator by operator and in the same floating-point order,
 * the reference's single-rank arithmetic:
 *   grid.cpp:18-25 (RK4 tables), 460-486 (Initialize tail), 595-614,
 *   776-900 (substep methods), poisson.cpp:391-462 (FFT iteration),
 *   tensor0.cpp:54-296,354-396,466-478, tensor1.cpp:24-311,320-325,
 *   particle.cpp:58-134,150-311,314-654, params.cpp:61-86,
 *   communicator.cpp:108-216 (ghost fill), 228-324 (ghost fold).
 * Fields are stored WITH the reference's ghost layers (width bs) and every
 * element-wise operation runs over ghosts too, exactly like tensor0.
 *
 * The one deliberate deviation: the reference's 3-D FFT (Sandia fft_3d over
 * FFTW 2.1.5, absent from the reference tree) is replaced by the radix-2 FFT
 * in or_fft.c, pinned against a brute-force DFT and numpy.fft in tests.
 *
 * How each part is pinned (tests/test_oracle_operators.py,
 * tests/test_oracle_physics.py), since the reference itself cannot be built
 * here and ships no golden vectors:
 *   stencils S1-S9, fold    single-Fourier-mode closed forms, summation by
 *                           parts, brute-force fold
 *   or_fft3d, ccp_fft       brute-force DFT, numpy fixture, mode / zero-mode
 *                           rule, exact inverse of the 7-point Laplacian
 *   or_solve_poisson        uniform-coefficient exactness, residual decrease
 *   or_update_rho (G1)      mass conservation
 *   or_update_p0 (G2)       exact RK4 of a linear ODE
 *   or_update_ru_wop (G4)   momentum/energy conservation of convection,
 *                           viscous single modes, forcing, body force, drag
 *   or_compute_div_u_new, or_compute_rhs_pois, or_update_ru_wp (G5, G6, G10)
 *                           single-mode closed forms
 *   or_make_mean_u0 (G12)   exact mean
 *   or_initialize (G14)     known temperature gap -> known heat source
 *   particle gathers and deposits (P1, P5, P6, P8, P9): exact transposes,
 *                           partition of unity, separable exactness
 *   particle RK4 (P2-P4)    frozen uniform gas = classical RK4 matrix map
 *   or_p_send_recv (P7)     ids kept, one-period shifts into [0, L]
 *   or_update_passive_scalar (grid.cpp:686-745)
 *                           RK4 amplification of a diffusing single mode and of
 *                           a mode carried by a uniform velocity (closed
 *                           forms), exact growth under either source form,
 *                           conservation of the scalar's sum
 *   or_step (the composition of simulation.cpp:57-93, params::update)
 *                           whole steps of cases that collapse to linear
 *                           systems with closed-form RK4 maps
 *                           (tests/test_oracle_step_composition.py): shear-
 *                           mode decay with forcing, the particle-temperature
 *                           / P0 pair with the lagged heat source, the drag
 *                           pair with make_mean_U0 once per step, the body
 *                           force on a z-varying density, the threshold
 *                           switch of A, g, I0 and Iteration, and a
 *                           divergence-free start staying divergence-free;
 *                           copies of this file with a planted composition
 *                           mistake (RU_new from RU_int, Update_P0 after
 *                           Update_Particle, make_mean_U0 per substep,
 *                           Rho_face not rebuilt, no threshold switch) each
 *                           fail at least one of them.
 * Parity unpinned: the order in which or_p_send_recv re-appends wrapped
 * particles (every comparison matches particles by flag, so it cannot change
 * a result).
 */
#ifndef TPX_ORACLE_H
#define TPX_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Input-file parameters (params.cpp:23-59), single rank. */
typedef struct or_params {
  int nx, ny, nz;
  double lx, ly, lz;
  int bs;
  double dt, threshold;
  int iteration1, iteration2;
  double epsilon;
  int twc, particle_gravity, cooling;
  double a1, a2, rho0;
  double u0, v0, w0;
  double t0, mu0, k;
  double gx1, gy1, gz1, gx2, gy2, gz2;
  double cp, cv, np0, dp, rhop, cvp, epsilonp, nu, i01, i02;
  int solve_for_scalar, s1_type;  /* passive scalar switch and source form */
  double d_m;                     /* its molecular diffusivity */
  int solve_for_vector, s2_type;  /* passive vector switch and source form (S2_type != 0) */
  double eta, rho_forv;           /* its diffusivity eta0 and constant density Rho_forV */
} or_params;

typedef struct or_state or_state;

/* field identifiers */
enum {
  OR_RU = 0, OR_RU_INT, OR_RU_NEW, OR_RU_NP1, OR_RHS_RU, OR_U, OR_RHO_FACE,
  OR_DUMMY, OR_DUMMY2, OR_PS_FACE,
  OR_RV, OR_RV_INT, OR_RV_NEW, OR_RV_NP1, OR_RHS_RV, OR_V, OR_S2,  /* tensor1 (component 0..2) */
  OR_P = 32, OR_RHS_POIS, OR_C, OR_RHO, OR_RHO_INT, OR_RHO_NEW, OR_RHO_NP1,
  OR_T, OR_DIVERGENCE, OR_RHS_PART_TEMP, OR_RHS_RHO,
  OR_PS, OR_PS_INT, OR_PS_NEW, OR_PS_NP1, OR_RHS_PS, OR_S1,  /* tensor0; PS the passive scalar */
  OR_Q, OR_RHS_POIS_Q                                      /* the passive vector's pressure */
};
/* particle stage identifiers */
enum { OR_BASE = 0, OR_INT = 1, OR_NEW = 2, OR_NP1 = 3 };
/* particle per-particle auxiliary arrays */
enum { OR_AUX_RHS_U = 0, OR_AUX_RHS_V, OR_AUX_RHS_W, OR_AUX_UG, OR_AUX_VG,
       OR_AUX_WG, OR_AUX_TG, OR_AUX_RHS_T, OR_AUX_TAU, OR_AUX_EPS };
/* scalar identifiers for or_get_scalars (fixed order) */
enum { OR_S_P0 = 0, OR_S_P0_INT, OR_S_P0_NEW, OR_S_P0_NP1, OR_S_T_CUR,
       OR_S_NUM_TIMESTEP, OR_S_NUM_ITERATION, OR_S_MEAN_ENERGY, OR_S_A,
       OR_S_I0, OR_S_ITERATION, OR_S_GX, OR_S_GY, OR_S_GZ, OR_S_DP0_DT,
       OR_S_TOTAL_ITERATIONS, OR_S_COUNT };

or_state *or_create(const or_params *p, int np_capacity);
void or_destroy(or_state *s);
int or_nthreads(int n); /* set OpenMP threads (n<=0: leave); returns count in use */

/* interior layout (nz, ny, nx) x fastest; set fills ghosts */
int or_set_field(or_state *s, int id, int comp, const double *interior);
int or_get_field(const or_state *s, int id, int comp, double *interior);
/* ghosted layout ((nz+2bs)(ny+2bs)(nx+2bs)) */
int or_get_field_ghosted(const or_state *s, int id, int comp, double *ghosted);

int or_set_particles(or_state *s, int n, const double *x, const double *y,
                     const double *z, const double *u, const double *v,
                     const double *w, const double *T, const int *flag,
                     const double *tau, const double *eps);
int or_get_np(const or_state *s);
/* q[7] = x,y,z,u,v,w,T of one stage; any pointer may be NULL */
int or_get_particles(const or_state *s, int stage, double *const *q, int *flag);
int or_get_particle_aux(const or_state *s, int which, double *out);
int or_set_particle_aux(or_state *s, int which, const double *in);

void or_get_scalars(const or_state *s, double *out /* OR_S_COUNT */);
void or_set_restart_numbers(or_state *s, double P0, double T_cur, int num_timestep);

/* reference methods (grid.h:112-134, particle.h:121-138, params::update) */
void or_initialize(or_state *s);              /* grid.cpp:460-486 */
void or_params_update(or_state *s);           /* params.cpp:75-86 with T_cur */
void or_set_rk4_count(or_state *s, int rk);
void or_update_rho(or_state *s);              /* grid.cpp:608-614 */
void or_update_p0(or_state *s);               /* grid.cpp:816-830 */
void or_update_passive_scalar(or_state *s);   /* grid.cpp:711-745 */
void or_update_rv_woq(or_state *s);           /* grid.cpp:646-692 (V_Source :616-641) */
void or_compute_rhs_pois_q(or_state *s);      /* grid.cpp:694-701 */
void or_solve_poisson_q(or_state *s);         /* grid.cpp:703-707 */
void or_update_rv_wq(or_state *s);            /* grid.cpp:709-715 */
void or_update_particle(or_state *s);         /* grid.cpp:865-888 */
void or_update_ru_wop(or_state *s);           /* grid.cpp:776-814 */
void or_compute_div_u_new(or_state *s);       /* grid.cpp:832-839 */
void or_compute_rhs_pois(or_state *s);        /* grid.cpp:841-849 */
void or_solve_poisson(or_state *s);           /* grid.cpp:851-855 */
void or_update_ru_wp(or_state *s);            /* grid.cpp:857-863 */
/* poisson::Solve with the caller's face coefficient (poisson.cpp:115-117,
 * 391-462); interior layout, P in/out */
void or_poisson_solve(or_state *s, const double *cx, const double *cy, const double *cz, double *P,
                      const double *rhs, int *iters);
void or_time_advance_rk4(or_state *s);        /* grid.cpp:890-900 */
void or_make_mean_u0(or_state *s, const double a[3]); /* tensor1.cpp:406-411 on RU_np1 */
void or_time_advance(or_state *s);            /* grid.cpp:595-606 */
/* the body of simulation.cpp:57-93 with Update_Particle enabled, nsteps times */
void or_step(or_state *s, int nsteps);

/* ---- primitives exposed for closed-form pins (operate on the state's
 * scratch tensors; inputs/outputs in interior layout) ---- */
enum { OR_OP_DIV_F2C = 0,   /* a=(x,y,z) faces -> out0 centre      S5 */
       OR_OP_GRAD_C2F,      /* a0 centre -> out(x,y,z) faces        S7 */
       OR_OP_DEL2,          /* a0 -> out0                            S6 */
       OR_OP_I_C2F,         /* a0 centre -> out(x,y,z)               S8 */
       OR_OP_I_F2C,         /* a=(x,y,z) -> out0 (1/3 mean)          S9 */
       OR_OP_IX_C2F, OR_OP_IY_C2F, OR_OP_IZ_C2F, /* a0 -> out0       S4 */
       OR_OP_IX_F2C, OR_OP_IY_F2C, OR_OP_IZ_F2C,
       OR_OP_FOLD,          /* ghosted a0 -> ghost fold -> interior out0 */
       OR_OP_CCP_FFT        /* a0 -> out0 (constant-coefficient solve) G9 */
};
int or_apply(or_state *s, int op, const double *a0, const double *a1,
             const double *a2, double *out0, double *out1, double *out2);
/* reductions over the interior in the reference's loop order (S11) */
double or_sum(or_state *s, const double *interior);
double or_max_abs(or_state *s, const double *interior);

/* particle primitives on the current particle set (P1, P5, P6, P8, P9) */
void or_p_gather_velocity(or_state *s);          /* U -> ug,vg,wg at x_int */
void or_p_deposit_velocity(or_state *s);         /* RHS_u.. at x_int -> RHS_RU */
void or_p_deposit_concentration(or_state *s);    /* x_int -> C */
void or_p_gather_temp(or_state *s, int stage);   /* T field -> Tg at x_int/x_new */
void or_p_deposit_temp(or_state *s, int stage);  /* RHS_T -> RHS_Part_Temp */
void or_p_send_recv(or_state *s);                /* wrap + reorder */

/* complex 3-D FFT on an interleaved (re,im) array, x fastest, unnormalised;
 * sign=-1 forward, +1 inverse */
void or_fft3d(double *data, int nx, int ny, int nz, int sign);

#ifdef __cplusplus
}
#endif
#endif
